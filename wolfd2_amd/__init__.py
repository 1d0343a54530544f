"""wolfd2_amd -- MI355X (gfx950) HIP implementation of the wolfd2 per-timestep hot path.

* :mod:`wolfd2_amd.case`  -- case set-up (grid metrics, region tables, scalings)
* :mod:`wolfd2_amd.api`   -- the reference's call interface on the GPU, and the
  device-resident :class:`~wolfd2_amd.api.Solver`
* ``libwolfd2_amd.so``    -- the HIP kernels behind the C ABI of include/wolfd2_amd.h
"""
from . import case  # noqa: F401

__all__ = ["case", "api"]
