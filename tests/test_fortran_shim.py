"""The Fortran binding (wolfd2_amd/fortran/wd2_shim.f) compiles on its own
against a caller's config.f and exports the reference's entry points with the
reference's symbol names, each calling the matching C entry point."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "wolfd2_amd", "fortran", "wd2_shim.f")
FC = shutil.which("amdflang") or shutil.which("flang")

# a caller's compile-time extents, as include/config.f of a wolfd2 build defines them
CONFIG = """\
      INTEGER mnx, mny, mgri, mgrj
      parameter ( mnx  = 1030 )
      parameter ( mny  = 1030 )
      parameter ( mgri =   20 )
      parameter ( mgrj =   10 )
"""

EXPORTS = {"nauxmomentum_": "wd2_naux_momentum", "ppe_": "wd2_ppe", "project_": "wd2_project",
           "velboundcond_": "wd2_vel_bound_cond", "presboundcond_": "wd2_pres_bound_cond",
           "veloutflowbcs_": "wd2_vel_outflow_bcs", "dmaxnorm_": "wd2_dmax_norm",
           "diffmaxnorm_": "wd2_diff_max_norm", "tempboundcond_": "wd2_temp_bound_cond",
           "thermenergy_": "wd2_therm_energy", "eqstate_": "wd2_eq_state", "filter_": "wd2_filter"}


@pytest.mark.skipif(FC is None, reason="no Fortran compiler (amdflang) in this image")
def test_shim_compiles_and_exports_reference_names(tmp_path):
    (tmp_path / "config.f").write_text(CONFIG)
    obj = tmp_path / "wd2_shim.o"
    r = subprocess.run([FC, "-ffixed-form", "-O2", "-c", SHIM, "-o", str(obj), f"-I{tmp_path}"],
                       cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", str(obj)], capture_output=True, text=True, check=True).stdout
    defined = {ln.split()[-1] for ln in nm.splitlines() if " T " in ln}
    undefined = {ln.split()[-1] for ln in nm.splitlines() if " U " in ln}
    for fsym, csym in EXPORTS.items():
        assert fsym in defined, fsym
        assert csym in undefined, csym
    assert "wd2_last_error" in undefined


@pytest.mark.skipif(FC is None, reason="no Fortran compiler (amdflang) in this image")
def test_shim_passes_the_callers_leading_dimension(tmp_path):
    """ld = mnx + 1 reaches the C ABI: with mnx = 1030 the constant 1031 is in
    the object code, with mnx = 302 the constant 303."""
    for mnx in (1030, 302):
        (tmp_path / "config.f").write_text(CONFIG.replace("1030", str(mnx)))
        asm = tmp_path / f"shim_{mnx}.s"
        r = subprocess.run([FC, "-ffixed-form", "-O2", "-S", SHIM, "-o", str(asm), f"-I{tmp_path}"],
                           cwd=tmp_path, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        text = asm.read_text()
        assert f"${mnx + 1}" in text, f"ld={mnx + 1} not passed"
