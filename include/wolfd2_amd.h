/*
 * wolfd2_amd.h -- C ABI of the MI355X (gfx950) implementation of the wolfd2
 * per-timestep hot path.
 *
 * Two layers:
 *
 * 1. Drop-in entry points, one per reference subroutine on the hot path, with
 *    the reference's argument lists (all arrays are the caller's host arrays,
 *    dimensioned (0:mnx,0:mny) column-major) plus the caller's compile-time
 *    extents made explicit: ld = mnx+1 (leading dimension of every field) and
 *    mgri, mgrj (extents of the region tables).  Each call uploads its inputs,
 *    runs on the GPU and downloads its outputs before returning; the library
 *    keeps device buffers (sized by nx, ny on first use) but no host pointer.
 *    The Fortran binding in wolfd2_amd/fortran/wd2_shim.f exports these under
 *    the reference's own names (nauxmomentum_, ppe_, project_, ...).
 *
 * 2. A device-resident context (wd2_ctx_*) that keeps the fields in HBM
 *    across time steps and runs whole steps of main.f:690-972 on the GPU.
 *
 * Status codes: WD2_OK, or a negative WD2_ERR_*; wd2_last_error() describes
 * the last failure.  A bad boundary or solver flag, where the reference
 * prints a message and STOPs, returns WD2_ERR_FLAG.
 *
 * Calls are synchronous and not reentrant (the reference keeps static
 * scratch too).  FP64 results equal the reference's bit for bit.
 */
#ifndef WOLFD2_AMD_H
#define WOLFD2_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define WD2_OK 0
#define WD2_ERR_FLAG (-1)
#define WD2_ERR_ARG (-2)
#define WD2_ERR_HIP (-3)
#define WD2_ERR_UNSUPPORTED (-4)

const char *wd2_last_error(void);
/* Number of times the parallel AltTridLU fell back to its sequential path. */
int wd2_trid_fallbacks(void);
/* Red-black SOR on Cartesian grids without blockages runs several sweeps per
 * launch on LDS tiles (on by default); 0 selects one launch per colour and
 * sweep.  Both give the reference's iterates bit for bit.  Returns the
 * previous setting. */
int wd2_set_tiled_sor(int enable);
/* Lexicographic Sor (nPpeSolver 1) on grids of up to 512 pressure rows runs
 * as one single-workgroup launch per solve, a grid row per lane (on by
 * default); 0 selects the tiled wavefront, one launch per sweep.  Both give
 * the reference's iterates bit for bit.  Returns the previous setting. */
int wd2_set_lex_rows(int enable);
/* Mixed precision only: solve the FP32 momentum chains of q-l iterations 2..
 * with the windowed partitioned-Thomas/PCR solver (enable = 1, the default;
 * results agree with the bit-exact FP32 solver to FP32 rounding; the first
 * iteration of each step keeps the bit-exact solver) or every chain with the
 * bit-exact speculative solver (0).  Returns the previous setting.  No
 * counterpart in the reference. */
int wd2_set_window_chains(int enable);
/* Non-zero (default): on a uniform Cartesian grid, where rau and rgv each hold
 * one value over the entries the red-black SOR reads, the tiled SOR kernels
 * take the two values as arguments instead of loading the arrays (the same
 * values, so the same bits); 0 always loads them.  Returns the previous
 * setting. */
int wd2_set_uniform_coefficients(int enable);
/* Non-zero (default): print the reference's warnings on standard output
 * (Ppe non-convergence, bad solver flag).  Returns the previous setting. */
int wd2_set_verbose(int enable);
/* Red-black SOR sweeps per launch of its dominant kernel (1 when not tiled),
 * FP64 path and mixed-precision path. */
int wd2_sor_sweeps_per_launch(void);
int wd2_sor_mixed_sweeps_per_launch(void);

/* momentum.f:nAuxMomentum 33-48.  *nQLiter receives the function value: the
 * number of q-l iterations, or -1 when not converged in mqiter iterations. */
int wd2_naux_momentum(int nx, int ny, int ld, int mgri, int mgrj, int mqiter,
                      const int *nReg, const int *nRegBrd, const int *nRegType, const int *nMomBdTp,
                      double dk, double re, double fr, double qtol,
                      const double *dPRporos, const double *dPRporc1, const double *dPRporc2,
                      const double *dBCVal,
                      const double *ran, const double *rbn, const double *rgn,
                      const double *rac, const double *rbc, const double *rgc,
                      const double *dju, const double *djv,
                      const double *xec, const double *yec, const double *xzn, const double *yzn,
                      const double *xen, const double *yen, const double *xzc, const double *yzc,
                      const double *xeu, const double *yeu, const double *xzu, const double *yzu,
                      const double *xev, const double *yev, const double *xzv, const double *yzv,
                      const double *d, const double *dn,
                      const double *un, const double *vn, double *us, double *vs, int *nQLiter);

/* momentum.f:XMomentum 199-514 / YMomentum 520-838 (one ADI pass; the
 * increment is written on the reference's output range only). */
int wd2_x_momentum(int nx, int ny, int ld, int mgri, int mgrj,
                   const int *nReg, const int *nRegBrd, const int *nRegType, const int *nMomBdTp,
                   double dk, double re,
                   const double *dPRporos, const double *dPRporc1, const double *dPRporc2,
                   const double *rbn, const double *rgn, const double *rac, const double *rbc, const double *dju,
                   const double *xec, const double *yec, const double *xzn, const double *yzn,
                   const double *xeu, const double *yeu, const double *xzu, const double *yzu,
                   const double *us, const double *vs, const double *un, const double *vn, double *dus);
int wd2_y_momentum(int nx, int ny, int ld, int mgri, int mgrj,
                   const int *nReg, const int *nRegBrd, const int *nRegType, const int *nMomBdTp,
                   double dk, double re, double fr,
                   const double *dPRporos, const double *dPRporc1, const double *dPRporc2,
                   const double *ran, const double *rbn, const double *rbc, const double *rgc, const double *djv,
                   const double *xen, const double *yen, const double *xzc, const double *yzc,
                   const double *xev, const double *yev, const double *xzv, const double *yzv,
                   const double *d, const double *dn,
                   const double *us, const double *vs, const double *un, const double *vn, double *dvs);

/* momentum.f:AltTridLU 1307-1339: a is a(3,n) (sub, diagonal, super per row),
 * b the right-hand side, overwritten by the solution.  a is not modified. */
int wd2_alt_trid_lu(int n, const double *a, double *b);
/* The mixed path's FP32 chain solve on host arrays (no counterpart in the
 * reference; for measurement and tests): a batch of chains, rows [0, n1) in
 * chains of L1 rows and rows [n1, n) in chains of L2 rows, each solved as
 * AltTridLU solves it (first-row quirk included), with the window solver and
 * its bit-exact fallback (window = 1: the default shape; 128, 256 or 512: that
 * shape, see wd2_ctx_trid_set_window) or the bit-exact solver (0).  a, d, c,
 * b: sub-diagonal, diagonal, super-diagonal, right-hand side; x: solution. */
int wd2_trid_f32(long n, long n1, long L1, long L2, const float *a, const float *d, const float *c, const float *b,
                 float *x, int window);

/* No counterpart in the reference: AltTridLU (momentum.f:1307-1339, the same
 * a(3,n) and b(n) arguments and the same bits as wd2_alt_trid_lu) for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  a and b are DEVICE
 * pointers; b is overwritten by the solution.  The work is queued on `stream`
 * (a hipStream_t; NULL = the default stream) and the call returns without
 * synchronising.  It calls no allocating routine: the solver's scratch is a
 * static device array of each GPU (held until the process ends), so
 * n <= WD2_TRID_NOALLOC_MAX_ROWS (larger n returns WD2_ERR_ARG), and calls
 * that overlap in time on one GPU race on that array: serialise them. */
#define WD2_TRID_NOALLOC_MAX_ROWS 266240
int wd2_alt_trid_lu_noalloc(int n, const double *a, double *b, void *stream);

/* pressure.f:Ppe 30-38.  lCartesGrid is a Fortran logical (non-zero = true).
 * nPpeSolver 1 sor, 2 lsor, 3 rb_lsor, 4 par_rb_lsor, 5 rb_sor, 6 par_rb_sor
 * (the reference's solvers, bit for bit), and 7 mg, which has no counterpart
 * in the reference: multigrid V-cycles (red-black Gauss-Seidel smoothing on
 * LDS tiles, cell-centred coarsening, FP64 or, with WD2_PREC_MIXED, FP32)
 * inside FP64 defect correction on the same frozen-ghost system
 * (pressure.f:97-106 with the cross terms of RhsPpe on non-Cartesian grids),
 * until the divergence the projection leaves, max |dk (b - A p)| over the
 * interior, is below sortol; msorit caps the V-cycles and sorrel is unused.
 * Grids with blockages return WD2_ERR_UNSUPPORTED for solver 7.
 * *nSorConv receives the iteration count (sweeps, or V-cycles for solver 7;
 * msorit when not converged, after the reference's warning). */
int wd2_ppe(int nx, int ny, int ld, int mgri, int mgrj,
            const int *nReg, const int *nRegBrd, const int *nRegType,
            int lCartesGrid, int nPpeSolver, int msorit, int *nSorConv,
            double dk, double sortol, double sorrel,
            const double *rau, const double *rbu, const double *rbv, const double *rgv,
            const double *xeu, const double *yeu, const double *xzv, const double *yzv,
            const double *u, const double *v, double *p);

/* pressure.f:Divergence 265-323 (nloc 1: pressure-cell centres, 2: nodes). */
int wd2_divergence(int nx, int ny, int ld, int nloc,
                   const double *xet, const double *yet, const double *xzi, const double *yzi,
                   const double *u, const double *v, double *div);

/* utility.f:Project 253-259 */
int wd2_project(int nx, int ny, int ld, int mgri, int mgrj,
                const int *nReg, const int *nRegBrd, const int *nRegType, const int *nMomBdTp,
                double dk, const double *dju, const double *djv,
                const double *yeu, const double *xzv, const double *yzu, const double *xev,
                const double *p, double *u, double *v);

/* bound_cond.f:VelBoundCond 511-513, PresBoundCond 853-856, VelOutflowBCs 1656-1659 */
int wd2_vel_bound_cond(int nx, int ny, int ld, int mgri, int mgrj,
                       const int *nReg, const int *nRegBrd, const int *nMomBdTp, const double *dBCVal,
                       double *u, double *v);
int wd2_pres_bound_cond(int nx, int ny, int ld, int mgri, int mgrj,
                        const int *nReg, const int *nRegBrd, const int *nRegType, const int *nMomBdTp,
                        const double *dBCVal, double *p);
int wd2_vel_outflow_bcs(int nx, int ny, int ld, int mgri, int mgrj,
                        const int *nReg, const int *nRegBrd, const int *nMomBdTp, const double *dBCVal,
                        double *u, double *v);

/* utility.f:DMaxNorm 479-507, DiffMaxNorm 446-473 (result in *out). */
int wd2_dmax_norm(int nx, int ny, int ld, const double *u, double *out);
int wd2_diff_max_norm(int nx, int ny, int ld, const double *un, const double *u, double *out);

/* Thermal energy.  Thermal tables: nTRgType(mgri,mgrj) region types
 * (0 no source, 1 heat generation, 2 fixed temperature), nTemBdTp(mgri,mgrj,4)
 * boundary types (0 internal, 1 fixed temperature, 2 heat flux),
 * dTRgVal(mgri,mgrj) region temperatures, dHGSTval(mgri,mgrj) heat sources;
 * boundary temperatures are dBCVal(.,.,side,4).  A bad nTemBdTp flag returns
 * WD2_ERR_FLAG (the reference STOPs). */
/* bound_cond.f:TempBoundCond 1030-1033 */
int wd2_temp_bound_cond(int nx, int ny, int ld, int mgri, int mgrj,
                        const int *nReg, const int *nRegBrd, const int *nTRgType, const int *nTemBdTp,
                        const double *dTRgVal, const double *dBCVal, double *t);
/* thermal.f:ThermEnergy 24-32: t is advanced in place (it also receives
 * TempBoundCond's fills); un, vn, tn are time level n, u, v the current
 * velocities. */
int wd2_therm_energy(int nx, int ny, int ld, int mgri, int mgrj,
                     const int *nReg, const int *nRegBrd, const int *nTRgType, const int *nTemBdTp,
                     double dk, double pe, const double *dTRgVal, const double *dHGSTval,
                     const double *dBCVal,
                     const double *rau, const double *rbu, const double *rbv, const double *rgv,
                     const double *djc, const double *xeu, const double *yeu, const double *xzv,
                     const double *yzv, const double *xec, const double *yec, const double *xzc,
                     const double *yzc, const double *un, const double *vn, const double *u,
                     const double *v, const double *tn, double *t);
/* thermal.f:EqState 280-282: den on i = 2..nx, j = 2..ny (the rest of den is
 * left as passed). */
int wd2_eq_state(int nx, int ny, int ld, double uref, double densref, double tmax, double tref,
                 double rconst, const double *p, const double *t, double *den);
/* utility.f:Filter 31-34: ncomp 1 (u), 2 (v) or 4 (t); qu filtered in place.
 * Another ncomp returns WD2_ERR_FLAG (the reference STOPs). */
int wd2_filter(int nx, int ny, int ld, int mgri, int mgrj, int ncomp,
               const int *nReg, const int *nRegBrd, const int *nRegType, const int *nMomBdTp,
               const int *nTRgType, double fp, double *qu);

/* ----------------------------------------------------------------------
 * Device-resident context.
 */
typedef struct wd2_ctx wd2_ctx;

/* Run parameters (non-dimensional, as file_manip.f:ReadParam leaves them). */
typedef struct {
    int lCartesGrid;   /* non-zero: Cartesian grid (RhsPpe once per solve) */
    int mqiter;        /* max_ql_iter */
    int nPpeSolver;    /* 1..6 as for wd2_ppe */
    int msorit;        /* max_sor_iter */
    double qtol, sortol, sorrel;
    double dk, re, fr;
} wd2_params;

/* metrics: the 30 arrays in grid.f:Grid argument order
 * (rau, rbu, rbv, rgv, ran, rbn, rgn, rac, rbc, rgc, dju, djv, djc, djn,
 *  xen, yen, xzn, yzn, xec, yec, xzc, yzc, xeu, yeu, xzv, yzv, xzu, yzu, xev, yev),
 * each a host (0:mnx,0:mny) array with leading dimension ld. */
int wd2_ctx_create(wd2_ctx **ctx, int nx, int ny, int ld, int mgri, int mgrj,
                   const int *nReg, const int *nRegBrd, const int *nRegType, const int *nMomBdTp,
                   const double *dBCVal, const double *dPRporos, const double *dPRporc1,
                   const double *dPRporc2, const double *const *metrics);
int wd2_ctx_destroy(wd2_ctx *ctx);
/* Upload u, v, p (host, leading dimension ld) to the context. */
int wd2_ctx_set_fields(wd2_ctx *ctx, int ld, const double *u, const double *v, const double *p);
int wd2_ctx_get_fields(wd2_ctx *ctx, int ld, double *u, double *v, double *p);
/* main.f:606-641: VelBoundCond, Ppe, PresBoundCond, Project, VelBoundCond. */
int wd2_ctx_cold_start(wd2_ctx *ctx, const wd2_params *prm, int *nSorConv);
/* nsteps of main.f:690-972 with the context's step options (default: cold
 * flow, one momentum/energy iteration, no filters).  Per step k (if the
 * pointers are non-null): ql_iter[k], ppe_iter[k] of the step's last
 * momentum/energy iteration, and dif[3k..3k+2] = DiffMaxNorm of p, u, v
 * against the previous step. */
int wd2_ctx_step(wd2_ctx *ctx, const wd2_params *prm, int nsteps, int *ql_iter, int *ppe_iter, double *dif);
/* As wd2_ctx_step, with dif[4k..4k+3] = DiffMaxNorm of p, u, v, t (the four
 * columns of PrintDiff when the energy equation is on). */
int wd2_ctx_step4(wd2_ctx *ctx, const wd2_params *prm, int nsteps, int *ql_iter, int *ppe_iter, double *dif);

/* Step options: the rest of main.f:690-972 (input deck keywords of
 * parse.f:117-193 in brackets). */
typedef struct {
    int nthermen;      /* [thermal_energy] ThermEnergy in every iteration, TempBoundCond at step end */
    int neqstate;      /* [buoyancy] EqState after ThermEnergy: the density d of YMomentum */
    int nmeiter;       /* [max_me_iter] momentum/energy iterations (1 when nthermen = 0) */
    int nfiltu, nfiltv, nfiltt;   /* [filter_u/v/t] Shuman filters of us, vs, t */
    double dmeittol;   /* [me_tolerance] stop when l > 1 and max(dif) < dmeittol */
    double fpu, fpv, fpt;         /* filter parameters */
    double pe;         /* Peclet number (file_manip.f:158) */
    double uref, densref, tmax, tref, rconst;   /* EqState arguments */
} wd2_step_options;
int wd2_ctx_set_options(wd2_ctx *ctx, const wd2_step_options *opt);
/* Thermal region tables of the context (see wd2_therm_energy); the boundary
 * temperatures are the context's dBCVal(.,.,side,4). */
int wd2_ctx_set_thermal(wd2_ctx *ctx, const int *nTRgType, const int *nTemBdTp, const double *dTRgVal,
                        const double *dHGSTval);
/* Temperature t and density d of the context (host, leading dimension ld;
 * a null pointer is skipped). */
int wd2_ctx_set_thermal_fields(wd2_ctx *ctx, int ld, const double *t, const double *d);
int wd2_ctx_get_thermal_fields(wd2_ctx *ctx, int ld, double *t, double *d);
/* Device pointers of the context's u, v, p fields and their leading
 * dimension (the device layout pads rows to a multiple of 16 elements). */
int wd2_ctx_device_fields(wd2_ctx *ctx, double **u, double **v, double **p, int *ld);
/* The HIP stream the context runs on. */
void *wd2_ctx_stream(wd2_ctx *ctx);

/* Arithmetic of the context's time steps (no counterpart in the reference).
 * WD2_PREC_FP64 (default): the reference's FP64 arithmetic, bit for bit.
 * WD2_PREC_MIXED: the fields stay FP64; the momentum stencils and the ADI
 * chain solves run in FP32 and their increments are added in FP64; the
 * red-black SOR Poisson solve (nPpeSolver 5/6 on a Cartesian grid without
 * blockages) runs as defect correction: FP64 residual and pressure
 * accumulation, FP32 relaxation sweeps.  The iteration (q-l loop, sweep
 * order, stopping tests, caps) is the reference's; the results differ from
 * FP64 by FP32 rounding.  The other Poisson solvers keep FP64. */
#define WD2_PREC_FP64 0
#define WD2_PREC_MIXED 1
int wd2_ctx_set_precision(wd2_ctx *ctx, int precision);

/* Boundary profile of one region side (no counterpart in the reference, whose
 * dBCVal(mgri,mgrj,4,4) holds one constant per region side and variable, set
 * from the deck at parse.f:1524-1540 for walls and bound_cond.f:585-593 for
 * inlets).  var is 1 (u), 2 (v) or 3 (p); values[k] replaces
 * dBCVal(ireg,jreg,side,var) on grid line k: k = j = 0..ny+1 on the west and
 * east sides (n = ny+2), k = i = 0..nx+1 on the south and north sides
 * (n = nx+2).  Every boundary formula of VelBoundCond, PresBoundCond and
 * VelOutflowBCs is kept, with the constant replaced by the line's value, so
 * a profile of constant values gives the constant's results bit for bit.
 * values = NULL removes the profile.  Applies to this context's later steps. */
int wd2_ctx_set_bc_profile(wd2_ctx *ctx, int ireg, int jreg, int side, int var, const double *values, int n);
/* Timing of one Poisson kernel with HIP events on the context's stream.
 * enable=1 starts recording the red-black SOR kernels: one event pair around
 * each batch of back-to-back launches of the tiled multi-sweep kernel, or
 * around each colour kernel of the untiled path (launches after a solve has
 * converged exit early and are counted too).  enable=2 / 3 records the
 * multigrid solver's level-0 down / up kernel, one pair per launch, keeping
 * only the launches that ran in full.  enable=0 stops and returns the mean
 * launch duration (microseconds: the bracketed time over the launches it
 * holds) and the launch count. */
int wd2_ctx_kernel_timing(wd2_ctx *ctx, int enable, double *avg_us, long *launches);
/* The multigrid solver's (nPpeSolver 7) stopping measure: max |dk (b - A p)|
 * over the interior when the last solve stopped (*last) and the largest such
 * value since the last reset (*max_since); reset != 0 then clears the
 * maximum.  No counterpart in the reference. */
int wd2_ctx_ppe_divergence(wd2_ctx *ctx, double *last, double *max_since, int reset);
/* Mean duration (microseconds) of the multigrid solver's level-0 down
 * (which = 2) or up (which = 3) kernel, or of its one-workgroup bottom kernel
 * (4), over `reps` back-to-back launches on the context's last solve (event
 * pairs around single launches in the solve's own cycles include the gap
 * between launches); p is restored afterwards.  No counterpart in the
 * reference. */
int wd2_ctx_mg_kernel_time(wd2_ctx *ctx, int which, int reps, double *avg_us);
/* Mean time (microseconds) of the chain solver's forward (which = 1) or
 * backward (2) kernel, or of the mixed path's window kernel (3, mixed
 * precision only), over `reps` back-to-back launches on the context's
 * stream, on the first-split chains of the context's last q-l iteration (the
 * X and Y chains as one two-chain batch) in the context's precision.  The
 * chains are solved once more afterwards; fields are unchanged.  No
 * counterpart in the reference (measurement only). */
int wd2_ctx_trid_kernel_time(wd2_ctx *ctx, int which, int reps, double *avg_us);
/* The mixed path's window shape (no counterpart in the reference): the
 * lanes of one window's workgroup, 128, 256 (the default) or 512 -- windows
 * of 1024, 2048 or 4096 rows, at most 8, 9 or 10 parallel-cyclic-reduction
 * levels.  The shapes round differently (FP32 results differ in the last
 * bits).  wd2_ctx_trid_autotune times `reps` launches of each on the
 * chains of the context's last q-l iteration (mixed precision, after a
 * step), keeps the fastest and returns it and the three times (us[3], may
 * be NULL); wd2_ctx_trid_set_window sets one. */
int wd2_ctx_trid_set_window(wd2_ctx *ctx, int window_threads);
/* Multigrid Poisson solves (nPpeSolver 7) from an extrapolated initial guess
 * on the interior cells, from the pressures p^(n-1), p^(n-2) at the start of
 * the steps before: order 1, 2 p^n - p^(n-1) (from the second step after the
 * fields were set); order 2, 3 p^n - 3 p^(n-1) + p^(n-2) (from the third; the
 * second step is linear); order 0, the plain start (the default).  The
 * stopping test and the frozen ghosts are the same, so the solve converges
 * to the same solution in fewer V-cycles; the iterates, and so the last bits
 * of the fields, differ from the plain start.  A solve from the extrapolated
 * start applies at least two V-cycles (msorit permitting) before its
 * tolerance may stop it: with one, the error the divergence test barely
 * sees grows over thousands of steps.  The adjoint's runs always use
 * the plain start (and the step after them restarts the history).  Returns
 * the previous order; WD2_ERR_ARG for an order outside 0..2.  No
 * counterpart in the reference. */
int wd2_ctx_set_mg_extrapolation(wd2_ctx *ctx, int order);
int wd2_ctx_trid_autotune(wd2_ctx *ctx, int reps, int *window_threads, double *us);
/* Kernel shapes of the multigrid solver (no counterpart in the reference;
 * none changes the arithmetic, so results are the same bits): nt0 threads
 * per level-0 tile workgroup (256 or 512), stage 1 to form the level-0
 * residual from an LDS copy of p (0: from memory), bside the widest level
 * the one-workgroup bottom kernel takes (32 or 64). */
int wd2_ctx_mg_set_tune(wd2_ctx *ctx, int nt0, int stage, int bside);
int wd2_ctx_mg_get_tune(wd2_ctx *ctx, int *nt0, int *stage, int *bside);
/* Autotuner: times `cycles` full V-cycles of every shape above on the
 * context's grid, precision and current right-hand side (p is restored),
 * keeps the fastest and returns its time per V-cycle in *best_us. */
int wd2_ctx_mg_autotune(wd2_ctx *ctx, double dk, int cycles, double *best_us);
/* hipGraph steps (no counterpart in the reference).  enable != 0: each
 * time step of a context on the multigrid path (nPpeSolver 7, no porous
 * regions or blockages, no energy equation or equation of state) is one
 * replay of a captured graph: the q-l iterations and V-cycles are queued to
 * a budget (the last step's counts plus margin; work after a stopping test
 * has fired returns at entry, so the fields are those of the host-driven
 * step, bit for bit), and a step whose budget ran out is restored and redone
 * host-driven.  Graphs are cached per budget.  Other setups step host-driven.
 * wd2_ctx_graph_stats: steps replayed and steps redone so far. */
int wd2_ctx_set_graph(wd2_ctx *ctx, int enable);
int wd2_ctx_graph_stats(wd2_ctx *ctx, long *steps, long *redos);

/* An ensemble of independent contexts (any sizes) advanced together (no
 * counterpart in the reference, which runs one case per process,
 * main.f:690-1212): every member's step is queued on its own stream before
 * any is waited for, so the members run side by side on the GPU (graph mode
 * members as one replay each).  prm holds one wd2_params per member; per
 * step k and member i, ql_iter[k*n+i], ppe_iter[k*n+i] and dif[3(k*n+i)..]
 * as for wd2_ctx_step.  The ensemble does not own the contexts. */
typedef struct wd2_ens wd2_ens;
int wd2_ens_create(wd2_ens **ens, int n, wd2_ctx *const *ctxs);
int wd2_ens_destroy(wd2_ens *ens);
int wd2_ens_step(wd2_ens *ens, const wd2_params *prm, int nsteps, int *ql_iter, int *ppe_iter, double *dif);

/* Mean time of one full V-cycle in the current shape over `cycles` cycles on
 * the context's current right-hand side (p is restored). */
int wd2_ctx_mg_cycle_time(wd2_ctx *ctx, double dk, int cycles, double *us);

/* ---- Adjoint of a run of time steps (no counterpart in the reference).
 * Reverse-mode gradient, FP64, of nsteps steps of wd2_ctx_step (main.f:690-972)
 * with each q-l loop and Poisson solve taken at its converged solution, on a
 * cavity: one fluid region covering the grid, no-slip walls (type 1) on every
 * side with any boundary profiles (wd2_ctx_set_bc_profile), nPpeSolver 7,
 * lCartesGrid set, no energy equation, density or filters.  The forward run
 * should converge its q-l loops: the gradient is that of the converged map.
 * A negative qtol runs every loop for max_ql_iter iterations, which avoids
 * the reference's stopping norm (DiffMaxNorm over j = 2..ny-1) ending a loop
 * whose increments sit on the rows next to the north or south wall (the
 * first steps of a lid started from rest).
 *
 * wd2_adj_create binds an adjoint object to a context (checkpoint_every:
 * steps between the saved states, e.g. 20; memory: 3 fields per checkpoint,
 * 5 fields per step of one segment, 15 work fields).  wd2_adj_forward runs
 * nsteps steps of the context from its current fields (as wd2_ctx_step),
 * saving the checkpoints.  wd2_adj_backward takes the adjoint of the final
 * u, v, p (host arrays with leading dimension ld; p_bar may be NULL) and
 * returns the adjoint of the initial u, v, p (any output may be NULL) and of
 * the boundary values: bc_bar[((side-1)*2 + (var-1))*W + k] for side 1..4
 * (W, E, S, N), var 1..2 (u, v) and grid line k (j on W/E, i on S/N), W =
 * max(nx, ny) + 2 -- for a side with a constant value, the gradient of that
 * constant is the sum over its lines.  The backward pass re-runs the
 * segments from their checkpoints with the boundary values in force when it
 * is called: keep them as they were for wd2_adj_forward.  The context's
 * fields are those after the forward run when it returns. */
typedef struct wd2_adj wd2_adj;
typedef struct {
    int steps;               /* steps differentiated by the last backward pass */
    int adj_iters_max;       /* most fixed-point iterations of one adjoint q-l loop */
    double adj_iters_mean;
    int ppe_cycles_max;      /* most V-cycles of one adjoint Poisson solve */
    double ppe_cycles_mean;
    int fwd_unconverged;     /* forward steps whose q-l loop (qtol > 0) or Poisson solve hit its cap */
    int adj_unconverged;     /* adjoint q-l loops or Poisson solves that hit their caps */
    double ms_forward;       /* wall time of the last wd2_adj_forward */
    double ms_recompute;     /* backward pass: re-running the segments from their checkpoints */
    double ms_reverse;       /* backward pass: the adjoint steps */
} wd2_adj_stats;
int wd2_adj_create(wd2_adj **adj, wd2_ctx *ctx, int checkpoint_every);
int wd2_adj_destroy(wd2_adj *adj);
/* fixed_point_tol: the adjoint q-l loop stops when max|update| <= tol *
 * max|nu| (default 1e-12, at most max_iter iterations, default 200);
 * poisson_tol: the adjoint Poisson solve stops when max|b - A lam| <= tol *
 * max|b| (default 1e-13).  Non-positive values keep the current setting. */
int wd2_adj_set_tolerances(wd2_adj *adj, double fixed_point_tol, int max_iter, double poisson_tol);
int wd2_adj_forward(wd2_adj *adj, const wd2_params *prm, int nsteps, int *ql_iter, int *ppe_iter);
int wd2_adj_backward(wd2_adj *adj, int ld, const double *u_bar, const double *v_bar, const double *p_bar,
                     double *u0_bar, double *v0_bar, double *p0_bar, double *bc_bar, wd2_adj_stats *stats);
/* Tangent-linear (forward-mode) derivative of the same run: the forward run
 * of the last wd2_adj_forward again from its initial state, each step
 * followed by its tangent, along the direction (du0, dv0, dp0: host arrays
 * of the initial u, v, p with leading dimension ld, any may be NULL for
 * zero; dbc: boundary values laid out as bc_bar, NULL for zero).  Returns
 * the derivative of the final u, v, p (any output may be NULL).  The q-l
 * loops' tangents stop on the adjoint's fixed-point tolerance and the
 * Poisson solves on its Poisson tolerance (wd2_adj_set_tolerances); stats
 * report them in the adjoint's fields (adj_iters_*, ppe_cycles_*,
 * adj_unconverged) and the pass's wall time in ms_reverse.  No counterpart
 * in the reference: a check of the adjoint (<J d, w> = <d, J^T w>).  Like
 * the backward pass it re-runs the forward steps with the boundary values in
 * force when it is called: keep them as they were for wd2_adj_forward.  The
 * context's fields are those after the forward run when it returns. */
int wd2_adj_tangent(wd2_adj *adj, int ld, const double *du0, const double *dv0, const double *dp0,
                    const double *dbc, double *duN, double *dvN, double *dpN, wd2_adj_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
