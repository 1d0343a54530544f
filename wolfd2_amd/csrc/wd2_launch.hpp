// wd2_launch.hpp -- host-side launchers of the wolfd2 HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "wd2_dev.hpp"
#include "wd2_dual.hpp"

namespace wd2 {

// Scratch of the parallel AltTridLU for chains of up to nmax rows (T: double
// for the reference's arithmetic, float for the mixed-precision momentum).
template <typename T>
struct TridWorkT {
    T *cp, *bp;              // forward-swept super-diagonal / right-hand side
    T *blk_c, *blk_b, *blk_x;  // per-workgroup boundary states
    T *blk_xe;                 // per-workgroup speculated x of the next block's first row
    int *fail, *nfallback;     // device flags
};
using TridWork = TridWorkT<double>;
int trid_blocks(long n);
// Row layout of a batch of independent chains: rows [0, n1) are chains of
// length L1, rows [n1, n) chains of length L2 (n1 a multiple of L1).
struct Chains {
    long n1, L1, L2;
};
// Chains of equal length L over rows [0, n).
// B2, Z (optional): also write B2[k] = Z[k] ? +0 : X[k], the right-hand side
// of a following system whose rows flagged in Z are identity rows.
template <typename T>
void launch_trid(const TridWorkT<T> &w, long n, long L, const T *A, const T *D, const T *C, const T *B, T *X,
                 hipStream_t s, T *B2 = nullptr, const unsigned char *Z = nullptr);
// Any layout (n rows in all).  stop (optional): every kernel returns at entry
// while *stop != 0.  rearm (optional): set to 1 once the solve has run.
template <typename T>
void launch_trid(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C, const T *B,
                 T *X, hipStream_t s, T *B2 = nullptr, const unsigned char *Z = nullptr, const int *stop = nullptr,
                 int *rearm = nullptr);
// Set the window solver's keep flag (the third int of a solver's flag block
// fail, nfallback, keep) to 1, its value between solves.
void launch_trid_keep_init(int *fail_block, hipStream_t s);

// The mixed path's FP32 chain solve (not bit-exact): overlapping windows
// solved by partitioned Thomas + PCR in LDS, the bit-exact solver as the
// fallback for a window that is not diagonally dominant enough; same
// arguments as launch_trid.
// wt: the window shape, the lanes of one window's workgroup (128, 256 or
// 512: windows of 1024, 2048 or 4096 rows).
template <typename T>
void launch_trid_window(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C,
                        const T *B, T *X, hipStream_t s, T *B2, const unsigned char *Z, const int *stop,
                        int wt = 256);
template <typename T>
void launch_trid_window_bench(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C,
                              const T *B, T *X, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int wt = 256);
int trid_window_shape_ok(int wt);
// reps back-to-back launches of the forward (which = 1) or backward (2)
// kernel of a solve of these chains between events e0 and e1.
template <typename T>
void launch_trid_kernel_bench(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C,
                              const T *B, T *X, int which, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1);

// Boundary fills (one workgroup).  err receives 1000+side for a bad flag.
template <typename T>
void launch_vel_bc(const Regions &R, int ld, T *u, T *v, bool outflow_only, int *err, hipStream_t s,
                   const int *stop = nullptr);   // stop: as for launch_trid
template <typename T>
void launch_pres_bc(const Regions &R, int ld, T *p, int *err, hipStream_t s);

// Device scalars shared by the reductions (one cache line each).
struct Scalars {
    double val[8];                 // reduction results
    unsigned int counter[8];       // last-workgroup tickets (always returned to 0)
    int sor_done, sor_nconv, sor_m, pad0;
    double sor_dif;
    int err;                       // bad-flag code from a kernel (0 = none)
    int redo_launch, redo_sweeps;  // temporally blocked SOR: launch to redo, sweeps to keep
    int ql_stop;                   // q-l iteration that met qtol (0 = none yet)
    int mg_done;                   // multigrid solve stopped
    int mg_nconv;                  // V-cycles applied when it converged, -1 not converged
    int mg_cycles;                 // V-cycles applied so far
    int pad1[9];
};

constexpr int RED_MAX_BLOCKS = 4096;   // partial-maximum slots per reduction

// utility.f:Project
void launch_project(const Regions &R, int ld, double dk, const Metrics &m, const double *p, double *u,
                    double *v, hipStream_t s);
// pressure.f:Divergence (nloc 1 or 2)
void launch_divergence(int nx, int ny, int ld, int nloc, const Metrics &m, const double *u, const double *v,
                       double *div, hipStream_t s);
// Ppe right-hand side: rhs = div/dk with blockage rows zeroed (prow != 0 marks them).
void launch_ppe_rhs(int nx, int ny, int ld, double dk, const double *div, const unsigned char *prow,
                    double *rhs, hipStream_t s);
// RhsPpe mixed-derivative terms for curvilinear grids: beff = rhs - cross(p).
void launch_rhs_cross(int nx, int ny, int ld, const Metrics &m, const double *rhs, const double *p,
                      double *beff, hipStream_t s);
// Five PPE coefficient fields with the blockage identity rows.
void launch_ppe_coef(int nx, int ny, int ld, const Metrics &m, const unsigned char *prow, double *a1,
                     double *a2, double *a3, double *a4, double *a5, hipStream_t s);
// utility.f:DMaxNorm / DiffMaxNorm into sc->val[slot] (b may be null for DMaxNorm).
void launch_max_norm(int nx, int ny, int ld, const double *a, const double *b, double *partials, Scalars *sc,
                     int slot, hipStream_t s);
// dst = src on i in [i0,i1], j in [j0,j1]
void launch_copy(int ld, int i0, int i1, int j0, int j1, const double *src, double *dst, hipStream_t s);
// dst[q] = src[q] over the same box for q < n (n <= 6), one launch
void launch_copy_set(int ld, int i0, int i1, int j0, int j1, const double *const *src, double *const *dst, int n,
                     hipStream_t s);
void launch_fill(int ld, int i0, int i1, int j0, int j1, double val, double *dst, hipStream_t s);
// p = 2 p - p1 (p2 null) or 3 (p - p1) + p2 on the pressure interior (the
// multigrid solve's extrapolated initial guess)
void launch_mg_extrap(int nx, int ny, int ld, double *p, const double *p1, const double *p2, hipStream_t s);

// SOR family.  coef: null -> coefficients from rau/rgv; else five fields.
struct PpeCoef {
    const double *a1, *a2, *a3, *a4, *a5;
};
// Sweep number `sweep` (1-based) of SorRB/SorRBP; partials holds two sweeps'
// worth of sor_rb_partials() maxima.  ev: null, or four events recorded around
// the black and the red kernel.  launch_sor_rb_check evaluates the stopping
// test of the last launched sweep.
int sor_rb_partials(int nx, int ny);
void launch_sor_rb_sweep(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *b,
                         double sorrel, double sortol, int sweep, double *p, double *partials, Scalars *sc,
                         hipStream_t s, hipEvent_t *ev);
void launch_sor_rb_check(int nx, int ny, int sweep, double sortol, int msorit, const double *partials,
                         Scalars *sc, hipStream_t s);
void launch_sor_reset(Scalars *sc, hipStream_t s);
// Temporally blocked SorRB/SorRBP (Cartesian, no blockage): nsweep <= sor_tiled_sweeps()
// sweeps of p_in into p_out; masks receives one stopping-test bit mask per
// tile.  The launch first applies the stopping test to the previous launch
// (prev_*; prev_nsweep = 0 for none).
int sor_tiled_tiles(int nx, int ny, int *ntx);
int sor_tiled_sweeps();
void launch_sor_tiled(int nx, int ny, int ld, const Metrics &m, const double *b, double sorrel, int nsweep,
                      const double *p_in, double *p_out, unsigned *masks, const unsigned *prev_masks,
                      int prev_nsweep, int prev_m0, int prev_launch, double sortol, int msorit, Scalars *sc,
                      int force, const double *uni, hipStream_t s);
// uni: null, or {rau, rgv} when every rau entry on i = 1..nx, j = 2..ny and
// every rgv entry on i = 2..nx, j = 1..ny has that one value (a uniform
// Cartesian grid): the kernels then use the values instead of loading them.
// Mixed precision (FP64 residual and accumulation, FP32 sweeps), same contract.
int sor_tiled_mx_tiles(int nx, int ny, int *ntx);
int sor_tiled_mx_sweeps();
void launch_sor_tiled_mx(int nx, int ny, int ld, const Metrics &m, const double *b, double sorrel, int nsweep,
                         const double *p_in, double *p_out, unsigned *masks, const unsigned *prev_masks,
                         int prev_nsweep, int prev_m0, int prev_launch, double sortol, int msorit, Scalars *sc,
                         int force, const double *uni, hipStream_t s);
// Stopping test of a launch over ntiles tile masks (the last launch of a batch).
void launch_sor_tiled_check(int ntiles, int nsweep, int m0, int launch, int msorit, const unsigned *masks,
                            Scalars *sc, hipStream_t s);
// Lexicographic SOR sweep (pressure.f:Sor), tiled wavefront.
void launch_sor_lex_sweep(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *b,
                          double sorrel, double sortol, int msorit, double *p, int *tile_flags, int sweep,
                          double *partials, Scalars *sc, hipStream_t s);
// Lexicographic Sor, whole solve in one launch (one workgroup; ny - 1 <= sor_lex_rows_max()).
int sor_lex_rows_max();
void launch_sor_lex_rows(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *rhs,
                         double *beff, int cart, double sorrel, double sortol, int msorit, double *p, Scalars *sc,
                         hipStream_t s);
// Line SOR: assemble the chains of one colour (first_line, step 2 lines, or one line), then relax.
void launch_slor_assemble(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *b,
                          const double *p, int first_line, int line_step, int nlines, double *ca, double *cd,
                          double *cc, double *cb, const Scalars *sc, hipStream_t s);
void launch_slor_relax(int nx, int ny, int ld, int first_line, int line_step, int nlines, const double *x,
                       double sorrel, int dif_mode, double *p, const double *pn, double *partials, Scalars *sc,
                       int slot, hipStream_t s);
void launch_slor_finish(int nparts, double sortol, int msorit, double *partials, Scalars *sc, hipStream_t s);
// SlorRBP: whole-grid relaxation of one colour's lines and |pn - p| max.
void launch_slor_rbp_relax(int nx, int ny, int ld, int first_line, double sorrel, const double *pn, double *p,
                           hipStream_t s);

// Momentum (momentum.f).  T: type of the temporaries, chains and arithmetic
// (double: the reference's; float: the mixed-precision path).
// cnvn / difn hold the x-momentum's time-level-n terms, cnvn2 / difn2 the
// y-momentum's, so both survive the q-l iterations of one nAuxMomentum call.
template <typename T>
struct MomTempsT {
    T *cj1, *cj2, *c1s, *c2s, *c1n, *c2n, *cnvs, *cnvn, *difs, *difn, *cps, *cpn, *cpj, *cnvn2, *difn2;
};
template <typename T>
struct ChainT {
    T *a, *d, *c, *b, *x;
};
using MomTemps = MomTempsT<double>;
using Chain = ChainT<double>;
// metric f of a Metrics m with its box
#define MARR(m, f) (MArr{(m).f, (m).box[MI_##f]})
template <typename T>
void launch_conv_coef(int nx, int ny, int ld, int ncomp, int njacob, const MArr &xzi, const MArr &xet,
                      const MArr &yzi, const MArr &yet, const double *u, const double *v, T *cc1, T *cc2,
                      hipStream_t s);
template <typename T>
void launch_porous_div(const Regions &R, int ld, int comp, const MomTempsT<T> &t, hipStream_t s);
template <typename T>
void launch_poros_coef(const Regions &R, int ld, int comp, const double *us, const double *vs, const double *un,
                       const double *vn, const MomTempsT<T> &t, hipStream_t s);
template <typename T>
void launch_conv_diff(int nx, int ny, int ld, int comp, const Metrics &m, const double *us, const double *un,
                      const MomTempsT<T> &t, hipStream_t s);
// Fused kernels for grids without porous regions: k_mom_n writes cnvn / difn
// of one component from un, vn; k_mom_asm_cb and k_mom_asm_gen assemble both
// splits of a component from us, vs (ch1: the first split's chain; ch2: the
// second split's coefficients; zm: its identity rows).
template <typename T>
void launch_mom_n(int nx, int ny, int ld, int comp, const Metrics &m, const double *un, const double *vn, T *cnvn,
                  T *difn, hipStream_t s);
template <typename T>
void launch_mom_asm(const Regions &R, int ld, int comp, double dk, double re, double fr, const Metrics &m,
                    const MomTempsT<T> &t, const double *us, const double *vs, const double *un, const double *vn,
                    const double *d, const double *dn, const ChainT<T> &ch1, const ChainT<T> &ch2, unsigned char *zm,
                    hipStream_t s, int *err, int live, const int *stop = nullptr);
// Both components at once (x-momentum: t1, ch1x, ch2x, zmx; y-momentum: t2,
// ch1y, ch2y, zmy), the same rows and bits as two launch_mom_asm calls.
template <typename T>
void launch_mom_asm_pair(const Regions &R, int ld, double dk, double re, double fr, const Metrics &m,
                         const MomTempsT<T> &t1, const MomTempsT<T> &t2, const double *us, const double *vs,
                         const double *un, const double *vn, const double *d, const double *dn,
                         const ChainT<T> &ch1x, const ChainT<T> &ch2x, unsigned char *zmx, const ChainT<T> &ch1y,
                         const ChainT<T> &ch2y, unsigned char *zmy, hipStream_t s, int *err, int live,
                         const int *stop = nullptr);
// live: bit 0, porous regions (load the Dupuit-Forchheimer coefficients);
// bit 1, a density field (load d, dn); otherwise those values are +0.
template <typename T>
void launch_assemble(const Regions &R, int ld, int comp, int split, double dk, double re, double fr,
                     const Metrics &m, const MomTempsT<T> &t, const double *us, const double *un, const double *d,
                     const double *dn, const T *xprev, const ChainT<T> &ch, hipStream_t s, int *err,
                     int live);
template <typename T>
void launch_scatter(int nx, int ny, int ld, int comp, const T *x, double *out, hipStream_t s);
void launch_update_norms(int nx, int ny, int ld, double *us, double *vs, const double *dus, const double *dvs,
                         double *partials, Scalars *sc, hipStream_t s);
// launch_scatter of both components (x1: the X chain solution, x2: the Y
// chain solution) followed by launch_update_norms, in one pass.  Both kernels
// return at entry while sc->ql_stop != 0; the fold then sets sc->ql_stop = m
// when max(val[0], val[1]) <= qtol (the q-l stopping test of iteration m).
template <typename T>
void launch_scatter_update(int nx, int ny, int ld, const T *x1, const T *x2, double *us, double *vs, double *dus,
                           double *dvs, double *partials, Scalars *sc, double qtol, int m, hipStream_t s);
// sc->ql_stop = 0 (a kernel, not a memset).
void launch_ql_reset(Scalars *sc, hipStream_t s);

// Thermal energy (k_thermal.hip).  err receives 2000+side for a bad
// temperature boundary flag.
struct ThermTemps {
    double *cu1, *cv1, *cun, *cvn;   // ConvCoef case 6 of (u, v), case 3 of (un, vn)
};
void launch_temp_bc(const Regions &R, const Thermal &TH, int ld, double *t, int *err, hipStream_t s);
void launch_therm_energy(const Regions &R, const Thermal &TH, int ld, double dk, double pe, const Metrics &m,
                         const double *un, const double *vn, const double *u, const double *v, const double *tn,
                         double *t, const ThermTemps &w, const ChainT<double> &ch, const TridWork &tw, int *err,
                         hipStream_t s);
void launch_eq_state(int nx, int ny, int ld, double uref, double densref, double tmax, double tref, double rconst,
                     const double *p, const double *t, double *den, hipStream_t s);
void launch_filter(const Regions &R, const Thermal &TH, int ld, int ncomp, double fp, double *qu, double *qh,
                   hipStream_t s);

// Multigrid Poisson solver (k_mg.hip).  A level of nx x ny cells: arrays of
// the arithmetic type (float or double) with row stride ld; on cells with a
// zero ring e (the correction after pre-smoothing), r (right-hand side) and
// c (the correction after the up pass, which the next finer level
// prolongates: a separate array, since an up launch reads e in its halos
// while it writes); ax / ay its x- and y-face coefficients.  Levels
// 0..ntiled-1 run as LDS tiles, the rest in the one-workgroup bottom kernel.
constexpr int MG_MAXLEV = 16;
struct MgLevel {
    int nx, ny, ld;
    void *e, *r, *ax, *ay, *c;
};
// Shapes an autotuner chooses among (wd2_ctx_mg_autotune): threads of a
// level-0 tile workgroup (256 or 512), whether the level-0 residual reads p
// from an LDS copy of the tile (1) or straight from memory (0), and the
// widest level the one-workgroup bottom kernel takes (32 or 64).
struct MgTune {
    int nt0, stage, bside;
};
struct MgHierarchy {
    int nlev, ntiled;
    MgLevel lev[MG_MAXLEV];
    MgTune tune;
};
int mg_tiles(int NX, int NY, int *ntx);
int mg_bottom_side();     // widest level the bottom kernel can take
int mg_bottom_cells();    // LDS cells per array of the bottom kernel
int mg_sweeps();          // smoothing sweeps before and after
template <typename T>
void launch_mg_coef(const MgHierarchy &H, int nx, int ny, int ld, const double *rau, const double *rgv,
                    const double *uni, hipStream_t s);
// One V-cycle of defect correction on b (1-based cycle count); part: one
// double per level-0 tile.  ev_down / ev_up (optional): an event pair around
// the level-0 down / up kernel.  minc: the tolerance stops the solve only
// once at least minc cycles are applied.
template <typename T>
void launch_mg_cycle(const MgHierarchy &H, int nx, int ny, int ld, const double *b, double *p, const double *rau,
                     const double *rgv, const double *uni, double dk, double tol, int cycle, int maxcyc,
                     double *part, Scalars *sc, hipStream_t s, hipEvent_t *ev_down, hipEvent_t *ev_up, int minc = 0);
void launch_mg_reset(Scalars *sc, hipStream_t s);
// reps back-to-back launches of the level-0 down (which = 2) or up (3) kernel
// between events e0 and e1 (clears the stop flag; the up kernel changes p).
template <typename T>
void launch_mg_kernel_bench(const MgHierarchy &H, int which, int reps, int nx, int ny, int ld, const double *b,
                            double *p, const double *rau, const double *rgv, const double *uni, double *part,
                            Scalars *sc, hipStream_t s, hipEvent_t e0, hipEvent_t e1);

// Adjoint kernels (k_adjoint.hip), FP64.  See that file for the maps each
// one transposes.  AdjCoefs: adjoints of one component's ConvCoef outputs
// (the upwind Jacobian's cj1, cj2 and the centred fluxes of us and un).
struct AdjCoefs {
    double *cj1, *cj2, *c1s, *c2s, *c1n, *c2n;
};
// The transposed system of a chain solve, in the form the chain solver
// (with its first-row quirk) solves exactly.
void launch_adj_tchain(long n, const Chains &cs, const double *A, const double *D, const double *C, double *At,
                       double *Dt, double *Ct, hipStream_t s);
// Coefficient adjoints of S x = b given mu = S^-T xbar (Z: rows without;
// XK: x read as +0 on the rows it does not flag).
void launch_adj_coef(long n, const Chains &cs, const double *A, const double *D, const double *C, const double *mu,
                     const double *x, const unsigned char *Z, const unsigned char *XK, double *Ab, double *Db,
                     double *Cb, hipStream_t s);
// Chain-ordered vector (X rows, then Y rows) from two fields, +0 on the rows
// flagged in Z (null: none); also clears red[slot..slot+1] (red null: none).
void launch_adj_gather(int nx, int ny, int ld, const double *ub, const double *vb, const unsigned char *Z,
                       double *lam, unsigned long long *red, int slot, hipStream_t s);
// nu += z, red[slot] = max |z|, red[slot + 1] = max |nu| (slots cleared before).
void launch_adj_update(long n, const double *z, double *nu, unsigned long long *red, int slot, hipStream_t s);
void launch_adj_zmask(long n, const unsigned char *Z, const double *in, double *out, hipStream_t s);
// One component's momentum assembly transposed (rows of that component's
// chain); with_n = 0 skips the time-level-n terms.
void launch_adj_asm(int comp, int nx, int ny, int ld, double dk, double re, const Metrics &m, const double *us,
                    const double *vs, const double *un, const double *vn, const double *a1b, const double *d1b,
                    const double *c1b, const double *bb, int with_n, double *fsb, double *fnb, const AdjCoefs &A,
                    hipStream_t s);
// (clears A as it reads it; with_n = 0: the time-level-n parts are zero)
void launch_adj_cc(int comp, int nx, int ny, int ld, const Metrics &m, const AdjCoefs &A, int with_n, double *usb,
                   double *vsb, double *unb, double *vnb, hipStream_t s);
// VelBoundCond / PresBoundCond transposed (one region, no-slip walls);
// bcb: 8 lines of max(nx,ny)+2 values, [side-1][var-1][k].
void launch_adj_vbc(const Regions &R, int ld, double *ub, double *vb, double *bcb, hipStream_t s);
void launch_adj_pbc(int nx, int ny, int ld, double *pb, hipStream_t s);
void launch_adj_project(int nx, int ny, int ld, double dk, const Metrics &m, const double *ub, const double *vb,
                        double *pb, hipStream_t s);
void launch_adj_div(int nx, int ny, int ld, double dk, const Metrics &m, const double *lam, double *ub, double *vb,
                    double *pb, hipStream_t s);
void launch_adj_axpy(long n, double alpha, const double *x, double *y, hipStream_t s);
void launch_adj_zero(long n, double *y, hipStream_t s);
void launch_adj_combine(int comp, int nx, int ny, int ld, const double *fsb, const double *fnb, double *out,
                        hipStream_t s);
void launch_adj_maxabs(long n, const double *x, unsigned long long *red, int slot, hipStream_t s);

// ------------------------------------------------------------ tangent-linear
// The momentum assembly on dual numbers (k_momentum.hip): values and
// directional derivatives of the time-level-n terms and of the chains.
void launch_mom_n_tl(int nx, int ny, int ld, int comp, const Metrics &m, const Dual *un, const Dual *vn, Dual *cnvn,
                     Dual *difn, hipStream_t s);
void launch_mom_asm_tl_pair(const Regions &R, int ld, double dk, double re, double fr, const Metrics &m,
                            const MomTempsT<Dual> &t1, const MomTempsT<Dual> &t2, const Dual *us, const Dual *vs,
                            const Dual *un, const Dual *vn, const ChainT<Dual> &ch1x, const ChainT<Dual> &ch2x,
                            unsigned char *zmx, const ChainT<Dual> &ch1y, const ChainT<Dual> &ch2y,
                            unsigned char *zmy, hipStream_t s, int *err);
// out = (v, d) element-wise
void launch_tl_pack(long n, const double *v, const double *d, Dual *out, hipStream_t s);
// r = db - dS1 x1 from the dual first-split chains (k_tangent.hip)
void launch_tl_rhs(long n, const Chains &cs, const ChainT<Dual> &ch, const double *x1, double *r, hipStream_t s);
// du, dv += the chain vector z on the chain rows' points
void launch_tl_scatter(int nx, int ny, int ld, const double *z, double *du, double *dv, hipStream_t s);

}  // namespace wd2
