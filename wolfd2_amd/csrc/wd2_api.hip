// wd2_api.hip -- host side of the wolfd2 HIP path: device context, the
// reference's call sequences (nAuxMomentum, Ppe, main.f time step) and the C
// ABI of include/wolfd2_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/wolfd2_amd.h"
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

using namespace wd2;

static thread_local std::string g_err;
static int g_fallbacks = 0;
static int g_tiled_sor = 1;   // temporally blocked SorRB when the grid allows it
static int g_verbose = 1;     // print the reference's warnings (pressure.f:242-246, 234)
static int g_lex_rows = 1;    // lexicographic Sor as one single-workgroup launch when the rows fit
static int g_uniform = 1;     // tiled SorRB takes rau/rgv as two values on uniform Cartesian grids
static int g_graph_epoch = 0; // bumped by the global switches a captured step bakes in
static int g_window_chains = 1;   // mixed path: FP32 chains by the windowed PCR solver

static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(WD2_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));   \
    } while (0)


struct wd2_ctx {
    int nx = 0, ny = 0, ld = 0;
    size_t nf = 0;                 // elements per field
    hipStream_t s = nullptr;
    // region tables
    int mgri = 0, mgrj = 0;
    std::vector<int> h_nreg, h_brd, h_rtype, h_btp;
    std::vector<double> h_bcv, h_por, h_pc1, h_pc2;
    int *d_int = nullptr;
    double *d_dbl = nullptr;
    Regions R{};
    // boundary profiles by their index in the (mgri,mgrj,4,3) offset table,
    // and the device copies of that table and of the values
    std::map<int, std::vector<double>> h_prof;
    // thermal region tables: host copies, device copies and their view
    std::vector<int> h_trt, h_tbt;
    std::vector<double> h_trv, h_hg;
    int *d_thi = nullptr;
    double *d_thd = nullptr;
    Thermal TH{};
    // step options (parse.f:117-193 defaults)
    wd2_step_options opt{0, 0, 1, 0, 0, 0, 1.0e-4, 5.0e2, 5.0e2, 5.0, 1.0, 1.0, 1.0, 1.0e3, 2.8357e2, 0.0};
    int *d_poff = nullptr;
    double *d_pval = nullptr;
    size_t pval_cap = 0;
    bool any_outlet = false, any_porous = false, any_blockage = false;
    unsigned char *prow = nullptr;
    // metrics
    double *metric_store = nullptr;
    double *met[M_COUNT] = {};
    Metrics M{};
    // fields
    double *store = nullptr;
    double *u, *v, *p, *un, *vn, *pn, *us, *vs, *dus, *dvs, *d, *dn, *div, *rhs, *beff, *pscr;
    double *t, *tn, *ts, *qh;      // temperature levels and the filter's copy
    // ThermEnergy's face coefficients: buffers of their own, since the mixed
    // path's FP32 momentum temporaries alias the FP64 ones and read the
    // entries they never write as zero
    ThermTemps tt{};
    double *pc[5];
    MomTemps T{};
    Chain ch{}, ch2{};   // ch2: the second split's chain of the fused momentum path
    unsigned char *zm = nullptr;   // its identity-row flags
    TridWork tw{};
    // FP32 views of the same buffers for the mixed-precision momentum
    MomTempsT<float> Tf{};
    ChainT<float> chf{}, chf2{};
    TridWorkT<float> twf{};
    double *partials = nullptr;
    Scalars *sc = nullptr;
    int *tile_flags = nullptr;
    int sweep_id = 0;
    bool coef_ready = false;
    // d (and dn) may be non-zero: set once EqState runs or a caller sends d;
    // until then both hold +0 and the momentum assembly does not load them
    bool d_live = false;
    bool sor_capped = false;   // the last tiled SOR solve ran to msorit
    int precision = WD2_PREC_FP64;
    int win_wt = 256;   // the mixed path's window shape (wd2_ctx_trid_autotune)
    // pinned copies of the stop flag and events of the SOR batches in flight
    // (h_flag[0..1]) and of the q-l iterations in flight (h_flag[2..3])
    int *h_flag = nullptr;
    hipEvent_t bev[2] = {nullptr, nullptr};
    hipEvent_t qev[2] = {nullptr, nullptr};
    // host transfers: two pinned staging buffers of one packed field each (the
    // caller's arrays are pageable and strided), and the last uploaded content
    // of every metric array of the drop-in calls (re-sent only when it changes)
    double *h_stage[2] = {nullptr, nullptr};
    hipEvent_t sev[2] = {nullptr, nullptr};
    int stage_i = 0;
    std::vector<std::vector<double>> met_shadow;
    int met_version = 0, uni_version = -1;
    // {rau, rgv} when each is one value over the entries the tiled red-black
    // SOR reads (a uniform Cartesian grid); uni_ok false otherwise
    bool uni_ok = false;
    double uni[2] = {0.0, 0.0};
    // multigrid Poisson solver (nPpeSolver 7): its levels and their storage,
    // the per-tile residual maxima, the precision and metric version its
    // coefficients were built for, the last solve's V-cycle count (the size
    // of the next solve's first batch), and max |dk (b - A p)| at the end of
    // the last solve and its maximum since the last query
    MgHierarchy mg{};
    double *mg_store = nullptr;
    size_t mg_bytes = 0;
    double *mg_part = nullptr;
    int mg_prec = -1, mg_ver = -1, mg_last = 4;
    double mg_div = 0.0, mg_div_max = 0.0;
    // hipGraph mode (wd2_ctx_set_graph): captured steps keyed by the q-l
    // iterations and V-cycles queued and the buffer u is in (u and us swap
    // every step), the pre-step u, v, p for a step that is redone eagerly,
    // the pinned copy of the scalars each graph ends with, the last step's
    // q-l count, and counts of replayed and redone steps
    // A graph bakes in the step's parameters, the precision, the multigrid
    // shapes, the region and profile tables and the metrics: it is cached
    // with the parameters and the global epoch, and every change of the rest
    // clears the cache (graph_clear).
    struct GraphEntry {
        int q, cyc;
        const double *u;
        wd2_params P;
        int epoch;
        int xp, xw;   // the extrapolated start's order (0: plain start) and history slot
        hipGraphExec_t exec;
    };
    bool graph_on = false;
    // Multigrid solves from an extrapolated initial guess
    // (wd2_ctx_set_mg_extrapolation): mg_extrap is the order asked for (0
    // off, 1 linear, 2 quadratic), extrap_off suspends it (the adjoint's
    // runs), p_hist counts the steps since the fields were set.  phist holds
    // two history slots: a step writes the pressure at the start of the step
    // before (pn) into slot xw, the other slot holding the one before that
    // from the step before, and xw flips when the step completes; x_run counts
    // the steps in a row that wrote a slot.
    int mg_extrap = 0;
    bool extrap_off = false;
    // V-cycles a multigrid solve applies before its tolerance may stop it: 2
    // for a solve from the extrapolated start (mg_min_cycles), else 0
    int mg_minc = 0;
    double *phist = nullptr;
    int p_hist = 0, xw = 0, x_run = 0, g_x = 0;
    std::vector<GraphEntry> graphs;
    double *gsave = nullptr;
    Scalars *h_sc = nullptr;
    int ql_last = 4, g_q = 0, g_c = 0;
    long graph_steps = 0, graph_redos = 0;
    // adjoint tape (wd2_adj): where a step copies its converged q-l
    // iterates us, vs (null: no copy)
    double *adj_us = nullptr, *adj_vs = nullptr;
    // optional HIP-event timing of one kernel (kt_on: 1 the red-black SOR
    // kernel, 2 / 3 the multigrid level-0 down / up kernel): event pairs,
    // each around kt_cnt[pair] back-to-back launches
    int kt_on = 0;
    std::vector<hipEvent_t> kt_ev;
    std::vector<int> kt_cnt;
    size_t kt_used = 0;
};

// npairs new event pairs for the timing of kernel `which` (nullptr when the
// timing is off or times another kernel)
static hipEvent_t *kt_events(wd2_ctx *c, int npairs, int which = 1) {
    if (c->kt_on != which) return nullptr;
    while (c->kt_ev.size() < c->kt_used + 2 * npairs) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->kt_ev.push_back(e);
    }
    hipEvent_t *ev = &c->kt_ev[c->kt_used];
    c->kt_used += 2 * npairs;
    for (int k = 0; k < npairs; k++) c->kt_cnt.push_back(1);
    return ev;
}

static inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Drop every cached step graph (after the stream has drained).
static void graph_clear(wd2_ctx *c) {
    if (c->graphs.empty()) return;
    (void)hipStreamSynchronize(c->s);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

static void ctx_free(wd2_ctx *c) {
    if (!c) return;
    for (hipEvent_t e : c->kt_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->bev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->qev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_flag) (void)hipHostFree(c->h_flag);
    for (int k = 0; k < 2; k++) {
        if (c->sev[k]) (void)hipEventDestroy(c->sev[k]);
        if (c->h_stage[k]) (void)hipHostFree(c->h_stage[k]);
    }
    (void)hipFree(c->d_int);
    (void)hipFree(c->d_dbl);
    (void)hipFree(c->d_poff);
    (void)hipFree(c->d_thi);
    (void)hipFree(c->d_thd);
    (void)hipFree(c->d_pval);
    (void)hipFree(c->prow);
    (void)hipFree(c->metric_store);
    (void)hipFree(c->store);
    (void)hipFree(c->partials);
    (void)hipFree(c->sc);
    (void)hipFree(c->tile_flags);
    (void)hipFree(c->mg_store);
    (void)hipFree(c->mg_part);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    (void)hipFree(c->gsave);
    (void)hipFree(c->phist);
    if (c->h_sc) (void)hipHostFree(c->h_sc);
    (void)hipFree(c->tw.cp);
    (void)hipFree(c->zm);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
}

// Rows of each chain buffer: every chain of the grid fits, and a multiple of
// 16 keeps each buffer 128-byte aligned (the chain solver reads 16-byte vectors).
// Chain capacity: the X- and Y-momentum chains of one q-l iteration side by
// side (nx(ny-1) + (nx-1)ny rows), padded.
static size_t chain_rows(int nx, int ny) { return (2 * (size_t)nx * ny + 64 + 15) / 16 * 16; }

static int ctx_alloc(wd2_ctx **out, int nx, int ny) {
    if (nx < 5 || ny < 5) return fail(WD2_ERR_ARG, "grid %dx%d too small (need nx, ny >= 5)", nx, ny);
    wd2_ctx *c = new wd2_ctx();
    c->nx = nx;
    c->ny = ny;
    c->ld = round_up(nx + 2, 16);
    c->nf = (size_t)c->ld * (ny + 2);
    HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc((void **)&c->h_flag, 4 * sizeof(int), hipHostMallocDefault));
    for (hipEvent_t &e : c->bev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (hipEvent_t &e : c->qev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (int k = 0; k < 2; k++) {
        HIPCHK(hipHostMalloc((void **)&c->h_stage[k], (size_t)(nx + 2) * (ny + 2) * sizeof(double),
                             hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&c->sev[k], hipEventDisableTiming));
    }
    const int NFIELDS = 16 + 4 + 4 + 5 + 15;
    const size_t nchain = chain_rows(nx, ny);
    const size_t total = NFIELDS * c->nf + 10 * nchain;
    HIPCHK(hipMalloc(&c->store, total * sizeof(double)));
    HIPCHK(hipMemsetAsync(c->store, 0, total * sizeof(double), c->s));
    double *q = c->store;
    auto take = [&](size_t n) { double *r = q; q += n; return r; };
    double **f[] = {&c->u, &c->v, &c->p, &c->un, &c->vn, &c->pn, &c->us, &c->vs, &c->dus, &c->dvs,
                    &c->d, &c->dn, &c->div, &c->rhs, &c->beff, &c->pscr, &c->t, &c->tn, &c->ts, &c->qh};
    for (double **pp : f) *pp = take(c->nf);
    c->tt = ThermTemps{take(c->nf), take(c->nf), take(c->nf), take(c->nf)};
    for (int k = 0; k < 5; k++) c->pc[k] = take(c->nf);
    double **t[] = {&c->T.cj1, &c->T.cj2, &c->T.c1s, &c->T.c2s, &c->T.c1n, &c->T.c2n, &c->T.cnvs,
                    &c->T.cnvn, &c->T.difs, &c->T.difn, &c->T.cps, &c->T.cpn, &c->T.cpj, &c->T.cnvn2,
                    &c->T.difn2};
    for (double **pp : t) *pp = take(c->nf);
    c->ch.a = take(nchain);
    c->ch.d = take(nchain);
    c->ch.c = take(nchain);
    c->ch.b = take(nchain);
    c->ch.x = take(nchain);
    c->ch2 = Chain{take(nchain), take(nchain), take(nchain), take(nchain), take(nchain)};
    HIPCHK(hipMalloc(&c->zm, nchain));
    const int nb = trid_blocks((long)nchain) + 2;
    HIPCHK(hipMalloc(&c->tw.cp, (2 * nchain + 4 * nb) * sizeof(double) + 4 * sizeof(int) + 64));
    c->tw.bp = c->tw.cp + nchain;
    c->tw.blk_c = c->tw.bp + nchain;
    c->tw.blk_b = c->tw.blk_c + nb;
    c->tw.blk_x = c->tw.blk_b + nb;
    c->tw.blk_xe = c->tw.blk_x + nb;
    c->tw.fail = (int *)(c->tw.blk_xe + nb);
    c->tw.nfallback = c->tw.fail + 1;
    auto F = [](double *q) { return reinterpret_cast<float *>(q); };
    c->Tf = MomTempsT<float>{F(c->T.cj1),  F(c->T.cj2),  F(c->T.c1s),  F(c->T.c2s), F(c->T.c1n),
                             F(c->T.c2n),  F(c->T.cnvs), F(c->T.cnvn), F(c->T.difs), F(c->T.difn),
                             F(c->T.cps),  F(c->T.cpn),  F(c->T.cpj),  F(c->T.cnvn2), F(c->T.difn2)};
    c->chf = ChainT<float>{F(c->ch.a), F(c->ch.d), F(c->ch.c), F(c->ch.b), F(c->ch.x)};
    c->chf2 = ChainT<float>{F(c->ch2.a), F(c->ch2.d), F(c->ch2.c), F(c->ch2.b), F(c->ch2.x)};
    c->twf = TridWorkT<float>{F(c->tw.cp),    F(c->tw.bp),     F(c->tw.blk_c), F(c->tw.blk_b),
                              F(c->tw.blk_x), F(c->tw.blk_xe), c->tw.fail,     c->tw.nfallback};
    HIPCHK(hipMemsetAsync(c->tw.fail, 0, 4 * sizeof(int), c->s));
    launch_trid_keep_init(c->tw.fail, c->s);
    const size_t npart = 4 * (size_t)RED_MAX_BLOCKS + 2 * (size_t)(c->nf / 256 + 64);
    HIPCHK(hipMalloc(&c->partials, npart * sizeof(double)));
    HIPCHK(hipMalloc(&c->sc, sizeof(Scalars)));
    HIPCHK(hipMemsetAsync(c->sc, 0, sizeof(Scalars), c->s));
    const int ntiles = ((nx + 62) / 64 + 1) * ((ny + 62) / 64 + 1);
    HIPCHK(hipMalloc(&c->tile_flags, ntiles * sizeof(int)));
    HIPCHK(hipMemsetAsync(c->tile_flags, 0, ntiles * sizeof(int), c->s));
    HIPCHK(hipMalloc(&c->metric_store, (size_t)M_COUNT * c->nf * sizeof(double)));
    HIPCHK(hipMemsetAsync(c->metric_store, 0, (size_t)M_COUNT * c->nf * sizeof(double), c->s));
    for (int k = 0; k < M_COUNT; k++) c->met[k] = c->metric_store + (size_t)k * c->nf;
    Metrics &M = c->M;
    M.ran = c->met[M_RAN]; M.rbn = c->met[M_RBN]; M.rgn = c->met[M_RGN];
    M.rac = c->met[M_RAC]; M.rbc = c->met[M_RBC]; M.rgc = c->met[M_RGC];
    M.rau = c->met[M_RAU]; M.rbu = c->met[M_RBU]; M.rbv = c->met[M_RBV]; M.rgv = c->met[M_RGV];
    M.dju = c->met[M_DJU]; M.djv = c->met[M_DJV]; M.djc = c->met[M_DJC];
    M.xen = c->met[M_XEN]; M.yen = c->met[M_YEN]; M.xzn = c->met[M_XZN]; M.yzn = c->met[M_YZN];
    M.xec = c->met[M_XEC]; M.yec = c->met[M_YEC]; M.xzc = c->met[M_XZC]; M.yzc = c->met[M_YZC];
    M.xeu = c->met[M_XEU]; M.yeu = c->met[M_YEU]; M.xzv = c->met[M_XZV]; M.yzv = c->met[M_YZV];
    M.xzu = c->met[M_XZU]; M.yzu = c->met[M_YZU]; M.xev = c->met[M_XEV]; M.yev = c->met[M_YEV];
    for (int k = 0; k < M_COUNT; k++) M.box[k] = MBox{0.0, 1, 0, 1, 0};
    HIPCHK(hipStreamSynchronize(c->s));
    *out = c;
    return WD2_OK;
}

// host (ld_h) -> device (c->ld) field copy of columns 0..ny+1, rows 0..nx+1
// The rows are packed into a pinned staging buffer on the host, then sent in
// one DMA copy; the two buffers alternate, so packing the next field overlaps
// the previous copy.
static int up(wd2_ctx *c, const double *h, int ld_h, double *dst) {
    if (!h) return WD2_OK;
    const size_t w = (size_t)c->nx + 2, rows = (size_t)c->ny + 2;
    const int k = c->stage_i;
    c->stage_i ^= 1;
    HIPCHK(hipEventSynchronize(c->sev[k]));   // the buffer's previous copy has finished
    double *st = c->h_stage[k];
    for (size_t j = 0; j < rows; j++) std::memcpy(st + j * w, h + j * (size_t)ld_h, w * sizeof(double));
    HIPCHK(hipMemcpy2DAsync(dst, (size_t)c->ld * 8, st, w * 8, w * 8, rows, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipEventRecord(c->sev[k], c->s));
    return WD2_OK;
}
// device (c->ld) -> host (ld_h); the host array is filled when this returns
static int down(wd2_ctx *c, const double *src, double *h, int ld_h) {
    const size_t w = (size_t)c->nx + 2, rows = (size_t)c->ny + 2;
    const int k = c->stage_i;
    c->stage_i ^= 1;
    HIPCHK(hipEventSynchronize(c->sev[k]));
    double *st = c->h_stage[k];
    HIPCHK(hipMemcpy2DAsync(st, w * 8, src, (size_t)c->ld * 8, w * 8, rows, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (size_t j = 0; j < rows; j++) std::memcpy(h + j * (size_t)ld_h, st + j * w, w * sizeof(double));
    return WD2_OK;
}

static bool valid_btype(int t) { return t >= BM_INTERN && t <= BM_OUTLT2; }

// Region tables (host copies + device copies + derived flags and the PPE
// identity-row mask of pressure.f:109-196).
static int set_regions(wd2_ctx *c, int mgri, int mgrj, const int *nReg, const int *nRegBrd, const int *nRegType,
                       const int *nMomBdTp, const double *dBCVal, const double *por, const double *pc1,
                       const double *pc2) {
    if (!nReg || !nRegBrd) return fail(WD2_ERR_ARG, "region tables missing");
    const int ni = nReg[0], nj = nReg[1];
    if (mgri < 1 || mgrj < 1 || ni < 1 || nj < 1 || ni > mgri || nj > mgrj)
        return fail(WD2_ERR_ARG, "bad region counts %d x %d (tables %d x %d)", ni, nj, mgri, mgrj);
    const size_t n2 = (size_t)mgri * mgrj;
    c->mgri = mgri;
    c->mgrj = mgrj;
    c->h_nreg.assign(nReg, nReg + 2);
    c->h_brd.assign(nRegBrd, nRegBrd + 4 * n2);
    c->h_rtype.assign(n2, RM_INTERN);
    if (nRegType) c->h_rtype.assign(nRegType, nRegType + n2);
    c->h_btp.assign(4 * n2, BM_INTERN);   // callers without momentum boundary types: all internal
    if (nMomBdTp) c->h_btp.assign(nMomBdTp, nMomBdTp + 4 * n2);
    c->h_bcv.assign(16 * n2, 0.0);
    if (dBCVal) c->h_bcv.assign(dBCVal, dBCVal + 16 * n2);
    c->h_por.assign(n2, 1.0);
    c->h_pc1.assign(n2, 0.0);
    c->h_pc2.assign(n2, 0.0);
    if (por) c->h_por.assign(por, por + n2);
    if (pc1) c->h_pc1.assign(pc1, pc1 + n2);
    if (pc2) c->h_pc2.assign(pc2, pc2 + n2);
    if (!c->d_int) {
        HIPCHK(hipMalloc(&c->d_int, (2 + 9 * n2) * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_dbl, 19 * n2 * sizeof(double)));
    }
    int *di = c->d_int;
    double *dd = c->d_dbl;
    HIPCHK(hipMemcpyAsync(di, c->h_brd.data(), 4 * n2 * sizeof(int), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(di + 4 * n2, c->h_rtype.data(), n2 * sizeof(int), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(di + 5 * n2, c->h_btp.data(), 4 * n2 * sizeof(int), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(dd, c->h_bcv.data(), 16 * n2 * sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(dd + 16 * n2, c->h_por.data(), n2 * sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(dd + 17 * n2, c->h_pc1.data(), n2 * sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(dd + 18 * n2, c->h_pc2.data(), n2 * sizeof(double), hipMemcpyHostToDevice, c->s));
    Regions &R = c->R;
    R.nx = c->nx; R.ny = c->ny; R.ni = ni; R.nj = nj; R.mgri = mgri; R.mgrj = mgrj;
    R.brd = di; R.rtype = di + 4 * n2; R.btp = di + 5 * n2;
    R.bcv = dd; R.poros = dd + 16 * n2; R.porc1 = dd + 17 * n2; R.porc2 = dd + 18 * n2;
    // host-side scans
    auto brd = [&](int ir, int jr, int s) { return c->h_brd[(ir - 1) + mgri * ((jr - 1) + mgrj * (s - 1))]; };
    auto btp = [&](int ir, int jr, int s) { return c->h_btp[(ir - 1) + mgri * ((jr - 1) + mgrj * (s - 1))]; };
    auto rty = [&](int ir, int jr) { return c->h_rtype[(ir - 1) + mgri * (jr - 1)]; };
    c->any_outlet = c->any_porous = c->any_blockage = false;
    for (int jr = 1; jr <= nj; jr++)
        for (int ir = 1; ir <= ni; ir++) {
            for (int s = 1; s <= 4; s++) {
                const int t = btp(ir, jr, s);
                if (!valid_btype(t)) return fail(WD2_ERR_FLAG, "Wrong nBdType flag %d in region %d,%d", t, ir, jr);
                if (t == BM_OUTLT1 || t == BM_OUTLT2) c->any_outlet = true;
                const int b = brd(ir, jr, s);
                if (b < 1 || b > ((s <= 2) ? c->nx : c->ny))
                    return fail(WD2_ERR_ARG, "region %d,%d border %d out of range", ir, jr, b);
            }
            if (rty(ir, jr) == RM_POROUS) c->any_porous = true;
            if (rty(ir, jr) == RM_BLOCKG) c->any_blockage = true;
        }
    if (c->any_blockage) {
        const int nx = c->nx, ny = c->ny, ld = c->ld;
        std::vector<unsigned char> m(c->nf, 0);
        auto mark = [&](int i, int j) {
            if (i >= 2 && i <= nx && j >= 2 && j <= ny) m[(size_t)i + (size_t)ld * j] = 1;
        };
        for (int jr = 1; jr <= nj; jr++)
            for (int ir = 1; ir <= ni; ir++) {
                if (rty(ir, jr) != RM_BLOCKG) continue;
                const int iW = brd(ir, jr, WEST), iE = brd(ir, jr, EAST), jS = brd(ir, jr, SOUTH),
                          jN = brd(ir, jr, NORTH);
                for (int j = jS + 2; j <= jN - 1; j++)
                    for (int i = iW + 2; i <= iE - 1; i++) mark(i, j);
                if (ir > 1 && rty(ir - 1, jr) == RM_BLOCKG)
                    for (int j = jS + 1; j <= jN; j++) mark(iW + 1, j);
                if (ir < ni && rty(ir + 1, jr) == RM_BLOCKG)
                    for (int j = jS + 1; j <= jN; j++) mark(iE, j);
                if (jr > 1 && rty(ir, jr - 1) == RM_BLOCKG)
                    for (int i = iW + 1; i <= iE; i++) mark(i, jS + 1);
                if (jr < nj && rty(ir, jr + 1) == RM_BLOCKG)
                    for (int i = iW + 1; i <= iE; i++) mark(i, jN);
            }
        if (!c->prow) HIPCHK(hipMalloc(&c->prow, c->nf));
        HIPCHK(hipMemcpyAsync(c->prow, m.data(), c->nf, hipMemcpyHostToDevice, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
    }
    c->coef_ready = false;
    return WD2_OK;
}

// Thermal region tables (any may be null: zeros).  Call after set_regions.
static int set_thermal(wd2_ctx *c, const int *nTRgType, const int *nTemBdTp, const double *dTRgVal,
                       const double *dHGSTval) {
    const size_t n2 = (size_t)c->mgri * c->mgrj;
    c->h_trt.assign(n2, RT_NOSRCE);
    c->h_tbt.assign(4 * n2, BT_INTERN);
    c->h_trv.assign(n2, 0.0);
    c->h_hg.assign(n2, 0.0);
    if (nTRgType) c->h_trt.assign(nTRgType, nTRgType + n2);
    if (nTemBdTp) c->h_tbt.assign(nTemBdTp, nTemBdTp + 4 * n2);
    if (dTRgVal) c->h_trv.assign(dTRgVal, dTRgVal + n2);
    if (dHGSTval) c->h_hg.assign(dHGSTval, dHGSTval + n2);
    if (!c->d_thi) {
        HIPCHK(hipMalloc(&c->d_thi, 5 * n2 * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_thd, 2 * n2 * sizeof(double)));
    }
    HIPCHK(hipMemcpyAsync(c->d_thi, c->h_trt.data(), n2 * sizeof(int), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->d_thi + n2, c->h_tbt.data(), 4 * n2 * sizeof(int), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->d_thd, c->h_trv.data(), n2 * sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->d_thd + n2, c->h_hg.data(), n2 * sizeof(double), hipMemcpyHostToDevice, c->s));
    Thermal &TH = c->TH;
    TH.mgri = c->mgri;
    TH.mgrj = c->mgrj;
    TH.trtype = c->d_thi;
    TH.tbtp = c->d_thi + n2;
    TH.trval = c->d_thd;
    TH.hgval = c->d_thd + n2;
    return WD2_OK;
}

// The constant box of a metric array (host, leading dimension ldh): one value
// over [2..nx-1] x [2..ny-1], grown a row or column at a time while the new
// line holds the same value (bit for bit); empty when the core is not one value.
static MBox scan_box(const double *h, int ldh, int nx, int ny) {
    const MBox empty{0.0, 1, 0, 1, 0};
    if (!h || nx < 4 || ny < 4) return empty;
    auto bits = [&](int i, int j) {
        uint64_t b;
        std::memcpy(&b, h + (size_t)i + (size_t)ldh * j, sizeof b);
        return b;
    };
    const uint64_t v = bits(2, 2);
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx - 1; i++)
            if (bits(i, j) != v) return empty;
    int i0 = 2, i1 = nx - 1, j0 = 2, j1 = ny - 1;
    auto col = [&](int i) {
        for (int j = j0; j <= j1; j++)
            if (bits(i, j) != v) return false;
        return true;
    };
    auto row = [&](int j) {
        for (int i = i0; i <= i1; i++)
            if (bits(i, j) != v) return false;
        return true;
    };
    for (bool grew = true; grew;) {
        grew = false;
        if (i0 > 0 && col(i0 - 1)) { i0--; grew = true; }
        if (i1 < nx + 1 && col(i1 + 1)) { i1++; grew = true; }
        if (j0 > 0 && row(j0 - 1)) { j0--; grew = true; }
        if (j1 < ny + 1 && row(j1 + 1)) { j1++; grew = true; }
    }
    double c;
    std::memcpy(&c, &v, sizeof c);
    return MBox{c, i0, i1, j0, j1};
}

// Uniform-coefficient scan: rau on i = 1..nx, j = 2..ny and rgv on
// i = 2..nx, j = 1..ny (host arrays, leading dimension ld_h).
static void scan_uniform(wd2_ctx *c, const double *rau, const double *rgv, int ld_h) {
    const int nx = c->nx, ny = c->ny;
    auto same = [&](const double *a, int i0, int i1, int j0, int j1) {
        const double v = a[(size_t)i0 + (size_t)ld_h * j0];
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++)
                if (std::memcmp(&a[(size_t)i + (size_t)ld_h * j], &v, sizeof v) != 0) return false;
        return true;
    };
    c->uni_ok = rau && rgv && same(rau, 1, nx, 2, ny) && same(rgv, 2, nx, 1, ny);
    if (c->uni_ok) {
        c->uni[0] = rau[1 + (size_t)ld_h * 2];
        c->uni[1] = rgv[2 + (size_t)ld_h * 1];
    }
}

static int check_err(wd2_ctx *c) {
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, &c->sc->err, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (e) {
        const int z = 0;
        HIPCHK(hipMemcpy(&c->sc->err, &z, sizeof(int), hipMemcpyHostToDevice));
        if (e >= 2000) return fail(WD2_ERR_FLAG, "Wrong nTemBdTp flag on side %d", e - 2000);
        return fail(WD2_ERR_FLAG, "bad boundary flag (code %d)", e);
    }
    return WD2_OK;
}

// -------------------------------------------------------------- momentum
// XMomentum (comp 1) / YMomentum (comp 2) on temporaries and chains of type T
// (double: the reference's arithmetic; float: the mixed-precision momentum).
// n_ready: this component's time-level-n terms (cnvn, difn) are already in
// place from an earlier call on the same un, vn and metrics.
template <typename T>
static int momentum(wd2_ctx *c, int comp, double dk, double re, double fr, const double *us, const double *vs,
                    const double *un, const double *vn, const MomTempsT<T> &t0, const ChainT<T> &ch,
                    const ChainT<T> &ch2, const TridWorkT<T> &tw, bool n_ready, double *out) {
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    const Metrics &M = c->M;
    hipStream_t s = c->s;
    MomTempsT<T> t = t0;
    if (comp == 2) {
        t.cnvn = t0.cnvn2;
        t.difn = t0.difn2;
    }
    const double *fs = comp == 1 ? us : vs, *fn = comp == 1 ? un : vn;
    const long n = comp == 1 ? (long)nx * (ny - 1) : (long)(nx - 1) * ny;
    const int live = (c->any_porous ? 1 : 0) | (c->d_live ? 2 : 0);
    if (!c->any_porous) {
        // fused: the convection coefficients are evaluated where they are
        // read; the first split's solve writes the second split's right-hand
        // side (identity rows zeroed)
        if (!n_ready) launch_mom_n<T>(nx, ny, ld, comp, M, un, vn, t.cnvn, t.difn, s);
        launch_mom_asm<T>(c->R, ld, comp, dk, re, fr, M, t, us, vs, un, vn, c->d, c->dn, ch, ch2, c->zm, s,
                          &c->sc->err, live);
        launch_trid<T>(tw, n, n, ch.a, ch.d, ch.c, ch.b, ch.x, s, ch2.b, c->zm);
        launch_trid<T>(tw, n, n, ch2.a, ch2.d, ch2.c, ch2.b, ch2.x, s);
        launch_scatter<T>(nx, ny, ld, comp, ch2.x, out, s);
        return WD2_OK;
    }
    if (comp == 1) {
        launch_conv_coef<T>(nx, ny, ld, 4, 1, MARR(M, xzu), MARR(M, xeu), MARR(M, yzu), MARR(M, yeu), us, vs, t.cj1, t.cj2, s);
        launch_conv_coef<T>(nx, ny, ld, 1, 0, MARR(M, xzn), MARR(M, xec), MARR(M, yzn), MARR(M, yec), us, vs, t.c1s, t.c2s, s);
        launch_conv_coef<T>(nx, ny, ld, 1, 0, MARR(M, xzn), MARR(M, xec), MARR(M, yzn), MARR(M, yec), un, vn, t.c1n, t.c2n, s);
    } else {
        launch_conv_coef<T>(nx, ny, ld, 5, 1, MARR(M, xzv), MARR(M, xev), MARR(M, yzv), MARR(M, yev), us, vs, t.cj1, t.cj2, s);
        launch_conv_coef<T>(nx, ny, ld, 2, 0, MARR(M, xzc), MARR(M, xen), MARR(M, yzc), MARR(M, yen), us, vs, t.c1s, t.c2s, s);
        launch_conv_coef<T>(nx, ny, ld, 2, 0, MARR(M, xzc), MARR(M, xen), MARR(M, yzc), MARR(M, yen), un, vn, t.c1n, t.c2n, s);
    }
    launch_porous_div<T>(c->R, ld, comp, t, s);
    launch_conv_diff<T>(nx, ny, ld, comp, M, fs, fn, t, s);
    launch_poros_coef<T>(c->R, ld, comp, us, vs, un, vn, t, s);
    launch_assemble<T>(c->R, ld, comp, 1, dk, re, fr, M, t, fs, fn, c->d, c->dn, nullptr, ch, s, &c->sc->err, live);
    launch_trid<T>(tw, n, n, ch.a, ch.d, ch.c, ch.b, ch.x, s);
    launch_assemble<T>(c->R, ld, comp, 2, dk, re, fr, M, t, fs, fn, c->d, c->dn, ch.x, ch, s, &c->sc->err, live);
    launch_trid<T>(tw, n, n, ch.a, ch.d, ch.c, ch.b, ch.x, s);
    launch_scatter<T>(nx, ny, ld, comp, ch.x, out, s);
    return WD2_OK;
}

// XMomentum and YMomentum of one q-l iteration without porous regions, then
// us += dus, vs += dvs and the increment norms (momentum.f:171-180).  The two
// components read the same us, vs, un, vn and write separate increments, so
// their chains are assembled side by side (X rows first) and each split is one
// solve of a batch of two chains, which gives each chain the bits it has alone.
// Iteration m's kernels return at entry once an earlier iteration has met
// qtol (sc->ql_stop != 0), and its fold sets sc->ql_stop = m when it does, so
// the host can queue iteration m before it reads iteration m-1's test.
template <typename T>
static void momentum_pair(wd2_ctx *c, double dk, double re, double fr, const MomTempsT<T> &t, const ChainT<T> &ch,
                          const ChainT<T> &ch2, const TridWorkT<T> &tw, int m, double qtol) {
    const bool n_ready = m > 1;
    const int *stop = &c->sc->ql_stop;
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    const Metrics &M = c->M;
    hipStream_t s = c->s;
    const long n1 = (long)nx * (ny - 1), n2 = (long)(nx - 1) * ny;
    const int live = c->d_live ? 2 : 0;
    MomTempsT<T> t2 = t;
    t2.cnvn = t.cnvn2;
    t2.difn = t.difn2;
    auto shift = [](const ChainT<T> &q, long o) { return ChainT<T>{q.a + o, q.d + o, q.c + o, q.b + o, q.x + o}; };
    if (!n_ready) {
        launch_mom_n<T>(nx, ny, ld, 1, M, c->un, c->vn, t.cnvn, t.difn, s);
        launch_mom_n<T>(nx, ny, ld, 2, M, c->un, c->vn, t2.cnvn, t2.difn, s);
    }
    launch_mom_asm_pair<T>(c->R, ld, dk, re, fr, M, t, t2, c->us, c->vs, c->un, c->vn, c->d, c->dn, ch, ch2, c->zm,
                           shift(ch, n1), shift(ch2, n1), c->zm + n1, s, &c->sc->err, live, stop);
    const Chains cs{n1, n1, n2};
    if constexpr (std::is_same<T, float>::value) {
        // The window solver serves iterations 2.. of the q-l loop.  The first
        // iteration's solve is the whole update from time level n and may be
        // the step's result (the test after one iteration can be met, as on
        // the first step from rest), so it takes the sequential solver.  Per
        // solve the two have errors of one size, but on a 1024^2 cavity from
        // rest the window solver in the first iteration leaves v 1.3e-6 from
        // the FP64 path after two steps (relative L2), against 2.8e-7 this way.
        if (g_window_chains && m > 1) {
            launch_trid_window<T>(tw, n1 + n2, cs, ch.a, ch.d, ch.c, ch.b, ch.x, s, ch2.b, c->zm, stop, c->win_wt);
            launch_trid_window<T>(tw, n1 + n2, cs, ch2.a, ch2.d, ch2.c, ch2.b, ch2.x, s, nullptr, nullptr, stop,
                                  c->win_wt);
        } else {
            launch_trid<T>(tw, n1 + n2, cs, ch.a, ch.d, ch.c, ch.b, ch.x, s, ch2.b, c->zm, stop);
            launch_trid<T>(tw, n1 + n2, cs, ch2.a, ch2.d, ch2.c, ch2.b, ch2.x, s, nullptr, nullptr, stop);
        }
    } else {
        launch_trid<T>(tw, n1 + n2, cs, ch.a, ch.d, ch.c, ch.b, ch.x, s, ch2.b, c->zm, stop);
        launch_trid<T>(tw, n1 + n2, cs, ch2.a, ch2.d, ch2.c, ch2.b, ch2.x, s, nullptr, nullptr, stop);
    }
    launch_scatter_update<T>(nx, ny, ld, ch2.x, ch2.x + n1, c->us, c->vs, c->dus, c->dvs, c->partials, c->sc, qtol,
                             m, s);
}

static int momentum(wd2_ctx *c, int comp, double dk, double re, double fr, const double *us, const double *vs,
                    const double *un, const double *vn, double *out, bool n_ready = false) {
    if (c->precision == WD2_PREC_MIXED)
        return momentum<float>(c, comp, dk, re, fr, us, vs, un, vn, c->Tf, c->chf, c->chf2, c->twf, n_ready, out);
    return momentum<double>(c, comp, dk, re, fr, us, vs, un, vn, c->T, c->ch, c->ch2, c->tw, n_ready, out);
}

// momentum.f:nAuxMomentum 33-193 on the context's un, vn -> us, vs.
static int naux_momentum(wd2_ctx *c, int mqiter, double dk, double re, double fr, double qtol, int *nql) {
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    {
        const double *src[2] = {c->un, c->vn};
        double *dst[2] = {c->us, c->vs};
        launch_copy_set(ld, 1, nx + 1, 1, ny + 1, src, dst, 2, c->s);
    }
    *nql = -1;
    if (!c->any_porous) {
        if (mqiter < 1) return check_err(c);
        // Pipelined by one iteration: iteration m is queued before the host
        // reads iteration m-1's stopping test.  Once a test is met, every
        // later iteration's kernels return at entry, so the fields are those
        // of the sequential loop.  un, vn and the metrics are fixed during
        // the iterations: the time-level-n terms of the first serve them all.
        hipStream_t s = c->s;
        const int *stop = &c->sc->ql_stop;
        launch_ql_reset(c->sc, s);
        int m = 1, last = 0;   // last: the iteration whose test was read last
        for (; m <= mqiter; m++) {
            if (c->any_outlet) launch_vel_bc<double>(c->R, ld, c->us, c->vs, true, &c->sc->err, s, stop);
            if (c->precision == WD2_PREC_MIXED)
                momentum_pair<float>(c, dk, re, fr, c->Tf, c->chf, c->chf2, c->twf, m, qtol);
            else
                momentum_pair<double>(c, dk, re, fr, c->T, c->ch, c->ch2, c->tw, m, qtol);
            HIPCHK(hipMemcpyAsync(&c->h_flag[2 + m % 2], &c->sc->ql_stop, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(hipEventRecord(c->qev[m % 2], s));
            if (m >= 2) {
                HIPCHK(hipEventSynchronize(c->qev[(m - 1) % 2]));
                last = m - 1;
                if (c->h_flag[2 + last % 2]) break;
            }
        }
        if (m > mqiter) {
            HIPCHK(hipEventSynchronize(c->qev[mqiter % 2]));
            last = mqiter;
        }
        if (c->h_flag[2 + last % 2]) *nql = c->h_flag[2 + last % 2];
        return check_err(c);
    }
    for (int m = 1; m <= mqiter; m++) {
        if (c->any_outlet) launch_vel_bc<double>(c->R, ld, c->us, c->vs, true, &c->sc->err, c->s);
        int st = momentum(c, 1, dk, re, fr, c->us, c->vs, c->un, c->vn, c->dus, m > 1);
        if (st) return st;
        st = momentum(c, 2, dk, re, fr, c->us, c->vs, c->un, c->vn, c->dvs, m > 1);
        if (st) return st;
        launch_update_norms(nx, ny, ld, c->us, c->vs, c->dus, c->dvs, c->partials, c->sc, c->s);
        double nv[2];
        HIPCHK(hipMemcpyAsync(nv, c->sc->val, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        const double difmax = nv[0] > nv[1] ? nv[0] : nv[1];
        if (difmax <= qtol) {
            *nql = m;
            break;
        }
    }
    return check_err(c);
}

// ------------------------------------------------------------------- Ppe
// The levels at most tune.bside cells wide (and at most mg_bottom_side())
// go to the bottom kernel; level 0 is always tiled.
static int mg_retile(MgHierarchy &H) {
    const int side = std::min(H.tune.bside, mg_bottom_side());
    H.ntiled = 1;
    while (H.ntiled < H.nlev && std::max(H.lev[H.ntiled].nx, H.lev[H.ntiled].ny) > side) H.ntiled++;
    size_t cells = 0;
    for (int l = H.ntiled; l < H.nlev; l++) cells += (size_t)(H.lev[l].nx + 2) * (H.lev[l].ny + 2);
    if (H.ntiled >= H.nlev || cells > (size_t)mg_bottom_cells())
        return fail(WD2_ERR_ARG, "multigrid: bottom levels do not fit (%zu cells)", cells);
    return WD2_OK;
}

// Levels of the multigrid solver: level 0 is the (nx-1) x (ny-1) pressure
// grid, each next level ceil-halves both sides down to at most 2 x 2 cells.
static int mg_setup(wd2_ctx *c) {
    if (c->mg_store) return WD2_OK;
    MgHierarchy &H = c->mg;
    int NX = c->nx - 1, NY = c->ny - 1;
    size_t off[MG_MAXLEV], total = 0;
    H.nlev = 0;
    for (;;) {
        if (H.nlev >= MG_MAXLEV) return fail(WD2_ERR_ARG, "multigrid: too many levels for %dx%d", c->nx, c->ny);
        MgLevel &L = H.lev[H.nlev];
        L.nx = NX;
        L.ny = NY;
        L.ld = round_up(NX + 2, 16);
        off[H.nlev] = total;
        total += 5 * (size_t)L.ld * (NY + 2);
        H.nlev++;
        if (std::max(NX, NY) <= 2) break;
        NX = (NX + 1) / 2;
        NY = (NY + 1) / 2;
    }
    H.tune = MgTune{256, 0, 64};
    int st = mg_retile(H);
    if (st) return st;
    HIPCHK(hipMalloc(&c->mg_store, total * sizeof(double)));
    c->mg_bytes = total * sizeof(double);
    HIPCHK(hipMemsetAsync(c->mg_store, 0, c->mg_bytes, c->s));
    for (int l = 0; l < H.nlev; l++) {
        MgLevel &L = H.lev[l];
        const size_t n = (size_t)L.ld * (L.ny + 2);
        double *b = c->mg_store + off[l];
        L.e = b;
        L.r = b + n;
        L.ax = b + 2 * n;
        L.ay = b + 3 * n;
        L.c = b + 4 * n;
    }
    const int nt = mg_tiles(H.lev[0].nx, H.lev[0].ny, nullptr);
    HIPCHK(hipMalloc(&c->mg_part, nt * sizeof(double)));
    return WD2_OK;
}

// Levels allocated and the face coefficients built for the context's
// precision and metrics.
static int mg_prepare(wd2_ctx *c) {
    int st = mg_setup(c);
    if (st) return st;
    if (c->mg_prec != c->precision || c->mg_ver != c->met_version) {
        const double *uni = (g_uniform && c->uni_ok) ? c->uni : nullptr;
        // float and double views share the storage: clear it (zero rings)
        if (c->mg_prec != -1 && c->mg_prec != c->precision)
            HIPCHK(hipMemsetAsync(c->mg_store, 0, c->mg_bytes, c->s));
        if (c->precision == WD2_PREC_MIXED)
            launch_mg_coef<float>(c->mg, c->nx, c->ny, c->ld, c->M.rau, c->M.rgv, uni, c->s);
        else
            launch_mg_coef<double>(c->mg, c->nx, c->ny, c->ld, c->M.rau, c->M.rgv, uni, c->s);
        c->mg_prec = c->precision;
        c->mg_ver = c->met_version;
    }
    return WD2_OK;
}

// Mean time of a full V-cycle (microseconds) over `cycles` cycles on the
// context's current right-hand side and p (tolerance 0, so none stops
// early); p is restored afterwards.
static int mg_cycle_time(wd2_ctx *c, int cycles, double dk, double *us) {
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    hipEvent_t e[2];
    HIPCHK(hipEventCreate(&e[0]));
    HIPCHK(hipEventCreate(&e[1]));
    const double *uni = (g_uniform && c->uni_ok) ? c->uni : nullptr;
    launch_copy(ld, 0, nx + 1, 0, ny + 1, c->p, c->pscr, s);
    launch_mg_reset(c->sc, s);
    HIPCHK(hipEventRecord(e[0], s));
    for (int k = 1; k <= cycles; k++) {
        if (c->precision == WD2_PREC_MIXED)
            launch_mg_cycle<float>(c->mg, nx, ny, ld, c->rhs, c->p, c->M.rau, c->M.rgv, uni, dk, 0.0, k, cycles + 1,
                                   c->mg_part, c->sc, s, nullptr, nullptr);
        else
            launch_mg_cycle<double>(c->mg, nx, ny, ld, c->rhs, c->p, c->M.rau, c->M.rgv, uni, dk, 0.0, k, cycles + 1,
                                    c->mg_part, c->sc, s, nullptr, nullptr);
    }
    HIPCHK(hipEventRecord(e[1], s));
    launch_copy(ld, 0, nx + 1, 0, ny + 1, c->pscr, c->p, s);
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e[0], e[1]));
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    *us = 1e3 * ms / cycles;
    return WD2_OK;
}

// nPpeSolver 7: multigrid V-cycles (FP64 or FP32 by the context's precision)
// in FP64 defect correction, until max |dk (b - A p)| < tol (the divergence
// Project leaves) or maxcyc cycles; *nconv_out = the V-cycles applied.
// bsrc: the right-hand side (null: the Ppe right-hand side c->rhs / c->beff).
static int ppe_mg(wd2_ctx *c, int cart, int maxcyc, double dk, double tol, double *p, int *nconv_out,
                  const double *bsrc = nullptr) {
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    if (c->any_blockage)
        return fail(WD2_ERR_UNSUPPORTED, "the multigrid Poisson solver (nPpeSolver 7) does not take blockages");
    int st = mg_prepare(c);
    if (st) return st;
    const bool mx = c->precision == WD2_PREC_MIXED;
    const double *uni = (g_uniform && c->uni_ok) ? c->uni : nullptr;
    launch_mg_reset(c->sc, s);
    const double *b = bsrc ? bsrc : cart ? c->rhs : c->beff;
    const size_t kt0 = c->kt_used;
    // Batches of V-cycles with a stopping check at the end of each, two in
    // flight (as the SOR solves): the first batch is the last solve's cycle
    // count plus the final check, so a steady run rarely overruns; the next
    // batches take 1, 2, 4, 8, then 16 cycles.  The one-cycle second batch is
    // queued before the first batch's test is read, so the GPU has it to run
    // (returning at entry once the solve has stopped) while the host reads
    // the test and queues what follows the solve.
    int cyc = 0, issued = 0, checked = 0, stop = 0;
    const int total = maxcyc + 1;   // the last cycle only tests
    while (!stop) {
        if (cyc < total && issued - checked < 2) {
            const int Q = issued == 0 ? std::max(1, c->mg_last + 1) : 1 << std::min(issued - 1, 4);
            for (int q = 0; q < Q && cyc < total; q++) {
                cyc++;
                if (!cart) launch_rhs_cross(nx, ny, ld, c->M, c->rhs, p, c->beff, s);
                hipEvent_t *ed = kt_events(c, 1, 2), *eu = kt_events(c, 1, 3);
                if (mx)
                    launch_mg_cycle<float>(c->mg, nx, ny, ld, b, p, c->M.rau, c->M.rgv, uni, dk, tol, cyc, maxcyc,
                                           c->mg_part, c->sc, s, ed, eu, c->mg_minc);
                else
                    launch_mg_cycle<double>(c->mg, nx, ny, ld, b, p, c->M.rau, c->M.rgv, uni, dk, tol, cyc, maxcyc,
                                            c->mg_part, c->sc, s, ed, eu, c->mg_minc);
            }
            HIPCHK(hipMemcpyAsync(&c->h_flag[issued % 2], &c->sc->mg_done, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(hipEventRecord(c->bev[issued % 2], s));
            issued++;
            continue;
        }
        if (checked == issued) break;
        HIPCHK(hipEventSynchronize(c->bev[checked % 2]));
        stop = c->h_flag[checked % 2];
        checked++;
    }
    Scalars h;
    HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    int nconv = h.mg_nconv;
    c->mg_div = h.val[7];
    if (!(c->mg_div <= c->mg_div_max)) c->mg_div_max = c->mg_div;
    // the timing keeps the launches that ran in full: the down kernel of every
    // applied cycle and of the final test, the up kernel of the applied ones
    if (c->kt_on == 2 || c->kt_on == 3) {
        const int applied = nconv < 0 ? maxcyc : nconv;
        const size_t keep = kt0 + 2 * (size_t)(c->kt_on == 2 ? applied + 1 : applied);
        if (c->kt_used > keep) {
            c->kt_used = keep;
            c->kt_cnt.resize(keep / 2);
        }
    }
    if (nconv < 0) {
        if (g_verbose) fprintf(stdout, " Warning: multigrid did not converge after %6d V-cycles.\n", maxcyc);
        nconv = maxcyc;
    }
    c->mg_last = std::min(nconv, 64);
    *nconv_out = nconv;
    return WD2_OK;
}

static int ppe(wd2_ctx *c, int cart, int solver, int msorit, double dk, double sortol, double sorrel,
               const double *u, const double *v, double *p, int *nconv_out) {
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    if (solver < 1 || solver > 7) {
        // pressure.f:234-236 prints and leaves p untouched; Ppe then reports msorit
        if (g_verbose) fprintf(stdout, " Wrong nPpeSolver flag passed to Ppe\n");
        *nconv_out = msorit;
        return WD2_OK;
    }
    launch_divergence(nx, ny, ld, 1, c->M, u, v, c->div, s);
    launch_ppe_rhs(nx, ny, ld, dk, c->div, c->any_blockage ? c->prow : nullptr, c->rhs, s);
    if (solver == 7) return ppe_mg(c, cart, msorit, dk, sortol, p, nconv_out);
    const PpeCoef cf{c->pc[0], c->pc[1], c->pc[2], c->pc[3], c->pc[4]};
    const PpeCoef *coef = nullptr;
    if (c->any_blockage || solver == 2 || solver == 3 || solver == 4) {
        if (!c->coef_ready) {
            launch_ppe_coef(nx, ny, ld, c->M, c->any_blockage ? c->prow : nullptr, c->pc[0], c->pc[1], c->pc[2],
                            c->pc[3], c->pc[4], s);
            c->coef_ready = true;
        }
        coef = &cf;
    }
    launch_sor_reset(c->sc, s);
    const double *b = c->rhs;
    if ((solver == 5 || solver == 6) && cart && !c->any_blockage && (g_tiled_sor || c->precision == WD2_PREC_MIXED)) {
        // temporally blocked sweeps, ping-pong between p and the scratch field;
        // mixed precision: FP64 residual + FP32 sweeps + FP64 accumulation
        const bool mx = c->precision == WD2_PREC_MIXED;
        auto sweep = mx ? launch_sor_tiled_mx : launch_sor_tiled;
        const double *uni = (g_uniform && c->uni_ok) ? c->uni : nullptr;
        const int K = mx ? sor_tiled_mx_sweeps() : sor_tiled_sweeps();
        launch_copy(ld, 0, nx + 1, 0, ny + 1, p, c->pscr, s);
        int m0 = 0, L = 0;
        const int nt = mx ? sor_tiled_mx_tiles(nx, ny, nullptr) : sor_tiled_tiles(nx, ny, nullptr);
        unsigned *part[2] = {(unsigned *)c->partials, (unsigned *)c->partials + nt};
        int prev_k = 0, prev_m0 = 0;
        // Batches of launches with a stopping check at the end of each, two
        // batches in flight: the host reads batch n's flag (copied to pinned
        // memory) while batch n+1 runs.  A batch issued after the solve
        // stopped returns at entry in every launch.  Batches grow from 8 to 64
        // launches, so a short solve is not overrun and a long one rarely
        // waits; after a solve that ran to msorit they are 128 launches from
        // the start (fewer checks; a solve that then stops early overruns by
        // at most two batches of launches that return at entry).
        int issued = 0, checked = 0, stop = 0;
        while (!stop) {
            if (m0 < msorit && issued - checked < 2) {
                const int Q = c->sor_capped ? 128 : std::min(64, 8 << std::min(issued, 3));
                // timing: one event pair around the batch's launches (no marker
                // between two launches, so their durations are not inflated)
                hipEvent_t *ev = kt_events(c, 1);
                if (ev) HIPCHK(hipEventRecord(ev[0], s));
                int nb = 0;
                for (int q = 0; q < Q && m0 < msorit; q++, L++, nb++) {
                    const int kr = std::min(K, msorit - m0);
                    const double *in = (L % 2 == 0) ? p : c->pscr;
                    double *out = (L % 2 == 0) ? c->pscr : p;
                    sweep(nx, ny, ld, c->M, b, sorrel, kr, in, out, part[L % 2], part[(L + 1) % 2], prev_k, prev_m0,
                          L - 1, sortol, msorit, c->sc, 0, uni, s);
                    prev_k = kr;
                    prev_m0 = m0;
                    m0 += kr;
                }
                if (ev) {
                    HIPCHK(hipEventRecord(ev[1], s));
                    c->kt_cnt.back() = nb;
                }
                // the stopping test of the batch's last launch
                launch_sor_tiled_check(nt, prev_k, prev_m0, L - 1, msorit, part[(L - 1) % 2], c->sc, s);
                HIPCHK(hipMemcpyAsync(&c->h_flag[issued % 2], &c->sc->sor_done, sizeof(int), hipMemcpyDeviceToHost,
                                      s));
                HIPCHK(hipEventRecord(c->bev[issued % 2], s));
                issued++;
                continue;
            }
            if (checked == issued) break;
            HIPCHK(hipEventSynchronize(c->bev[checked % 2]));
            stop = c->h_flag[checked % 2];
            checked++;
        }
        Scalars h;
        HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
        const int Ld = h.redo_launch, keep = h.redo_sweeps;
        const int krd = std::min(K, msorit - K * Ld);
        const double *in = (Ld % 2 == 0) ? p : c->pscr;
        double *out = (Ld % 2 == 0) ? c->pscr : p;
        if (L > 0) {
            if (keep < krd)   // stopped inside the launch: redo it with exactly `keep` sweeps
                sweep(nx, ny, ld, c->M, b, sorrel, keep, in, out, part[Ld % 2], nullptr, 0, 0, 0, sortol, msorit,
                      c->sc, 1, uni, s);
            if (out != p) launch_copy(ld, 0, nx + 1, 0, ny + 1, out, p, s);
        }
        int nconv = h.sor_nconv;
        c->sor_capped = nconv < 1;
        if (nconv < 1) {
            if (g_verbose) fprintf(stdout, " Warning: SOR iterations did not converge after %6d iterations.\n", msorit);
            nconv = msorit;
        }
        *nconv_out = nconv;
        return WD2_OK;
    }
    const int batch = (solver == 2) ? 1 : 64;
    int done = 0;
    if (solver == 1 && g_lex_rows && ny - 1 <= sor_lex_rows_max()) {
        launch_sor_lex_rows(nx, ny, ld, c->M, coef, c->rhs, c->beff, cart ? 1 : 0, sorrel, sortol, msorit, p, c->sc,
                            s);
        done = 1;
    }
    for (int it = 0; it < msorit && !done;) {
        for (int q = 0; q < batch && it < msorit; q++, it++) {
            if (solver == 5 || solver == 6) {
                if (!cart) { launch_rhs_cross(nx, ny, ld, c->M, c->rhs, p, c->beff, s); b = c->beff; }
                hipEvent_t *ev = kt_events(c, 2);
                launch_sor_rb_sweep(nx, ny, ld, c->M, coef, b, sorrel, sortol, it + 1, p, c->partials, c->sc, s, ev);
            } else if (solver == 1) {
                if (!cart) { launch_rhs_cross(nx, ny, ld, c->M, c->rhs, p, c->beff, s); b = c->beff; }
                launch_sor_lex_sweep(nx, ny, ld, c->M, coef, b, sorrel, sortol, msorit, p, c->tile_flags,
                                     ++c->sweep_id, c->partials, c->sc, s);
            } else if (solver == 3 || solver == 4) {
                for (int k = 2; k <= 3; k++) {
                    if (!cart) { launch_rhs_cross(nx, ny, ld, c->M, c->rhs, p, c->beff, s); b = c->beff; }
                    const int nl = (ny - k) / 2 + 1;
                    if (nl < 1) continue;
                    const long n = (long)(nx - 1) * nl;
                    launch_slor_assemble(nx, ny, ld, c->M, coef, b, p, k, 2, nl, c->ch.a, c->ch.d, c->ch.c,
                                         c->ch.b, c->sc, s);
                    launch_trid(c->tw, n, nx - 1, c->ch.a, c->ch.d, c->ch.c, c->ch.b, c->ch.x, s);
                    launch_slor_relax(nx, ny, ld, k, 2, nl, c->ch.x, sorrel, solver == 4 ? 1 : 0, p, nullptr,
                                      nullptr, c->sc, 6, s);
                }
                launch_slor_finish(0, sortol, msorit, nullptr, c->sc, s);
            } else {   // solver 2: lines in order j = 2..ny
                if (!cart) { launch_rhs_cross(nx, ny, ld, c->M, c->rhs, p, c->beff, s); b = c->beff; }
                for (int j = 2; j <= ny; j++) {
                    launch_slor_assemble(nx, ny, ld, c->M, coef, b, p, j, 1, 1, c->ch.a, c->ch.d, c->ch.c,
                                         c->ch.b, c->sc, s);
                    launch_trid(c->tw, nx - 1, nx - 1, c->ch.a, c->ch.d, c->ch.c, c->ch.b, c->ch.x, s);
                    launch_slor_relax(nx, ny, ld, j, 1, 1, c->ch.x, sorrel, 0, p, nullptr, nullptr, c->sc, 6, s);
                }
                launch_slor_finish(0, sortol, msorit, nullptr, c->sc, s);
            }
        }
        if (solver == 5 || solver == 6) launch_sor_rb_check(nx, ny, it, sortol, msorit, c->partials, c->sc, s);
        HIPCHK(hipMemcpyAsync(&done, &c->sc->sor_done, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    int nconv = 0;
    HIPCHK(hipMemcpyAsync(&nconv, &c->sc->sor_nconv, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (nconv < 1) {
        if (g_verbose) fprintf(stdout, " Warning: SOR iterations did not converge after %6d iterations.\n", msorit);
        nconv = msorit;
    }
    *nconv_out = nconv;
    return WD2_OK;
}

// ------------------------------------------------------------- main.f step
static int cold_start(wd2_ctx *c, const wd2_params *P, int *nsor) {
    launch_vel_bc<double>(c->R, c->ld, c->u, c->v, false, &c->sc->err, c->s);
    int st = ppe(c, P->lCartesGrid, P->nPpeSolver, P->msorit, P->dk, P->sortol, P->sorrel, c->u, c->v, c->p, nsor);
    if (st) return st;
    launch_pres_bc<double>(c->R, c->ld, c->p, &c->sc->err, c->s);
    launch_project(c->R, c->ld, P->dk, c->M, c->p, c->u, c->v, c->s);
    launch_vel_bc<double>(c->R, c->ld, c->u, c->v, false, &c->sc->err, c->s);
    return check_err(c);
}

// One time step of main.f:690-972: the time-level copies, the momentum/energy
// iterations (nAuxMomentum, filters, boundary fills, Ppe, Project and, with
// the energy equation, ThermEnergy and EqState), the temperature filter, the
// step-end fills and the PrintDiff norms dif[0..3] = |p-pn|, |u-un|, |v-vn|,
// |t-tn| (dif may be null; only its first three are written without the
// energy equation).
// The order of this step's extrapolated start: 0 (the plain start) unless
// asked for, not suspended, the multigrid solver and a step of history since
// the fields were set; then 1, p = 2 p^n - p^(n-1) on the interior cells, or
// 2, p = 3 p^n - 3 p^(n-1) + p^(n-2), when asked for and the step before
// wrote its history slot.  The converged solve is the same either way (the
// ghosts stay the values PresBoundCond set); only the number of V-cycles
// changes.
static int mg_extrap_mode(const wd2_ctx *c, const wd2_params *P) {
    if (!(c->mg_extrap > 0 && !c->extrap_off && P->nPpeSolver == 7 && c->p_hist > 0)) return 0;
    return (c->mg_extrap >= 2 && c->x_run > 0) ? 2 : 1;
}
static int mg_extrap_alloc(wd2_ctx *c) {
    if (!c->phist) HIPCHK(hipMalloc(&c->phist, 2 * c->nf * sizeof(double)));
    return WD2_OK;
}
// The history slots of a step of order x: p^(n-1) goes to slot xw, p^(n-2)
// is the other slot (order 2 only).
static double *extrap_p1(const wd2_ctx *c) { return c->phist + (size_t)c->xw * c->nf; }
static const double *extrap_p2(const wd2_ctx *c, int x) {
    return x == 2 ? c->phist + (size_t)(c->xw ^ 1) * c->nf : nullptr;
}
// The V-cycles a solve from the extrapolated start applies before its
// tolerance may stop it.  A guess so close that one cycle meets the
// tolerance leaves the error in the smooth modes, which the divergence test
// barely sees, damped by one cycle only; the next steps extrapolate that
// error again, and over thousands of such steps it grows (a 1024^2 cavity:
// mixed against FP64 p 2e-4 relative L2 after 5000 steps, from 1e-7).  Two
// cycles damp it more than the extrapolation can amplify it.
static int mg_min_cycles(const wd2_params *P) { return std::min(2, std::max(P->msorit, 0)); }
// A step of order x completed: flip the slot, count the history.
static void extrap_advance(wd2_ctx *c, int x) {
    if (x) {
        c->xw ^= 1;
        c->x_run++;
    } else {
        c->x_run = 0;
    }
    c->p_hist++;
}

// The adjoint's runs (forward, segment re-runs, tangent re-runs) solve from
// the plain initial guess, so that a re-run from a checkpoint repeats the
// forward run bit for bit without the extrapolation's history.
struct ExtrapOff {
    wd2_ctx *c;
    bool old;
    explicit ExtrapOff(wd2_ctx *c_) : c(c_), old(c_->extrap_off) { c->extrap_off = true; }
    ~ExtrapOff() { c->extrap_off = old; }
};

static int one_step(wd2_ctx *c, const wd2_params *P, int *nql, int *nsor, double *dif) {
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    const wd2_step_options &O = c->opt;
    const bool therm = O.nthermen == 1;
    hipStream_t s = c->s;
    const int xp = mg_extrap_mode(c, P);
    if (xp) {
        int st0 = mg_extrap_alloc(c);
        if (st0) return st0;
        launch_copy(ld, 0, nx + 1, 0, ny + 1, c->pn, extrap_p1(c), s);   // the pressure of the step before
    }
    // main.f:696-704 and the first iteration's us = u, vs = v (main.f:741-747)
    // in one launch.  Without a density field d and dn are both still all +0
    // (nothing has written d), so that copy is skipped.
    {
        const double *src[6] = {c->p, c->u, c->v, c->u, c->v, nullptr};
        double *dst[6] = {c->pn, c->un, c->vn, c->us, c->vs, nullptr};
        int n = 5;
        if (c->d_live) { src[n] = c->d; dst[n] = c->dn; n++; }
        launch_copy_set(ld, 0, nx + 1, 0, ny + 1, src, dst, n, s);
        if (therm) launch_copy(ld, 0, nx + 1, 0, ny + 1, c->t, c->tn, s);
    }
    // main.f:735: one iteration without the energy equation
    const int nme = (!therm && O.nmeiter > 0) ? 1 : O.nmeiter;
    int st;
    for (int l = 1; l <= nme; l++) {
        // main.f:741-747 (the first iteration's us, vs come with the copies above)
        if (l > 1) {
            const double *src[2] = {c->u, c->v};
            double *dst[2] = {c->us, c->vs};
            launch_copy_set(ld, 0, nx + 1, 0, ny + 1, src, dst, 2, s);
        }
        if (therm) launch_copy(ld, 0, nx + 1, 0, ny + 1, c->t, c->ts, s);
        if ((st = naux_momentum(c, P->mqiter, P->dk, P->re, P->fr, P->qtol, nql))) return st;
        if (c->adj_us) {
            const double *src[2] = {c->us, c->vs};
            double *dst[2] = {c->adj_us, c->adj_vs};
            launch_copy_set(ld, 0, nx + 1, 0, ny + 1, src, dst, 2, s);
        }
        // main.f:783-791
        if (O.nfiltu == 1) launch_filter(c->R, c->TH, ld, VU, O.fpu, c->us, c->qh, s);
        if (O.nfiltv == 1) launch_filter(c->R, c->TH, ld, VV, O.fpv, c->vs, c->qh, s);
        launch_vel_bc<double>(c->R, ld, c->us, c->vs, false, &c->sc->err, s);
        launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
        if (xp && l == 1) launch_mg_extrap(nx, ny, ld, c->p, extrap_p1(c), extrap_p2(c, xp), s);
        c->mg_minc = (xp && l == 1) ? mg_min_cycles(P) : 0;
        st = ppe(c, P->lCartesGrid, P->nPpeSolver, P->msorit, P->dk, P->sortol, P->sorrel, c->us, c->vs, c->p, nsor);
        c->mg_minc = 0;
        if (st) return st;
        launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
        launch_project(c->R, ld, P->dk, c->M, c->p, c->us, c->vs, s);
        launch_vel_bc<double>(c->R, ld, c->us, c->vs, false, &c->sc->err, s);
        launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
        // main.f:840-855
        if (therm)
            launch_therm_energy(c->R, c->TH, ld, P->dk, O.pe, c->M, c->un, c->vn, c->us, c->vs, c->tn, c->ts, c->tt,
                                c->ch, c->tw, &c->sc->err, s);
        if (O.neqstate == 1)
            launch_eq_state(nx, ny, ld, O.uref, O.densref, O.tmax, O.tref, O.rconst, c->p, therm ? c->ts : c->t,
                            c->d, s);
        // main.f:857-875: the iteration test needs DiffMaxNorm of (u,us),
        // (v,vs), (t,ts) from the second iteration on (after the last one the
        // loop ends either way)
        const bool test = l > 1 && l < nme;
        if (test) {
            launch_max_norm(nx, ny, ld, c->u, c->us, c->partials, c->sc, 5, s);
            launch_max_norm(nx, ny, ld, c->v, c->vs, c->partials + RED_MAX_BLOCKS, c->sc, 6, s);
            launch_max_norm(nx, ny, ld, therm ? c->t : c->u, therm ? c->ts : c->u, c->partials + 2 * RED_MAX_BLOCKS,
                            c->sc, 7, s);
        }
        // u = us, v = vs, t = ts (swap the buffers: same values)
        std::swap(c->u, c->us);
        std::swap(c->v, c->vs);
        if (therm) std::swap(c->t, c->ts);
        if (test) {
            double dm[3];
            HIPCHK(hipMemcpyAsync(dm, &c->sc->val[5], 3 * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            // main.f:873: difmax = max(dif(1),dif(2),dif(3)) (Fortran MAX)
            double dmax = dm[0];
            if (dm[1] > dmax) dmax = dm[1];
            if (dm[2] > dmax) dmax = dm[2];
            if (dmax < O.dmeittol) break;
        }
    }
    // main.f:890-895
    if (therm && O.nfiltt == 1) launch_filter(c->R, c->TH, ld, VT, O.fpt, c->t, c->qh, s);
    launch_vel_bc<double>(c->R, ld, c->u, c->v, false, &c->sc->err, s);
    launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
    if (therm) launch_temp_bc(c->R, c->TH, ld, c->t, &c->sc->err, s);
    // main.f:962-965
    launch_max_norm(nx, ny, ld, c->pn, c->p, c->partials, c->sc, 2, s);
    launch_max_norm(nx, ny, ld, c->un, c->u, c->partials + RED_MAX_BLOCKS, c->sc, 3, s);
    launch_max_norm(nx, ny, ld, c->vn, c->v, c->partials + 2 * RED_MAX_BLOCKS, c->sc, 4, s);
    if (therm) launch_max_norm(nx, ny, ld, c->tn, c->t, c->partials + 3 * RED_MAX_BLOCKS, c->sc, 5, s);
    if (dif) HIPCHK(hipMemcpyAsync(dif, &c->sc->val[2], (therm ? 4 : 3) * sizeof(double), hipMemcpyDeviceToHost, s));
    extrap_advance(c, xp);
    return check_err(c);
}

namespace wd2 {
int api_fail(int code, const char *msg) { return fail(code, "%s", msg); }
int noalloc_trid_fallbacks();
}  // namespace wd2

// ------------------------------------------------------ hipGraph steps
// A time step on the multigrid path as one graph replay.  The host-driven
// step reads two stopping tests (the q-l loop's and the Poisson solve's)
// while it queues work; in a graph both loops are queued to a budget instead
// (the last step's count plus margin): iterations after a stopping test has
// fired return at entry, exactly as in the host-driven step, so the fields
// are the same.  If a budget ran out before its test fired, the step is
// restored from the copies the graph took first and redone host-driven.
static bool graph_ok(const wd2_ctx *c, const wd2_params *P) {
    return c->graph_on && P->nPpeSolver == 7 && P->mqiter >= 1 && !c->any_porous && !c->any_blockage &&
           c->opt.nthermen != 1 && c->opt.neqstate != 1 && !c->d_live && c->kt_on == 0;
}

// The step's work on the stream (the body of a capture): main.f:690-972 with
// Q q-l iterations and C V-cycles queued.
static void step_body(wd2_ctx *c, const wd2_params *P, int Q, int C, int xp) {
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    const wd2_step_options &O = c->opt;
    {   // the step's start, for a redo (pn: the extrapolation's history)
        const double *src[4] = {c->u, c->v, c->p, c->pn};
        double *dst[4] = {c->gsave, c->gsave + c->nf, c->gsave + 2 * c->nf, c->gsave + 3 * c->nf};
        launch_copy_set(ld, 0, nx + 1, 0, ny + 1, src, dst, 4, s);
    }
    if (xp) launch_copy(ld, 0, nx + 1, 0, ny + 1, c->pn, extrap_p1(c), s);
    {
        const double *src[5] = {c->p, c->u, c->v, c->u, c->v};
        double *dst[5] = {c->pn, c->un, c->vn, c->us, c->vs};
        launch_copy_set(ld, 0, nx + 1, 0, ny + 1, src, dst, 5, s);
    }
    {   // momentum.f:114-119
        const double *src[2] = {c->un, c->vn};
        double *dst[2] = {c->us, c->vs};
        launch_copy_set(ld, 1, nx + 1, 1, ny + 1, src, dst, 2, s);
    }
    const int *stop = &c->sc->ql_stop;
    launch_ql_reset(c->sc, s);
    for (int m = 1; m <= Q; m++) {
        if (c->any_outlet) launch_vel_bc<double>(c->R, ld, c->us, c->vs, true, &c->sc->err, s, stop);
        if (c->precision == WD2_PREC_MIXED)
            momentum_pair<float>(c, P->dk, P->re, P->fr, c->Tf, c->chf, c->chf2, c->twf, m, P->qtol);
        else
            momentum_pair<double>(c, P->dk, P->re, P->fr, c->T, c->ch, c->ch2, c->tw, m, P->qtol);
    }
    if (O.nfiltu == 1) launch_filter(c->R, c->TH, ld, VU, O.fpu, c->us, c->qh, s);
    if (O.nfiltv == 1) launch_filter(c->R, c->TH, ld, VV, O.fpv, c->vs, c->qh, s);
    launch_vel_bc<double>(c->R, ld, c->us, c->vs, false, &c->sc->err, s);
    launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
    if (xp) launch_mg_extrap(nx, ny, ld, c->p, extrap_p1(c), extrap_p2(c, xp), s);
    // Ppe (pressure.f:30) with the multigrid solver
    launch_divergence(nx, ny, ld, 1, c->M, c->us, c->vs, c->div, s);
    launch_ppe_rhs(nx, ny, ld, P->dk, c->div, nullptr, c->rhs, s);
    const double *uni = (g_uniform && c->uni_ok) ? c->uni : nullptr;
    const double *b = P->lCartesGrid ? c->rhs : c->beff;
    launch_mg_reset(c->sc, s);
    for (int cyc = 1; cyc <= C; cyc++) {
        if (!P->lCartesGrid) launch_rhs_cross(nx, ny, ld, c->M, c->rhs, c->p, c->beff, s);
        if (c->precision == WD2_PREC_MIXED)
            launch_mg_cycle<float>(c->mg, nx, ny, ld, b, c->p, c->M.rau, c->M.rgv, uni, P->dk, P->sortol, cyc,
                                   P->msorit, c->mg_part, c->sc, s, nullptr, nullptr, xp ? mg_min_cycles(P) : 0);
        else
            launch_mg_cycle<double>(c->mg, nx, ny, ld, b, c->p, c->M.rau, c->M.rgv, uni, P->dk, P->sortol, cyc,
                                    P->msorit, c->mg_part, c->sc, s, nullptr, nullptr, xp ? mg_min_cycles(P) : 0);
    }
    launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
    launch_project(c->R, ld, P->dk, c->M, c->p, c->us, c->vs, s);
    launch_vel_bc<double>(c->R, ld, c->us, c->vs, false, &c->sc->err, s);
    launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
    // u = us, v = vs: the host swaps the buffers after the step, so the rest
    // of the step works on us, vs as the new u, v (main.f:946-965)
    launch_vel_bc<double>(c->R, ld, c->us, c->vs, false, &c->sc->err, s);
    launch_pres_bc<double>(c->R, ld, c->p, &c->sc->err, s);
    launch_max_norm(nx, ny, ld, c->pn, c->p, c->partials, c->sc, 2, s);
    launch_max_norm(nx, ny, ld, c->un, c->us, c->partials + RED_MAX_BLOCKS, c->sc, 3, s);
    launch_max_norm(nx, ny, ld, c->vn, c->vs, c->partials + 2 * RED_MAX_BLOCKS, c->sc, 4, s);
    (void)hipMemcpyAsync(c->h_sc, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, s);
}

// Queue one graph step (capturing it first if no graph of this shape is
// cached); graph_end completes it.
static int graph_begin(wd2_ctx *c, const wd2_params *P) {
    int st = mg_prepare(c);
    if (st) return st;
    if (!c->gsave) HIPCHK(hipMalloc(&c->gsave, 4 * c->nf * sizeof(double)));
    const int xp = mg_extrap_mode(c, P);
    if (xp && (st = mg_extrap_alloc(c))) return st;
    if (!c->h_sc) HIPCHK(hipHostMalloc((void **)&c->h_sc, sizeof(Scalars), hipHostMallocDefault));
    const int Q = std::min(P->mqiter, c->ql_last + 2);
    const int C = std::min(P->msorit + 1, c->mg_last + 2);
    hipGraphExec_t exec = nullptr;
    for (auto &g : c->graphs)
        if (g.q == Q && g.cyc == C && g.u == c->u && g.epoch == g_graph_epoch && g.xp == xp && g.xw == c->xw &&
            std::memcmp(&g.P, P, sizeof(wd2_params)) == 0)
            exec = g.exec;
    if (!exec) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        step_body(c, P, Q, C, xp);
        const hipError_t ec = hipStreamEndCapture(c->s, &graph);
        if (ec != hipSuccess) return fail(WD2_ERR_HIP, "step capture: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(WD2_ERR_HIP, "step graph instantiation: %s", hipGetErrorString(ei));
        if (c->graphs.size() >= 16) {
            (void)hipGraphExecDestroy(c->graphs.front().exec);
            c->graphs.erase(c->graphs.begin());
        }
        c->graphs.push_back({Q, C, c->u, *P, g_graph_epoch, xp, c->xw, exec});
    }
    HIPCHK(hipGraphLaunch(exec, c->s));
    c->g_q = Q;
    c->g_c = C;
    c->g_x = xp;
    return WD2_OK;
}

static int graph_end(wd2_ctx *c, const wd2_params *P, int *nql, int *nsor, double *dif) {
    HIPCHK(hipStreamSynchronize(c->s));
    const Scalars h = *c->h_sc;
    if (h.err) return check_err(c);
    const bool ql_short = h.ql_stop == 0 && c->g_q < P->mqiter;
    const bool mg_short = h.mg_done == 0;
    if (ql_short || mg_short) {
        // a budget ran out: restore the step's start and redo it host-driven
        const double *src[4] = {c->gsave, c->gsave + c->nf, c->gsave + 2 * c->nf, c->gsave + 3 * c->nf};
        double *dst[4] = {c->u, c->v, c->p, c->pn};
        launch_copy_set(c->ld, 0, c->nx + 1, 0, c->ny + 1, src, dst, 4, c->s);
        c->graph_redos++;
        int st = one_step(c, P, nql, nsor, dif);
        if (st) return st;
        HIPCHK(hipStreamSynchronize(c->s));
        c->ql_last = *nql > 0 ? *nql : P->mqiter;
        return WD2_OK;
    }
    std::swap(c->u, c->us);
    std::swap(c->v, c->vs);
    extrap_advance(c, c->g_x);
    *nql = h.ql_stop ? h.ql_stop : -1;
    int nconv = h.mg_nconv;
    c->mg_div = h.val[7];
    if (!(c->mg_div <= c->mg_div_max)) c->mg_div_max = c->mg_div;
    if (nconv < 0) {
        if (g_verbose) fprintf(stdout, " Warning: multigrid did not converge after %6d V-cycles.\n", P->msorit);
        nconv = P->msorit;
    }
    *nsor = nconv;
    c->mg_last = std::min(nconv, 64);
    c->ql_last = *nql > 0 ? *nql : P->mqiter;
    if (dif) for (int q = 0; q < 3; q++) dif[q] = h.val[2 + q];
    c->graph_steps++;
    return WD2_OK;
}

static int fallbacks(wd2_ctx *c) {
    int n = 0;
    if (hipMemcpy(&n, c->tw.nfallback, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return n;
}

// ================================================================ adjoint
// Reverse-mode gradient of a run of time steps (wd2_adj_*), FP64, on a
// one-region cavity with no-slip walls (boundary profiles allowed), the
// multigrid Poisson solver and lCartesGrid set.  The map differentiated is
// main.f's step with each q-l loop and Poisson solve at its converged
// solution:
// * q-l loop: the iterates stop where the second split's increment vanishes,
//   Pa S1(w)^-1 b(w) = 0 (Pa: rows that are not identity rows of the second
//   split, S1: the first split's matrix with the chain solver's first-row
//   quirk, b: the assembled right-hand side).  Its adjoint is the fixed point
//   of nu <- nu + S1^-T Pa S2^-T Pa (G^T nu + wbar), the transpose of the
//   forward iteration (G = db/dw - dS1/dw x1, x1 = S1^-1 b); the converged
//   nu gives the adjoints of un, vn and of the points the loop leaves alone;
// * Ppe: p on the interior cells solves A p = div(us, vs)/dk - (frozen ghost
//   couplings); A is symmetric on the Cartesian five-point form, so the
//   adjoint solve A lam = pbar is another multigrid solve;
// * boundary fills and Project are linear and replayed in reverse order.
// Checkpoints of (u, v, p) every `every` steps bound the memory; the
// backward pass re-runs each segment from its checkpoint, keeping (u, v, p)
// and the converged (us, vs) of its steps, then walks the segment backwards.
struct wd2_adj {
    wd2_ctx *c = nullptr;
    int every = 20, nsteps = 0, nseg_cap = 0;
    bool have_fwd = false;
    wd2_params P{};
    double fp_tol = 1e-12, ppe_tol = 1e-13;
    int fp_max = 200;
    int mg_last = 4;   // V-cycles of the last adjoint Poisson solve (first batch size)
    double *ckpt = nullptr;   // nseg_cap x (u, v, p)
    double *tape = nullptr;   // every x (u, v, p, us, vs)
    double *fin = nullptr;    // (u, v, p) after the run
    // adjoint fields: the state's (ub, vb, pb); scratch tu, tv (w.r.t. us,
    // vs), nub, nvb (w.r.t. un, vn); the Poisson adjoint lam (zero ghosts)
    // and its right-hand side; the ConvCoef adjoints co (6 fields)
    double *fstore = nullptr;
    double *ub, *vb, *pb, *tu, *tv, *nub, *nvb, *lam, *arhs;
    AdjCoefs co{};
    double *bcb = nullptr;    // 8 x W boundary-value adjoints
    int W = 0;
    // chain vectors: S1^T and S2^T (A, D, C), nu, r, y, z, coefficient adjoints
    double *cstore = nullptr;
    double *At1, *Dt1, *Ct1, *At2, *Dt2, *Ct2, *nu, *r, *y, *z, *a1b, *d1b, *c1b;
    // tangent-linear pass (wd2_adj_tangent): dual fields us, vs, un, vn, the
    // dual time-level-n terms of both components, the dual chains of both
    // splits; the tangent boundary values (a Regions with zero constants and
    // one profile per wall side and velocity component)
    Dual *tl_store = nullptr;
    Dual *tUS, *tVS, *tUN, *tVN, *tC1, *tD1, *tC2, *tD2;
    ChainT<Dual> tch1{}, tch2{};
    double *tl_bcv = nullptr, *tl_pval = nullptr;
    int *tl_poff = nullptr;
    // maxima of the adjoint q-l updates (two slots per iteration parity) and
    // their pinned copies; qev: the events after each parity's copy
    unsigned long long *red = nullptr, *h_red = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    hipEvent_t qev[2] = {nullptr, nullptr};
    wd2_adj_stats st{};
};

static void adj_free(wd2_adj *a) {
    if (!a) return;
    (void)hipFree(a->ckpt);
    (void)hipFree(a->tape);
    (void)hipFree(a->fin);
    (void)hipFree(a->fstore);
    (void)hipFree(a->bcb);
    (void)hipFree(a->cstore);
    (void)hipFree(a->tl_store);
    (void)hipFree(a->tl_bcv);
    (void)hipFree(a->tl_pval);
    (void)hipFree(a->tl_poff);
    (void)hipFree(a->red);
    if (a->h_red) (void)hipHostFree(a->h_red);
    for (hipEvent_t e : a->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : a->qev)
        if (e) (void)hipEventDestroy(e);
    delete a;
}

static int adj_check(const wd2_ctx *c, const wd2_params *P) {
    if (c->precision != WD2_PREC_FP64) return fail(WD2_ERR_UNSUPPORTED, "adjoint: FP64 precision only");
    if (c->R.ni != 1 || c->R.nj != 1 || c->h_rtype[0] != RM_INTERN)
        return fail(WD2_ERR_UNSUPPORTED, "adjoint: one fluid region only");
    const int want[4] = {1, c->nx, 1, c->ny};
    for (int side = WEST; side <= NORTH; side++) {
        if (c->R.border(1, 1, side) != want[side - 1])
            return fail(WD2_ERR_UNSUPPORTED, "adjoint: the region must cover the grid");
        if (c->R.btype(1, 1, side) != BM_WALL1)
            return fail(WD2_ERR_UNSUPPORTED, "adjoint: no-slip walls (type 1) on every side only");
    }
    const wd2_step_options &O = c->opt;
    if (c->d_live || O.nthermen == 1 || O.neqstate == 1 || O.nfiltu == 1 || O.nfiltv == 1)
        return fail(WD2_ERR_UNSUPPORTED, "adjoint: no energy equation, density or filters");
    if (P->nPpeSolver != 7 || !P->lCartesGrid || P->mqiter < 1)
        return fail(WD2_ERR_UNSUPPORTED, "adjoint: nPpeSolver 7, lCartesGrid and max_ql_iter >= 1 only");
    return WD2_OK;
}

static int adj_alloc(wd2_adj *a) {
    wd2_ctx *c = a->c;
    const size_t nf = c->nf, nch = chain_rows(c->nx, c->ny);
    HIPCHK(hipMalloc(&a->tape, 5 * (size_t)a->every * nf * sizeof(double)));
    HIPCHK(hipMalloc(&a->fin, 3 * nf * sizeof(double)));
    HIPCHK(hipMalloc(&a->fstore, 15 * nf * sizeof(double)));
    HIPCHK(hipMemsetAsync(a->fstore, 0, 15 * nf * sizeof(double), c->s));
    double *f = a->fstore;
    double **fp[9] = {&a->ub, &a->vb, &a->pb, &a->tu, &a->tv, &a->nub, &a->nvb, &a->lam, &a->arhs};
    for (int k = 0; k < 9; k++) *fp[k] = f + k * nf;
    a->co = AdjCoefs{f + 9 * nf, f + 10 * nf, f + 11 * nf, f + 12 * nf, f + 13 * nf, f + 14 * nf};
    a->W = std::max(c->nx, c->ny) + 2;
    HIPCHK(hipMalloc(&a->bcb, 8 * (size_t)a->W * sizeof(double)));
    HIPCHK(hipMalloc(&a->cstore, 13 * nch * sizeof(double)));
    HIPCHK(hipMemsetAsync(a->cstore, 0, 13 * nch * sizeof(double), c->s));
    double **cp[13] = {&a->At1, &a->Dt1, &a->Ct1, &a->At2, &a->Dt2, &a->Ct2, &a->nu,
                       &a->r,   &a->y,   &a->z,   &a->a1b, &a->d1b, &a->c1b};
    for (int k = 0; k < 13; k++) *cp[k] = a->cstore + k * nch;
    HIPCHK(hipMalloc(&a->red, 4 * sizeof(unsigned long long)));
    HIPCHK(hipHostMalloc((void **)&a->h_red, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    for (auto &e : a->ev) HIPCHK(hipEventCreate(&e));
    for (auto &e : a->qev) HIPCHK(hipEventCreate(&e));
    return WD2_OK;
}

static void adj_copy3(wd2_ctx *c, const double *u, const double *v, const double *p, double *du, double *dv,
                      double *dp) {
    const double *src[3] = {u, v, p};
    double *dst[3] = {du, dv, dp};
    launch_copy_set(c->ld, 0, c->nx + 1, 0, c->ny + 1, src, dst, 3, c->s);
}

// max |x| over n elements of up to two vectors, read back on the host
static int adj_max2(wd2_adj *a, long n, const double *x0, const double *x1, double *m0, double *m1) {
    hipStream_t s = a->c->s;
    HIPCHK(hipMemsetAsync(a->red, 0, 2 * sizeof(unsigned long long), s));
    launch_adj_maxabs(n, x0, a->red, 0, s);
    if (x1) launch_adj_maxabs(n, x1, a->red, 1, s);
    HIPCHK(hipMemcpyAsync(a->h_red, a->red, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memcpy(m0, &a->h_red[0], sizeof(double));
    if (m1) std::memcpy(m1, &a->h_red[1], sizeof(double));
    return WD2_OK;
}

// tu, tv += G_s^T nu and (with_n) nub, nvb += G_n^T nu at the converged
// iterate (us, vs) of the step from (un, vn); x1 in the first chain's x
static void adj_gtrans(wd2_adj *a, const double *us, const double *vs, const double *un, const double *vn,
                       int with_n) {
    wd2_ctx *c = a->c;
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    const long n1 = (long)nx * (ny - 1), n2 = (long)(nx - 1) * ny, n = n1 + n2;
    const Chains cs{n1, n1, n2};
    launch_adj_coef(n, cs, c->ch.a, c->ch.d, c->ch.c, a->nu, c->ch.x, nullptr, c->zm, a->a1b, a->d1b, a->c1b, s);
    for (int comp = 1; comp <= 2; comp++) {
        const long off = comp == 1 ? 0 : n1;
        // co is all zero here: k_adj_cc clears what it reads
        launch_adj_asm(comp, nx, ny, ld, a->P.dk, a->P.re, c->M, us, vs, un, vn, a->a1b + off, a->d1b + off,
                       a->c1b + off, a->nu + off, with_n, comp == 1 ? a->tu : a->tv, comp == 1 ? a->nub : a->nvb,
                       a->co, s);
        launch_adj_cc(comp, nx, ny, ld, c->M, a->co, with_n, a->tu, a->tv, a->nub, a->nvb, s);
    }
}

// One step backwards: (ub, vb, pb) from the adjoint of the step's end state
// to that of its start state; bcb accumulates the boundary values' adjoints.
static int adj_step(wd2_adj *a, const double *un, const double *vn, const double *us, const double *vs) {
    wd2_ctx *c = a->c;
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    const wd2_params &P = a->P;
    const long n1 = (long)nx * (ny - 1), n2 = (long)(nx - 1) * ny, n = n1 + n2;
    const Chains cs{n1, n1, n2};
    // main.f:946-965 step-end fills, then (in reverse) PresBoundCond,
    // Project, VelBoundCond, PresBoundCond after the solve
    launch_adj_pbc(nx, ny, ld, a->pb, s);
    launch_adj_vbc(c->R, ld, a->ub, a->vb, a->bcb, s);
    launch_adj_pbc(nx, ny, ld, a->pb, s);
    launch_adj_vbc(c->R, ld, a->ub, a->vb, a->bcb, s);
    launch_adj_project(nx, ny, ld, P.dk, c->M, a->ub, a->vb, a->pb, s);
    launch_adj_pbc(nx, ny, ld, a->pb, s);
    // Ppe: lam = A^-1 pbar on the interior cells (warm-started from the last
    // step's lam), then the divergence and the frozen ghosts take it
    launch_copy(ld, 2, nx, 2, ny, a->pb, a->arhs, s);
    launch_fill(ld, 2, nx, 2, ny, 0.0, a->pb, s);
    double bmax = 0.0;
    int st = adj_max2(a, (long)c->nf, a->arhs, nullptr, &bmax, nullptr);
    if (st) return st;
    int ncyc = 0;
    if (bmax > 0.0) {
        const int ml = c->mg_last;
        const double md = c->mg_div, mdm = c->mg_div_max;
        c->mg_last = a->mg_last;
        st = ppe_mg(c, 1, P.msorit, 1.0, a->ppe_tol * bmax, a->lam, &ncyc, a->arhs);
        a->mg_last = c->mg_last;
        const bool conv = c->mg_div < a->ppe_tol * bmax;
        c->mg_last = ml;
        c->mg_div = md;
        c->mg_div_max = mdm;
        if (st) return st;
        if (!conv) a->st.adj_unconverged++;
        launch_adj_div(nx, ny, ld, P.dk, c->M, a->lam, a->ub, a->vb, a->pb, s);
    } else {
        launch_adj_zero((long)c->nf, a->lam, s);
    }
    a->st.ppe_cycles_max = std::max(a->st.ppe_cycles_max, ncyc);
    a->st.ppe_cycles_mean += ncyc;
    // PresBoundCond and VelBoundCond before the solve
    launch_adj_pbc(nx, ny, ld, a->pb, s);
    launch_adj_vbc(c->R, ld, a->ub, a->vb, a->bcb, s);
    // q-l loop: the chains at the converged iterate, x1 = S1^-1 b
    MomTemps t = c->T, t2 = c->T;
    t2.cnvn = c->T.cnvn2;
    t2.difn = c->T.difn2;
    auto shift = [](const Chain &q, long o) { return Chain{q.a + o, q.d + o, q.c + o, q.b + o, q.x + o}; };
    launch_mom_n<double>(nx, ny, ld, 1, c->M, un, vn, t.cnvn, t.difn, s);
    launch_mom_n<double>(nx, ny, ld, 2, c->M, un, vn, t2.cnvn, t2.difn, s);
    launch_mom_asm_pair<double>(c->R, ld, P.dk, P.re, P.fr, c->M, t, t2, us, vs, un, vn, c->d, c->dn, c->ch, c->ch2,
                                c->zm, shift(c->ch, n1), shift(c->ch2, n1), c->zm + n1, s, &c->sc->err, 0);
    launch_trid<double>(c->tw, n, cs, c->ch.a, c->ch.d, c->ch.c, c->ch.b, c->ch.x, s);
    launch_adj_tchain(n, cs, c->ch.a, c->ch.d, c->ch.c, a->At1, a->Dt1, a->Ct1, s);
    launch_adj_tchain(n, cs, c->ch2.a, c->ch2.d, c->ch2.c, a->At2, a->Dt2, a->Ct2, s);
    launch_adj_zero(n, a->nu, s);
    const double *fsrc[2] = {a->ub, a->vb};
    double *fdst[2] = {a->tu, a->tv};
    // Pipelined by one iteration, as the forward q-l loop: iteration m is
    // queued before the host reads iteration m-1's maxima; the extra
    // iteration after the test is met only refines nu.
    auto converged = [&](int m) {
        double zmax, numax;
        std::memcpy(&zmax, &a->h_red[2 * (m % 2)], sizeof(double));
        std::memcpy(&numax, &a->h_red[2 * (m % 2) + 1], sizeof(double));
        return zmax <= a->fp_tol * numax;
    };
    int m = 1;
    bool conv = false;
    for (; m <= a->fp_max; m++) {
        const int sl = 2 * (m % 2);
        launch_copy_set(ld, 0, nx + 1, 0, ny + 1, fsrc, fdst, 2, s);
        if (m > 1) adj_gtrans(a, us, vs, un, vn, 0);
        launch_adj_gather(nx, ny, ld, a->tu, a->tv, c->zm, a->r, a->red, sl, s);
        launch_trid<double>(c->tw, n, cs, a->At2, a->Dt2, a->Ct2, a->r, a->y, s);
        launch_adj_zmask(n, c->zm, a->y, a->y, s);
        launch_trid<double>(c->tw, n, cs, a->At1, a->Dt1, a->Ct1, a->y, a->z, s);
        launch_adj_update(n, a->z, a->nu, a->red, sl, s);
        HIPCHK(hipMemcpyAsync(&a->h_red[sl], &a->red[sl], 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              s));
        HIPCHK(hipEventRecord(a->qev[m % 2], s));
        if (m >= 2) {
            HIPCHK(hipEventSynchronize(a->qev[(m - 1) % 2]));
            if (converged(m - 1)) {
                conv = true;
                break;
            }
        }
    }
    if (!conv) {
        m = std::min(m, a->fp_max);
        HIPCHK(hipEventSynchronize(a->qev[m % 2]));
        conv = converged(m);
        if (!conv) a->st.adj_unconverged++;
    }
    a->st.adj_iters_max = std::max(a->st.adj_iters_max, m);
    a->st.adj_iters_mean += m;
    // the adjoints of un, vn and of the points the loop leaves at un, vn
    launch_copy_set(ld, 0, nx + 1, 0, ny + 1, fsrc, fdst, 2, s);
    launch_adj_zero(2 * (long)c->nf, a->nub, s);   // nub, nvb are adjacent
    adj_gtrans(a, us, vs, un, vn, 1);
    launch_adj_combine(1, nx, ny, ld, a->tu, a->nub, a->ub, s);
    launch_adj_combine(2, nx, ny, ld, a->tv, a->nvb, a->vb, s);
    return check_err(c);
}

// ================================================================ tangent-linear
// Forward-mode derivative of the same step (wd2_adj_tangent): the linear
// pieces run as the forward kernels on the tangent fields (the boundary
// fills with the tangent boundary values), the Poisson solve is a multigrid
// solve of the tangent right-hand side, and the q-l loop's tangent is the
// fixed point of dw <- dw + S2^-1 Pa S1^-1 (db - dS1 x1) at the converged
// iterate, with db, dS1 from the assembly on dual numbers (k_tangent.hip).
static int tl_alloc(wd2_adj *a) {
    if (a->tl_store) return WD2_OK;
    wd2_ctx *c = a->c;
    const size_t nf = c->nf, nch = chain_rows(c->nx, c->ny);
    HIPCHK(hipMalloc(&a->tl_store, (8 * nf + 10 * nch) * sizeof(Dual)));
    HIPCHK(hipMemsetAsync(a->tl_store, 0, (8 * nf + 10 * nch) * sizeof(Dual), c->s));
    Dual *f = a->tl_store;
    Dual **fp[8] = {&a->tUS, &a->tVS, &a->tUN, &a->tVN, &a->tC1, &a->tD1, &a->tC2, &a->tD2};
    for (int k = 0; k < 8; k++) *fp[k] = f + k * nf;
    Dual *q = f + 8 * nf;
    a->tch1 = ChainT<Dual>{q, q + nch, q + 2 * nch, q + 3 * nch, q + 4 * nch};
    q += 5 * nch;
    a->tch2 = ChainT<Dual>{q, q + nch, q + 2 * nch, q + 3 * nch, q + 4 * nch};
    // tangent boundary values: zero constants, a profile on every side for u
    // and v (offsets into tl_pval, laid out as bc_bar: [side][var][line]),
    // none for p
    const int mr = c->mgri * c->mgrj;
    HIPCHK(hipMalloc(&a->tl_bcv, 16 * (size_t)mr * sizeof(double)));
    HIPCHK(hipMemsetAsync(a->tl_bcv, 0, 16 * (size_t)mr * sizeof(double), c->s));
    HIPCHK(hipMalloc(&a->tl_pval, 8 * (size_t)a->W * sizeof(double)));
    HIPCHK(hipMemsetAsync(a->tl_pval, 0, 8 * (size_t)a->W * sizeof(double), c->s));
    std::vector<int> off(12 * (size_t)mr, -1);
    for (int side = WEST; side <= NORTH; side++)
        for (int var = VU; var <= VV; var++)
            off[(size_t)mr * ((side - 1) + 4 * (var - 1))] = ((side - 1) * 2 + (var - 1)) * a->W;
    HIPCHK(hipMalloc(&a->tl_poff, off.size() * sizeof(int)));
    HIPCHK(hipMemcpy(a->tl_poff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    return WD2_OK;
}

// One step forwards: the tangent state (ub, vb, pb) of the step's start to
// that of its end, given the forward step's start (un, vn) and converged
// iterate (us, vs); Rt carries the tangent boundary values.
static int tl_step(wd2_adj *a, const Regions &Rt, const double *un, const double *vn, const double *us,
                   const double *vs) {
    wd2_ctx *c = a->c;
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    const wd2_params &P = a->P;
    const long n1 = (long)nx * (ny - 1), n2 = (long)(nx - 1) * ny, n = n1 + n2;
    const Chains cs{n1, n1, n2};
    const long nf = (long)c->nf;
    double *du = a->ub, *dv = a->vb, *dp = a->pb;
    double *dun = a->nub, *dvn = a->nvb, *dus = a->tu, *dvs = a->tv;
    {
        const double *src[4] = {du, dv, du, dv};
        double *dst[4] = {dun, dvn, dus, dvs};
        launch_copy_set(ld, 0, nx + 1, 0, ny + 1, src, dst, 4, s);
    }
    // the chains' values at the converged iterate: S1, S2 and x1 = S1^-1 b
    MomTemps t = c->T, t2 = c->T;
    t2.cnvn = c->T.cnvn2;
    t2.difn = c->T.difn2;
    auto shift = [](const Chain &q, long o) { return Chain{q.a + o, q.d + o, q.c + o, q.b + o, q.x + o}; };
    auto dshift = [](const ChainT<Dual> &q, long o) {
        return ChainT<Dual>{q.a + o, q.d + o, q.c + o, q.b + o, q.x + o};
    };
    launch_mom_n<double>(nx, ny, ld, 1, c->M, un, vn, t.cnvn, t.difn, s);
    launch_mom_n<double>(nx, ny, ld, 2, c->M, un, vn, t2.cnvn, t2.difn, s);
    launch_mom_asm_pair<double>(c->R, ld, P.dk, P.re, P.fr, c->M, t, t2, us, vs, un, vn, c->d, c->dn, c->ch, c->ch2,
                                c->zm, shift(c->ch, n1), shift(c->ch2, n1), c->zm + n1, s, &c->sc->err, 0);
    launch_trid<double>(c->tw, n, cs, c->ch.a, c->ch.d, c->ch.c, c->ch.b, c->ch.x, s);
    // the time-level-n terms on dual numbers
    launch_tl_pack(nf, un, dun, a->tUN, s);
    launch_tl_pack(nf, vn, dvn, a->tVN, s);
    launch_mom_n_tl(nx, ny, ld, 1, c->M, a->tUN, a->tVN, a->tC1, a->tD1, s);
    launch_mom_n_tl(nx, ny, ld, 2, c->M, a->tUN, a->tVN, a->tC2, a->tD2, s);
    MomTempsT<Dual> tt1{}, tt2{};
    tt1.cnvn = a->tC1;
    tt1.difn = a->tD1;
    tt2.cnvn = a->tC2;
    tt2.difn = a->tD2;
    int m = 1;
    bool conv = false;
    for (; m <= a->fp_max; m++) {
        launch_tl_pack(nf, us, dus, a->tUS, s);
        launch_tl_pack(nf, vs, dvs, a->tVS, s);
        launch_mom_asm_tl_pair(c->R, ld, P.dk, P.re, P.fr, c->M, tt1, tt2, a->tUS, a->tVS, a->tUN, a->tVN, a->tch1,
                               a->tch2, c->zm, dshift(a->tch1, n1), dshift(a->tch2, n1), c->zm + n1, s, &c->sc->err);
        launch_tl_rhs(n, cs, a->tch1, c->ch.x, a->r, s);
        launch_trid<double>(c->tw, n, cs, c->ch.a, c->ch.d, c->ch.c, a->r, a->y, s);
        launch_adj_zmask(n, c->zm, a->y, a->y, s);
        launch_trid<double>(c->tw, n, cs, c->ch2.a, c->ch2.d, c->ch2.c, a->y, a->z, s);
        launch_tl_scatter(nx, ny, ld, a->z, dus, dvs, s);
        double zmax = 0.0, umax = 0.0, vmax = 0.0;
        int st = adj_max2(a, n, a->z, nullptr, &zmax, nullptr);
        if (!st) st = adj_max2(a, nf, dus, dvs, &umax, &vmax);
        if (st) return st;
        if (zmax <= a->fp_tol * std::max(umax, vmax)) {
            conv = true;
            break;
        }
    }
    if (!conv) {
        m = a->fp_max;
        a->st.adj_unconverged++;
    }
    a->st.adj_iters_max = std::max(a->st.adj_iters_max, m);
    a->st.adj_iters_mean += m;
    // main.f: VelBoundCond, PresBoundCond, Ppe, PresBoundCond, Project,
    // VelBoundCond, PresBoundCond, then u = us and the step-end fills
    launch_vel_bc<double>(Rt, ld, dus, dvs, false, &c->sc->err, s);
    launch_pres_bc<double>(Rt, ld, dp, &c->sc->err, s);
    launch_divergence(nx, ny, ld, 1, c->M, dus, dvs, c->div, s);
    launch_ppe_rhs(nx, ny, ld, P.dk, c->div, nullptr, c->rhs, s);
    double bmax = 0.0, pmax = 0.0;
    int st = adj_max2(a, nf, c->rhs, dp, &bmax, &pmax);
    if (st) return st;
    int ncyc = 0;
    if (bmax > 0.0 || pmax > 0.0) {
        const int ml = c->mg_last;
        const double md = c->mg_div, mdm = c->mg_div_max;
        c->mg_last = a->mg_last;
        const double tol = a->ppe_tol * std::fabs(P.dk) * (bmax > 0.0 ? bmax : pmax);
        st = ppe_mg(c, 1, P.msorit, P.dk, tol, dp, &ncyc);
        a->mg_last = c->mg_last;
        const bool pconv = c->mg_div < tol;
        c->mg_last = ml;
        c->mg_div = md;
        c->mg_div_max = mdm;
        if (st) return st;
        if (!pconv) a->st.adj_unconverged++;
    }
    a->st.ppe_cycles_max = std::max(a->st.ppe_cycles_max, ncyc);
    a->st.ppe_cycles_mean += ncyc;
    launch_pres_bc<double>(Rt, ld, dp, &c->sc->err, s);
    launch_project(Rt, ld, P.dk, c->M, dp, dus, dvs, s);
    launch_vel_bc<double>(Rt, ld, dus, dvs, false, &c->sc->err, s);
    launch_pres_bc<double>(Rt, ld, dp, &c->sc->err, s);
    {
        const double *src[2] = {dus, dvs};
        double *dst[2] = {du, dv};
        launch_copy_set(ld, 0, nx + 1, 0, ny + 1, src, dst, 2, s);
    }
    launch_vel_bc<double>(Rt, ld, du, dv, false, &c->sc->err, s);
    launch_pres_bc<double>(Rt, ld, dp, &c->sc->err, s);
    return check_err(c);
}

// ================================================================ C ABI

static wd2_ctx *g_ctx = nullptr;   // drop-in calls share one context per grid size

static int dropin_ctx(int nx, int ny, int ld) {
    if (ld < nx + 2) return fail(WD2_ERR_ARG, "ld=%d < nx+2=%d", ld, nx + 2);
    if (g_ctx && g_ctx->nx == nx && g_ctx->ny == ny) return WD2_OK;
    if (g_ctx) {
        g_fallbacks += fallbacks(g_ctx);
        ctx_free(g_ctx);
        g_ctx = nullptr;
    }
    return ctx_alloc(&g_ctx, nx, ny);
}

// Metric arrays of the drop-in calls: compared with the content last sent and
// uploaded only when it differs (the driver's metrics never change, but the
// library cannot assume it).
static int up_metric(int k, const double *h, int ld) {
    wd2_ctx *c = g_ctx;
    if (!h) return WD2_OK;
    const size_t w = (size_t)c->nx + 2, rows = (size_t)c->ny + 2;
    if (c->met_shadow.size() < (size_t)M_COUNT) c->met_shadow.resize(M_COUNT);
    std::vector<double> &sh = c->met_shadow[k];
    if (sh.size() == w * rows) {
        bool same = true;
        for (size_t j = 0; j < rows && same; j++)
            same = std::memcmp(h + j * (size_t)ld, sh.data() + j * w, w * sizeof(double)) == 0;
        if (same) return WD2_OK;
    }
    sh.resize(w * rows);
    for (size_t j = 0; j < rows; j++) std::memcpy(sh.data() + j * w, h + j * (size_t)ld, w * sizeof(double));
    c->met_version++;
    graph_clear(c);
    c->M.box[k] = scan_box(h, ld, c->nx, c->ny);
    return up(c, h, ld, c->met[k]);
}

extern "C" {

const char *wd2_last_error(void) { return g_err.c_str(); }

int wd2_sor_sweeps_per_launch(void) { return g_tiled_sor ? sor_tiled_sweeps() : 1; }
int wd2_sor_mixed_sweeps_per_launch(void) { return sor_tiled_mx_sweeps(); }

int wd2_set_verbose(int enable) {
    const int old = g_verbose;
    g_verbose = enable ? 1 : 0;
    return old;
}

int wd2_set_window_chains(int enable) {
    const int old = g_window_chains;
    g_window_chains = enable ? 1 : 0;
    if (g_window_chains != old) g_graph_epoch++;
    return old;
}

int wd2_set_uniform_coefficients(int enable) {
    const int old = g_uniform;
    g_uniform = enable ? 1 : 0;
    if (g_uniform != old) g_graph_epoch++;
    return old;
}

int wd2_set_lex_rows(int enable) {
    const int old = g_lex_rows;
    g_lex_rows = enable ? 1 : 0;
    return old;
}

int wd2_set_tiled_sor(int enable) {
    const int old = g_tiled_sor;
    g_tiled_sor = enable ? 1 : 0;
    return old;
}

int wd2_trid_fallbacks(void) { return g_fallbacks + (g_ctx ? fallbacks(g_ctx) : 0) + noalloc_trid_fallbacks(); }

int wd2_vel_bound_cond(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                       const int *nMomBdTp, const double *dBCVal, double *u, double *v) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nullptr, nMomBdTp, dBCVal, nullptr, nullptr, nullptr)))
        return st;
    if ((st = up(c, u, ld, c->u)) || (st = up(c, v, ld, c->v))) return st;
    launch_vel_bc<double>(c->R, c->ld, c->u, c->v, false, &c->sc->err, c->s);
    if ((st = check_err(c))) return st;
    if ((st = down(c, c->u, u, ld)) || (st = down(c, c->v, v, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_vel_outflow_bcs(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                        const int *nMomBdTp, const double *dBCVal, double *u, double *v) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nullptr, nMomBdTp, dBCVal, nullptr, nullptr, nullptr)))
        return st;
    if ((st = up(c, u, ld, c->u)) || (st = up(c, v, ld, c->v))) return st;
    launch_vel_bc<double>(c->R, c->ld, c->u, c->v, true, &c->sc->err, c->s);
    if ((st = check_err(c))) return st;
    if ((st = down(c, c->u, u, ld)) || (st = down(c, c->v, v, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_pres_bound_cond(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                        const int *nRegType, const int *nMomBdTp, const double *dBCVal, double *p) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, nMomBdTp, dBCVal, nullptr, nullptr, nullptr)))
        return st;
    if ((st = up(c, p, ld, c->p))) return st;
    launch_pres_bc<double>(c->R, c->ld, c->p, &c->sc->err, c->s);
    if ((st = check_err(c))) return st;
    if ((st = down(c, c->p, p, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_project(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                const int *nRegType, const int *nMomBdTp, double dk, const double *dju, const double *djv,
                const double *yeu, const double *xzv, const double *yzu, const double *xev, const double *p,
                double *u, double *v) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, nMomBdTp, nullptr, nullptr, nullptr, nullptr)))
        return st;
    if ((st = up_metric(M_DJU, dju, ld)) || (st = up_metric(M_DJV, djv, ld)) || (st = up_metric(M_YEU, yeu, ld)) ||
        (st = up_metric(M_XZV, xzv, ld)) || (st = up_metric(M_YZU, yzu, ld)) || (st = up_metric(M_XEV, xev, ld)))
        return st;
    if ((st = up(c, p, ld, c->p)) || (st = up(c, u, ld, c->u)) || (st = up(c, v, ld, c->v))) return st;
    launch_project(c->R, c->ld, dk, c->M, c->p, c->u, c->v, c->s);
    if ((st = down(c, c->u, u, ld)) || (st = down(c, c->v, v, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_divergence(int nx, int ny, int ld, int nloc, const double *xet, const double *yet, const double *xzi,
                   const double *yzi, const double *u, const double *v, double *div) {
    int st;
    if (nloc != 1 && nloc != 2) return fail(WD2_ERR_FLAG, "Wrong location flag passed to Divergence: %d", nloc);
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    const int kx = nloc == 1 ? M_XEU : M_XEV, ky = nloc == 1 ? M_YEU : M_YEV;
    const int kzx = nloc == 1 ? M_XZV : M_XZU, kzy = nloc == 1 ? M_YZV : M_YZU;
    if ((st = up_metric(kx, xet, ld)) || (st = up_metric(ky, yet, ld)) || (st = up_metric(kzx, xzi, ld)) ||
        (st = up_metric(kzy, yzi, ld)))
        return st;
    if ((st = up(c, u, ld, c->u)) || (st = up(c, v, ld, c->v)) || (st = up(c, div, ld, c->div))) return st;
    launch_divergence(nx, ny, c->ld, nloc, c->M, c->u, c->v, c->div, c->s);
    if ((st = down(c, c->div, div, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_ppe(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd, const int *nRegType,
            int lCartesGrid, int nPpeSolver, int msorit, int *nSorConv, double dk, double sortol, double sorrel,
            const double *rau, const double *rbu, const double *rbv, const double *rgv, const double *xeu,
            const double *yeu, const double *xzv, const double *yzv, const double *u, const double *v,
            double *p) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    std::vector<int> btp(4 * (size_t)mgri * mgrj, BM_INTERN);   // Ppe takes no boundary types
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, btp.data(), nullptr, nullptr, nullptr, nullptr)))
        return st;
    if ((st = up_metric(M_RAU, rau, ld)) || (st = up_metric(M_RBU, rbu, ld)) || (st = up_metric(M_RBV, rbv, ld)) ||
        (st = up_metric(M_RGV, rgv, ld)) || (st = up_metric(M_XEU, xeu, ld)) || (st = up_metric(M_YEU, yeu, ld)) ||
        (st = up_metric(M_XZV, xzv, ld)) || (st = up_metric(M_YZV, yzv, ld)))
        return st;
    if (c->met_version != c->uni_version) {
        scan_uniform(c, rau, rgv, ld);
        c->uni_version = c->met_version;
    }
    if ((st = up(c, u, ld, c->u)) || (st = up(c, v, ld, c->v)) || (st = up(c, p, ld, c->p))) return st;
    if ((st = ppe(c, lCartesGrid != 0, nPpeSolver, msorit, dk, sortol, sorrel, c->u, c->v, c->p, nSorConv)))
        return st;
    if ((st = down(c, c->p, p, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_x_momentum(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                   const int *nRegType, const int *nMomBdTp, double dk, double re, const double *dPRporos,
                   const double *dPRporc1, const double *dPRporc2, const double *rbn, const double *rgn,
                   const double *rac, const double *rbc, const double *dju, const double *xec, const double *yec,
                   const double *xzn, const double *yzn, const double *xeu, const double *yeu, const double *xzu,
                   const double *yzu, const double *us, const double *vs, const double *un, const double *vn,
                   double *dus) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, nMomBdTp, nullptr, dPRporos, dPRporc1,
                          dPRporc2)))
        return st;
    const int ks[] = {M_RBN, M_RGN, M_RAC, M_RBC, M_DJU, M_XEC, M_YEC, M_XZN, M_YZN, M_XEU, M_YEU, M_XZU, M_YZU};
    const double *hs[] = {rbn, rgn, rac, rbc, dju, xec, yec, xzn, yzn, xeu, yeu, xzu, yzu};
    for (int k = 0; k < 13; k++)
        if ((st = up_metric(ks[k], hs[k], ld))) return st;
    if ((st = up(c, us, ld, c->us)) || (st = up(c, vs, ld, c->vs)) || (st = up(c, un, ld, c->un)) ||
        (st = up(c, vn, ld, c->vn)) || (st = up(c, dus, ld, c->dus)))
        return st;
    if ((st = momentum(c, 1, dk, re, 1.0, c->us, c->vs, c->un, c->vn, c->dus))) return st;
    if ((st = check_err(c))) return st;
    if ((st = down(c, c->dus, dus, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_y_momentum(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                   const int *nRegType, const int *nMomBdTp, double dk, double re, double fr,
                   const double *dPRporos, const double *dPRporc1, const double *dPRporc2, const double *ran,
                   const double *rbn, const double *rbc, const double *rgc, const double *djv, const double *xen,
                   const double *yen, const double *xzc, const double *yzc, const double *xev, const double *yev,
                   const double *xzv, const double *yzv, const double *d, const double *dn, const double *us,
                   const double *vs, const double *un, const double *vn, double *dvs) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, nMomBdTp, nullptr, dPRporos, dPRporc1,
                          dPRporc2)))
        return st;
    const int ks[] = {M_RAN, M_RBN, M_RBC, M_RGC, M_DJV, M_XEN, M_YEN, M_XZC, M_YZC, M_XEV, M_YEV, M_XZV, M_YZV};
    const double *hs[] = {ran, rbn, rbc, rgc, djv, xen, yen, xzc, yzc, xev, yev, xzv, yzv};
    for (int k = 0; k < 13; k++)
        if ((st = up_metric(ks[k], hs[k], ld))) return st;
    if ((st = up(c, us, ld, c->us)) || (st = up(c, vs, ld, c->vs)) || (st = up(c, un, ld, c->un)) ||
        (st = up(c, vn, ld, c->vn)) || (st = up(c, d, ld, c->d)) || (st = up(c, dn, ld, c->dn)) ||
        (st = up(c, dvs, ld, c->dvs)))
        return st;
    if (d || dn) c->d_live = true;
    if ((st = momentum(c, 2, dk, re, fr, c->us, c->vs, c->un, c->vn, c->dvs))) return st;
    if ((st = check_err(c))) return st;
    if ((st = down(c, c->dvs, dvs, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_naux_momentum(int nx, int ny, int ld, int mgri, int mgrj, int mqiter, const int *nReg, const int *nRegBrd,
                      const int *nRegType, const int *nMomBdTp, double dk, double re, double fr, double qtol,
                      const double *dPRporos, const double *dPRporc1, const double *dPRporc2,
                      const double *dBCVal, const double *ran, const double *rbn, const double *rgn,
                      const double *rac, const double *rbc, const double *rgc, const double *dju, const double *djv,
                      const double *xec, const double *yec, const double *xzn, const double *yzn, const double *xen,
                      const double *yen, const double *xzc, const double *yzc, const double *xeu, const double *yeu,
                      const double *xzu, const double *yzu, const double *xev, const double *yev, const double *xzv,
                      const double *yzv, const double *d, const double *dn, const double *un, const double *vn,
                      double *us, double *vs, int *nQLiter) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, nMomBdTp, dBCVal, dPRporos, dPRporc1,
                          dPRporc2)))
        return st;
    const int ks[] = {M_RAN, M_RBN, M_RGN, M_RAC, M_RBC, M_RGC, M_DJU, M_DJV, M_XEC, M_YEC, M_XZN, M_YZN,
                      M_XEN, M_YEN, M_XZC, M_YZC, M_XEU, M_YEU, M_XZU, M_YZU, M_XEV, M_YEV, M_XZV, M_YZV};
    const double *hs[] = {ran, rbn, rgn, rac, rbc, rgc, dju, djv, xec, yec, xzn, yzn,
                          xen, yen, xzc, yzc, xeu, yeu, xzu, yzu, xev, yev, xzv, yzv};
    for (int k = 0; k < 24; k++)
        if ((st = up_metric(ks[k], hs[k], ld))) return st;
    if ((st = up(c, d, ld, c->d)) || (st = up(c, dn, ld, c->dn)) || (st = up(c, un, ld, c->un)) ||
        (st = up(c, vn, ld, c->vn)) || (st = up(c, us, ld, c->us)) || (st = up(c, vs, ld, c->vs)))
        return st;
    if (d || dn) c->d_live = true;
    // nAuxMomentum zeroes the increments on 0..nx+1 x 0..ny+1 each iteration
    launch_fill(c->ld, 0, nx + 1, 0, ny + 1, 0.0, c->dus, c->s);
    launch_fill(c->ld, 0, nx + 1, 0, ny + 1, 0.0, c->dvs, c->s);
    int nql = -1;
    if ((st = naux_momentum(c, mqiter, dk, re, fr, qtol, &nql))) return st;
    if (nQLiter) *nQLiter = nql;
    if ((st = down(c, c->us, us, ld)) || (st = down(c, c->vs, vs, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_alt_trid_lu(int n, const double *a, double *b) {
    if (n < 2) return fail(WD2_ERR_ARG, "AltTridLU needs n >= 2");
    int st;
    const int side = (int)std::ceil(std::sqrt((double)n)) + 8;
    if ((st = dropin_ctx(side, side, side + 2))) return st;
    wd2_ctx *c = g_ctx;
    if ((size_t)n > (size_t)c->nx * c->ny) return fail(WD2_ERR_ARG, "chain too long");
    std::vector<double> ad(n), dd(n), cd(n);
    for (int k = 0; k < n; k++) {
        ad[k] = a[3 * k];
        dd[k] = a[3 * k + 1];
        cd[k] = a[3 * k + 2];
    }
    HIPCHK(hipMemcpyAsync(c->ch.a, ad.data(), n * 8, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->ch.d, dd.data(), n * 8, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->ch.c, cd.data(), n * 8, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->ch.b, b, n * 8, hipMemcpyHostToDevice, c->s));
    launch_trid(c->tw, n, n, c->ch.a, c->ch.d, c->ch.c, c->ch.b, c->ch.x, c->s);
    HIPCHK(hipMemcpyAsync(b, c->ch.x, n * 8, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_trid_f32(long n, long n1, long L1, long L2, const float *a, const float *d, const float *cc, const float *b,
                 float *x, int window) {
    if (n < 1 || n1 < 0 || n1 > n || L1 < 1 || L2 < 1 || (n1 > 0 && n1 % L1) || (n - n1) % L2)
        return fail(WD2_ERR_ARG, "wd2_trid_f32: chains of %ld and %ld rows do not tile %ld and %ld rows", L1, L2, n1,
                    n - n1);
    if (window != 0 && window != 1 && !trid_window_shape_ok(window))
        return fail(WD2_ERR_ARG, "wd2_trid_f32: window %d (0, 1, or a window shape of 128, 256 or 512 lanes)", window);
    int st;
    const int side = (int)std::ceil(std::sqrt((double)n)) + 8;
    if ((st = dropin_ctx(side, side, side + 2))) return st;
    wd2_ctx *c = g_ctx;
    if ((size_t)n > (size_t)c->nx * c->ny) return fail(WD2_ERR_ARG, "chain too long");
    const ChainT<float> &ch = c->chf;
    HIPCHK(hipMemcpyAsync(ch.a, a, n * 4, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(ch.d, d, n * 4, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(ch.c, cc, n * 4, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(ch.b, b, n * 4, hipMemcpyHostToDevice, c->s));
    const Chains cs{n1, n1 > 0 ? L1 : L2, L2};
    if (window)
        launch_trid_window<float>(c->twf, n, cs, ch.a, ch.d, ch.c, ch.b, ch.x, c->s, nullptr, nullptr, nullptr,
                                  window == 1 ? 256 : window);
    else
        launch_trid<float>(c->twf, n, cs, ch.a, ch.d, ch.c, ch.b, ch.x, c->s);
    HIPCHK(hipMemcpyAsync(x, ch.x, n * 4, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

static int norm_common(int nx, int ny, int ld, const double *a, const double *b, double *out) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = up(c, a, ld, c->un)) || (st = up(c, b, ld, c->u))) return st;
    launch_max_norm(nx, ny, c->ld, c->un, b ? c->u : nullptr, c->partials, c->sc, 2, c->s);
    HIPCHK(hipMemcpyAsync(out, &c->sc->val[2], sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_dmax_norm(int nx, int ny, int ld, const double *u, double *out) { return norm_common(nx, ny, ld, u, nullptr, out); }
int wd2_diff_max_norm(int nx, int ny, int ld, const double *un, const double *u, double *out) {
    return norm_common(nx, ny, ld, un, u, out);
}

// ---------------------------------------------------------------- thermal
int wd2_temp_bound_cond(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                        const int *nTRgType, const int *nTemBdTp, const double *dTRgVal, const double *dBCVal,
                        double *t) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nullptr, nullptr, dBCVal, nullptr, nullptr, nullptr)) ||
        (st = set_thermal(c, nTRgType, nTemBdTp, dTRgVal, nullptr)))
        return st;
    if ((st = up(c, t, ld, c->t))) return st;
    launch_temp_bc(c->R, c->TH, c->ld, c->t, &c->sc->err, c->s);
    if ((st = check_err(c))) return st;
    return down(c, c->t, t, ld);
}

int wd2_therm_energy(int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                     const int *nTRgType, const int *nTemBdTp, double dk, double pe, const double *dTRgVal,
                     const double *dHGSTval, const double *dBCVal, const double *rau, const double *rbu,
                     const double *rbv, const double *rgv, const double *djc, const double *xeu, const double *yeu,
                     const double *xzv, const double *yzv, const double *xec, const double *yec, const double *xzc,
                     const double *yzc, const double *un, const double *vn, const double *u, const double *v,
                     const double *tn, double *t) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nullptr, nullptr, dBCVal, nullptr, nullptr, nullptr)) ||
        (st = set_thermal(c, nTRgType, nTemBdTp, dTRgVal, dHGSTval)))
        return st;
    const int ks[] = {M_RAU, M_RBU, M_RBV, M_RGV, M_DJC, M_XEU, M_YEU, M_XZV, M_YZV, M_XEC, M_YEC, M_XZC, M_YZC};
    const double *hs[] = {rau, rbu, rbv, rgv, djc, xeu, yeu, xzv, yzv, xec, yec, xzc, yzc};
    for (int k = 0; k < 13; k++)
        if ((st = up_metric(ks[k], hs[k], ld))) return st;
    if ((st = up(c, un, ld, c->un)) || (st = up(c, vn, ld, c->vn)) || (st = up(c, u, ld, c->us)) ||
        (st = up(c, v, ld, c->vs)) || (st = up(c, tn, ld, c->tn)) || (st = up(c, t, ld, c->t)))
        return st;
    launch_therm_energy(c->R, c->TH, c->ld, dk, pe, c->M, c->un, c->vn, c->us, c->vs, c->tn, c->t, c->tt, c->ch,
                        c->tw, &c->sc->err, c->s);
    if ((st = check_err(c))) return st;
    return down(c, c->t, t, ld);
}

int wd2_eq_state(int nx, int ny, int ld, double uref, double densref, double tmax, double tref, double rconst,
                 const double *p, const double *t, double *den) {
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    // den is written on the interior only: its other entries come back as sent
    if ((st = up(c, p, ld, c->p)) || (st = up(c, t, ld, c->t)) || (st = up(c, den, ld, c->d))) return st;
    c->d_live = true;
    launch_eq_state(nx, ny, c->ld, uref, densref, tmax, tref, rconst, c->p, c->t, c->d, c->s);
    return down(c, c->d, den, ld);
}

int wd2_filter(int nx, int ny, int ld, int mgri, int mgrj, int ncomp, const int *nReg, const int *nRegBrd,
               const int *nRegType, const int *nMomBdTp, const int *nTRgType, double fp, double *qu) {
    if (ncomp != VU && ncomp != VV && ncomp != VT) return fail(WD2_ERR_FLAG, "Wrong ncomp flag passed to Filter");
    int st;
    if ((st = dropin_ctx(nx, ny, ld))) return st;
    wd2_ctx *c = g_ctx;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, nMomBdTp, nullptr, nullptr, nullptr, nullptr)) ||
        (st = set_thermal(c, nTRgType, nullptr, nullptr, nullptr)))
        return st;
    if ((st = up(c, qu, ld, c->t))) return st;
    launch_filter(c->R, c->TH, c->ld, ncomp, fp, c->t, c->qh, c->s);
    return down(c, c->t, qu, ld);
}

// ------------------------------------------------------------ context API
int wd2_ctx_create(wd2_ctx **out, int nx, int ny, int ld, int mgri, int mgrj, const int *nReg, const int *nRegBrd,
                   const int *nRegType, const int *nMomBdTp, const double *dBCVal, const double *dPRporos,
                   const double *dPRporc1, const double *dPRporc2, const double *const *metrics) {
    if (!out || !metrics) return fail(WD2_ERR_ARG, "null argument");
    if (ld < nx + 2) return fail(WD2_ERR_ARG, "ld=%d < nx+2=%d", ld, nx + 2);
    wd2_ctx *c = nullptr;
    int st = ctx_alloc(&c, nx, ny);
    if (st) return st;
    if ((st = set_regions(c, mgri, mgrj, nReg, nRegBrd, nRegType, nMomBdTp, dBCVal, dPRporos, dPRporc1, dPRporc2))) {
        ctx_free(c);
        return st;
    }
    if ((st = set_thermal(c, nullptr, nullptr, nullptr, nullptr))) {
        ctx_free(c);
        return st;
    }
    for (int k = 0; k < M_COUNT; k++)
        if ((st = up(c, metrics[k], ld, c->met[k]))) {
            ctx_free(c);
            return st;
        }
    if (hipStreamSynchronize(c->s) != hipSuccess) {
        ctx_free(c);
        return fail(WD2_ERR_HIP, "metric upload failed");
    }
    scan_uniform(c, metrics[M_RAU], metrics[M_RGV], ld);
    for (int k = 0; k < M_COUNT; k++) c->M.box[k] = scan_box(metrics[k], ld, nx, ny);
    *out = c;
    return WD2_OK;
}

int wd2_ctx_destroy(wd2_ctx *c) {
    if (c) g_fallbacks += fallbacks(c);
    ctx_free(c);
    return WD2_OK;
}

int wd2_ctx_set_fields(wd2_ctx *c, int ld, const double *u, const double *v, const double *p) {
    int st;
    if ((st = up(c, u, ld, c->u)) || (st = up(c, v, ld, c->v)) || (st = up(c, p, ld, c->p))) return st;
    c->p_hist = c->x_run = 0;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_ctx_get_fields(wd2_ctx *c, int ld, double *u, double *v, double *p) {
    int st;
    if (u && (st = down(c, c->u, u, ld))) return st;
    if (v && (st = down(c, c->v, v, ld))) return st;
    if (p && (st = down(c, c->p, p, ld))) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_ctx_cold_start(wd2_ctx *c, const wd2_params *P, int *nSorConv) {
    int n = 0;
    c->p_hist = c->x_run = 0;
    int st = cold_start(c, P, &n);
    if (nSorConv) *nSorConv = n;
    return st;
}

static int ctx_steps(wd2_ctx *c, const wd2_params *P, int nsteps, int *ql_iter, int *ppe_iter, double *dif,
                     int ndif) {
    double d4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < nsteps; k++) {
        int nql = 0, nsor = 0;
        int st;
        if (ndif == 3 && graph_ok(c, P)) {
            if ((st = graph_begin(c, P)) || (st = graph_end(c, P, &nql, &nsor, d4))) return st;
        } else if ((st = one_step(c, P, &nql, &nsor, dif ? d4 : nullptr))) {
            return st;
        }
        if (dif) {
            HIPCHK(hipStreamSynchronize(c->s));   // d4 is written by the step's last copy
            for (int q = 0; q < ndif; q++) dif[ndif * k + q] = d4[q];
        }
        if (ql_iter) ql_iter[k] = nql;
        if (ppe_iter) ppe_iter[k] = nsor;
    }
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_ctx_step(wd2_ctx *c, const wd2_params *P, int nsteps, int *ql_iter, int *ppe_iter, double *dif) {
    return ctx_steps(c, P, nsteps, ql_iter, ppe_iter, dif, 3);
}

int wd2_ctx_step4(wd2_ctx *c, const wd2_params *P, int nsteps, int *ql_iter, int *ppe_iter, double *dif) {
    return ctx_steps(c, P, nsteps, ql_iter, ppe_iter, dif, 4);
}

int wd2_ctx_set_options(wd2_ctx *c, const wd2_step_options *opt) {
    if (!c || !opt) return fail(WD2_ERR_ARG, "wd2_ctx_set_options: null argument");
    if (opt->nthermen == 1 && !(opt->pe > 0.0))
        return fail(WD2_ERR_ARG, "wd2_ctx_set_options: the energy equation needs a Peclet number > 0");
    c->opt = *opt;
    if (opt->neqstate == 1) c->d_live = true;   // EqState writes d
    graph_clear(c);
    return WD2_OK;
}

int wd2_ctx_set_thermal(wd2_ctx *c, const int *nTRgType, const int *nTemBdTp, const double *dTRgVal,
                        const double *dHGSTval) {
    if (!c) return fail(WD2_ERR_ARG, "wd2_ctx_set_thermal: null context");
    graph_clear(c);
    int st = set_thermal(c, nTRgType, nTemBdTp, dTRgVal, dHGSTval);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_ctx_set_thermal_fields(wd2_ctx *c, int ld, const double *t, const double *d) {
    int st;
    if ((st = up(c, t, ld, c->t)) || (st = up(c, d, ld, c->d))) return st;
    if (d) c->d_live = true;
    HIPCHK(hipStreamSynchronize(c->s));
    return WD2_OK;
}

int wd2_ctx_get_thermal_fields(wd2_ctx *c, int ld, double *t, double *d) {
    int st;
    if (t && (st = down(c, c->t, t, ld))) return st;
    if (d && (st = down(c, c->d, d, ld))) return st;
    return WD2_OK;
}

int wd2_ctx_device_fields(wd2_ctx *c, double **u, double **v, double **p, int *ld) {
    if (u) *u = c->u;
    if (v) *v = c->v;
    if (p) *p = c->p;
    if (ld) *ld = c->ld;
    return WD2_OK;
}

void *wd2_ctx_stream(wd2_ctx *c) { return (void *)c->s; }

int wd2_ctx_set_precision(wd2_ctx *c, int precision) {
    if (!c) return fail(WD2_ERR_ARG, "wd2_ctx_set_precision: null context");
    if (precision != WD2_PREC_FP64 && precision != WD2_PREC_MIXED)
        return fail(WD2_ERR_ARG, "wd2_ctx_set_precision: unknown precision %d", precision);
    if (precision != c->precision) {
        graph_clear(c);
        // The FP32 momentum views alias the FP64 temporaries and chains.  Entries
        // the momentum kernels never write are read as 0 (as in the reference),
        // so clear the whole block when the arithmetic changes.
        const size_t nchain = chain_rows(c->nx, c->ny);
        HIPCHK(hipMemsetAsync(c->T.cj1, 0, (15 * c->nf + 10 * nchain) * sizeof(double), c->s));
        HIPCHK(hipStreamSynchronize(c->s));
    }
    c->precision = precision;
    return WD2_OK;
}

int wd2_ctx_set_bc_profile(wd2_ctx *c, int ireg, int jreg, int side, int var, const double *values, int n) {
    if (!c) return fail(WD2_ERR_ARG, "wd2_ctx_set_bc_profile: null context");
    if (ireg < 1 || ireg > c->h_nreg[0] || jreg < 1 || jreg > c->h_nreg[1])
        return fail(WD2_ERR_ARG, "wd2_ctx_set_bc_profile: region %d,%d outside the %d x %d regions", ireg, jreg,
                    c->h_nreg[0], c->h_nreg[1]);
    if (side < WEST || side > NORTH || var < VU || var > VP)
        return fail(WD2_ERR_ARG, "wd2_ctx_set_bc_profile: bad side %d or variable %d", side, var);
    const int want = side <= EAST ? c->ny + 2 : c->nx + 2;
    const int key = (ireg - 1) + c->mgri * ((jreg - 1) + c->mgrj * ((side - 1) + 4 * (var - 1)));
    if (values && n != want)
        return fail(WD2_ERR_ARG, "wd2_ctx_set_bc_profile: side %d needs %d values (one per grid line), got %d", side,
                    want, n);
    // kernels in flight may still read the current tables; cached step graphs
    // hold the table pointers
    HIPCHK(hipStreamSynchronize(c->s));
    graph_clear(c);
    if (values)
        c->h_prof[key].assign(values, values + n);
    else
        c->h_prof.erase(key);
    if (c->h_prof.empty()) {
        c->R.poff = nullptr;
        c->R.pval = nullptr;
        return WD2_OK;
    }
    const size_t n2 = (size_t)c->mgri * c->mgrj;
    std::vector<int> off(12 * n2, -1);
    std::vector<double> val;
    for (const auto &kv : c->h_prof) {
        off[kv.first] = (int)val.size();
        val.insert(val.end(), kv.second.begin(), kv.second.end());
    }
    if (!c->d_poff) HIPCHK(hipMalloc(&c->d_poff, 12 * n2 * sizeof(int)));
    if (val.size() > c->pval_cap) {
        (void)hipFree(c->d_pval);
        c->d_pval = nullptr;
        c->pval_cap = 0;
        HIPCHK(hipMalloc(&c->d_pval, val.size() * sizeof(double)));
        c->pval_cap = val.size();
    }
    HIPCHK(hipMemcpyAsync(c->d_poff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->d_pval, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->R.poff = c->d_poff;
    c->R.pval = c->d_pval;
    return WD2_OK;
}

int wd2_ctx_mg_set_tune(wd2_ctx *c, int nt0, int stage, int bside) {
    if ((nt0 != 256 && nt0 != 512) || (stage != 0 && stage != 1) || (bside != 32 && bside != 64))
        return fail(WD2_ERR_ARG, "multigrid shape: nt0 256|512, stage 0|1, bside 32|64");
    int st = mg_setup(c);
    if (st) return st;
    graph_clear(c);
    const MgTune old = c->mg.tune;
    c->mg.tune = MgTune{nt0, stage, bside};
    if ((st = mg_retile(c->mg))) {
        c->mg.tune = old;
        (void)mg_retile(c->mg);
        return st;
    }
    return WD2_OK;
}

int wd2_ctx_mg_get_tune(wd2_ctx *c, int *nt0, int *stage, int *bside) {
    int st = mg_setup(c);
    if (st) return st;
    if (nt0) *nt0 = c->mg.tune.nt0;
    if (stage) *stage = c->mg.tune.stage;
    if (bside) *bside = c->mg.tune.bside;
    return WD2_OK;
}

int wd2_ctx_mg_cycle_time(wd2_ctx *c, double dk, int cycles, double *us) {
    if (cycles < 1) return fail(WD2_ERR_ARG, "cycle timing needs cycles >= 1");
    int st = mg_prepare(c);
    if (st) return st;
    return mg_cycle_time(c, cycles, dk, us);
}

int wd2_ctx_mg_autotune(wd2_ctx *c, double dk, int cycles, double *best_us) {
    if (cycles < 1) return fail(WD2_ERR_ARG, "autotune needs cycles >= 1");
    int st = mg_prepare(c);
    if (st) return st;
    graph_clear(c);
    MgTune best = c->mg.tune;
    double tbest = 0.0;
    bool have = false;
    for (int nt0 : {256, 512})
        for (int stage : {0, 1})
            for (int bside : {32, 64}) {
                c->mg.tune = MgTune{nt0, stage, bside};
                if (mg_retile(c->mg)) continue;
                double us = 0.0;
                (void)mg_cycle_time(c, 1, dk, &us);   // warm
                if ((st = mg_cycle_time(c, cycles, dk, &us))) return st;
                if (!have || us < tbest) {
                    tbest = us;
                    best = c->mg.tune;
                    have = true;
                }
            }
    c->mg.tune = best;
    if ((st = mg_retile(c->mg))) return st;
    if (best_us) *best_us = tbest;
    return WD2_OK;
}

int wd2_ctx_mg_kernel_time(wd2_ctx *c, int which, int reps, double *avg_us) {
    if (!c->mg_store || c->mg_prec < 0) return fail(WD2_ERR_ARG, "no multigrid solve has run on this context");
    if (which < 2 || which > 4 || reps < 1) return fail(WD2_ERR_ARG, "mg kernel timing: which %d reps %d", which, reps);
    const int nx = c->nx, ny = c->ny, ld = c->ld;
    hipStream_t s = c->s;
    hipEvent_t e[2];
    HIPCHK(hipEventCreate(&e[0]));
    HIPCHK(hipEventCreate(&e[1]));
    const double *uni = (g_uniform && c->uni_ok) ? c->uni : nullptr;
    launch_copy(ld, 0, nx + 1, 0, ny + 1, c->p, c->pscr, s);   // the up kernel changes p
    if (c->mg_prec == WD2_PREC_MIXED)
        launch_mg_kernel_bench<float>(c->mg, which, reps, nx, ny, ld, c->rhs, c->p, c->M.rau, c->M.rgv, uni,
                                      c->mg_part, c->sc, s, e[0], e[1]);
    else
        launch_mg_kernel_bench<double>(c->mg, which, reps, nx, ny, ld, c->rhs, c->p, c->M.rau, c->M.rgv, uni,
                                       c->mg_part, c->sc, s, e[0], e[1]);
    launch_copy(ld, 0, nx + 1, 0, ny + 1, c->pscr, c->p, s);
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e[0], e[1]));
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    if (avg_us) *avg_us = 1e3 * ms / reps;
    return WD2_OK;
}

int wd2_ctx_trid_kernel_time(wd2_ctx *c, int which, int reps, double *avg_us) {
    if (which < 1 || which > 3 || reps < 1)
        return fail(WD2_ERR_ARG, "chain solver timing: which %d (1 forward, 2 backward, 3 window), reps %d", which,
                    reps);
    if (which == 3 && c->precision != WD2_PREC_MIXED)
        return fail(WD2_ERR_ARG, "chain solver timing: the window solver serves the mixed path only");
    if (c->any_porous) return fail(WD2_ERR_UNSUPPORTED, "chain solver timing: the two-chain batches only");
    const long n1 = (long)c->nx * (c->ny - 1), n2 = (long)(c->nx - 1) * c->ny;
    const Chains cs{n1, n1, n2};
    hipEvent_t e[2];
    HIPCHK(hipEventCreate(&e[0]));
    HIPCHK(hipEventCreate(&e[1]));
    if (which == 3)
        launch_trid_window_bench<float>(c->twf, n1 + n2, cs, c->chf.a, c->chf.d, c->chf.c, c->chf.b, c->chf.x, reps,
                                        c->s, e[0], e[1], c->win_wt);
    else if (c->precision == WD2_PREC_MIXED)
        launch_trid_kernel_bench<float>(c->twf, n1 + n2, cs, c->chf.a, c->chf.d, c->chf.c, c->chf.b, c->chf.x, which,
                                        reps, c->s, e[0], e[1]);
    else
        launch_trid_kernel_bench<double>(c->tw, n1 + n2, cs, c->ch.a, c->ch.d, c->ch.c, c->ch.b, c->ch.x, which,
                                         reps, c->s, e[0], e[1]);
    HIPCHK(hipStreamSynchronize(c->s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e[0], e[1]));
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    if (avg_us) *avg_us = 1e3 * ms / reps;
    return check_err(c);
}

int wd2_ctx_set_mg_extrapolation(wd2_ctx *c, int order) {
    if (!c) return fail(WD2_ERR_ARG, "wd2_ctx_set_mg_extrapolation: null context");
    if (order < 0 || order > 2) return fail(WD2_ERR_ARG, "extrapolation order %d (0, 1 or 2)", order);
    const int old = c->mg_extrap;
    c->mg_extrap = order;
    return old;
}

int wd2_ctx_trid_set_window(wd2_ctx *c, int wt) {
    if (!c || !trid_window_shape_ok(wt)) return fail(WD2_ERR_ARG, "window shape %d (128, 256 or 512 lanes)", wt);
    if (wt != c->win_wt) graph_clear(c);
    c->win_wt = wt;
    return WD2_OK;
}

int wd2_ctx_trid_autotune(wd2_ctx *c, int reps, int *wt_out, double *us_out) {
    if (!c || reps < 1) return fail(WD2_ERR_ARG, "window autotune: reps %d", reps);
    if (c->precision != WD2_PREC_MIXED) return fail(WD2_ERR_ARG, "window autotune: the mixed path only");
    if (c->any_porous) return fail(WD2_ERR_UNSUPPORTED, "window autotune: the two-chain batches only");
    const long n1 = (long)c->nx * (c->ny - 1), n2 = (long)(c->nx - 1) * c->ny;
    const Chains cs{n1, n1, n2};
    hipEvent_t e[2];
    HIPCHK(hipEventCreate(&e[0]));
    HIPCHK(hipEventCreate(&e[1]));
    const int shapes[3] = {128, 256, 512};
    int best = c->win_wt;
    double tbest = 0.0;
    for (int q = 0; q < 3; q++) {
        // one launch first (code load, caches), then the timed ones
        launch_trid_window_bench<float>(c->twf, n1 + n2, cs, c->chf.a, c->chf.d, c->chf.c, c->chf.b, c->chf.x, 1,
                                        c->s, e[0], e[1], shapes[q]);
        launch_trid_window_bench<float>(c->twf, n1 + n2, cs, c->chf.a, c->chf.d, c->chf.c, c->chf.b, c->chf.x, reps,
                                        c->s, e[0], e[1], shapes[q]);
        HIPCHK(hipStreamSynchronize(c->s));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e[0], e[1]));
        const double us = 1e3 * ms / reps;
        if (us_out) us_out[q] = us;
        if (q == 0 || us < tbest) {
            tbest = us;
            best = shapes[q];
        }
    }
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    if (best != c->win_wt) graph_clear(c);
    c->win_wt = best;
    if (wt_out) *wt_out = best;
    return WD2_OK;
}

int wd2_ctx_set_graph(wd2_ctx *c, int enable) {
    c->graph_on = enable != 0;
    return WD2_OK;
}

int wd2_ctx_graph_stats(wd2_ctx *c, long *steps, long *redos) {
    if (steps) *steps = c->graph_steps;
    if (redos) *redos = c->graph_redos;
    return WD2_OK;
}

// ------------------------------------------------------------ ensembles
struct wd2_ens {
    std::vector<wd2_ctx *> m;
};

int wd2_ens_create(wd2_ens **out, int n, wd2_ctx *const *ctxs) {
    if (!out || n < 1 || !ctxs) return fail(WD2_ERR_ARG, "ensemble needs n >= 1 contexts");
    wd2_ens *e = new wd2_ens();
    e->m.assign(ctxs, ctxs + n);
    *out = e;
    return WD2_OK;
}

int wd2_ens_destroy(wd2_ens *e) {
    delete e;
    return WD2_OK;
}

// Every member's step is queued on its own stream before any is waited for,
// so the members' graphs run side by side on the GPU; a member whose setup
// has no graph path steps host-driven after the others are queued.
int wd2_ens_step(wd2_ens *e, const wd2_params *P, int nsteps, int *ql_iter, int *ppe_iter, double *dif) {
    const int n = (int)e->m.size();
    std::vector<char> g(n);
    for (int k = 0; k < nsteps; k++) {
        for (int i = 0; i < n; i++) {
            g[i] = graph_ok(e->m[i], &P[i]);
            if (g[i]) {
                int st = graph_begin(e->m[i], &P[i]);
                if (st) return st;
            }
        }
        for (int i = 0; i < n; i++) {
            wd2_ctx *c = e->m[i];
            int nql = 0, nsor = 0;
            double d3[3] = {0.0, 0.0, 0.0};
            int st = g[i] ? graph_end(c, &P[i], &nql, &nsor, d3) : one_step(c, &P[i], &nql, &nsor, d3);
            if (st) return st;
            if (!g[i]) HIPCHK(hipStreamSynchronize(c->s));
            const size_t r = (size_t)k * n + i;
            if (ql_iter) ql_iter[r] = nql;
            if (ppe_iter) ppe_iter[r] = nsor;
            if (dif)
                for (int q = 0; q < 3; q++) dif[3 * r + q] = d3[q];
        }
    }
    return WD2_OK;
}

int wd2_ctx_ppe_divergence(wd2_ctx *c, double *last, double *max_since, int reset) {
    if (last) *last = c->mg_div;
    if (max_since) *max_since = c->mg_div_max;
    if (reset) c->mg_div_max = 0.0;
    return WD2_OK;
}

int wd2_ctx_kernel_timing(wd2_ctx *c, int enable, double *avg_us, long *launches) {
    if (enable) {
        if (enable < 1 || enable > 3) return fail(WD2_ERR_ARG, "kernel timing selector %d (1..3)", enable);
        c->kt_on = enable;
        c->kt_used = 0;
        c->kt_cnt.clear();
        return WD2_OK;
    }
    HIPCHK(hipStreamSynchronize(c->s));
    double tot = 0.0;
    long n = 0;
    for (size_t k = 0; k + 1 < c->kt_used; k += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->kt_ev[k], c->kt_ev[k + 1]));
        tot += ms;
        n += c->kt_cnt[k / 2];
    }
    c->kt_on = 0;
    c->kt_used = 0;
    c->kt_cnt.clear();
    if (avg_us) *avg_us = n ? 1e3 * tot / n : 0.0;
    if (launches) *launches = n;
    return WD2_OK;
}

}  // extern "C"

// ---------------------------------------------------------- adjoint C ABI
extern "C" {

int wd2_adj_create(wd2_adj **out, wd2_ctx *c, int every) {
    if (!out || !c) return fail(WD2_ERR_ARG, "null argument");
    if (every < 1) return fail(WD2_ERR_ARG, "checkpoint interval %d < 1", every);
    wd2_adj *a = new wd2_adj;
    a->c = c;
    a->every = every;
    int st = adj_alloc(a);
    if (st) {
        adj_free(a);
        return st;
    }
    *out = a;
    return WD2_OK;
}

int wd2_adj_destroy(wd2_adj *a) {
    adj_free(a);
    return WD2_OK;
}

int wd2_adj_set_tolerances(wd2_adj *a, double fp_tol, int max_iter, double ppe_tol) {
    if (!a) return fail(WD2_ERR_ARG, "null argument");
    if (fp_tol > 0.0) a->fp_tol = fp_tol;
    if (max_iter > 0) a->fp_max = max_iter;
    if (ppe_tol > 0.0) a->ppe_tol = ppe_tol;
    return WD2_OK;
}

int wd2_adj_forward(wd2_adj *a, const wd2_params *P, int nsteps, int *ql_iter, int *ppe_iter) {
    if (!a || !P) return fail(WD2_ERR_ARG, "null argument");
    if (nsteps < 1) return fail(WD2_ERR_ARG, "nsteps %d < 1", nsteps);
    wd2_ctx *c = a->c;
    int st = adj_check(c, P);
    if (st) return st;
    const ExtrapOff plain(c);
    const int nseg = (nsteps + a->every - 1) / a->every;
    if (nseg > a->nseg_cap) {
        (void)hipFree(a->ckpt);
        a->ckpt = nullptr;
        HIPCHK(hipMalloc(&a->ckpt, 3 * (size_t)nseg * c->nf * sizeof(double)));
        a->nseg_cap = nseg;
    }
    a->P = *P;
    a->have_fwd = false;
    a->st.fwd_unconverged = 0;
    HIPCHK(hipEventRecord(a->ev[0], c->s));
    for (int k = 0; k < nsteps; k++) {
        if (k % a->every == 0) {
            double *ck = a->ckpt + 3 * (size_t)(k / a->every) * c->nf;
            adj_copy3(c, c->u, c->v, c->p, ck, ck + c->nf, ck + 2 * c->nf);
        }
        int nql = 0, nsor = 0;
        if ((st = one_step(c, P, &nql, &nsor, nullptr))) return st;
        if ((P->qtol > 0.0 && nql < 0) || !(c->mg_div < P->sortol)) a->st.fwd_unconverged++;
        if (ql_iter) ql_iter[k] = nql;
        if (ppe_iter) ppe_iter[k] = nsor;
    }
    adj_copy3(c, c->u, c->v, c->p, a->fin, a->fin + c->nf, a->fin + 2 * c->nf);
    HIPCHK(hipEventRecord(a->ev[1], c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
    a->st.ms_forward = ms;
    a->nsteps = nsteps;
    a->have_fwd = true;
    return WD2_OK;
}

int wd2_adj_backward(wd2_adj *a, int ld_h, const double *u_bar, const double *v_bar, const double *p_bar,
                     double *u0_bar, double *v0_bar, double *p0_bar, double *bc_bar, wd2_adj_stats *stats) {
    if (!a || !u_bar || !v_bar) return fail(WD2_ERR_ARG, "null argument");
    if (!a->have_fwd) return fail(WD2_ERR_ARG, "wd2_adj_backward before wd2_adj_forward");
    wd2_ctx *c = a->c;
    if (ld_h < c->nx + 2) return fail(WD2_ERR_ARG, "ld=%d < nx+2=%d", ld_h, c->nx + 2);
    int st = adj_check(c, &a->P);
    if (st) return st;
    const ExtrapOff plain(c);
    hipStream_t s = c->s;
    const size_t nf = c->nf;
    launch_adj_zero(3 * (long)nf, a->ub, s);   // ub, vb, pb are adjacent
    if ((st = up(c, u_bar, ld_h, a->ub)) || (st = up(c, v_bar, ld_h, a->vb))) return st;
    if (p_bar && (st = up(c, p_bar, ld_h, a->pb))) return st;
    launch_adj_zero(8 * (long)a->W, a->bcb, s);
    launch_adj_zero((long)nf, a->lam, s);
    const int fu = a->st.fwd_unconverged;
    const double msf = a->st.ms_forward;
    a->st = wd2_adj_stats{};
    a->st.fwd_unconverged = fu;
    a->st.ms_forward = msf;
    const int nseg = (a->nsteps + a->every - 1) / a->every;
    for (int seg = nseg - 1; seg >= 0; seg--) {
        const int k0 = seg * a->every, k1 = std::min(a->nsteps, k0 + a->every);
        const double *ck = a->ckpt + 3 * (size_t)seg * nf;
        HIPCHK(hipEventRecord(a->ev[0], s));
        adj_copy3(c, ck, ck + nf, ck + 2 * nf, c->u, c->v, c->p);
        for (int k = k0; k < k1; k++) {
            double *tp = a->tape + 5 * (size_t)(k - k0) * nf;
            adj_copy3(c, c->u, c->v, c->p, tp, tp + nf, tp + 2 * nf);
            c->adj_us = tp + 3 * nf;
            c->adj_vs = tp + 4 * nf;
            int nql = 0, nsor = 0;
            st = one_step(c, &a->P, &nql, &nsor, nullptr);
            c->adj_us = c->adj_vs = nullptr;
            if (st) return st;
        }
        HIPCHK(hipEventRecord(a->ev[1], s));
        for (int k = k1 - 1; k >= k0; k--) {
            const double *tp = a->tape + 5 * (size_t)(k - k0) * nf;
            if ((st = adj_step(a, tp, tp + nf, tp + 3 * nf, tp + 4 * nf))) return st;
        }
        HIPCHK(hipEventRecord(a->ev[2], s));
        HIPCHK(hipStreamSynchronize(s));
        float m1 = 0.f, m2 = 0.f;
        HIPCHK(hipEventElapsedTime(&m1, a->ev[0], a->ev[1]));
        HIPCHK(hipEventElapsedTime(&m2, a->ev[1], a->ev[2]));
        a->st.ms_recompute += m1;
        a->st.ms_reverse += m2;
    }
    adj_copy3(c, a->fin, a->fin + nf, a->fin + 2 * nf, c->u, c->v, c->p);
    a->st.steps = a->nsteps;
    a->st.adj_iters_mean /= a->nsteps;
    a->st.ppe_cycles_mean /= a->nsteps;
    if (u0_bar && (st = down(c, a->ub, u0_bar, ld_h))) return st;
    if (v0_bar && (st = down(c, a->vb, v0_bar, ld_h))) return st;
    if (p0_bar && (st = down(c, a->pb, p0_bar, ld_h))) return st;
    if (bc_bar) HIPCHK(hipMemcpy(bc_bar, a->bcb, 8 * (size_t)a->W * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipStreamSynchronize(s));
    if (stats) *stats = a->st;
    return WD2_OK;
}

int wd2_adj_tangent(wd2_adj *a, int ld_h, const double *du0, const double *dv0, const double *dp0, const double *dbc,
                    double *duN, double *dvN, double *dpN, wd2_adj_stats *stats) {
    if (!a) return fail(WD2_ERR_ARG, "null argument");
    if (!a->have_fwd) return fail(WD2_ERR_ARG, "wd2_adj_tangent before wd2_adj_forward");
    wd2_ctx *c = a->c;
    if (ld_h < c->nx + 2) return fail(WD2_ERR_ARG, "ld=%d < nx+2=%d", ld_h, c->nx + 2);
    int st = adj_check(c, &a->P);
    if (st || (st = tl_alloc(a))) return st;
    const ExtrapOff plain(c);
    hipStream_t s = c->s;
    const size_t nf = c->nf;
    launch_adj_zero(3 * (long)nf, a->ub, s);   // the tangent state: ub, vb, pb are adjacent
    if (du0 && (st = up(c, du0, ld_h, a->ub))) return st;
    if (dv0 && (st = up(c, dv0, ld_h, a->vb))) return st;
    if (dp0 && (st = up(c, dp0, ld_h, a->pb))) return st;
    if (dbc)
        HIPCHK(hipMemcpyAsync(a->tl_pval, dbc, 8 * (size_t)a->W * sizeof(double), hipMemcpyHostToDevice, s));
    else
        launch_adj_zero(8 * (long)a->W, a->tl_pval, s);
    Regions Rt = c->R;
    Rt.bcv = a->tl_bcv;
    Rt.poff = a->tl_poff;
    Rt.pval = a->tl_pval;
    const int fu = a->st.fwd_unconverged;
    const double msf = a->st.ms_forward;
    a->st = wd2_adj_stats{};
    a->st.fwd_unconverged = fu;
    a->st.ms_forward = msf;
    // the forward run again from its first checkpoint, each step followed by
    // its tangent (the converged iterate kept in the tape's first slot)
    HIPCHK(hipEventRecord(a->ev[0], s));
    adj_copy3(c, a->ckpt, a->ckpt + nf, a->ckpt + 2 * nf, c->u, c->v, c->p);
    double *tus = a->tape + 3 * nf, *tvs = a->tape + 4 * nf;
    for (int k = 0; k < a->nsteps; k++) {
        c->adj_us = tus;
        c->adj_vs = tvs;
        int nql = 0, nsor = 0;
        st = one_step(c, &a->P, &nql, &nsor, nullptr);
        c->adj_us = c->adj_vs = nullptr;
        if (st || (st = tl_step(a, Rt, c->un, c->vn, tus, tvs))) return st;
    }
    adj_copy3(c, a->fin, a->fin + nf, a->fin + 2 * nf, c->u, c->v, c->p);
    HIPCHK(hipEventRecord(a->ev[1], s));
    if (duN && (st = down(c, a->ub, duN, ld_h))) return st;
    if (dvN && (st = down(c, a->vb, dvN, ld_h))) return st;
    if (dpN && (st = down(c, a->pb, dpN, ld_h))) return st;
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
    a->st.ms_reverse = ms;
    a->st.steps = a->nsteps;
    a->st.adj_iters_mean /= a->nsteps;
    a->st.ppe_cycles_mean /= a->nsteps;
    if (stats) *stats = a->st;
    return WD2_OK;
}

}  // extern "C"
