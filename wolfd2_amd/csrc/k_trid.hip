// k_trid.hip -- bit-exact parallel AltTridLU (momentum.f:1307-1339).
//
// The reference solves each tridiagonal system with one sequential Thomas
// sweep: forward  d'_k = d_k - a_k c'_{k-1},  c'_k = c_k / d'_k,
//                 b'_k = (b_k - a_k b'_{k-1}) / d'_k,
// backward        x_k = b'_k - c'_k x_{k+1},
// with the first row's super-diagonal divided by the SECOND row's original
// diagonal (c'_1 = c_1 / d_2) and no c' for the last row.  The momentum
// chains are ~nx*ny rows long, so a sequential sweep is far too slow on a GPU.
//
// Both recurrences contract (the systems are diagonally dominant), so two
// runs of a recurrence started from different states agree bit for bit after
// a few dozen rows.  Each thread therefore runs its own 8-row segment from a
// guessed state (speculation), then re-runs it from its predecessor's end
// state until its values coincide bit for bit with the speculative ones
// (fix-up).  From the coalescence point on the two runs are identical, so the
// result equals the sequential sweep exactly.  Fix-ups are iterated inside a
// workgroup until no segment end changes.  Across workgroups, the backward
// kernel repairs each block's forward head from its predecessor's end state
// before using it, and a one-workgroup kernel repairs each block's backward
// tail from its successor's first value.  A repair that never coalesces raises
// a flag and the same kernel recomputes the whole system sequentially, so the
// answer is exact in every case.  A thread loads and stores its segment as
// 16-byte vectors.
//
// A batch of independent chains is solved at once: rows [0, n1) are chains
// of length L1 and rows [n1, n) chains of length L2 (Chains).  No state
// crosses a chain start, so a batch gives each chain the bits it has alone.
#include <type_traits>

#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int TR_T = 256;   // threads per workgroup
constexpr int TR_S = 8;     // rows per thread
constexpr int TR_R = TR_T * TR_S;
constexpr int TR_P = TR_S + 1;   // padded LDS stride of a thread's rows

__device__ inline bool same(double x, double y) { return dbits(x) == dbits(y); }
__device__ inline bool same(float x, float y) { return __float_as_uint(x) == __float_as_uint(y); }

// One forward row.  kind: 0 interior, 1 chain start, 2 chain end.
template <typename T>
__device__ inline void fwd_row(int kind, T a, T d, T c, T b, T dnext, T cpv, T bpv, T &cpk, T &bpk) {
    if (kind == 1) {
        cpk = c / dnext;
        bpk = b / d;
    } else {
        const T dp = d - (a * cpv);
        cpk = (kind == 2) ? T(0) : c / dp;
        bpk = (b - a * bpv) / dp;
    }
}

// Position r of row k in its chain, and that chain's length Lk.
__device__ inline long chain_pos(const Chains &cs, long k, long &Lk) {
    if (k < cs.n1) {
        Lk = cs.L1;
        return k % cs.L1;
    }
    Lk = cs.L2;
    return (k - cs.n1) % cs.L2;
}
// (r, Lk) of row k + 1 from those of row k, and of row k - 1.
__device__ inline void chain_up(const Chains &cs, long k, long &r, long &Lk) {
    if (++r == Lk) {
        r = 0;
        if (k + 1 == cs.n1) Lk = cs.L2;
    }
}
__device__ inline void chain_down(const Chains &cs, long k, long &r, long &Lk) {
    if (r == 0) {
        if (k == cs.n1) Lk = cs.L1;
        r = Lk - 1;
    } else {
        r--;
    }
}
__device__ inline int kind_at(long r, long Lk) { return r == 0 ? 1 : (r == Lk - 1 ? 2 : 0); }
__device__ inline bool chain_end(long r, long Lk) { return r == Lk - 1 && r != 0; }
__device__ inline int row_kind(long k, const Chains &cs) {
    long Lk;
    const long r = chain_pos(cs, k, Lk);
    return kind_at(r, Lk);
}

// A thread's TR_S consecutive rows are read and written as 16-byte vectors
// (chain arrays start 128-byte aligned and a segment starts at a multiple of
// TR_S rows): a wavefront covers one contiguous span, every load in flight at
// once.  A partial last segment goes row by row.
template <typename T>
__device__ inline void seg_load(const T *__restrict__ G, long seg0, int len, T (&r)[TR_S]) {
    if (len == TR_S) {
        constexpr int V = 16 / sizeof(T);
        using V4 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
        const V4 *g = reinterpret_cast<const V4 *>(G + seg0);
#pragma unroll
        for (int q = 0; q < TR_S / V; q++) {
            const V4 v = g[q];
            const T *e = reinterpret_cast<const T *>(&v);
#pragma unroll
            for (int u = 0; u < V; u++) r[q * V + u] = e[u];
        }
    } else {
#pragma unroll
        for (int q = 0; q < TR_S; q++) r[q] = q < len ? G[seg0 + q] : T(0);
    }
}
template <typename T>
__device__ inline void seg_store(T *__restrict__ G, long seg0, int len, const T (&r)[TR_S]) {
    if (len == TR_S) {
        constexpr int V = 16 / sizeof(T);
        using V4 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
        V4 *g = reinterpret_cast<V4 *>(G + seg0);
#pragma unroll
        for (int q = 0; q < TR_S / V; q++) {
            V4 v;
            T *e = reinterpret_cast<T *>(&v);
#pragma unroll
            for (int u = 0; u < V; u++) e[u] = r[q * V + u];
            g[q] = v;
        }
    } else {
#pragma unroll
        for (int q = 0; q < TR_S; q++)
            if (q < len) G[seg0 + q] = r[q];
    }
}


// Kinds of the TR_S rows from seg0 on, two bits per row (one 64-bit modulo).
__device__ inline unsigned seg_kinds(long seg0, const Chains &cs) {
    long Lk;
    long r = chain_pos(cs, seg0, Lk);
    unsigned m = 0;
#pragma unroll
    for (int s = 0; s < TR_S; s++) {
        m |= (unsigned)kind_at(r, Lk) << (2 * s);
        chain_up(cs, seg0 + s, r, Lk);
    }
    return m;
}
#define SKIND(m, s) ((int)(((m) >> (2 * (s))) & 3u))

// Speculated states at block boundaries: the forward state entering row blk0
// and the backward value at row k1 are guessed by running the recurrence over
// the TR_W (forward) or TR_WB (backward) rows beyond the boundary from a zero
// state.  The recurrences contract, so the guess is usually the exact state
// bit for bit; it is always verified (and repaired if not) against the
// neighbour block's own result.  The rows a warm-up reads are first loaded by
// as many lanes at once into LDS (forward: w[0..3][q] = A, D, C, B of row
// k0 + q, and D of row k0 + q + 1 in w[4]), so the one lane that runs the
// recurrence waits for one load, not one per row.
constexpr int TR_W = 16;
template <typename T>
__device__ inline void warm_fwd_load(const T *__restrict__ A, const T *__restrict__ D, const T *__restrict__ Cc,
                                     const T *__restrict__ B, long blk0, T (*w)[TR_W]) {
    const long k0 = blk0 > TR_W ? blk0 - TR_W : 0;
    const int q = threadIdx.x;
    if (q < TR_W && k0 + q < blk0) {
        w[0][q] = A[k0 + q];
        w[1][q] = D[k0 + q];
        w[2][q] = Cc[k0 + q];
        w[3][q] = B[k0 + q];
        w[4][q] = D[k0 + q + 1];   // row k0 + q + 1 <= blk0 < n
    }
}
template <typename T>
__device__ inline void warm_fwd(long blk0, const Chains &cs, const T (*w)[TR_W], T &cpv, T &bpv) {
    const long k0 = blk0 > TR_W ? blk0 - TR_W : 0;
    long Lk;
    long r = chain_pos(cs, k0, Lk);
    T c = T(0), bb = T(0);
    for (int q = 0; k0 + q < blk0; chain_up(cs, k0 + q, r, Lk), q++) {
        const int kind = kind_at(r, Lk);
        T ck, bk;
        fwd_row<T>(kind, w[0][q], w[1][q], w[2][q], w[3][q], kind == 1 ? w[4][q] : T(1), c, bb, ck, bk);
        c = ck;
        bb = bk;
    }
    cpv = c;
    bpv = bb;
}
// Backward warm-up over the rows k1 .. k1+TR_WB-1 (cp in w[0], bp in w[1]).
constexpr int TR_WB = 16;
template <typename T>
__device__ inline void warm_bwd_load(const T *CP, const T *BP, long n, long k1, T (*w)[TR_WB]) {
    const int q = threadIdx.x;
    if (q < TR_WB && k1 + q < n) {
        w[0][q] = CP[k1 + q];
        w[1][q] = BP[k1 + q];
    }
}
template <typename T>
__device__ inline T warm_bwd(const T (*w)[TR_WB], long n, long k1, const Chains &cs) {
    const long k2 = min(n, k1 + TR_WB);
    long Lk;
    long r = chain_pos(cs, k2 - 1, Lk);
    T xw = T(0);
    for (long k = k2 - 1; k >= k1; chain_down(cs, k, r, Lk), k--) {
        const int q = (int)(k - k1);
        xw = chain_end(r, Lk) ? w[1][q] : w[1][q] - (w[0][q] * xw);
    }
    return xw;
}

// Bit queries on a workgroup's per-lane flags (TR_T bits in 64-bit words).
__device__ inline int lowest_bit_in(const unsigned long long *m, int lo, int hi) {   // -1 if none
    for (int w = lo >> 6; lo <= hi && w <= (hi >> 6); w++) {
        unsigned long long x = m[w];
        if (w == (lo >> 6)) x &= ~0ull << (lo & 63);
        if (w == (hi >> 6) && (hi & 63) != 63) x &= (1ull << ((hi & 63) + 1)) - 1;
        if (x) return (w << 6) + __ffsll((long long)x) - 1;
    }
    return -1;
}
__device__ inline int highest_clear_at_or_below(const unsigned long long *m, int hi) {   // -1 if none
    for (int w = hi >> 6; w >= 0; w--) {
        unsigned long long x = ~m[w];
        if (w == (hi >> 6) && (hi & 63) != 63) x &= (1ull << ((hi & 63) + 1)) - 1;
        if (x) return (w << 6) + 63 - __clzll((long long)x);
    }
    return -1;
}
__device__ inline int highest_bit_in(const unsigned long long *m, int lo, int hi) {   // -1 if none
    for (int w = hi >> 6; lo <= hi && w >= (lo >> 6); w--) {
        unsigned long long x = m[w];
        if (w == (lo >> 6)) x &= ~0ull << (lo & 63);
        if (w == (hi >> 6) && (hi & 63) != 63) x &= (1ull << ((hi & 63) + 1)) - 1;
        if (x) return (w << 6) + 63 - __clzll((long long)x);
    }
    return -1;
}
__device__ inline int lowest_clear_at_or_above(const unsigned long long *m, int lo, int n) {   // n if none
    for (int w = lo >> 6; w < n / 64; w++) {
        unsigned long long x = ~m[w];
        if (w == (lo >> 6)) x &= ~0ull << (lo & 63);
        if (x) return (w << 6) + __ffsll((long long)x) - 1;
    }
    return n;
}
__device__ inline bool bit_at(const unsigned long long *m, int i) { return (m[i >> 6] >> (i & 63)) & 1ull; }

// Rounding cycles.  Over a stretch of rows whose coefficients repeat with
// period two (identical rows included), the rounded forward map has fixed
// points or two-cycles, and runs from different states can settle on
// different ones and never coincide: the fix-up rounds would then carry the
// exact state through the stretch one segment per round.  Call a lane
// "constant-even" (CE) if its TR_S rows are interior rows whose coefficients
// equal those of the row two before (rows -2 and -1 being its predecessor's
// last two rows).  If lanes v+1 .. t-1 are CE and lane v's end state S is on
// the cycle of their row pair (f7(f6(S)) == S for lane v's last two rows,
// which are the pair of every row pair after them), each of those lanes maps
// S to itself (an even number of rows), so the exact state entering lane t
// is S.  After TR_ACC rounds a block still changing marks its CE lanes, and
// from then on a lane whose predecessor is CE takes its entering state from
// the first lane v of that CE run (or the lane before it) whose end state is
// on the cycle.  At convergence every lane's entering state then equals its
// predecessor's end state (by induction along the run), so the result is the
// sequential sweep's; only the number of rounds changes.
constexpr int TR_ACC = 3;

template <typename T>
__global__ __launch_bounds__(TR_T) void k_trid_fwd(long n, Chains cs, const T *__restrict__ A, const T *__restrict__ D,
                                                   const T *__restrict__ Cc, const T *__restrict__ B,
                                                   T *__restrict__ CP, T *__restrict__ BP, T *__restrict__ blk_c,
                                                   T *__restrict__ blk_b, const int *stop) {
    if (stop && *stop) return;
    __shared__ T sc[TR_T], sb[TR_T];
    __shared__ T sw[5][TR_W];
    __shared__ int changed;
    __shared__ unsigned long long ce_mask[TR_T / 64], cyc_mask[TR_T / 64];
    __shared__ T s67[8][TR_T];   // each lane's last two rows (A, D, C, B), for the CE test
    const int t = threadIdx.x;
    const long blk0 = (long)blockIdx.x * TR_R;
    const int nrow = (int)min((long)TR_R, n - blk0);
    const long seg0 = blk0 + (long)t * TR_S;
    const int len = max(0, min(TR_S, nrow - t * TR_S));
    T a[TR_S], d[TR_S], c[TR_S], b[TR_S], cp[TR_S], bp[TR_S];
    if (blk0 > 0) warm_fwd_load<T>(A, D, Cc, B, blk0, sw);
    seg_load(A, seg0, len, a);
    seg_load(D, seg0, len, d);
    seg_load(Cc, seg0, len, c);
    seg_load(B, seg0, len, b);
    __syncthreads();   // sw is complete
    const unsigned kinds = seg_kinds(seg0, cs);
    // the first-pivot quirk needs the original diagonal of the row after a chain start
#define DNEXT(s) (SKIND(kinds, s) == 1 ? ((s) + 1 < len ? d[((s) + 1) % TR_S] : D[seg0 + (s) + 1]) : T(1))
    // speculative run: lane 0 from the warmed-up state entering the block,
    // the others from the zero state
    T cpv = T(0), bpv = T(0);
    if (t == 0 && blk0 > 0) warm_fwd<T>(blk0, cs, sw, cpv, bpv);
#pragma unroll
    for (int s = 0; s < TR_S; s++) {
        if (s < len) {
            fwd_row<T>(SKIND(kinds, s), a[s], d[s], c[s], b[s], DNEXT(s), cpv, bpv, cp[s], bp[s]);
            cpv = cp[s]; bpv = bp[s];
        } else {
            cp[s] = T(0); bp[s] = T(0);
        }
    }
    // fix-up rounds inside the workgroup
    const bool dep = len > 0 && t > 0 && SKIND(kinds, 0) != 1;
    bool accel = false;   // the CE shortcut is on (uniform)
    bool cen = false;     // lane t+1 is CE
    int runlo = -1;       // lane t-1 is CE: the first lane v may take (else -1)
    for (int round = 0; round <= TR_T; round++) {
        if (len > 0) { sc[t] = cp[len - 1]; sb[t] = bp[len - 1]; }
        if (t == 0) changed = 0;
        if (accel) {
            bool on = false;
            if (cen) {
                T c6, b6, c7, b7;
                fwd_row<T>(0, a[6], d[6], c[6], b[6], T(1), cp[7], bp[7], c6, b6);
                fwd_row<T>(0, a[7], d[7], c[7], b[7], T(1), c6, b6, c7, b7);
                on = same(c7, cp[7]) && same(b7, bp[7]);
            }
            const unsigned long long m = __ballot(on);
            if ((t & 63) == 0) cyc_mask[t >> 6] = m;
        }
        __syncthreads();
        if (dep) {
            T sc0 = sc[t - 1], sb0 = sb[t - 1];
            if (runlo >= 0) {
                const int v = lowest_bit_in(cyc_mask, runlo, t - 2);
                if (v >= 0) { sc0 = sc[v]; sb0 = sb[v]; }
            }
            bool live = true;
#pragma unroll
            for (int s = 0; s < TR_S; s++) {
                if (live && s < len) {
                    T ck, bk;
                    fwd_row<T>(SKIND(kinds, s), a[s], d[s], c[s], b[s], DNEXT(s), sc0, sb0, ck, bk);
                    if (same(ck, cp[s]) && same(bk, bp[s])) {
                        live = false;
                    } else {
                        cp[s] = ck; bp[s] = bk; sc0 = ck; sb0 = bk;
                    }
                }
            }
            if (live) changed = 1;
        }
        __syncthreads();
        if (!changed) break;
        if (round == TR_ACC) {
            // mark the CE lanes (rows -2, -1 from the predecessor lane; lane 0
            // of a block is never CE)
            if (len == TR_S) {
                s67[0][t] = a[6]; s67[1][t] = d[6]; s67[2][t] = c[6]; s67[3][t] = b[6];
                s67[4][t] = a[7]; s67[5][t] = d[7]; s67[6][t] = c[7]; s67[7][t] = b[7];
            }
            __syncthreads();
            bool ce = len == TR_S && kinds == 0u && t > 0 && SKIND(seg_kinds(seg0 - TR_S, cs), 6) == 0 &&
                      SKIND(seg_kinds(seg0 - TR_S, cs), 7) == 0;
#pragma unroll
            for (int s = 2; s < TR_S; s++)
                ce = ce && same(a[s], a[s - 2]) && same(d[s], d[s - 2]) && same(c[s], c[s - 2]) &&
                     same(b[s], b[s - 2]);
            ce = ce && same(a[0], s67[0][t - 1]) && same(d[0], s67[1][t - 1]) && same(c[0], s67[2][t - 1]) &&
                 same(b[0], s67[3][t - 1]) && same(a[1], s67[4][t - 1]) && same(d[1], s67[5][t - 1]) &&
                 same(c[1], s67[6][t - 1]) && same(b[1], s67[7][t - 1]);
            const unsigned long long m = __ballot(ce);
            if ((t & 63) == 0) ce_mask[t >> 6] = m;
            __syncthreads();
            cen = t + 1 < TR_T && bit_at(ce_mask, t + 1);
            if (t >= 2 && bit_at(ce_mask, t - 1)) runlo = max(0, highest_clear_at_or_below(ce_mask, t - 1));
            accel = true;
        }
        __syncthreads();
    }
#undef DNEXT
    // block end state = state after the block's last valid row
    if (len > 0 && t * TR_S + len == nrow) {
        blk_c[blockIdx.x] = cp[len - 1];
        blk_b[blockIdx.x] = bp[len - 1];
    }
    seg_store(CP, seg0, len, cp);
    seg_store(BP, seg0, len, bp);
}

// Backward pass.  A workgroup first repairs the head of its block's forward
// values from the predecessor block's end state (one lane, row by row until
// the recomputed values coincide bit for bit with the speculative ones; the
// repaired rows are written back), then back-substitutes its segments with
// the same speculation and in-workgroup fix-up rounds.
template <typename T>
__global__ __launch_bounds__(TR_T) void k_trid_bwd(long n, Chains cs, const T *__restrict__ A, const T *__restrict__ D,
                                                   const T *__restrict__ Cc, const T *__restrict__ B,
                                                   T *__restrict__ CP, T *__restrict__ BP,
                                                   const T *__restrict__ blk_c, const T *__restrict__ blk_b,
                                                   T *__restrict__ X, T *__restrict__ blk_x, T *__restrict__ blk_xe,
                                                   int *__restrict__ fail, T *__restrict__ B2,
                                                   const unsigned char *__restrict__ Z, const int *stop) {
    if (stop && *stop) return;
    __shared__ T scp[TR_T * TR_P], sbp[TR_T * TR_P];
    __shared__ T sx[TR_T];
    __shared__ unsigned long long ce_mask[TR_T / 64], cyc_mask[TR_T / 64];
    __shared__ T sw[2][TR_WB];
    // the head rows a repair usually needs (A, D, C, B of rows blk0 .. +TR_H-1
    // and D of the row after each), loaded by TR_H lanes at once
    constexpr int TR_H = 64;
    __shared__ T sh[5][TR_H];
    __shared__ int changed;
    const int t = threadIdx.x;
    const long blk0 = (long)blockIdx.x * TR_R;
    const int nrow = (int)min((long)TR_R, n - blk0);
    const long seg0 = blk0 + (long)t * TR_S;
    const int len = max(0, min(TR_S, nrow - t * TR_S));
#pragma unroll
    for (int m = 0; m < TR_S; m++) {
        const int k = t + TR_T * m;
        if (k < nrow) {
            scp[(k / TR_S) * TR_P + k % TR_S] = CP[blk0 + k];
            sbp[(k / TR_S) * TR_P + k % TR_S] = BP[blk0 + k];
        }
    }
    if (blockIdx.x > 0 && t < TR_H && t < nrow) {
        const long g = blk0 + t;
        sh[0][t] = A[g];
        sh[1][t] = D[g];
        sh[2][t] = Cc[g];
        sh[3][t] = B[g];
        sh[4][t] = g + 1 < n ? D[g + 1] : T(1);
    }
    if (blk0 + nrow < n) warm_bwd_load<T>(CP, BP, n, blk0 + nrow, sw);
    __syncthreads();
    if (t == 0 && blockIdx.x > 0) {
        long Lk;
        long r = chain_pos(cs, blk0, Lk);   // position of the block's first row in its chain
        if (r != 0) {
            T cpv = blk_c[blockIdx.x - 1], bpv = blk_b[blockIdx.x - 1];
            int k = 0;
            for (; k < nrow; chain_up(cs, blk0 + k, r, Lk), k++) {
                const int kind = kind_at(r, Lk);
                const long g = blk0 + k;
                const bool inl = k < TR_H;
                const T dn = (kind == 1) ? (inl ? sh[4][k] : D[g + 1]) : T(1);
                T ck, bk;
                fwd_row<T>(kind, inl ? sh[0][k] : A[g], inl ? sh[1][k] : D[g], inl ? sh[2][k] : Cc[g],
                           inl ? sh[3][k] : B[g], dn, cpv, bpv, ck, bk);
                const int q = (k / TR_S) * TR_P + k % TR_S;
                if (same(ck, scp[q]) && same(bk, sbp[q])) break;
                scp[q] = ck; sbp[q] = bk; CP[g] = ck; BP[g] = bk;
                cpv = ck; bpv = bk;
            }
            if (k == nrow) atomicOr(fail, 1);
        }
    }
    __syncthreads();
    T cp[TR_S], bp[TR_S], x[TR_S];
#pragma unroll
    for (int q = 0; q < TR_S; q++) { cp[q] = scp[t * TR_P + q]; bp[q] = sbp[t * TR_P + q]; }
    const unsigned kinds = seg_kinds(seg0, cs);
    // the lane with the block's last row starts from the warmed-up value of the
    // next block's first row (recorded for the check in k_trid_finish)
    T xn = T(0);
    if (len > 0 && t * TR_S + len == nrow && blk0 + nrow < n) {
        xn = warm_bwd<T>(sw, n, blk0 + nrow, cs);
        blk_xe[blockIdx.x] = xn;
    }
#pragma unroll
    for (int s = TR_S - 1; s >= 0; s--) {
        if (s < len) {
            x[s] = (SKIND(kinds, s) == 2) ? bp[s] : bp[s] - (cp[s] * xn);
            xn = x[s];
        } else {
            x[s] = T(0);
        }
    }
    // Fix-up rounds, with the CE shortcut of the forward pass mirrored: a lane
    // is CE if its rows are interior rows whose (c', b') equal those of the
    // row two after (rows 8, 9 being its successor's first two rows); if
    // lanes t+1 .. v-1 are CE and lane v's first value x0 is on the cycle of
    // its first two rows (g0(g1(x0)) == x0, g_s(x) = b'_s - c'_s x), each of
    // those lanes maps x0 to itself, so the exact value after lane t is x0.
    const bool dep = len > 0 && SKIND(kinds, len - 1) != 2 && t * TR_S + len < nrow;
    bool accel = false;   // the CE shortcut is on (uniform)
    bool cpv = false;     // lane t-1 is CE
    int runhi = -1;       // lane t+1 is CE: the last lane v may take (else -1)
    for (int round = 0; round <= TR_T; round++) {
        if (len > 0) sx[t] = x[0];
        if (t == 0) changed = 0;
        if (accel) {
            bool on = false;
            if (cpv) {
                const T y = bp[1] - (cp[1] * x[0]);
                on = same(bp[0] - (cp[0] * y), x[0]);
            }
            const unsigned long long m = __ballot(on);
            if ((t & 63) == 0) cyc_mask[t >> 6] = m;
        }
        __syncthreads();
        if (dep) {
            T xv = sx[t + 1];
            if (runhi >= 0) {
                const int v = highest_bit_in(cyc_mask, t + 2, runhi);
                if (v >= 0) xv = sx[v];
            }
            bool live = true;
#pragma unroll
            for (int s = TR_S - 1; s >= 0; s--) {
                if (live && s < len) {
                    const T xk = (SKIND(kinds, s) == 2) ? bp[s] : bp[s] - (cp[s] * xv);
                    if (same(xk, x[s])) live = false; else { x[s] = xk; xv = xk; }
                }
            }
            if (live) changed = 1;
        }
        __syncthreads();
        if (!changed) break;
        if (round == TR_ACC) {
            // mark the CE lanes (rows 8, 9 from the successor's staged c', b')
            bool ce = len == TR_S && kinds == 0u && (t + 1) * TR_S + 2 <= nrow;
            if (ce) {
                const unsigned kn = seg_kinds(seg0 + TR_S, cs);
                ce = SKIND(kn, 0) == 0 && SKIND(kn, 1) == 0;
            }
#pragma unroll
            for (int s = 0; s + 2 < TR_S; s++) ce = ce && same(cp[s], cp[s + 2]) && same(bp[s], bp[s + 2]);
            ce = ce && same(cp[6], scp[(t + 1) * TR_P]) && same(bp[6], sbp[(t + 1) * TR_P]) &&
                 same(cp[7], scp[(t + 1) * TR_P + 1]) && same(bp[7], sbp[(t + 1) * TR_P + 1]);
            const unsigned long long m = __ballot(ce);
            if ((t & 63) == 0) ce_mask[t >> 6] = m;
            __syncthreads();
            cpv = t >= 1 && bit_at(ce_mask, t - 1);
            if (t + 1 < TR_T && bit_at(ce_mask, t + 1)) runhi = min(TR_T - 1, lowest_clear_at_or_above(ce_mask, t + 1, TR_T));
            accel = true;
        }
        __syncthreads();
    }
    if (t == 0 && len > 0) blk_x[blockIdx.x] = x[0];
    seg_store(X, seg0, len, x);
    if (B2) {
        unsigned long long zb = 0;   // the segment's flags, one byte per row
        if (len == TR_S) {
            zb = *reinterpret_cast<const unsigned long long *>(Z + seg0);
        } else {
            for (int q = 0; q < len; q++) zb |= (unsigned long long)Z[seg0 + q] << (8 * q);
        }
#pragma unroll
        for (int q = 0; q < TR_S; q++)
            if ((zb >> (8 * q)) & 0xffu) x[q] = T(0);
        seg_store(B2, seg0, len, x);
    }
}

// One workgroup: repair each block's tail from its successor's first value,
// then, if any repair failed to coalesce inside its block, solve every chain
// sequentially (exact in every case).  rearm (may be null) is set to 1 once
// the solve has run: the window solver's keep flag, which is this solve's
// stop flag, so it is set again for the next window solve.
constexpr int TRF_T = 1024;
template <typename T>
__global__ __launch_bounds__(TRF_T) void k_trid_finish(long n, Chains cs, int nblk, const T *__restrict__ A,
                                                       const T *__restrict__ D, const T *__restrict__ Cc,
                                                       const T *__restrict__ B, T *__restrict__ CP,
                                                       T *__restrict__ BP, T *__restrict__ X,
                                                       const T *__restrict__ blk_x, const T *__restrict__ blk_xe,
                                                       int *__restrict__ fail, int *__restrict__ nfallback,
                                                       T *__restrict__ B2, const unsigned char *__restrict__ Z,
                                                       const int *stop, int *rearm) {
    if (stop && *stop) return;
    for (int blk = threadIdx.x; blk < nblk - 1; blk += TRF_T) {
        const long k0 = (long)blk * TR_R, k1 = min(k0 + TR_R, n);
        long Lk;
        long r = chain_pos(cs, k1 - 1, Lk);   // position of row k in its chain
        if (chain_end(r, Lk)) continue;
        T xv = blk_x[blk + 1];
        if (same(xv, blk_xe[blk])) continue;   // the block started from the exact value
        long k = k1 - 1;
        for (; k >= k0; chain_down(cs, k, r, Lk), k--) {
            const T xk = chain_end(r, Lk) ? BP[k] : BP[k] - (CP[k] * xv);
            if (same(xk, X[k])) break;
            X[k] = xk;
            if (B2) B2[k] = Z[k] ? T(0) : xk;
            xv = xk;
        }
        if (k < k0) atomicOr(fail, 1);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (rearm) *rearm = 1;   // every thread read *stop at entry, before the barrier
    if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    for (long c0 = 0; c0 < n;) {
        const long L = c0 < cs.n1 ? cs.L1 : cs.L2;
        T cpv = T(0), bpv = T(0);
        for (long k = c0; k < c0 + L; k++) {
            const int kind = row_kind(k, cs);
            fwd_row<T>(kind, A[k], D[k], Cc[k], B[k], kind == 1 ? D[k + 1] : T(1), cpv, bpv, CP[k], BP[k]);
            cpv = CP[k]; bpv = BP[k];
        }
        T xv = T(0);
        for (long k = c0 + L - 1; k >= c0; k--) {
            X[k] = (row_kind(k, cs) == 2) ? BP[k] : BP[k] - (CP[k] * xv);
            xv = X[k];
            if (B2) B2[k] = Z[k] ? T(0) : xv;
        }
        c0 += L;
    }
    *fail = 0;
    *nfallback += 1;
}

// ------------------------------------------------------------------------
// Windowed chain solve for the mixed path (FP32, not bit-exact).  The FP32
// chains need only agree with the FP64 solve to FP32 accuracy, and the
// momentum chains are strongly diagonally dominant, so the coupling of a
// row to rows far away decays geometrically.  Each workgroup solves a window
// of TW_N rows with the coupling at the window's two edges dropped and keeps
// the TW_O rows in its middle (TW_H rows from either edge): for a dominance
// ratio (|a| + |c|) / |d| <= TW_DOM the dropped coupling changes the kept
// rows by less than about 0.63^TW_H = 1e-13 relative.  The window system is
// the chain's own matrix, with the first-row quirk (S(1,2) = c1 d1 / d2) and
// no coupling across chain starts and ends.
//
// Inside the window (partitioned Thomas + PCR): each thread holds TW_S rows
// in registers and expresses its interior rows through its first and last
// row (three 6-row Thomas solves: particular, left and right spikes), which
// leaves two interface rows per thread, a tridiagonal system of 2 TW_T rows;
// parallel cyclic reduction in LDS solves it (each level l combines a row
// with the rows 2^l away; at most log2(2 TW_T) levels, fewer once the
// couplings left are negligible), and the interior rows follow from the
// interface values.  A window whose dominance ratio exceeds TW_DOM clears
// *keep, and the bit-exact speculative solver (launched after this kernel
// with stop = keep) solves the chains instead.
// Window shapes: TW_T lanes (128, 256 or 512: windows of 1024, 2048 or
// 4096 rows, at most 8, 9 or 10 PCR levels), chosen by
// wd2_ctx_trid_autotune; 256 unless tuned.
constexpr int TW_S = 8, TW_H = 64;
template <int TW_T>
struct TwShape {
    static constexpr int N = TW_T * TW_S, O = N - 2 * TW_H, M = 2 * TW_T;
};
constexpr float TW_DOM = 0.9f;
constexpr float TW_EPS = 0x1p-48f;   // PCR stops once every coupling is below this share of its diagonal

__global__ void k_trid_keep_set(int *keep) { *keep = 1; }

// Hardware reciprocal (1 ulp) for the window solver, which need not round
// like the reference.
__device__ inline float wrcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T, int TW_T>
__global__ __launch_bounds__(TW_T) void k_trid_window(long n, Chains cs, const T *__restrict__ A,
                                                      const T *__restrict__ D, const T *__restrict__ C,
                                                      const T *__restrict__ B, T *__restrict__ X, T *__restrict__ B2,
                                                      const unsigned char *__restrict__ Z, const int *stop,
                                                      int *keep) {
    static_assert(std::is_same<T, float>::value, "the window solver serves the FP32 chains");
    constexpr int TW_N = TwShape<TW_T>::N, TW_O = TwShape<TW_T>::O, TW_M = TwShape<TW_T>::M;
    if (stop && *stop) return;
    // interface rows in LDS: sub- and super-diagonal, right-hand side and
    // the reciprocal of the diagonal (each row's diagonal stays with its thread)
    __shared__ T ra[2][TW_M], rc[2][TW_M], rb[2][TW_M], rq[2][TW_M];
    const int t = threadIdx.x;
    const long own0 = (long)blockIdx.x * TW_O, w0 = own0 - TW_H;
    T a[TW_S], d[TW_S], c[TW_S], b[TW_S];
    const long k0 = w0 + (long)t * TW_S;   // a multiple of TW_S: 16-byte aligned segments
    if (k0 >= 0 && k0 + TW_S <= n) {
        seg_load(A, k0, TW_S, a);
        seg_load(D, k0, TW_S, d);
        seg_load(C, k0, TW_S, c);
        seg_load(B, k0, TW_S, b);
    } else {
#pragma unroll
        for (int s = 0; s < TW_S; s++) {
            const long k = k0 + s;
            const bool in = k >= 0 && k < n;   // outside the chains: x = 0, uncoupled
            a[s] = in ? A[k] : T(0);
            d[s] = in ? D[k] : T(1);
            c[s] = in ? C[k] : T(0);
            b[s] = in ? B[k] : T(0);
        }
    }
    // chain starts (no sub-diagonal; the first-row quirk on the
    // super-diagonal) and ends (no super-diagonal)
    bool weak = false;
    {
        long Lk = 1, r = 0;
        if (k0 >= 0 && k0 < n) r = chain_pos(cs, k0, Lk);
#pragma unroll
        for (int s = 0; s < TW_S; s++) {
            const long k = k0 + s;
            if (k >= 0 && k < n) {
                if (r == 0) {
                    a[s] = T(0);
                    if (Lk > 1) c[s] = c[s] * d[s] / D[k + 1];
                }
                if (r == Lk - 1) c[s] = T(0);
                if (!(fabsf(a[s]) + fabsf(c[s]) <= TW_DOM * fabsf(d[s]))) weak = true;
                chain_up(cs, k, r, Lk);
            } else if (k == -1) {
                r = 0;
                Lk = cs.n1 > 0 ? cs.L1 : cs.L2;
            }
        }
    }
    // the window's edges: drop the coupling out of the window
    if (t == 0) a[0] = T(0);
    if (t == TW_T - 1) c[TW_S - 1] = T(0);
    if (__any(weak)) *keep = 0;
    // interior rows 1..TW_S-2 through x0 and x7: y (particular), v (x0
    // spike, from -a1 x0), w (x7 spike, from -c6 x7)
    T cp[TW_S], y[TW_S], v[TW_S], w[TW_S];
    {
        T q = wrcp(d[1]);
        cp[1] = c[1] * q;
        y[1] = b[1] * q;
        v[1] = -a[1] * q;
        w[1] = T(0);
#pragma unroll
        for (int s = 2; s <= TW_S - 2; s++) {
            q = wrcp(d[s] - a[s] * cp[s - 1]);
            cp[s] = c[s] * q;
            y[s] = (b[s] - a[s] * y[s - 1]) * q;
            v[s] = -a[s] * v[s - 1] * q;
            w[s] = ((s == TW_S - 2 ? -c[s] : T(0)) - a[s] * w[s - 1]) * q;
        }
#pragma unroll
        for (int s = TW_S - 3; s >= 1; s--) {
            y[s] -= cp[s] * y[s + 1];
            v[s] -= cp[s] * v[s + 1];
            w[s] -= cp[s] * w[s + 1];
        }
    }
    // interface rows 2t (the segment's first row) and 2t+1 (its last)
    T di[2];
    {
        const int i0 = 2 * t, i1 = 2 * t + 1;
        di[0] = d[0] + c[0] * v[1];
        di[1] = d[TW_S - 1] + a[TW_S - 1] * w[TW_S - 2];
        ra[0][i0] = a[0];
        rc[0][i0] = c[0] * w[1];
        rb[0][i0] = b[0] - c[0] * y[1];
        rq[0][i0] = wrcp(di[0]);
        ra[0][i1] = a[TW_S - 1] * v[TW_S - 2];
        rc[0][i1] = c[TW_S - 1];
        rb[0][i1] = b[TW_S - 1] - a[TW_S - 1] * y[TW_S - 2];
        rq[0][i1] = wrcp(di[1]);
    }
    __syncthreads();
    // PCR on the interface rows: row i eliminates its coupling to rows i-h
    // and i+h with those rows' equations.  Each level squares the couplings
    // of a dominant system (the interface rows alternate the coupling of two
    // adjacent rows with that across a segment, the product of TW_S - 1
    // dominant rows), so the reduction stops at the first level after which
    // every row's remaining coupling is below TW_EPS of its diagonal; the
    // dropped terms change x by less than TW_EPS times x 2^l segments away.
    int cur = 0;
    for (int h = 1; h < TW_M; h <<= 1) {
        bool coupled = false;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int i = 2 * t + e;
            const T ai = ra[cur][i], ci = rc[cur][i], bi = rb[cur][i];
            T an = T(0), dn = di[e], cn = T(0), bn = bi;
            if (i - h >= 0) {
                const T k1 = ai * rq[cur][i - h];
                an = -ra[cur][i - h] * k1;
                dn -= rc[cur][i - h] * k1;
                bn -= rb[cur][i - h] * k1;
            }
            if (i + h < TW_M) {
                const T k2 = ci * rq[cur][i + h];
                cn = -rc[cur][i + h] * k2;
                dn -= ra[cur][i + h] * k2;
                bn -= rb[cur][i + h] * k2;
            }
            di[e] = dn;
            ra[cur ^ 1][i] = an;
            rc[cur ^ 1][i] = cn;
            rb[cur ^ 1][i] = bn;
            rq[cur ^ 1][i] = wrcp(dn);
            coupled |= !(fabsf(an) + fabsf(cn) <= TW_EPS * fabsf(dn));
        }
        const int more = __syncthreads_or(coupled);
        cur ^= 1;
        if (!more) break;
    }
    // the halo is TW_H / TW_S whole segments at either end, so a lane's rows
    // are all kept or all dropped; kept segments are stored as 16-byte vectors
    if (t < TW_H / TW_S || t >= TW_T - TW_H / TW_S || k0 >= n) return;
    const T x0 = rb[cur][2 * t] * rq[cur][2 * t], x7 = rb[cur][2 * t + 1] * rq[cur][2 * t + 1];
    T x[TW_S];
#pragma unroll
    for (int s = 0; s < TW_S; s++) x[s] = s == 0 ? x0 : (s == TW_S - 1 ? x7 : y[s] + v[s] * x0 + w[s] * x7);
    const int len = (int)(n - k0 < TW_S ? n - k0 : TW_S);
    seg_store(X, k0, len, x);
    if (B2) {
        unsigned long long zb = 0;   // the segment's identity-row flags, one byte per row
        if (Z) {
            if (len == TW_S) {
                zb = *reinterpret_cast<const unsigned long long *>(Z + k0);
            } else {
                for (int q = 0; q < len; q++) zb |= (unsigned long long)Z[k0 + q] << (8 * q);
            }
        }
#pragma unroll
        for (int q = 0; q < TW_S; q++)
            if ((zb >> (8 * q)) & 0xffu) x[q] = T(0);
        seg_store(B2, k0, len, x);
    }
}

template <typename T, int TW_T>
static void launch_window_kernel(long n, const Chains &cs, const T *A, const T *D, const T *C, const T *B, T *X,
                                 hipStream_t s, T *B2, const unsigned char *Z, const int *stop, int *keep) {
    const unsigned nb = (unsigned)((n + TwShape<TW_T>::O - 1) / TwShape<TW_T>::O);
    hipLaunchKernelGGL((k_trid_window<T, TW_T>), dim3(nb), dim3(TW_T), 0, s, n, cs, A, D, C, B, X, B2, Z, stop, keep);
}
template <typename T>
static void launch_window_shape(int wt, long n, const Chains &cs, const T *A, const T *D, const T *C, const T *B,
                                T *X, hipStream_t s, T *B2, const unsigned char *Z, const int *stop, int *keep) {
    if (wt == 128)
        launch_window_kernel<T, 128>(n, cs, A, D, C, B, X, s, B2, Z, stop, keep);
    else if (wt == 512)
        launch_window_kernel<T, 512>(n, cs, A, D, C, B, X, s, B2, Z, stop, keep);
    else
        launch_window_kernel<T, 256>(n, cs, A, D, C, B, X, s, B2, Z, stop, keep);
}
int trid_window_shape_ok(int wt) { return wt == 128 || wt == 256 || wt == 512; }

template <typename T>
void launch_trid_window(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C,
                        const T *B, T *X, hipStream_t s, T *B2, const unsigned char *Z, const int *stop, int wt) {
    // the solver's flag block: fail, nfallback, keep; keep is 1 between solves
    // (set at allocation, and again by each fallback solve that runs)
    int *keep = w.fail + 2;
    launch_window_shape<T>(wt, n, cs, A, D, C, B, X, s, B2, Z, stop, keep);
    // the bit-exact solver runs only where a window was too weakly dominant
    // (keep cleared) and then sets keep again; it also returns at entry once
    // the q-l loop has stopped, since then the window kernel left keep set
    launch_trid<T>(w, n, cs, A, D, C, B, X, s, B2, Z, keep, keep);
}
void launch_trid_keep_init(int *fail_block, hipStream_t s) {
    hipLaunchKernelGGL(k_trid_keep_set, dim3(1), dim3(1), 0, s, fail_block + 2);
}
// reps back-to-back launches of the window kernel of shape wt between events
// e0 and e1 (X written; no second right-hand side).
template <typename T>
void launch_trid_window_bench(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C,
                              const T *B, T *X, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int wt) {
    int *keep = w.fail + 2;
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < reps; k++) launch_window_shape<T>(wt, n, cs, A, D, C, B, X, s, nullptr, nullptr, nullptr, keep);
    (void)hipEventRecord(e1, s);
    // no fallback runs here: set keep again for the next solve
    hipLaunchKernelGGL(k_trid_keep_set, dim3(1), dim3(1), 0, s, keep);
}
template void launch_trid_window_bench<float>(const TridWorkT<float> &, long, const Chains &, const float *,
                                              const float *, const float *, const float *, float *, int, hipStream_t,
                                              hipEvent_t, hipEvent_t, int);
template void launch_trid_window<float>(const TridWorkT<float> &, long, const Chains &, const float *, const float *,
                                        const float *, const float *, float *, hipStream_t, float *,
                                        const unsigned char *, const int *, int);

int trid_blocks(long n) { return (int)((n + TR_R - 1) / TR_R); }

template <typename T>
void launch_trid(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C, const T *B,
                 T *X, hipStream_t s, T *B2, const unsigned char *Z, const int *stop, int *rearm) {
    const int nb = trid_blocks(n);
    hipLaunchKernelGGL(k_trid_fwd<T>, dim3(nb), dim3(TR_T), 0, s, n, cs, A, D, C, B, w.cp, w.bp, w.blk_c, w.blk_b,
                       stop);
    hipLaunchKernelGGL(k_trid_bwd<T>, dim3(nb), dim3(TR_T), 0, s, n, cs, A, D, C, B, w.cp, w.bp, w.blk_c, w.blk_b,
                       X, w.blk_x, w.blk_xe, w.fail, B2, Z, stop);
    hipLaunchKernelGGL(k_trid_finish<T>, dim3(1), dim3(TRF_T), 0, s, n, cs, nb, A, D, C, B, w.cp, w.bp, X, w.blk_x,
                       w.blk_xe, w.fail, w.nfallback, B2, Z, stop, rearm);
}
// reps back-to-back launches of the forward (which = 1) or backward (2)
// kernel of one solve between events e0 and e1, after a full solve that puts
// the forward kernel's outputs in place; a full solve afterwards leaves the
// flags and X consistent.
template <typename T>
void launch_trid_kernel_bench(const TridWorkT<T> &w, long n, const Chains &cs, const T *A, const T *D, const T *C,
                              const T *B, T *X, int which, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const int nb = trid_blocks(n);
    launch_trid<T>(w, n, cs, A, D, C, B, X, s);
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < reps; k++) {
        if (which == 1)
            hipLaunchKernelGGL(k_trid_fwd<T>, dim3(nb), dim3(TR_T), 0, s, n, cs, A, D, C, B, w.cp, w.bp, w.blk_c,
                               w.blk_b, nullptr);
        else
            hipLaunchKernelGGL(k_trid_bwd<T>, dim3(nb), dim3(TR_T), 0, s, n, cs, A, D, C, B, w.cp, w.bp, w.blk_c,
                               w.blk_b, X, w.blk_x, w.blk_xe, w.fail, nullptr, nullptr, nullptr);
    }
    (void)hipEventRecord(e1, s);
    launch_trid<T>(w, n, cs, A, D, C, B, X, s);
}
template void launch_trid_kernel_bench<double>(const TridWorkT<double> &, long, const Chains &, const double *,
                                               const double *, const double *, const double *, double *, int, int,
                                               hipStream_t, hipEvent_t, hipEvent_t);
template void launch_trid_kernel_bench<float>(const TridWorkT<float> &, long, const Chains &, const float *,
                                              const float *, const float *, const float *, float *, int, int,
                                              hipStream_t, hipEvent_t, hipEvent_t);

template <typename T>
void launch_trid(const TridWorkT<T> &w, long n, long L, const T *A, const T *D, const T *C, const T *B, T *X,
                 hipStream_t s, T *B2, const unsigned char *Z) {
    launch_trid<T>(w, n, Chains{n, L, L}, A, D, C, B, X, s, B2, Z);
}
template void launch_trid<double>(const TridWorkT<double> &, long, long, const double *, const double *,
                                  const double *, const double *, double *, hipStream_t, double *,
                                  const unsigned char *);
template void launch_trid<float>(const TridWorkT<float> &, long, long, const float *, const float *, const float *,
                                 const float *, float *, hipStream_t, float *, const unsigned char *);
template void launch_trid<double>(const TridWorkT<double> &, long, const Chains &, const double *, const double *,
                                  const double *, const double *, double *, hipStream_t, double *,
                                  const unsigned char *, const int *, int *);
template void launch_trid<float>(const TridWorkT<float> &, long, const Chains &, const float *, const float *,
                                 const float *, const float *, float *, hipStream_t, float *, const unsigned char *,
                                 const int *, int *);

}  // namespace wd2
