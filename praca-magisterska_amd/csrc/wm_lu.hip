// Batched dense complex LU with partial pivoting for gfx950.
//
// Replaces the reference's LAPACKE_zgetrf ("wersja referencyjna/Wire_MoM.cu":372,
// MKL) and cublasZgetrfBatched ("wersja gpu/Wire_MoM.cu":468).  Same contract
// as LAPACK zgetrf: column-major, pivot = first row with the largest
// |re|+|im| in the column (izamax), 1-based ipiv, info = first exactly-zero
// pivot, rows interchanged across the whole matrix.
//
// Blocked right-looking algorithm with panel width NB = 64, in outer blocks
// of kSub panels.  Per panel k:
//   1. k_panel      : GEPP of the n-k x NB panel.  A cluster of P workgroups
//                     per matrix holds the panel in registers (one row per
//                     thread, or two threads per row for complex128); per
//                     column the workgroups exchange their pivot candidates
//                     through tagged packets, all reduce them to the same
//                     pivot, swap and apply the rank-1 update locally.  It
//                     also emits the composite row permutation of the panel.
//   2. k_swap_trsm  : applies that permutation to the columns outside the
//                     panel and computes U12 = L11^{-1} A12 for the outer
//                     block's remaining columns (forward substitution, L11
//                     from LDS).
//   3. k_gemm_*     : A -= L U on the matrix cores, with three real products
//                     per complex product: complex128 with
//                     v_mfma_f64_16x16x4f64, complex64 with
//                     v_mfma_f32_32x32x2f32, or with split-bf16 operands and
//                     v_mfma_f32_32x32x16_bf16.  Inside an outer block a
//                     short product brings the later panels' rows up to
//                     date; one product with K = 64 kSub updates the
//                     trailing matrix per outer block.  The MFMA is issued on
//                     the transposed problem (C^T = U^T L^T) so each lane's
//                     accumulator column is a run of consecutive rows of C:
//                     the C read-modify-write is coalesced in column-major
//                     storage.
#include <algorithm>

#include "wm_cplx.h"
#include "wm_internal.h"

namespace wm {

namespace {

constexpr int NB = 64;
// Panels per outer block of the factorization (tunable per context).  Deeper
// blocks give the trailing update a longer K (better MFMA use) at the cost of
// more narrow updates inside the block; measured on MI355X at N = 8192, 8
// panels beat 4 by 1.4-6.5 % for every factor precision.
constexpr int kSubDefault = 4, kSubLarge = 8, kSubLargeN = 8192;
constexpr int kTprF32 = 1;      // parts per panel row (waves sharing a row) for complex64 factors: one
                                // part keeps the cluster small enough for 16 matrices per launch at N = 8192

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// Panel exchange buffers.  The workgroups of one panel cluster talk through
// 8-byte packets {32-bit payload, 32-bit tag}; the tag names the column (a
// per-launch base plus the column index, never reused while the buffer lives),
// so a reader polls each packet until its tag matches and no fence or counter
// is needed: 8-byte loads and stores are single-copy atomic, and agent-scope
// relaxed atomics go to the device coherence point without flushing caches.
// Packets alternate between two buffers by column parity: a workgroup
// publishes column j+2 only after reading all packets of column j+1, which
// every workgroup published after its reads of column j.
//
// Per launch member and parity:
//   slot [pmax][4] : |candidate| as two 32-bit halves of a double, its position, unused
//   cand [pmax][NW]: each workgroup's candidate row (NW 32-bit words)
// followed by the panel permutation: cnt, rows[2NB], src[2NB].
template <typename V>
struct XLayout {
    static constexpr int NW = NB * (int)sizeof(V) / 4;  // 32-bit words per panel row
    size_t par_bytes, perm_off, mat_bytes;
    int pmax;
    __host__ __device__ XLayout(int pmax_) : pmax(pmax_)
    {
        par_bytes = ((size_t)pmax * 4 + (size_t)pmax * NW) * sizeof(uint64_t);
        par_bytes = (par_bytes + 255) & ~(size_t)255;
        perm_off = 2 * par_bytes;
        mat_bytes = perm_off + (1 + 4 * NB) * sizeof(int);
        mat_bytes = (mat_bytes + 255) & ~(size_t)255;
    }
    __device__ uint64_t *slot(char *m, int par) const { return (uint64_t *)(m + par * par_bytes); }
    __device__ uint64_t *cand(char *m, int par) const { return slot(m, par) + (size_t)pmax * 4; }
    __device__ int *perm(char *m) const { return (int *)(m + perm_off); }
};

template <typename T>
struct PanelArgs {
    typename CT<T>::V *A;
    int lda;
    int64_t strideA;
    int n, k, P, pmax;
    int *ipiv;
    int64_t strideP;
    int *info;
    int *err;
    char *xbuf;
    const int *bidx;  // launch member -> matrix index (null: identity)
    const uint32_t *tagp;  // device tag base of the current factorization (bumped per launch chunk)
    uint32_t toff;         // packet tags of this launch are *tagp + toff + 1 .. *tagp + toff + NB
    const int *nmem;       // per matrix index: its own size (rows/columns beyond are the identity), or null
};

// Argmax combine: larger value wins, equal values go to the lower row.
__device__ inline bool better(double v2, int r2, double v1, int r1) { return v2 > v1 || (v2 == v1 && r2 < r1); }

// Cross-lane moves within rows of 16 lanes (DPP: quad permutations and row
// rotations, every source lane valid).
template <int CTRL>
__device__ inline double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ inline int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
__device__ inline double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Maximum over the wave (quad_perm [1,0,3,2], [2,3,0,1], row_ror 4, row_ror 8,
// then the four row maxima); every lane gets the result.
__device__ inline double wave_max_f64(double v)
{
    v = fmax(v, dpp_f64<0xB1>(v));
    v = fmax(v, dpp_f64<0x4E>(v));
    v = fmax(v, dpp_f64<0x124>(v));
    v = fmax(v, dpp_f64<0x128>(v));
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}
__device__ inline int wave_min_i32(int v)
{
    v = min(v, dpp_i32<0xB1>(v));
    v = min(v, dpp_i32<0x4E>(v));
    v = min(v, dpp_i32<0x124>(v));
    v = min(v, dpp_i32<0x128>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// GEPP of the (n-k) x NB panel at column k for one matrix per blockIdx.y.
// Workgroup w holds rows k + w*RW .. k + (w+1)*RW - 1 in registers.  A row is
// split into TPR parts of W = NB/TPR columns held by different waves ("chunk"
// c = wave / NWC), so that within a wave the column index is uniform:
// finished columns are skipped by uniform branches, only the waves owning
// column j take part in its argmax (a DPP reduction), and each thread carries
// its row's value in the next pivot column (nxr, nxi), taken from the rank-1
// update that produces it, so registers are only indexed statically.
//
// Interchanges move labels, not data: each thread keeps its row and carries
// the row's current position lr (initially the row itself).  Exchanging rows
// k+j and gr swaps the two labels; the row labelled k+j is the pivot row and
// is final from then on, the rows labelled above k+j take the rank-1 update.
// The arithmetic and the pivots are those of the interchanging algorithm
// (ties go to the lowest position, as izamax), and each row is stored at its
// final position at the end.
template <typename T, int TPR>
__global__ __launch_bounds__(256, 3) void k_panel(PanelArgs<T> a)
{
    using V = typename CT<T>::V;
    constexpr int W = NB / TPR;     // columns per thread
    constexpr int NWC = 4 / TPR;    // waves per chunk
    constexpr int RW = 64 * NWC;    // rows per workgroup
    constexpr int NW = XLayout<V>::NW;
    constexpr int kNone = 0x7fffffff;
    static_assert(NW <= 256 && NW % 64 == 0 && RW >= NB, "panel layout");
    const int bl = blockIdx.y, w = blockIdx.x;          // bl: member of this launch
    const int b = a.bidx ? a.bidx[bl] : bl;              // b: matrix index
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = wave / NWC, rb = wave % NWC;           // chunk (columns c*W ..), row block
    const int k = a.k, base = k + w * RW;
    const int row = base + rb * 64 + lane;
    const bool valid = row < a.n;
    V *Ab = a.A + (size_t)b * a.strideA;
    const XLayout<V> X(a.pmax);
    char *xm = a.xbuf + (size_t)bl * X.mat_bytes;
    if (a.nmem && k >= a.nmem[b]) {
        // the member's own matrix ends above this panel: an identity panel, no
        // interchanges (pivots k+1 .. k+NB) and an empty permutation
        if (w == 0 && tid < NB) a.ipiv[(size_t)b * a.strideP + k + tid] = k + tid + 1;
        if (w == 0 && tid == 0) X.perm(xm)[0] = 0;
        return;
    }

    T re[W], im[W];
#pragma unroll
    for (int i = 0; i < W; ++i) {
        V x = V{0, 0};
        if (valid) x = Ab[row + (size_t)(k + c * W + i) * a.lda];
        re[i] = x.x;
        im[i] = x.y;
    }

    __shared__ double s_wv[NWC];
    __shared__ int s_wr[NWC];
    __shared__ V s_cand[NB];  // this workgroup's best row
    __shared__ V s_u[NB];     // pivot row
    __shared__ V s_l[TPR > 1 ? RW : 1];  // multipliers for the other chunks
    __shared__ double s_gv;
    __shared__ int s_gr;
    __shared__ double2 s_rinv;  // reciprocal of the pivot
    __shared__ int s_piv[NB];  // pivot row of each column (workgroup 0)

    const uint32_t tag0 = *a.tagp + a.toff;
    int lr = row;  // current position of this thread's row
    // this row's value in the column of the next step (column 0 is in chunk 0)
    T nxr = re[0], nxi = im[0];
    for (int j = 0; j < NB; ++j) {
        const int par = j & 1;
        const int pr = k + j;
        const int cj = j / W;
        const uint32_t tag = tag0 + (uint32_t)j + 1u;

        // 1. best row of each wave of the chunk holding column j: largest
        //    |re|+|im| among positions >= k+j, the lowest position on ties
        if (c == cj) {
            const double v = (valid && lr >= pr) ? fabs((double)nxr) + fabs((double)nxi) : -1.0;
            const double m = wave_max_f64(v);
            const int r = wave_min_i32(v == m && v >= 0.0 ? lr : kNone);
            if (lane == 0) {
                s_wv[rb] = m;
                s_wr[rb] = r;
            }
        }
        __syncthreads();

        // 2. this workgroup's candidate: its slot is published at once, so the
        //    other workgroups' reductions start while the row is staged in LDS
        double wv = s_wv[0];
        int wr = s_wr[0];
#pragma unroll
        for (int q = 1; q < NWC; ++q)
            if (better(s_wv[q], s_wr[q], wv, wr)) { wv = s_wv[q]; wr = s_wr[q]; }
        uint64_t *slot = X.slot(xm, par);
        uint64_t *cand = X.cand(xm, par);
        if (tid < 3) {
            const uint64_t bits = (uint64_t)__double_as_longlong(wv);
            const uint32_t word = tid == 0 ? (uint32_t)bits : tid == 1 ? (uint32_t)(bits >> 32) : (uint32_t)wr;
            put_pkt(slot + (size_t)w * 4 + tid, tag, word);
        }
        if (wr != kNone && lr == wr) {
#pragma unroll
            for (int i = 0; i < W; ++i) s_cand[c * W + i] = V{re[i], im[i]};
        }
        __syncthreads();

        // 3. publish the candidate row
        if (tid < NW) put_pkt(cand + (size_t)w * NW + tid, tag, ((const uint32_t *)s_cand)[tid]);

        // 4. wave 0 reduces all candidates to the same pivot (first max), then
        //    fetches the pivot row from the workgroup holding it and takes the
        //    pivot's reciprocal, so one barrier hands everything to the others
        if (wave == 0) {
            double gv = -1.0;
            int gr = kNone, gs = 0;
            for (int s = lane; s < a.P; s += 64) {
                const uint64_t *ps = slot + (size_t)s * 4;
                uint64_t w0, w1, w2;
                long it = 0;
                for (;;) {
                    w0 = ld_pkt(ps);
                    w1 = ld_pkt(ps + 1);
                    w2 = ld_pkt(ps + 2);
                    if ((uint32_t)(w0 >> 32) == tag && (uint32_t)(w1 >> 32) == tag && (uint32_t)(w2 >> 32) == tag)
                        break;
                    if (spin_give_up(it, a.err)) break;
                }
                const double v2 = __longlong_as_double((long long)(((w1 & 0xffffffffull) << 32) | (w0 & 0xffffffffull)));
                const int r2 = (int)(uint32_t)w2;
                if (better(v2, r2, gv, gr)) { gv = v2; gr = r2; gs = s; }
            }
            const double m = wave_max_f64(gv);
            const int r = wave_min_i32(gv == m ? gr : kNone);
            const unsigned long long win = __ballot(gv == m && gr == r);
            const int ws = __builtin_amdgcn_readlane(gs, win ? (int)__builtin_ctzll(win) : 0);
            // no usable candidate (non-finite column): a zero pivot in place
            const bool ok = r >= pr && r < a.n;
            constexpr int PL = NW / 64;  // pivot-row words per lane, loaded together
            const uint64_t *pu = cand + (size_t)(ok ? ws : 0) * NW + lane;
            uint64_t wu[PL];
            long it = 0;
            for (;;) {
                bool got = true;
#pragma unroll
                for (int e = 0; e < PL; ++e) {
                    wu[e] = ld_pkt(pu + 64 * e);
                    got = got && (uint32_t)(wu[e] >> 32) == tag;
                }
                if (got) break;
                if (spin_give_up(it, a.err)) break;
            }
#pragma unroll
            for (int e = 0; e < PL; ++e) ((uint32_t *)s_u)[lane + 64 * e] = (uint32_t)wu[e];
            if (lane == 0) {
                s_gv = ok ? m : 0.0;
                s_gr = ok ? r : pr;
                const V p = s_u[j];  // written by this wave above
                s_rinv = ok && m != 0.0 ? crecip((double)p.x, (double)p.y) : make_double2(0.0, 0.0);
            }
        }
        __syncthreads();
        const double gv = s_gv;
        const int gr = s_gr;

        // 5. interchange positions k+j and gr, scale column j, rank-1 update
        const double2 rinv = s_rinv;
        if (w == 0 && tid == 0) {
            s_piv[j] = gr;
            a.ipiv[(size_t)b * a.strideP + pr] = gr + 1;
            if (gv == 0.0 && a.info[b] == 0) a.info[b] = pr + 1;
        }
        if (lr == gr) lr = pr;
        else if (lr == pr) lr = gr;
        const bool act = gv != 0.0 && valid && lr > pr;
        T lre = 0, lim = 0;
        if (c == cj) {
            if (act) {
                lre = (T)((double)nxr * rinv.x - (double)nxi * rinv.y);
                lim = (T)((double)nxr * rinv.y + (double)nxi * rinv.x);
            }
            if (TPR > 1) s_l[rb * 64 + lane] = V{lre, lim};
        }
        if (TPR > 1) {
            __syncthreads();
            if (c != cj) {
                const V l = s_l[rb * 64 + lane];
                lre = l.x;
                lim = l.y;
            }
        }
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const int col = c * W + i;  // uniform
            if (col == j) {             // the multiplier replaces column j
                if (act) {
                    re[i] = lre;
                    im[i] = lim;
                }
            } else if (col > j) {
                // complex64 factors: rows that take no update have a zero
                // multiplier, so the update needs no per-row branch (a
                // non-finite pivot row then spreads to finished rows, and the
                // refinement of such factors fails over to FP64 factors);
                // complex128 factors keep LAPACK's untouched finished rows
                if (sizeof(T) == 4 || act) {
                    const V u = s_u[col];
                    re[i] -= lre * u.x - lim * u.y;
                    im[i] -= lre * u.y + lim * u.x;
                }
                if (col == j + 1) {
                    nxr = re[i];
                    nxi = im[i];
                }
            }
            // keep the scheduler from hoisting every column's pivot-row load at once
            if ((i & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
    }

    // each row to its final position
#pragma unroll
    for (int i = 0; i < W; ++i)
        if (valid) Ab[lr + (size_t)(k + c * W + i) * a.lda] = V{re[i], im[i]};

    // composite permutation of the panel's interchanges, for the other columns:
    // after applying them in order, row rows[t] holds the old content of src[t].
    // Slots t < NB are rows k + t; the rows from below the panel that take
    // part get slots NB, NB+1, ... in order of first use.  One wave of
    // workgroup 0 keeps slot t and slot NB + t in lane t (NB is the wave
    // width), so each interchange is a ballot and two lane reads.
    __syncthreads();  // s_piv complete
    if (w == 0 && wave == 0) {
        static_assert(NB == 64, "one slot pair per lane");
        const int rlo = k + lane;
        int slo = k + lane, rhi = -1, shi = -1;
        int cnt = NB;
        for (int jx = 0; jx < NB; ++jx) {
            const int pp = s_piv[jx];
            if (pp == k + jx) continue;
            int tp;
            if (pp < k + NB) {
                tp = pp - k;
            } else {
                const unsigned long long hit = __ballot(NB + lane < cnt && rhi == pp);
                if (hit) {
                    tp = NB + (int)__builtin_ctzll(hit);
                } else {
                    tp = cnt++;
                    if (lane == tp - NB) { rhi = pp; shi = pp; }
                }
            }
            const int sj = __builtin_amdgcn_readlane(slo, jx);
            const int st = tp < NB ? __builtin_amdgcn_readlane(slo, tp) : __builtin_amdgcn_readlane(shi, tp - NB);
            if (lane == jx) slo = st;
            if (tp < NB) {
                if (lane == tp) slo = sj;
            } else if (lane == tp - NB) {
                shi = sj;
            }
        }
        int *pm = X.perm(xm);
        int *rows = pm + 1, *src = pm + 1 + 2 * NB;
        rows[lane] = rlo;
        src[lane] = slo;
        if (NB + lane < cnt) {
            rows[NB + lane] = rhi;
            src[NB + lane] = shi;
        }
        if (lane == 0) pm[0] = cnt;
    }
}

// ---------------------------------------------------------------------------
// Row interchanges outside the panel and U12 = L11^{-1} A12.  Tiles of CB
// columns, three kinds by tile index:
//   [0, ntl)           columns [tile*CB, k): interchanges only (L to the left)
//   [ntl, ntl+ntr)     columns [c_tr0, c_tr1): interchanges (if do_swap) and
//                      the triangular solve of rows k..k+NB-1
//   [ntl+ntr, ...)     columns [c_sw0, n): interchanges only (columns whose
//                      rows k..k+NB-1 still await earlier panels' updates)
struct SwapArgs {
    int n, k, pmax, ntl, ntr, c_tr0, c_tr1, c_sw0, do_swap;
    const int *nmem;  // per matrix index: its own size (identity beyond), or null
};

template <typename T, int CB>
__global__ __launch_bounds__(256) void k_swap_trsm(typename CT<T>::V *A, int lda, int64_t strideA, SwapArgs a,
                                                   char *xbuf, const int *bidx)
{
    using V = typename CT<T>::V;
    constexpr int RG = 256 / CB;        // row groups for the triangular solve
    constexpr int RPT = NB / RG;        // rows per thread in the solve
    constexpr int EPT = 2 * NB * CB / 256;  // permuted elements per thread (max)
    const int bl = blockIdx.y, tile = blockIdx.x, tid = threadIdx.x;
    const int b = bidx ? bidx[bl] : bl;
    const int n = a.n, k = a.k;
    V *Ab = A + (size_t)b * strideA;
    const XLayout<V> X(a.pmax);
    char *xm = xbuf + (size_t)bl * X.mat_bytes;

    const bool trsm = tile >= a.ntl && tile < a.ntl + a.ntr;
    const int c0 = tile < a.ntl ? tile * CB : trsm ? a.c_tr0 + (tile - a.ntl) * CB
                                                   : a.c_sw0 + (tile - a.ntl - a.ntr) * CB;
    const int cend = tile < a.ntl ? k : trsm ? a.c_tr1 : n;
    // beyond the member's own size: no interchanges (its panel was the
    // identity) and zero blocks right of it (their solve changes nothing)
    if (a.nmem && (k >= a.nmem[b] || c0 >= a.nmem[b])) return;

    // gather the old contents of the touched rows, keep the new top block
    // (rows k..k+NB-1) of solved columns in LDS, then scatter.  Consecutive
    // lanes take consecutive row slots of one column: the NB rows k..k+NB-1,
    // source or target of most moves, are then one contiguous run per column
    // (coalesced in column-major storage); s_top is kept column by column,
    // padded so the solve's reads across columns spread over the LDS banks.
    __shared__ V s_top[CB][NB + 1];
    if (a.do_swap) {
        __shared__ int s_rows[2 * NB], s_src[2 * NB];
        __shared__ int s_cnt;
        const int *pm = X.perm(xm);
        if (tid == 0) s_cnt = pm[0];
        if (tid < 2 * NB) { s_rows[tid] = pm[1 + tid]; s_src[tid] = pm[1 + 2 * NB + tid]; }
        __syncthreads();
        const int tcnt = s_cnt;
        V val[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int idx = tid + 256 * e;
            const int t = idx % (2 * NB), cc = idx / (2 * NB), col = c0 + cc;
            val[e] = V{0, 0};
            if (t < tcnt && col < cend) {
                val[e] = Ab[s_src[t] + (size_t)col * lda];
                if (trsm && t < NB) s_top[cc][t] = val[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int idx = tid + 256 * e;
            const int t = idx % (2 * NB), col = c0 + idx / (2 * NB);
            if (t < tcnt && col < cend) Ab[s_rows[t] + (size_t)col * lda] = val[e];
        }
    } else if (trsm) {
        for (int idx = tid; idx < NB * CB; idx += 256) {
            const int t = idx % NB, cc = idx / NB, col = c0 + cc;
            if (col < cend) s_top[cc][t] = Ab[(k + t) + (size_t)col * lda];
        }
    }
    if (!trsm) return;

    // forward substitution with the unit lower L11 (strictly lower part in LDS)
    __shared__ V s_l[NB * (NB - 1) / 2];
    __shared__ V s_x[2][CB];
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e % NB, jj = e / NB;
        if (i > jj) s_l[i * (i - 1) / 2 + jj] = Ab[(k + i) + (size_t)(k + jj) * lda];
    }
    __syncthreads();
    const int c = tid % CB, rg = tid / CB;
    const bool cok = (c0 + c) < cend;
    T xr[RPT], xi[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const V v = cok ? s_top[c][rg + q * RG] : V{0, 0};
        xr[q] = v.x;
        xi[q] = v.y;
    }
    for (int jj = 0; jj < NB; ++jj) {
        const int par = jj & 1;
        if (rg == jj % RG) {
#pragma unroll
            for (int q = 0; q < RPT; ++q)
                if (q == jj / RG) s_x[par][c] = V{xr[q], xi[q]};
        }
        __syncthreads();
        const V xj = s_x[par][c];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int i = rg + q * RG;
            if (i > jj) {
                const V l = s_l[i * (i - 1) / 2 + jj];
                xr[q] -= l.x * xj.x - l.y * xj.y;
                xi[q] -= l.x * xj.y + l.y * xj.x;
            }
        }
    }
    if (cok) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) Ab[(k + rg + q * RG) + (size_t)(c0 + c) * lda] = V{xr[q], xi[q]};
    }
}

// ---------------------------------------------------------------------------
// Block update  C (M x N) -= A (M x K) * B (K x N)  on the matrix cores.
//
// Both kernels issue the MFMA on the transposed problem, C^T = B^T A^T:
// MFMA rows are columns of C and MFMA columns are rows of C, so each
// accumulator register of a lane maps to a run of consecutive rows of one
// column of C and the read-modify-write of C is coalesced.  Complex products
// take three real products (below).  K is walked in chunks staged through
// LDS (8 for complex128, 16 for complex64); the next chunk is fetched from
// global memory into registers while the current one is multiplied.
//
// Tiles are dealt to the XCDs in contiguous runs (workgroup ids go to the 8
// XCDs round-robin) and ordered in groups of row tiles (tile_coords), so the
// tiles an XCD runs at once share their L and U blocks in its L2.

// C = A[r0.., c0..] (mrows x ncols) -= A[r0.., kk..kk+K) * A[kk..kk+K, c0..); K a multiple of 64.
struct GemmArgs {
    int r0, c0, mrows, ncols, kk, K;
    const int *nmem;  // per matrix index: its own size (identity beyond, so its updates there are zero), or null
};

// Tile of workgroup g out of T, grouped per XCD; >= T means "no tile".
__device__ inline int xcd_tile(int g, int T)
{
    const int per = (T + 7) / 8;
    return (g % 8) * per + g / 8;
}

inline int xcd_grid(int T) { return 8 * ((T + 7) / 8); }

// Tile t (of TI x TJ) -> (row tile, column tile), in groups of kGroupM row
// tiles: consecutive t walk down a group's rows, then across its columns.
// With the contiguous per-XCD runs of xcd_tile, the 64 tiles an XCD runs at
// once cover about 8 row blocks x 8 column blocks, so their L and U blocks
// are shared through that XCD's L2 instead of each tile streaming its own.
constexpr int kGroupM = 8;
__device__ inline void tile_coords(int t, int TI, int TJ, int &ti, int &tj)
{
    const int per = kGroupM * TJ;
    const int g = t / per, r0 = g * kGroupM;
    const int rows = (TI - r0 < kGroupM) ? TI - r0 : kGroupM;
    const int u = t - g * per;
    ti = r0 + u % rows;
    tj = u / rows;
}

// complex128 with three real products per complex product (Gauss): over the
// K terms, P = sum Br Ar, Q = sum Bi Ai and S = sum (Br + Bi)(Ar + Ai), so
// Re(LU) = P - Q and Im(LU) = S - P - Q.  Three accumulators instead of two
// make the wave tile 32 x 32 (2 x 2 tiles of v_mfma_f64_16x16x4f64, lane l
// holding srcA[l%16][l/16], srcB[l/16][l%16] and D[l/16 + 4*v][l%16]);
// workgroup tile 64 x 64.  Against four products this takes a quarter fewer
// MFMA cycles per update.  The error bound per entry is that of the
// four-product form with |a||b| replaced by (|Ar| + |Ai|)(|Br| + |Bi|), at
// most twice as large.
// K is staged through LDS in chunks of 8: the smaller chunk's prefetch registers
// keep the kernel at three waves per SIMD, which hides the LDS and global
// latency better than two waves with chunks of 16 (83 % against 78 % of the
// FP64 matrix-core peak issued at K = 512, N = 8192).  Two chunks are in flight
// in registers (the next to stage and the one after): 168 VGPRs, four dwords of
// tile bookkeeping spilled outside the loop, 1 % faster than one chunk.
constexpr int GZ_BM = 64, GZ_BN = 64, GZ_KC = 8;

__global__ __launch_bounds__(256, 3) void k_gemm_f64(double2 *Abase, int lda, int64_t strideA, GemmArgs g,
                                                     const int *bidx, int TI, int TJ)
{
    const int t = xcd_tile(blockIdx.x, TI * TJ);
    if (t >= TI * TJ) return;
    const int b = bidx ? bidx[blockIdx.z] : (int)blockIdx.z;
    double2 *M = Abase + (size_t)b * strideA;
    const int mrows = g.mrows, ncols = g.ncols, K = g.K;
    int ti, tj;
    tile_coords(t, TI, TJ, ti, tj);
    const int i0 = ti * GZ_BM, j0 = tj * GZ_BN;
    if (g.nmem && (g.r0 + i0 >= g.nmem[b] || g.c0 + j0 >= g.nmem[b])) return;
    const double2 *A = M + g.r0 + (size_t)g.kk * lda;  // L block
    const double2 *B = M + g.kk + (size_t)g.c0 * lda;  // U block
    double2 *C = M + g.r0 + (size_t)g.c0 * lda;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;

    constexpr int EA = GZ_KC / 4, BJS = 256 / GZ_KC, EB = GZ_BN / BJS;
    __shared__ double2 sA[GZ_KC][GZ_BM + 1];
    __shared__ double2 sB[GZ_KC][GZ_BN + 1];
    // two chunks in flight in registers: the one to stage next and the one after
    double2 pa0[EA], pb0[EB], pa1[EA], pb1[EB];
    // chunk elements moved by this thread: L rows ai, k = ak + 4e; U k = bk, columns bj + BJS e
    const int ai = tid % GZ_BM, ak = tid / GZ_BM;
    const int bk = tid % GZ_KC, bj = tid / GZ_KC;
    const bool arow = i0 + ai < mrows;
    auto fetch = [&](int kc, double2 *pa, double2 *pb) {
        const double2 *pA = A + (i0 + ai) + (size_t)(kc + ak) * lda;
#pragma unroll
        for (int e = 0; e < EA; ++e) pa[e] = arow ? pA[(size_t)(4 * e) * lda] : double2{0.0, 0.0};
        const double2 *pB = B + (kc + bk) + (size_t)(j0 + bj) * lda;
#pragma unroll
        for (int e = 0; e < EB; ++e) pb[e] = (j0 + bj + BJS * e < ncols) ? pB[(size_t)(BJS * e) * lda] : double2{0.0, 0.0};
    };
    auto stash = [&](const double2 *pa, const double2 *pb) {
#pragma unroll
        for (int e = 0; e < EA; ++e) sA[ak + 4 * e][ai] = pa[e];
#pragma unroll
        for (int e = 0; e < EB; ++e) sB[bk][bj + BJS * e] = pb[e];
    };
    v4d P[2][2], Q[2][2], S[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) { P[x][y] = v4d{0, 0, 0, 0}; Q[x][y] = P[x][y]; S[x][y] = P[x][y]; }
    auto compute = [&]() {
#pragma unroll
        for (int ks = 0; ks < GZ_KC; ks += 4) {
            const int kk = ks + lane / 16;
            double2 bv[2], av[2];
            double bs[2], as[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                bv[q] = sB[kk][wj + q * 16 + lane % 16];
                av[q] = sA[kk][wi + q * 16 + lane % 16];
                bs[q] = bv[q].x + bv[q].y;
                as[q] = av[q].x + av[q].y;
            }
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti) P[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[tj].x, av[ti].x, P[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti) Q[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[tj].y, av[ti].y, Q[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti) S[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(bs[tj], as[ti], S[tj][ti], 0, 0, 0);
        }
    };

    // K is a multiple of 2 chunks: the loop takes two chunks per trip, chunk
    // kc from LDS while chunk kc + GZ_KC waits in pa1/pb1 and kc + 2 GZ_KC loads
    fetch(0, pa0, pb0);
    stash(pa0, pb0);
    fetch(GZ_KC, pa1, pb1);
    __syncthreads();
    for (int kc = 0; kc < K; kc += 2 * GZ_KC) {
        if (kc + 2 * GZ_KC < K) fetch(kc + 2 * GZ_KC, pa0, pb0);
        compute();
        __syncthreads();
        stash(pa1, pb1);
        __syncthreads();
        if (kc + 3 * GZ_KC < K) fetch(kc + 3 * GZ_KC, pa1, pb1);
        compute();
        if (kc + 2 * GZ_KC < K) {
            __syncthreads();
            stash(pa0, pb0);
            __syncthreads();
        }
    }
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = i0 + wi + ti * 16 + lane % 16;
                const int gj = j0 + wj + tj * 16 + (lane / 16) + 4 * v;
                if (gi < mrows && gj < ncols) {
                    const double p = P[tj][ti][v], q = Q[tj][ti][v];
                    double2 c = C[gi + (size_t)gj * lda];
                    c.x -= p - q;
                    c.y -= S[tj][ti][v] - p - q;
                    C[gi + (size_t)gj * lda] = c;
                }
            }
}

// complex64 with three real products per complex product, as k_gemm_f64:
// P = sum Br Ar, Q = sum Bi Ai, S = sum (Br + Bi)(Ar + Ai).  Workgroup tile
// 128 x 64, four waves of 64 x 32 (2 x 1 tiles of v_mfma_f32_32x32x2f32).
// Lane l holds srcA[l%32][l/32], srcB[l/32][l%32] and
// D[8*(v/4)+4*(l/32)+v%4][l%32].
constexpr int GC_BM = 128, GC_BN = 64, GC_KC = 16;

__global__ __launch_bounds__(256, 2) void k_gemm_f32(float2 *Abase, int lda, int64_t strideA, GemmArgs g,
                                                     const int *bidx, int TI, int TJ)
{
    const int t = xcd_tile(blockIdx.x, TI * TJ);
    if (t >= TI * TJ) return;
    const int b = bidx ? bidx[blockIdx.z] : (int)blockIdx.z;
    float2 *M = Abase + (size_t)b * strideA;
    const int mrows = g.mrows, ncols = g.ncols, K = g.K;
    int ti, tj;
    tile_coords(t, TI, TJ, ti, tj);
    const int i0 = ti * GC_BM, j0 = tj * GC_BN;
    if (g.nmem && (g.r0 + i0 >= g.nmem[b] || g.c0 + j0 >= g.nmem[b])) return;
    const float2 *A = M + g.r0 + (size_t)g.kk * lda;
    const float2 *B = M + g.kk + (size_t)g.c0 * lda;
    float2 *C = M + g.r0 + (size_t)g.c0 * lda;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 64, wj = (wave >> 1) * 32;

    __shared__ float2 sA[GC_KC][GC_BM + 2];
    __shared__ float2 sB[GC_KC][GC_BN + 2];
    float2 pa[8], pb[4];
    // chunk elements moved by this thread: L rows ai, k = ak + 2e; U k = bk, columns bj + 16e
    const int ai = tid % GC_BM, ak = tid / GC_BM;
    const int bk = tid % GC_KC, bj = tid / GC_KC;
    const bool arow = i0 + ai < mrows;
    auto fetch = [&](int kc) {
        const float2 *pA = A + (i0 + ai) + (size_t)(kc + ak) * lda;
#pragma unroll
        for (int e = 0; e < 8; ++e) pa[e] = arow ? pA[(size_t)(2 * e) * lda] : float2{0.f, 0.f};
        const float2 *pB = B + (kc + bk) + (size_t)(j0 + bj) * lda;
#pragma unroll
        for (int e = 0; e < 4; ++e) pb[e] = (j0 + bj + 16 * e < ncols) ? pB[(size_t)(16 * e) * lda] : float2{0.f, 0.f};
    };
    auto stash = [&]() {
#pragma unroll
        for (int e = 0; e < 8; ++e) sA[ak + 2 * e][ai] = pa[e];
#pragma unroll
        for (int e = 0; e < 4; ++e) sB[bk][bj + 16 * e] = pb[e];
    };
    v16f P[2], Q[2], S[2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int v = 0; v < 16; ++v) { P[x][v] = 0.f; Q[x][v] = 0.f; S[x][v] = 0.f; }

    fetch(0);
    stash();
    __syncthreads();
    for (int kc = 0; kc < K; kc += GC_KC) {
        if (kc + GC_KC < K) fetch(kc + GC_KC);
#pragma unroll
        for (int ks = 0; ks < GC_KC; ks += 2) {
            const int kk = ks + lane / 32;
            const float2 bv = sB[kk][wj + lane % 32];
            const float bs = bv.x + bv.y;
            float2 av[2];
            float as[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                av[q] = sA[kk][wi + q * 32 + lane % 32];
                as[q] = av[q].x + av[q].y;
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) P[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(bv.x, av[q].x, P[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 2; ++q) Q[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(bv.y, av[q].y, Q[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 2; ++q) S[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(bs, as[q], S[q], 0, 0, 0);
        }
        if (kc + GC_KC < K) {
            __syncthreads();
            stash();
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int gi = i0 + wi + q * 32 + lane % 32;
            const int gj = j0 + wj + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
            if (gi < mrows && gj < ncols) {
                const float p = P[q][v], r = Q[q][v];
                float2 c = C[gi + (size_t)gj * lda];
                c.x -= p - r;
                c.y -= S[q][v] - p - r;
                C[gi + (size_t)gj * lda] = c;
            }
        }
}

// complex64 with split-BF16x3 products: every FP32 operand x is split into
// bf16 hi = rn(x) and lo = rn(x - hi), and each real product X*Y is taken as
// Xhi*Yhi + Xhi*Ylo + Xlo*Yhi on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation (relative error about 2^-16 per product instead of 2^-24).
// Workgroup tile 128 x 128, four waves of 64 x 64 (2 x 2 MFMA tiles), K in
// chunks of 16.  Lane l holds srcA[l%32][8*(l/32)+i], srcB[8*(l/32)+i][l%32]
// (i = 0..7) and D[8*(v/4)+4*(l/32)+v%4][l%32].  The chunk is staged in LDS as
// four bf16 planes (re hi, re lo, im hi, im lo) with k contiguous per row of L
// and per column of U, so each operand is one 16-byte LDS read.  A row is 32
// bytes, two 16-byte halves, stored swapped in rows 8-15 of every 16 (an XOR
// swizzle instead of padding), so rows 0 and 8 of a pass no longer share banks
// (bank-conflict cycles 43 -> 20 % of the LDS's active cycles).  Two LDS buffers
// (64 KiB, two workgroups per CU): while the waves multiply one chunk they
// split and store the next (already in registers) into the other buffer and
// load the one after, with one barrier per chunk.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int GB_BM = 128, GB_BN = 128, GB_KC = 16;

__device__ inline void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// k offset (bf16) of half h (0: k 0-7, 1: k 8-15) of LDS row r of a chunk
__device__ inline int b3_half(int r, int h) { return 8 * (h ^ ((r >> 3) & 1)); }

__device__ inline bf16x8 neg_bf16x8(bf16x8 v)
{
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

__global__ __launch_bounds__(256, 2) void k_gemm_b3(float2 *Abase, int lda, int64_t strideA, GemmArgs g,
                                                    const int *bidx, int TI, int TJ)
{
    const int t = xcd_tile(blockIdx.x, TI * TJ);
    if (t >= TI * TJ) return;
    const int b = bidx ? bidx[blockIdx.z] : (int)blockIdx.z;
    float2 *M = Abase + (size_t)b * strideA;
    const int mrows = g.mrows, ncols = g.ncols, K = g.K;
    int ti, tj;
    tile_coords(t, TI, TJ, ti, tj);
    const int i0 = ti * GB_BM, j0 = tj * GB_BN;
    if (g.nmem && (g.r0 + i0 >= g.nmem[b] || g.c0 + j0 >= g.nmem[b])) return;
    const float2 *A = M + g.r0 + (size_t)g.kk * lda;  // L block
    const float2 *B = M + g.kk + (size_t)g.c0 * lda;  // U block
    float2 *C = M + g.r0 + (size_t)g.c0 * lda;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 64, wj = (wave >> 1) * 64;

    // [buffer][plane: 0 re hi, 1 re lo, 2 im hi, 3 im lo][row][k]
    __shared__ __attribute__((aligned(16))) __bf16 sL[2][4][GB_BM][GB_KC];
    __shared__ __attribute__((aligned(16))) __bf16 sU[2][4][GB_BN][GB_KC];
    // chunk elements moved by this thread: 8 consecutive k of one row of L and of one column of U
    const int lr = tid % GB_BM, lh = tid / GB_BM;  // L: row lr, k = 8*lh + e
    const int uc = tid / 2, uh = tid % 2;           // U: column uc, k = 8*uh + e
    const bool lok = i0 + lr < mrows, uok = j0 + uc < ncols;
    float2 pa[8], pb[8];
    auto fetch = [&](int kc) {
        const float2 *pA = A + (i0 + lr) + (size_t)(kc + 8 * lh) * lda;
#pragma unroll
        for (int e = 0; e < 8; ++e) pa[e] = lok ? pA[(size_t)e * lda] : float2{0.f, 0.f};
        const float2 *pB = B + (kc + 8 * uh) + (size_t)(j0 + uc) * lda;
#pragma unroll
        for (int e = 0; e < 8; ++e) pb[e] = uok ? pB[e] : float2{0.f, 0.f};
    };
    auto stash = [&](int buf) {
        bf16x8 v[4];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            __bf16 h, l;
            split_bf16(pa[e].x, h, l);
            v[0][e] = h;
            v[1][e] = l;
            split_bf16(pa[e].y, h, l);
            v[2][e] = h;
            v[3][e] = l;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) *(bf16x8 *)&sL[buf][q][lr][b3_half(lr, lh)] = v[q];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            __bf16 h, l;
            split_bf16(pb[e].x, h, l);
            v[0][e] = h;
            v[1][e] = l;
            split_bf16(pb[e].y, h, l);
            v[2][e] = h;
            v[3][e] = l;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) *(bf16x8 *)&sU[buf][q][uc][b3_half(uc, uh)] = v[q];
    };
    v16f accR[2][2], accI[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) { accR[x][y][v] = 0.f; accI[x][y][v] = 0.f; }

    auto compute = [&](int buf) {
        // U operands (MFMA rows = columns of C), L operands (MFMA columns = rows of C)
        bf16x8 ur[2][2], ui[2][2], un[2][2], lre[2][2], lim[2][2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int cj = wj + q * 32 + lane % 32, ri = wi + q * 32 + lane % 32;
            const int ku = b3_half(cj, lane / 32), kl = b3_half(ri, lane / 32);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                ur[q][h] = *(const bf16x8 *)&sU[buf][0 + h][cj][ku];
                ui[q][h] = *(const bf16x8 *)&sU[buf][2 + h][cj][ku];
                un[q][h] = neg_bf16x8(ui[q][h]);
                lre[q][h] = *(const bf16x8 *)&sL[buf][0 + h][ri][kl];
                lim[q][h] = *(const bf16x8 *)&sL[buf][2 + h][ri][kl];
            }
        }
        // split terms (hi,hi), (hi,lo), (lo,hi); real and imaginary accumulators alternate
#pragma unroll
        for (int term = 0; term < 3; ++term) {
            const int hu = term == 2 ? 1 : 0, hl = term == 1 ? 1 : 0;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    accR[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ur[x][hu], lre[y][hl], accR[x][y], 0, 0, 0);
                    accI[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ur[x][hu], lim[y][hl], accI[x][y], 0, 0, 0);
                }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    accR[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(un[x][hu], lim[y][hl], accR[x][y], 0, 0, 0);
                    accI[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ui[x][hu], lre[y][hl], accI[x][y], 0, 0, 0);
                }
        }
    };

    fetch(0);
    stash(0);
    if (GB_KC < K) fetch(GB_KC);
    __syncthreads();
    int buf = 0;
    for (int kc = 0; kc < K; kc += GB_KC, buf ^= 1) {
        // the next chunk (in registers) to the other buffer, the one after
        // from memory, then this chunk's products
        if (kc + GB_KC < K) {
            stash(buf ^ 1);
            if (kc + 2 * GB_KC < K) fetch(kc + 2 * GB_KC);
        }
        compute(buf);
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int gi = i0 + wi + y * 32 + lane % 32;
                const int gj = j0 + wj + x * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
                if (gi < mrows && gj < ncols) {
                    float2 c = C[gi + (size_t)gj * lda];
                    c.x -= accR[x][y][v];
                    c.y -= accI[x][y][v];
                    C[gi + (size_t)gj * lda] = c;
                }
            }
}

// Launches C -= L U for nb_ matrices, C = A[r0.., c0..] (mrows x ncols),
// L = A[r0.., kk..kk+K), U = A[kk..kk+K, c0..).  gk: 8 complex128, 4 complex64,
// 2 complex64 with split-BF16x3 products.
void launch_gemm(int gk, void *A, int lda, int64_t strideA, const GemmArgs &g, int nb_, const int *bidx,
                 hipStream_t s)
{
    if (g.mrows <= 0 || g.ncols <= 0 || g.K <= 0) return;
    if (gk == 8) {
        const int TI = (g.mrows + GZ_BM - 1) / GZ_BM, TJ = (g.ncols + GZ_BN - 1) / GZ_BN;
        hipLaunchKernelGGL(k_gemm_f64, dim3(xcd_grid(TI * TJ), 1, nb_), dim3(256), 0, s, (double2 *)A, lda, strideA,
                           g, bidx, TI, TJ);
    } else if (gk == 4) {
        const int TI = (g.mrows + GC_BM - 1) / GC_BM, TJ = (g.ncols + GC_BN - 1) / GC_BN;
        hipLaunchKernelGGL(k_gemm_f32, dim3(xcd_grid(TI * TJ), 1, nb_), dim3(256), 0, s, (float2 *)A, lda, strideA,
                           g, bidx, TI, TJ);
    } else {
        const int TI = (g.mrows + GB_BM - 1) / GB_BM, TJ = (g.ncols + GB_BN - 1) / GB_BN;
        hipLaunchKernelGGL(k_gemm_b3, dim3(xcd_grid(TI * TJ), 1, nb_), dim3(256), 0, s, (float2 *)A, lda, strideA,
                           g, bidx, TI, TJ);
    }
}

__global__ void k_zero_int(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_idx(int *p, const int *idx, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[idx[i]] = 0;
}

__global__ void k_zero_u64(uint64_t *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_tag_bump(uint32_t *tag, uint32_t inc)
{
    if (threadIdx.x == 0) *tag += inc;
}

__global__ void k_tag_set(uint32_t *tag, uint32_t v)
{
    if (threadIdx.x == 0) *tag = v;
}

// Exchange buffer of at least `bytes` for a factorization that uses `ntags`
// packet tags.  Tags come from the device counter ctx->d_tag, which every launch
// chunk bumps by n (so replays of a captured factorization get fresh tags); the
// host tracks the same count.  The buffer is cleared, and the counter reset,
// when the buffer is (re)allocated and before the counter could wrap, so no
// stale packet can carry a current tag.  A reallocation drops the captured
// factorizations, which hold the old buffer.
int prepare_xbuf(wm_ctx *ctx, size_t bytes, uint64_t ntags, hipStream_t s)
{
    if (ntags > 0x40000000u) {
        set_error("getrf: %llu panel tags in one call", (unsigned long long)ntags);
        return WM_ERR_ARG;
    }
    const size_t had = ctx->ws_lu.bytes;
    WM_TRY(ctx->ws_lu.ensure(bytes));
    const bool fresh = ctx->ws_lu.bytes != had;
    if (fresh) drop_graphs(ctx);
    if (fresh || ctx->panel_tag + ntags + NB > 0xFFFFFFFFull) {
        const size_t nw = ctx->ws_lu.bytes / sizeof(uint64_t);
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s,
                           (uint64_t *)ctx->ws_lu.ptr, nw);
        hipLaunchKernelGGL(k_tag_set, dim3(1), dim3(64), 0, s, ctx->d_tag, 0u);
        ctx->panel_tag = 0;
    }
    return WM_OK;
}

template <typename T, int TPR>
int occupancy_panel(int *blocks_per_cu)
{
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_panel<T, TPR>, 256, 0);
    if (e != hipSuccess) {
        set_error("occupancy query failed: %s", hipGetErrorString(e));
        return WM_ERR_HIP;
    }
    return WM_OK;
}

// Launch plan of a batched factorization.  All members step through the
// panels together; a panel launch holds as many members as keep every panel
// cluster co-resident at that panel's height (resident / P workgroups), so
// tall panels go in several launches and short ones in one, and each member
// has its own exchange buffer.
struct GetrfPlan {
    int pmax = 0, resident = 0, per_launch = 0;
    size_t xbytes = 0;
};

template <typename T, int TPR>
int getrf_plan(wm_ctx *ctx, int n, int batch, GetrfPlan *pl)
{
    using V = typename CT<T>::V;
    constexpr int RW = 256 / TPR;
    const int pmax = (n + RW - 1) / RW;
    int bpc = 0;
    WM_TRY((occupancy_panel<T, TPR>(&bpc)));
    // panel workgroups of this context that may be resident at once: the
    // occupancy, shared with the other contexts that factor on the device at
    // the same time (panel_share), so that every cluster of every concurrent
    // panel launch fits on the device together
    int resident = bpc * ctx->num_cu / std::max(1, ctx->panel_share);
    if (ctx->panel_resident_cap > 0 && ctx->panel_resident_cap < resident) resident = ctx->panel_resident_cap;
    if (pmax > resident || bpc <= 0) {
        set_error("getrf: panel cluster of %d workgroups exceeds %d resident", pmax, resident);
        return WM_ERR_ARG;
    }
    pl->pmax = pmax;
    pl->resident = resident;
    pl->per_launch = std::min(batch, resident / pmax);
    pl->xbytes = (size_t)batch * XLayout<V>(pmax).mat_bytes;
    return WM_OK;
}

// Enqueues the launches of a planned factorization on s (kernel launches only,
// so the sequence can be captured into a graph).
template <typename T, int TPR>
int getrf_impl(wm_ctx *ctx, const GetrfPlan &pl, int n, typename CT<T>::V *A, int lda, int64_t strideA, int *ipiv,
               int64_t strideP, int batch, int *d_info, const int *d_bidx, hipStream_t s, int gk, int kSub,
               const int *d_nmem)
{
    using V = typename CT<T>::V;
    constexpr int RW = 256 / TPR;
    constexpr int CB = (sizeof(T) == 8) ? 16 : 32;
    const int pmax = pl.pmax;
    const XLayout<V> X(pmax);
    char *xbase = (char *)ctx->ws_lu.ptr;
    const double esz = 2.0 * sizeof(T);
    const int kPanelClass = sizeof(T) == 8 ? K_PANEL : K_PANEL32;
    const int kSwapClass = sizeof(T) == 8 ? K_SWAP : K_SWAP32;
    const int kGemmClass = sizeof(T) == 8 ? K_GEMM : gk == 2 ? K_GEMMB3 : K_GEMM32;

    // One launch group: m matrices factored together on one stream.
    struct Grp {
        V *Ab;
        int *ipb, *infob;
        const int *bidx;
        int m;
        char *xbuf;
        hipStream_t st;
    };
    // panel at column c for every member, in launches of as many members as
    // keep their clusters co-resident; members b0.. of a launch use exchange
    // buffers b0.. (the interchange kernel reads member b's permutation there)
    auto panel = [&](const Grp &G, int c) {
        const int P = (n - c + RW - 1) / RW;
        const int per = std::max(1, std::min(G.m, pl.resident / P));
        const double mr = n - c;
        for (int b0 = 0; b0 < G.m; b0 += per) {
            const int m = std::min(per, G.m - b0);
            PanelArgs<T> pa;
            pa.A = G.bidx ? G.Ab : G.Ab + (size_t)b0 * strideA;
            pa.lda = lda; pa.strideA = strideA; pa.n = n; pa.k = c;
            pa.P = P; pa.pmax = pmax;
            pa.ipiv = G.bidx ? G.ipb : G.ipb + (size_t)b0 * strideP;
            pa.strideP = strideP;
            pa.info = G.bidx ? G.infob : G.infob + b0;
            pa.err = ctx->d_err;
            pa.xbuf = G.xbuf + (size_t)b0 * X.mat_bytes;
            pa.bidx = G.bidx ? G.bidx + b0 : nullptr;
            pa.tagp = ctx->d_tag; pa.toff = (uint32_t)c;
            pa.nmem = d_nmem;
            // per column j, (mr-j) scalings and (mr-j)(NB-j-1) complex multiply-adds
            prof_begin(ctx, kPanelClass, G.st);
            hipLaunchKernelGGL((k_panel<T, TPR>), dim3(P, m), dim3(256), 0, G.st, pa);
            prof_end(ctx, kPanelClass, G.st, m * 8.0 * (mr * NB * NB / 2.0 - (double)NB * NB * NB / 6.0),
                     m * 2.0 * mr * NB * esz);
        }
    };
    // interchanges of the panel at c left of c; solve for [tr0, tr1); interchanges only for [tr1, n)
    auto swap = [&](const Grp &G, int c, int tr1, int do_swap, int tr0) {
        SwapArgs sa;
        sa.n = n; sa.k = c; sa.pmax = pmax; sa.do_swap = do_swap;
        sa.nmem = d_nmem;
        sa.ntl = do_swap ? (c + CB - 1) / CB : 0;
        sa.c_tr0 = tr0; sa.c_tr1 = tr1;
        sa.ntr = tr1 > tr0 ? (tr1 - tr0 + CB - 1) / CB : 0;
        sa.c_sw0 = tr1;
        const int nsw = do_swap && n > tr1 ? (n - tr1 + CB - 1) / CB : 0;
        if (sa.ntl + sa.ntr + nsw == 0) return;
        prof_begin(ctx, kSwapClass, G.st);
        hipLaunchKernelGGL((k_swap_trsm<T, CB>), dim3(sa.ntl + sa.ntr + nsw, G.m), dim3(256), 0, G.st, G.Ab, lda,
                           strideA, sa, G.xbuf, G.bidx);
        prof_end(ctx, kSwapClass, G.st, G.m * 4.0 * NB * NB * (tr1 - tr0), G.m * 4.0 * NB * n * esz);
    };
    auto gemm = [&](const Grp &G, int r0, int c0, int mrows, int ncols, int kk, int K) {
        if (mrows <= 0 || ncols <= 0) return;
        const GemmArgs g{r0, c0, mrows, ncols, kk, K, d_nmem};
        // C read and written, L and U read once; flops of the kernel's algorithm:
        // three real products per complex multiply-add (6 flops) for complex128
        // and complex64, four FP32-equivalent real products (8 flops, each taken
        // as three bf16 products) for split-BF16x3
        const double fpc = gk == 2 ? 8.0 : 6.0;
        prof_begin(ctx, kGemmClass, G.st);
        launch_gemm(gk, G.Ab, lda, strideA, g, G.m, G.bidx, G.st);
        prof_end(ctx, kGemmClass, G.st, G.m * fpc * mrows * ncols * K,
                 G.m * (2.0 * mrows * ncols + (double)(mrows + ncols) * K) * esz);
    };

    {
        Grp G;
        G.m = batch;
        G.Ab = A;
        G.ipb = ipiv;
        G.infob = d_info;
        G.bidx = d_bidx;
        G.xbuf = xbase;
        G.st = s;
        hipLaunchKernelGGL(k_tag_bump, dim3(1), dim3(64), 0, s, ctx->d_tag, (uint32_t)n);
        // Outer blocks of kSub panels: inside a block the panels are applied to
        // the block's own columns at once, while the columns to the right get
        // their interchanges at once and their update from the whole block in
        // one product with K = kSub * NB (rows of the later panels' solves are
        // updated first by a short product).
        for (int k = 0; k < n; k += kSub * NB) {
            const int kend = (n - k < kSub * NB) ? n : k + kSub * NB;
            for (int c = k; c < kend; c += NB) {
                if (c > k) gemm(G, c, c, n - c, kend - c, c - NB, NB);
                panel(G, c);
                if (c == k) {
                    swap(G, c, n, 1, c + NB);
                } else {
                    swap(G, c, kend, 1, c + NB);
                    if (kend < n) {
                        gemm(G, c, kend, NB, n - kend, k, c - k);
                        swap(G, c, n, 0, kend);
                    }
                }
            }
            gemm(G, kend, kend, n - kend, n - kend, k, kend - k);
        }
        WM_HIP_TRY(hipGetLastError());
    }
    return WM_OK;
}

}  // namespace

int getrf_block(const wm_ctx *ctx, int prec, int n)
{
    const int dflt = n >= kSubLargeN ? kSubLarge : kSubDefault;
    if (prec < 0 || prec > 2) return dflt;
    const int sc = std::min(n / 1024, wm_ctx::kTuneSizes - 1);
    const int v = ctx->getrf_block[prec][sc];
    return v > 0 ? v : dflt;
}

int getrf_width(wm_ctx *ctx, int prec, int n)
{
    int bpc = 0;
    const bool f64 = prec == WM_PREC_F64;  // complex64 factors (FP32, BF16X3) share the FP32 panel
    if ((f64 ? occupancy_panel<double, 2>(&bpc) : occupancy_panel<float, kTprF32>(&bpc)) != WM_OK || bpc <= 0) return 1;
    const int RW = f64 ? 128 : 256 / kTprF32;
    const int pmax = (n + RW - 1) / RW;
    return std::max(1, bpc * ctx->num_cu / std::max(1, ctx->panel_share) / pmax);
}

namespace {

// Enqueues a whole factorization (info reset, tag bumps, all launch chunks).
template <typename T, int TPR>
int getrf_enqueue(wm_ctx *ctx, const GetrfPlan &pl, int n, void *A, int lda, int64_t strideA, int *ipiv,
                  int64_t strideP, int batch, int *d_info, const int *d_bidx, hipStream_t s, int gk, int ks,
                  const int *d_nmem)
{
    if (!d_bidx) hipLaunchKernelGGL(k_zero_int, dim3((batch + 255) / 256), dim3(256), 0, s, d_info, batch);
    else hipLaunchKernelGGL(k_zero_idx, dim3((batch + 255) / 256), dim3(256), 0, s, d_info, d_bidx, batch);
    return getrf_impl<T, TPR>(ctx, pl, n, (typename CT<T>::V *)A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx,
                              s, gk, ks, d_nmem);
}

template <typename T, int TPR>
int getrf_run(wm_ctx *ctx, int prec, int n, void *A, int lda, int64_t strideA, int *ipiv, int64_t strideP, int batch,
              int *d_info, const int *d_bidx, hipStream_t s, int gk, const int *d_nmem)
{
    const int ks = getrf_block(ctx, prec, n);
    GetrfPlan pl;
    WM_TRY((getrf_plan<T, TPR>(ctx, n, batch, &pl)));
    const uint64_t ntags = (uint64_t)n;
    WM_TRY(prepare_xbuf(ctx, pl.xbytes, ntags, s));
    if (!ctx->use_graphs || ctx->prof.on) {
        WM_TRY((getrf_enqueue<T, TPR>(ctx, pl, n, A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx, s, gk, ks,
                                      d_nmem)));
    } else {
        // Replay a captured factorization of the same shape and buffers, or
        // capture one.  Its kernels read their tags from the device counter, so
        // every replay uses fresh tags.
        const GraphKey key{(int64_t)prec, (int64_t)n, (int64_t)lda, strideA, (int64_t)(intptr_t)A,
                           (int64_t)(intptr_t)ipiv, strideP, (int64_t)batch, (int64_t)(intptr_t)d_info,
                           (int64_t)(intptr_t)d_bidx, (int64_t)ks, (int64_t)(intptr_t)ctx->ws_lu.ptr,
                           (int64_t)(intptr_t)d_nmem};
        hipGraphExec_t exec = find_graph(ctx, key);
        if (!exec) {
            WM_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = getrf_enqueue<T, TPR>(ctx, pl, n, A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx, s,
                                                 gk, ks, d_nmem);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &graph);
            if (rc != WM_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            WM_HIP_TRY(ec);
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            WM_HIP_TRY(ei);
            keep_graph(ctx, key, exec);
        }
        WM_HIP_TRY(hipGraphLaunch(exec, s));
    }
    ctx->panel_tag += ntags;
    return WM_OK;
}

}  // namespace

// Destroys the captured factorizations.  A replay may still be running on the
// stream it was launched on (nothing waits for it before the exchange buffer
// grows or the cache fills), and HIP does not promise that an executable graph
// outlives its destruction while in flight, so the device is drained first.
void drop_graphs(wm_ctx *ctx)
{
    if (ctx->graphs.empty()) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
}

hipGraphExec_t find_graph(wm_ctx *ctx, const GraphKey &key)
{
    auto it = ctx->graphs.find(key);
    return it == ctx->graphs.end() ? nullptr : it->second;
}

void keep_graph(wm_ctx *ctx, const GraphKey &key, hipGraphExec_t exec)
{
    if (ctx->graphs.size() >= wm_ctx::kGraphCap) drop_graphs(ctx);
    ctx->graphs[key] = exec;
}

int getrf_batched(wm_ctx *ctx, int prec, int n, void *A, int lda, int64_t strideA, int *ipiv, int64_t strideP,
                  int batch, int *d_info, hipStream_t s, const int *d_bidx, const int *d_nmem)
{
    if (n <= 0 || n % NB != 0 || lda < n || batch <= 0 || !A || !ipiv || !d_info) {
        set_error("getrf: bad arguments (n=%d must be a positive multiple of %d, lda=%d, batch=%d)", n, NB, lda,
                  batch);
        return WM_ERR_ARG;
    }
    if (prec == WM_PREC_F64)
        return getrf_run<double, 2>(ctx, prec, n, A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx, s, 8, d_nmem);
    if (prec == WM_PREC_F32)
        return getrf_run<float, kTprF32>(ctx, prec, n, A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx, s, 4,
                                         d_nmem);
    if (prec == WM_PREC_BF16X3)
        return getrf_run<float, kTprF32>(ctx, prec, n, A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx, s, 2,
                                         d_nmem);
    set_error("getrf: precision %d not available", prec);
    return WM_ERR_ARG;
}

}  // namespace wm

// ---------------------------------------------------------------------------
// Internal test hooks (not part of include/wiremom.h): run one step of the
// blocked factorization on a device matrix so tests can check each kernel
// against numpy.  A is n x n (lda = n), complex128 for prec 0, complex64 else.
// C -= L U with C = A[K.., K..], L = A[K.., 0..K), U = A[0..K, K..] (no factorization).
// Caps the panel workgroups taken as co-resident (0 restores the occupancy
// query's count), so small factorizations split their panel launches as large
// ones do; drops captured factorizations, whose launch split it changes.
extern "C" int wmx_set_panel_resident(wm_ctx *ctx, int cap)
{
    if (!ctx || cap < 0) return WM_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->panel_resident_cap = cap;
    wm::drop_graphs(ctx);
    return WM_OK;
}

// Factors the members idx[0..count) of a batch through an index table (as the
// refinement levels do); info[m] receives member m's info word.
extern "C" int wmx_getrf_batched_idx(wm_ctx *ctx, int prec, int n, void *A, int lda, int64_t strideA, int *ipiv,
                                     int64_t strideP, int count, int *info, const int *idx)
{
    using namespace wm;
    if (!ctx || count <= 0 || !idx || !info) return WM_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    int nm = 0;
    for (int i = 0; i < count; ++i) nm = std::max(nm, idx[i] + 1);
    int *d = nullptr;
    WM_HIP_TRY(hipMalloc(&d, sizeof(int) * (size_t)(count + nm)));
    int rc = copy_h2d(d, idx, sizeof(int) * count, ctx->stream);
    if (rc == WM_OK) rc = getrf_batched(ctx, prec, n, A, lda, strideA, ipiv, strideP, count, d + count, ctx->stream, d);
    if (rc == WM_OK) rc = check_device_error(ctx, ctx->stream);
    std::vector<int> h(nm, 0);
    if (rc == WM_OK) rc = copy_d2h(h.data(), d + count, sizeof(int) * nm, ctx->stream);
    for (int i = 0; i < count && rc == WM_OK; ++i) info[i] = h[idx[i]];
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    return rc;
}

extern "C" int wmx_gemm_k(wm_ctx *ctx, int prec, int n, void *A, int K, int reps)
{
    using namespace wm;
    if (K <= 0 || K >= n || K % NB || reps <= 0) return WM_ERR_ARG;
    const GemmArgs g{K, K, n - K, n - K, 0, K, nullptr};
    for (int r = 0; r < reps; ++r)
        launch_gemm(prec == WM_PREC_F64 ? 8 : prec == WM_PREC_BF16X3 ? 2 : 4, A, n, 0, g, 1, nullptr, ctx->stream);
    WM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return WM_OK;
}

extern "C" int wmx_gemm_step(wm_ctx *ctx, int prec, int n, void *A, int k)
{
    using namespace wm;
    const int m2 = n - k - NB;
    if (m2 <= 0 || n % NB || k % NB) return WM_ERR_ARG;
    const GemmArgs g{k + NB, k + NB, m2, m2, k, NB, nullptr};
    launch_gemm(prec == WM_PREC_F64 ? 8 : prec == WM_PREC_BF16X3 ? 2 : 4, A, n, 0, g, 1, nullptr, ctx->stream);
    WM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return WM_OK;
}

// One panel factorization at column k followed by the interchanges and U12.
extern "C" int wmx_panel_step(wm_ctx *ctx, int prec, int n, void *A, int *ipiv, int k, int do_swap)
{
    using namespace wm;
    if (n % NB || k % NB || k + NB > n) return WM_ERR_ARG;
    int *info = nullptr;
    WM_HIP_TRY(hipMalloc(&info, sizeof(int)));
    WM_HIP_TRY(hipMemset(info, 0, sizeof(int)));
    auto run = [&](auto tag, auto *Ap, int TPRv) -> int {
        using T = decltype(tag);
        using V = typename CT<T>::V;
        const int RW = 256 / TPRv;
        const int pmax = (n + RW - 1) / RW;
        const XLayout<V> X(pmax);
        WM_TRY(prepare_xbuf(ctx, X.mat_bytes, (uint64_t)n, ctx->stream));
        char *xbuf = (char *)ctx->ws_lu.ptr;
        hipLaunchKernelGGL(k_tag_bump, dim3(1), dim3(64), 0, ctx->stream, ctx->d_tag, (uint32_t)n);
        ctx->panel_tag += (uint64_t)n;
        PanelArgs<T> pa;
        pa.A = (V *)Ap; pa.lda = n; pa.strideA = 0; pa.n = n; pa.k = k;
        pa.P = (n - k + RW - 1) / RW; pa.pmax = pmax;
        pa.ipiv = ipiv; pa.strideP = 0; pa.info = info;
        pa.err = ctx->d_err; pa.xbuf = xbuf; pa.bidx = nullptr; pa.tagp = ctx->d_tag; pa.toff = (uint32_t)k;
        pa.nmem = nullptr;
        constexpr int CB = (sizeof(T) == 8) ? 16 : 32;
        if (sizeof(T) == 8)
            hipLaunchKernelGGL((k_panel<double, 2>), dim3(pa.P, 1), dim3(256), 0, ctx->stream, *(PanelArgs<double> *)&pa);
        else
            hipLaunchKernelGGL((k_panel<float, kTprF32>), dim3(pa.P, 1), dim3(256), 0, ctx->stream, *(PanelArgs<float> *)&pa);
        if (do_swap) {
            SwapArgs sa;
            sa.n = n; sa.k = k; sa.pmax = pmax; sa.do_swap = 1; sa.nmem = nullptr;
            sa.ntl = (k + CB - 1) / CB;
            sa.ntr = (n - k - NB + CB - 1) / CB;
            sa.c_tr0 = k + NB; sa.c_tr1 = n; sa.c_sw0 = n;
            if (sa.ntl + sa.ntr > 0)
                hipLaunchKernelGGL((k_swap_trsm<T, CB>), dim3(sa.ntl + sa.ntr, 1), dim3(256), 0, ctx->stream, (V *)Ap,
                                   n, (int64_t)0, sa, xbuf, (const int *)nullptr);
        }
        return WM_OK;
    };
    int rc = (prec == WM_PREC_F64) ? run(double(0), (double2 *)A, 2) : run(float(0), (float2 *)A, kTprF32);
    WM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    (void)hipFree(info);
    if (rc != WM_OK) return rc;
    return check_device_error(ctx, ctx->stream);
}
