// Batched dense complex LU with partial pivoting for gfx950.
//
// Replaces the reference's LAPACKE_zgetrf ("wersja referencyjna/Wire_MoM.cu":372,
// MKL) and cublasZgetrfBatched ("wersja gpu/Wire_MoM.cu":468).  Same contract
// as LAPACK zgetrf: column-major, pivot = first row with the largest
// |re|+|im| in the column (izamax), 1-based ipiv, info = first exactly-zero
// pivot, rows interchanged across the whole matrix.
//
// Blocked right-looking algorithm with panel width NB = 64.  Per panel k:
//   1. k_panel      : GEPP of the n-k x NB panel.  A cluster of P workgroups
//                     per matrix holds the panel in registers (one row per
//                     thread, or two threads per row for complex128); per
//                     column the cluster agrees on the pivot through one
//                     bounded cross-workgroup barrier, swaps and applies the
//                     rank-1 update locally.  It also emits the composite row
//                     permutation of the panel.
//   2. k_swap_trsm  : applies that permutation to every column outside the
//                     panel and computes U12 = L11^{-1} A12 (forward
//                     substitution, L11 from LDS).
//   3. k_gemm_*     : A22 -= L21 U12 on the matrix cores: complex128 with
//                     v_mfma_f64_16x16x4f64, complex64 with
//                     v_mfma_f32_32x32x2f32, four real products per complex
//                     product.  The MFMA is issued on the transposed problem
//                     (C^T = U12^T L21^T) so each lane's accumulator column is
//                     a run of consecutive rows of C: the C read-modify-write is
//                     coalesced in column-major storage.
#include "wm_cplx.h"
#include "wm_internal.h"

namespace wm {

namespace {

constexpr int NB = 64;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// exchange buffers for the panel: per matrix, two parities of
//   slots [Pmax] {double v; int row; int wg}, cand [Pmax][NB], diag [NB]
// followed by the panel permutation: cnt, rows[2NB], src[2NB].
struct Slot {
    double v;
    int row;
    int pad;
};

template <typename V>
struct XLayout {
    size_t par_bytes, perm_off, mat_bytes;
    int pmax;
    __host__ __device__ XLayout(int pmax_) : pmax(pmax_)
    {
        par_bytes = (size_t)pmax * sizeof(Slot) + (size_t)pmax * NB * sizeof(V) + NB * sizeof(V);
        par_bytes = (par_bytes + 255) & ~(size_t)255;
        perm_off = 2 * par_bytes;
        mat_bytes = perm_off + (1 + 4 * NB) * sizeof(int);
        mat_bytes = (mat_bytes + 255) & ~(size_t)255;
    }
    __device__ Slot *slots(char *m, int par) const { return (Slot *)(m + par * par_bytes); }
    __device__ V *cand(char *m, int par) const { return (V *)(m + par * par_bytes + (size_t)pmax * sizeof(Slot)); }
    __device__ V *diag(char *m, int par) const
    {
        return (V *)(m + par * par_bytes + (size_t)pmax * sizeof(Slot) + (size_t)pmax * NB * sizeof(V));
    }
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
    int *cnt;
    int *err;
    char *xbuf;
    const int *bidx;  // launch member -> matrix index (null: identity)
};

// Argmax combine: larger value wins, equal values go to the lower row.
__device__ inline bool better(double v2, int r2, double v1, int r1) { return v2 > v1 || (v2 == v1 && r2 < r1); }

template <typename T, int TPR>
__global__ __launch_bounds__(256) void k_panel(PanelArgs<T> a)
{
    using V = typename CT<T>::V;
    constexpr int W = NB / TPR;    // columns per thread
    constexpr int RPW = 64 / TPR;  // rows per wave
    constexpr int RW = 4 * RPW;    // rows per workgroup
    const int bl = blockIdx.y, w = blockIdx.x;          // bl: member of this launch
    const int b = a.bidx ? a.bidx[bl] : bl;              // b: matrix index
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int chunk = (TPR == 1) ? 0 : lane / RPW;
    const int row = a.k + w * RW + wave * RPW + (lane % RPW);
    const bool valid = row < a.n;
    const int k = a.k;
    V *Ab = a.A + (size_t)b * a.strideA;
    const XLayout<V> X(a.pmax);
    char *xm = a.xbuf + (size_t)bl * X.mat_bytes;

    T re[W], im[W];
#pragma unroll
    for (int i = 0; i < W; ++i) {
        if (valid) {
            const V v = Ab[row + (size_t)(k + chunk * W + i) * a.lda];
            re[i] = v.x;
            im[i] = v.y;
        } else {
            re[i] = 0;
            im[i] = 0;
        }
    }

    __shared__ double s_wv[4];
    __shared__ int s_wr[4];
    __shared__ V s_u[NB];
    __shared__ V s_d[NB];
    __shared__ double s_gv;
    __shared__ int s_gr, s_gw;
    __shared__ double2 s_rinv;

    for (int j = 0; j < NB; ++j) {
        const int par = j & 1;
        const int pr = k + j;
        const int cj = j / W;
        // current value of column j in this thread's row (owner chunk only)
        T xr = 0, xi = 0;
#pragma unroll
        for (int i = 0; i < W; ++i)
            if (chunk * W + i == j) { xr = re[i]; xi = im[i]; }
        double v = -1.0;
        int r = 0x7fffffff;
        if (valid && row >= pr && chunk == cj) {
            v = fabs((double)xr) + fabs((double)xi);
            r = row;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double v2 = __shfl_xor(v, off);
            const int r2 = __shfl_xor(r, off);
            if (better(v2, r2, v, r)) { v = v2; r = r2; }
        }
        if (lane == 0) { s_wv[wave] = v; s_wr[wave] = r; }
        __syncthreads();
        double wv = s_wv[0];
        int wr = s_wr[0];
#pragma unroll
        for (int q = 1; q < 4; ++q)
            if (better(s_wv[q], s_wr[q], wv, wr)) { wv = s_wv[q]; wr = s_wr[q]; }

        // publish this workgroup's candidate, its row, and (workgroup 0) row pr
        Slot *slots = X.slots(xm, par);
        V *cand = X.cand(xm, par);
        V *diag = X.diag(xm, par);
        bool wrote = false;
        if (tid == 0) {
            Slot s;
            s.v = wv;
            s.row = wr;
            s.pad = 0;
            slots[w] = s;
            wrote = true;
        }
        if (valid && row == wr) {
#pragma unroll
            for (int i = 0; i < W; ++i) cand[(size_t)w * NB + chunk * W + i] = V{re[i], im[i]};
            wrote = true;
        }
        if (valid && row == pr) {
#pragma unroll
            for (int i = 0; i < W; ++i) diag[chunk * W + i] = V{re[i], im[i]};
            wrote = true;
        }
        if (wrote) __threadfence();
        cluster_barrier(a.cnt + bl, (j + 1) * a.P, a.err);

        if (wave == 0) {
            double gv = -1.0;
            int gr = 0x7fffffff, gw = 0;
            for (int s = lane; s < a.P; s += 64) {
                const double v2 = __hip_atomic_load(&slots[s].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const int r2 = __hip_atomic_load(&slots[s].row, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (better(v2, r2, gv, gr)) { gv = v2; gr = r2; gw = s; }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double v2 = __shfl_xor(gv, off);
                const int r2 = __shfl_xor(gr, off);
                const int w2 = __shfl_xor(gw, off);
                if (better(v2, r2, gv, gr)) { gv = v2; gr = r2; gw = w2; }
            }
            if (lane == 0) { s_gv = gv; s_gr = gr; s_gw = gw; }
        }
        __syncthreads();
        const double gv = s_gv;
        const int gr = s_gr, gw = s_gw;
        const bool own_gr = gr >= k + w * RW && gr < k + (w + 1) * RW;
        if (tid < NB) {
            s_u[tid] = cand[(size_t)gw * NB + tid];
            if (own_gr) s_d[tid] = diag[tid];
        }
        __syncthreads();
        if (tid == 0) {
            const V p = s_u[j];
            s_rinv = (gv != 0.0) ? crecip((double)p.x, (double)p.y) : make_double2(0.0, 0.0);
            if (w == 0) {
                a.ipiv[(size_t)b * a.strideP + pr] = gr + 1;
                if (gv == 0.0 && a.info[b] == 0) a.info[b] = pr + 1;
            }
        }
        __syncthreads();
        if (gr != pr && valid) {
            if (row == gr) {
#pragma unroll
                for (int i = 0; i < W; ++i) { const V d = s_d[chunk * W + i]; re[i] = d.x; im[i] = d.y; }
            } else if (row == pr) {
#pragma unroll
                for (int i = 0; i < W; ++i) { const V u = s_u[chunk * W + i]; re[i] = u.x; im[i] = u.y; }
            }
        }
        // multiplier: column j of this row, held by chunk cj of the row
        T cr = 0, ci = 0;
#pragma unroll
        for (int i = 0; i < W; ++i)
            if (chunk * W + i == j) { cr = re[i]; ci = im[i]; }
        if (TPR > 1) {
            const int src = (lane % RPW) + RPW * cj;
            cr = __shfl(cr, src);
            ci = __shfl(ci, src);
        }
        if (gv != 0.0 && valid && row > pr) {
            const double2 ri = s_rinv;
            const T lr = (T)((double)cr * ri.x - (double)ci * ri.y);
            const T li = (T)((double)cr * ri.y + (double)ci * ri.x);
#pragma unroll
            for (int i = 0; i < W; ++i) {
                const int col = chunk * W + i;
                if (col == j) { re[i] = lr; im[i] = li; }
                if (col > j) {
                    const V u = s_u[col];
                    re[i] -= lr * u.x - li * u.y;
                    im[i] -= lr * u.y + li * u.x;
                }
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < W; ++i)
        if (valid) Ab[row + (size_t)(k + chunk * W + i) * a.lda] = V{re[i], im[i]};

    // composite permutation of the panel's interchanges, for the other columns:
    // after applying them in order, row rows[t] holds the old content of src[t].
    if (w == 0 && tid == 0) {
        int *pm = X.perm(xm);
        int *rows = pm + 1, *src = pm + 1 + 2 * NB;
        int cnt = NB;
        for (int t = 0; t < NB; ++t) { rows[t] = k + t; src[t] = k + t; }
        for (int j = 0; j < NB; ++j) {
            const int p = a.ipiv[(size_t)b * a.strideP + k + j] - 1;
            if (p == k + j) continue;
            int tp = -1;
            if (p < k + NB) tp = p - k;
            else {
                for (int t = NB; t < cnt; ++t)
                    if (rows[t] == p) { tp = t; break; }
                if (tp < 0) { tp = cnt++; rows[tp] = p; src[tp] = p; }
            }
            const int tmp = src[j];
            src[j] = src[tp];
            src[tp] = tmp;
        }
        pm[0] = cnt;
    }
}

// ---------------------------------------------------------------------------
// Row interchanges outside the panel and U12 = L11^{-1} A12.
template <typename T, int CB>
__global__ __launch_bounds__(256) void k_swap_trsm(typename CT<T>::V *A, int lda, int64_t strideA, int n, int k,
                                                   int pmax, char *xbuf, int ntile_left, int *cnt, const int *bidx)
{
    using V = typename CT<T>::V;
    constexpr int RG = 256 / CB;        // row groups for the triangular solve
    constexpr int RPT = NB / RG;        // rows per thread in the solve
    constexpr int EPT = 2 * NB * CB / 256;  // permuted elements per thread (max)
    const int bl = blockIdx.y, tile = blockIdx.x, tid = threadIdx.x;
    const int b = bidx ? bidx[bl] : bl;
    V *Ab = A + (size_t)b * strideA;
    const XLayout<V> X(pmax);
    char *xm = xbuf + (size_t)bl * X.mat_bytes;
    if (tile == 0 && tid == 0) cnt[bl] = 0;  // panel barrier counter, for the next panel

    const bool left = tile < ntile_left;
    const int c0 = left ? tile * CB : k + NB + (tile - ntile_left) * CB;
    const int cend = left ? k : n;

    __shared__ int s_rows[2 * NB], s_src[2 * NB];
    __shared__ int s_cnt;
    const int *pm = X.perm(xm);
    if (tid == 0) s_cnt = pm[0];
    if (tid < 2 * NB) { s_rows[tid] = pm[1 + tid]; s_src[tid] = pm[1 + 2 * NB + tid]; }
    __syncthreads();
    const int tcnt = s_cnt;

    // gather the old contents of the touched rows, keep the new top block
    // (rows k..k+NB-1) of right-hand columns in LDS, then scatter
    __shared__ V s_top[NB][CB];
    V val[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int idx = tid + 256 * e;
        const int t = idx / CB, cc = idx % CB, col = c0 + cc;
        if (t < tcnt && col < cend) {
            val[e] = Ab[s_src[t] + (size_t)col * lda];
            if (!left && t < NB) s_top[t][cc] = val[e];
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int idx = tid + 256 * e;
        const int t = idx / CB, col = c0 + idx % CB;
        if (t < tcnt && col < cend) Ab[s_rows[t] + (size_t)col * lda] = val[e];
    }
    if (left) return;

    // forward substitution with the unit lower L11 (strictly lower part in LDS)
    __shared__ V s_l[NB * (NB - 1) / 2];
    __shared__ V s_x[2][CB];
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e % NB, jj = e / NB;
        if (i > jj) s_l[i * (i - 1) / 2 + jj] = Ab[(k + i) + (size_t)(k + jj) * lda];
    }
    __syncthreads();
    const int c = tid % CB, rg = tid / CB;
    const bool cok = (c0 + c) < n;
    T xr[RPT], xi[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const V v = cok ? s_top[rg + q * RG][c] : V{0, 0};
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
// Trailing update  C (M x M) -= A (M x NB) * B (NB x M)  on the matrix cores.
//
// Both kernels issue the MFMA on the transposed problem, C^T = B^T A^T:
// MFMA rows are columns of C and MFMA columns are rows of C, so each
// accumulator register of a lane maps to a run of consecutive rows of one
// column of C and the read-modify-write of C is coalesced.  Complex products
// take four real MFMAs (re*re - im*im, re*im + im*re).  K = NB = 64 is walked
// in chunks of 16 staged through LDS; the next chunk is fetched from global
// memory into registers while the current one is multiplied.
//
// Tiles are ordered column by column and dealt to the XCDs in contiguous runs
// (workgroup ids go to the 8 XCDs round-robin), so tiles that share a block
// column of U12 share an L2.

// Tile of workgroup g out of T, grouped per XCD; >= T means "no tile".
__device__ inline int xcd_tile(int g, int T)
{
    const int per = (T + 7) / 8;
    return (g % 8) * per + g / 8;
}

inline int xcd_grid(int T) { return 8 * ((T + 7) / 8); }

// complex128: workgroup tile 128 (rows) x 64 (columns), four waves of 64 x 32,
// v_mfma_f64_16x16x4f64.  Lane l holds srcA[l%16][l/16], srcB[l/16][l%16] and
// D[l/16 + 4*v][l%16] (measured on gfx950).
constexpr int GD_BM = 128, GD_BN = 64, GD_KC = 16;

__global__ __launch_bounds__(256, 2) void k_gemm_f64(double2 *Abase, int lda, int64_t strideA, int n, int k,
                                                     const int *bidx, int TI, int TJ)
{
    const int t = xcd_tile(blockIdx.x, TI * TJ);
    if (t >= TI * TJ) return;
    const int b = bidx ? bidx[blockIdx.z] : (int)blockIdx.z;
    double2 *M = Abase + (size_t)b * strideA;
    const int r0 = k + NB, mrows = n - r0;
    const int i0 = (t % TI) * GD_BM, j0 = (t / TI) * GD_BN;
    const double2 *A = M + r0 + (size_t)k * lda;  // L21
    const double2 *B = M + k + (size_t)r0 * lda;  // U12
    double2 *C = M + r0 + (size_t)r0 * lda;       // A22
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 64, wj = (wave >> 1) * 32;

    __shared__ double2 sA[GD_KC][GD_BM + 1];
    __shared__ double2 sB[GD_KC][GD_BN + 1];
    double2 pa[8], pb[4];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid + 256 * e, i = idx % GD_BM, kk = idx / GD_BM, gi = i0 + i;
            pa[e] = gi < mrows ? A[gi + (size_t)(kc + kk) * lda] : double2{0.0, 0.0};
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid + 256 * e, kk = idx % GD_KC, j = idx / GD_KC, gj = j0 + j;
            pb[e] = gj < mrows ? B[(kc + kk) + (size_t)gj * lda] : double2{0.0, 0.0};
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid + 256 * e;
            sA[idx / GD_BM][idx % GD_BM] = pa[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid + 256 * e;
            sB[idx % GD_KC][idx / GD_KC] = pb[e];
        }
    };
    v4d accR[2][4], accI[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) { accR[x][y] = v4d{0, 0, 0, 0}; accI[x][y] = v4d{0, 0, 0, 0}; }

    fetch(0);
    stash();
    __syncthreads();
    for (int kc = 0; kc < NB; kc += GD_KC) {
        if (kc + GD_KC < NB) fetch(kc + GD_KC);
#pragma unroll
        for (int ks = 0; ks < GD_KC; ks += 4) {
            const int kk = ks + lane / 16;
            double2 bv[2], av[4];
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) bv[tj] = sB[kk][wj + tj * 16 + lane % 16];
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) av[ti] = sA[kk][wi + ti * 16 + lane % 16];
            // four passes over the tiles so that consecutive MFMAs never share an accumulator
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 4; ++ti)
                    accR[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[tj].x, av[ti].x, accR[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 4; ++ti)
                    accI[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[tj].x, av[ti].y, accI[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 4; ++ti)
                    accR[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bv[tj].y, av[ti].y, accR[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 4; ++ti)
                    accI[tj][ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[tj].y, av[ti].x, accI[tj][ti], 0, 0, 0);
        }
        if (kc + GD_KC < NB) {
            __syncthreads();
            stash();
            __syncthreads();
        }
    }
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = i0 + wi + ti * 16 + lane % 16;
                const int gj = j0 + wj + tj * 16 + (lane / 16) + 4 * v;
                if (gi < mrows && gj < mrows) {
                    double2 c = C[gi + (size_t)gj * lda];
                    c.x -= accR[tj][ti][v];
                    c.y -= accI[tj][ti][v];
                    C[gi + (size_t)gj * lda] = c;
                }
            }
}

// complex64: workgroup tile 128 x 128, four waves of 64 x 64 (2 x 2 MFMA
// tiles), v_mfma_f32_32x32x2f32.  Lane l holds srcA[l%32][l/32],
// srcB[l/32][l%32] and D[8*(v/4)+4*(l/32)+v%4][l%32].
constexpr int GS_BM = 128, GS_BN = 128, GS_KC = 16;

__global__ __launch_bounds__(256, 2) void k_gemm_f32(float2 *Abase, int lda, int64_t strideA, int n, int k,
                                                     const int *bidx, int TI, int TJ)
{
    const int t = xcd_tile(blockIdx.x, TI * TJ);
    if (t >= TI * TJ) return;
    const int b = bidx ? bidx[blockIdx.z] : (int)blockIdx.z;
    float2 *M = Abase + (size_t)b * strideA;
    const int r0 = k + NB, mrows = n - r0;
    const int i0 = (t % TI) * GS_BM, j0 = (t / TI) * GS_BN;
    const float2 *A = M + r0 + (size_t)k * lda;
    const float2 *B = M + k + (size_t)r0 * lda;
    float2 *C = M + r0 + (size_t)r0 * lda;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 1) * 64, wj = (wave >> 1) * 64;

    __shared__ float2 sA[GS_KC][GS_BM + 2];
    __shared__ float2 sB[GS_KC][GS_BN + 2];
    float2 pa[8], pb[8];
    // per-thread fixed offsets of the chunk elements it moves
    const int ai = tid % GS_BM, ak = tid / GS_BM;  // A: rows ai, k = ak + 2e
    const int bk = tid % GS_KC, bj = tid / GS_KC;  // B: k = bk, columns bj + 16e
    const bool arow = i0 + ai < mrows;
    auto fetch = [&](int kc) {
        const float2 *pA = A + (i0 + ai) + (size_t)(kc + ak) * lda;
#pragma unroll
        for (int e = 0; e < 8; ++e) pa[e] = arow ? pA[(size_t)(2 * e) * lda] : float2{0.f, 0.f};
        const float2 *pB = B + (kc + bk) + (size_t)(j0 + bj) * lda;
#pragma unroll
        for (int e = 0; e < 8; ++e) pb[e] = (j0 + bj + 16 * e < mrows) ? pB[(size_t)(16 * e) * lda] : float2{0.f, 0.f};
    };
    auto stash = [&]() {
#pragma unroll
        for (int e = 0; e < 8; ++e) sA[ak + 2 * e][ai] = pa[e];
#pragma unroll
        for (int e = 0; e < 8; ++e) sB[bk][bj + 16 * e] = pb[e];
    };
    v16f accR[2][2], accI[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) { accR[x][y][v] = 0.f; accI[x][y][v] = 0.f; }

    fetch(0);
    stash();
    __syncthreads();
    for (int kc = 0; kc < NB; kc += GS_KC) {
        if (kc + GS_KC < NB) fetch(kc + GS_KC);
#pragma unroll
        for (int ks = 0; ks < GS_KC; ks += 2) {
            const int kk = ks + lane / 32;
            float2 bv[2], av[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                bv[q] = sB[kk][wj + q * 32 + lane % 32];
                av[q] = sA[kk][wi + q * 32 + lane % 32];
            }
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
                    accR[tj][ti] = __builtin_amdgcn_mfma_f32_32x32x2f32(bv[tj].x, av[ti].x, accR[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
                    accI[tj][ti] = __builtin_amdgcn_mfma_f32_32x32x2f32(bv[tj].x, av[ti].y, accI[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
                    accR[tj][ti] = __builtin_amdgcn_mfma_f32_32x32x2f32(-bv[tj].y, av[ti].y, accR[tj][ti], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
                    accI[tj][ti] = __builtin_amdgcn_mfma_f32_32x32x2f32(bv[tj].y, av[ti].x, accI[tj][ti], 0, 0, 0);
        }
        if (kc + GS_KC < NB) {
            __syncthreads();
            stash();
            __syncthreads();
        }
    }
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int gi = i0 + wi + ti * 32 + lane % 32;
                const int gj = j0 + wj + tj * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
                if (gi < mrows && gj < mrows) {
                    float2 c = C[gi + (size_t)gj * lda];
                    c.x -= accR[tj][ti][v];
                    c.y -= accI[tj][ti][v];
                    C[gi + (size_t)gj * lda] = c;
                }
            }
}

// Launches the trailing update of step k for nb_ matrices.
void launch_gemm(int esz, void *A, int lda, int64_t strideA, int n, int k, int nb_, const int *bidx, hipStream_t s)
{
    const int m2 = n - k - NB;
    if (m2 <= 0) return;
    if (esz == 8) {
        const int TI = (m2 + GD_BM - 1) / GD_BM, TJ = (m2 + GD_BN - 1) / GD_BN;
        hipLaunchKernelGGL(k_gemm_f64, dim3(xcd_grid(TI * TJ), 1, nb_), dim3(256), 0, s, (double2 *)A, lda, strideA,
                           n, k, bidx, TI, TJ);
    } else {
        const int TI = (m2 + GS_BM - 1) / GS_BM, TJ = (m2 + GS_BN - 1) / GS_BN;
        hipLaunchKernelGGL(k_gemm_f32, dim3(xcd_grid(TI * TJ), 1, nb_), dim3(256), 0, s, (float2 *)A, lda, strideA,
                           n, k, bidx, TI, TJ);
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

template <typename T, int TPR>
int getrf_impl(wm_ctx *ctx, int n, typename CT<T>::V *A, int lda, int64_t strideA, int *ipiv, int64_t strideP,
               int batch, int *d_info, const int *d_bidx, hipStream_t s)
{
    using V = typename CT<T>::V;
    constexpr int RW = 256 / TPR;
    constexpr int CB = (sizeof(T) == 8) ? 16 : 32;
    const int pmax = (n + RW - 1) / RW;
    const XLayout<V> X(pmax);
    int bpc = 0;
    WM_TRY((occupancy_panel<T, TPR>(&bpc)));
    const int resident = bpc * ctx->num_cu;
    if (pmax > resident || bpc <= 0) {
        set_error("getrf: panel cluster of %d workgroups exceeds %d resident", pmax, resident);
        return WM_ERR_ARG;
    }
    // matrices per launch so that every cluster is co-resident
    int per_launch = resident / pmax;
    if (per_launch > batch) per_launch = batch;
    if (per_launch > wm_ctx::kSyncCap) per_launch = wm_ctx::kSyncCap;
    WM_TRY(ctx->ws_lu.ensure((size_t)per_launch * X.mat_bytes));
    char *xbuf = (char *)ctx->ws_lu.ptr;

    for (int b0 = 0; b0 < batch; b0 += per_launch) {
        const int nb_ = (batch - b0 < per_launch) ? batch - b0 : per_launch;
        // with an index table the bases stay put and members are looked up
        V *Ab = d_bidx ? A : A + (size_t)b0 * strideA;
        int *ipb = d_bidx ? ipiv : ipiv + (size_t)b0 * strideP;
        int *infob = d_bidx ? d_info : d_info + b0;
        const int *bidx = d_bidx ? d_bidx + b0 : nullptr;
        hipLaunchKernelGGL(k_zero_int, dim3((nb_ + 255) / 256), dim3(256), 0, s, ctx->d_sync, nb_);
        for (int k = 0; k < n; k += NB) {
            PanelArgs<T> pa;
            pa.A = Ab; pa.lda = lda; pa.strideA = strideA; pa.n = n; pa.k = k;
            pa.P = (n - k + RW - 1) / RW; pa.pmax = pmax;
            pa.ipiv = ipb; pa.strideP = strideP; pa.info = infob;
            pa.cnt = ctx->d_sync; pa.err = ctx->d_err; pa.xbuf = xbuf; pa.bidx = bidx;
            const double mr = n - k, m2 = n - k - NB, esz = 2.0 * sizeof(T);
            // panel: per column j, (mr-j) scalings and (mr-j)(NB-j-1) complex multiply-adds
            prof_begin(ctx, K_PANEL, s);
            hipLaunchKernelGGL((k_panel<T, TPR>), dim3(pa.P, nb_), dim3(256), 0, s, pa);
            prof_end(ctx, K_PANEL, s, nb_ * 8.0 * (mr * NB * NB / 2.0 - (double)NB * NB * NB / 6.0),
                     nb_ * 2.0 * mr * NB * esz);
            const int ntl = (k + CB - 1) / CB;
            const int ntr = (n - k - NB + CB - 1) / CB;
            if (ntl + ntr > 0) {
                prof_begin(ctx, K_SWAP, s);
                hipLaunchKernelGGL((k_swap_trsm<T, CB>), dim3(ntl + ntr, nb_), dim3(256), 0, s, Ab, lda, strideA, n,
                                   k, pmax, xbuf, ntl, ctx->d_sync, bidx);
                prof_end(ctx, K_SWAP, s, nb_ * 4.0 * NB * NB * m2, nb_ * 4.0 * NB * n * esz);
            }
            if (m2 > 0) {
                // trailing update: 8 M^2 NB real flops; C read and written, L21 and U12 read once
                prof_begin(ctx, K_GEMM, s);
                launch_gemm((int)sizeof(T), Ab, lda, strideA, n, k, nb_, bidx, s);
                prof_end(ctx, K_GEMM, s, nb_ * 8.0 * m2 * m2 * NB, nb_ * (2.0 * m2 * m2 + 2.0 * m2 * NB) * esz);
            }
        }
        WM_HIP_TRY(hipGetLastError());
    }
    return WM_OK;
}

}  // namespace

int getrf_batched(wm_ctx *ctx, int prec, int n, void *A, int lda, int64_t strideA, int *ipiv, int64_t strideP,
                  int batch, int *d_info, hipStream_t s, const int *d_bidx)
{
    if (n <= 0 || n % NB != 0 || lda < n || batch <= 0 || !A || !ipiv || !d_info) {
        set_error("getrf: bad arguments (n=%d must be a positive multiple of %d, lda=%d, batch=%d)", n, NB, lda,
                  batch);
        return WM_ERR_ARG;
    }
    if (!d_bidx) hipLaunchKernelGGL(k_zero_int, dim3((batch + 255) / 256), dim3(256), 0, s, d_info, batch);
    else hipLaunchKernelGGL(k_zero_idx, dim3((batch + 255) / 256), dim3(256), 0, s, d_info, d_bidx, batch);
    if (prec == WM_PREC_F64)
        return getrf_impl<double, 2>(ctx, n, (double2 *)A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx, s);
    if (prec == WM_PREC_F32)
        return getrf_impl<float, 1>(ctx, n, (float2 *)A, lda, strideA, ipiv, strideP, batch, d_info, d_bidx, s);
    set_error("getrf: precision %d not available", prec);
    return WM_ERR_ARG;
}

}  // namespace wm

// ---------------------------------------------------------------------------
// Internal test hooks (not part of include/wiremom.h): run one step of the
// blocked factorization on a device matrix so tests can check each kernel
// against numpy.  A is n x n (lda = n), complex128 for prec 0, complex64 else.
extern "C" int wmx_gemm_step(wm_ctx *ctx, int prec, int n, void *A, int k)
{
    using namespace wm;
    const int m2 = n - k - NB;
    if (m2 <= 0 || n % NB || k % NB) return WM_ERR_ARG;
    launch_gemm(prec == WM_PREC_F64 ? 8 : 4, A, n, 0, n, k, 1, nullptr, ctx->stream);
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
    WM_HIP_TRY(hipMemset(ctx->d_sync, 0, sizeof(int)));
    auto run = [&](auto tag, auto *Ap, int TPRv) -> int {
        using T = decltype(tag);
        using V = typename CT<T>::V;
        const int RW = 256 / TPRv;
        const int pmax = (n + RW - 1) / RW;
        const XLayout<V> X(pmax);
        WM_TRY(ctx->ws_lu.ensure(X.mat_bytes));
        PanelArgs<T> pa;
        pa.A = (V *)Ap; pa.lda = n; pa.strideA = 0; pa.n = n; pa.k = k;
        pa.P = (n - k + RW - 1) / RW; pa.pmax = pmax;
        pa.ipiv = ipiv; pa.strideP = 0; pa.info = info;
        pa.cnt = ctx->d_sync; pa.err = ctx->d_err; pa.xbuf = (char *)ctx->ws_lu.ptr; pa.bidx = nullptr;
        constexpr int CB = (sizeof(T) == 8) ? 16 : 32;
        if (sizeof(T) == 8)
            hipLaunchKernelGGL((k_panel<double, 2>), dim3(pa.P, 1), dim3(256), 0, ctx->stream, *(PanelArgs<double> *)&pa);
        else
            hipLaunchKernelGGL((k_panel<float, 1>), dim3(pa.P, 1), dim3(256), 0, ctx->stream, *(PanelArgs<float> *)&pa);
        if (do_swap) {
            const int ntl = (k + CB - 1) / CB, ntr = (n - k - NB + CB - 1) / CB;
            if (ntl + ntr > 0)
                hipLaunchKernelGGL((k_swap_trsm<T, CB>), dim3(ntl + ntr, 1), dim3(256), 0, ctx->stream, (V *)Ap, n,
                                   (int64_t)0, n, k, pmax, (char *)ctx->ws_lu.ptr, ntl, ctx->d_sync,
                                   (const int *)nullptr);
        }
        return WM_OK;
    };
    int rc = (prec == WM_PREC_F64) ? run(double(0), (double2 *)A, 2) : run(float(0), (float2 *)A, 1);
    WM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    (void)hipFree(info);
    if (rc != WM_OK) return rc;
    return check_device_error(ctx, ctx->stream);
}
