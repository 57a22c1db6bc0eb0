// Triangular solves with LU factors, FP64 residuals and refinement updates.
//
// Replaces LAPACKE_zgetrs ("wersja referencyjna/Wire_MoM.cu":382) and
// cublasZgetrsBatched ("wersja gpu/Wire_MoM.cu":477): B is permuted by the
// pivot sequence, then L y = P b (unit lower) and U x = y are solved.
//
// Each triangular solve is one launch.  Workgroup i owns row block i (64
// rows).  It subtracts the contributions of the blocks already solved, as
// each one becomes ready (its values arrive as tagged packets, written by the
// block's workgroup after its solve), and then solves its diagonal block
// inside one wave with lane broadcasts.  A workgroup only ever waits on
// blocks owned by workgroups with a smaller id in the same grid, which are
// dispatched before it, so the chain always makes progress; every wait is
// bounded.  Arithmetic is FP64; complex64 factors are widened on load.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <vector>

#include "wm_cplx.h"
#include "wm_internal.h"

namespace wm {

namespace {

constexpr int TB = 64;  // rows per block

template <typename T>
__device__ inline double2 ld2(const typename CT<T>::V *p)
{
    const typename CT<T>::V v = *p;
    return make_double2((double)v.x, (double)v.y);
}

// acc -= a*b with every rounding written out (fused multiply-adds), so a
// value does not depend on how the compiler contracts the instantiation it
// is inlined into (the right-hand-side count of a triangular solve).
__device__ inline void cfms(double2 &acc, double2 a, double2 b)
{
    acc.x = fma(-a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(-a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}

// acc += a*b, roundings written out as in cfms.
__device__ inline void cfma(double2 &acc, double2 a, double2 b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}

// perm[t] = source row of position t after applying ipiv in order.
__global__ void k_perm(const int *ipiv, int64_t strideP, int n, int *perm, int64_t strideQ, const int *bidx)
{
    extern __shared__ int s_p[];
    const int b = bidx ? bidx[blockIdx.x] : (int)blockIdx.x;
    const int *ip = ipiv + (size_t)b * strideP;
    int *pq = perm + (size_t)b * strideQ;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; ++i) {
            const int p = ip[i] - 1;
            if (p != i) { const int t = s_p[i]; s_p[i] = s_p[p]; s_p[p] = t; }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) pq[i] = s_p[i];
}

struct TrsvArgs {
    const void *A;
    int lda;
    int64_t strideA;
    int n, nrhs;
    const int *perm;
    int64_t strideQ;
    const double2 *Bin;   // lower solve: right-hand sides (gathered through perm), per launch member
    int ldb;
    int64_t strideB;
    double2 *Y;           // output (lower: y, upper: x in place of y)
    uint64_t *pkt;        // solved values as tagged packets: per member [block][rhs][row][4 words]
    int epoch;            // packet tag of this launch
    int *err;
    const int *bidx;      // launch member -> matrix index of B and Y (null: identity)
    const int *aidx;      // launch member -> matrix index of the factors and permutation (null: as bidx)
    const double2 *dinv;  // inverted diagonal blocks (k_diag_inv), used instead of substitution, or null
    const int *dslot;     // launch member -> slot of its factors' blocks in dinv
};

// Packed entries of one 64 x 64 triangle with its diagonal.
constexpr int kPack = TB * (TB + 1) / 2;

__device__ inline double bcast(double v, int lane)  // value of `lane` to the whole wave
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)(u & 0xffffffffu), lane);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The double whose low and high words lanes `l` and `l + 1` hold, to the whole wave.
__device__ inline double word_pair(uint32_t w, int l)
{
    const unsigned lo = __builtin_amdgcn_readlane(w, l), hi = __builtin_amdgcn_readlane(w, l + 1);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Packed index of element (i, j) of the 64 x 64 diagonal block, column by
// column (so the lanes of a wave, one row each, read consecutive entries of
// a column: no LDS bank conflicts): strictly lower part for L (i > j), upper
// part with the diagonal for U (i <= j).
__device__ inline int dl_idx(int i, int j) { return j * (2 * TB - 1 - j) / 2 + (i - j - 1); }
__device__ inline int du_idx(int i, int j) { return j * (j + 1) / 2 + i; }

// Triangular solve, one launch.  Workgroup (member, g) owns row block bi of
// that launch member (lower: g, upper: nblk-1-g); lower solves L y = P b,
// upper solves U x = y in place, and also publishes each solved value as four
// tagged packets for the later blocks.  The grid is (members, blocks), so
// workgroup ids run over the members first: the resident workgroups hold the
// leading blocks of every member and all members' chains advance together.
// Waits only on blocks of workgroups with smaller ids (dispatched earlier).
//
// TRANS solves with the conjugate transposes of the factors (A^H = U^H L^H P):
// the lower pass takes U^H (lower, non-unit diagonal conj(u_ii)) on the
// right-hand sides as staged (no gather), the upper pass L^H (upper, unit
// diagonal) in place; element (i, j) of the triangle is conj(A(j, i)).  The
// caller scatters the result through the permutation (x = P^T w).
template <typename T, int NR, bool UPPER, bool TRANS, bool INV = false>
__global__ __launch_bounds__(256) void k_trsv(TrsvArgs a, int r_off)
{
    // no contraction: a right-hand side's arithmetic is the same in every
    // instantiation (NR), so grouping columns never changes their values
#pragma clang fp contract(off)
    using V = typename CT<T>::V;
    constexpr bool UNIT = TRANS ? UPPER : !UPPER;  // unit diagonal (L or L^H)
    // element (i, j) of the triangle being solved
    auto elem = [&](const V *A, int i, int j) -> double2 {
        if (!TRANS) return ld2<T>(A + i + (size_t)j * a.lda);
        const double2 v = ld2<T>(A + j + (size_t)i * a.lda);
        return make_double2(v.x, -v.y);
    };
    const int bl = blockIdx.x, tid = threadIdx.x, row = tid & 63;
    const int b = a.bidx ? a.bidx[bl] : bl;
    const int fb = a.aidx ? a.aidx[bl] : b;  // factors (and their permutation) of this member
    const int part = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nblk = a.n / TB;
    const int bi = UPPER ? nblk - 1 - (int)blockIdx.y : (int)blockIdx.y, r0 = bi * TB;
    const V *A = (const V *)a.A + (size_t)fb * a.strideA;
    double2 *Y = a.Y + (size_t)b * a.strideB + (size_t)r_off * a.ldb;
    uint64_t *pk = a.pkt + (size_t)bl * nblk * NR * 256;
    const uint32_t tag = (uint32_t)a.epoch;
    const int nr = a.nrhs - r_off;

    // diagonal block to LDS (independent of other blocks, so before any wait)
    __shared__ double2 s_d[kPack];
    __shared__ double2 s_rd[TB];
    if (INV) {
        const double2 *D = a.dinv + (((size_t)a.dslot[bl] * 2 + (UPPER ? 1 : 0)) * nblk + bi) * kPack;
        for (int e = tid; e < kPack; e += 256) s_d[e] = D[e];
    }
    for (int e = tid; !INV && e < TB * TB; e += 256) {
        // consecutive threads walk down a column of the stored factor
        const int i = TRANS ? e / TB : e % TB, j = TRANS ? e % TB : e / TB;
        if (!UPPER && i > j) s_d[dl_idx(i, j)] = elem(A, r0 + i, r0 + j);
        if (UPPER && i < j) s_d[du_idx(i, j)] = elem(A, r0 + i, r0 + j);
        if (!UNIT && i == j) {
            const double2 u = elem(A, r0 + i, r0 + j);
            if (UPPER) s_d[du_idx(i, j)] = u;
            s_rd[i] = crecip(u.x, u.y);
        }
    }

    double2 acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = make_double2(0.0, 0.0);
    if (part == 0) {
        const int src = (UPPER || TRANS) ? r0 + row : a.perm[(size_t)fb * a.strideQ + r0 + row];
        const double2 *in = UPPER ? Y : a.Bin + (size_t)bl * a.strideB + (size_t)r_off * a.ldb;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (r < nr) acc[r] = in[src + (size_t)r * a.ldb];
    }

    // off-diagonal blocks in dependency order, next block prefetched
    const int nb_off = UPPER ? nblk - 1 - bi : bi;
    auto blk = [&](int q) { return UPPER ? nblk - 1 - q : q; };
    double2 lcur[16], lnext[16];
    if (nb_off > 0) {
        const int c0 = blk(0) * TB + part * 16;
#pragma unroll
        for (int c = 0; c < 16; ++c) lcur[c] = elem(A, r0 + row, c0 + c);
    }
    for (int q = 0; q < nb_off; ++q) {
        // The part's 16 solved values of block q: 64 packets per right-hand
        // side, one per lane, polled until the block's workgroup has tagged
        // them (a value and its readiness in one memory round trip), and all
        // loaded before the next block's L prefetch.
        const uint64_t *pq = pk + (size_t)blk(q) * NR * 256 + part * 64 + row;
        uint32_t yw[NR];
        long it = 0;
        for (;;) {
            bool got = true;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                yw[r] = 0u;
                if (r < nr) {
                    const uint64_t v = ld_pkt(pq + r * 256);
                    yw[r] = (uint32_t)v;
                    got = got && (uint32_t)(v >> 32) == tag;
                }
            }
            if (__all(got)) break;
            if (spin_give_up(it, a.err)) break;
        }
        __builtin_amdgcn_sched_barrier(0);
        if (q + 1 < nb_off) {
            const int c1 = blk(q + 1) * TB + part * 16;
#pragma unroll
            for (int c = 0; c < 16; ++c) lnext[c] = elem(A, r0 + row, c1 + c);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < 16; ++c) {
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (r < nr) cfms(acc[r], lcur[c], make_double2(word_pair(yw[r], 4 * c), word_pair(yw[r], 4 * c + 2)));
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) lcur[c] = lnext[c];
    }

    // The four parts' sums meet in LDS; wave p then solves the right-hand
    // sides p, p + 4, ... (each sum in the order part 0 + 1 + 2 + 3, so a
    // value does not depend on NR or on which wave solves it).
    __shared__ double2 s_acc[4][TB][NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) s_acc[part][row][r] = acc[r];
    __syncthreads();
    constexpr int RW = (NR + 3) / 4;  // right-hand sides per wave
    if (part < NR) {
        double2 x[RW];
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int r = part + 4 * k;
            x[k] = make_double2(0.0, 0.0);
            if (r < NR) {
                x[k] = s_acc[0][row][r];
#pragma unroll
                for (int p = 1; p < 4; ++p) { x[k].x += s_acc[p][row][r].x; x[k].y += s_acc[p][row][r].y; }
            }
        }
        if (INV) {
            // x = D^-1 a with the inverted diagonal block: a product instead of
            // a 64-step substitution (the wave's sums go through its own LDS
            // entries; the block of L^-1 has a unit diagonal)
#pragma unroll
            for (int k = 0; k < RW; ++k)
                if (part + 4 * k < NR) s_acc[0][row][part + 4 * k] = x[k];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int r = part + 4 * k;
                if (r >= NR) break;
                double2 v = UPPER ? make_double2(0.0, 0.0) : x[k];
#pragma unroll
                for (int j = 0; j < TB; ++j) {
                    if (UPPER ? j >= row : j < row)
                        cfma(v, s_d[UPPER ? du_idx(row, j) : dl_idx(row, j)], s_acc[0][j][r]);
                }
                x[k] = v;
            }
        } else if (!UPPER) {
        // the substitution is the launch's critical path: unrolled, so each
        // step's LDS read of the block is issued ahead of the step
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                const double2 l = row > j ? s_d[dl_idx(row, j)] : make_double2(0.0, 0.0);
                const double2 ri = UNIT ? make_double2(1.0, 0.0) : s_rd[j];
#pragma unroll
                for (int k = 0; k < RW; ++k) {
                    if (part + 4 * k >= NR) break;
                    if (!UNIT && row == j) {
                        const double2 v = x[k];
                        x[k] = make_double2(v.x * ri.x - v.y * ri.y, v.x * ri.y + v.y * ri.x);
                    }
                    if (j < TB - 1) {
                        const double2 yj = make_double2(bcast(x[k].x, j), bcast(x[k].y, j));
                        if (row > j) cfms(x[k], l, yj);
                    }
                }
            }
        } else {
#pragma unroll
            for (int j = TB - 1; j >= 0; --j) {
                const double2 u = row < j ? s_d[du_idx(row, j)] : make_double2(0.0, 0.0);
                const double2 ri = UNIT ? make_double2(1.0, 0.0) : s_rd[j];
#pragma unroll
                for (int k = 0; k < RW; ++k) {
                    if (part + 4 * k >= NR) break;
                    if (!UNIT && row == j) {
                        const double2 v = x[k];
                        x[k] = make_double2(v.x * ri.x - v.y * ri.y, v.x * ri.y + v.y * ri.x);
                    }
                    const double2 xj = make_double2(bcast(x[k].x, j), bcast(x[k].y, j));
                    if (row < j) cfms(x[k], u, xj);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int r = part + 4 * k;
            if (r < nr) {
                Y[(r0 + row) + (size_t)r * a.ldb] = x[k];
                uint64_t *pw = pk + ((size_t)bi * NR + r) * 256 + row * 4;
                const uint64_t xb = (uint64_t)__double_as_longlong(x[k].x), yb = (uint64_t)__double_as_longlong(x[k].y);
                put_pkt(pw, tag, (uint32_t)xb);
                put_pkt(pw + 1, tag, (uint32_t)(xb >> 32));
                put_pkt(pw + 2, tag, (uint32_t)yb);
                put_pkt(pw + 3, tag, (uint32_t)(yb >> 32));
            }
        }
    }
}

// Partial products P[split][row][r] = sum_{j in split} Z[row][j] X[j][r].
// Lane = row (each column of Z is one coalesced 1 KiB read per wave); the
// wave's column range is wave-uniform, so X arrives through scalar loads.  Z
// streams through in batches of kZB columns (all loads of a batch issued
// before its products, non-temporal: Z is read once per product), and each
// column's term is added in column order with fused multiply-adds, so a
// row's sum depends on n only.
typedef double d2v __attribute__((ext_vector_type(2)));

// ABS: also the split's partial row sums of |Z| (the infinity norm of Z, taken
// in the same pass over Z), Pabs[(member * nsplit + split) * n + row].
template <int NR, bool ABS>
__global__ __launch_bounds__(256) void k_resid_part(const double2 *Z, int ldz, int64_t strideZ, int n, int nrhs,
                                                    const double2 *X, int ldb, int64_t strideB, int cols_per_split,
                                                    int nsplit, double2 *P, int r_off, const int *bidx, double *Pabs)
{
    constexpr int kZB = 8;
    const int bl = blockIdx.z, split = blockIdx.y, lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = bidx ? bidx[bl] : bl;
    const int row = blockIdx.x * 64 + lane;
    const double2 *Zb = Z + (size_t)b * strideZ;
    const double2 *Xb = X + (size_t)b * strideB + (size_t)r_off * ldb;
    const int cw = cols_per_split / 4;
    const int c0 = split * cols_per_split + wave * cw;
    // NR <= nrhs - r_off: the launcher runs right-hand sides in chunks of 8, 4, 2, 1
    double2 acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = make_double2(0.0, 0.0);
    double asum = 0.0;
    const double2 *zp = Zb + row + (size_t)c0 * ldz;
    for (int c = 0; c < cw; c += kZB) {  // cw is a multiple of kZB
        double2 z[kZB];
#pragma unroll
        for (int e = 0; e < kZB; ++e) {
            const d2v v = __builtin_nontemporal_load((const d2v *)(zp + (size_t)(c + e) * ldz));
            z[e] = make_double2(v.x, v.y);
        }
#pragma unroll
        for (int e = 0; e < kZB; ++e) {
#pragma unroll
            for (int r = 0; r < NR; ++r) cfma(acc[r], z[e], Xb[c0 + c + e + (size_t)r * ldb]);
            if (ABS) asum += hypot(z[e].x, z[e].y);
        }
    }
    __shared__ double2 s[3][64][NR];
    __shared__ double sa[3][64];
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) s[wave - 1][lane][r] = acc[r];
        if (ABS) sa[wave - 1][lane] = asum;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
#pragma unroll
            for (int p = 0; p < 3; ++p) { acc[r].x += s[p][lane][r].x; acc[r].y += s[p][lane][r].y; }
            P[(((size_t)bl * nsplit + split) * NR + r) * n + row] = acc[r];
        }
        if (ABS) {
#pragma unroll
            for (int p = 0; p < 3; ++p) asum += sa[p][lane];
            Pabs[((size_t)bl * nsplit + split) * n + row] = asum;
        }
    }
}

// R = B - sum_split P, in split order.  With Pabs: the row sums of |Z| (in
// split order) over the leading np rows, their maximum into znorm[b] as the
// ordered bit pattern of a non-negative double (non-finite as +inf).
template <int NR>
__global__ void k_resid_final(const double2 *P, int nsplit, int n, int nrhs, const double2 *B, int ldb,
                              int64_t strideB, double2 *R, int r_off, const int *bidx, const double *Pabs, int np,
                              unsigned long long *znorm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int bl = blockIdx.y;
    const int b = bidx ? bidx[bl] : bl;
    if (Pabs) {  // uniform per launch
        double a = 0.0;
        if (i < np) {
            for (int sp = 0; sp < nsplit; ++sp) a += Pabs[((size_t)bl * nsplit + sp) * n + i];
            if (!(a <= DBL_MAX)) a = INFINITY;
        }
        for (int off = 32; off >= 1; off >>= 1) a = fmax(a, __shfl_xor(a, off));
        if ((threadIdx.x & 63) == 0) atomicMax(&znorm[b], (unsigned long long)__double_as_longlong(a));
    }
    if (i >= n) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (r >= nrhs - r_off) break;
        double2 s = make_double2(0.0, 0.0);
        for (int sp = 0; sp < nsplit; ++sp) {
            const double2 p = P[(((size_t)bl * nsplit + sp) * NR + r) * n + i];
            s.x += p.x;
            s.y += p.y;
        }
        const size_t o = (size_t)b * strideB + (size_t)(r_off + r) * ldb + i;
        R[o] = B ? make_double2(B[o].x - s.x, B[o].y - s.y) : s;  // no B: the product Z X
    }
}

template <typename T, int NR, bool TRANS>
void launch_trsv(const TrsvArgs &a, int batch, int r_off, bool upper, hipStream_t s)
{
    const dim3 grid(batch, a.n / TB);
    if (a.dinv && !TRANS) {
        if (upper) hipLaunchKernelGGL((k_trsv<T, NR, true, false, true>), grid, dim3(256), 0, s, a, r_off);
        else hipLaunchKernelGGL((k_trsv<T, NR, false, false, true>), grid, dim3(256), 0, s, a, r_off);
        return;
    }
    if (upper) hipLaunchKernelGGL((k_trsv<T, NR, true, TRANS>), grid, dim3(256), 0, s, a, r_off);
    else hipLaunchKernelGGL((k_trsv<T, NR, false, TRANS>), grid, dim3(256), 0, s, a, r_off);
}

// Inverses of the 64 x 64 diagonal blocks of LU factors, FP64 arithmetic:
// blockIdx = (block, triangle 0 = L (unit lower) / 1 = U, slot); slot z
// takes the factors of index fidx[z].  Lane j solves for column j of the
// inverse by substitution against the block staged in LDS; the inverse is
// stored packed as the triangular solve stages it (strictly lower part of
// L^-1, upper part of U^-1 with its diagonal), kPack entries per block.
template <typename T>
__global__ __launch_bounds__(64) void k_diag_inv(const void *F, int lda, int64_t strideF, const int *fidx, int nblk,
                                                 double2 *Dinv)
{
#pragma clang fp contract(off)
    using V = typename CT<T>::V;
    const int bi = blockIdx.x, upper = blockIdx.y, j = threadIdx.x, r0 = bi * TB;
    const V *A = (const V *)F + (size_t)fidx[blockIdx.z] * strideF;
    __shared__ double2 s_t[TB][TB + 1];  // the block, s_t[i][k] = A(r0 + i, r0 + k)
    __shared__ double2 s_x[TB][TB + 1];  // the inverse, column j by lane j
    for (int e = j; e < TB * TB; e += TB) {
        const int i = e % TB, k = e / TB;
        s_t[i][k] = ld2<T>(A + (r0 + i) + (size_t)(r0 + k) * lda);
    }
    __syncthreads();
    if (!upper) {
        for (int i = 0; i < TB; ++i) {
            double2 v = make_double2(i == j ? 1.0 : 0.0, 0.0);
            if (i > j)
                for (int k = j; k < i; ++k) cfms(v, s_t[i][k], s_x[k][j]);
            s_x[i][j] = v;
        }
    } else {
        for (int i = TB - 1; i >= 0; --i) {
            double2 v = make_double2(i == j ? 1.0 : 0.0, 0.0);
            if (i <= j) {
                for (int k = i + 1; k <= j; ++k) cfms(v, s_t[i][k], s_x[k][j]);
                const double2 d = s_t[i][i], rd = crecip(d.x, d.y);
                v = make_double2(v.x * rd.x - v.y * rd.y, v.x * rd.y + v.y * rd.x);
            }
            s_x[i][j] = v;
        }
    }
    __syncthreads();
    double2 *D = Dinv + (((size_t)blockIdx.z * 2 + upper) * nblk + bi) * kPack;
    for (int e = j; e < TB * TB; e += TB) {
        const int i = e % TB, k = e / TB;
        if (!upper && i > k) D[dl_idx(i, k)] = s_x[i][k];
        if (upper && i <= k) D[du_idx(i, k)] = s_x[i][k];
    }
}

// Right-hand sides handled by one pass starting at r_off: 8, 4, 2 or 1.
inline int rhs_chunk(int left) { return left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1; }

template <typename T, bool TRANS = false>
void trsv_dispatch(const TrsvArgs &a, int batch, int r_off, bool upper, hipStream_t s)
{
    switch (rhs_chunk(a.nrhs - r_off)) {
        case 8: launch_trsv<T, 8, TRANS>(a, batch, r_off, upper, s); break;
        case 4: launch_trsv<T, 4, TRANS>(a, batch, r_off, upper, s); break;
        case 2: launch_trsv<T, 2, TRANS>(a, batch, r_off, upper, s); break;
        default: launch_trsv<T, 1, TRANS>(a, batch, r_off, upper, s); break;
    }
}

// Bstage[member] = B[matrix] (n x nrhs columns with leading dimension ldb)
__global__ void k_stage(const double2 *B, int ldb, int64_t strideB, int n, int nrhs, double2 *S, const int *bidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, bl = blockIdx.y;
    const int b = bidx ? bidx[bl] : bl;
    if (i >= n) return;
    for (int r = 0; r < nrhs; ++r) S[(size_t)bl * strideB + (size_t)r * ldb + i] = B[(size_t)b * strideB + (size_t)r * ldb + i];
}

// S[member][perm[t]] = B[matrix][t] (the scatter x = P^T w of a transposed solve)
__global__ void k_unperm(const double2 *B, int ldb, int64_t strideB, int n, int nrhs, const int *perm,
                         int64_t strideQ, double2 *S, const int *bidx, const int *aidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, bl = blockIdx.y;
    const int b = bidx ? bidx[bl] : bl;
    if (i >= n) return;
    const int t = perm[(size_t)(aidx ? aidx[bl] : b) * strideQ + i];
    for (int r = 0; r < nrhs; ++r) S[(size_t)bl * strideB + (size_t)r * ldb + t] = B[(size_t)b * strideB + (size_t)r * ldb + i];
}

// B[matrix] = S[member]
__global__ void k_unstage(const double2 *S, int ldb, int64_t strideB, int n, int nrhs, double2 *B, const int *bidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, bl = blockIdx.y;
    const int b = bidx ? bidx[bl] : bl;
    if (i >= n) return;
    for (int r = 0; r < nrhs; ++r) B[(size_t)b * strideB + (size_t)r * ldb + i] = S[(size_t)bl * strideB + (size_t)r * ldb + i];
}

__global__ void k_zero_words(int *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

int perm_batched(wm_ctx *ctx, int n, const int *ipiv, int64_t strideP, int *perm, int64_t strideQ, int batch,
                 hipStream_t s, const int *d_bidx)
{
    if ((size_t)n * sizeof(int) > ctx->max_lds) {
        set_error("permutation: n=%d exceeds the LDS capacity", n);
        return WM_ERR_ARG;
    }
    hipLaunchKernelGGL(k_perm, dim3(batch), dim3(256), n * sizeof(int), s, ipiv, strideP, n, perm, strideQ, d_bidx);
    WM_HIP_TRY(hipGetLastError());
    return WM_OK;
}

namespace {

// Buffers of a triangular solve supplied by the caller instead of the
// context's (the allocation-free entry point): staging of B (batch * strideB
// complex128) and the packets of both triangles, zeroed by the caller, whose
// epochs then start at 1.
struct SolveBufs {
    double2 *stage = nullptr;
    uint64_t *pkt = nullptr;
    int epoch = 0;
};

// Packets of the two triangles of a batched solve.
size_t solve_pkts(int n, int nrhs, int batch) { return 2 * (size_t)batch * (n / TB) * rhs_chunk(nrhs) * 256; }

// Staging of B per launch member in ctx->ws_solve; the solved values'
// packets in ctx->ws_pkt, per triangle and member nblk x NR x 256 packets.
// trans: solve A^H X = B instead of A X = B.  ext: caller-supplied buffers
// (nothing allocated, no profiler events).
int getrs_any(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA, const int *perm,
              int64_t strideQ, void *B, int ldb, int64_t strideB, int batch, hipStream_t s, const int *d_bidx,
              bool trans, const int *d_aidx, SolveBufs *ext = nullptr, const double2 *dinv = nullptr,
              const int *d_dslot = nullptr)
{
    if (n <= 0 || n % TB != 0 || nrhs <= 0 || batch <= 0 || ldb < n) {
        set_error("getrs: bad arguments (n=%d multiple of %d, nrhs=%d, batch=%d)", n, TB, nrhs, batch);
        return WM_ERR_ARG;
    }
    const int nblk = n / TB;
    const int nr_max = rhs_chunk(nrhs);
    const size_t tri_pkts = (size_t)batch * nblk * nr_max * 256;  // packets per triangle
    double2 *Ystage = nullptr;
    uint64_t *pkt = nullptr;
    int *epoch = nullptr;
    if (ext) {
        Ystage = ext->stage;
        pkt = ext->pkt;
        epoch = &ext->epoch;
    } else {
        WM_TRY(ctx->ws_solve.ensure((size_t)batch * strideB * sizeof(double2)));
        Ystage = (double2 *)ctx->ws_solve.ptr;
        // The packets live in their own buffer and only ever carry epoch tags of
        // this context, so a wait can never be satisfied by stale data.  The
        // buffer is zeroed whenever it is (re)allocated.
        const size_t pkt_bytes = 2 * tri_pkts * sizeof(uint64_t);
        if (ctx->ws_pkt.bytes < pkt_bytes || !ctx->ws_pkt.ptr) {
            WM_TRY(ctx->ws_pkt.ensure(pkt_bytes));
            const int64_t nf = (int64_t)(ctx->ws_pkt.bytes / sizeof(int));
            hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, (int *)ctx->ws_pkt.ptr, nf);
            ctx->solve_epoch = 0;
        }
        pkt = (uint64_t *)ctx->ws_pkt.ptr;
        if (ctx->solve_epoch > (1 << 30)) {  // keep the tags unique: restart from a zeroed buffer
            const int64_t nf = (int64_t)(ctx->ws_pkt.bytes / sizeof(int));
            hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, (int *)pkt, nf);
            ctx->solve_epoch = 0;
        }
        epoch = &ctx->solve_epoch;
    }
    // the lower solve reads B (staged per member) and writes y into B; the upper solves in place
    hipLaunchKernelGGL(k_stage, dim3((n + 255) / 256, batch), dim3(256), 0, s, (const double2 *)B, ldb, strideB, n, nrhs,
                       Ystage, d_bidx);
    for (int r0 = 0; r0 < nrhs; r0 += rhs_chunk(nrhs - r0)) {
        TrsvArgs a;
        a.A = A; a.lda = lda; a.strideA = strideA; a.n = n; a.nrhs = nrhs;
        a.perm = perm; a.strideQ = strideQ; a.ldb = ldb; a.strideB = strideB; a.err = ctx->d_err;
        a.bidx = d_bidx;
        a.aidx = d_aidx;
        a.dinv = trans ? nullptr : dinv;
        a.dslot = d_dslot;
        a.Bin = Ystage; a.Y = (double2 *)B; a.pkt = pkt;
        a.epoch = ++*epoch;
        const int nr = rhs_chunk(nrhs - r0);
        const double esz = prec == WM_PREC_F64 ? 16.0 : 8.0;
        const double fl = batch * 4.0 * (double)n * n * nr, by = batch * 0.5 * (double)n * n * esz;
        for (int up = 0; up < 2; ++up) {
            if (up) { a.Bin = nullptr; a.pkt = pkt + tri_pkts; }
            if (!ext) prof_begin(ctx, K_TRSV, s);
            if (prec == WM_PREC_F64) {
                if (trans) trsv_dispatch<double, true>(a, batch, r0, up, s);
                else trsv_dispatch<double>(a, batch, r0, up, s);
            } else {
                if (trans) trsv_dispatch<float, true>(a, batch, r0, up, s);
                else trsv_dispatch<float>(a, batch, r0, up, s);
            }
            if (!ext) prof_end(ctx, K_TRSV, s, fl, by);
        }
    }
    if (trans) {  // x = P^T w, through the staging buffer
        hipLaunchKernelGGL(k_unperm, dim3((n + 255) / 256, batch), dim3(256), 0, s, (const double2 *)B, ldb, strideB, n,
                           nrhs, perm, strideQ, Ystage, d_bidx, d_aidx);
        hipLaunchKernelGGL(k_unstage, dim3((n + 255) / 256, batch), dim3(256), 0, s, (const double2 *)Ystage, ldb,
                           strideB, n, nrhs, (double2 *)B, d_bidx);
    }
    WM_HIP_TRY(hipGetLastError());
    return WM_OK;
}

}  // namespace

int getrs_perm_batched(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                       const int *perm, int64_t strideQ, void *B, int ldb, int64_t strideB, int batch, hipStream_t s,
                       const int *d_bidx, const int *d_aidx)
{
    return getrs_any(ctx, prec, n, nrhs, A, lda, strideA, perm, strideQ, B, ldb, strideB, batch, s, d_bidx, false,
                     d_aidx);
}

int getrsh_perm_batched(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                        const int *perm, int64_t strideQ, void *B, int ldb, int64_t strideB, int batch, hipStream_t s,
                        const int *d_bidx)
{
    return getrs_any(ctx, prec, n, nrhs, A, lda, strideA, perm, strideQ, B, ldb, strideB, batch, s, d_bidx, true,
                     nullptr);
}

namespace {

// S[member * W + j] = column cols[member * W + j] of V (element offset; -1: zeros).
__global__ void k_gather_cols(const double2 *V, int n, const int64_t *cols, double2 *S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, mj = blockIdx.y;
    if (i >= n) return;
    const int64_t c = cols[mj];
    S[(size_t)mj * n + i] = c >= 0 ? V[c + i] : make_double2(0.0, 0.0);
}

// The columns back: V[cols[member * W + j]] = S[member * W + j].
__global__ void k_scatter_cols(const double2 *S, int n, const int64_t *cols, double2 *V)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, mj = blockIdx.y;
    if (i >= n) return;
    const int64_t c = cols[mj];
    if (c >= 0) V[c + i] = S[(size_t)mj * n + i];
}

}  // namespace

size_t diag_inv_bytes(int n, int nfac) { return (size_t)nfac * 2 * (n / TB) * kPack * sizeof(double2); }

int diag_inv_batched(wm_ctx *ctx, int prec, int n, const void *F, int lda, int64_t strideF,
                     const std::vector<int> &factors, double2 *Dinv, hipStream_t s)
{
    if (factors.empty()) return WM_OK;
    if (n % TB != 0) {
        set_error("diag_inv: n=%d not a multiple of %d", n, TB);
        return WM_ERR_ARG;
    }
    WM_TRY(ctx->ws_dinv_idx.ensure(factors.size() * sizeof(int)));
    WM_TRY(copy_h2d(ctx->ws_dinv_idx.ptr, factors.data(), factors.size() * sizeof(int), s));
    const dim3 grid((unsigned)(n / TB), 2, (unsigned)factors.size());
    const int *fidx = (const int *)ctx->ws_dinv_idx.ptr;
    if (prec == WM_PREC_F64)
        hipLaunchKernelGGL(k_diag_inv<double>, grid, dim3(TB), 0, s, F, lda, strideF, fidx, n / TB, Dinv);
    else
        hipLaunchKernelGGL(k_diag_inv<float>, grid, dim3(TB), 0, s, F, lda, strideF, fidx, n / TB, Dinv);
    WM_HIP_TRY(hipGetLastError());
    return WM_OK;
}

int getrs_grouped(wm_ctx *ctx, int prec, int n, int nr, const void *F, int lda, int64_t strideF, const int *perm,
                  int64_t strideQ, double2 *V, int64_t strideV, const std::vector<int> &items,
                  const std::vector<int> &fac, hipStream_t s, const double2 *dinv, const std::vector<int> *slot_of,
                  int max_w)
{
    if (items.empty()) return WM_OK;
    if (items.size() != fac.size() || nr <= 0 || strideV < (int64_t)n * nr) {
        set_error("getrs_grouped: bad arguments");
        return WM_ERR_ARG;
    }
    // the columns of each factor, factors in the order of their first item
    std::vector<int> order;
    std::vector<std::vector<int64_t>> cols_of;
    std::vector<int> pos_of;  // factor index -> position in order (grown on demand)
    for (size_t q = 0; q < items.size(); ++q) {
        const int f = fac[q];
        if (f < 0) {
            set_error("getrs_grouped: item %d has no factors", items[q]);
            return WM_ERR_ARG;
        }
        if ((size_t)f >= pos_of.size()) pos_of.resize((size_t)f + 1, -1);
        if (pos_of[f] < 0) {
            pos_of[f] = (int)order.size();
            order.push_back(f);
            cols_of.emplace_back();
        }
        for (int r = 0; r < nr; ++r) cols_of[pos_of[f]].push_back((int64_t)items[q] * strideV + (int64_t)r * n);
    }
    size_t widest = 0;
    for (const auto &c : cols_of) widest = std::max(widest, c.size());
    int W = 1;  // right-hand sides per launch member: a power of two, at most max_w (<= 8)
    while (W < std::min(max_w, 8) && (size_t)W < widest) W *= 2;
    std::vector<int> aidx;
    std::vector<int64_t> cols;
    for (size_t g = 0; g < order.size(); ++g)
        for (size_t j0 = 0; j0 < cols_of[g].size(); j0 += W) {
            aidx.push_back(order[g]);
            for (int j = 0; j < W; ++j)
                cols.push_back(j0 + j < cols_of[g].size() ? cols_of[g][j0 + j] : (int64_t)-1);
        }
    const int members = (int)aidx.size();
    std::vector<int> dslot(members, 0);
    if (dinv) {
        for (int m = 0; m < members; ++m) {
            const int f = aidx[m];
            if (!slot_of || f >= (int)slot_of->size() || (*slot_of)[f] < 0) {
                set_error("getrs_grouped: factor %d has no inverted diagonal blocks", f);
                return WM_ERR_ARG;
            }
            dslot[m] = (*slot_of)[f];
        }
    }
    const size_t aidx_b = ((size_t)members * sizeof(int) + 255) & ~(size_t)255;
    const size_t cols_b = ((size_t)members * W * sizeof(int64_t) + 255) & ~(size_t)255;
    const size_t stage_b = (size_t)members * W * n * sizeof(double2);
    WM_TRY(ctx->ws_group.ensure(2 * aidx_b + cols_b + stage_b));
    char *base = (char *)ctx->ws_group.ptr;
    std::vector<char> host(2 * aidx_b + cols_b, 0);
    std::memcpy(host.data(), aidx.data(), aidx.size() * sizeof(int));
    std::memcpy(host.data() + aidx_b, dslot.data(), dslot.size() * sizeof(int));
    std::memcpy(host.data() + 2 * aidx_b, cols.data(), cols.size() * sizeof(int64_t));
    WM_TRY(copy_h2d(base, host.data(), host.size(), s));
    const int *d_aidx = (const int *)base;
    const int *d_dslot = (const int *)(base + aidx_b);
    const int64_t *d_cols = (const int64_t *)(base + 2 * aidx_b);
    double2 *S = (double2 *)(base + 2 * aidx_b + cols_b);
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)(members * W));
    hipLaunchKernelGGL(k_gather_cols, grid, dim3(256), 0, s, (const double2 *)V, n, d_cols, S);
    WM_TRY(getrs_any(ctx, prec, n, W, F, lda, strideF, perm, strideQ, S, n, (int64_t)W * n, members, s, nullptr,
                     false, d_aidx, nullptr, dinv, dinv ? d_dslot : nullptr));
    hipLaunchKernelGGL(k_scatter_cols, grid, dim3(256), 0, s, (const double2 *)S, n, d_cols, V);
    WM_HIP_TRY(hipGetLastError());
    return WM_OK;
}

int getrs_batched(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA, const int *ipiv,
                  int64_t strideP, void *B, int ldb, int64_t strideB, int batch, hipStream_t s)
{
    WM_TRY(ctx->ws_perm.ensure((size_t)batch * n * sizeof(int)));
    int *perm = (int *)ctx->ws_perm.ptr;
    WM_TRY(perm_batched(ctx, n, ipiv, strideP, perm, n, batch, s));
    return getrs_perm_batched(ctx, prec, n, nrhs, A, lda, strideA, perm, n, B, ldb, strideB, batch, s);
}

int residual_batched(wm_ctx *ctx, int n, int nrhs, const void *Z, int ldz, int64_t strideZ, const void *X,
                     const void *B, int ldb, int64_t strideB, void *R, int batch, hipStream_t s, const int *d_bidx,
                     unsigned long long *znorm, int np)
{
    if (n % 64 != 0) {
        set_error("residual: n=%d not a multiple of 64", n);
        return WM_ERR_ARG;
    }
    // column splits depend on n only, so an item's rounding never depends on the batch
    int nsplit = 1;
    while (nsplit < 16 && n % (nsplit * 2 * 64) == 0 && n / (nsplit * 2) >= 256) nsplit *= 2;
    const int cps = n / nsplit;
    const int nrp = nrhs >= 8 ? 8 : nrhs >= 4 ? 4 : nrhs >= 2 ? 2 : 1;
    DevBuf &partial = ctx->ws_resid;
    const size_t pbytes = (size_t)batch * nsplit * nrp * n * sizeof(double2);
    const size_t need = pbytes + (znorm ? (size_t)batch * nsplit * n * sizeof(double) : 0);
    WM_TRY(partial.ensure(need));
    double *pabs = znorm ? (double *)((char *)partial.ptr + pbytes) : nullptr;
    for (int r0 = 0; r0 < nrhs; r0 += rhs_chunk(nrhs - r0)) {
        const int left = rhs_chunk(nrhs - r0);
        const dim3 g1(n / 64, nsplit, batch), g2((n + 255) / 256, batch);
        // the norm of Z rides along the first right-hand-side chunk only
        double *pa = r0 == 0 ? pabs : nullptr;
#define WM_RESID(NR)                                                                                          \
    if (pa)                                                                                                   \
        hipLaunchKernelGGL((k_resid_part<NR, true>), g1, dim3(256), 0, s, (const double2 *)Z, ldz, strideZ, n, \
                           nrhs, (const double2 *)X, ldb, strideB, cps, nsplit, (double2 *)partial.ptr, r0,     \
                           d_bidx, pa);                                                                       \
    else                                                                                                      \
        hipLaunchKernelGGL((k_resid_part<NR, false>), g1, dim3(256), 0, s, (const double2 *)Z, ldz, strideZ, n, \
                           nrhs, (const double2 *)X, ldb, strideB, cps, nsplit, (double2 *)partial.ptr, r0,     \
                           d_bidx, (double *)nullptr);                                                         \
    hipLaunchKernelGGL(k_resid_final<NR>, g2, dim3(256), 0, s, (const double2 *)partial.ptr, nsplit, n, nrhs,   \
                       (const double2 *)B, ldb, strideB, (double2 *)R, r0, d_bidx, (const double *)pa, np, znorm)
        prof_begin(ctx, K_RESID, s);
        if (left == 8) { WM_RESID(8); }
        else if (left == 4) { WM_RESID(4); }
        else if (left == 2) { WM_RESID(2); }
        else { WM_RESID(1); }
        prof_end(ctx, K_RESID, s, batch * 8.0 * (double)n * n * left, batch * 16.0 * (double)n * n);
#undef WM_RESID
    }
    WM_HIP_TRY(hipGetLastError());
    return WM_OK;
}

}  // namespace wm

// ---------------------------------------------------------------------------
// Allocation-free triangular solves: the staging, permutation and packet
// buffers of wm_getrs_batched live in `scratch`, a file-scope device array of
// kNoAllocBytes that every GPU the library runs on holds until the process
// ends.  Its device address is taken once per device (wm::noalloc_init, from
// wm_init with that device current); host code uses that address, never the
// array's name.  Calls on one GPU share the array: overlapping calls race.
namespace {
constexpr size_t kNoAllocBytes = (size_t)32 << 20;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
void *g_scratch_dev[kNoAllocDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

namespace wm {

int noalloc_init(int device)
{
    if (device < 0 || device >= kNoAllocDevices) {
        set_error("scratch: device %d beyond %d", device, kNoAllocDevices);
        return WM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return WM_OK;
    void *p = nullptr;
    WM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = p;
    return WM_OK;
}

}  // namespace wm

extern "C" int wm_getrs_batched_noalloc(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda,
                                        int64_t strideA, const int *ipiv, int64_t strideP, wm_c128 *B, int ldb,
                                        int64_t strideB, int batch)
{
    using namespace wm;
    if (!ctx || !A || !ipiv || !B || n <= 0 || n % TB != 0 || nrhs <= 0 || batch <= 0 || ldb < n || lda < n ||
        strideB < (int64_t)ldb * nrhs || (prec != WM_PREC_F64 && prec != WM_PREC_F32 && prec != WM_PREC_BF16X3)) {
        set_error("wm_getrs_batched_noalloc: bad arguments (n=%d a multiple of %d, nrhs=%d, batch=%d, ldb=%d, "
                  "strideB=%lld)", n, TB, nrhs, batch, ldb, (long long)strideB);
        return WM_ERR_ARG;
    }
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t perm_b = up((size_t)batch * n * sizeof(int));
    const size_t stage_b = up((size_t)batch * strideB * sizeof(double2));
    const size_t pkt_b = up(solve_pkts(n, nrhs, batch) * sizeof(uint64_t));
    if (perm_b + stage_b + pkt_b > kNoAllocBytes) {
        set_error("wm_getrs_batched_noalloc: n=%d, nrhs=%d, batch=%d need %zu bytes of scratch, the maximum is %zu "
                  "(use wm_getrs_batched)", n, nrhs, batch, perm_b + stage_b + pkt_b, kNoAllocBytes);
        return WM_ERR_NOMEM;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    char *base = (char *)g_scratch_dev[ctx->device];
    if (!base) {
        set_error("wm_getrs_batched_noalloc: no scratch address for device %d", ctx->device);
        return WM_ERR_STATE;
    }
    const hipStream_t s = ctx->stream;
    int *perm = (int *)base;
    SolveBufs sb;
    sb.stage = (double2 *)(base + perm_b);
    sb.pkt = (uint64_t *)(base + perm_b + stage_b);
    // the previous call's packets carry the same epochs: start from zero
    const int64_t nw = (int64_t)(pkt_b / sizeof(int));
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, (int *)sb.pkt, nw);
    WM_TRY(perm_batched(ctx, n, ipiv, strideP, perm, n, batch, s));
    WM_TRY(getrs_any(ctx, prec, n, nrhs, A, lda, strideA, perm, n, B, ldb, strideB, batch, s, nullptr, false, nullptr,
                     &sb));
    return check_device_error(ctx, s);
}

// Internal test hook (not part of include/wiremom.h): grouped solves.  Item q
// (vectors at V + items[q] * strideV, nr columns of n) is solved in place with
// the factors fac[q] of wm_getrf_batched (device buffers, stride n * n; pivots
// stride n).
extern "C" int wmx_getrs_grouped(wm_ctx *ctx, int prec, int n, int nr, const void *A, int nfac, const int *ipiv,
                                 wm_c128 *V, int64_t strideV, int nitems, const int *items, const int *fac, int inv,
                                 int max_w)
{
    using namespace wm;
    if (!ctx || !A || !ipiv || !V || nitems <= 0 || nfac <= 0) {
        set_error("wmx_getrs_grouped: bad arguments");
        return WM_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    WM_TRY(ctx->ws_perm.ensure((size_t)nfac * n * sizeof(int)));
    int *perm = (int *)ctx->ws_perm.ptr;
    WM_TRY(perm_batched(ctx, n, ipiv, n, perm, n, nfac, ctx->stream));
    const double2 *dinv = nullptr;
    std::vector<int> facs(nfac), slot(nfac);
    for (int f = 0; f < nfac; ++f) facs[f] = slot[f] = f;
    if (inv) {
        WM_TRY(ctx->ws_dinv.ensure(diag_inv_bytes(n, nfac)));
        dinv = (const double2 *)ctx->ws_dinv.ptr;
        WM_TRY(diag_inv_batched(ctx, prec, n, A, n, (int64_t)n * n, facs, (double2 *)dinv, ctx->stream));
    }
    WM_TRY(getrs_grouped(ctx, prec, n, nr, A, n, (int64_t)n * n, perm, n, (double2 *)V, strideV,
                         std::vector<int>(items, items + nitems), std::vector<int>(fac, fac + nitems), ctx->stream,
                         dinv, inv ? &slot : nullptr, max_w > 0 ? max_w : kGroupCols));
    return check_device_error(ctx, ctx->stream);
}

// Internal test hook (not part of include/wiremom.h): solves A^H X = B with the
// factors and pivots of wm_getrf_batched (device buffers, as wm_getrs_batched).
extern "C" int wmx_getrs_trans(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                               const int *ipiv, int64_t strideP, wm_c128 *B, int ldb, int64_t strideB, int batch)
{
    using namespace wm;
    if (!ctx || !A || !ipiv || !B || batch <= 0) return WM_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    WM_TRY(ctx->ws_perm.ensure((size_t)batch * n * sizeof(int)));
    int *perm = (int *)ctx->ws_perm.ptr;
    WM_TRY(perm_batched(ctx, n, ipiv, strideP, perm, n, batch, ctx->stream));
    WM_TRY(getrsh_perm_batched(ctx, prec, n, nrhs, A, lda, strideA, perm, n, B, ldb, strideB, batch, ctx->stream));
    WM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return check_device_error(ctx, ctx->stream);
}
