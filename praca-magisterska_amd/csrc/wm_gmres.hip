// FP64 GMRES iterative refinement with another matrix's LU factors as the
// preconditioner: the solve of a frequency sweep that factors only some
// frequencies ("anchors") and solves the others with their nearest anchor's
// factors.
//
// The reference factors every frequency from scratch (LAPACKE_zgetrf,
// "wersja referencyjna/Wire_MoM.cu":372, inside the loop over frequencies at
// 335-400).  Within a sweep the matrices of nearby frequencies are close: below
// the structure's resonances Z(f) = j w L + P / (j w) with L and P nearly
// independent of f, so with M = Z(f_a) the eigenvalues of M^-1 Z(f) lie between
// f_a / f and f / f_a, and GMRES on Z(f) preconditioned by the LU of M
// converges in a handful of iterations; near resonances a few more.
//
// Outer loop (iterative refinement, as for complex64 factors): R = B - Z X in
// FP64, D = GMRES(Z, R) preconditioned on the right by M^-1 = (L U)^-1 P of the
// anchor, X += D, until ||D|| / ||X|| <= tol (wm::refine_decide).  Inner loop:
// flexible GMRES (the preconditioned vectors z_k = M^-1 v_k are kept, so the
// correction is sum_k y_k z_k), Arnoldi with classical Gram-Schmidt applied
// twice, Givens rotations on the host; it stops when the estimated residual
// falls below kInnerTol0 ||R|| on the first outer iteration (X = 0: the solve
// itself) and kInnerTol ||R|| later (a correction to a few digits is enough to
// measure ||D||), or after m iterations.  All items of a batch iterate in
// lockstep, one launch per step, each with its own anchor's factors.  Every
// reduction has a fixed order, so an item's result does not depend on its
// batch neighbours.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "wm_gmres.h"

namespace wm {

namespace {

// inner tolerance with X = 0 (relative to ||B||): the acceptance test is the
// backward error (below), which ||Z|| ||X|| >> ||B|| makes reachable from a
// looser GMRES residual.  On the lightning sweep at N = 8192 (4 anchors):
// GMRES steps 592 / 572 / 554 / 570 / 592 at 3e-13 / 1e-12 / 3e-12 / 1e-11 /
// 3e-11, the same results; looser than 3e-12 more items need a second outer
// iteration
constexpr double kInnerTol0 = 3e-12;
constexpr double kInnerTol = 1e-3;    // inner tolerance of later corrections
// A result whose normwise backward error ||R|| / (||Z|| ||X|| + ||B||) (infinity
// norms, R the FP64 residual) is at most this is accepted at once: an LU
// solution's is a small multiple of u64 = 1.1e-16.
constexpr double kBerrTol = 1e-15;

// out[(member * nsets + j) * nr + r] = sum_i conj(V_j[b][r][i]) * W[b][r][i]
// for j < nsets, member -> item b = bidx[member]; V_j at Vb + j * setStride.
__global__ __launch_bounds__(256) void k_dots(const double2 *Vb, int64_t setStride, int nsets, const double2 *W,
                                              int64_t sV, int n, int nr, const int *bidx, double2 *out)
{
    const int mem = blockIdx.x, j = blockIdx.y, r = blockIdx.z, tid = threadIdx.x;
    const int b = bidx[mem];
    const double2 *v = Vb + (size_t)j * setStride + (size_t)b * sV + (size_t)r * n;
    const double2 *w = W + (size_t)b * sV + (size_t)r * n;
    double sx = 0.0, sy = 0.0;
    for (int i = tid; i < n; i += 256) {
        const double2 a = v[i], c = w[i];
        sx += a.x * c.x + a.y * c.y;
        sy += a.x * c.y - a.y * c.x;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
    }
    __shared__ double s[2][4];
    if ((tid & 63) == 0) {
        s[0][tid >> 6] = sx;
        s[1][tid >> 6] = sy;
    }
    __syncthreads();
    if (tid == 0)
        out[((size_t)mem * nsets + j) * nr + r] =
            make_double2((s[0][0] + s[0][1]) + (s[0][2] + s[0][3]), (s[1][0] + s[1][1]) + (s[1][2] + s[1][3]));
}

// W[b][r] -= sum_{j < nsets} h[(member * hstride + j) * nr + r] * V_j[b][r]
__global__ void k_axpy(double2 *W, const double2 *Vb, int64_t setStride, int nsets, const double2 *h, int hstride,
                       int64_t sV, int n, int nr, const int *bidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int mem = blockIdx.y / nr, r = blockIdx.y % nr;
    if (i >= n) return;
    const int b = bidx[mem];
    double2 w = W[(size_t)b * sV + (size_t)r * n + i];
    for (int j = 0; j < nsets; ++j) {
        const double2 c = h[((size_t)mem * hstride + j) * nr + r];
        const double2 v = Vb[(size_t)j * setStride + (size_t)b * sV + (size_t)r * n + i];
        w.x -= c.x * v.x - c.y * v.y;
        w.y -= c.x * v.y + c.y * v.x;
    }
    W[(size_t)b * sV + (size_t)r * n + i] = w;
}

// Dst[b][r] = Src[b][r] / sqrt(nrm2[member * nr + r].x) (columns of zero norm stay zero)
__global__ void k_normalize(double2 *Dst, const double2 *Src, const double2 *nrm2, int64_t sV, int n, int nr,
                            const int *bidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int mem = blockIdx.y / nr, r = blockIdx.y % nr;
    if (i >= n) return;
    const int b = bidx[mem];
    const double q = nrm2[(size_t)mem * nr + r].x;
    const double sc = q > 0.0 ? 1.0 / sqrt(q) : 0.0;
    const double2 v = Src[(size_t)b * sV + (size_t)r * n + i];
    Dst[(size_t)b * sV + (size_t)r * n + i] = make_double2(v.x * sc, v.y * sc);
}

// D[b][r] = sum_{j < k} y[(member * ystride + j) * nr + r] * Zk_j[b][r]
__global__ void k_combine(double2 *D, const double2 *Zb, int64_t setStride, int k, const double2 *y, int ystride,
                          int64_t sV, int n, int nr, const int *bidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int mem = blockIdx.y / nr, r = blockIdx.y % nr;
    if (i >= n) return;
    const int b = bidx[mem];
    double2 d = make_double2(0.0, 0.0);
    for (int j = 0; j < k; ++j) {
        const double2 c = y[((size_t)mem * ystride + j) * nr + r];
        // beyond this column's own Arnoldi steps the coefficients are zero and
        // the sets may hold another call's data: skipped, not multiplied by 0
        if (c.x == 0.0 && c.y == 0.0) continue;
        const double2 v = Zb[(size_t)j * setStride + (size_t)b * sV + (size_t)r * n + i];
        d.x += c.x * v.x - c.y * v.y;
        d.y += c.x * v.y + c.y * v.x;
    }
    D[(size_t)b * sV + (size_t)r * n + i] = d;
}

// Dst[b] = Src[b] for the listed items (sV entries each)
__global__ void k_copy_set(double2 *Dst, const double2 *Src, int64_t sV, const int *bidx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sV) return;
    const size_t o = (size_t)bidx[blockIdx.y] * sV + i;
    Dst[o] = Src[o];
}

__global__ void k_zero_set(double2 *Dst, int64_t sV, const int *bidx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < sV) Dst[(size_t)bidx[blockIdx.y] * sV + i] = make_double2(0.0, 0.0);
}

// X += D for the listed items; max |D|, |X| per (item, rhs) as ordered bit
// patterns (non-finite values as +inf).
__global__ void k_update_max(double2 *X, const double2 *D, int n, int64_t sV, int nr, const int *bidx,
                             unsigned long long *mx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int br = blockIdx.y, b = bidx[br / nr], r = br % nr;
    double dm = 0.0, xm = 0.0;
    if (i < n) {
        const size_t o = (size_t)b * sV + (size_t)r * n + i;
        const double2 d = D[o];
        double2 x = X[o];
        x.x += d.x;
        x.y += d.y;
        X[o] = x;
        dm = hypot(d.x, d.y);
        xm = hypot(x.x, x.y);
        if (!(dm <= DBL_MAX)) dm = INFINITY;
        if (!(xm <= DBL_MAX)) xm = INFINITY;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        dm = fmax(dm, __shfl_xor(dm, off));
        xm = fmax(xm, __shfl_xor(xm, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&mx[2 * (b * nr + r)], (unsigned long long)__double_as_longlong(dm));
        atomicMax(&mx[2 * (b * nr + r) + 1], (unsigned long long)__double_as_longlong(xm));
    }
}

// max |V[b][r]| per (item, rhs) as ordered bit patterns (non-finite as +inf)
__global__ void k_maxabs(const double2 *V, int n, int64_t sV, int nr, const int *bidx, unsigned long long *mx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int br = blockIdx.y, b = bidx[br / nr], r = br % nr;
    double m = 0.0;
    if (i < n) {
        const double2 v = V[(size_t)b * sV + (size_t)r * n + i];
        m = hypot(v.x, v.y);
        if (!(m <= DBL_MAX)) m = INFINITY;
    }
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&mx[b * nr + r], (unsigned long long)__double_as_longlong(m));
}

__global__ void k_zero_words64(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

typedef std::complex<double> cd;

// Givens state of one (item, right-hand side) GMRES run.
struct Arnoldi {
    std::vector<cd> H;    // (m + 1) x m, column-major
    std::vector<cd> cs, sn, g;
    double beta = 0.0;
    bool done = false;
    int k = 0;            // columns built
};

}  // namespace

size_t gmres_bytes(int n, int nr, int m)
{
    // Krylov vectors V_0..V_m and preconditioned vectors z_0..z_{m-1}
    return (size_t)(2 * m + 1) * n * nr * sizeof(double2);
}

int gmres_ir(wm_ctx *c, const wm_solve_opts &o, const GmresArgs &a, const std::vector<int> &items,
             const std::vector<int> &fac, GmresOut &out, hipStream_t s)
{
    const int n = a.n, nr = a.nr, m = a.m, cnt = a.count;
    const int64_t sV = a.sV, setS = (int64_t)cnt * sV;
    double2 *Vb = a.K, *Zb = a.K + (size_t)(m + 1) * setS;
    std::vector<int> active = items;
    const int nit = (int)items.size();
    std::vector<double> prev((size_t)cnt * nr, 1e300), first(cnt, 0.0), last(cnt, HUGE_VAL);
    // fac per item (indexed by item)
    std::vector<int> facOf(cnt, -1);
    for (int q = 0; q < nit; ++q) facOf[items[q]] = fac[q];
    const int outer_max = std::max(1, std::min(o.max_iter, 8));
    std::vector<unsigned long long> hm((size_t)2 * cnt * nr);
    std::vector<double2> hred;

    // the inverted diagonal blocks of every factor in use: the preconditioner
    // applies them as products (its solves need not be backward stable; the
    // FP64 residuals decide the result)
    std::vector<int> facs, slot_of;
    for (int f : fac) {
        if (f < 0) continue;
        if ((size_t)f >= slot_of.size()) slot_of.resize((size_t)f + 1, -1);
        if (slot_of[f] < 0) {
            slot_of[f] = (int)facs.size();
            facs.push_back(f);
        }
    }
    WM_TRY(c->ws_dinv.ensure(diag_inv_bytes(n, std::max<int>(1, (int)facs.size()))));
    double2 *dinv = (double2 *)c->ws_dinv.ptr;
    WM_TRY(diag_inv_batched(c, a.fprec, n, a.F, n, a.sF, facs, dinv, s));

    std::vector<int> uploaded;  // the list a.dact (and a.dfac) hold
    auto upload_idx = [&](const std::vector<int> &lst) -> int {
        if (lst == uploaded) return WM_OK;
        std::vector<int> f(lst.size());
        for (size_t q = 0; q < lst.size(); ++q) f[q] = facOf[lst[q]];
        uploaded.clear();
        WM_TRY(copy_h2d(a.dact, lst.data(), lst.size() * sizeof(int), s));
        WM_TRY(copy_h2d(a.dfac, f.data(), f.size() * sizeof(int), s));
        uploaded = lst;
        return WM_OK;
    };
    const dim3 blk(256);
    auto gridv = [&](int members) { return dim3((unsigned)((n + 255) / 256), (unsigned)(members * nr)); };

    // X = 0, R = B; ||B||_inf per item for the backward error now, ||Z||_inf
    // from the first GMRES step's product Z z_0 (the same pass over Z)
    std::vector<double> znorm(cnt, 0.0), bnorm((size_t)cnt * nr, 0.0);
    std::vector<unsigned long long> hm2((size_t)cnt + (size_t)cnt * nr);
    auto bits_to = [](unsigned long long u) { double d; std::memcpy(&d, &u, 8); return d; };
    if (!active.empty()) {
        WM_TRY(upload_idx(active));
        const int na = (int)active.size();
        hipLaunchKernelGGL(k_zero_set, dim3((unsigned)((sV + 255) / 256), na), blk, 0, s, a.X, sV, a.dact);
        hipLaunchKernelGGL(k_copy_set, dim3((unsigned)((sV + 255) / 256), na), blk, 0, s, a.R, a.B, sV, a.dact);
        hipLaunchKernelGGL(k_zero_words64, dim3((unsigned)((hm2.size() + 255) / 256)), blk, 0, s, a.mx2, (int)hm2.size());
        hipLaunchKernelGGL(k_maxabs, gridv(na), blk, 0, s, a.B, n, sV, nr, a.dact, a.mx2 + cnt);
        WM_TRY(copy_d2h(hm2.data(), a.mx2, hm2.size() * sizeof(unsigned long long), s));
        for (int b : active)
            for (int r = 0; r < nr; ++r) bnorm[(size_t)b * nr + r] = bits_to(hm2[cnt + (size_t)b * nr + r]);
    }
    for (int outer = 0; outer < outer_max && !active.empty(); ++outer) {
        const int na = (int)active.size();
        const double itol = outer == 0 ? kInnerTol0 : kInnerTol;
        WM_TRY(upload_idx(active));
        // v_0 = R / ||R||
        hipLaunchKernelGGL(k_dots, dim3(na, 1, nr), blk, 0, s, a.R, (int64_t)0, 1, a.R, sV, n, nr, a.dact, a.red);
        hipLaunchKernelGGL(k_normalize, gridv(na), blk, 0, s, Vb, a.R, a.red, sV, n, nr, a.dact);
        hred.resize((size_t)na * nr);
        WM_TRY(copy_d2h(hred.data(), a.red, hred.size() * sizeof(double2), s));
        std::vector<Arnoldi> ar((size_t)na * nr);
        for (int q = 0; q < na * nr; ++q) {
            Arnoldi &A = ar[q];
            A.H.assign((size_t)(m + 1) * m, cd(0, 0));
            A.cs.assign(m, cd(0, 0));
            A.sn.assign(m, cd(0, 0));
            A.g.assign(m + 1, cd(0, 0));
            A.beta = std::sqrt(std::max(0.0, hred[q].x));
            A.g[0] = cd(A.beta, 0);
            A.done = !(A.beta > 0.0) || !std::isfinite(A.beta);
        }
        // members whose columns still iterate; all columns of a member iterate together
        std::vector<int> inner(na);
        for (int q = 0; q < na; ++q) inner[q] = q;
        auto member_done = [&](int q) {
            for (int r = 0; r < nr; ++r)
                if (!ar[(size_t)q * nr + r].done) return false;
            return true;
        };
        for (int k = 0; k < m; ++k) {
            std::vector<int> live, lst;
            for (int q : inner)
                if (!member_done(q)) { live.push_back(q); lst.push_back(active[q]); }
            inner.swap(live);
            if (inner.empty()) break;
            const int nl = (int)inner.size();
            WM_TRY(upload_idx(lst));
            double2 *vk = Vb + (size_t)k * setS, *zk = Zb + (size_t)k * setS, *w = Vb + (size_t)(k + 1) * setS;
            // z_k = M^-1 v_k (the anchor's factors), w = Z z_k
            hipLaunchKernelGGL(k_copy_set, dim3((unsigned)((sV + 255) / 256), nl), blk, 0, s, zk, vk, sV, a.dact);
            std::vector<int> fl(lst.size());
            for (size_t q = 0; q < lst.size(); ++q) fl[q] = facOf[lst[q]];
            WM_TRY(getrs_grouped(c, a.fprec, n, nr, a.F, n, a.sF, a.perm, n, zk, sV, lst, fl, s, dinv, &slot_of,
                                 c->group_cols));
            WM_TRY(residual_batched(c, n, nr, a.Z, n, a.sZ, zk, nullptr, n, sV, w, nl, s, a.dact,
                                    outer == 0 && k == 0 ? a.mx2 : nullptr, a.np));
            // classical Gram-Schmidt, twice; the coefficients stay on the device
            double2 *h1 = a.red, *h2 = a.red + (size_t)nl * (m + 1) * nr, *nw = h2 + (size_t)nl * (m + 1) * nr;
            hipLaunchKernelGGL(k_dots, dim3(nl, k + 1, nr), blk, 0, s, Vb, setS, k + 1, w, sV, n, nr, a.dact, h1);
            hipLaunchKernelGGL(k_axpy, gridv(nl), blk, 0, s, w, Vb, setS, k + 1, h1, k + 1, sV, n, nr, a.dact);
            hipLaunchKernelGGL(k_dots, dim3(nl, k + 1, nr), blk, 0, s, Vb, setS, k + 1, w, sV, n, nr, a.dact, h2);
            hipLaunchKernelGGL(k_axpy, gridv(nl), blk, 0, s, w, Vb, setS, k + 1, h2, k + 1, sV, n, nr, a.dact);
            hipLaunchKernelGGL(k_dots, dim3(nl, 1, nr), blk, 0, s, w, (int64_t)0, 1, w, sV, n, nr, a.dact, nw);
            hipLaunchKernelGGL(k_normalize, gridv(nl), blk, 0, s, w, w, nw, sV, n, nr, a.dact);
            WM_TRY(check_device_error(c, s));
            // h1, h2 and the norms in one copy (they lie in that order in a.red)
            const size_t hs = (size_t)nl * (m + 1) * nr;
            std::vector<double2> hh(2 * hs + (size_t)nl * nr);
            WM_TRY(copy_d2h(hh.data(), h1, hh.size() * sizeof(double2), s));
            const double2 *a1 = hh.data(), *a2 = hh.data() + hs, *an = hh.data() + 2 * hs;
            for (int p = 0; p < nl; ++p) {
                const int q = inner[p];
                for (int r = 0; r < nr; ++r) {
                    Arnoldi &A = ar[(size_t)q * nr + r];
                    if (A.done) continue;
                    cd *col = &A.H[(size_t)k * (m + 1)];
                    for (int j = 0; j <= k; ++j) {
                        const double2 x = a1[((size_t)p * (k + 1) + j) * nr + r], y = a2[((size_t)p * (k + 1) + j) * nr + r];
                        col[j] = cd(x.x + y.x, x.y + y.y);
                    }
                    const double hk = std::sqrt(std::max(0.0, an[(size_t)p * nr + r].x));
                    col[k + 1] = cd(hk, 0);
                    for (int j = 0; j < k; ++j) {  // earlier rotations
                        const cd t = std::conj(A.cs[j]) * col[j] + std::conj(A.sn[j]) * col[j + 1];
                        col[j + 1] = -A.sn[j] * col[j] + A.cs[j] * col[j + 1];
                        col[j] = t;
                    }
                    const double den = std::hypot(std::abs(col[k]), hk);
                    if (!(den > 0.0) || !std::isfinite(den)) {
                        A.done = true;  // breakdown: the space is exhausted (or the run is not finite)
                        A.k = k;
                        continue;
                    }
                    A.cs[k] = col[k] / den;
                    A.sn[k] = cd(hk / den, 0);
                    col[k] = cd(den, 0);
                    col[k + 1] = cd(0, 0);
                    A.g[k + 1] = -A.sn[k] * A.g[k];
                    A.g[k] = std::conj(A.cs[k]) * A.g[k];
                    A.k = k + 1;
                    if (std::abs(A.g[k + 1]) <= itol * A.beta) A.done = true;
                }
            }
            out.inner_iters += nl;
            for (int q : inner) ++out.iters[active[q]];
        }
        if (outer == 0) {  // ||Z||_inf of the items, from the first step's product
            WM_TRY(copy_d2h(hm2.data(), a.mx2, (size_t)cnt * sizeof(unsigned long long), s));
            for (int b : active) znorm[b] = bits_to(hm2[b]);
        }
        // correction D = sum_j y_j z_j per member, y from the triangular system R y = g
        std::vector<double2> yv((size_t)na * m * nr, make_double2(0.0, 0.0));
        int kmax = 0;
        for (int q = 0; q < na; ++q)
            for (int r = 0; r < nr; ++r) {
                Arnoldi &A = ar[(size_t)q * nr + r];
                const int k = A.k;
                kmax = std::max(kmax, k);
                std::vector<cd> y(k);
                for (int i = k - 1; i >= 0; --i) {
                    cd t = A.g[i];
                    for (int j = i + 1; j < k; ++j) t -= A.H[(size_t)j * (m + 1) + i] * y[j];
                    y[i] = t / A.H[(size_t)i * (m + 1) + i];
                }
                for (int j = 0; j < k; ++j) yv[((size_t)q * m + j) * nr + r] = make_double2(y[j].real(), y[j].imag());
            }
        WM_TRY(upload_idx(active));
        WM_TRY(copy_h2d(a.red, yv.data(), yv.size() * sizeof(double2), s));
        hipLaunchKernelGGL(k_combine, gridv(na), blk, 0, s, a.D, Zb, setS, kmax, a.red, m, sV, n, nr, a.dact);
        hipLaunchKernelGGL(k_zero_words64, dim3((2 * cnt * nr + 255) / 256), blk, 0, s, a.mx, 2 * cnt * nr);
        hipLaunchKernelGGL(k_update_max, gridv(na), blk, 0, s, a.X, a.D, n, sV, nr, a.dact, a.mx);
        WM_TRY(residual_batched(c, n, nr, a.Z, n, a.sZ, a.X, a.B, n, sV, a.R, na, s, a.dact));
        hipLaunchKernelGGL(k_zero_words64, dim3((unsigned)((cnt * nr + 255) / 256)), blk, 0, s, a.mx2 + cnt, cnt * nr);
        hipLaunchKernelGGL(k_maxabs, gridv(na), blk, 0, s, a.R, n, sV, nr, a.dact, a.mx2 + cnt);
        WM_TRY(check_device_error(c, s));
        WM_TRY(copy_d2h(hm.data(), a.mx, hm.size() * sizeof(unsigned long long), s));
        WM_TRY(copy_d2h(hm2.data() + cnt, a.mx2 + cnt, (size_t)cnt * nr * sizeof(unsigned long long), s));
        std::vector<int> still;
        for (int q = 0; q < na; ++q) {
            const int b = active[q];
            double dm[64], xm[64];
            for (int r = 0; r < nr; ++r) {
                std::memcpy(&dm[r], &hm[2 * ((size_t)b * nr + r)], 8);
                std::memcpy(&xm[r], &hm[2 * ((size_t)b * nr + r) + 1], 8);
            }
            bool inner_ok = true;  // every column's inner solve reached its tolerance
            for (int r = 0; r < nr; ++r) {
                const Arnoldi &A = ar[(size_t)q * nr + r];
                inner_ok = inner_ok && (A.beta == 0.0 || (A.k > 0 && std::abs(A.g[A.k]) <= itol * A.beta));
            }
            // normwise backward error of X (the worst right-hand side)
            double berr = 0.0;
            for (int r = 0; r < nr; ++r) {
                const double rn = bits_to(hm2[cnt + (size_t)b * nr + r]);
                const double den = znorm[b] * xm[r] + bnorm[(size_t)b * nr + r];
                berr = std::max(berr, den > 0.0 ? rn / den : (rn > 0.0 ? HUGE_VAL : 0.0));
            }
            if (!std::isfinite(berr)) berr = HUGE_VAL;
            out.berr[b] = berr;
            double rate = -1.0;
            const int v = refine_decide(nr, dm, xm, &prev[(size_t)b * nr], outer, o.tol, outer_max, first[b], rate,
                                        out.dx_rel[b]);
            if (v >= 0 && inner_ok && berr <= kBerrTol) {
                out.conv[b] = 1;  // as accurate as the matrix allows in FP64
                last[b] = out.dx_rel[b];
                continue;
            }
            // An inner solve to kInnerTol contracts the error by about that factor,
            // so a correction that does not halve after one marks the FP64 floor:
            // the solution is as accurate as the working precision allows (where
            // kappa u64 > tol, the correction cannot fall below tol).
            const bool floor = v <= 0 && outer >= 1 && inner_ok && std::isfinite(out.dx_rel[b]) &&
                               out.dx_rel[b] > 0.5 * last[b];
            last[b] = out.dx_rel[b];
            if (v > 0 || floor) out.conv[b] = 1;
            else if (v == 0) still.push_back(b);
        }
        active.swap(still);
    }
    return WM_OK;
}

}  // namespace wm
