// Batch solver: fill, factor and solve a set of (geometry, frequency, feeds)
// items of one padded size on one stream.  Used by wm_solve_sweep (one
// geometry, many frequencies) and by the request server (many requests).
//
// Replaces the per-frequency body of the reference's loop, CalculateByFreq
// ("wersja gpu/Wire_MoM.cu":355-413): wavenumber, fill, delta-gap excitation,
// LU solve, Zin = V/I at the feed.
//
// Precision forms:
//   WM_PREC_F64    : complex128 fill, complex128 LU and solve.
//   WM_PREC_F32    : complex128 fill kept for residuals, complex64 copy
//                    factored on FP32 matrix cores; FP64 iterative
//                    refinement x += LU32^{-1}(b - Z x) until
//                    ||dx||/||x|| <= tol or the correction stalls at the
//                    FP64 floor; items that do not converge are refactored in
//                    FP64 (fallback).
//   WM_PREC_BF16X3 : as WM_PREC_F32, with the factorization's trailing
//                    updates taken as split-BF16x3 products (panels in FP32).
// Each item's arithmetic depends only on that item (per-item convergence
// mask, fixed reduction splits), so an item's result is the same alone or
// in any batch.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "wm_gmres.h"
#include "wm_internal.h"

namespace wm {

namespace {

constexpr double kPi = 3.14159265358979323846;

// Z[f] (n x n, leading dimension n) = T[f] (lg x lg) padded with the identity,
// and the complex64 copy alike when Z32 is given.
__global__ void k_expand(double2 *Z, int64_t sZ, float2 *Z32, int n, const double2 *T, const float2 *T32, int64_t sT,
                         int lg)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int f = blockIdx.y;
    const int i = (int)(e % n), j = (int)(e / n);
    const bool in = i < lg && j < lg;
    const int64_t t = (int64_t)f * sT + i + (int64_t)j * lg;
    const double2 v = in ? T[t] : make_double2(i == j ? 1.0 : 0.0, 0.0);
    Z[(int64_t)f * sZ + e] = v;
    if (Z32) Z32[(int64_t)f * sZ + e] = in ? T32[t] : make_float2(i == j ? 1.f : 0.f, 0.f);
}

__global__ void k_zero2(double2 *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}

// Delta-gap right-hand sides: up to ne entries (column, index, value) per
// item, index -1 for none; entries of one item name distinct positions.
__global__ void k_set_rhs(double2 *B, int64_t strideB, int ldb, int ne, const int *col, const int *idx,
                          const double2 *val, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * ne) return;
    const int b = t / ne;
    const int i = idx[t];
    if (i >= 0) B[(size_t)b * strideB + (size_t)col[t] * ldb + i] = val[t];
}

// Lumped loads: Z[b][i][i] += val for up to nl entries (index, value) per
// item (index -1 for none, distinct indices per item); the complex64 copy,
// if any, takes the rounded new diagonal.
__global__ void k_add_diag(double2 *Z, int64_t strideZ, int ld, float2 *Z32, const int *idx, const double2 *val,
                           int nl, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nl) return;
    const int b = t / nl, i = idx[t];
    if (i < 0) return;
    const size_t o = (size_t)b * strideZ + (size_t)i * ld + i;
    double2 z = Z[o];
    z.x += val[t].x;
    z.y += val[t].y;
    Z[o] = z;
    if (Z32) Z32[o] = make_float2((float)z.x, (float)z.y);
}

// Incident plane wave: B[m] += sum_h l_h . E_inc(r_h) over the two test points
// of basis m (obs: rc1, l1, rc2, l2 as 12 arrays of ld doubles), E_inc(r) =
// E exp(j k rh.r) with E = e_theta th + e_phi ph (complex vector, 6 doubles).
struct PlaneWaveArgs {
    double rh[3];
    double ev[6];  // re x, re y, re z, im x, im y, im z
    double ka;
};

__global__ void k_plane_wave(double2 *B, const double *__restrict__ obs, int ld, int npls, PlaneWaveArgs a)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= npls) return;
    double vr = 0.0, vi = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const double rx = obs[(size_t)(6 * h + 0) * ld + m], ry = obs[(size_t)(6 * h + 1) * ld + m],
                     rz = obs[(size_t)(6 * h + 2) * ld + m];
        const double lx = obs[(size_t)(6 * h + 3) * ld + m], ly = obs[(size_t)(6 * h + 4) * ld + m],
                     lz = obs[(size_t)(6 * h + 5) * ld + m];
        double sn, cs;
        sincos(a.ka * (a.rh[0] * rx + a.rh[1] * ry + a.rh[2] * rz), &sn, &cs);
        const double dr = lx * a.ev[0] + ly * a.ev[1] + lz * a.ev[2];
        const double di = lx * a.ev[3] + ly * a.ev[4] + lz * a.ev[5];
        vr += dr * cs - di * sn;
        vi += dr * sn + di * cs;
    }
    double2 b = B[m];
    b.x += vr;
    b.y += vi;
    B[m] = b;
}

PlaneWaveArgs plane_wave_args(const wm_plane_wave &p, double ka)
{
    const double d2r = kPi / 180.0;
    const double t = p.theta_deg * d2r, f = p.phi_deg * d2r;
    const double st = std::sin(t), ct = std::cos(t), sf = std::sin(f), cf = std::cos(f);
    const double th[3] = {ct * cf, ct * sf, -st}, ph[3] = {-sf, cf, 0.0};
    PlaneWaveArgs a;
    a.rh[0] = st * cf;
    a.rh[1] = st * sf;
    a.rh[2] = ct;
    for (int d = 0; d < 3; ++d) {
        a.ev[d] = p.e_theta.re * th[d] + p.e_phi.re * ph[d];
        a.ev[3 + d] = p.e_theta.im * th[d] + p.e_phi.im * ph[d];
    }
    a.ka = ka;
    return a;
}

double2 source_value(float ampl, float phas)  // ampl*exp(j*pi*phas/180), "wersja referencyjna/Wire_MoM.cu":307-309
{
    const double arg = kPi * (double)phas / 180.;
    return make_double2((double)ampl * std::cos(arg), (double)ampl * std::sin(arg));
}

// Series load impedance at f MHz: r + j(2 pi f l - 1/(2 pi f c)), l in uH, c in uF.
double2 load_value(const wm_load &l, float f)
{
    double x = 2. * kPi * (double)f * l.l_uh;
    if (l.c_uf != 0.0) x -= 1.0 / (2. * kPi * (double)f * l.c_uf);
    return make_double2(l.r_ohm, x);
}

// X += D for the listed items; max |D|, |X| per (item, rhs) as ordered bit patterns.
__global__ void k_update(double2 *X, const double2 *D, int n, int ldb, int64_t strideB, int nr, const int *bidx,
                         unsigned long long *mx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int br = blockIdx.y, b = bidx[br / nr], r = br % nr;
    const int bb = b * nr + r;
    double dm = 0.0, xm = 0.0;
    if (i < n) {
        const size_t o = (size_t)b * strideB + (size_t)r * ldb + i;
        const double2 d = D[o];
        double2 x = X[o];
        x.x += d.x;
        x.y += d.y;
        X[o] = x;
        dm = hypot(d.x, d.y);
        xm = hypot(x.x, x.y);
        // fmax drops NaN operands: a non-finite correction or iterate is
        // reported as +inf, which the host takes as divergence
        if (!(dm <= DBL_MAX)) dm = INFINITY;
        if (!(xm <= DBL_MAX)) xm = INFINITY;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        dm = fmax(dm, __shfl_xor(dm, off));
        xm = fmax(xm, __shfl_xor(xm, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&mx[2 * bb], (unsigned long long)__double_as_longlong(dm));
        atomicMax(&mx[2 * bb + 1], (unsigned long long)__double_as_longlong(xm));
    }
}

// Dst[b] = Src[b] for the listed items (n x nr complex128 per item).
__global__ void k_copy_items(double2 *Dst, const double2 *Src, int64_t stride, int64_t len, const int *bidx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const size_t o = (size_t)bidx[blockIdx.y] * stride + i;
    Dst[o] = Src[o];
}

// Dst[bidx[y]] = 0 over len entries (items at a fixed stride).
__global__ void k_zero_items(double2 *Dst, int64_t stride, int64_t len, const int *bidx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) Dst[(int64_t)bidx[blockIdx.y] * stride + i] = make_double2(0.0, 0.0);
}

// Complex64 copy of complex128 matrices bidx[y] (nn entries each).
__global__ void k_to_c64(float2 *Z32, const double2 *Z, int64_t nn, int64_t stride, const int *bidx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int64_t o = (int64_t)bidx[blockIdx.y] * stride + i;
    const double2 z = Z[o];
    Z32[o] = make_float2((float)z.x, (float)z.y);
}

// 1-norm of Z (largest column sum of moduli over the leading np[b] x np[b]
// block) per item, as the ordered bit pattern of a non-negative double.  One
// wave per column.
__global__ void k_norm1(const double2 *Z, int64_t strideZ, int ld, const int *np, const int *bidx,
                        unsigned long long *out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = blockIdx.x * 4 + wave, bl = blockIdx.y, b = bidx[bl];
    const int n = np[bl];
    if (col >= n) return;  // wave-uniform
    const double2 *zc = Z + (size_t)b * strideZ + (size_t)col * ld;
    double sum = 0.0;
    for (int i = lane; i < n; i += 64) sum += hypot(zc[i].x, zc[i].y);
    for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) atomicMax(&out[bl], (unsigned long long)__double_as_longlong(sum));
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

wm_c128 cdiv(wm_c128 a, wm_c128 b)  // Smith (1962)
{
    wm_c128 q;
    if (std::fabs(b.re) >= std::fabs(b.im)) {
        const double t = b.im / b.re, d = b.re + b.im * t;
        q.re = (a.re + a.im * t) / d;
        q.im = (a.im - a.re * t) / d;
    } else {
        const double t = b.re / b.im, d = b.im + b.re * t;
        q.re = (a.re * t + a.im) / d;
        q.im = (a.im * t - a.re) / d;
    }
    return q;
}

}  // namespace

// Estimates ||A^-1||_1 for the items lst (factors F, complex128 for prec
// WM_PREC_F64, complex64 otherwise, permutations perm; A restricted to its
// leading np[i] x np[i] block, the padded identity excluded) with Hager's
// method as refined by Higham (ACM TOMS 14:381-396, 1988; the algorithm of
// LAPACK's zlacon/zlacn2): from x = (1/n, ..., 1/n), alternate y = A^-1 x and
// z = A^-H sign(y), moving x to the unit vector at the largest |z_j| while the
// estimate ||y||_1 grows (at most 5 steps), then take the larger of that and
// 2 ||A^-1 v||_1 / (3n) for v_i = (-1)^i (1 + i / (n - 1)).  Items advance in
// lockstep; each round solves, in one batched launch per direction, the items
// that need it.  V (device, stride sV) holds the vectors; dact is the index
// table buffer.
int inv_norm1_estimate(wm_ctx *c, int prec, int n, const void *F, int64_t sF, const int *perm, double2 *V,
                       int64_t sV, const std::vector<int> &lst, const std::vector<int> &np, int *dact, hipStream_t s,
                       std::vector<double> &est)
{
    const int m = (int)lst.size();
    est.assign(m, 0.0);
    enum Step { FIRST, SIGN_Z, UNIT_Y, UNIT_Z, ALT, DONE };
    struct St {
        Step step = FIRST;
        int j = 0, iter = 0;
        double estold = 0.0;
    };
    std::vector<St> st(m);
    std::vector<std::vector<double2>> vec(m, std::vector<double2>((size_t)n, make_double2(0.0, 0.0)));
    auto set_first = [&](int i) {
        for (int r = 0; r < n; ++r) vec[i][r] = make_double2(r < np[i] ? 1.0 / np[i] : 0.0, 0.0);
    };
    auto set_unit = [&](int i, int j) {
        for (int r = 0; r < n; ++r) vec[i][r] = make_double2(r == j ? 1.0 : 0.0, 0.0);
    };
    auto set_alt = [&](int i) {
        const int k = np[i];
        for (int r = 0; r < n; ++r)
            vec[i][r] = make_double2(r < k ? (r % 2 ? -1.0 : 1.0) * (1.0 + (k > 1 ? (double)r / (k - 1) : 0.0)) : 0.0, 0.0);
    };
    auto norm1 = [&](int i) {
        double t = 0.0;
        for (int r = 0; r < np[i]; ++r) t += std::hypot(vec[i][r].x, vec[i][r].y);
        return t;
    };
    auto to_sign = [&](int i) {  // y_r / |y_r| (1 where y_r = 0), zero on the padding
        for (int r = 0; r < n; ++r) {
            const double a = std::hypot(vec[i][r].x, vec[i][r].y);
            vec[i][r] = r >= np[i] ? make_double2(0.0, 0.0)
                        : a > DBL_MIN ? make_double2(vec[i][r].x / a, vec[i][r].y / a) : make_double2(1.0, 0.0);
        }
    };
    auto argmax = [&](int i) {
        int j = 0;
        double best = -1.0;
        for (int r = 0; r < np[i]; ++r) {
            const double a = std::hypot(vec[i][r].x, vec[i][r].y);
            if (a > best) { best = a; j = r; }
        }
        return j;
    };
    for (int i = 0; i < m; ++i) set_first(i);
    for (int round = 0; round < 32; ++round) {
        // FIRST, UNIT_Y, ALT want y = A^-1 x; SIGN_Z, UNIT_Z want z = A^-H x
        for (int h = 0; h < 2; ++h) {
            std::vector<int> sel, mem;
            for (int i = 0; i < m; ++i) {
                const bool wants_h = st[i].step == SIGN_Z || st[i].step == UNIT_Z;
                if (st[i].step != DONE && wants_h == (h == 1)) { sel.push_back(i); mem.push_back(lst[i]); }
            }
            if (sel.empty()) continue;
            const int k = (int)sel.size();
            for (int q = 0; q < k; ++q)
                WM_TRY(copy_h2d(V + (size_t)mem[q] * sV, vec[sel[q]].data(), (size_t)n * sizeof(double2), s));
            WM_TRY(copy_h2d(dact, mem.data(), k * sizeof(int), s));
            if (h == 0) WM_TRY(getrs_perm_batched(c, prec, n, 1, F, n, sF, perm, n, V, n, sV, k, s, dact));
            else WM_TRY(getrsh_perm_batched(c, prec, n, 1, F, n, sF, perm, n, V, n, sV, k, s, dact));
            WM_TRY(check_device_error(c, s));
            for (int q = 0; q < k; ++q)
                WM_TRY(copy_d2h(vec[sel[q]].data(), V + (size_t)mem[q] * sV, (size_t)n * sizeof(double2), s));
            for (int i : sel) {
                St &t = st[i];
                switch (t.step) {
                    case FIRST:
                        est[i] = norm1(i);
                        if (np[i] == 1) { t.step = DONE; break; }
                        to_sign(i);
                        t.step = SIGN_Z;
                        break;
                    case SIGN_Z:
                        t.j = argmax(i);
                        t.iter = 2;
                        set_unit(i, t.j);
                        t.step = UNIT_Y;
                        break;
                    case UNIT_Y:
                        t.estold = est[i];
                        est[i] = norm1(i);
                        if (est[i] <= t.estold) { est[i] = std::max(est[i], t.estold); set_alt(i); t.step = ALT; break; }
                        to_sign(i);
                        t.step = UNIT_Z;
                        break;
                    case UNIT_Z: {
                        const int jlast = t.j;
                        t.j = argmax(i);
                        const double zl = std::hypot(vec[i][jlast].x, vec[i][jlast].y);
                        const double zj = std::hypot(vec[i][t.j].x, vec[i][t.j].y);
                        if (zl != zj && t.iter < 5) {
                            ++t.iter;
                            set_unit(i, t.j);
                            t.step = UNIT_Y;
                        } else {
                            set_alt(i);
                            t.step = ALT;
                        }
                        break;
                    }
                    case ALT: {
                        const double temp = 2.0 * norm1(i) / (3.0 * np[i]);
                        if (temp > est[i]) est[i] = temp;
                        t.step = DONE;
                        break;
                    }
                    default: break;
                }
            }
        }
        bool all = true;
        for (int i = 0; i < m; ++i) all = all && st[i].step == DONE;
        if (all) break;
    }
    return WM_OK;
}

int refine_decide(int nrhs, const double *dm, const double *xm, double *prev, int it, double tol, int max_iter,
                  double &first, double &rate, double &worst)
{
    worst = 0.0;
    bool stalled = true, finite = true;
    for (int r = 0; r < nrhs; ++r) {
        if (!std::isfinite(dm[r]) || !std::isfinite(xm[r])) finite = false;
        const double rel = xm[r] > 0 ? dm[r] / xm[r] : (dm[r] > 0 ? HUGE_VAL : 0.0);
        worst = std::max(worst, rel);
        double &p = prev[r];
        if (!(rel > 0.5 * p && p < 1e-9)) stalled = false;
        p = rel;
    }
    if (!finite || !(worst < HUGE_VAL)) {  // a non-finite iterate: these factors cannot refine
        rate = HUGE_VAL;
        return -1;
    }
    // recorded rate: geometric mean contraction since the first correction
    if (it == 0) first = worst;
    if (it >= 1 && !stalled && first > 0) rate = std::pow(worst / first, 1.0 / it);
    else if (it == 0 && worst <= tol) rate = worst;  // converged at once: rate at most this
    if (worst <= tol || stalled) return 1;
    // contraction too slow to reach tol within max_iter: stop now
    if (it >= 1) {
        const double need = rate > 0 && rate < 1 ? std::log(tol / worst) / std::log(rate) : 1e9;
        if (it + 1 + need > max_iter) return -1;
    }
    if (it + 1 >= max_iter) return -1;  // out of iterations
    return 0;
}

double wavenumber_mhz(float freq_mhz)  // "wersja referencyjna/Wire_MoM.cu":5-6,338-340
{
    const double ambda = 2.99792448E+8 / (double)freq_mhz * 1E-6;
    return (kPi + kPi) / ambda;
}

size_t item_bytes(int ld, int nrhs, int prec)
{
    const size_t nn = (size_t)ld * ld;
    size_t b = nn * sizeof(double2);
    if (prec != WM_PREC_F64) b += nn * sizeof(float2);
    b += (size_t)ld * nrhs * sizeof(double2) * 4 + (size_t)ld * sizeof(int) + 1024;
    return b;
}

int auto_batch(wm_ctx *ctx, int ld, int nrhs, int prec, double frac)
{
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return 1;
    if (!(frac > 0 && frac <= 1)) frac = 0.8;
    const size_t have = (size_t)(frac * (double)(freeb + ctx->ws_sweep.bytes));
    int b = (int)std::max<size_t>(1, have / item_bytes(ld, nrhs, prec));
    return std::min(b, 64);
}

double predict_rate(const wm_geom *g, float f, int prec)
{
    std::lock_guard<std::mutex> lk(g->rate_mu);
    for (int same = 1; same >= 0; --same) {
        const wm_geom::RateObs *above = nullptr, *below = nullptr;
        for (const auto &ob : g->rate_obs) {
            if ((ob.prec == prec) != (same == 1)) continue;
            if (ob.freq >= f) {
                if (!above || ob.freq < above->freq) above = &ob;
            } else if (!below || ob.freq > below->freq) {
                below = &ob;
            }
        }
        const auto *ob = above ? above : below;
        if (!ob) continue;
        const double sc = (double)ob->freq / (double)f;
        double r = ob->rate * sc * sc * sc;
        if (!same) r = prec == WM_PREC_BF16X3 ? r * kBf16x3RateFactor : r / kBf16x3RateFactor;
        return r;
    }
    return -1.0;
}

void record_rates(const WorkItem *items, int count)
{
    auto put = [](const wm_geom *g, float f, int prec, double rate) {
        std::lock_guard<std::mutex> lk(g->rate_mu);
        for (auto &ob : g->rate_obs)
            if (ob.freq == f && ob.prec == prec) {
                ob.rate = rate;
                return;
            }
        g->rate_obs.push_back(wm_geom::RateObs{f, prec, rate});
    };
    for (int b = 0; b < count; ++b) {
        const WorkItem &w = items[b];
        if (w.rate_b3 >= 0) put(w.g, w.freq, WM_PREC_BF16X3, w.rate_b3);
        if (w.rate >= 0 && (w.fprec == WM_PREC_BF16X3 || w.fprec == WM_PREC_F32)) put(w.g, w.freq, w.fprec, w.rate);
    }
}

// Panel launches of a lockstep factorization of `count` matrices of size n
// when `width` full-height panels fit in one launch: the panel at row s*64
// (height n - 64 s) goes in ceil(count (n - 64 s) / (width n)) launches.
double panel_launches(int n, int count, int width)
{
    double t = 0.0;
    for (int h = n; h > 0; h -= 64) t += std::ceil((double)count * h / ((double)width * n));
    return t;
}

// Cost constants measured on MI355X at N = 8192 (lockstep batches of 48
// complex128 and 16 complex64 matrices): a panel launch costs 64 columns x
// 7.4 us (complex128) or x 7.3 us (complex64) whatever the cluster size; the
// products and interchanges of a factorization run at 76 TF/s (FP64),
// 122 TF/s (FP32) and 180 TF/s FP32-equivalent (split-BF16x3), all counted at
// 8 real flops per complex multiply-add; a refinement iteration costs two
// triangular-solve latency chains per batch (n/64 blocks x 8 us each) plus
// 24 N^2 bytes per member at about 5 TB/s; about 6 iterations.
constexpr double kPanelColF64 = 7.4e-6, kPanelColC64 = 7.3e-6;
constexpr double kRateF64 = 76e12, kRateF32 = 122e12, kRateB3 = 180e12;
constexpr double kComplex64Margin = 1.2;

double refine_cost(double n, int count)
{
    return 6.0 * (2.0 * n / 64 * 8e-6 + count * 24.0 * n * n / 5e12);
}

bool prefer_f64(wm_ctx *c, int prec, int n, int count)
{
    if (prec == WM_PREC_F64 || count <= 0) return false;
    const double nn = (double)n, flu = 8.0 / 3.0 * nn * nn * nn;
    const int w64 = getrf_width(c, WM_PREC_F64, n), w32 = getrf_width(c, prec, n);
    const double t64 = 64 * kPanelColF64 * panel_launches(n, count, w64) + count * flu / kRateF64;
    const double rate = prec == WM_PREC_BF16X3 ? kRateB3 : kRateF32;
    const double tmix = 64 * kPanelColC64 * panel_launches(n, count, w32) + count * flu / rate + refine_cost(nn, count);
    // complex64 factors must be predicted clearly cheaper: the first batch of a
    // structure is routed before any contraction rate is known, a member that
    // then fails costs an FP64 factorization on top, and the rates above were
    // measured at n = 8192 (smaller updates run slower); at n = 4096 with 12
    // frequencies the model's 9 % margin for BF16x3 turned into a 1.9x loss
    return t64 <= kComplex64Margin * tmix;
}

bool join_f64(wm_ctx *c, int prec, int n, int r)
{
    if (prec == WM_PREC_F64 || r <= 0) return false;
    // complex128 members of a large lockstep batch cost their products plus a
    // share of its panel launches (n x 7.4 us / (2 x launch width)); a separate
    // complex64 batch pays its own panel launches, products and refinement
    const double nn = (double)n, flu = 8.0 / 3.0 * nn * nn * nn;
    const int w64 = getrf_width(c, WM_PREC_F64, n), w32 = getrf_width(c, prec, n);
    const double t_join = r * (flu / kRateF64 + nn * kPanelColF64 / (2.0 * w64));
    const double rate = prec == WM_PREC_BF16X3 ? kRateB3 : kRateF32;
    const double t_sep = 64 * kPanelColC64 * panel_launches(n, r, w32) + r * flu / rate + refine_cost(nn, r);
    return t_join <= t_sep;
}

void route_items(const wm_solve_opts &o, WorkItem *items, int count)
{
    // rate that reaches tol from a first correction of relative size 1 within
    // max_iter: the iterations of a batch share their triangular-solve launches
    // (about 3 ms per iteration for a batch at N = 8192), so even slow
    // contraction costs less than FP64 factors in a batch of their own; an item
    // whose measured contraction falls short stops early and is refactored
    const double rmax = o.max_iter > 1 ? std::pow(o.tol, 1.0 / o.max_iter) : 0.0;
    for (int b = 0; b < count; ++b) {
        int p = o.prec;
        if (p == WM_PREC_BF16X3) {
            const double r = predict_rate(items[b].g, items[b].freq, WM_PREC_BF16X3);
            if (r > rmax) p = WM_PREC_F32;
        }
        if (p == WM_PREC_F32) {
            const double r = predict_rate(items[b].g, items[b].freq, WM_PREC_F32);
            if (r > rmax) p = WM_PREC_F64;
        }
        items[b].fprec = p;
    }
}

int solve_items(wm_ctx *c, const wm_solve_opts &o, WorkItem *items, int count, hipStream_t s)
{
    if (count <= 0) return WM_OK;
    // items of different sizes share the largest padded size: each matrix is
    // the identity beyond its own (the factorization skips that part)
    int n = 0;
    for (int b = 0; b < count; ++b) n = std::max(n, items[b].g->ld);
    bool uneven = false;
    int ld_small = 0;  // largest padded size below n
    for (int b = 0; b < count; ++b)
        if (items[b].g->ld != n) {
            uneven = true;
            ld_small = std::max(ld_small, items[b].g->ld);
        }
    int nr = 0;
    for (int b = 0; b < count; ++b) {
        const WorkItem &w = items[b];
        const int np = w.g->npls;
        if (!w.gaps) {  // generators and plane waves
            for (int q = 0; q < w.nsrc; ++q)
                if (w.src[q].basis < 1 || w.src[q].basis > np || w.src[q].rhs < 0 || w.src[q].rhs >= w.nrhs) {
                    set_error("generator %d: basis %d outside 1..%d or right-hand side %d outside 0..%d", q,
                              w.src[q].basis, np, w.src[q].rhs, w.nrhs - 1);
                    return WM_ERR_ARG;
                }
            for (int q = 0; q < w.npw; ++q)
                if (w.pw[q].rhs < 0 || w.pw[q].rhs >= w.nrhs) {
                    set_error("plane wave %d: right-hand side %d outside 0..%d", q, w.pw[q].rhs, w.nrhs - 1);
                    return WM_ERR_ARG;
                }
        } else {
            for (int r = 0; r < w.nrhs; ++r)
                if (w.gaps[r] < 1 || w.gaps[r] > np) {
                    set_error("feed basis %d outside 1..%d", w.gaps[r], np);
                    return WM_ERR_ARG;
                }
        }
        for (int q = 0; q < w.nload; ++q)
            if (w.loads[q].basis < 1 || w.loads[q].basis > np) {
                set_error("load %d on basis %d outside 1..%d", q, w.loads[q].basis, np);
                return WM_ERR_ARG;
            }
        nr = std::max(nr, w.nrhs);
    }
    // Right-hand-side entries per item (generators on one basis and column
    // summed), the ports whose V/I is reported, and the loads per basis (summed).
    struct Port {
        int col, idx;
        double2 v;
    };
    std::vector<std::vector<Port>> ent(count), ports(count), lds(count);
    auto merge = [](std::vector<Port> &v, const Port &p) {
        for (auto &e : v)
            if (e.col == p.col && e.idx == p.idx) {
                e.v.x += p.v.x;
                e.v.y += p.v.y;
                return;
            }
        v.push_back(p);
    };
    int ne = 1, nl = 0;
    for (int b = 0; b < count; ++b) {
        const WorkItem &w = items[b];
        if (!w.gaps) {
            for (int q = 0; q < w.nsrc; ++q)
                ports[b].push_back(Port{w.src[q].rhs, w.src[q].basis - 1, source_value(w.src[q].ampl, w.src[q].phas)});
        } else {
            for (int r = 0; r < w.nrhs; ++r) ports[b].push_back(Port{r, w.gaps[r] - 1, source_value(w.ampl, w.phas)});
        }
        for (const Port &p : ports[b]) merge(ent[b], p);
        for (int q = 0; q < w.nload; ++q) merge(lds[b], Port{0, w.loads[q].basis - 1, load_value(w.loads[q], w.freq)});
        ne = std::max(ne, (int)ent[b].size());
        nl = std::max(nl, (int)lds[b].size());
    }
    // recycled sweep: items with an anchor (a batch index) are solved by GMRES
    // refinement with their anchor's FP64 factors; anchors are factored
    bool recycle = false;
    for (int b = 0; b < count; ++b)
        if (items[b].anchor >= 0) {
            recycle = true;
            if (items[b].anchor >= count || items[items[b].anchor].anchor != items[b].anchor) {
                set_error("solve_items: item %d names anchor %d, which is not an anchor of the batch", b,
                          items[b].anchor);
                return WM_ERR_ARG;
            }
        }
    const bool mixed = o.prec != WM_PREC_F64 && !recycle;
    if (o.prec != WM_PREC_F64 && o.prec != WM_PREC_F32 && o.prec != WM_PREC_BF16X3) {
        set_error("unknown precision %d", o.prec);
        return WM_ERR_ARG;
    }
    const size_t nn = (size_t)n * n;
    const int64_t sZ = (int64_t)nn, sV = (int64_t)n * nr;
    size_t off = 0;
    auto carve = [&](size_t bytes) { const size_t o0 = off; off += (bytes + 255) & ~(size_t)255; return o0; };
    const size_t oZ = carve(nn * sizeof(double2) * count);
    const size_t oZ32 = carve(mixed ? nn * sizeof(float2) * count : 0);
    const size_t szV = (size_t)sV * count * sizeof(double2);
    const size_t oB = carve(szV), oX = carve(szV), oD = carve(mixed || recycle || o.cond_est ? szV : 0),
                 oR = carve(mixed || recycle ? szV : 0);
    const int gm = kGmresRestart;
    const size_t oK = carve(recycle ? gmres_bytes(n, nr, gm) * count : 0);
    const size_t oRed = carve(recycle ? (size_t)(3 * gm + 4) * count * nr * sizeof(double2) : 0);
    const size_t oF = carve(recycle ? count * sizeof(int) : 0);
    const size_t oM2 = carve(recycle ? (size_t)count * (1 + nr) * sizeof(unsigned long long) : 0);
    // uneven sizes: per-item sizes, and the fill's staging of up to kFillChunk
    // smaller matrices before they are expanded to the batch's layout
    constexpr int kFillChunk = 4;
    const size_t oN = carve(uneven ? count * sizeof(int) : 0);
    const size_t tZ = (size_t)ld_small * ld_small;
    const size_t oT = carve(uneven ? tZ * sizeof(double2) * kFillChunk : 0);
    const size_t oT32 = carve(uneven && mixed ? tZ * sizeof(float2) * kFillChunk : 0);
    const size_t oP = carve((size_t)n * count * sizeof(int)), oI = carve(count * sizeof(int));
    const size_t oM = carve((size_t)2 * count * nr * sizeof(unsigned long long));
    const size_t oA = carve(count * sizeof(int));
    const size_t oQ = carve((size_t)n * count * sizeof(int));
    const size_t oG = carve((size_t)count * ne * (2 * sizeof(int) + sizeof(double2)));
    const size_t oL = carve((size_t)count * nl * (sizeof(int) + sizeof(double2)));
    WM_TRY(c->ws_sweep.ensure(off));
    char *base = (char *)c->ws_sweep.ptr;
    double2 *Z = (double2 *)(base + oZ);
    float2 *Z32 = mixed ? (float2 *)(base + oZ32) : nullptr;
    double2 *B = (double2 *)(base + oB), *X = (double2 *)(base + oX);
    double2 *D = (double2 *)(base + oD), *R = (double2 *)(base + oR);
    int *ipiv = (int *)(base + oP), *dinfo = (int *)(base + oI), *dact = (int *)(base + oA);
    int *perm = (int *)(base + oQ);
    unsigned long long *mx = (unsigned long long *)(base + oM);
    const size_t cne = (size_t)count * ne, cnl = (size_t)count * nl;
    int *gcol = (int *)(base + oG), *gidx = gcol + cne;
    double2 *gval = (double2 *)(base + oG + cne * 2 * sizeof(int));
    int *lidx = (int *)(base + oL);
    double2 *lval = (double2 *)(base + oL + cnl * sizeof(int));
    double2 *Kry = (double2 *)(base + oK), *red = (double2 *)(base + oRed);
    int *dfac = (int *)(base + oF);
    unsigned long long *mx2 = (unsigned long long *)(base + oM2);
    int *dnmem = uneven ? (int *)(base + oN) : nullptr;
    double2 *Tst = (double2 *)(base + oT);
    float2 *Tst32 = (float2 *)(base + oT32);

    // excitation: ampl*exp(j*pi*phas/180) at each feed ("wersja referencyjna/Wire_MoM.cu":307-309)
    std::vector<int> hcol(cne, 0), hidx(cne, -1), hlidx(cnl, -1);
    std::vector<double2> hval(cne, make_double2(0.0, 0.0)), hlval(cnl, make_double2(0.0, 0.0));
    for (int b = 0; b < count; ++b) {
        for (size_t e = 0; e < ent[b].size(); ++e) {
            hcol[(size_t)b * ne + e] = ent[b][e].col;
            hidx[(size_t)b * ne + e] = ent[b][e].idx;
            hval[(size_t)b * ne + e] = ent[b][e].v;
        }
        for (size_t e = 0; e < lds[b].size(); ++e) {
            hlidx[(size_t)b * nl + e] = lds[b][e].idx;
            hlval[(size_t)b * nl + e] = lds[b][e].v;
        }
    }
    WM_TRY(copy_h2d(gcol, hcol.data(), cne * sizeof(int), s));
    WM_TRY(copy_h2d(gidx, hidx.data(), cne * sizeof(int), s));
    WM_TRY(copy_h2d(gval, hval.data(), cne * sizeof(double2), s));
    const int64_t nv = sV * count;
    hipLaunchKernelGGL(k_zero2, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, B, nv);
    hipLaunchKernelGGL(k_set_rhs, dim3((unsigned)((cne + 255) / 256)), dim3(256), 0, s, B, sV, n, ne, gcol, gidx, gval,
                       count);
    for (int b = 0; b < count; ++b)
        for (int q = 0; q < items[b].npw; ++q) {
            const WorkItem &w = items[b];
            const PlaneWaveArgs pa = plane_wave_args(w.pw[q], wavenumber_mhz(w.freq));
            hipLaunchKernelGGL(k_plane_wave, dim3((unsigned)((w.g->npls + 255) / 256)), dim3(256), 0, s,
                               B + (size_t)b * sV + (size_t)w.pw[q].rhs * n, w.g->d_obs, w.g->ld, w.g->npls, pa);
        }

    // fill, one launch group per run of items sharing a geometry; a smaller
    // structure's matrices are filled at their own padded size and expanded
    if (uneven) {
        std::vector<int> hn(count);
        for (int b = 0; b < count; ++b) hn[b] = items[b].g->ld;
        WM_TRY(copy_h2d(dnmem, hn.data(), count * sizeof(int), s));
    }
    // Recycled sweeps fill the frequencies that GMRES will solve on a second
    // stream, so their fills run while the anchors are factored (whose panels
    // leave most CUs idle); the first GMRES step waits for them.  The profiler
    // keeps every launch on one stream, so each kernel's time is its own.
    const bool overlap = recycle && !uneven && nl == 0 && !o.cond_est && c->overlap_fill && !c->prof.on;
    auto is_own = [&](int b) { return items[b].anchor < 0 || items[b].anchor == b; };
    if (overlap) {
        if (!c->aux) {  // events first, the stream last: c->aux set means all exist
            for (auto &e : c->aux_ev)
                if (!e) WM_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            WM_HIP_TRY(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
        }
        WM_HIP_TRY(hipEventRecord(c->aux_ev[0], s));  // after everything already queued
        WM_HIP_TRY(hipStreamWaitEvent(c->aux, c->aux_ev[0], 0));
    }
    // the launch stream waits for the second one before the matrices are used,
    // and on every early return, so no later call overlaps fills still running
    struct AuxJoin {
        hipStream_t s, aux;  // aux null: nothing to join
        hipEvent_t e;
        int join()
        {
            if (!aux) return WM_OK;
            const hipError_t r1 = hipEventRecord(e, aux), r2 = hipStreamWaitEvent(s, e, 0);
            aux = nullptr;
            return r1 == hipSuccess && r2 == hipSuccess ? WM_OK : WM_ERR_HIP;
        }
        ~AuxJoin() { (void)join(); }
    } aux_join{s, overlap ? c->aux : nullptr, c->aux_ev[1]};
    if (overlap) {
        // the anchors on the launch stream, the others on the second one, each
        // side in launches of up to four frequencies of one structure (the
        // matrices need not be adjacent); every item's structure has ld = n
        for (int side = 0; side < 2; ++side) {
            std::vector<char> done(count, 0);
            for (int b0 = 0; b0 < count; ++b0) {
                if (done[b0] || is_own(b0) != (side == 0)) continue;
                std::vector<int> slots;
                std::vector<double> ka;
                for (int b = b0; b < count; ++b)
                    if (!done[b] && is_own(b) == (side == 0) && items[b].g == items[b0].g) {
                        done[b] = 1;
                        slots.push_back(b);
                        ka.push_back(wavenumber_mhz(items[b].freq));
                    }
                WM_TRY(fill_launch(c, items[b0].g, (int)slots.size(), ka.data(), Z, sZ, nullptr, 0,
                                   side ? c->aux : s, side ? &c->ws_fill2 : nullptr, slots.data()));
            }
        }
    }
    for (int b0 = 0; b0 < count && !overlap;) {
        int b1 = b0 + 1;
        while (b1 < count && items[b1].g == items[b0].g) ++b1;
        const wm_geom *g = items[b0].g;
        std::vector<double> ka(b1 - b0);
        for (int b = b0; b < b1; ++b) ka[b - b0] = wavenumber_mhz(items[b].freq);
        if (g->ld == n) {
            WM_TRY(fill_launch(c, g, b1 - b0, ka.data(), Z + (size_t)b0 * sZ, sZ,
                               mixed ? (void *)(Z32 + (size_t)b0 * sZ) : nullptr, sZ, s));
        } else {
            const int lg = g->ld;
            const int64_t sT = (int64_t)lg * lg;
            for (int f0 = b0; f0 < b1; f0 += kFillChunk) {
                const int nf = std::min(kFillChunk, b1 - f0);
                WM_TRY(fill_launch(c, g, nf, ka.data() + (f0 - b0), Tst, sT, mixed ? (void *)Tst32 : nullptr, sT, s));
                hipLaunchKernelGGL(k_expand, dim3((unsigned)((nn + 255) / 256), (unsigned)nf), dim3(256), 0, s,
                                   Z + (size_t)f0 * sZ, sZ, mixed ? Z32 + (size_t)f0 * sZ : nullptr, n,
                                   (const double2 *)Tst, (const float2 *)Tst32, sT, lg);
            }
        }
        b0 = b1;
    }
    // lumped loads on the diagonals ("wersja referencyjna/Wire_MoM.cu":344-363)
    if (nl > 0) {
        WM_TRY(copy_h2d(lidx, hlidx.data(), cnl * sizeof(int), s));
        WM_TRY(copy_h2d(lval, hlval.data(), cnl * sizeof(double2), s));
        hipLaunchKernelGGL(k_add_diag, dim3((unsigned)((cnl + 255) / 256)), dim3(256), 0, s, Z, sZ, n, Z32, lidx, lval,
                           nl, count);
    }

    // 1-norms of the matrices before they are factored (condition estimate)
    std::vector<unsigned long long> hnorm(count, 0ull);
    if (o.cond_est) {
        std::vector<int> all(count), np(count);
        for (int b = 0; b < count; ++b) { all[b] = b; np[b] = items[b].g->npls; }
        int *dnp = perm;  // free until the first factorization
        WM_TRY(copy_h2d(dact, all.data(), count * sizeof(int), s));
        WM_TRY(copy_h2d(dnp, np.data(), count * sizeof(int), s));
        hipLaunchKernelGGL(k_zero_u64, dim3((count + 255) / 256), dim3(256), 0, s, mx, count);
        hipLaunchKernelGGL(k_norm1, dim3((unsigned)(n / 4), (unsigned)count), dim3(256), 0, s, Z, sZ, n, dnp, dact, mx);
        WM_TRY(copy_d2h(hnorm.data(), mx, count * sizeof(unsigned long long), s));
    }

    std::vector<int> hinfo(count, 0), iters(count, 0), fell(count, 0);
    std::vector<double> dxr(count, 0.0), berr(count, 0.0);
    int noconv = 0;
    auto upload_list = [&](const std::vector<int> &lst) { return copy_h2d(dact, lst.data(), lst.size() * sizeof(int), s); };
    if (recycle) {
        // anchors (and items without one): FP64 factors, direct solve
        std::vector<int> own, deps, fac;
        for (int b = 0; b < count; ++b) {
            if (items[b].anchor < 0 || items[b].anchor == b) own.push_back(b);
            else { deps.push_back(b); fac.push_back(items[b].anchor); }
        }
        auto factor_solve = [&](const std::vector<int> &lst) -> int {
            const int k = (int)lst.size();
            WM_TRY(upload_list(lst));
            WM_TRY(getrf_batched(c, WM_PREC_F64, n, Z, n, sZ, ipiv, n, k, dinfo, s, dact, dnmem));
            WM_TRY(perm_batched(c, n, ipiv, n, perm, n, k, s, dact));
            hipLaunchKernelGGL(k_copy_items, dim3((unsigned)((sV + 255) / 256), k), dim3(256), 0, s, X, B, sV, sV, dact);
            WM_TRY(getrs_perm_batched(c, WM_PREC_F64, n, nr, Z, n, sZ, perm, n, X, n, sV, k, s, dact));
            WM_TRY(check_device_error(c, s));
            std::vector<int> h(count, 0);  // info words of the listed items only
            WM_TRY(copy_d2h(h.data(), dinfo, sizeof(int) * count, s));
            for (int b : lst) hinfo[b] = h[b];
            return WM_OK;
        };
        if (!own.empty()) WM_TRY(factor_solve(own));
        for (int b : own) items[b].anchor_used = b;
        WM_TRY(aux_join.join());  // the other items' matrices are filled
        // the others: FP64 GMRES refinement preconditioned by the anchor's factors
        // (an anchor with a zero pivot preconditions nothing: its dependents are factored)
        std::vector<int> gdeps, gfac, redo;
        for (size_t q = 0; q < deps.size(); ++q) {
            if (hinfo[fac[q]] != 0) redo.push_back(deps[q]);
            else { gdeps.push_back(deps[q]); gfac.push_back(fac[q]); }
        }
        if (!gdeps.empty()) {
            GmresArgs ga;
            ga.n = n; ga.nr = nr; ga.m = gm; ga.count = count;
            ga.Z = Z; ga.sZ = sZ; ga.F = Z; ga.fprec = WM_PREC_F64; ga.sF = sZ; ga.perm = perm;
            ga.B = B; ga.X = X; ga.R = R; ga.D = D; ga.sV = sV;
            ga.K = Kry; ga.red = red; ga.dact = dact; ga.dfac = dfac; ga.mx = mx; ga.mx2 = mx2;
            ga.np = 0;
            for (int b : gdeps) ga.np = std::max(ga.np, items[b].g->npls);
            GmresOut go(count);
            WM_TRY(gmres_ir(c, o, ga, gdeps, gfac, go, s));
            for (size_t q = 0; q < gdeps.size(); ++q) {
                const int b = gdeps[q];
                iters[b] = go.iters[b];
                dxr[b] = go.dx_rel[b];
                berr[b] = go.berr[b];
                if (go.conv[b]) items[b].anchor_used = gfac[q];
                else redo.push_back(b);
            }
        }
        // items whose refinement did not converge: factored themselves
        if (!redo.empty()) {
            std::sort(redo.begin(), redo.end());
            WM_TRY(factor_solve(redo));
            for (int b : redo) {
                fell[b] = 1;
                items[b].anchor_used = b;
            }
        }
        for (int b = 0; b < count; ++b) items[b].fprec = WM_PREC_F64;
    } else if (!mixed) {
        WM_TRY(getrf_batched(c, WM_PREC_F64, n, Z, n, sZ, ipiv, n, count, dinfo, s, nullptr, dnmem));
        WM_TRY(perm_batched(c, n, ipiv, n, perm, n, count, s));
        WM_HIP_TRY(hipMemcpyAsync(X, B, szV, hipMemcpyDeviceToDevice, s));
        WM_TRY(getrs_perm_batched(c, WM_PREC_F64, n, nr, Z, n, sZ, perm, n, X, n, sV, count, s));
        WM_TRY(check_device_error(c, s));
        WM_TRY(copy_d2h(hinfo.data(), dinfo, sizeof(int) * count, s));
    } else {
        // Complex64 factors and FP64 refinement, level by level: split-BF16x3
        // updates, then FP32 updates (items routed there, and BF16x3 attempts
        // that did not converge, refactored from the complex128 matrix), then
        // the FP64 fallback.
        std::vector<int> lvl[2], conv(count, 0);
        for (int b = 0; b < count; ++b) {
            items[b].rate = items[b].rate_b3 = -1.0;
            if (items[b].fprec < 0) items[b].fprec = o.prec;
            if (items[b].fprec == WM_PREC_BF16X3) lvl[0].push_back(b);
            else if (items[b].fprec == WM_PREC_F32) lvl[1].push_back(b);
        }
        std::vector<double> prev((size_t)count * nr), first(count);
        std::vector<unsigned long long> hm((size_t)2 * count * nr);
        std::vector<int> retry;  // level-0 items refactored at level 1
        std::vector<char> tried(count, 0);  // entered a complex64 refinement level
        for (int q = 0; q < 2; ++q) {
            std::vector<int> &L = lvl[q];
            if (L.empty()) continue;
            std::sort(L.begin(), L.end());
            const int pq = q == 0 ? WM_PREC_BF16X3 : WM_PREC_F32, np = (int)L.size();
            if (!retry.empty()) {
                // the failed attempts' factors overwrote their complex64 copies
                WM_TRY(upload_list(retry));
                hipLaunchKernelGGL(k_to_c64, dim3((unsigned)((nn + 255) / 256), (unsigned)retry.size()), dim3(256), 0, s,
                                   Z32, Z, (int64_t)nn, sZ, dact);
            }
            WM_TRY(upload_list(L));
            if (np == count) {
                WM_TRY(getrf_batched(c, pq, n, Z32, n, sZ, ipiv, n, count, dinfo, s, nullptr, dnmem));
                WM_TRY(perm_batched(c, n, ipiv, n, perm, n, count, s));
            } else {
                WM_TRY(getrf_batched(c, pq, n, Z32, n, sZ, ipiv, n, np, dinfo, s, dact, dnmem));
                WM_TRY(perm_batched(c, n, ipiv, n, perm, n, np, s, dact));
            }
            // X = 0, R = B for the items of this level
            hipLaunchKernelGGL(k_zero_items, dim3((unsigned)((sV + 255) / 256), np), dim3(256), 0, s, X, sV, sV, dact);
            hipLaunchKernelGGL(k_copy_items, dim3((unsigned)((sV + 255) / 256), np), dim3(256), 0, s, R, B, sV, sV, dact);
            WM_TRY(check_device_error(c, s));
            WM_TRY(copy_d2h(hinfo.data(), dinfo, sizeof(int) * count, s));
            std::vector<int> active, failed;
            for (int b : L) {
                for (int r = 0; r < nr; ++r) prev[(size_t)b * nr + r] = 1e300;
                items[b].fprec = pq;
                tried[b] = 1;
                if (hinfo[b] == 0) active.push_back(b);
                else {
                    items[b].rate = HUGE_VAL;  // singular complex64 factor
                    failed.push_back(b);
                }
            }
            for (int it = 0; it < o.max_iter && !active.empty(); ++it) {
                const int na = (int)active.size();
                WM_TRY(upload_list(active));
                // D = R for the active items, then D = LU32^{-1} D (FP64 arithmetic on widened factors)
                hipLaunchKernelGGL(k_copy_items, dim3((unsigned)((sV + 255) / 256), na), dim3(256), 0, s, D, R, sV, sV,
                                   dact);
                WM_TRY(getrs_perm_batched(c, WM_PREC_F32, n, nr, Z32, n, sZ, perm, n, D, n, sV, na, s, dact));
                hipLaunchKernelGGL(k_zero_u64, dim3((2 * count * nr + 255) / 256), dim3(256), 0, s, mx, 2 * count * nr);
                hipLaunchKernelGGL(k_update, dim3((n + 255) / 256, na * nr), dim3(256), 0, s, X, D, n, n, sV, nr, dact, mx);
                WM_TRY(residual_batched(c, n, nr, Z, n, sZ, X, B, n, sV, R, na, s, dact));
                WM_TRY(check_device_error(c, s));
                WM_TRY(copy_d2h(hm.data(), mx, hm.size() * sizeof(unsigned long long), s));
                std::vector<int> still;
                for (int b : active) {
                    double dm[64], xm[64];
                    for (int r = 0; r < items[b].nrhs; ++r) {
                        std::memcpy(&dm[r], &hm[2 * ((size_t)b * nr + r)], 8);
                        std::memcpy(&xm[r], &hm[2 * ((size_t)b * nr + r) + 1], 8);
                    }
                    iters[b] += 1;
                    const int v = refine_decide(items[b].nrhs, dm, xm, &prev[(size_t)b * nr], it, o.tol, o.max_iter,
                                                first[b], items[b].rate, dxr[b]);
                    if (v > 0) conv[b] = 1;
                    else if (v < 0) failed.push_back(b);
                    else still.push_back(b);
                }
                active.swap(still);
            }
            for (int b : active) failed.push_back(b);  // ran out of iterations
            retry.clear();
            if (q == 0) {
                for (int b : failed) {
                    items[b].rate_b3 = items[b].rate;
                    items[b].rate = -1.0;
                    lvl[1].push_back(b);
                    retry.push_back(b);
                }
            }
        }
        // Items routed to FP64 factors, and (with fallback_f64) items whose
        // refinement did not converge or whose complex64 factors were singular:
        // one batched FP64 factorization.  Without fallback_f64 the items that
        // failed refinement keep their last iterate and the call reports
        // WM_ERR_NOCONV after every other item is finished.
        std::vector<int> bad;
        int nfail = 0;
        for (int b = 0; b < count; ++b) {
            if (conv[b]) continue;
            if (tried[b] && !o.fallback_f64) {
                ++nfail;
                fell[b] = -1;
                continue;
            }
            bad.push_back(b);
        }
        if (nfail > 0) noconv = nfail;
        if (!bad.empty()) {
            const int nb = (int)bad.size();
            WM_TRY(upload_list(bad));
            WM_TRY(getrf_batched(c, WM_PREC_F64, n, Z, n, sZ, ipiv, n, nb, dinfo, s, dact, dnmem));
            WM_TRY(perm_batched(c, n, ipiv, n, perm, n, nb, s, dact));
            hipLaunchKernelGGL(k_copy_items, dim3((unsigned)((sV + 255) / 256), nb), dim3(256), 0, s, X, B, sV, sV, dact);
            WM_TRY(getrs_perm_batched(c, WM_PREC_F64, n, nr, Z, n, sZ, perm, n, X, n, sV, nb, s, dact));
            WM_TRY(check_device_error(c, s));
            WM_TRY(copy_d2h(hinfo.data(), dinfo, sizeof(int) * count, s));
            for (int b : bad) {
                fell[b] = tried[b] ? 1 : 2;
                items[b].fprec = WM_PREC_F64;
            }
        }
    }

    // condition estimates with the final factors of every item
    std::vector<double> cond(count, 0.0), anorm(count, 0.0);
    if (o.cond_est) {
        for (int b = 0; b < count; ++b) {
            unsigned long long bits = hnorm[b];
            std::memcpy(&anorm[b], &bits, 8);
        }
        for (int f64 = 0; f64 < 2; ++f64) {
            std::vector<int> lst;
            for (int b = 0; b < count; ++b) {
                if (recycle && items[b].anchor_used != b) continue;  // solved with another item's factors
                if ((mixed ? items[b].fprec == WM_PREC_F64 : true) == (f64 == 1)) lst.push_back(b);
            }
            if (lst.empty()) continue;
            std::vector<int> np(lst.size());
            for (size_t i = 0; i < lst.size(); ++i) np[i] = items[lst[i]].g->npls;
            std::vector<double> est;
            WM_TRY(inv_norm1_estimate(c, f64 ? WM_PREC_F64 : WM_PREC_F32, n, f64 ? (const void *)Z : (const void *)Z32,
                                      sZ, perm, D, sV, lst, np, dact, s, est));
            for (size_t i = 0; i < lst.size(); ++i) cond[lst[i]] = anorm[lst[i]] * est[i];
        }
    }

    std::vector<double2> hX((size_t)sV * count);
    WM_TRY(check_device_error(c, s));
    WM_TRY(copy_d2h(hX.data(), X, szV, s));
    int status = WM_OK;
    for (int b = 0; b < count; ++b) {
        const WorkItem &w = items[b];
        const int N = w.g->npls;
        if (hinfo[b] != 0 && status == WM_OK) {
            set_error("zero pivot at column %d (%g MHz)", hinfo[b], (double)w.freq);
            status = hinfo[b];
        }
        if (w.I)
            for (int r = 0; r < w.nrhs; ++r) {
                const double2 *xv = &hX[(size_t)b * sV + (size_t)r * n];
                for (int i = 0; i < N; ++i) w.I[(size_t)r * N + i] = wm_c128{xv[i].x, xv[i].y};
            }
        if (w.zin)
            for (size_t q = 0; q < ports[b].size(); ++q) {
                const Port &p = ports[b][q];
                const double2 x = hX[(size_t)b * sV + (size_t)p.col * n + p.idx];
                w.zin[q] = cdiv(wm_c128{p.v.x, p.v.y}, wm_c128{x.x, x.y});
            }
        if (w.st) {
            w.st->iters = iters[b];
            w.st->info = hinfo[b];
            w.st->fell_back = fell[b];
            w.st->prec = mixed ? w.fprec : WM_PREC_F64;
            w.st->rate = mixed ? w.rate : -1.0;
            w.st->dx_rel = dxr[b];
            w.st->cond1 = cond[b];
            w.st->anorm1 = anorm[b];
            w.st->anchor_mhz = recycle && w.anchor_used >= 0 ? (double)items[w.anchor_used].freq : 0.0;
            w.st->berr = berr[b];
        }
    }
    if (noconv > 0 && status == WM_OK) {
        set_error("refinement did not converge for %d of %d items (fallback to FP64 factors disabled)", noconv, count);
        return WM_ERR_NOCONV;
    }
    return status;
}

}  // namespace wm

// Internal test hook (not part of include/wiremom.h): the host convergence test
// of one refinement iteration, wm::refine_decide, callable without a device.
extern "C" int wmx_refine_decide(int nrhs, const double *dm, const double *xm, double *prev, int it, double tol,
                                 int max_iter, double *first, double *rate, double *worst)
{
    if (nrhs <= 0 || !dm || !xm || !prev || !first || !rate || !worst) return -100;
    return wm::refine_decide(nrhs, dm, xm, prev, it, tol, max_iter, *first, *rate, *worst);
}

// Internal hook: right-hand sides per launch member of the GMRES
// preconditioner's grouped solves (1, 2, 4 or 8).
extern "C" int wmx_set_group_cols(wm_ctx *ctx, int w)
{
    if (!ctx || (w != 1 && w != 2 && w != 4 && w != 8)) return WM_ERR_ARG;
    ctx->group_cols = w;
    return WM_OK;
}
