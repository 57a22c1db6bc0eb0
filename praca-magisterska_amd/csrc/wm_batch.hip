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
#include <cmath>
#include <cstring>
#include <vector>

#include "wm_internal.h"

namespace wm {

namespace {

constexpr double kPi = 3.14159265358979323846;

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
    return t64 <= tmix;
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
    const int n = items[0].g->ld;
    int nr = 0;
    for (int b = 0; b < count; ++b) {
        const WorkItem &w = items[b];
        const int np = w.g->npls;
        if (w.g->ld != n) {
            set_error("solve_items: items of different sizes in one batch");
            return WM_ERR_ARG;
        }
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
    const bool mixed = o.prec != WM_PREC_F64;
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
    const size_t oB = carve(szV), oX = carve(szV), oD = carve(mixed ? szV : 0), oR = carve(mixed ? szV : 0);
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
                               B + (size_t)b * sV + (size_t)w.pw[q].rhs * n, w.g->d_obs, n, w.g->npls, pa);
        }

    // fill, one launch group per run of items sharing a geometry
    for (int b0 = 0; b0 < count;) {
        int b1 = b0 + 1;
        while (b1 < count && items[b1].g == items[b0].g) ++b1;
        std::vector<double> ka(b1 - b0);
        for (int b = b0; b < b1; ++b) ka[b - b0] = wavenumber_mhz(items[b].freq);
        WM_TRY(fill_launch(c, items[b0].g, b1 - b0, ka.data(), Z + (size_t)b0 * sZ, sZ,
                           mixed ? (void *)(Z32 + (size_t)b0 * sZ) : nullptr, sZ, s));
        b0 = b1;
    }
    // lumped loads on the diagonals ("wersja referencyjna/Wire_MoM.cu":344-363)
    if (nl > 0) {
        WM_TRY(copy_h2d(lidx, hlidx.data(), cnl * sizeof(int), s));
        WM_TRY(copy_h2d(lval, hlval.data(), cnl * sizeof(double2), s));
        hipLaunchKernelGGL(k_add_diag, dim3((unsigned)((cnl + 255) / 256)), dim3(256), 0, s, Z, sZ, n, Z32, lidx, lval,
                           nl, count);
    }

    std::vector<int> hinfo(count, 0), iters(count, 0), fell(count, 0);
    std::vector<double> dxr(count, 0.0);
    auto upload_list = [&](const std::vector<int> &lst) { return copy_h2d(dact, lst.data(), lst.size() * sizeof(int), s); };
    if (!mixed) {
        WM_TRY(getrf_batched(c, WM_PREC_F64, n, Z, n, sZ, ipiv, n, count, dinfo, s));
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
                WM_TRY(getrf_batched(c, pq, n, Z32, n, sZ, ipiv, n, count, dinfo, s));
                WM_TRY(perm_batched(c, n, ipiv, n, perm, n, count, s));
            } else {
                WM_TRY(getrf_batched(c, pq, n, Z32, n, sZ, ipiv, n, np, dinfo, s, dact));
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
                    double worst = 0.0, rate = 0.0;
                    bool stalled = true;
                    for (int r = 0; r < items[b].nrhs; ++r) {
                        double dm, xm;
                        std::memcpy(&dm, &hm[2 * ((size_t)b * nr + r)], 8);
                        std::memcpy(&xm, &hm[2 * ((size_t)b * nr + r) + 1], 8);
                        const double rel = xm > 0 ? dm / xm : 0.0;
                        worst = std::max(worst, rel);
                        double &p = prev[(size_t)b * nr + r];
                        rate = std::max(rate, p < 1e299 && p > 0 ? rel / p : 0.0);
                        if (!(rel > 0.5 * p && p < 1e-9)) stalled = false;
                        p = rel;
                    }
                    dxr[b] = worst;
                    iters[b] += 1;
                    // recorded rate: geometric mean contraction since the first correction
                    if (it == 0) first[b] = worst;
                    if (it >= 1 && !stalled && first[b] > 0)
                        items[b].rate = std::pow(worst / first[b], 1.0 / it);
                    else if (it == 0 && worst <= o.tol) items[b].rate = worst;  // converged at once: rate at most this
                    if (worst <= o.tol || stalled) {
                        conv[b] = 1;
                        continue;
                    }
                    // contraction too slow to reach tol within max_iter: stop now
                    if (it >= 1) {
                        const double gm = items[b].rate;
                        const double need = gm > 0 && gm < 1 ? std::log(o.tol / worst) / std::log(gm) : 1e9;
                        if (it + 1 + need > o.max_iter) {
                            failed.push_back(b);
                            continue;
                        }
                    }
                    still.push_back(b);
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
        // items that did not converge (or had singular complex64 factors): one batched FP64 refactorization
        std::vector<int> bad;
        for (int b = 0; b < count; ++b)
            if (!conv[b]) bad.push_back(b);
        if (!bad.empty()) {
            if (!o.fallback_f64) {
                set_error("refinement did not converge for %zu of %d items", bad.size(), count);
                return WM_ERR_NOCONV;
            }
            const int nb = (int)bad.size();
            WM_TRY(upload_list(bad));
            WM_TRY(getrf_batched(c, WM_PREC_F64, n, Z, n, sZ, ipiv, n, nb, dinfo, s, dact));
            WM_TRY(perm_batched(c, n, ipiv, n, perm, n, nb, s, dact));
            hipLaunchKernelGGL(k_copy_items, dim3((unsigned)((sV + 255) / 256), nb), dim3(256), 0, s, X, B, sV, sV, dact);
            WM_TRY(getrs_perm_batched(c, WM_PREC_F64, n, nr, Z, n, sZ, perm, n, X, n, sV, nb, s, dact));
            WM_TRY(check_device_error(c, s));
            WM_TRY(copy_d2h(hinfo.data(), dinfo, sizeof(int) * count, s));
            for (int b : bad) {
                fell[b] = items[b].fprec == WM_PREC_F64 ? 2 : 1;
                items[b].fprec = WM_PREC_F64;
            }
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
        }
    }
    return status;
}

}  // namespace wm
