// Impedance-matrix fill for gfx950.
//
// Replaces the reference's CPU fill zmatrix()/wim() ("wersja referencyjna/
// Zmatrix.cpp":27-233) and its six-pass CUDA fill (zmatrix_gpu,
// "wersja referencyjna/Wire_MoM.cu":4052-4197; kernels at 520, 851, 1856, 2189,
// 3200, 3407).  One launch writes every entry of up to kMaxF frequency
// matrices exactly once:
//
//   Z[m,n] = j*eta0/(4 pi k) * sum_{h=obs half of m} ( k^2 * sum_g (l_g . l_h) V_hg  -+  sum_g +-S_hg / dl_g )
//
// where for source half g of basis n, V_hg / S_hg are the vector / scalar
// potential integrals (weights s/dl, (dl-s)/dl and 1) of exp(-jkR)/R taken
// with the reference's 2-point Gauss-Legendre rule, except on the self
// segment, where the singularity-extracted FP64 integrals (4-point rules on
// both halves plus the analytic log term, Zmatrix.cpp:279-323, 695-770) are
// precomputed per (basis, half, frequency) by k_self_terms.
//
// Mapping: one lane per test basis m (64 consecutive rows per wave, so each
// column store is one coalesced 1 KiB run), each wave walks kColsPerWave
// source bases n; n is wave-uniform, so the source records arrive through
// scalar loads.  R and 1/R are shared by every frequency of the launch: the
// kernel e^{-jkR}/R is taken as (cos kR, -sin kR) * (1/R), 1/R from one
// reciprocal square root of R^2 per quadrature point and R = R^2 (1/R), and
// the divisions by a segment length as products with its reciprocal, where
// the reference takes a square root and divides ("wersja referencyjna/
// Zmatrix.cpp":679-688); the results differ from its by about an ulp.
// cos kR and sin kR come from sincos_phase (wm_sincos.h: Cody-Waite reduction
// and Taylor polynomials, within about 1.5 ulp of the exact values for kR up
// to 100 and 2.5 ulp up to its limit), which needs a fraction of the
// operations and registers of a general-purpose sincos.
//
// PEC ground (ienv 2, "wersja referencyjna/Zmatrix.cpp":94-145, 174-224): every
// source half is followed by its image in z = 0 (coordinates scaled by the
// environment signs, weight sgnenv), which doubles the kernel evaluations.
#include <algorithm>

#include "wm_internal.h"
#include "wm_sincos.h"

namespace wm {

namespace {

constexpr int kMaxF = 4;          // frequencies per launch (with or without the image terms)
constexpr int kColsPerWave = 4;   // source bases per wave
constexpr int kWaves = 4;         // waves per block
constexpr double kPi = 3.14159265358979323846;
constexpr double kGL2 = 0.288675134594812882;  // 2-point node offset / length
constexpr double kGL4c1 = 0.43056815579702629, kGL4w1 = 0.17392742256872693;
constexpr double kGL4c2 = 0.16999052179242813, kGL4w2 = 0.32607257743127307;

struct FreqArgs {
    double ka[kMaxF];
    double ka2[kMaxF];
    double q[kMaxF];  // eta0 / (4 pi k)
    int64_t zo[kMaxF];  // offset of each frequency's matrix from Z (elements)
};

// Complete elliptic K from the Abramowitz & Stegun 17.3.34 polynomial in
// m1 = 1 - bet^2 with the reference's truncation thresholds (Zmatrix.cpp:731-770).
__device__ double elliptic_k_as(double bet)
{
    const double m1 = 1.0 - bet * bet;
    double a = 1.38629436112 + 0.09666344259 * m1;
    double b = 0.5 + 0.12498593597 * m1;
    double m12 = 0.0, m13 = 0.0;
    if (m1 >= 1.0e-18) { m12 = m1 * m1; a += 0.03590092383 * m12; b += 0.06880248576 * m12; }
    if (m1 >= 1.0e-12) { m13 = m12 * m1; a += 0.03742563713 * m13; b += 0.03328355346 * m13; }
    if (m1 >= 1.0e-9) { const double m14 = m13 * m1; a += 0.01451196212 * m14; b += 0.00441787012 * m14; }
    return a - b * log(m1);
}

// Singularity-extracted self kernel at offset s along a segment of length dl
// and radius an, times the ifun weight (Zmatrix.cpp:695-728).
__device__ double2 self_kernel(double s, int ifun, double ka, double dl, double an)
{
    const double r = 0.5 * dl - s;
    const double rr = sqrt(an * an + r * r);
    double sn, cn;
    sincos(ka * rr, &sn, &cn);
    const double c1r = (cn - 1.0) / rr, c1i = (-sn) / rr;
    const double bet = (an + an) / sqrt(4.0 * an * an + r * r);
    const double bnd2 = (bet * elliptic_k_as(bet) + log(fabs(r / (8.0 * an)))) / (kPi * an);
    double vr = c1r + bnd2, vi = c1i;
    if (ifun == 1) { const double w = s / dl; vr *= w; vi *= w; }
    else if (ifun == 2) { const double w = (dl - s) / dl; vr *= w; vi *= w; }
    return make_double2(vr, vi);
}

__device__ double2 gl4_self(double xl, double xu, int ifun, double ka, double dl, double an)
{
    const double a = 0.5 * (xl + xu), b = xu - xl;
    double c = kGL4c1 * b;
    double2 p = self_kernel(a + c, ifun, ka, dl, an), q = self_kernel(a - c, ifun, ka, dl, an);
    const double sr = (p.x + q.x) * kGL4w1, si = (p.y + q.y) * kGL4w1;
    c = kGL4c2 * b;
    p = self_kernel(a + c, ifun, ka, dl, an);
    q = self_kernel(a - c, ifun, ka, dl, an);
    const double tr = (p.x + q.x) * kGL4w2, ti = (p.y + q.y) * kGL4w2;
    return make_double2((sr + tr) * b, (si + ti) * b);
}

// Integral over the self segment (Zmatrix.cpp:279-323).
__device__ double2 self_integral(int ifun, double ka, double dl, double an)
{
    const double dlh = 0.5 * dl;
    const double2 a = gl4_self(0.0, dlh, ifun, ka, dl, an);
    const double2 b = gl4_self(dlh, dl, ifun, ka, dl, an);
    double res = dl * (1.0 + log(16.0 * an / dl)) / (kPi * an);
    if (ifun != 3) res = 0.5 * dl * (1.0 + log(16.0 * an / dl)) / (kPi * an);
    return make_double2((a.x + b.x) + res, a.y + b.y);
}

// self[(f*ld + n)*4 + 2*g + {0: vector, 1: scalar}] for source half g of n.
__global__ void k_self_terms(const double *__restrict__ src, int npls, int ld, int nf, FreqArgs fa,
                             double2 *__restrict__ self)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = t >> 1, g = t & 1;
    if (n >= ld) return;
    for (int f = 0; f < nf; ++f) {
        double2 v = make_double2(0.0, 0.0), s = make_double2(0.0, 0.0);
        if (n < npls) {
            const double dl = src[(size_t)n * 16 + g * 8 + 6];
            const double an = src[(size_t)n * 16 + g * 8 + 7];
            v = self_integral(g + 1, fa.ka[f], dl, an);
            s = self_integral(3, fa.ka[f], dl, an);
        }
        self[((size_t)f * ld + n) * 4 + 2 * g] = v;
        self[((size_t)f * ld + n) * 4 + 2 * g + 1] = s;
    }
}

// Image source halves (PEC ground, IMG): after each source half g its image
// is taken with the same rule -- the self rule on the image of the
// observation segment itself when that segment lies in the ground plane --
// and added with weight sgnenv, in the reference's order ("wersja
// referencyjna/Zmatrix.cpp":86-151, 167-231).
struct ImgArgs {
    const double *rec;  // image records, as src
    const int *key;     // image self keys
    double sgnenv;
};

template <int F, bool kW32, bool IMG>
__global__ __launch_bounds__(256) void k_fill(const double *__restrict__ obs, const int *__restrict__ obs_key,
                                              const double *__restrict__ src, const int *__restrict__ src_key,
                                              ImgArgs im, const double2 *__restrict__ self, int npls, int ld,
                                              FreqArgs fa, double2 *__restrict__ Z, int64_t strideZ,
                                              float2 *__restrict__ Z32, int64_t strideZ32)
{
    // Every rounding is spelled out (fma where fused), so an entry does not
    // depend on how many frequencies share the launch (F): the compiler may
    // not contract products differently per instantiation.
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = blockIdx.x * 64 + lane;

    double rc[2][3], lm[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            rc[h][d] = obs[(size_t)(6 * h + d) * ld + m];
            lm[h][d] = obs[(size_t)(6 * h + 3 + d) * ld + m];
        }
    const int okey[2] = {obs_key[m], obs_key[ld + m]};

    for (int c = 0; c < kColsPerWave; ++c) {
        const int n = (blockIdx.y * kWaves + wave) * kColsPerWave + c;  // wave-uniform
        double2 acc[F];
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = make_double2(0.0, 0.0);

#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double2 cv[F], cs[F];
#pragma unroll
            for (int f = 0; f < F; ++f) { cv[f] = make_double2(0.0, 0.0); cs[f] = make_double2(0.0, 0.0); }
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                // the image pass is a loop, not unrolled: unrolled, four
                // frequencies need more than 256 VGPRs and the compiler parks
                // values in AGPRs; as a loop the image kernel needs 189 VGPRs
                // (build.py rejects a fill kernel that uses AGPRs or scratch)
#pragma unroll 1
                for (int pass = 0; pass < (IMG ? 2 : 1); ++pass) {
                    const double *sr = (pass ? im.rec : src) + (size_t)n * 16;
                    const int skey = pass ? im.key[2 * n + g] : src_key[2 * n + g];
                    const double rsx = sr[8 * g + 0], rsy = sr[8 * g + 1], rsz = sr[8 * g + 2];
                    const double dcx = sr[8 * g + 3], dcy = sr[8 * g + 4], dcz = sr[8 * g + 5];
                    const double dl = sr[8 * g + 6], an = sr[8 * g + 7];
                    const bool is_self = (okey[h] == skey);
                    const double half = 0.5 * dl, off = kGL2 * dl, rdl = 1.0 / dl;
                    const double bx = rc[h][0] - rsx, by = rc[h][1] - rsy, bz = rc[h][2] - rsz;
                    double2 V[F], S[F];
#pragma unroll
                    for (int f = 0; f < F; ++f) { V[f] = make_double2(0.0, 0.0); S[f] = make_double2(0.0, 0.0); }
#pragma unroll
                    for (int qp = 0; qp < 2; ++qp) {
                        const double s = qp == 0 ? half + off : half - off;
                        const double dx = fma(-s, dcx, bx), dy = fma(-s, dcy, by), dz = fma(-s, dcz, bz);
                        const double R2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, an * an)));
                        const double rR = rsqrt(R2), R = R2 * rR;
                        const double w = (g == 0 ? s : dl - s) * rdl;
#pragma unroll
                        for (int f = 0; f < F; ++f) {
                            double sn, cn;
                            sincos_phase(fa.ka[f] * R, sn, cn);
                            const double gr = cn * rR, gi = -sn * rR;
                            V[f].x = fma(w, gr, V[f].x); V[f].y = fma(w, gi, V[f].y);
                            S[f].x += gr;                S[f].y += gi;
                        }
                    }
                    double t = fma(dcz, lm[h][2], fma(dcy, lm[h][1], dcx * lm[h][0]));
                    if (pass) t *= im.sgnenv;
#pragma unroll
                    for (int f = 0; f < F; ++f) {
                        double2 v = make_double2(V[f].x * half, V[f].y * half);
                        double2 sc = make_double2(S[f].x * half, S[f].y * half);
                        if (is_self) {
                            v = self[((size_t)f * ld + n) * 4 + 2 * g];
                            sc = self[((size_t)f * ld + n) * 4 + 2 * g + 1];
                        }
                        cv[f].x = fma(v.x, t, cv[f].x); cv[f].y = fma(v.y, t, cv[f].y);
                        const double wq = pass ? im.sgnenv * rdl : rdl;
                        const double ws = g == 0 ? wq : -wq;
                        cs[f].x = fma(sc.x, ws, cs[f].x); cs[f].y = fma(sc.y, ws, cs[f].y);
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const double sx = h == 0 ? -cs[f].x : cs[f].x, sy = h == 0 ? -cs[f].y : cs[f].y;
                acc[f].x += fma(cv[f].x, fa.ka2[f], sx);
                acc[f].y += fma(cv[f].y, fa.ka2[f], sy);
            }
        }
        const bool inside = (m < npls) && (n < npls);
        const size_t idx = (size_t)m + (size_t)n * ld;
#pragma unroll
        for (int f = 0; f < F; ++f) {
            double2 z = make_double2(-fa.q[f] * acc[f].y, fa.q[f] * acc[f].x);
            if (!inside) z = make_double2(m == n ? 1.0 : 0.0, 0.0);
            Z[fa.zo[f] + idx] = z;
            if (kW32) Z32[f * strideZ32 + idx] = make_float2((float)z.x, (float)z.y);
        }
    }
}

template <int F, bool IMG>
void launch_fill_fi(dim3 grid, hipStream_t s, const wm_geom *g, const double2 *self, const FreqArgs &fa,
                    double2 *Z, int64_t sZ, float2 *Z32, int64_t sZ32)
{
    const ImgArgs im{g->d_img, g->d_img_key, g->sgnenv};
    if (Z32)
        hipLaunchKernelGGL((k_fill<F, true, IMG>), grid, dim3(256), 0, s, g->d_obs, g->d_obs_key, g->d_src,
                           g->d_src_key, im, self, g->npls, g->ld, fa, Z, sZ, Z32, sZ32);
    else
        hipLaunchKernelGGL((k_fill<F, false, IMG>), grid, dim3(256), 0, s, g->d_obs, g->d_obs_key, g->d_src,
                           g->d_src_key, im, self, g->npls, g->ld, fa, Z, sZ, (float2 *)nullptr, (int64_t)0);
}

template <int F>
void launch_fill_f(dim3 grid, hipStream_t s, const wm_geom *g, const double2 *self, const FreqArgs &fa,
                   double2 *Z, int64_t sZ, float2 *Z32, int64_t sZ32)
{
    if (g->ienv == 2) launch_fill_fi<F, true>(grid, s, g, self, fa, Z, sZ, Z32, sZ32);
    else launch_fill_fi<F, false>(grid, s, g, self, fa, Z, sZ, Z32, sZ32);
}

}  // namespace

int fill_launch(wm_ctx *ctx, const wm_geom *g, int nf, const double *ka, void *Z, int64_t strideZ,
                void *Z32, int64_t strideZ32, hipStream_t s, DevBuf *self_ws, const int *slots)
{
    if (!g || nf <= 0 || !ka || !Z || (slots && Z32)) {
        set_error("fill: bad arguments");
        return WM_ERR_ARG;
    }
    const int ld = g->ld;
    if (ld % (kWaves * kColsPerWave) != 0 || ld % 64 != 0) {
        set_error("fill: ld %d not a multiple of 64", ld);
        return WM_ERR_ARG;
    }
    // self-term workspace: kMaxF * ld * 4 complex128
    const size_t self_bytes = (size_t)kMaxF * ld * 4 * sizeof(double2);
    DevBuf &ws = self_ws ? *self_ws : ctx->ws_fill;
    WM_TRY(ws.ensure(self_bytes));
    double2 *self = (double2 *)ws.ptr;
    const dim3 grid(ld / 64, ld / (kWaves * kColsPerWave));
    // the phases k R stay inside the range of the fill's argument reduction
    // (R is at most the structure's diameter, images included, plus a radius)
    for (int f = 0; f < nf; ++f)
        if (!(ka[f] > 0.0) || ka[f] * (2.0 * g->diam + 1.0) >= kSinCosMax) {
            set_error("fill: wavenumber %g on a structure of diameter %g m is outside the fill's phase range", ka[f],
                      g->diam);
            return WM_ERR_ARG;
        }
    // frequencies per launch: the tuned setting for the environment and size, else kMaxF
    const int tuned = ctx->fill_freqs[g->ienv == 2 ? 1 : 0][std::min(ld / 1024, wm_ctx::kTuneSizes - 1)];
    const int fmax = tuned >= 1 && tuned <= kMaxF ? tuned : kMaxF;
    for (int f0 = 0; f0 < nf; f0 += fmax) {
        const int F = (nf - f0 < fmax) ? nf - f0 : fmax;
        FreqArgs fa;
        for (int f = 0; f < kMaxF; ++f) {
            const double k = ka[f0 + (f < F ? f : 0)];
            fa.ka[f] = k;
            fa.ka2[f] = k * k;
            fa.q[f] = (120.0 * kPi) / (4.0 * kPi * k);
            fa.zo[f] = (int64_t)(slots ? slots[f0 + (f < F ? f : 0)] : f0 + f) * strideZ;
        }
        hipLaunchKernelGGL(k_self_terms, dim3((2 * ld + 255) / 256), dim3(256), 0, s, g->d_src, g->npls, ld, F,
                           fa, self);
        double2 *Zf = (double2 *)Z;  // each frequency's matrix at fa.zo[f]
        float2 *Z32f = Z32 ? (float2 *)Z32 + (size_t)f0 * strideZ32 : nullptr;
        // algorithmic count per frequency: 220 N^2 FP64 operations (440 N^2 with
        // the image terms of the PEC ground), 16 N^2 bytes written
        const double nn = (double)g->npls * g->npls;
        const double ops = g->ienv == 2 ? 440.0 : 220.0;
        prof_begin(ctx, K_FILL, s);
        switch (F) {
            case 1: launch_fill_f<1>(grid, s, g, self, fa, Zf, strideZ, Z32f, strideZ32); break;
            case 2: launch_fill_f<2>(grid, s, g, self, fa, Zf, strideZ, Z32f, strideZ32); break;
            case 3: launch_fill_f<3>(grid, s, g, self, fa, Zf, strideZ, Z32f, strideZ32); break;
            default: launch_fill_f<4>(grid, s, g, self, fa, Zf, strideZ, Z32f, strideZ32); break;
        }
        prof_end(ctx, K_FILL, s, ops * nn * F, (Z32 ? 24.0 : 16.0) * nn * F);
        WM_HIP_TRY(hipGetLastError());
    }
    return WM_OK;
}

}  // namespace wm
