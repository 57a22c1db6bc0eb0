// Time-domain responses from a frequency sweep: the step after the path that
// the reference prompts for but leaves out ("spectrum weight", "wersja
// referencyjna/Wire_MoM.cu":232-248; "multiply the current by the spectrum
// -- for the samples of the time analysis", 390), over sweeps such as its
// 127-point lightning grid freq1.txt (1 kHz - 10 MHz, piecewise uniform).
//
// For outputs m (bases, or any linear response) with transfer H_m(f) sampled
// at increasing frequencies f_0 < ... < f_{K-1} and a source spectrum S(f),
// the real response is
//   r_m(t) = 2 Re int_0^inf H_m(f) S(f) e^{j 2 pi f t} df
// taken with the trapezoid rule over [0, f_0, ..., f_{K-1}], the DC value of
// the (real) response taken as Re(H_m(f_0) S(f_0)):
//   r_m(t_j) = f_0 Re(X_m0) + 2 sum_k w_k Re(X_mk e^{j 2 pi f_k t_j}),  X_mk = H_m(f_k) S(f_k),
// w_0 = f_1 / 2, w_k = (f_{k+1} - f_{k-1}) / 2, w_{K-1} = (f_{K-1} - f_{K-2}) / 2.
// On a uniform grid f_k = (k+1) df this is the inverse DFT, i.e. the
// response periodized with period 1/df, up to the DC extrapolation.
// One workgroup per (output, block of 256 times); the K products
// C_mk = w_k H_mk S_k are staged in LDS in chunks, and each thread sums its
// time sample over them (FP64, one sincos per term).
#include <cmath>
#include <vector>

#include "wm_internal.h"

namespace wm {

namespace {

constexpr int kTrChunk = 1024;  // frequency terms staged per LDS pass (16 KB of complex128)
constexpr double kTwoPi = 6.28318530717958647692;

__global__ __launch_bounds__(256) void k_transient(const double2 *__restrict__ C, const double *__restrict__ fhz,
                                                   const double *__restrict__ dc, int K, int nt, double t0,
                                                   double dt, double *__restrict__ r)
{
    __shared__ double2 sC[kTrChunk];
    __shared__ double sF[kTrChunk];
    const int m = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    const double t = t0 + dt * (double)j;
    double acc = 0.0;
    for (int k0 = 0; k0 < K; k0 += kTrChunk) {
        const int kn = (K - k0 < kTrChunk) ? K - k0 : kTrChunk;
        __syncthreads();
        for (int i = threadIdx.x; i < kn; i += 256) {
            sC[i] = C[(size_t)m * K + k0 + i];
            sF[i] = fhz[k0 + i];
        }
        __syncthreads();
        if (j < nt) {
            for (int i = 0; i < kn; ++i) {
                double s, c;
                sincos(kTwoPi * sF[i] * t, &s, &c);
                acc += sC[i].x * c - sC[i].y * s;
            }
        }
    }
    if (j < nt) r[(size_t)m * nt + j] = dc[m] + 2.0 * acc;
}

}  // namespace

}  // namespace wm

using namespace wm;

extern "C" {

int wm_transient(wm_ctx *c, int nfreq, const float *freq_mhz, int nout, const wm_c128 *H, const wm_c128 *S, int nt,
                 double t0, double dt, double *r)
{
    if (!c || nfreq <= 0 || !freq_mhz || nout <= 0 || !H || nt <= 0 || !(dt > 0.0) || !r) {
        set_error("wm_transient: bad arguments");
        return WM_ERR_ARG;
    }
    for (int k = 0; k < nfreq; ++k)
        if (!(freq_mhz[k] > 0.f) || (k > 0 && !(freq_mhz[k] > freq_mhz[k - 1]))) {
            set_error("wm_transient: frequencies must be positive and increasing (index %d)", k);
            return WM_ERR_ARG;
        }
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    // trapezoid weights over [0, f_0, ..., f_{K-1}]; the DC sample (real) is Re X(f_0)
    std::vector<double> f(nfreq), w(nfreq), dc(nout);
    for (int k = 0; k < nfreq; ++k) f[k] = (double)freq_mhz[k] * 1e6;
    for (int k = 0; k < nfreq; ++k) {
        const double lo = 0.5 * (k > 0 ? f[k] - f[k - 1] : f[0]);
        const double hi = k + 1 < nfreq ? 0.5 * (f[k + 1] - f[k]) : 0.0;
        w[k] = lo + hi;
    }
    std::vector<double2> hc((size_t)nout * nfreq);
    for (int m = 0; m < nout; ++m)
        for (int k = 0; k < nfreq; ++k) {
            const wm_c128 h = H[(size_t)k * nout + m];
            const wm_c128 s = S ? S[k] : wm_c128{1.0, 0.0};
            const double xr = h.re * s.re - h.im * s.im, xi = h.re * s.im + h.im * s.re;
            hc[(size_t)m * nfreq + k] = make_double2(w[k] * xr, w[k] * xi);
            if (k == 0) dc[m] = f[0] * xr;
        }
    double2 *dC = nullptr;
    double *dF = nullptr, *dR = nullptr, *dD = nullptr;
    auto release = [&]() {
        (void)hipFree(dC);
        (void)hipFree(dF);
        (void)hipFree(dR);
        (void)hipFree(dD);
    };
    const hipStream_t s = c->stream;
    int rc = WM_OK;
    if (hipMalloc(&dC, hc.size() * sizeof(double2)) != hipSuccess || hipMalloc(&dF, f.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&dR, (size_t)nout * nt * sizeof(double)) != hipSuccess || hipMalloc(&dD, dc.size() * sizeof(double)) != hipSuccess) {
        release();
        set_error("wm_transient: device allocation failed");
        return WM_ERR_HIP;
    }
    rc = copy_h2d(dC, hc.data(), hc.size() * sizeof(double2), s);
    if (rc == WM_OK) rc = copy_h2d(dF, f.data(), f.size() * sizeof(double), s);
    if (rc == WM_OK) rc = copy_h2d(dD, dc.data(), dc.size() * sizeof(double), s);
    if (rc == WM_OK) {
        hipLaunchKernelGGL(k_transient, dim3((unsigned)((nt + 255) / 256), (unsigned)nout), dim3(256), 0, s, dC, dF,
                           dD, nfreq, nt, t0, dt, dR);
        if (hipGetLastError() != hipSuccess) {
            set_error("wm_transient: launch failed");
            rc = WM_ERR_HIP;
        }
    }
    if (rc == WM_OK) rc = copy_d2h(r, dR, (size_t)nout * nt * sizeof(double), s);
    (void)hipStreamSynchronize(s);
    release();
    return rc;
}

int wm_spectrum_dexp(int nfreq, const float *freq_mhz, double i0, double alpha, double beta, wm_c128 *S)
{
    if (nfreq <= 0 || !freq_mhz || !S || !(alpha > 0.0) || !(beta > 0.0)) {
        set_error("wm_spectrum_dexp: bad arguments");
        return WM_ERR_ARG;
    }
    // i0 (e^{-alpha t} - e^{-beta t}) u(t)  <->  i0 (1/(alpha + j w) - 1/(beta + j w)),  w = 2 pi f
    for (int k = 0; k < nfreq; ++k) {
        const double w = kTwoPi * (double)freq_mhz[k] * 1e6;
        const double da = alpha * alpha + w * w, db = beta * beta + w * w;
        S[k] = wm_c128{i0 * (alpha / da - beta / db), i0 * (-w / da + w / db)};
    }
    return WM_OK;
}

}  // extern "C"
