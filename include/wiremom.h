/*
 * wiremom.h -- C ABI of libwiremom.so, the MI355X (gfx950) wire-MoM hot path:
 * impedance-matrix fill, dense complex LU factor/solve, current and Zin.
 *
 * Two layers share one implementation:
 *
 *  1. Drop-in entry points with the reference program's names and parameter
 *     lists (zmatrix, zmatrix_gpu, CalculateByFreq, LU_cuBLAS, geoinp,
 *     config, configwire).  They take and return host arrays exactly as the
 *     reference's do; complex values are 16-byte {double re, im} pairs
 *     (wm_c128, layout-identical to the reference's clDoubleComplex).
 *
 *  2. The native wm_* API: a device context, a geometry object resident in
 *     HBM, batched fill / getrf / getrs on device buffers, whole frequency
 *     sweeps in FP64 or in mixed precision with FP64 iterative refinement,
 *     and a continuous-batching request server.
 *
 * Status codes: every wm_* call returns int.  0 = ok; >0 = LAPACK-style info
 * (1-based column of the first exactly-zero pivot); <0 = error, see WM_ERR_*
 * and wm_last_error() for a message (thread-local).
 *
 * Matrices are column-major complex (interleaved re, im), element (i, j) of
 * a batch member b at  A + b*strideA + (i + j*lda)  in complex units.
 */
#ifndef WIREMOM_H
#define WIREMOM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } wm_c128;
typedef struct { float re, im; } wm_c64;

/* ---- status ------------------------------------------------------------ */
#define WM_OK 0
#define WM_ERR_ARG (-1)       /* bad argument */
#define WM_ERR_HIP (-2)       /* HIP runtime error */
#define WM_ERR_NOMEM (-3)     /* device or host allocation failed / over budget */
#define WM_ERR_SPIN (-4)      /* a bounded device-side wait expired */
#define WM_ERR_NOCONV (-5)    /* iterative refinement did not converge */
#define WM_ERR_IO (-6)        /* file could not be read or written */
#define WM_ERR_GEOMETRY (-7)  /* geometry capacity or topology error */
#define WM_ERR_STATE (-8)     /* call not valid in the object's state */

/* Precision of the factorization. */
#define WM_PREC_F64 0         /* complex128 LU, FP64 MFMA trailing update */
#define WM_PREC_F32 1         /* complex64 LU, FP32 MFMA update, FP64 refinement */
#define WM_PREC_BF16X3 2      /* complex64 LU, split-BF16x3 MFMA update, FP64 refinement */

const char *wm_last_error(void);
const char *wm_version(void);

/* ---- context ------------------------------------------------------------ */
typedef struct wm_ctx wm_ctx;

/* Opens device `device` (HIP ordinal).  One context per device and host
 * thread group; calls on one context are serialised internally. */
int wm_init(int device, wm_ctx **out);
int wm_shutdown(wm_ctx *ctx);
/* Synchronises the context's stream. */
int wm_sync(wm_ctx *ctx);
/* Device name, CU count and HBM bytes. */
int wm_device_info(wm_ctx *ctx, char *name, int name_len, int *num_cu, size_t *hbm_bytes);

/* Profiler: when enabled, every launch of the timed kernel classes ("fill",
 * "panel", "swap_trsm", "gemm", "trsv", "resid") is bracketed by HIP events on
 * its own stream, and its algorithmic flops and bytes are recorded.
 * wm_prof_enable(ctx, 1) also clears earlier records.  wm_prof_get returns
 * the summed event time (ms), the launch count, and the summed algorithmic
 * flops and bytes of one class. */
int wm_prof_enable(wm_ctx *ctx, int on);
int wm_prof_get(wm_ctx *ctx, const char *kernel, double *ms, int64_t *launches, double *flops, double *bytes);

/* Device memory helpers so that callers need no HIP headers. */
int wm_dev_alloc(wm_ctx *ctx, size_t bytes, void **ptr);
int wm_dev_free(wm_ctx *ctx, void *ptr);
int wm_memcpy_h2d(wm_ctx *ctx, void *dst, const void *src, size_t bytes);
int wm_memcpy_d2h(wm_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy on the context's stream; returns when it is done. */
int wm_memcpy_d2d(wm_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- geometry ------------------------------------------------------------ */
typedef struct wm_geom wm_geom;

/* Builds the device geometry from the arrays configwire produces (0-based
 * positions, 1-based stored ids, two-column arrays with stride nsmax), for
 * basis functions 0..npls-1.  Free space (ienv = 1). */
int wm_geom_create(wm_ctx *ctx, const float *x, const float *y, const float *z, const float *rad,
                   const int *ipc, const int *ips, const int *ise, int nsmax, int npls, wm_geom **out);
/* As wm_geom_create with the environment arguments of the reference's
 * zmatrix ("wersja referencyjna/Zmatrix.h":7-9): ienv 1 free space, 2 PEC
 * ground z = 0 by images, whose source coordinates are scaled by
 * (sgnx, sgny, sgnz) and whose terms are weighted by sgnenv (the reference
 * uses 1, 1, -1, -1: "wersja referencyjna/Wire_MoM.cu":107-122).  With
 * ienv 2 the fill evaluates twice the kernels (Zmatrix.cpp:94-145, 174-224).
 * Other ienv values are rejected (WM_ERR_ARG). */
int wm_geom_create_env(wm_ctx *ctx, const float *x, const float *y, const float *z, const float *rad,
                       const int *ipc, const int *ips, const int *ise, int nsmax, int npls, int ienv,
                       float sgnx, float sgny, float sgnz, float sgnenv, wm_geom **out);
/* Reads a .geo file and runs the reference segmentation (configwire) on it;
 * writes configwire.txt to `wirefile` unless it is NULL or empty. */
int wm_geom_load(wm_ctx *ctx, const char *geo_path, const char *wirefile, wm_geom **out);
/* As wm_geom_load for environment ienv (2: PEC ground; configwire then adds
 * the virtual ground wires at nodes with z = 0, "wersja referencyjna/
 * Configuration.cpp":131-202, and the signs are the reference's). */
int wm_geom_load_env(wm_ctx *ctx, const char *geo_path, const char *wirefile, int ienv, wm_geom **out);
int wm_geom_destroy(wm_geom *g);
/* Number of unknowns N, and the padded leading dimension used on device. */
int wm_geom_npls(const wm_geom *g);
int wm_geom_ld(const wm_geom *g);

/* ---- fill ----------------------------------------------------------------- */
/* Fills nfreq complex128 matrices Z_f (ld = wm_geom_ld, padded rows/columns
 * set to the identity) on device, f = 0..nfreq-1, Z_f at Z + f*strideZ.
 * ka[f] is the wavenumber (host array).  If Z32 is not NULL, a complex64
 * copy is written there as well (same ld, stride strideZ32). */
int wm_zfill(wm_ctx *ctx, const wm_geom *g, int nfreq, const double *ka,
             wm_c128 *Z, int64_t strideZ, wm_c64 *Z32, int64_t strideZ32);

/* ---- dense LU (device buffers) ---------------------------------------------- */
/* Batched LU with partial pivoting (pivot = first max of |re|+|im|).
 * A: n x n per member (n a multiple of 64), complex128 for WM_PREC_F64 and
 * complex64 otherwise.  ipiv: device int32, n per member, 1-based.
 * info (host, batch entries, may be NULL). */
int wm_getrf_batched(wm_ctx *ctx, int prec, int n, void *A, int lda, int64_t strideA,
                     int *ipiv, int64_t strideP, int batch, int *info);
/* Batched factorizations are captured into HIP graphs per (shape, buffers)
 * and replayed (default on; the profiler, when enabled, times direct launches).
 * wm_set_graphs(ctx, 0) launches directly and drops the captured graphs.  A
 * capture is invalidated by device-synchronizing calls (allocations, frees,
 * synchronous copies) that another host thread makes while it is open: a
 * context used while other threads drive the same device should launch
 * directly (the request server's lanes do). */
int wm_set_graphs(wm_ctx *ctx, int on);
/* Recycled sweeps (wm_solve_opts.recycle > 1) fill the matrices of the
 * frequencies refined by GMRES on a second stream of the context while the
 * anchor frequencies are factored on its own stream (default on; off while
 * the profiler is enabled, and for sweeps with lumped loads, condition
 * estimates or structures of several sizes).  Results are identical either
 * way.  wm_set_fill_overlap(ctx, 0) keeps every launch on one stream. */
int wm_set_fill_overlap(wm_ctx *ctx, int on);
/* Tuning (replaces the reference's fixed (8, 8) launch shape of the fill,
 * "wersja referencyjna/Wire_MoM.cu":4110-4113, and its fixed cuBLAS calls).
 *
 * Factorization: the blocked LU applies a trailing update after every
 * `panels` panels of 64 columns (K = 64 * panels), per factor precision and
 * size class (n / 1024).  wm_set_getrf_block sets it (0 = default: 4, and 8
 * at n >= 8192; else 1, 2, 4, 8 or 16); wm_tune_getrf times one batch of
 * `batch` (0: one launch width) synthetic n x n factorizations for each
 * choice, keeps the fastest for this context and returns it with its time. */
int wm_set_getrf_block(wm_ctx *ctx, int prec, int n, int panels);
int wm_get_getrf_block(wm_ctx *ctx, int prec, int n);
int wm_tune_getrf(wm_ctx *ctx, int prec, int n, int batch, int *best_panels, double *ms_best);
/* Fill: frequencies per fill launch (1-4; 0 = default, 4), per environment
 * (ienv 1 free space, 2 PEC ground) and size class (ld / 1024).  A matrix's
 * entries do not depend on it (bit for bit), only the time does.
 * wm_tune_fill times fills of four frequencies of `g` at each setting, keeps
 * the fastest for g's environment and size class and returns it with the time
 * per frequency. */
int wm_set_fill_freqs(wm_ctx *ctx, int ienv, int n, int freqs);
int wm_get_fill_freqs(wm_ctx *ctx, int ienv, int n);
int wm_tune_fill(wm_ctx *ctx, const wm_geom *g, int *best_freqs, double *ms_per_freq);
/* Tuning cache: wm_tune_save writes every setting made on this context
 * (tuned or set) as JSON lines keyed by (architecture, knob, precision or
 * environment, size class); wm_tune_load applies the entries of a file that
 * were made on this device's architecture and returns how many it applied
 * (negative: error).  Unknown keys and other architectures are skipped. */
int wm_tune_save(wm_ctx *ctx, const char *path);
int wm_tune_load(wm_ctx *ctx, const char *path);
/* Solves A X = B with the factors of wm_getrf_batched; B complex128
 * (n x nrhs, ldb) per member, overwritten by X.  The factors may be complex64
 * (prec != WM_PREC_F64); the solve then runs in FP64 on widened values. */
int wm_getrs_batched(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                     const int *ipiv, int64_t strideP, wm_c128 *B, int ldb, int64_t strideB, int batch);

/* wm_getrs_batched without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Its staging,
 * permutation and packet buffers live in a 32 MiB device array of the library
 * that each GPU holds until the process ends; sizes needing more than that
 * (batch * (4 n + 16 strideB + 64 n nr') bytes with nr' = min(nrhs, 8)
 * rounded down to a power of two, each term rounded up to 256) are rejected
 * with WM_ERR_NOMEM.  Same arguments and results as wm_getrs_batched.  Every
 * call on one GPU uses the same array, so calls that overlap on one GPU race:
 * serialise them (one context's calls already are).  The profiler does not
 * time its launches. */
int wm_getrs_batched_noalloc(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                             const int *ipiv, int64_t strideP, wm_c128 *B, int ldb, int64_t strideB, int batch);

/* ---- whole solves ------------------------------------------------------------ */
typedef struct {
    int prec;              /* WM_PREC_* */
    int max_iter;          /* refinement cap (mixed precision), default 20; an item whose
                              contraction rate predicts more is refactored in FP64 early */
    double tol;            /* stop when ||dx||/||x|| <= tol, default 1e-12 */
    int batch;             /* frequencies in flight, 0 = auto by HBM budget */
    double hbm_fraction;   /* share of free HBM the sweep may use, default 0.8 */
    int fallback_f64;      /* refactor in FP64 if refinement fails, default 1 */
    int size_policy;       /* factor a batch in FP64 where that is cheaper at its size, default 1 */
    int cond_est;          /* estimate the 1-norm condition number of every Z (wm_freq_stats.cond1),
                              default 0: Hager-Higham's estimator (LAPACK zlacon's algorithm)
                              with the final factors, about ten extra triangular solves per batch */
    double recycle;        /* recycled sweep, default 0 (off).  A value r > 1 factors (in FP64) only
                              "anchor" frequencies, chosen so that every frequency f of the sweep has
                              an anchor within a factor r(f) of it -- r^(1 + log10(1 / (k D))) while
                              the structure is electrically small (k D <= 1, D its diameter: one more
                              power of r per decade below k D = 1), 1 + (r - 1) / (k D) above -- and
                              solves each other frequency by FP64 GMRES iterative
                              refinement preconditioned with its anchor's factors, until the
                              correction is below tol or stops shrinking at the FP64 floor; a frequency
                              whose refinement does not converge is factored itself.  prec is
                              ignored. */
} wm_solve_opts;

void wm_solve_opts_default(wm_solve_opts *o);

/* Per-frequency statistics of a sweep. */
typedef struct {
    int iters;             /* refinement iterations (0 for FP64) */
    int info;              /* LU info */
    int fell_back;         /* FP64 factors produced the result: 1 after refinement failed,
                              2 routed there directly (predicted not to refine); -1: refinement
                              failed with fallback_f64 off (the result is its last iterate) */
    int prec;              /* WM_PREC_* of the factors that produced the result */
    double rate;           /* last refinement contraction ||dx_k|| / ||dx_k-1|| (-1: none) */
    double dx_rel;         /* last ||dx||_inf / ||x||_inf: the refined result's relative
                              distance from the FP64 solution is about dx_rel * rate / (1 - rate) */
    double cond1;          /* with cond_est: estimate of ||Z||_1 ||Z^-1||_1 (a lower bound, in
                              practice within a factor 3), else 0.  Two FP64 solutions of Z x = b
                              with backward errors e1, e2 differ by up to about cond1 * (e1 + e2)
                              relative: where cond1 * 1e-16 approaches 1e-10, agreement to 1e-10
                              needs identical arithmetic */
    double anorm1;         /* with cond_est: ||Z||_1 (largest column sum of moduli), else 0 */
    double anchor_mhz;     /* recycled sweep (wm_solve_opts.recycle): the frequency whose FP64 factors
                              produced the result (this one's own for anchors); else 0 */
    double berr;           /* recycled sweep, GMRES-refined frequencies: normwise backward error
                              ||b - Z x||_inf / (||Z||_inf ||x||_inf + ||b||_inf) of the result; else 0 */
} wm_freq_stats;

/* Sweep over nfreq frequencies (MHz, float32 like the reference's freq.txt)
 * with nrhs delta-gap feeds (1-based basis indices gaps[r], source value
 * ampl*exp(j*pi*phas/180)).  Outputs (host, may be NULL):
 *   I   : complex128 [nfreq][nrhs][npls]   current coefficients
 *   zin : complex128 [nfreq][nrhs]          V/I at each feed
 *   st  : wm_freq_stats [nfreq]
 * An FP64 sweep factors as many frequencies per batch as the HBM budget
 * allows (members of a batch go through the panels together).  In mixed
 * precision the frequencies are solved highest first, in batches at most one
 * full-height panel launch wide; the refinement contraction rates already
 * seen, scaled by (f_obs/f)^3, predict the rate at each lower frequency, and
 * each is factored in the cheapest precision predicted to reach tol within
 * max_iter: the requested one, else FP32 (for BF16X3), else FP64 directly
 * (fell_back = 2).  Batches are cut where the routed precision changes; the
 * run of frequencies routed to FP64 goes as one batch, and a short run of
 * complex64 frequencies just above it joins that batch when its own panel
 * launches and refinement are predicted to cost more.  A batch whose size
 * makes FP64 factors cheaper than complex64 factors plus refinement (small N,
 * where both factorizations are bound by panel latency) is factored in FP64
 * directly as well.  A split-BF16x3 attempt that does not converge is retried
 * with FP32 factors, an FP32 one with FP64 factors (fell_back = 1).
 * Returns the first nonzero status. */
int wm_solve_sweep(wm_ctx *ctx, const wm_geom *g, const wm_solve_opts *opts, int nfreq,
                   const float *freq_mhz, int nrhs, const int *gaps, float ampl, float phas,
                   wm_c128 *I, wm_c128 *zin, wm_freq_stats *st);

/* Lumped series load on basis function `basis` (1-based), added to Z's
 * diagonal before the factorization: Z_L = r + j(2*pi*f*l - 1/(2*pi*f*c)) with
 * f in MHz, l in uH and c in uF (c = 0: no capacitor) -- the units of the
 * reference's hard-wired load block, "wersja referencyjna/Wire_MoM.cu":346-363
 * (iload == 1: 1 + j*2*pi*f*4.5 on bases 592..2581, 30 on 2613, 1e12 on 2582,
 * which is this list with l = 4.5, r = 30 and r = 1e12).  Loads on one basis
 * are summed before they are added. */
typedef struct {
    int basis;
    double r_ohm, l_uh, c_uf;
} wm_load;

/* Delta-gap generator: voltage ampl*exp(j*pi*phas/180) across basis `basis`
 * (1-based) in right-hand side `rhs` (0-based).  Several generators may share
 * a right-hand side (an array driven at several points); the reference has
 * one generator ("wersja referencyjna/Wire_MoM.cu":183, 300-309). */
typedef struct {
    int basis, rhs;
    float ampl, phas;
} wm_source;

/* Incident plane wave in right-hand side `rhs` (0-based): the reference's
 * excitation 2, prompted for but not built ("wersja referencyjna/Wire_MoM.cu":
 * 171-181).  The wave arrives from direction (theta, phi) [degrees], unit
 * vector rh = (sin t cos p, sin t sin p, cos t):
 *   E_inc(r) = (e_theta th + e_phi ph) exp(+j k rh.r)  [V/m],
 * th = (cos t cos p, cos t sin p, -sin t), ph = (-sin p, cos p, 0).  Basis m
 * receives sum_h l_h . E_inc(r_h) over its two half-segment midpoints r_h with
 * half vectors l_h, the points and vectors at which the fill tests the field
 * (the same convention as a delta gap: a field along the basis direction acts
 * as a positive gap voltage).  Over the PEC ground add the reflected wave as a
 * second plane wave. */
typedef struct {
    int rhs;
    double theta_deg, phi_deg;
    wm_c128 e_theta, e_phi;
} wm_plane_wave;

/* wm_solve_sweep with general excitations and lumped loads: nrhs right-hand
 * sides built from nsrc generators and npw plane waves, nload loads (counts may
 * be 0 with NULL arrays; at least one generator or plane wave).
 *   I   : complex128 [nfreq][nrhs][npls]
 *   zin : complex128 [nfreq][nsrc]   V_s / I at generator s (its active impedance)
 * Routing, batching and statistics as wm_solve_sweep. */
int wm_solve_sweep_ex(wm_ctx *ctx, const wm_geom *g, const wm_solve_opts *opts, int nfreq,
                      const float *freq_mhz, int nrhs, int nsrc, const wm_source *src, int npw,
                      const wm_plane_wave *pw, int nload, const wm_load *loads, wm_c128 *I, wm_c128 *zin,
                      wm_freq_stats *st);

/* ---- time domain ------------------------------------------------------------------ */
/* Real responses from a sweep, the time analysis the reference prompts for
 * but leaves out ("spectrum weight", "wersja referencyjna/Wire_MoM.cu":232-248,
 * 390): for outputs m with transfer H[k][m] at increasing frequencies
 * freq_mhz[k] (MHz, any spacing, e.g. the piecewise-uniform freq1.txt) and
 * source spectrum S[k] (NULL: 1),
 *   r[m][j] = f_0 Re X[0][m] + 2 sum_k w_k Re(X[k][m] exp(j 2 pi f_k t_j)),  X = H S,
 * t_j = t0 + j dt [s]: the trapezoid rule for 2 Re int_0^inf X e^{j 2 pi f t} df
 * (f in Hz) over [0, f_0, ..., f_{K-1}] with the DC value taken as Re X(f_0)
 * (w_k = (f_{k+1} - f_{k-1}) / 2, f_{-1} = 0, f_K = f_{K-1}).  A sweep's currents for a
 * unit delta gap give H in A/V; divided by the current at the feed they give
 * the current transfer of a current injection.  Runs on the GPU. */
int wm_transient(wm_ctx *ctx, int nfreq, const float *freq_mhz, int nout, const wm_c128 *H, const wm_c128 *S,
                 int nt, double t0, double dt, double *r);
/* Spectrum of the double-exponential source i0 (e^{-alpha t} - e^{-beta t}),
 * t >= 0 (alpha, beta in 1/s): S = i0 (1/(alpha + j w) - 1/(beta + j w)), w = 2 pi f. */
int wm_spectrum_dexp(int nfreq, const float *freq_mhz, double i0, double alpha, double beta, wm_c128 *S);

/* ---- serving: continuous batching ------------------------------------------------ */
typedef struct wm_server wm_server;

typedef struct {
    size_t hbm_budget;     /* bytes the server may hold for requests, 0 = 80% of free HBM */
    int max_inflight;      /* frequencies in flight at once, 0 = auto */
    int prec;              /* WM_PREC_* for every request */
    int streams;           /* lanes: worker threads with their own device stream and workspaces, whose
                              batches run on the device at the same time; default 2, at most 4 */
    int size_policy;       /* as wm_solve_opts.size_policy, default 1 */
} wm_server_opts;

void wm_server_opts_default(wm_server_opts *o);
int wm_server_create(wm_ctx *ctx, const wm_server_opts *o, wm_server **out);
/* Queues a request: a geometry (the server keeps a reference; do not destroy
 * it before the request completes), its frequencies and feeds.  Returns a
 * request id (>= 0) in *id.  The request is admitted once its HBM need fits
 * the budget; admission never blocks the caller. */
int wm_server_submit(wm_server *s, const wm_geom *g, int nfreq, const float *freq_mhz, int nrhs,
                     const int *gaps, float ampl, float phas, int64_t *id);
/* Drives the device: retires the batches the lanes finished, admits queued
 * requests, hands a batch to every idle lane and completes the requests whose
 * frequencies are all done; the number completed by this call is in *done.
 * Non-blocking unless `wait` is nonzero: then it returns once a request has
 * completed, a request has been submitted, or nothing is pending.  Call it
 * from one thread.  A batch that fails completes its requests with the error
 * status (wm_server_result returns it). */
int wm_server_step(wm_server *s, int wait, int *done);
/* Batches solved so far, the host time during which at least one lane was
 * busy, and the time during which two or more were (device overlap). */
int wm_server_overlap(wm_server *s, int64_t *batches, double *busy_s, double *overlap_s);
/* Results of a completed request (host arrays sized as in wm_solve_sweep);
 * latency = completion - submission in seconds.  Releases the request.
 * Returns the request's status: negative if a batch it belonged to failed. */
int wm_server_result(wm_server *s, int64_t id, wm_c128 *I, wm_c128 *zin, wm_freq_stats *st,
                     double *latency_s);
/* Ids of completed requests whose results have not been read (ascending, at
 * most cap of them). */
int wm_server_done(wm_server *s, int64_t *ids, int cap, int *count);
/* Queued, admitted and completed-but-unread request counts, HBM in use. */
int wm_server_status(wm_server *s, int *queued, int *running, int *completed, size_t *hbm_used);
int wm_server_destroy(wm_server *s);

/* ---- drop-in entry points (reference names and parameter lists) --------------------- */

/* Replaces geoinp, "wersja referencyjna/Configuration.cpp":5-75 and
 * Configuration.h:10-11.  The reference reads the fixed file name dipol.geo;
 * this reads WM_GEO_PATH if set, else dipol.geo.  Arrays are 1-based. */
void geoinp(int nnmax, int nwmax, int ntmax, int *nnod, float *xnod, float *ynod, float *znod,
            int *nw, int *iwn, float *wrad, int *nsw, int *nt, int *iwt);

/* Replaces config, "wersja referencyjna/Configuration.cpp":78-108 and
 * Configuration.h:13-15 (npls = nplsw).  Writes configwire.txt. */
void config(int *npnt, int nnod, float *xnod, float *ynod, float *znod, float *x, float *y, float *z,
            int nw, int *iwn, float *wrad, int *nsw, int *ise, int *isp, int *ips, int *ipc, int *iwe,
            float *rad, int *npls, int *nplsw, int *nseg, int nsmax, int nwmax, int ienv);

/* Replaces configwire, "wersja referencyjna/Configuration.cpp":112-520 and
 * Configuration.h:17-19.  Writes configwire.txt in the working directory. */
void configwire(int *npnt, int nnod, float *xnod, float *ynod, float *znod, float *x, float *y, float *z,
                int nw, int *iwn, float *wrad, int *nsw, int *ise, int *isp, int *ips, int *ipc, int *iwe,
                float *rad, int *nplsw, int *nseg, int nsmax, int nwmax, int ienv);

/* Replaces tablice, "wersja referencyjna/Configuration.cpp":524-556 and
 * Configuration.h:21: writes the geometry tables (points, basis peak nodes,
 * segment nodes, basis segments) to tablice.txt in the working directory. */
void tablice(int npnt, int nseg, int nsmax, float *x, float *y, float *z, int nplsw, int *ise, int *isp,
             int *ips, int *ipc);

/* Replaces zmatrix, "wersja referencyjna/Zmatrix.cpp":27-52 and Zmatrix.h:7-9:
 * fills cz (host, npls x npls, column-major) for wavenumber ka.  The fill
 * runs on the GPU.  ienv 1 (free space) or 2 (PEC ground z = 0 with the
 * image signs sgn*, as wm_geom_create_env); other values print an error. */
void zmatrix(wm_c128 *cz, float *x, float *y, float *z, int *ipc, int *ips, int *ise, float *rad,
             int npls, double ka, int nsmax, int ienv, float sgnx, float sgny, float sgnz, float sgnenv);

/* Replaces zmatrix_gpu, "wersja referencyjna/Wire_MoM.cu":4052-4197 and
 * Wire_MoM.h:49-51 (the "wersja gpu/Wire_MoM.h":45-47 form with `rank`
 * is zmatrix_gpu_rank).  Same contract as zmatrix. */
void zmatrix_gpu(wm_c128 *cz_h, float *x_h, float *y_h, float *z_h, float *rad_h, int *ipc_h,
                 int *ips_h, int *ise_h, int nsmax, double ka, int npls, int ienv, float sgnx,
                 float sgny, float sgnz, float sgnenv);
void zmatrix_gpu_rank(wm_c128 *cz_h, float *x_h, float *y_h, float *z_h, float *rad_h, int *ipc_h,
                      int *ips_h, int *ise_h, int nsmax, double ka, int npls, int ienv, float sgnx,
                      float sgny, float sgnz, float sgnenv, int rank);

/* Replaces LU_cuBLAS, "wersja gpu/Wire_MoM.cu":417-504 and
 * "wersja gpu/Wire_MoM.h":43: solves cz * X = ci_cu in FP64 on the GPU,
 * ci_cu (host, n x nrhs) in, solution out; cz (host) is not modified.
 * cu_cu is unused, as in the reference. batchSize must be 1. */
void LU_cuBLAS(int n, int nrhs, wm_c128 *ci_cu, wm_c128 *cu_cu, wm_c128 *cz, int batchSize, int rank);

/* Replaces CalculateByFreq, "wersja gpu/Wire_MoM.cu":355-413 and
 * "wersja gpu/Wire_MoM.h":49-52: one frequency (MHz) end to end -- fill,
 * optional hard-wired loads (iload == 1, reference behaviour), delta-gap
 * excitation at basis ngap, FP64 LU solve on the GPU -- and the reference's
 * "Process - r Freq = f Re Im" line on stdout.  cz receives Z (host),
 * cu_cu the excitation, ci_cu the currents. */
void CalculateByFreq(int i, int npls, int nfreq, float freq, float *x, float *y, float *z, float *rad,
                     int *ipc, int *ips, int *ise, const int nsmax, int ienv, float sgnx, float sgny,
                     float sgnz, float sgnenv, int iload, float phas, int ngap, float ampl, wm_c128 *cz,
                     wm_c128 *cu_cu, wm_c128 *ci_cu, int rank);

#ifdef __cplusplus
}
#endif

#endif /* WIREMOM_H */
