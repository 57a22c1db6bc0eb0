// Internal declarations shared by the libwiremom sources (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wiremom.h"

namespace wm {

void set_error(const char *fmt, ...);

#define WM_HIP_TRY(expr)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::wm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),       \
                            __FILE__, __LINE__);                                         \
            return WM_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)

#define WM_TRY(expr)                 \
    do {                             \
        int r_ = (expr);             \
        if (r_ != WM_OK) return r_;  \
    } while (0)

// Device buffer that only grows; reused across calls on one context.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
};

// Device-side error word written by kernels with bounded waits.
enum DevErr : int { DEV_OK = 0, DEV_SPIN_TIMEOUT = 1 };

}  // namespace wm

namespace wm {
// Kernel classes timed by the optional profiler (events on the launch stream).
enum KClass { K_FILL = 0, K_PANEL, K_SWAP, K_GEMM, K_TRSV, K_RESID, K_PANEL32, K_SWAP32, K_GEMM32, K_GEMMB3, K_NCLASS };
struct ProfRec {
    hipEvent_t a, b;
    double flops, bytes;
};
struct Prof {
    bool on = false;
    std::vector<ProfRec> rec[K_NCLASS];
    std::vector<hipEvent_t> pool;
};
}  // namespace wm

namespace wm {
// Captured factorization: precision, n, lda, strideA, A, ipiv, strideP, batch,
// info, index table, outer-block depth, exchange buffer.
typedef std::array<int64_t, 13> GraphKey;
}  // namespace wm

struct wm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    size_t hbm_bytes = 0;
    size_t max_lds = 0;
    char name[256] = {0};
    std::recursive_mutex mu;
    // small persistent device scratch: sync counters and the error word
    int *d_sync = nullptr;  // kSyncCap counters
    int *d_err = nullptr;   // 1 int
    static const int kSyncCap = 4096;
    wm::DevBuf ws_info;     // info words of wm_getrf_batched
    wm::DevBuf ws_fill;     // per-frequency self-term integrals
    wm::DevBuf ws_lu;       // panel exchange buffers
    wm::DevBuf ws_solve;    // triangular-solve permutation and staging
    wm::DevBuf ws_pkt;    // triangular solves: solved values as packets (epoch tags only)
    wm::DevBuf ws_perm;     // row permutations for getrs on raw pivots
    wm::DevBuf ws_resid;    // residual partial sums
    wm::DevBuf ws_sweep;    // matrices and vectors of a sweep batch
    wm::DevBuf ws_group;    // grouped triangular solves: factor and column tables, staging
    wm::DevBuf ws_dinv;     // inverted diagonal blocks of preconditioner factors
    wm::DevBuf ws_dinv_idx; // their factor indices
    int solve_epoch = 0;    // monotonic tag of triangular-solve packets
    uint64_t panel_tag = 0; // packet tags used since the exchange buffer was last cleared
    uint32_t *d_tag = nullptr;  // device tag base, bumped by every factorization launch chunk
    bool use_graphs = true;     // replay captured factorizations (wm_set_graphs)
    static const size_t kGraphCap = 64;
    std::map<wm::GraphKey, hipGraphExec_t> graphs;
    // factorization outer-block depth (panels per trailing update) per factor
    // precision and size class (ld / 1024, capped), 0 = default; set by
    // wm_tune_getrf or wm_set_getrf_block
    static const int kTuneSizes = 32;
    int getrf_block[3][kTuneSizes] = {};
    // frequencies per fill launch per environment (free space, PEC ground)
    // and size class (ld / 1024), 0 = default; wm_tune_fill or wm_set_fill_freqs
    int fill_freqs[2][kTuneSizes] = {};
    char arch[64] = {0};    // gcnArchName of the device (tuning cache key)
    wm::Prof prof;          // per-kernel-class event timing (off by default)
    // panel workgroups assumed co-resident (0 = the occupancy query's count);
    // lowered only by tests, to split panel launches at small sizes
    int panel_resident_cap = 0;
    // contexts that may run panel launches on the device at the same time (the
    // request server's lanes): each takes this share of the resident workgroups
    int panel_share = 1;
    // recycled sweeps: the dependents' matrices are filled on this second
    // stream (own self-term workspace) while the anchors are factored on
    // `stream`; created on first use
    // second stream of recycled sweeps (wm_set_fill_overlap): the GMRES items'
    // fills, with their own self-term workspace; created on first use
    bool overlap_fill = true;
    int group_cols = 4;  // preconditioner solves: right-hand sides per launch member (1, 2, 4 or 8)
    hipStream_t aux = nullptr;
    hipEvent_t aux_ev[2] = {nullptr, nullptr};  // launch stream -> aux, aux -> launch stream
    wm::DevBuf ws_fill2;
};

// Geometry as the fill kernel reads it.  Observation data per basis (SoA,
// one lane per test basis) and source data per basis (AoS, wave-uniform).
struct wm_geom {
    wm_ctx *ctx = nullptr;
    int npls = 0;     // N, unknowns
    int ld = 0;       // padded size (multiple of 128)
    double diam = 0;  // bounding-box diagonal of the structure (with its image over the PEC ground), m
    double *d_obs = nullptr;   // 12 arrays of ld doubles: rc1xyz, lm1xyz, rc2xyz, lm2xyz
    int *d_obs_key = nullptr;  // 2 arrays of ld ints: segment key of obs half 1, 2
    double *d_src = nullptr;   // ld records of 16 doubles: rs1[3] dc1[3] dl1 an1 rs2[3] dc2[3] dl2 an2
    int *d_src_key = nullptr;  // ld records of 2 ints
    // PEC ground (ienv 2): image records (coordinates and directions scaled by
    // sgnx, sgny, sgnz) laid out as d_src, image terms weighted by sgnenv; an
    // image half's key is its segment key when the self rule applies to it
    // (both nodes in the ground plane), else -3 (never an observation key)
    int ienv = 1;
    double sgnenv = 1.0;
    double *d_img = nullptr;
    int *d_img_key = nullptr;
    std::vector<double> h_src; // host copy of the source records (self-term precompute)
    // observed refinement contraction rates of mixed-precision solves on this
    // structure; see wm::predict_rate
    struct RateObs {
        float freq;   // MHz
        int prec;     // WM_PREC_BF16X3 or WM_PREC_F32 factors
        double rate;  // last contraction ||dx_k|| / ||dx_k-1|| (inf: singular factor)
    };
    mutable std::vector<RateObs> rate_obs;
    mutable std::mutex rate_mu;  // guards rate_obs (the server's lanes record rates concurrently)
};

namespace wm {

// Fill: nf frequencies, complex128 output (+ optional complex64 copy).
// Frequency f's matrix goes to Z + (slots ? slots[f] : f) * strideZ (slots:
// complex128 only); self_ws: the self-term workspace (default ctx->ws_fill).
int fill_launch(wm_ctx *ctx, const wm_geom *g, int nf, const double *ka, void *Z, int64_t strideZ,
                void *Z32, int64_t strideZ32, hipStream_t s,
                DevBuf *self_ws = nullptr, const int *slots = nullptr);

// LU: batched blocked right-looking factorization with partial pivoting.
// d_bidx (device, optional): member i of the launch works on matrix d_bidx[i]
// (the matrix of index m lives at base + m * stride); null means 0..batch-1.
// d_nmem (device, optional, per matrix index): the matrix's own size, a
// multiple of 64 not above n, rows and columns beyond it the identity; the
// factorization skips the work there (its result is the identity's).
int getrf_batched(wm_ctx *ctx, int prec, int n, void *A, int lda, int64_t strideA, int *ipiv,
                  int64_t strideP, int batch, int *d_info, hipStream_t s, const int *d_bidx = nullptr,
                  const int *d_nmem = nullptr);
// Row permutation of each factorization: perm[t] = source row of position t.
int perm_batched(wm_ctx *ctx, int n, const int *ipiv, int64_t strideP, int *perm, int64_t strideQ, int batch,
                 hipStream_t s, const int *d_bidx = nullptr);
// Triangular solves with the LU factors and a precomputed permutation; B complex128.
// d_aidx (optional): member i solves with the factors and permutation of
// matrix d_aidx[i] (e.g. another frequency's) on the right-hand sides of
// matrix d_bidx[i].
int getrs_perm_batched(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                       const int *perm, int64_t strideQ, void *B, int ldb, int64_t strideB, int batch,
                       hipStream_t s, const int *d_bidx = nullptr, const int *d_aidx = nullptr);
// Solves with shared factors: column r < nr of item items[q] (at V + items[q] *
// strideV + r * n) is solved in place with the factors and permutation of
// index fac[q].  The columns of items that share a factor are the right-hand
// sides of one launch member (up to max_w per member), so each factor is read
// once per member rather than once per column; a column's arithmetic is the
// same whatever it is grouped with.  A triangular solve's time per 64-row
// block grows with its right-hand sides while its chain of blocks is
// latency-bound, so members of a few columns each run faster than one wide
// member (kGroupCols).
// With dinv (the inverted diagonal blocks of diag_inv_batched, factor f in slot
// (*slot_of)[f]) each diagonal block is applied as a product instead of a
// substitution: not backward stable like the substitution, for preconditioner
// applications only.
constexpr int kGroupCols = 4;
int getrs_grouped(wm_ctx *ctx, int prec, int n, int nr, const void *F, int lda, int64_t strideF, const int *perm,
                  int64_t strideQ, double2 *V, int64_t strideV, const std::vector<int> &items,
                  const std::vector<int> &fac, hipStream_t s, const double2 *dinv = nullptr,
                  const std::vector<int> *slot_of = nullptr, int max_w = kGroupCols);
// Inverses of the 64 x 64 diagonal blocks of the factors of index factors[z]
// into slot z of Dinv (diag_inv_bytes(n, factors.size()) bytes), FP64.
size_t diag_inv_bytes(int n, int nfac);
int diag_inv_batched(wm_ctx *ctx, int prec, int n, const void *F, int lda, int64_t strideF,
                     const std::vector<int> &factors, double2 *Dinv, hipStream_t s);
// Solves A^H X = B with the same factors and permutation (A = P^T L U, so
// X = P^T L^-H U^-H B); used by the condition estimate.
int getrsh_perm_batched(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                        const int *perm, int64_t strideQ, void *B, int ldb, int64_t strideB, int batch,
                        hipStream_t s, const int *d_bidx = nullptr);
// Same, computing the permutation from ipiv first.
int getrs_batched(wm_ctx *ctx, int prec, int n, int nrhs, const void *A, int lda, int64_t strideA,
                  const int *ipiv, int64_t strideP, void *B, int ldb, int64_t strideB, int batch,
                  hipStream_t s);
// R = B - Z X for complex128 Z (n x n), X, B (n x nrhs); batched.  B null: R = Z X.
// znorm (optional): ||Z||_inf of each matrix over its leading np rows, from the
// same pass over Z, atomically maxed into znorm[matrix] as ordered bit
// patterns (the caller zeroes it).
int residual_batched(wm_ctx *ctx, int n, int nrhs, const void *Z, int ldz, int64_t strideZ,
                     const void *X, const void *B, int ldb, int64_t strideB, void *R, int batch,
                     hipStream_t s, const int *d_bidx = nullptr, unsigned long long *znorm = nullptr,
                     int np = 0);

int check_device_error(wm_ctx *ctx, hipStream_t s);

// Device address of the allocation-free solve's scratch array on `device`
// (must be current), taken once per device.
int noalloc_init(int device);

// Captured factorizations of a context (see getrf_batched).
void drop_graphs(wm_ctx *ctx);
hipGraphExec_t find_graph(wm_ctx *ctx, const GraphKey &key);
void keep_graph(wm_ctx *ctx, const GraphKey &key, hipGraphExec_t exec);

// Profiler: bracket one launch of class k on stream s with events and record
// its algorithmic flops and bytes.  No-ops unless ctx->prof.on.
void prof_begin(wm_ctx *ctx, int k, hipStream_t s);
void prof_end(wm_ctx *ctx, int k, hipStream_t s, double flops, double bytes);

// Copies between device memory and pageable host memory.  Both wait for the
// stream first and copy synchronously, so they are ordered after every kernel
// already queued on s and complete on return.
int copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
int copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);
int copy2d_h2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height, hipStream_t s);
int copy2d_d2h(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height, hipStream_t s);

// One (geometry, frequency, feeds) solve.  Items of one solve_items call
// share the padded size ld; outputs are host pointers (may be null).
struct WorkItem {
    const wm_geom *g = nullptr;
    float freq = 0.f;          // MHz
    int nrhs = 0;
    const int *gaps = nullptr; // 1-based feed bases, nrhs of them
    float ampl = 1.f, phas = 0.f;
    wm_c128 *I = nullptr;      // [nrhs][npls]
    wm_c128 *zin = nullptr;    // [nrhs], or [nsrc] with generators
    const wm_source *src = nullptr;  // generators (null: one delta gap per column, gaps/ampl/phas)
    int nsrc = 0;
    const wm_load *loads = nullptr;  // lumped series loads on Z's diagonal
    int nload = 0;
    const wm_plane_wave *pw = nullptr;  // incident plane waves added to the right-hand sides
    int npw = 0;
    wm_freq_stats *st = nullptr;
    int fprec = -1;            // mixed precision: factor precision chosen by routing (-1: the batch's);
                               // out: the precision of the factors that produced the result
    double rate = -1.0;        // out: last refinement contraction rate (inf: complex64 factor singular)
    double rate_b3 = -1.0;     // out: contraction rate of a failed split-BF16x3 attempt
    int anchor = -1;           // recycled sweep: batch index of the item whose FP64 factors precondition
                               // this one's GMRES refinement (itself: factored and solved directly)
    int anchor_used = -1;      // out: batch index of the item whose factors produced the result
};

// GMRES restart length of recycled sweeps.
constexpr int kGmresRestart = 32;

// Without an observation in one precision, split-BF16x3 factors are taken to
// contract this many times slower than FP32 factors on the same matrix.  The
// rounding of Z to complex64, common to both, dominates their backward error
// (measured: equal iteration counts within one or two on the lightning
// structure at N = 8192), so the factor is small.
constexpr double kBf16x3RateFactor = 2.0;

// Refinement contraction rate of factors in precision prec (WM_PREC_BF16X3 or
// WM_PREC_F32) predicted for frequency f on geometry g from its recorded
// observations, or -1 if there are none.  Below the structure's resonances
// the EFIE condition number grows as 1/(k dl)^2, and the contraction rate
// with it, more steeply as the rate nears 1 (measured exponent about 3 near
// the FP32 limit of the lightning structure): the nearest observation at or
// above f (the nearest below, if none) is scaled by (f_obs / f)^3;
// observations in the same precision are preferred, else the other
// precision's are converted by kBf16x3RateFactor.
double predict_rate(const wm_geom *g, float f, int prec);
// Records the refinement rates of the items of a batch on their geometries.
void record_rates(const WorkItem *items, int count);
// Routes the items of a batch to the cheapest factor precision (o.prec, then
// FP32, then FP64) whose predicted rate reaches tol within max_iter.
void route_items(const wm_solve_opts &o, WorkItem *items, int count);
// True when a batch of `count` items of size n costs less with FP64 factors
// than with complex64 factors plus refinement (small n: the factorization is
// bound by the panels' per-column latency in both precisions).
bool prefer_f64(wm_ctx *c, int prec, int n, int count);
// A run of r complex64-routed frequencies followed by FP64-routed ones: true if
// factoring the run with the FP64 batch is predicted cheaper than on its own.
bool join_f64(wm_ctx *c, int prec, int n, int r);
// Matrices per launch of the batched factorization (panel cluster residency).
int getrf_width(wm_ctx *ctx, int prec, int n);
// Panels per outer block for a factorization of size n in precision prec.
int getrf_block(const wm_ctx *ctx, int prec, int n);

// Convergence test of one item after refinement iteration `it` (0-based):
// dm[r], xm[r] = max |dx|, max |x| over right-hand side r (non-finite values
// as +inf), prev[r] the previous relative correction (1e300 before the
// first).  Updates prev, first (the first correction), rate (geometric-mean
// contraction, HUGE_VAL for a non-finite iterate) and worst (largest relative
// correction).  Returns 1 converged (worst <= tol, or stalled at the FP64
// floor), -1 failed (non-finite, or too slow to reach tol within max_iter), 0
// continue.
int refine_decide(int nrhs, const double *dm, const double *xm, double *prev, int it, double tol, int max_iter,
                  double &first, double &rate, double &worst);
// Fill, factor and solve a batch of items on stream s (synchronous).
// Returns <0 on error, >0 (first zero pivot) if a matrix was singular.
int solve_items(wm_ctx *ctx, const wm_solve_opts &o, WorkItem *items, int count, hipStream_t s);
// Device bytes one item of size ld needs in solve_items.
size_t item_bytes(int ld, int nrhs, int prec);
// Items of size ld that fit the given share of free HBM (capped at 64).
int auto_batch(wm_ctx *ctx, int ld, int nrhs, int prec, double hbm_fraction);
double wavenumber_mhz(float freq_mhz);

}  // namespace wm
