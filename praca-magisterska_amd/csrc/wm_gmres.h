// FP64 GMRES iterative refinement with shared (anchor) LU factors.
#pragma once

#include <vector>

#include "wm_internal.h"

namespace wm {

// Buffers of one gmres_ir call.  Item-indexed data (matrices Z, vectors B, X,
// R, D, and each Krylov set) lives at base + item * stride; factor-indexed
// data (LU factors F and permutations perm) at base + factor * stride.
struct GmresArgs {
    int n = 0, nr = 0, m = 0;      // padded size, right-hand sides per item, restart length
    int count = 0;                 // items in the batch (the Krylov sets hold count members each)
    const double2 *Z = nullptr;    // complex128 matrices of the items
    int64_t sZ = 0;
    const void *F = nullptr;       // LU factors: complex128 for fprec WM_PREC_F64, complex64 otherwise
    int fprec = WM_PREC_F64;
    int64_t sF = 0;
    const int *perm = nullptr;     // row permutation of each factor, n ints per factor
    const double2 *B = nullptr;    // right-hand sides
    double2 *X = nullptr, *R = nullptr, *D = nullptr;
    int64_t sV = 0;                // per-item stride of the vectors (n x nr, column stride n)
    double2 *K = nullptr;          // 2m + 1 Krylov sets of count * sV (gmres_bytes per item)
    double2 *red = nullptr;        // reductions: (3m + 4) * count * nr complex128
    int *dact = nullptr, *dfac = nullptr;  // index tables, count ints each
    unsigned long long *mx = nullptr;      // 2 * count * nr
    unsigned long long *mx2 = nullptr;     // count + count * nr
    int np = 0;                    // unknowns of the items (rows/columns of Z before the padding)
};

struct GmresOut {
    std::vector<int> iters, conv;  // per item: inner iterations, converged flag
    std::vector<double> dx_rel;    // per item: last relative correction
    std::vector<double> berr;      // per item: normwise backward error of the result (infinity norms)
    long inner_iters = 0;          // member-iterations over the call
    explicit GmresOut(int count) : iters(count, 0), conv(count, 0), dx_rel(count, 0.0), berr(count, 0.0) {}
};

// Device bytes of the Krylov sets of one item.
size_t gmres_bytes(int n, int nr, int m);
// Solves Z_b X_b = B_b for the items `items` (batch indices), item items[q]
// preconditioned by the factors of index fac[q]; X and R are overwritten.
// out.conv[b] = 1 for items whose refinement converged (o.tol), the others
// are left to the caller (refactor).  Returns a status.
int gmres_ir(wm_ctx *c, const wm_solve_opts &o, const GmresArgs &a, const std::vector<int> &items,
             const std::vector<int> &fac, GmresOut &out, hipStream_t s);

}  // namespace wm
