// tridiag_dc_launch.hpp - host side of the batched tridiagonal divide and
// conquer (tridiag_dc.hip): the merge plan of an order-n problem and the
// device work arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace imago {

// largest order: the deflation's sort keeps 2 n doubles in LDS
constexpr int kDcMaxOrder = 2048;

// One height of the merge tree, as offsets into DcPlan::host / dev (ints):
// merges (6 ints each: r0, n1, n2, hL, hR, h), items (2 ints: merge index,
// column x; one per eigenpair of every merge), GEMM jobs (8 ints: A offset,
// B offset, C offset, M, N, K, source buffer, 0) and 64 x 64 output tiles
// (4 ints: job, tile row, tile column, 0).
struct DcLevel {
    int h, nmerge, nitem, njob, ntile;
    size_t merge_off, item_off, job_off, tile_off;
};

struct DcPlan {
    int n = 0, H = 0;
    std::vector<DcLevel> levels;
    std::vector<int> host;
    int *dev = nullptr;   // a device copy of host, owned by the caller
};

// Device arrays of one batched solve of nb problems of order n (carved out
// of one allocation of dc_work_bytes(n, nb) bytes).
struct DcWork {
    double *Zb[3], *Vu, *wb[3], *scale;
    double *dl, *zz, *dv, *val, *rc, *rs, *tau, *zh, *rho;
    int *perm, *kpos, *dpos, *rp, *rq, *org, *rank, *K, *nrot;
};

void dc_make_plan(int n, DcPlan &plan);
size_t dc_work_bytes(int n, int nb);
void dc_work_carve(void *base, int n, int nb, DcWork &w);

// Eigenvalues (ascending, into evals[b * n + i]) and eigenvectors (column-
// major n x n per problem into X, real or complex with zero imaginary part)
// of nb symmetric tridiagonals: diagonals d[b * n + i], off-diagonals
// e[b * n + i] for i < n - 1.  plan.dev must hold plan.host.
hipError_t launch_tridiag_dc(const DcPlan &plan, int nb, const double *d, const double *e,
                             const DcWork &w, double *evals, void *X, bool real,
                             hipStream_t stream);

}   // namespace imago
