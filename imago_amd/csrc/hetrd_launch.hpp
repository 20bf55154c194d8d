// hetrd_launch.hpp - host interface of the cooperative tridiagonal
// reduction (k_hetrd_coop in hetrd_coop.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>

namespace imago {

constexpr int kTrdMaxOrder = 1024;   // 32 column groups x 32 registers

struct TrdLaunch {
    int n, nb;              // order, number of matrices
    const double2 *A;       // [nb][n * n] full Hermitian, column-major
    double2 *V;             // [nb][n * n] reflectors (zhetrd 'L' layout)
    double *d, *e;          // [nb][n]
    double2 *tau;           // [nb][n]
    double2 *vbuf, *pbuf;   // [nb][n] exchange vectors
    double2 *dbuf;          // [nb][64]
    unsigned *bar;          // [nb]
    int *err;               // [1], zeroed by the caller
    hipStream_t stream;
};

bool hetrd_coop_supported(int n);
// A = Q T Q^H for every matrix of the batch; matrices are processed in
// groups whose work-groups are co-resident on the device.
hipError_t launch_hetrd_coop(const TrdLaunch &L);

}  // namespace imago
