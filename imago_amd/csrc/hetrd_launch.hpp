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
    double2 *vbuf, *pbuf;   // [nb][2][n] exchanged pivot rows and p
    double2 *dbuf;          // [nb][2][64] exchanged dot partials
    unsigned *bar;          // trd_flag_words(n, nb) barrier flags
    int *err;               // [1], zeroed by the caller
    hipStream_t stream;
};

constexpr int kTrdBlock = 32;   // reflectors per compact-WY block of the back-transform

bool hetrd_coop_supported(int n);
// barrier words the launch needs: one 64-byte flag per work-group, an abort word
inline size_t trd_flag_words(int n, int nb) { return (size_t)nb * ((n + 15) / 16) * 16 + 1; }
// Eigenvectors of nb tridiagonal problems solved as one block-diagonal
// problem of order N = nb * n (eigenvalues ascending over all blocks, Z of
// order N): owner[q] = the block holding eigenvector q (its largest
// entry), then for every block its eigenvalues in order into w[b0 + b][0..n)
// and its vectors, as complex, into X[b0 + b] (n x n, column-major).
hipError_t launch_trd_unmix(const double *Z, const double *wall, int n, int nb, int b0,
                            int *owner, int *rank, double *w, double2 *X, hipStream_t stream);
// Compact-WY factors of the reflectors of every matrix: T[m][b] (kTrdBlock
// x kTrdBlock, upper, column-major) with H(j0) ... H(j0 + nbk - 1) =
// I - V_b T V_b^H for block b (j0 = kTrdBlock b), zlarft forward/columnwise.
hipError_t launch_trd_larft(const double2 *V, const double2 *tau, int n, int nb, double2 *T,
                            hipStream_t stream);
// A = Q T Q^H for every matrix of the batch; matrices are processed in
// groups whose work-groups are co-resident on the device.
hipError_t launch_hetrd_coop(const TrdLaunch &L);

}  // namespace imago
