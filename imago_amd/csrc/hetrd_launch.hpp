// hetrd_launch.hpp - host interface of the cooperative tridiagonal
// reduction (k_hetrd_coop in hetrd_coop.hpp) and of the kernels around it
// in the eigensolver (eigenvector sort, compact-WY factors).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>

namespace imago {

// largest orders: 32 column groups x 32 complex / 64 real registers a lane
constexpr int kTrdMaxOrderComplex = 1024;
constexpr int kTrdMaxOrderReal = 2048;
constexpr int kTrdBlock = 64;   // reflectors per compact-WY block of the back-transform
constexpr int kPotrfBlock = 32;  // columns per diagonal block of the blocked Cholesky

// Buffers of one batch; T = double2 (complex Hermitian) or double (real
// symmetric) for every pointer typed void here.
struct TrdLaunch {
    int n, nb;              // order, number of matrices
    bool real;              // scalar type double instead of double2
    const void *A;          // [nb][n * n] full Hermitian / symmetric, column-major
    void *V;                // [nb][n * n] reflectors (LAPACK 'L' layout), zeroed by the caller
    double *d, *e;          // [nb][n]
    void *tau;              // [nb][n]
    void *rbuf, *pbuf;      // [nb][n][n] exchanged pivot rows and p, a slot per step
    void *dbuf;             // [nb][n][128] exchanged dot partials, a slot per step
    unsigned *abortw;       // [1] set when a wait gives up
    int *err;               // [1], zeroed by the caller
    hipStream_t stream;
};

bool hetrd_coop_supported(int n, bool real);
// The unblocked Cholesky (upper, A = U^H U) of the diagonal block
// A(j:j+jb, j:j+jb) of each of nb matrices (n x n, column-major, stride
// n*n), jb <= kPotrfBlock; info[m] as zpotrf (the first leading minor that
// is not positive definite, 1-based), cleared by the launch with j = 0.
hipError_t launch_potf2_diag(void *A, int n, int nb, int j, int jb, bool real, int *info,
                             hipStream_t stream);
// A = Q T Q^H for every matrix of the batch; matrices are processed in
// groups whose work-groups are co-resident on the device.
hipError_t launch_hetrd_coop(const TrdLaunch &L);
// Eigenvectors of nb tridiagonal problems solved as one block-diagonal
// problem of order N = nb * n (eigenvalues ascending over all blocks, Z of
// order N): owner[q] = the block holding eigenvector q (its largest
// entry), then for every block its eigenvalues in order into w[b0 + b][0..n)
// and its vectors into X[b0 + b] (n x n, column-major; complex with zero
// imaginary parts, or real when real is set).  Bit 2 of *err is set when a
// block does not own exactly n eigenvectors.
hipError_t launch_trd_unmix(const double *Z, const double *wall, int n, int nb, int b0,
                            int *owner, int *rank, double *w, void *X, bool real, int *err,
                            hipStream_t stream);
// Compact-WY factors of the reflectors of every matrix: T[m][b] (kTrdBlock
// x kTrdBlock, upper, column-major) with H(j0) ... H(j0 + nbk - 1) =
// I - V_b T V_b^H for block b (j0 = kTrdBlock b), zlarft / dlarft forward,
// columnwise; T holds G = V_b^H V_b on entry (the strict upper triangle is
// read) and T on exit.
hipError_t launch_trd_larft(const void *tau, int n, int nb, bool real, void *T, hipStream_t stream);

// The UT form of all n - 1 reflectors of every matrix, Q = I - V T V^H with
// T^-1 = diag(1 / tau) + strict upper triangle of V^H V: launch_ut_null
// zeroes the columns of V whose tau is 0 (identity reflectors), and after
// the caller has formed G = V^H V (order n - 1, leading dimension n, stride
// n * n), launch_ut_tinv turns G into T^-1 in place (zero below the
// diagonal, 1 / tau on it, 1 where tau = 0).
hipError_t launch_ut_null(const void *tau, int n, int nb, bool real, void *V, hipStream_t stream);
hipError_t launch_ut_tinv(const void *tau, int n, int nb, bool real, void *G, hipStream_t stream);

// Per-matrix eigensolver codes from the divide and conquer of one half of
// a batch.  The half's nh matrices of order n were solved as one
// block-diagonal problem of order N = nh * n; a nonzero dstedc info names
// the failing submatrix as rows info / (N + 1) .. mod(info, N + 1) (1-based,
// LAPACK dstedc).  Every matrix of the half that overlaps those rows gets n
// (zheevd's range 1..n for "did not converge"), the others keep 0; an info
// that names no valid row range marks the whole half.  kinfo points at the
// half's first matrix.
inline void map_stedc_info(int n, int nh, int info, int *kinfo)
{
    for (int m = 0; m < nh; ++m) kinfo[m] = 0;
    if (info == 0 || nh <= 0 || n <= 0) return;
    const long N = (long)n * nh;
    long r0 = (long)info / (N + 1), r1 = (long)info % (N + 1);
    if (info < 0 || r0 < 1 || r1 < r0 || r1 > N) {
        r0 = 1;
        r1 = N;
    }
    for (long m = (r0 - 1) / n; m <= (r1 - 1) / n; ++m) kinfo[m] = n;
}

}  // namespace imago
