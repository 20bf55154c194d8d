// gemm_mfma_launch.hpp - host interface of the batched FP64 MFMA GEMM
// (gemm_mfma.hip) used around the eigensolve: the Cholesky transform
// C = U^-H H U^-1, the solve X = U^-1 Y of the finished eigenvector columns
// and the density W W^H.
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

// Zero pattern of op(A) (amask) and op(B) (bmask): entries in the zero
// triangle are never read (the memory there may hold anything) and the
// K loop of a tile covers only the range where both can be non-zero.
enum GemmTri : int {
    kTriNone = 0,
    kTriUpper = 1,   // op(X)(r, c) = 0 for r > c
    kTriLower = 2,   // op(X)(r, c) = 0 for r < c
};

// C_b = alpha op(A_b) op(B_b) + beta C_b for every batch entry b, column-
// major, op(X) = X or X^H (X^T for real data).  Entry b = (outer * nb2 +
// inner) * nb3 + chunk takes its matrices at X + outer * sx1 + inner * sx2
// (in elements), and C at C + outer * sc1 + inner * sc2 + chunk * sc3.
// herm: C is Hermitian (symmetric): only its lower 64 x 64 tiles are formed
// and each element is written to (i, j) and, conjugated, to (j, i), with
// the diagonal made real; beta must be 0 then and M == N.
struct GemmArgs {
    int M = 0, N = 0, K = 0;
    int lda = 0, ldb = 0, ldc = 0;
    int batch = 1, nb2 = 1;
    long sa1 = 0, sa2 = 0, sb1 = 0, sb2 = 0, sc1 = 0, sc2 = 0;
    const void *A = nullptr, *B = nullptr;
    void *C = nullptr;
    double alpha_re = 1.0, alpha_im = 0.0, beta_re = 0.0, beta_im = 0.0;
    int amask = kTriNone, bmask = kTriNone;
    bool herm = false;
    // the K range of entry (outer, inner, chunk) starts at inner * koff2
    // (the operands are zero before it) plus, with klen2 > 0, chunk * klen2
    // and holds at most klen2 values (a split of K, each part into its own
    // C, to be summed by the caller); koff2 and klen2 multiples of 16 then
    int koff2 = 0;
    int klen2 = 0;
    int nb3 = 1;
    long sc3 = 0;
};

// complex: double2 elements (interleaved re, im), else double; conj_a /
// conj_b select op(A) = A^H / op(B) = B^H.
hipError_t launch_gemm_mfma(const GemmArgs &g, bool complex, bool conj_a, bool conj_b,
                            hipStream_t stream);

}  // namespace imago
