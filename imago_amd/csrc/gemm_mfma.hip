// gemm_mfma.hip - batched FP64 GEMM on the matrix cores
// (v_mfma_f64_16x16x4f64) for the dense products around the eigensolve.
//
// One 64 x 64 tile of C per work-group of four waves; each wave owns a
// 32 x 32 quarter as 2 x 2 blocks of 16 x 16 accumulators.  op(A) and
// op(B) are staged through LDS in K steps of 16, split into real and
// imaginary planes, in two stages: the next step's global loads are in
// flight during the current step's MFMAs and land in the other stage.  A complex product takes four real MFMAs per block
// (Re += ar br - ai bi, Im += ar bi + ai br); the eight accumulators of a
// wave are independent, so consecutive MFMAs never wait on each other.
//
// The MFMA forms the transposed tile, D = op(B)^T op(A)^T: its result sits
// in lane l, register r at row l / 16 + 4 r, column l % 16 of D, which is
// element (i0 + l % 16, j0 + l / 16 + 4 r) of C, so the 16 lanes of a group
// write 16 consecutive rows of a column.
//
// Triangular operands (amask / bmask) shorten each tile's K range to where
// both factors can be non-zero and read zeros for the other triangle; a
// Hermitian C (herm) is formed on its lower tiles and mirrored.
#include "gemm_mfma_launch.hpp"

namespace imago {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kGT = 64;    // tile edge
constexpr int kGK = 16;    // K per LDS stage
constexpr int kGLd = 65;   // LDS row (odd: the column-wise stores spread over the banks)

template <bool CX>
struct Elem;
template <>
struct Elem<true> {
    using T = double2;
    static __device__ __forceinline__ T zero() { return make_double2(0.0, 0.0); }
    static __device__ __forceinline__ T conj(T x) { return make_double2(x.x, -x.y); }
    static __device__ __forceinline__ double re(T x) { return x.x; }
    static __device__ __forceinline__ double im(T x) { return x.y; }
};
template <>
struct Elem<false> {
    using T = double;
    static __device__ __forceinline__ T zero() { return 0.0; }
    static __device__ __forceinline__ T conj(T x) { return x; }
    static __device__ __forceinline__ double re(T x) { return x; }
    static __device__ __forceinline__ double im(T) { return 0.0; }
};

struct TileRange {
    int i0, j0, klo, khi;
    bool ok;
};

__device__ __forceinline__ TileRange tile_of(const GemmArgs &g, int idx, int inner, int chunk)
{
    const int tm = (g.M + kGT - 1) / kGT, tn = (g.N + kGT - 1) / kGT;
    // Tiles whose K range is longest are handed out first (the work-groups
    // are dispatched in index order), so the short ones fill the tail.
    int ti, tj;
    if (g.herm) {
        if (g.amask == kTriLower) {
            // K grows with the row: lower tiles from the bottom row up, row
            // tm - 1 - r holding tiles (., 0 .. tm - 1 - r)
            int r = 0, s0 = 0;
            while (s0 + (tm - r) <= idx) {
                s0 += tm - r;
                ++r;
            }
            ti = tm - 1 - r;
            tj = idx - s0;
        } else {
            // lower tiles in row order: row ti holds tiles (ti, 0 .. ti)
            ti = 0;
            while ((ti + 1) * (ti + 2) / 2 <= idx) ++ti;
            tj = idx - ti * (ti + 1) / 2;
        }
    } else {
        ti = idx % tm;
        tj = idx / tm;
        if (g.bmask == kTriUpper) tj = tn - 1 - tj;   // K grows with the column
        if (g.amask == kTriLower) ti = tm - 1 - ti;   // K grows with the row
    }
    TileRange t;
    t.i0 = ti * kGT;
    t.j0 = tj * kGT;
    t.ok = ti < tm && tj < tn;
    int klo = inner * g.koff2, khi = g.K;
    if (g.klen2 > 0) {
        klo += chunk * g.klen2;
        khi = min(khi, klo + g.klen2);
    }
    if (g.amask == kTriUpper) klo = max(klo, t.i0);
    if (g.amask == kTriLower) khi = min(khi, t.i0 + kGT);
    if (g.bmask == kTriUpper) khi = min(khi, t.j0 + kGT);
    if (g.bmask == kTriLower) klo = max(klo, t.j0);
    t.klo = klo & ~(kGK - 1);
    t.khi = khi;
    return t;
}

template <bool CX, bool CA, bool CB>
__global__ __launch_bounds__(256) void k_gemm_mfma(GemmArgs g)
{
    using E = Elem<CX>;
    using T = typename E::T;
    // two stages: the MFMAs read one while the next K step is written to the other
    __shared__ double As[2][CX ? 2 : 1][kGK][kGLd];
    __shared__ double Bs[2][CX ? 2 : 1][kGK][kGLd];
    const int chunk = blockIdx.y % g.nb3, bi = blockIdx.y / g.nb3;
    const int outer = bi / g.nb2, inner = bi % g.nb2;
    const TileRange tr = tile_of(g, blockIdx.x, inner, chunk);
    if (!tr.ok) return;
    const T *A = static_cast<const T *>(g.A) + outer * g.sa1 + inner * g.sa2;
    const T *B = static_cast<const T *>(g.B) + outer * g.sb1 + inner * g.sb2;
    T *C = static_cast<T *>(g.C) + outer * g.sc1 + inner * g.sc2 + chunk * g.sc3;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wi = wv & 1, wj = wv >> 1;
    const int i0 = tr.i0, j0 = tr.j0;

    // staging: thread tid loads 4 elements of op(A) (64 x 16) and 4 of
    // op(B) (16 x 64); the index order follows the contiguous direction of
    // each operand in memory
    T ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int i, k;
            if (!CA) {
                i = tid & 63;
                k = (tid >> 6) + 4 * q;
            } else {
                k = tid & 15;
                i = (tid >> 4) + 16 * q;
            }
            const int gi = i0 + i, gk = k0 + k;
            bool ok = gi < g.M && gk < tr.khi;
            if (g.amask == kTriUpper) ok = ok && gk >= gi;
            if (g.amask == kTriLower) ok = ok && gk <= gi;
            T v = E::zero();
            if (ok) v = CA ? E::conj(A[(size_t)gi * g.lda + gk]) : A[(size_t)gk * g.lda + gi];
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int j, k;
            if (!CB) {
                k = tid & 15;
                j = (tid >> 4) + 16 * q;
            } else {
                j = tid & 63;
                k = (tid >> 6) + 4 * q;
            }
            const int gj = j0 + j, gk = k0 + k;
            bool ok = gj < g.N && gk < tr.khi;
            if (g.bmask == kTriUpper) ok = ok && gk <= gj;
            if (g.bmask == kTriLower) ok = ok && gk >= gj;
            T v = E::zero();
            if (ok) v = CB ? E::conj(B[(size_t)gk * g.ldb + gj]) : B[(size_t)gj * g.ldb + gk];
            rb[q] = v;
        }
    };
    auto stash = [&](int st) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int i, k;
            if (!CA) {
                i = tid & 63;
                k = (tid >> 6) + 4 * q;
            } else {
                k = tid & 15;
                i = (tid >> 4) + 16 * q;
            }
            As[st][0][k][i] = E::re(ra[q]);
            if (CX) As[st][CX ? 1 : 0][k][i] = E::im(ra[q]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int j, k;
            if (!CB) {
                k = tid & 15;
                j = (tid >> 4) + 16 * q;
            } else {
                j = tid & 63;
                k = (tid >> 6) + 4 * q;
            }
            Bs[st][0][k][j] = E::re(rb[q]);
            if (CX) Bs[st][CX ? 1 : 0][k][j] = E::im(rb[q]);
        }
    };

    dbl4 accR[2][2], accI[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            accR[p][q] = dbl4{0.0, 0.0, 0.0, 0.0};
            accI[p][q] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
    const int l16 = lane & 15;
    if (tr.klo < tr.khi) {
        load(tr.klo);
        stash(0);
    }
    __syncthreads();
    int cur = 0;
    for (int k0 = tr.klo; k0 < tr.khi; k0 += kGK) {
        const bool more = k0 + kGK < tr.khi;
        if (more) load(k0 + kGK);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int kk = 4 * s4 + (lane >> 4);
            double ar[2], br[2];
#pragma unroll
            for (int p = 0; p < 2; ++p) ar[p] = As[cur][0][kk][wi * 32 + 16 * p + l16];
#pragma unroll
            for (int q = 0; q < 2; ++q) br[q] = Bs[cur][0][kk][wj * 32 + 16 * q + l16];
            if (!CX) {
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        accR[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[q], ar[p], accR[p][q],
                                                                          0, 0, 0);
            } else {
                double ai[2], bi_[2], nbi[2];
#pragma unroll
                for (int p = 0; p < 2; ++p) ai[p] = As[cur][CX ? 1 : 0][kk][wi * 32 + 16 * p + l16];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    bi_[q] = Bs[cur][CX ? 1 : 0][kk][wj * 32 + 16 * q + l16];
                    nbi[q] = -bi_[q];
                }
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        accR[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[q], ar[p], accR[p][q],
                                                                          0, 0, 0);
                        accI[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi_[q], ar[p],
                                                                          accI[p][q], 0, 0, 0);
                    }
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        accR[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(nbi[q], ai[p], accR[p][q],
                                                                          0, 0, 0);
                        accI[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[q], ai[p], accI[p][q],
                                                                          0, 0, 0);
                    }
            }
        }
        // the next step into the other stage; one barrier orders both the
        // reads of this stage before its rewrite two steps on and the writes
        // of the next before its reads
        if (more) stash(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }

    // epilogue: lane l, register r holds C(i0 + .. + l % 16, j0 + .. + l / 16 + 4 r)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int i = i0 + wi * 32 + 16 * p + l16;
            if (i >= g.M) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + wj * 32 + 16 * q + (lane >> 4) + 4 * r;
                if (j >= g.N) continue;
                const double xr = accR[p][q][r], xi = CX ? accI[p][q][r] : 0.0;
                double vr = g.alpha_re * xr - g.alpha_im * xi;
                double vi = g.alpha_re * xi + g.alpha_im * xr;
                if (g.herm) {
                    if (i < j) continue;
                    if (CX && i == j) vi = 0.0;
                }
                T *cp = C + (size_t)j * g.ldc + i;
                if (!g.herm && (g.beta_re != 0.0 || g.beta_im != 0.0)) {
                    const T old = *cp;
                    const double orr = E::re(old), oi = E::im(old);
                    vr += g.beta_re * orr - g.beta_im * oi;
                    vi += g.beta_re * oi + g.beta_im * orr;
                }
                T v;
                if constexpr (CX) v = make_double2(vr, vi);
                else v = vr;
                *cp = v;
                if (g.herm && i != j) C[(size_t)i * g.ldc + j] = E::conj(v);
            }
        }
}

template <bool CX, bool CA, bool CB>
hipError_t launch_t(const GemmArgs &g, hipStream_t st)
{
    const int tm = (g.M + kGT - 1) / kGT, tn = (g.N + kGT - 1) / kGT;
    const int ntile = g.herm ? tm * (tm + 1) / 2 : tm * tn;
    hipLaunchKernelGGL((k_gemm_mfma<CX, CA, CB>), dim3((unsigned)ntile, (unsigned)g.batch),
                       dim3(256), 0, st, g);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm_mfma(const GemmArgs &g, bool complex, bool conj_a, bool conj_b,
                            hipStream_t stream)
{
    if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return hipSuccess;
    if (g.nb2 <= 0 || g.batch > 65535) return hipErrorInvalidValue;
    if (g.nb3 <= 0 || g.batch % g.nb3 != 0) return hipErrorInvalidValue;
    if (g.klen2 > 0 && (g.koff2 % 16 != 0 || g.klen2 % 16 != 0)) return hipErrorInvalidValue;
    if (g.nb3 > 1 && g.klen2 <= 0) return hipErrorInvalidValue;
    if (g.herm && (g.M != g.N || g.beta_re != 0.0 || g.beta_im != 0.0))
        return hipErrorInvalidValue;
    if (complex) {
        if (!conj_a && !conj_b) return launch_t<true, false, false>(g, stream);
        if (conj_a && !conj_b) return launch_t<true, true, false>(g, stream);
        if (!conj_a && conj_b) return launch_t<true, false, true>(g, stream);
        return launch_t<true, true, true>(g, stream);
    }
    if (!conj_a && !conj_b) return launch_t<false, false, false>(g, stream);
    if (conj_a && !conj_b) return launch_t<false, true, false>(g, stream);
    if (!conj_a && conj_b) return launch_t<false, false, true>(g, stream);
    return launch_t<false, true, true>(g, stream);
}

}  // namespace imago
