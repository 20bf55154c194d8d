// xcmesh.hip - exchange-correlation mesh of the potential sites
// (exchCorr.f90): the ray lengths of getECMeshParameters (:63-342) and the
// operator of every ray point of makeECMeshAndOverlap (:345-875), plus the
// pieces of its overlap sum exchCorrOverlap = sum_p w_p op_p op_p^T that
// this library does on the device (row compaction, gather, scatter-add;
// the rank-k products are a rocBLAS strided-batched dgemm).
#include <hip/hip_runtime.h>

#include "geom_kernel.hpp"
#include "xcmesh_launch.hpp"

namespace imago {

namespace {

constexpr double kXcSmall = 1.0e-8;   // smallThresh (constants.f90:49)

// One thread per (site i, vector j): the distance along the vector from
// the site to the nearest of the planes that cut the segments to its
// neighbours' images at the fraction cov_i / (cov_i + cov_k), capped at
// sqrt(rSampleOut / minPotAlpha) (exchCorr.f90:221-248).  The images are
// the first min(64, numCells) cells.
__global__ __launch_bounds__(256) void k_xc_maxrad(XcArgs X, double *__restrict__ maxrad)
{
#pragma clang fp contract(off)
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)X.nsites * X.nvec) return;
    const int i = (int)(g / X.nvec), j = (int)(g % X.nvec);
    const double ci = X.site_cov[i];
    if (ci < kXcSmall) {
        maxrad[g] = 0.0;
        return;
    }
    const double v0 = X.vec[3 * j], v1 = X.vec[3 * j + 1], v2 = X.vec[3 * j + 2];
    const double *pi = X.pls + 3 * i;
    const double x10 = pi[0] + X.cell_dims[0], x11 = pi[1] + X.cell_dims[1],
                 x12 = pi[2] + X.cell_dims[2];
    double cmax = X.max_radius;
    for (int k = 0; k < X.nsites; ++k) {
        const double ck = X.site_cov[k];
        if (ck < kXcSmall) continue;
        const double f = ci / (ci + ck);
        const double *pk = X.pls + 3 * k;
        for (int l = 0; l < X.nreps; ++l) {
            const double *R = X.cell_dims + 3 * l;
            const double s0 = f * ((pk[0] + R[0]) - x10), s1 = f * ((pk[1] + R[1]) - x11),
                         s2 = f * ((pk[2] + R[2]) - x12);
            const double dot = s0 * v0 + s1 * v1 + s2 * v2;
            if (dot <= kXcSmall) continue;
            const double q = (s0 * s0 + s1 * s1 + s2 * s2) / dot;
            if (q < cmax) cmax = q;
        }
    }
    maxrad[g] = cmax;
}

// One thread per (ray point p, potential type t): the sites of type t in
// order, their cells within (sqrt(th / alpha_1) + |offset|)^2 and within
// th / alpha_1 of the point, the type's alphas up to the first with
// alpha r^2 > th (exchCorr.f90:587-760).  Each thread owns its terms of
// its point, so the sums keep the reference's order.  GGA: the
// reference's recurrences op_o = alpha op_o + d_o e (o = x, y, z) and
// alpha^2 op_o + d_o e (o = xx .. zz), as written.
template <bool GGA>
__global__ __launch_bounds__(256) void k_xc_op(XcArgs X, const double *__restrict__ pts,
                                               int npts, int ld, double *__restrict__ op,
                                               unsigned char *__restrict__ nz,
                                               const int *__restrict__ pt_slot)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)npts * X.npt) return;
    const int t = (int)(g % X.npt), p = (int)(g / X.npt);
    const int na = X.type_nalpha[t];
    const double *a = X.type_alphas + (size_t)t * X.maxpa;
    const double negl = X.th / a[0];
    const int term0 = X.type_term0[t];
    const long P = X.pot_dim;
    const double pt0 = pts[3 * p], pt1 = pts[3 * p + 1], pt2 = pts[3 * p + 2];
    unsigned long long touched = 0;
    double *o = op + (size_t)p * P + term0;
    const size_t ostride = (size_t)ld * P;
    const int l0 = X.type_first[t], l1 = l0 + X.type_mult[t];
    for (int l = l0; l < l1; ++l) {
        double off[3], osq, negl_plus;
        {
#pragma clang fp contract(off)
            const double *pl = X.site_pos + 3 * l;
            const double sv[3] = {pt0 - pl[0], pt1 - pl[1], pt2 - pl[2]};
            double L[3];
            find_lattice_vector(X.rv, X.inv, X.cell_dims, X.cell_sizes, X.ncells, sv, L);
            for (int x = 0; x < 3; ++x) off[x] = sv[x] - L[x];
            osq = off[0] * off[0] + off[1] * off[1] + off[2] * off[2];
            negl_plus = negl + osq + 2.0 * sqrt(negl * osq);
        }
        for (int m = 0; m < X.ncells; ++m) {
            if (X.cell_sizes[m] > negl_plus) break;
            double dx, dy, dz, rad;
            {
#pragma clang fp contract(off)
                const double *R = X.cell_dims + 3 * m;
                dx = off[0] - R[0];
                dy = off[1] - R[1];
                dz = off[2] - R[2];
                rad = dx * dx + dy * dy + dz * dz;
            }
            if (rad > negl) continue;
            for (int n = 0; n < na; ++n) {
                double am;
                {
#pragma clang fp contract(off)
                    am = a[n] * rad;
                }
                if (am > X.th) break;
                const double e = exp(-am);
                o[n] = o[n] + e;
                touched |= 1ull << n;
                if (GGA) {
                    const double dd[9] = {dx, dy, dz, dx * dx, dx * dy, dx * dz,
                                          dy * dy, dy * dz, dz * dz};
#pragma unroll
                    for (int q = 1; q < 10; ++q) {
                        double *oq = o + q * ostride + n;
                        const double f = q < 4 ? a[n] : a[n] * a[n];
                        *oq = f * *oq + dd[q - 1] * e;
                    }
                }
            }
        }
    }
    if (nz) {
        unsigned char *fl = nz + (size_t)pt_slot[p] * P + term0;
        while (touched) {
            const int n = __ffsll((long long)touched) - 1;
            touched &= touched - 1;
            fl[n] = 1;
        }
    }
}

// One work-group of 1024 threads per flag row: an exclusive scan of the
// flags, chunk by chunk.
__global__ __launch_bounds__(1024) void k_xc_compact(const unsigned char *__restrict__ nz,
                                                     int pot_dim, int *__restrict__ rows,
                                                     int *__restrict__ count)
{
    __shared__ int sc[1024];
    const int s = blockIdx.x, tid = threadIdx.x;
    const unsigned char *f = nz + (size_t)s * pot_dim;
    int *r = rows + (size_t)s * pot_dim;
    int base = 0;
    for (int c0 = 0; c0 < pot_dim; c0 += 1024) {
        const int c = c0 + tid;
        const int v = c < pot_dim && f[c] ? 1 : 0;
        sc[tid] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int x = tid >= d ? sc[tid - d] : 0;
            __syncthreads();
            sc[tid] += x;
            __syncthreads();
        }
        if (v) r[base + sc[tid] - 1] = c;
        const int tot = sc[1023];
        __syncthreads();
        base += tot;
    }
    if (tid == 0) count[s] = base;
}

__global__ __launch_bounds__(256) void k_xc_gather(const double *__restrict__ op,
                                                   const double *__restrict__ w, int pot_dim,
                                                   const int *__restrict__ rows, int m, int npts,
                                                   int lda, double *__restrict__ A)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)m * npts) return;
    const int c = (int)(g % m), p = (int)(g / m);
    A[(size_t)p * lda + c] = op[(size_t)p * pot_dim + rows[c]] * sqrt(w[p]);
}

__global__ __launch_bounds__(256) void k_xc_scatter(const double *__restrict__ Cs, int ldc,
                                                    const int *__restrict__ rows, int m,
                                                    int pot_dim, double *__restrict__ C)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)m * m) return;
    const int c1 = (int)(g % m), c2 = (int)(g / m);
    if (c1 > c2) return;
    double *at = C + (size_t)rows[c2] * pot_dim + rows[c1];
    *at = *at + Cs[(size_t)c2 * ldc + c1];
}

inline unsigned xc_blocks(size_t n) { return (unsigned)((n + 255) / 256); }

inline bool xc_grid_ok(size_t n) { return n / 256 < 0x7fffffffUL; }

}  // namespace

hipError_t launch_xc_maxrad(const XcArgs &X, double *maxrad, hipStream_t stream)
{
    const size_t n = (size_t)X.nsites * X.nvec;
    if (n == 0) return hipSuccess;
    if (!xc_grid_ok(n)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xc_maxrad, dim3(xc_blocks(n)), dim3(256), 0, stream, X, maxrad);
    return hipGetLastError();
}

hipError_t launch_xc_op(const XcArgs &X, bool gga, const double *pts, int npts, int ld,
                        double *op, unsigned char *nz, const int *pt_slot, hipStream_t stream)
{
    const size_t n = (size_t)npts * X.npt;
    if (n == 0) return hipSuccess;
    if (!xc_grid_ok(n) || X.maxpa > 64 || ld < npts) return hipErrorInvalidValue;
    if (gga)
        hipLaunchKernelGGL(k_xc_op<true>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, npts,
                           ld, op, nz, pt_slot);
    else
        hipLaunchKernelGGL(k_xc_op<false>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, npts,
                           ld, op, nz, pt_slot);
    return hipGetLastError();
}

hipError_t launch_xc_compact(const unsigned char *nz, int nslots, int pot_dim, int *rows,
                             int *count, hipStream_t stream)
{
    if (nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xc_compact, dim3((unsigned)nslots), dim3(1024), 0, stream, nz, pot_dim,
                       rows, count);
    return hipGetLastError();
}

hipError_t launch_xc_gather(const double *op, const double *w, int pot_dim, const int *rows,
                            int m, int npts, int lda, double *A, hipStream_t stream)
{
    const size_t n = (size_t)m * npts;
    if (n == 0) return hipSuccess;
    if (!xc_grid_ok(n) || lda < m) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xc_gather, dim3(xc_blocks(n)), dim3(256), 0, stream, op, w, pot_dim, rows,
                       m, npts, lda, A);
    return hipGetLastError();
}

hipError_t launch_xc_scatter(const double *Cs, int ldc, const int *rows, int m, int pot_dim,
                             double *C, hipStream_t stream)
{
    const size_t n = (size_t)m * m;
    if (n == 0) return hipSuccess;
    if (!xc_grid_ok(n) || ldc < m) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xc_scatter, dim3(xc_blocks(n)), dim3(256), 0, stream, Cs, ldc, rows, m,
                       pot_dim, C);
    return hipGetLastError();
}

}  // namespace imago
