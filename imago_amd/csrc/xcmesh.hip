// xcmesh.hip - exchange-correlation mesh of the potential sites
// (exchCorr.f90): the ray lengths of getECMeshParameters (:63-342) and the
// operator of every ray point of makeECMeshAndOverlap (:345-875), plus the
// pieces of its overlap sum exchCorrOverlap = sum_p w_p op_p op_p^T that
// this library does on the device (the rows each site's points can touch,
// the operator written straight into the product operands, scatter-add;
// the products themselves are a rocBLAS strided-batched dgemm).
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
    __shared__ double lreps[3 * 64];
    for (int e = threadIdx.x; e < 3 * X.nreps; e += blockDim.x) lreps[e] = X.cell_dims[e];
    __syncthreads();
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
            const double *R = lreps + 3 * l;
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

struct XcCells {
    const double *dims, *sizes;
};

// Site l's offset from a point: (p - r_l) minus findLatticeVector of it,
// its squared length, and the cell-list bound (sqrt(negl) + |offset|)^2
// (exchCorr.f90:603-612).
__device__ __forceinline__ void xc_offset(const XcArgs &X, const XcCells &C, double pt0,
                                          double pt1, double pt2, int l, double negl,
                                          double off[3], double *negl_plus)
{
#pragma clang fp contract(off)
    const double *pl = X.site_pos + 3 * l;
    const double sv[3] = {pt0 - pl[0], pt1 - pl[1], pt2 - pl[2]};
    double L[3];
    find_lattice_vector(X.rv, X.inv, C.dims, C.sizes, X.ncells, sv, L);
    for (int x = 0; x < 3; ++x) off[x] = sv[x] - L[x];
    const double osq = off[0] * off[0] + off[1] * off[1] + off[2] * off[2];
    *negl_plus = negl + osq + 2.0 * sqrt(negl * osq);
}

// One thread per (ray point p, potential type t): the sites of type t in
// order, their cells within (sqrt(th / alpha_1) + |offset|)^2 and within
// th / alpha_1 of the point, the type's alphas up to the first with
// alpha r^2 > th (exchCorr.f90:587-760).  Each thread owns its terms of
// its point and sums them in registers (NA >= the type's alphas) in the
// reference's order.
template <int NA>
__global__ __launch_bounds__(256) void k_xc_op(XcArgs X, const double *__restrict__ pts,
                                               int npts, double *__restrict__ op)
{
    const XcCells C{X.cell_dims, X.cell_sizes};
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)npts * X.npt) return;
    const int t = (int)(g % X.npt), p = (int)(g / X.npt);
    const int na = X.type_nalpha[t];
    const double *a = X.type_alphas + (size_t)t * X.maxpa;
    double al[NA], acc[NA];
#pragma unroll
    for (int n = 0; n < NA; ++n) {
        al[n] = n < na ? a[n] : 0.0;
        acc[n] = 0.0;
    }
    const double negl = X.th / al[0];
    const double pt0 = pts[3 * p], pt1 = pts[3 * p + 1], pt2 = pts[3 * p + 2];
    const int l0 = X.type_first[t], l1 = l0 + X.type_mult[t];
    for (int l = l0; l < l1; ++l) {
        double off[3], negl_plus;
        xc_offset(X, C, pt0, pt1, pt2, l, negl, off, &negl_plus);
        for (int m = 0; m < X.ncells; ++m) {
            if (C.sizes[m] > negl_plus) break;
            double rad;
            {
#pragma clang fp contract(off)
                const double *R = C.dims + 3 * m;
                const double dx = off[0] - R[0], dy = off[1] - R[1], dz = off[2] - R[2];
                rad = dx * dx + dy * dy + dz * dz;
            }
            if (rad > negl) continue;
            bool go = true;
#pragma unroll
            for (int n = 0; n < NA; ++n) {
                double am;
                {
#pragma clang fp contract(off)
                    am = al[n] * rad;
                }
                go = go && n < na && !(am > X.th);
                if (go) acc[n] = acc[n] + exp(-am);
            }
        }
    }
    double *o = op + (size_t)p * X.pot_dim + X.type_term0[t];
#pragma unroll
    for (int n = 0; n < NA; ++n)
        if (n < na) o[n] = acc[n];
}

// Rows (terms) that the points of site i can touch, a superset: every
// point lies within D_i = max_j maxrad[i][j] of pls_i, so a term (t, n) is
// 0 at all of them when alpha_n (d - D_i)^2 > th for the nearest image, at
// distance d, of every site of type t.  One thread per (batch slot, type).
__global__ __launch_bounds__(256) void k_xc_rows(XcArgs X, int s0, int nb,
                                                 const double *__restrict__ dmax,
                                                 unsigned char *__restrict__ nz)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)nb * X.npt) return;
    const int t = (int)(g % X.npt), b = (int)(g / X.npt), i = s0 + b;
    const double *pi = X.pls + 3 * i;
    double dmin = 1e300;
    const int l0 = X.type_first[t], l1 = l0 + X.type_mult[t];
    for (int l = l0; l < l1; ++l) {
        const double *pl = X.site_pos + 3 * l;
        const double v[3] = {pi[0] - pl[0], pi[1] - pl[1], pi[2] - pl[2]};
        double L[3];
        find_lattice_vector(X.rv, X.inv, X.cell_dims, X.cell_sizes, X.ncells, v, L);
        const double d0 = v[0] - L[0], d1 = v[1] - L[1], d2 = v[2] - L[2];
        dmin = fmin(dmin, sqrt(d0 * d0 + d1 * d1 + d2 * d2));
    }
    // a little slack for the rounding of the points and of the distances
    const double dd = fmax(0.0, dmin - dmax[i] * (1.0 + 1e-12) - 1e-9);
    const int na = X.type_nalpha[t];
    const double *a = X.type_alphas + (size_t)t * X.maxpa;
    unsigned char *fl = nz + (size_t)b * X.pot_dim + X.type_term0[t];
    for (int n = 0; n < na; ++n) fl[n] = a[n] * dd * dd <= X.th ? 1 : 0;
}

// rank[b][rows[b][c]] = c for c < count[b] (rank preset to -1)
__global__ __launch_bounds__(256) void k_xc_rank(const int *__restrict__ rows,
                                                 const int *__restrict__ count, int pot_dim,
                                                 int *__restrict__ rank)
{
    const int b = blockIdx.y;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= count[b]) return;
    rank[(size_t)b * pot_dim + rows[(size_t)b * pot_dim + c]] = c;
}

// The LDA operator of a batch of sites' points written straight into the
// GEMM operands: B_b(p - first point of b, rank(term)) = op * sqrt(w_p)
// (B_b ldn x lda, column-major, one block per batch slot b).  Threads run
// over the points fastest: neighbouring lanes hold neighbouring points of
// one ray and walk the same cells.  Terms outside the slot's rows must be
// 0; err is set if one is not.
template <int NA>
__global__ __launch_bounds__(256) void k_xc_op_rows(XcArgs X, const double *__restrict__ pts,
                                                    const double *__restrict__ w, int npts,
                                                    const int *__restrict__ pt_slot,
                                                    const int *__restrict__ slot_first,
                                                    const int *__restrict__ rank, int lda,
                                                    int ldn, double *__restrict__ B,
                                                    int *__restrict__ err)
{
    const XcCells C{X.cell_dims, X.cell_sizes};
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)npts * X.npt) return;
    const int p = (int)(g % npts), t = (int)(g / npts);
    const int na = X.type_nalpha[t];
    const double *a = X.type_alphas + (size_t)t * X.maxpa;
    double al[NA], acc[NA];
#pragma unroll
    for (int n = 0; n < NA; ++n) {
        al[n] = n < na ? a[n] : 0.0;
        acc[n] = 0.0;
    }
    const double negl = X.th / al[0];
    const double pt0 = pts[3 * p], pt1 = pts[3 * p + 1], pt2 = pts[3 * p + 2];
    const int l0 = X.type_first[t], l1 = l0 + X.type_mult[t];
    for (int l = l0; l < l1; ++l) {
        double off[3], negl_plus;
        xc_offset(X, C, pt0, pt1, pt2, l, negl, off, &negl_plus);
        for (int m = 0; m < X.ncells; ++m) {
            if (C.sizes[m] > negl_plus) break;
            double rad;
            {
#pragma clang fp contract(off)
                const double *R = C.dims + 3 * m;
                const double dx = off[0] - R[0], dy = off[1] - R[1], dz = off[2] - R[2];
                rad = dx * dx + dy * dy + dz * dz;
            }
            if (rad > negl) continue;
            bool go = true;
#pragma unroll
            for (int n = 0; n < NA; ++n) {
                double am;
                {
#pragma clang fp contract(off)
                    am = al[n] * rad;
                }
                go = go && n < na && !(am > X.th);
                if (go) acc[n] = acc[n] + exp(-am);
            }
        }
    }
    const int b = pt_slot[p];
    const double sw = w ? sqrt(w[p]) : 1.0;
    const int *rk = rank + (size_t)b * X.pot_dim + X.type_term0[t];
    double *Bb = B + (size_t)b * lda * ldn + (p - slot_first[b]);
    bool bad = false;
#pragma unroll
    for (int n = 0; n < NA; ++n)
        if (n < na) {
            const int c = rk[n];
            if (c >= 0)
                Bb[(size_t)c * ldn] = acc[n] * sw;
            else if (acc[n] != 0.0)
                bad = true;
        }
    if (bad) *err = 1;
}

// wignerXC / wignerXCEnergy (potentialUpdate.F90:2835-2894), Hartree:
// the potential and energy density of the Wigner interpolation.
__device__ __forceinline__ void xc_wigner(double rho, double *pot, double *en)
{
    const double r13 = pow(rho, 0.3333333333333333333);
    const double f = 1.0 + 12.57 * r13;
    *pot = -r13 * (0.984 + (0.943656 + 8.8963 * r13) / (f * f));
    *en = -0.738 * r13 * (1.0 + 0.959 / f);
}

// One thread per ray point of a batch: the total and core densities
// op . rho over the slot's rows in term order (negative values set to 0),
// then V_b(p, k) = w_p v_k for the Wigner potential and energy density of
// the total and the energy density of the core (potentialUpdate.F90:
// 855-1007).  B_b holds the unscaled operator (points x rows); V is
// [nb][3][ldn], zero beyond each slot's points.
__global__ __launch_bounds__(256) void k_xc_point_pot(const double *__restrict__ B, int lda,
                                                      int ldn, const int *__restrict__ pt_slot,
                                                      const int *__restrict__ slot_first,
                                                      const int *__restrict__ rows,
                                                      const int *__restrict__ count, int pot_dim,
                                                      const double *__restrict__ w,
                                                      const double *__restrict__ rho_total,
                                                      const double *__restrict__ rho_core,
                                                      int npts, double *__restrict__ V)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npts) return;
    const int b = pt_slot[p], m = count[b], q = p - slot_first[b];
    const double *Bb = B + (size_t)b * lda * ldn + q;
    const int *rb = rows + (size_t)b * pot_dim;
    // each operator value is read once for both sums (each sum keeps its
    // own order, term by term)
    double core = 0.0, total = 0.0;
    for (int c = 0; c < m; ++c) {
        const double op = Bb[(size_t)c * ldn];
        const int row = rb[c];
        core = core + op * rho_core[row];
        total = total + op * rho_total[row];
    }
    if (core < 0.0) core = 0.0;
    if (total < 0.0) total = 0.0;
    double pot, en;
    xc_wigner(total, &pot, &en);
    const double r13 = pow(core, 0.3333333333333333333);
    const double enc = -0.738 * r13 * (1.0 + 0.959 / (1.0 + 12.57 * r13));
    double *Vb = V + (size_t)b * 3 * ldn + q;
    Vb[0] = w[p] * pot;
    Vb[ldn] = w[p] * en;
    Vb[2 * (size_t)ldn] = w[p] * enc;
}

// pot[k][r] += part_b(rank[b][r], k) for the slots b of the batch in order
// (part_b lda x 3, column-major, one block per slot)
__global__ __launch_bounds__(256) void k_xc_pot_add(const double *__restrict__ part, int lda,
                                                    const int *__restrict__ rank, int nb,
                                                    int pot_dim, double *__restrict__ pot)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= pot_dim) return;
    double a0 = pot[r], a1 = pot[pot_dim + r], a2 = pot[2 * (size_t)pot_dim + r];
    for (int b = 0; b < nb; ++b) {
        const int c = rank[(size_t)b * pot_dim + r];
        if (c < 0) continue;
        const double *pc = part + (size_t)b * 3 * lda + c;
        a0 += pc[0];
        a1 += pc[lda];
        a2 += pc[2 * (size_t)lda];
    }
    pot[r] = a0;
    pot[pot_dim + r] = a1;
    pot[2 * (size_t)pot_dim + r] = a2;
}

// The GGA operator of one site's points: as k_xc_op for the first value,
// then the reference's recurrences op_o = alpha op_o + d_o e (o = x, y, z)
// and alpha^2 op_o + d_o e (o = xx .. zz), as written, accumulated in
// place (op zeroed by the caller; values o at stride ld * potDim).
__global__ __launch_bounds__(256) void k_xc_op_gga(XcArgs X, const double *__restrict__ pts,
                                                   int npts, int ld, double *__restrict__ op)
{
    const XcCells C{X.cell_dims, X.cell_sizes};
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)npts * X.npt) return;
    const int t = (int)(g % X.npt), p = (int)(g / X.npt);
    const int na = X.type_nalpha[t];
    const double *a = X.type_alphas + (size_t)t * X.maxpa;
    const double negl = X.th / a[0];
    const double pt0 = pts[3 * p], pt1 = pts[3 * p + 1], pt2 = pts[3 * p + 2];
    double *o = op + (size_t)p * X.pot_dim + X.type_term0[t];
    const size_t ostride = (size_t)ld * X.pot_dim;
    const int l0 = X.type_first[t], l1 = l0 + X.type_mult[t];
    for (int l = l0; l < l1; ++l) {
        double off[3], negl_plus;
        xc_offset(X, C, pt0, pt1, pt2, l, negl, off, &negl_plus);
        for (int m = 0; m < X.ncells; ++m) {
            if (C.sizes[m] > negl_plus) break;
            double dx, dy, dz, rad;
            {
#pragma clang fp contract(off)
                const double *R = C.dims + 3 * m;
                dx = off[0] - R[0];
                dy = off[1] - R[1];
                dz = off[2] - R[2];
                rad = dx * dx + dy * dy + dz * dz;
            }
            if (rad > negl) continue;
            const double dd[9] = {dx, dy, dz, dx * dx, dx * dy, dx * dz, dy * dy, dy * dz, dz * dz};
            for (int n = 0; n < na; ++n) {
                double am;
                {
#pragma clang fp contract(off)
                    am = a[n] * rad;
                }
                if (am > X.th) break;
                const double e = exp(-am);
                o[n] = o[n] + e;
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
                        double *op, hipStream_t stream)
{
    const size_t n = (size_t)npts * X.npt;
    if (n == 0) return hipSuccess;
    if (!xc_grid_ok(n) || X.maxpa > 64 || ld < npts) return hipErrorInvalidValue;
    if (gga) {
        hipLaunchKernelGGL(k_xc_op_gga, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, npts, ld,
                           op);
        return hipGetLastError();
    }
    if (X.maxpa <= 16)
        hipLaunchKernelGGL(k_xc_op<16>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, npts, op);
    else if (X.maxpa <= 32)
        hipLaunchKernelGGL(k_xc_op<32>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, npts, op);
    else
        hipLaunchKernelGGL(k_xc_op<64>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, npts, op);
    return hipGetLastError();
}

hipError_t launch_xc_rows(const XcArgs &X, int s0, int nb, const double *dmax,
                          unsigned char *nz, hipStream_t stream)
{
    const size_t n = (size_t)nb * X.npt;
    if (n == 0) return hipSuccess;
    if (!xc_grid_ok(n)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xc_rows, dim3(xc_blocks(n)), dim3(256), 0, stream, X, s0, nb, dmax, nz);
    return hipGetLastError();
}

hipError_t launch_xc_rank(const int *rows, const int *count, int nb, int pot_dim, int *rank,
                          hipStream_t stream)
{
    if (nb <= 0 || nb > 65535) return nb <= 0 ? hipSuccess : hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xc_rank, dim3((unsigned)((pot_dim + 255) / 256), (unsigned)nb), dim3(256),
                       0, stream, rows, count, pot_dim, rank);
    return hipGetLastError();
}

hipError_t launch_xc_op_rows(const XcArgs &X, const double *pts, const double *w, int npts,
                             const int *pt_slot, const int *slot_first, const int *rank, int lda,
                             int ldn, double *B, int *err, hipStream_t stream)
{
    const size_t n = (size_t)npts * X.npt;
    if (n == 0) return hipSuccess;
    if (!xc_grid_ok(n) || X.maxpa > 64) return hipErrorInvalidValue;
    if (X.maxpa <= 16)
        hipLaunchKernelGGL(k_xc_op_rows<16>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, w,
                           npts, pt_slot, slot_first, rank, lda, ldn, B, err);
    else if (X.maxpa <= 32)
        hipLaunchKernelGGL(k_xc_op_rows<32>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, w,
                           npts, pt_slot, slot_first, rank, lda, ldn, B, err);
    else
        hipLaunchKernelGGL(k_xc_op_rows<64>, dim3(xc_blocks(n)), dim3(256), 0, stream, X, pts, w,
                           npts, pt_slot, slot_first, rank, lda, ldn, B, err);
    return hipGetLastError();
}

hipError_t launch_xc_point_pot(const double *B, int lda, int ldn, const int *pt_slot,
                               const int *slot_first, const int *rows, const int *count,
                               int pot_dim, const double *w, const double *rho_total,
                               const double *rho_core, int npts, double *V, hipStream_t stream)
{
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xc_point_pot, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, stream,
                       B, lda, ldn, pt_slot, slot_first, rows, count, pot_dim, w, rho_total,
                       rho_core, npts, V);
    return hipGetLastError();
}

hipError_t launch_xc_pot_add(const double *part, int lda, const int *rank, int nb, int pot_dim,
                             double *pot, hipStream_t stream)
{
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xc_pot_add, dim3((unsigned)((pot_dim + 255) / 256)), dim3(256), 0,
                       stream, part, lda, rank, nb, pot_dim, pot);
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
