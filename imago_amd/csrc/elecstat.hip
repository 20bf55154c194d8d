// elecstat.hip - electrostatics of the potential-side Gaussians
// (elecStat.f90).
//
// neutralAndNuclearQPot (:59-664) sums, over every pair of potential sites
// and every real-space cell, Gaussian-pair overlaps and screened Coulomb
// terms into potDim x potDim matrices and two potDim vectors; residualQ
// (:665-1043) forms the Ewald real- and reciprocal-space sums of every
// potential alpha against the system's smallest one.  Every matrix element
// is owned by one thread, which adds its terms in the reference's (site i,
// site j, cell) order (the nuclear vector adds sixteen interleaved partial
// sums over the sites j, the reciprocal sums their cell chunks, in a fixed
// order); the screening comparisons are made on the same values as the
// reference's (no contraction into fused multiply-adds).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "elecstat_launch.hpp"
#include "geom_kernel.hpp"

namespace imago {

namespace {

constexpr double kEsPi = 3.141592653589793;
constexpr double kEsSmall = 1.0e-8;   // smallThresh (constants.f90:49)
constexpr int kEsMaxAlphas = 64;      // alphas of a potential type (one wavefront)

// error_fn (mathSubs.f90:22-161): the Sandia rational approximation of erf
// (SAND77-1441; W. J. Cody, Math. Comp. 23 (1969) 631) in three ranges,
// |x| <= 0.46875, <= 4, <= 6, and sign(x) beyond, with the published
// coefficients; eleven significant digits.
__device__ __forceinline__ double es_erf(double x)
{
    const double ax = fabs(x);
    if (ax > 6.0) return x / ax;
    const double x2 = ax * ax;
    if (ax <= 0.46875) {
        double a = ax * (242.6679552305318 +
                         x2 * (21.97926161829415 +
                               x2 * (6.996383488619136 + x2 * -0.03560984370181539)));
        a = a / (215.0588758698612 +
                 x2 * (91.16490540451490 + x2 * (15.08279763040779 + x2)));
        return x < 0.0 ? -a : a;
    }
    double a;
    if (ax <= 4.0) {
        a = exp(-x2) *
            (22.898992851659 +
             ax * (26.094746956075 +
                   ax * (14.571898596926 +
                         ax * (4.2677201070898 +
                               ax * (0.56437160686381 + ax * -0.0000060858151959688)))));
        a = a / (22.898985749891 +
                 ax * (51.933570687552 +
                       ax * (50.273202863803 +
                             ax * (26.288795758761 + ax * (7.5688482293618 + ax)))));
    } else {
        const double ix2 = 1.0 / x2;
        a = ix2 * (-0.0121308276389978 +
                   ix2 * (-0.1199039552681460 +
                          ix2 * (-0.243911029488626 + ix2 * -0.0324319519277746))) /
            (0.0430026643452770 + ix2 * (0.489552441961437 + ix2 * (1.43771227937118 + ix2)));
        a = exp(-x2) * (0.564189583547756 + a) / ax;
    }
    const double v = fabs(1.0 - a);
    return x < 0.0 ? -v : v;
}

// ms[i][j] = (r_i - r_j) - findLatticeVector(r_i - r_j) and |r_i - r_j|^2
// (elecStat.f90:398-420), once per site pair for all the kernels below
__global__ __launch_bounds__(256) void k_es_minsep(EsArgs A)
{
#pragma clang fp contract(off)
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)A.nsites * A.nsites) return;
    const int i = (int)(g / A.nsites), j = (int)(g % A.nsites);
    const double *pi = A.site_pos + 3 * i, *pj = A.site_pos + 3 * j;
    const double sep[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
    double L[3];
    find_lattice_vector(A.rv, A.inv, A.cell_dims, A.cell_sizes, A.ncells, sep, L);
    A.msep[g] = make_double4(sep[0] - L[0], sep[1] - L[1], sep[2] - L[2],
                             sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2]);
}

__device__ __forceinline__ void es_min_sep(const EsArgs &A, int i, int j, double ms[3],
                                           double *sq)
{
    const double4 m = A.msep[(size_t)i * A.nsites + j];
    ms[0] = m.x;
    ms[1] = m.y;
    ms[2] = m.z;
    *sq = m.w;
}

__device__ __forceinline__ double es_r2(const EsArgs &A, const double ms[3], int k)
{
#pragma clang fp contract(off)
    const double *R = A.cell_dims + 3 * k;
    const double d0 = ms[0] - R[0], d1 = ms[1] - R[1], d2 = ms[2] - R[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}


// Entry e of the L-shaped shells of an n1 x n2 alpha block: shell q holds
// (q, q), (q, q+1 .. n2-1) and (q+1 .. n1-1, q), the shells in order.
__device__ __forceinline__ void es_shell_entry(int e, int n1, int n2, int *a, int *b, int *q)
{
    int s = 0, k = 0;
    while (e >= s + (n1 + n2 - 1 - 2 * k)) {
        s += n1 + n2 - 1 - 2 * k;
        ++k;
    }
    const int o = e - s;
    if (o < n2 - k) {
        *a = k;
        *b = k + o;
    } else {
        *a = k + 1 + (o - (n2 - k));
        *b = k;
    }
    *q = k;
}

// One work-group per (type of site i, type of site j); each thread owns
// an entry (alpha a of site i, alpha b of site j), the entries taken shell
// by shell (es_shell_entry).  The reference walks the alpha pairs of a
// site pair and cell in those L-shaped shells along the diagonal and stops
// at the first diagonal pair l with reducedPotAlpha(l,l) r^2 >
// logElecThresh, except that the last shell (l = min(n1, n2)) is always
// added when the walk reaches it (elecStat.f90:425-537); so (a, b) is
// added exactly when min(a, b, min(n1, n2) - 2) (0-based) is below f, the
// first failing diagonal pair (min(n1, n2) - 1 if none fails).  For each
// block of kEsNeutCells cells the work-group first finds r^2 and f of every
// cell (one a thread); then each wavefront visits, in cell order, only the
// cells whose f exceeds its smallest shell.
constexpr int kEsNeutCells = 256;

__global__ __launch_bounds__(256) void k_es_neut(EsArgs A, double *__restrict__ ovl,
                                                 double *__restrict__ nl,
                                                 double *__restrict__ lo)
{
    __shared__ double rdiag[kEsMaxAlphas];
    __shared__ double cr2[kEsNeutCells];
    __shared__ int cf[kEsNeutCells];
    const int t1 = blockIdx.x / A.npt, t2 = blockIdx.x % A.npt;
    const int n1 = A.type_nalpha[t1], n2 = A.type_nalpha[t2];
    const int minN = min(n1, n2);
    const double *a1 = A.type_alphas + (size_t)t1 * A.maxpa;
    const double *a2 = A.type_alphas + (size_t)t2 * A.maxpa;
    if ((int)threadIdx.x < minN) {
        const double x = a1[threadIdx.x], y = a2[threadIdx.x];
        rdiag[threadIdx.x] = x * y / (x + y);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long P = A.pot_dim;
    const int col0 = A.type_term0[t1], row0 = A.type_term0[t2];
    const int i0 = A.type_first[t1], i1 = i0 + A.type_mult[t1];
    const int j0 = A.type_first[t2], j1 = j0 + A.type_mult[t2];
    const double pi3 = kEsPi * kEsPi * kEsPi;
    const double x2isp = 2.0 / sqrt(kEsPi);
    const int ne = n1 * n2;
    // uniform trip counts: every thread takes part in the barriers and ballots
    for (int e0 = 0; e0 < ne; e0 += blockDim.x) {
        const int e = e0 + threadIdx.x;
        const bool act = e < ne;
        int a = 0, b = 0, qs = kEsMaxAlphas;
        if (act) {
            es_shell_entry(e, n1, n2, &a, &b, &qs);
            qs = min(qs, minN - 2);
        }
        // entries go by shell, so lane 0 holds the wavefront's smallest
        const int qmin = __builtin_amdgcn_readfirstlane(qs);
        const double x = a1[a], y = a2[b];
        const double sx = sqrt(x), sy = sqrt(y);
        const double red = x * y / (x + y);
        const double sred = sqrt(red);
        const double nqc = pi3 / (x * sx * y * sy);
        const double poc = pow(kEsPi / (x + y), 1.5);
        double ov = 0.0, vn = 0.0, vl = 0.0;
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j) {
                double ms[3], sq;
                es_min_sep(A, i, j, ms, &sq);
                for (int k0 = 0; k0 < A.ncells; k0 += kEsNeutCells) {
                    const int nk = min(kEsNeutCells, A.ncells - k0);
                    __syncthreads();
                    if ((int)threadIdx.x < nk) {
                        const double r2 = es_r2(A, ms, k0 + threadIdx.x);
                        int f = 0;
                        while (f < minN - 1 && !(rdiag[f] * r2 > A.th)) ++f;
                        cr2[threadIdx.x] = r2;
                        cf[threadIdx.x] = f;
                    }
                    __syncthreads();
                    for (int s0 = 0; s0 < nk; s0 += 64) {
                        const int u = s0 + lane;
                        unsigned long long mask = __ballot(u < nk && cf[min(u, nk - 1)] > qmin);
                        while (mask) {
                            const int k = s0 + __ffsll((long long)mask) - 1;
                            mask &= mask - 1;
                            if (qs >= cf[k]) continue;
                            const double r2 = cr2[k];
                            const double r = sqrt(r2);
                            if (r2 >= kEsSmall)
                                vn = vn + nqc / r * (-es_erf(sx * r) + es_erf(sred * r));
                            else
                                vl = vl + nqc * x2isp * sred;
                            ov = ov + poc * exp(-red * r2);
                        }
                    }
                }
            }
        if (act) {
            const size_t at = (size_t)(col0 + a) * P + (row0 + b);
            ovl[at] = ov;
            nl[at] = vn;
            lo[at] = vl;
        }
    }
}

// One work-group per type of site i; lane m of each wavefront owns alpha
// m, and the kEsNucWaves wavefronts take every kEsNucWaves-th site j, their
// partial sums added in wavefront order at the end.  For each site pair
// and cell the reference adds alphas in order until the first whose screen
// r^2 a_m nucAlpha / (a_m + nucAlpha) reaches logElecThresh
// (elecStat.f90:541-578): the wave finds that alpha with a ballot.  The
// bare nuclear part is added for coincident sites (:583-596).
constexpr int kEsNucWaves = 16;

__global__ __launch_bounds__(64 * kEsNucWaves) void k_es_nuc(EsArgs A, double *__restrict__ nl,
                                                              double *__restrict__ lo)
{
    __shared__ double part[kEsNucWaves][2][64];
    const int t1 = blockIdx.x, m = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n1 = A.type_nalpha[t1];
    const bool act = m < n1;
    const double am = act ? A.type_alphas[(size_t)t1 * A.maxpa + m] : 1.0;
    const double z1 = A.type_z[t1];
    const double x2pi = 2.0 * kEsPi, pi32 = pow(kEsPi, 1.5);
    const int i0 = A.type_first[t1], i1 = i0 + A.type_mult[t1];
    double vn = 0.0, vl = 0.0;
    for (int i = i0; i < i1; ++i)
        for (int j = w; j < A.nsites; j += kEsNucWaves) {
            const int t2 = A.site_type[j];
            const double nuc = A.type_nuc[t2], z2 = A.type_z[t2];
            const double nf = am + nuc;
            double ms[3], sq;
            es_min_sep(A, i, j, ms, &sq);
            for (int k = 0; k < A.ncells; ++k) {
                const double r2 = es_r2(A, ms, k);
                double scr;
                {
#pragma clang fp contract(off)
                    scr = r2 * am * nuc / nf;
                }
                const unsigned long long stop = __ballot(act && !(scr < A.th));
                const int first = stop ? __ffsll((long long)stop) - 1 : 64;
                if (act && m < first) {
                    if (fabs(scr) >= kEsSmall) {
                        const double r = sqrt(r2);
                        const double cf = am * r / sqrt(nf);
                        vn = vn - z2 * pi32 * exp(-scr) * es_erf(cf) / cf / nf;
                    } else {
                        vn = vn - z2 * exp(-scr) * x2pi / nf;
                    }
                }
            }
            if (sq < kEsSmall) vl = vl - x2pi * z1 / am;
        }
    part[w][0][m] = vn;
    part[w][1][m] = vl;
    __syncthreads();
    if (w == 0 && act) {
        double sn = part[0][0][m], sl = part[0][1][m];
        for (int q = 1; q < kEsNucWaves; ++q) {
            sn += part[q][0][m];
            sl += part[q][1][m];
        }
        const int at = A.type_term0[t1] + m;
        nl[at] = sn;
        lo[at] = sl;
    }
}

// E[s][c][k] = w_c exp(-|G_c|^2 / (4 mra_k)) / |G_c|^2 with mra_k = a_k
// minPot / (a_k + minPot) for the alphas of run s (elecStat.f90:843-856);
// w_c = 2 for a cell that stands for the pair +G, -G; 0 for padding.
__global__ __launch_bounds__(256) void k_es_recip_exp(EsArgs A, EsRecip R)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)R.nsrc * R.ncr * R.nk) return;
    const int k = (int)(g % R.nk);
    const size_t q = g / R.nk;
    const int c = (int)(q % R.ncr), s = (int)(q / R.ncr);
    const int t = R.src_type[s];
    double v = 0.0;
    if (k < A.type_nalpha[t]) {
        const double a = A.type_alphas[(size_t)t * A.maxpa + k];
        const double mra = a * R.min_pot / (a + R.min_pot);
        const double inv4 = 0.25 / mra;
        v = R.cw[c] * (exp(-R.csr[c] * inv4) / R.csr[c]);
    }
    R.E[g] = v;
}

// CS[c][s] = (cos, sin)(G_c . r_s); cos(G.(r_i - r_j)) is then
// cos_i cos_j + sin_i sin_j.
__global__ __launch_bounds__(256) void k_es_sincos(EsArgs A, EsRecip R)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)R.ncr * A.nsites) return;
    const int s = (int)(g % A.nsites), c = (int)(g / A.nsites);
    const double *G = R.cdr + 3 * c, *r = A.site_pos + 3 * s;
    double sn, cs;
    sincos(G[0] * r[0] + G[1] * r[1] + G[2] * r[2], &sn, &cs);
    R.CS[g] = make_double2(cs, sn);
}

// Reciprocal sums sum_c E[k][c] cos(G_c . (r_i - r_j)) for a block of BI
// sites i of one run (wave-uniform: the tables come through scalar loads)
// against 64 sites j (one a lane) over one chunk of the cells.
template <int NK, int BI>
__global__ __launch_bounds__(64) void k_es_recip(EsArgs A, EsRecip R)
{
    const int lane = threadIdx.x;
    const int j = blockIdx.x * 64 + lane;
    const int jj = min(j, A.nsites - 1);
    const int *bk = R.blk + 3 * blockIdx.y;
    const int i0 = bk[0], ni = bk[1], s = bk[2];
    const int per = (R.ncr + R.chunks - 1) / R.chunks;
    const int c0 = (int)blockIdx.z * per, c1 = min(R.ncr, ((int)blockIdx.z + 1) * per);
    double acc[BI][NK];
#pragma unroll
    for (int b = 0; b < BI; ++b)
#pragma unroll
        for (int k = 0; k < NK; ++k) acc[b][k] = 0.0;
    const double2 *__restrict__ CS = R.CS;
    const double *__restrict__ E = R.E + (size_t)s * R.ncr * NK;
#pragma unroll 2
    for (int c = c0; c < c1; ++c) {
        const double2 cj = CS[(size_t)c * A.nsites + jj];
        const double *e = E + (size_t)c * NK;
#pragma unroll
        for (int b = 0; b < BI; ++b) {
            const double2 ci = CS[(size_t)c * A.nsites + i0 + min(b, ni - 1)];
            const double ph = ci.x * cj.x + ci.y * cj.y;
#pragma unroll
            for (int k = 0; k < NK; ++k) acc[b][k] += e[k] * ph;
        }
    }
    if (j >= A.nsites) return;
#pragma unroll
    for (int b = 0; b < BI; ++b) {
        if (b >= ni) break;
        double *o = R.S + (((size_t)blockIdx.z * A.nsites + i0 + b) * NK) * A.nsites + j;
#pragma unroll
        for (int k = 0; k < NK; ++k) o[(size_t)k * A.nsites] = acc[b][k];
    }
}

// One thread per element (ordinal of site j's type, term of site i) of
// nonLocalResidualQ: the real-space Ewald sum over the cells, stopped at
// the first cell beyond both ranges (elecStat.f90:954-980), then the
// reciprocal sum of the site pair and its zero-point term (:986-990).
__global__ __launch_bounds__(256) void k_es_resid(EsArgs A, EsRecip R, double *__restrict__ out)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (size_t)A.npt * A.pot_dim) return;
    const int row = (int)(g % A.npt), col = (int)(g / A.npt);
    const int t1 = A.term_type[col], k = A.term_alpha[col];
    const int t2 = R.ord_type[row];
    const double a1 = A.type_alphas[(size_t)t1 * A.maxpa + k];
    const double a2 = A.type_alphas[(size_t)t2 * A.maxpa];
    const double pi32 = pow(kEsPi, 1.5), x2isp = 2.0 / sqrt(kEsPi);
    const double af = pi32 / (a1 * sqrt(a1)) * (pi32 / (a2 * sqrt(a2)));
    const double sred = sqrt(a1 * a2 / (a1 + a2));
    const double avf = 4.0 * kEsPi * af / R.volume;
    const double sqth = sqrt(A.th);
    const double zp = 0.25 / R.min_pot;
    const int i0 = A.type_first[t1], i1 = i0 + A.type_mult[t1];
    const int j0 = A.type_first[t2], j1 = j0 + A.type_mult[t2];
    double acc = 0.0;
    for (int i = i0; i < i1; ++i) {
        const int ts = R.src_type[R.site_src[i]];
        const double as = A.type_alphas[(size_t)ts * A.maxpa + k];
        const double mras = sqrt(as * R.min_pot / (as + R.min_pot));
        const double asf = x2isp * af * (sred - mras);
        for (int j = j0; j < j1; ++j) {
            double ms[3], sq;
            es_min_sep(A, i, j, ms, &sq);
            for (int l = 0; l < A.ncells; ++l) {
                const double r2 = es_r2(A, ms, l);
                if (r2 < kEsSmall) {
                    acc = acc + asf;
                } else {
                    const double r = sqrt(r2);
                    bool stop;
                    {
#pragma clang fp contract(off)
                        stop = mras * r > sqth && sred * r > sqth;
                    }
                    if (stop) break;
                    acc = acc + af * (-es_erf(mras * r) + es_erf(sred * r)) / r;
                }
            }
            double sum = 0.0;
            for (int z = 0; z < R.chunks; ++z)
                sum += R.S[(((size_t)z * A.nsites + i) * R.nk + k) * A.nsites + j];
            acc = acc + avf * (sum - zp);
        }
    }
    out[g] = acc;
}

inline unsigned es_blocks(size_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

hipError_t launch_es_neutral_nuclear(const EsArgs &A, double *ovl, double *nl_neut,
                                     double *l_neut, double *nl_nuc, double *l_nuc,
                                     hipStream_t stream)
{
    if (A.maxpa > kEsMaxAlphas || A.npt < 1 || (long)A.npt * A.npt > 0x7fffffffL)
        return hipErrorInvalidValue;
    const size_t npair = (size_t)A.nsites * A.nsites;
    if (npair / 256 >= 0x7fffffffUL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_es_minsep, dim3(es_blocks(npair, 256)), dim3(256), 0, stream, A);
    hipLaunchKernelGGL(k_es_neut, dim3((unsigned)(A.npt * A.npt)), dim3(256), 0, stream, A, ovl,
                       nl_neut, l_neut);
    hipLaunchKernelGGL(k_es_nuc, dim3((unsigned)A.npt), dim3(64 * kEsNucWaves), 0, stream, A,
                       nl_nuc, l_nuc);
    return hipGetLastError();
}

hipError_t launch_es_residual_q(const EsArgs &A, const EsRecip &R, double *out,
                                hipStream_t stream, hipEvent_t recip0, hipEvent_t recip1)
{
    if (A.maxpa > kEsMaxAlphas || R.nk != es_recip_nk(A.maxpa) || R.nblk < 1 ||
        R.nblk > 65535 || R.chunks < 1 || R.chunks > 65535 || R.ncr < 1)
        return hipErrorInvalidValue;
    const size_t ne = (size_t)R.nsrc * R.ncr * R.nk;
    const size_t ncs = (size_t)R.ncr * A.nsites;
    const size_t nout = (size_t)A.npt * A.pot_dim;
    const size_t npair = (size_t)A.nsites * A.nsites;
    if (es_blocks(ne, 256) == 0 || ne / 256 >= 0x7fffffffUL || ncs / 256 >= 0x7fffffffUL ||
        nout / 256 >= 0x7fffffffUL || npair / 256 >= 0x7fffffffUL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_es_minsep, dim3(es_blocks(npair, 256)), dim3(256), 0, stream, A);
    hipLaunchKernelGGL(k_es_recip_exp, dim3(es_blocks(ne, 256)), dim3(256), 0, stream, A, R);
    hipLaunchKernelGGL(k_es_sincos, dim3(es_blocks(ncs, 256)), dim3(256), 0, stream, A, R);
    const dim3 grid((unsigned)((A.nsites + 63) / 64), (unsigned)R.nblk, (unsigned)R.chunks);
    if (recip0) {
        const hipError_t e = hipEventRecord(recip0, stream);
        if (e != hipSuccess) return e;
    }
    if (R.nk == 16)
        hipLaunchKernelGGL((k_es_recip<16, 4>), grid, dim3(64), 0, stream, A, R);
    else if (R.nk == 32)
        hipLaunchKernelGGL((k_es_recip<32, 2>), grid, dim3(64), 0, stream, A, R);
    else
        hipLaunchKernelGGL((k_es_recip<64, 1>), grid, dim3(64), 0, stream, A, R);
    if (recip1) {
        const hipError_t e = hipEventRecord(recip1, stream);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_es_resid, dim3(es_blocks(nout, 256)), dim3(256), 0, stream, A, R, out);
    return hipGetLastError();
}

}  // namespace imago
