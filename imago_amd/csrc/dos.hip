// dos.hip - the Gaussian-broadened total and partial density of states of
// the reference's computeDOS (src/imago/dos.F90:160-1231, the
// kPointIntgCode /= 1 path) on the eigenpairs the library holds:
//   k_dos_mulliken  q_kjv = Re(conj(C_vj) (S C)_vj), the Mulliken gross
//                   population of basis function v in state j at k-point k
//                   (dos.F90:640-679: waveFnSqrd times the overlap column);
//   k_dos_gauss     G_(kj),e = w_k exp(-x^2) / (sigma sqrt(pi)) / hartree
//                   with x = (E_kj - e_e) / sigma, zero where x^2 >= 50
//                   (dos.F90:707-743);
//   k_dos_counts    the electron number of every basis function (states at
//                   or below zero counted with 1/spin, dos.F90:627-632,
//                   682-684) and the localisation index of every state
//                   (dos.F90:697-700).
// The channel sums and the broadening itself are GEMMs (P = Q^T M with M
// the basis-function -> channel map, PDOS = P^T G), issued by the caller.
#include <hip/hip_runtime.h>

#include "dos_launch.hpp"

namespace imago {

namespace {

constexpr double kHartreeEv = 27.211386245988;
constexpr double kSqrtPi = 1.7724538509055160273;

// Q[k][j][v] (v fastest) from C [k][j][v] (column j of the state matrix,
// stride cstride per k-point) and SC [k][j][v]
template <bool CX>
__global__ __launch_bounds__(256) void k_dos_mulliken(const double *__restrict__ C, long cstride,
                                                      const double *__restrict__ SC, int n,
                                                      int ns, int nk, double *__restrict__ Q)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    const long per = (long)n * ns;
    if (g >= per * nk) return;
    const long k = g / per, q = g % per;
    if (CX) {
        const double2 c = reinterpret_cast<const double2 *>(C)[k * cstride + q];
        const double2 s = reinterpret_cast<const double2 *>(SC)[g];
        Q[g] = c.x * s.x + c.y * s.y;
    } else {
        Q[g] = C[k * cstride + q] * SC[g];
    }
}

__global__ __launch_bounds__(256) void k_dos_gauss(const double *__restrict__ E,
                                                   const double *__restrict__ w, int ns, int nk,
                                                   double emin, double delta, int ne,
                                                   double sigma, double *__restrict__ G)
{
    const long g = (long)blockIdx.x * 256 + threadIdx.x;
    const long rows = (long)nk * ns;
    if (g >= rows * ne) return;
    const long r = g % rows;
    const int e = (int)(g / rows), k = (int)(r / ns);
    const double x = (E[r] - (emin + e * delta)) / sigma;
    const double x2 = x * x;
    G[g] = x2 < 50.0 ? exp(-x2) / (sigma * kSqrtPi) / kHartreeEv * w[k] : 0.0;
}

// en[v] = sum_k sum_{j: E_kj <= 0} Q_kjv w_k / spin; loc[j] = sum_k sum_v
// Q_kjv^2 w_k / spin: one thread per v, then one per j
__global__ __launch_bounds__(256) void k_dos_counts(const double *__restrict__ Q,
                                                    const double *__restrict__ E,
                                                    const double *__restrict__ w, int n, int ns,
                                                    int nk, int spin, double *__restrict__ en,
                                                    double *__restrict__ loc)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const double inv = 1.0 / spin;
    if (t < n) {
        double a = 0.0;
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < ns; ++j)
                if (E[(size_t)k * ns + j] <= 0.0) a += Q[((size_t)k * ns + j) * n + t] * w[k];
        en[t] = a * inv;
    } else if (t < n + ns) {
        const int j = t - n;
        double a = 0.0;
        for (int k = 0; k < nk; ++k) {
            const double *q = Q + ((size_t)k * ns + j) * n;
            double s = 0.0;
            for (int v = 0; v < n; ++v) s += q[v] * q[v];
            a += s * w[k];
        }
        loc[j] = a * inv;
    }
}

}   // namespace

hipError_t launch_dos_mulliken(const double *C, long cstride, const double *SC, int n, int ns,
                               int nk, bool cx, double *Q, hipStream_t stream)
{
    const long tot = (long)n * ns * nk;
    const unsigned g = (unsigned)((tot + 255) / 256);
    if (cx)
        hipLaunchKernelGGL(k_dos_mulliken<true>, dim3(g), dim3(256), 0, stream, C, cstride, SC, n,
                           ns, nk, Q);
    else
        hipLaunchKernelGGL(k_dos_mulliken<false>, dim3(g), dim3(256), 0, stream, C, cstride, SC,
                           n, ns, nk, Q);
    return hipGetLastError();
}

hipError_t launch_dos_gauss(const double *E, const double *w, int ns, int nk, double emin,
                            double delta, int ne, double sigma, double *G, hipStream_t stream)
{
    const long tot = (long)nk * ns * ne;
    hipLaunchKernelGGL(k_dos_gauss, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, E,
                       w, ns, nk, emin, delta, ne, sigma, G);
    return hipGetLastError();
}

hipError_t launch_dos_counts(const double *Q, const double *E, const double *w, int n, int ns,
                             int nk, int spin, double *en, double *loc, hipStream_t stream)
{
    hipLaunchKernelGGL(k_dos_counts, dim3((unsigned)((n + ns + 255) / 256)), dim3(256), 0, stream,
                       Q, E, w, n, ns, nk, spin, en, loc);
    return hipGetLastError();
}

}   // namespace imago
