// core_ortho.hip - the orthogonalisation of the valence functions against
// the core functions (orthoOL / ortho, src/imago/integrals.F90:3608-3910;
// intgOrtho.F90:23-213), around the GEMMs the host issues:
//
//   X'vv = Xvv - Scv^H Xcv - Xcv^H Scv + Scv^H (Xcc Scv)
//
// for every operator X (X = S for the overlap, where it reads
// Svv - 2 Scv^H Scv + Scv^H Scc Scv as zherk + makeValeVale write it).
// The extended matrices, each type's valence then core states, are
// assembled by the ordinary kernels into an extended block layout;
// k_core_unpack spreads them into dense [valence | core] matrices per
// k-point, and k_core_pack returns the upper valence triangle to the
// valence block layout with makeValeVale's flush of components below
// 1e-10 (intgOrtho.F90:137-190).
#include <hip/hip_runtime.h>

#include "assemble.hpp"
#include "core_ortho_launch.hpp"

namespace imago {

// dense[k] (N x N, column-major) entry (r, q) of the extended matrix; r, q
// count the valence functions first, then the core functions.  Entries
// below the block diagonal come from the conjugate of their mirror (the
// block layout holds the upper blocks, a diagonal block's upper triangle).
template <bool CX>
__global__ __launch_bounds__(256) void k_core_unpack(
    const double *__restrict__ blk, long nbe, int N, int nk, const PairRec *__restrict__ pairs,
    const int *__restrict__ fn_atom, const int *__restrict__ fn_state,
    const int *__restrict__ pair_index, const int *__restrict__ atom_nstates, int natoms,
    double *__restrict__ dense)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nn = (long)N * N;
    if (t >= nn * nk) return;
    const int k = (int)(t / nn);
    const long rc = t % nn;
    const int r = (int)(rc % N), q = (int)(rc / N);
    int a1 = fn_atom[r], s1 = fn_state[r], a2 = fn_atom[q], s2 = fn_state[q];
    const bool mirror = a1 > a2 || (a1 == a2 && s1 > s2);
    if (mirror) {
        const int ta = a1, ts = s1;
        a1 = a2;
        s1 = s2;
        a2 = ta;
        s2 = ts;
    }
    const PairRec pr = pairs[pair_index[(size_t)a1 * natoms + a2]];
    const long e = (long)pr.blk + s1 + (long)atom_nstates[a1] * s2;
    if (CX) {
        const double2 v = reinterpret_cast<const double2 *>(blk)[(size_t)k * nbe + e];
        reinterpret_cast<double2 *>(dense)[t] = make_double2(v.x, mirror ? -v.y : v.y);
    } else {
        dense[t] = blk[(size_t)k * nbe + e];
    }
}

__device__ __forceinline__ double flush10(double x) { return fabs(x) < 1.0e-10 ? 0.0 : x; }

// the upper n x n triangle of dense[k] (leading dimension N) into the
// valence block layout; a diagonal block's strict lower triangle is zero
template <bool CX>
__global__ __launch_bounds__(256) void k_core_pack(
    const double *__restrict__ dense, int N, int n, int nk, const PairRec *__restrict__ pairs,
    long nbe, const int *__restrict__ atom_of, const int *__restrict__ pair_index,
    const int *__restrict__ cumul, int natoms, double *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nn = (long)n * n;
    if (t >= nn * nk) return;
    const int k = (int)(t / nn);
    const long rc = t % nn;
    const int r = (int)(rc % n), q = (int)(rc / n);
    const int a = atom_of[r], b = atom_of[q];
    if (r > q && a != b) return;   // the mirror of an upper block entry
    const int lo = r <= q ? a : b, hi = r <= q ? b : a;
    const PairRec pr = pairs[pair_index[(size_t)lo * natoms + hi]];
    const int ns = (a + 1 < natoms ? cumul[a + 1] : n) - cumul[a];
    const long e = (long)pr.blk + (r - cumul[a]) + (long)ns * (q - cumul[b]);
    const size_t src = (size_t)k * N * N + (size_t)q * N + r;
    if (CX) {
        double2 v = make_double2(0.0, 0.0);
        if (r <= q) {
            const double2 d = reinterpret_cast<const double2 *>(dense)[src];
            v = make_double2(flush10(d.x), flush10(d.y));
        }
        reinterpret_cast<double2 *>(out)[(size_t)k * nbe + e] = v;
    } else {
        out[(size_t)k * nbe + e] = r <= q ? flush10(dense[src]) : 0.0;
    }
}

hipError_t launch_core_unpack(const double *blk, long nbe, int N, int nk, bool cx,
                              const PairRec *pairs, const int *fn_atom, const int *fn_state,
                              const int *pair_index, const int *atom_nstates, int natoms,
                              double *dense, hipStream_t stream)
{
    const long total = (long)N * N * nk;
    const unsigned g = (unsigned)((total + 255) / 256);
    if (cx)
        hipLaunchKernelGGL(k_core_unpack<true>, dim3(g), dim3(256), 0, stream, blk, nbe, N, nk,
                           pairs, fn_atom, fn_state, pair_index, atom_nstates, natoms, dense);
    else
        hipLaunchKernelGGL(k_core_unpack<false>, dim3(g), dim3(256), 0, stream, blk, nbe, N, nk,
                           pairs, fn_atom, fn_state, pair_index, atom_nstates, natoms, dense);
    return hipGetLastError();
}

hipError_t launch_core_pack(const double *dense, int N, int n, int nk, bool cx,
                            const PairRec *pairs, long nbe, const int *atom_of,
                            const int *pair_index, const int *cumul, int natoms, double *out,
                            hipStream_t stream)
{
    const long total = (long)n * n * nk;
    const unsigned g = (unsigned)((total + 255) / 256);
    if (cx)
        hipLaunchKernelGGL(k_core_pack<true>, dim3(g), dim3(256), 0, stream, dense, N, n, nk, pairs,
                           nbe, atom_of, pair_index, cumul, natoms, out);
    else
        hipLaunchKernelGGL(k_core_pack<false>, dim3(g), dim3(256), 0, stream, dense, N, n, nk,
                           pairs, nbe, atom_of, pair_index, cumul, natoms, out);
    return hipGetLastError();
}

}   // namespace imago
