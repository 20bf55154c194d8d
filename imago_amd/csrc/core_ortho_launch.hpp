// core_ortho_launch.hpp - launchers of the core-orthogonalisation kernels
// (core_ortho.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "assemble.hpp"

namespace imago {

// dense[k] (N x N, column-major, complex when cx) from the extended block
// layout (nbe entries per k-point): functions ordered valence first, then
// core; fn_atom / fn_state give each one's atom and its state within the
// atom's extended state list (valence states, then core states),
// atom_nstates each atom's extended state count.
hipError_t launch_core_unpack(const double *blk, long nbe, int N, int nk, bool cx,
                              const PairRec *pairs, const int *fn_atom, const int *fn_state,
                              const int *pair_index, const int *atom_nstates, int natoms,
                              double *dense, hipStream_t stream);

// the upper n x n valence triangle of dense[k] (leading dimension N) into
// the valence block layout (nbe entries per k-point), components below
// 1e-10 flushed to zero.
hipError_t launch_core_pack(const double *dense, int N, int n, int nk, bool cx,
                            const PairRec *pairs, long nbe, const int *atom_of,
                            const int *pair_index, const int *cumul, int natoms, double *out,
                            hipStream_t stream);

}   // namespace imago
