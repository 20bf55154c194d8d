// dos_launch.hpp - launchers of the density-of-states kernels (dos.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

// Q [nk][ns][n] = Re(conj(C) .* SC); C column j of k-point k at
// C + (k * cstride + j * n) scalars (complex when cx), SC [nk][ns][n].
hipError_t launch_dos_mulliken(const double *C, long cstride, const double *SC, int n, int ns,
                               int nk, bool cx, double *Q, hipStream_t stream);
// G [ne][nk * ns] (column-major (nk * ns) x ne): Gaussian weights of every
// state at every energy point.
hipError_t launch_dos_gauss(const double *E, const double *w, int ns, int nk, double emin,
                            double delta, int ne, double sigma, double *G, hipStream_t stream);
// electron number per basis function, localisation index per state
hipError_t launch_dos_counts(const double *Q, const double *E, const double *w, int n, int ns,
                             int nk, int spin, double *en, double *loc, hipStream_t stream);

}   // namespace imago
