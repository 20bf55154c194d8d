// xcmesh_launch.hpp - host interface of the exchange-correlation mesh
// kernels (xcmesh.hip): getECMeshParameters and makeECMeshAndOverlap of
// exchCorr.f90.
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

// What the kernels read of the system; device pointers unless noted.
struct XcArgs {
    int nsites, npt, pot_dim, ncells, maxpa, nreps, nvec;
    const double *site_pos;     // [site][3]
    const int *site_type;       // [site]
    const double *site_cov;     // [site] covalent radius of the site's type
    const int *type_first;      // [type] first site
    const int *type_mult;       // [type] sites
    const int *type_nalpha;     // [type]
    const double *type_alphas;  // [type][maxpa] ascending
    const int *type_term0;      // [type] first term in potDim
    const double *cell_dims;    // [cell][3], |R|^2 ascending
    const double *cell_sizes;   // [cell]
    const double *pls;          // [site][3] site minus its nearest lattice point
    const double *vec;          // [nvec][3] angular sample vectors
    double rv[9], inv[9];       // by value, for findLatticeVector
    double th;                  // logElecThresh
    double max_radius;          // sqrt(rSampleOut / minPotAlpha)
};

// siteSampleMaxRadius (exchCorr.f90:221-248): maxrad[site][vector].
hipError_t launch_xc_maxrad(const XcArgs &X, double *maxrad, hipStream_t stream);

// The operator of npts ray points (exchCorr.f90:587-760): op[o][p][term]
// for o < (gga ? 10 : 1), p < npts, leading dimension ld points per
// operator value; zeroed by the caller.  When nz is not null,
// nz[pt_slot[p]][term] is set to 1 for every term a point adds to.
hipError_t launch_xc_op(const XcArgs &X, bool gga, const double *pts, int npts, int ld,
                        double *op, unsigned char *nz, const int *pt_slot,
                        hipStream_t stream);

// Row lists of nslots flag rows: rows[s][0 .. count[s]) = the terms whose
// flag nz[s][term] is set, ascending.
hipError_t launch_xc_compact(const unsigned char *nz, int nslots, int pot_dim, int *rows,
                             int *count, hipStream_t stream);

// A(c, p) = op[p][rows[c]] * sqrt(w[p]) for p < npts, c < m (A column-major
// with leading dimension lda >= m, op with leading dimension pot_dim).
hipError_t launch_xc_gather(const double *op, const double *w, int pot_dim, const int *rows,
                            int m, int npts, int lda, double *A, hipStream_t stream);

// C(rows[c1], rows[c2]) += Cs(c1, c2) for c1 <= c2 < m (C pot_dim x pot_dim,
// Cs with leading dimension ldc; both column-major).
hipError_t launch_xc_scatter(const double *Cs, int ldc, const int *rows, int m, int pot_dim,
                             double *C, hipStream_t stream);

}  // namespace imago
