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
// for o < (gga ? 10 : 1), p < npts, ld points per operator value.  LDA
// writes every term of every point; GGA accumulates into op zeroed by the
// caller.
hipError_t launch_xc_op(const XcArgs &X, bool gga, const double *pts, int npts, int ld,
                        double *op, hipStream_t stream);

// nz[b][term] = 1 for the terms the points of site s0 + b can touch (a
// superset, from the nearest image of every site and the largest ray
// length dmax[site]), b < nb.
hipError_t launch_xc_rows(const XcArgs &X, int s0, int nb, const double *dmax,
                          unsigned char *nz, hipStream_t stream);

// rank[b][rows[b][c]] = c for c < count[b]; rank preset to -1 by the caller.
hipError_t launch_xc_rank(const int *rows, const int *count, int nb, int pot_dim, int *rank,
                          hipStream_t stream);

// The LDA operator of npts points of a batch of sites written into the
// product operands: B_b(p - slot_first[b], rank[b][term]) = op(term, p) *
// sqrt(w[p]) (w null: unscaled), b = pt_slot[p], B_b = B + b * lda * ldn (column-major, ldn
// rows = points, lda columns = terms).  *err is set if a term outside a
// slot's rows is not 0.
hipError_t launch_xc_op_rows(const XcArgs &X, const double *pts, const double *w, int npts,
                             const int *pt_slot, const int *slot_first, const int *rank, int lda,
                             int ldn, double *B, int *err, hipStream_t stream);

// The exchange-correlation terms of one batch (Wigner, spin 1, LDA;
// potentialUpdate.F90:833-1008), in two steps around a batched GEMM part_b
// = B_b^T V_b: V [nb][3][ldn] = w_p times (potential, energy density of
// the total, energy density of the core) at every point, zero beyond a
// slot's points (zeroed by the caller); then pot[k][r] += part_b(rank, k)
// slot by slot.  B holds the unscaled operator (k_xc_op_rows, w null).
hipError_t launch_xc_point_pot(const double *B, int lda, int ldn, const int *pt_slot,
                               const int *slot_first, const int *rows, const int *count,
                               int pot_dim, const double *w, const double *rho_total,
                               const double *rho_core, int npts, double *V, hipStream_t stream);
hipError_t launch_xc_pot_add(const double *part, int lda, const int *rank, int nb, int pot_dim,
                             double *pot, hipStream_t stream);

// Row lists of nslots flag rows: rows[s][0 .. count[s]) = the terms whose
// flag nz[s][term] is set, ascending.
hipError_t launch_xc_compact(const unsigned char *nz, int nslots, int pot_dim, int *rows,
                             int *count, hipStream_t stream);

// C(rows[c1], rows[c2]) += Cs(c1, c2) for c1 <= c2 < m (C pot_dim x pot_dim,
// Cs with leading dimension ldc; both column-major).
hipError_t launch_xc_scatter(const double *Cs, int ldc, const int *rows, int m, int pot_dim,
                             double *C, hipStream_t stream);

}  // namespace imago
