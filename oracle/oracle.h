/*
 * oracle.h - CPU restatement of the reference's integral + secular path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU library
 * is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never does.
 *
 * Every function cites the reference routine it restates
 * (UMKC-CPG/imago, paths relative to the repository root).
 */
#ifndef IMAGO_ORACLE_H
#define IMAGO_ORACLE_H

#include "../include/imago_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Primitive integrals, one tuple, sh(16,16) column-major; entries outside
 * the leading lmAlphaIndex block are set to 0. */
void orc_overlap2c(double a1, double a2, const double *A, const double *B,
                   int l1l2switch, double *sh);
void orc_kinetic2c(double a1, double a2, const double *A, const double *B,
                   int l1l2switch, double *sh);
void orc_electron3c(double a1, double a2, double a3, const double *A,
                    const double *B, const double *C, int l1l2switch,
                    double *sh);
void orc_nuclear3c(double a1, double a2, double a3, const double *A,
                   const double *B, const double *C, int l1l2switch,
                   double *sh);
void orc_boys(double T, double *F7);

/* Primitive evaluator selection for the assembly loops: 0 = this file's
 * restatement, 1 = external (set with orc_set_external_primitives, e.g.
 * the reference's own compiled module). */
typedef void (*orc_prim2c_fn)(double, double, const double *, const double *,
                              int, double *);
typedef void (*orc_prim3c_fn)(double, double, double, const double *,
                              const double *, const double *, int, double *);
void orc_set_external_primitives(orc_prim2c_fn ol, orc_prim2c_fn ke,
                                 orc_prim3c_fn ep, orc_prim3c_fn np);
void orc_use_external_primitives(int on);

/* Screening tables of gaussRelations.f90 for element pair / triple.
 * Returns 0 on success. out arrays sized by the caller:
 *   alpha_dist[max_alphas*max_alphas*nel*nel] in Fortran order
 *   alphaDist(a1,a2,e1,e2). */
int orc_alpha_dist(const imago_system_desc *sys, double *alpha_dist,
                   double *alpha_center, int *num_elements);

/* findLatticeVector (lattice.f90:784-890): lattice vector closest to v. */
void orc_find_lattice_vector(const imago_system_desc *sys, const double *v,
                             double *L);

/* Lattice-summed assembly of one kind, loop nest of integrals.F90.
 * out: complex packed per k, [nk][2*npack] (or [npack] for gamma).
 * For IMAGO_EP: terms [term_begin, term_end) in the reference's order,
 * out is [term][nk][...].  Returns 0, or IMAGO_E_LATTICE when the cell
 * list is too small (integrals.F90:378-383). */
int orc_assemble(const imago_system_desc *sys, int kind, int term_begin,
                 int term_end, double *out);

/* readPackedMatrixAccum semantics (matrixSubs.F90:339-376) over a whole
 * H build: H = NP + KE + sum_j c_j V_j for packed arrays of length len. */
void orc_build_hamiltonian(long len, const double *np, const double *ke,
                           int pot_dim, const double *const *v,
                           const double *coeffs, double *h);

/* matrixElementMult (matrixSubs.F90:66-112): returns
 * 2*(sum_offdiag re*re+im*im + sum_diag re*re/2) for complex packed arrays,
 * dim1 = 2 (complex) or 1 (real, matrixElementMultGamma). */
double orc_matrix_element_mult(int dim1, int n, const double *m1,
                               const double *m2);

/* threads of the assembly's outer atom loop (default 1: serial) */
void orc_set_threads(int n);

/* number of primitive tuples evaluated by the last orc_assemble call */
long orc_last_tuple_count(void);

/* error_fn (mathSubs.f90:22-161): the Sandia/Cody rational erf. */
double orc_error_fn(double x);

/* neutralAndNuclearQPot (elecStat.f90:59-664) at log threshold th
 * (-ln electroConvg).  ovl, nl_neut, l_neut: potDim x potDim column-major,
 * element (term of site j, term of site i); nl_nuc, l_nuc: [potDim]. */
int orc_elecstat(const imago_system_desc *sys, double th, double *ovl,
                 double *nl_neut, double *l_neut, double *nl_nuc,
                 double *l_nuc);

/* residualQ (elecStat.f90:665-1043).  Reciprocal cells cdr[ncr][3] with
 * sizes csr[ncr] (|G|^2 ascending, cell 0 the origin); out numPotTypes x
 * potDim column-major.  Returns 0 or IMAGO_E_ARG. */
int orc_residual_q(const imago_system_desc *sys, double th, int ncr,
                   const double *cdr, const double *csr, double *out);
/* The same for the outer-loop sites [i_begin, i_end) only: the columns of
 * their types, other columns 0. */
int orc_residual_q_sites(const imago_system_desc *sys, double th, int ncr,
                         const double *cdr, const double *csr, int i_begin,
                         int i_end, double *out);

/* makeSampleVectors (exchCorr.f90:920-980): v[n][3]. */
void orc_xc_sample_vectors(int n, double *v);

/* getECMeshParameters (exchCorr.f90:63-342): pls[site][3], maxrad
 * [site][nvec], npts[site]; returns maxNumRayPoints. */
int orc_xc_mesh(const imago_system_desc *sys, const imago_xc_mesh_params *p,
                const double *cov, const double *vec, double *pls,
                double *maxrad, int *npts);

/* One site's mesh of makeECMeshAndOverlap (exchCorr.f90:492-775):
 * rw[maxpts], op[numOp][maxpts][potDim]; returns its ray points. */
int orc_xc_site(const imago_system_desc *sys, const imago_xc_mesh_params *p,
                double th, const double *cov, const double *vec,
                const double *pls, const double *maxrad, int maxpts, int site,
                int gga, double *rw, double *op);

/* exchCorrOverlap (exchCorr.f90:805-810), potDim x potDim column-major,
 * upper triangle. */
int orc_xc_overlap(const imago_system_desc *sys, const imago_xc_mesh_params *p,
                   double th, const double *cov, const double *vec,
                   const double *pls, const double *maxrad, int maxpts,
                   double *ovl);
int orc_xc_overlap_sites(const imago_system_desc *sys,
                         const imago_xc_mesh_params *p, double th,
                         const double *cov, const double *vec,
                         const double *pls, const double *maxrad, int maxpts,
                         int i_begin, int i_end, double *ovl);

/* wignerXC (potentialUpdate.F90:2835-2867): potential and energy density. */
void orc_xc_wigner(double rho, double *pot, double *en);

/* exchCorrPot(:, 1:3) of makeSCFPot before its solve, Wigner functional,
 * spin 1, LDA (potentialUpdate.F90:833-1008); out[3][potDim]. */
int orc_xc_potential(const imago_system_desc *sys, const imago_xc_mesh_params *p,
                     double th, const double *cov, const double *vec,
                     const double *pls, const double *maxrad, int maxpts,
                     const double *rho_total, const double *rho_core,
                     double *out);

#ifdef __cplusplus
}
#endif
#endif
