/*
 * imago_hip.h - C ABI of the MI355X integral + k-point eigensolve library.
 *
 * Replaces the body of the Imago/OLCAO engine's setup-time integral stage
 * and its per-iteration secular-equation stage (UMKC-CPG/imago):
 *   src/imago/gaussIntegrals.f90  primitive Obara-Saika integrals + Boys
 *   src/imago/integrals.F90       lattice-summed OL/KE/NP/EP assembly
 *   src/imago/intgSaving.F90      contraction, Bloch phases, block save
 *   src/imago/secularEqn.F90      H(k) = NP + KE + sum_j c_j V_j, solve
 *   src/imago/interfaces.F90      zhegv / dsygv boundary
 *   src/imago/valeCharge.F90      density matrix and Tr(rho M) traces
 *
 * Conventions (all entry points):
 *   - every function returns IMAGO_OK (0) or a negative IMAGO_E* code;
 *     imago_last_error() gives the message of the most recent failure
 *     (the reference writes a message to unit 20 and stops instead);
 *   - arrays are column-major like the Fortran host's, caller-owned, host
 *     memory unless the name says _dev;
 *   - "packed" = upper triangle column by column, element (r,c), r<=c,
 *     1-based, at index c(c-1)/2 + r; complex packed matrices interleave
 *     (re,im) pairs exactly like the reference's real(2, n(n+1)/2) arrays
 *     (src/imago/matrixSubs.F90:506-535, hdf5SCFIntg.F90:134);
 *   - a context is not thread-safe; use one per host thread.
 *
 * No symbol here takes a torch or HIP type: the library owns its device
 * memory and stream.
 */
#ifndef IMAGO_HIP_H
#define IMAGO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define IMAGO_OK 0
#define IMAGO_E_ARG (-1)        /* invalid argument / inconsistent sizes   */
#define IMAGO_E_HIP (-2)        /* HIP runtime or library failure          */
#define IMAGO_E_LATTICE (-3)    /* cell list too small (integrals.F90:378) */
#define IMAGO_E_NOTPD (-4)      /* overlap not positive definite (zhegv
                                   info > n, interfaces.F90:106-109)       */
#define IMAGO_E_SOLVE (-5)      /* eigensolver did not converge (0<info<=n)*/
#define IMAGO_E_STATE (-6)      /* call order: e.g. assemble before set    */
#define IMAGO_E_SCREEN (-7)     /* screening table not monotone: the
                                   reference's alpha walk would differ     */
#define IMAGO_E_NOMEM (-8)      /* device allocation failed                */

/* integral kinds (opcodes of integrals.F90 / ortho()) */
#define IMAGO_OL 1   /* overlap            gaussOverlapOL  integrals.F90:151  */
#define IMAGO_KE 2   /* kinetic            gaussOverlapKE  integrals.F90:656  */
#define IMAGO_NP 3   /* nuclear potential  gaussOverlapNP  integrals.F90:1531 */
#define IMAGO_EP 4   /* electronic pot.    elecPotGaussOverlap :2357          */
#define IMAGO_EP_DIRECT 5 /* EP without panels: see imago_assemble           */

/* Boys-function modes for the nuclear integral */
#define IMAGO_BOYS_REFERENCE 0  /* gaussIntegrals.f90:16080-16221 verbatim */

/*
 * System description: what the reference keeps as module state after
 * setupSCF's input/lattice/k-point/basis steps (imago.F90:103-256).
 * All indices are 0-based.  Elements are small integers shared by types
 * of the same element: the reference keys its screening tables by element
 * and takes the exponents of the FIRST type of each element
 * (gaussRelations.f90:64-278); this library does the same.
 */
typedef struct imago_system_desc {
    /* atom types (atomicTypes.f90; basis.f90:55-243) */
    int num_atom_types;
    int max_alphas;               /* leading dimension of type_alphas      */
    int max_states;               /* leading dimension of type_basis       */
    const int *type_element;      /* [num_atom_types]                      */
    const int *type_num_orb_alphas; /* [num_atom_types][4] numOrbAlphas(l+1),
                                     non-increasing in l                   */
    const double *type_alphas;    /* [num_atom_types][max_alphas] ascending */
    const int *type_num_states;   /* [num_atom_types] valence states        */
    const double *type_basis;     /* [num_atom_types][max_states][max_alphas]
                                     radial coefficient * angNorm(lm); each
                                     type's valence states, then its core
                                     states (type_num_core_states)         */
    const int *type_lm_index;     /* [num_atom_types][max_states] 1..16     */

    /* atom sites (atomicSites.f90) */
    int num_atoms;
    const int *atom_type;         /* [num_atoms]                            */
    const double *atom_pos;       /* [num_atoms][3] Cartesian, bohr         */

    /* potential types and sites (potTypes.f90, potSites.f90) */
    int num_pot_types;
    int max_pot_alphas;
    const int *pot_type_element;  /* [num_pot_types]                        */
    const int *pot_type_num_alphas; /* [num_pot_types]                      */
    const double *pot_type_alphas;  /* [num_pot_types][max_pot_alphas] asc. */
    const double *pot_type_nuc_charge; /* [num_pot_types] Z                 */
    const double *pot_type_nuc_alpha;  /* [num_pot_types] nucAlpha          */
    const int *pot_type_first_site;    /* [num_pot_types] first site index  */
    const int *pot_type_multiplicity;  /* [num_pot_types] contiguous sites  */
    int num_pot_sites;
    const int *pot_site_type;     /* [num_pot_sites]                        */
    const double *pot_site_pos;   /* [num_pot_sites][3]                     */

    /* real-space lattice (lattice.f90:310-696): cells sorted by |R|^2 */
    double real_vectors[9];       /* a,b,c as columns: real_vectors[3*i+x]  */
    int num_cells;
    const double *cell_dims;      /* [num_cells][3]                         */
    const double *cell_sizes;     /* [num_cells] |R|^2 ascending            */

    /* k-points, Cartesian with the 2*pi included (kpoints.f90:543-576) */
    int num_kpoints;
    const double *kpoints;        /* [num_kpoints][3]                       */
    const double *kpoint_weights; /* [num_kpoints]                          */
    int gamma;                    /* 1: real Gamma-only build (imagoG)      */

    double log_basis_fn_thresh;   /* -ln(basisFnConvg) (lattice.f90:239)    */

    /* core states (numCoreStates, basis.f90:160-243), NULL when there are
     * none (coreDim = 0).  With core states the assembled matrices are
     * orthogonalised against them as orthoOL / ortho do (integrals.F90:
     * 3608-3910, intgOrtho.F90:23-213): IMAGO_OL must be assembled before
     * the other kinds, IMAGO_EP_DIRECT is refused. */
    const int *type_num_core_states; /* [num_atom_types]                    */
} imago_system_desc;

/*
 * Exchange-correlation mesh parameters, the input block of
 * readExchCorrMeshParameters (exchCorr.f90:876-913).
 */
typedef struct {
    int num_sample_vectors;       /* NUM_ANGULAR_SAMPLE_VECTORS             */
    double weight_in, weight_out; /* WTIN_WTOUT (angSampleWeights(1:2))     */
    double r_sample_in;           /* RADIAL_SAMPLE-IN_OUT_SPACING (1)       */
    double r_sample_out;          /*                              (2)       */
    double r_sample_space;        /*                              (3), < 1  */
} imago_xc_mesh_params;

typedef struct imago_ctx imago_ctx;

/* ---- context ---------------------------------------------------------- */
int imago_ctx_create(int device, imago_ctx **ctx);
int imago_ctx_destroy(imago_ctx *ctx);
const char *imago_last_error(void);

/* Upload a system.  Replaces the module state that initializeAtomSite /
 * initializePotSite read (basis.f90:257-383), the screening tables of
 * makeAlphaDist/makeAlphaNucDist/makeAlphaPotDist (gaussRelations.f90:
 * 64-278) and the phase factors of computePhaseFactors (kpoints.f90:578).
 * Returns valeDim and potDim through the out pointers. */
int imago_set_system(imago_ctx *ctx, const imago_system_desc *sys,
                     int *vale_dim, int *pot_dim);

/* ---- primitive seam: gaussIntegrals.f90, evaluated on the GPU ---------- *
 * Batched versions of overlap2CIntg (:26), kinetic2CIntg (:4422),
 * electron3CIntg (:11624), nuclear3CIntg (:16223) and boys (:16080).
 * For n tuples: a[3n] exponents (a3 unused for 2-centre kinds), centres
 * xyz[9n] as A,B,C, l1l2switch[n] as in integrals.F90:533-536.  sh[256n]
 * receives each sh(16,16) column-major; entries outside the leading
 * lmAlphaIndex(a1) x lmAlphaIndex(a2) block are written as 0. */
int imago_primitive_batch(imago_ctx *ctx, int kind, int n, const double *a,
                          const double *xyz, const int *l1l2switch,
                          double *sh);
int imago_boys_batch(imago_ctx *ctx, int n, const double *T, double *F7);

/* ---- assembly: gaussOverlapOL/KE/NP, elecPotGaussOverlap -------------- *
 * Computes the lattice-summed, Bloch-phased matrices of one kind into
 * device-resident panels (per k for the complex build; one real panel for
 * gamma); imago_get_packed returns any of them in the reference's packed
 * layout.  For IMAGO_EP all potDim terms are built, term j in the
 * reference's order (potential type, then alpha; integrals.F90:2411-2476).
 * IMAGO_EP_DIRECT builds no EP panels: it screens the EP tuples and stores
 * their primitive blocks once, and imago_secular / imago_density_traces
 * then form sum_j c_j V_j and Tr(rho V_j) from those blocks each call (the
 * fused accumulation of the reference's post-SCF path, integrals.F90:
 * 2860-3605), any number of k-points.  When both are assembled the
 * stored panels are used.
 * With core states (type_num_core_states) every kind is orthogonalised
 * against them as orthoOL / ortho do (integrals.F90:3608-3910,
 * intgOrtho.F90): IMAGO_OL first, then the others; IMAGO_EP_DIRECT is
 * refused.  Without them orthoOL / ortho reduce to packing. */
int imago_assemble(imago_ctx *ctx, int kind);

/* Copy one assembled packed matrix to the host: kind, term (EP only, else
 * 0), k-point.  out holds 2*n(n+1)/2 doubles (complex) or n(n+1)/2 (gamma).
 * This is the array the reference hands to h5dwrite_f in ortho(). */
int imago_get_packed(imago_ctx *ctx, int kind, int term, int kpoint,
                     double *out);

/* The inverse of imago_get_packed: load one packed matrix (same layout)
 * into the device panel of kind, term, k-point, in place of assembling it.
 * This is the read side of the reference's integral checkpoint: a run
 * resumed on an integral file whose group status is 1 reads the matrices
 * instead of computing them (integrals.F90:265-294; readPackedMatrix,
 * matrixSubs.F90:259-302).  The kind counts as assembled once every term
 * and k-point is loaded (or imago_assemble built it).  Entries of atom
 * pairs that have no lattice image within the system's cutoff must be
 * zero (IMAGO_E_ARG otherwise), as the assembly leaves them. */
int imago_set_packed(imago_ctx *ctx, int kind, int term, int kpoint,
                     const double *in);

/* ---- secular equation: secularEqnSCF (secularEqn.F90:37-328) ---------- *
 * H(k) = NP(k) + KE(k) + sum_j c_j V_j(k) for every k-point, with the
 * reference's rule that a coefficient of exactly 0.0 adds V_j unscaled
 * (readPackedMatrixAccum, matrixSubs.F90:369-373); then the generalized
 * problem H c = e S c of every k-point is solved on the GPU (zhegv/dsygv
 * itype=1, 'V', 'U'; interfaces.F90:73-200).  eigenvalues[nk*num_states]
 * (k-major) receive the lowest num_states values; eigenvectors stay on the
 * device for imago_density_traces.  info_out[nk] gets LAPACK-style info. */
int imago_secular(imago_ctx *ctx, const double *pot_coeffs, int num_states,
                  double *eigenvalues, int *info_out);

/* Copy the eigenvectors of one k-point (n x num_states, column-major,
 * complex interleaved or real for gamma) to the host. */
int imago_get_eigenvectors(imago_ctx *ctx, int kpoint, double *out);

/* ---- valence density and traces: makeValenceRho (valeCharge.F90) ----- *
 * rho(k) = sum_n f(n,k) c_n c_n^H (zher, upper), then Re Tr(rho M) via
 * matrixElementMult (matrixSubs.F90:66-112) summed over k for
 * M = S, NP, KE (traces[0..2]) and every V_j (pot_rho[j]).
 * occupations[nk*num_states] are the structured electron populations. */
int imago_density_traces(imago_ctx *ctx, const double *occupations,
                         double *traces3, double *pot_rho);

/* ---- single generalized eigenproblem on the GPU ----------------------- *
 * solveZHEGV (interfaces.F90:73-120): A (n x n complex, interleaved,
 * column-major, upper triangle read) is overwritten by the eigenvectors,
 * B by its Cholesky factor, w[n] gets all eigenvalues, *info as LAPACK. */
int imago_zhegv(imago_ctx *ctx, int n, double *A, double *B, double *w,
                int *info);
/* solveDSYGV (interfaces.F90:152-200), real symmetric. */
int imago_dsygv(imago_ctx *ctx, int n, double *A, double *B, double *w,
                int *info);

/* ---- density of states (post-SCF) ------------------------------------ *
 * computeDOS's Gaussian path (dos.F90:160-1231, kPointIntgCode /= 1) on the
 * eigenpairs of the last imago_secular and the assembled overlap:
 * energies[k * num_states + j] are the eigenvalues already shifted by the
 * highest occupied energy (shiftEnergyEigenValues, Hartree); channel[v] the
 * PDOS channel (0-based) of basis function v (the reference's pdosIndex for
 * its detailCodePDOS); the energy grid is emin + e * delta, e < num_points
 * (Hartree); sigma the broadening (Hartree).  Writes pdos[ch + num_channels
 * * e] (states / eV / cell, before the reference's electron-count factor),
 * and when not null electron_number[v] (sum over k and states at or below
 * zero of the Mulliken population times w_k / spin) and localization[j]
 * (sum over k and v of the squared population times w_k / spin). */
int imago_dos(imago_ctx *ctx, const double *energies, int spin, const int *channel,
              int num_channels, double emin, double delta, int num_points, double sigma,
              double *pdos, double *electron_number, double *localization);

/* ---- run-time choices of the batched eigensolve ------------------------ *
 * For the autotuner (imago_amd/tune.py): IMAGO_TUNE_POTRF_BLOCK, columns per
 * diagonal block of the blocked Cholesky, 1..32 (default 32);
 * IMAGO_TUNE_EIGENSOLVER, 0 = this library's tridiagonal reduction and
 * back-transform (default where the order allows it), 1 = rocSOLVER's
 * heevd / syevd; IMAGO_TUNE_CHOLESKY, 0 = this library's blocked Cholesky,
 * 1 = rocSOLVER's potrf; IMAGO_TUNE_TRIDIAG, the tridiagonal stage of this
 * library's eigensolver, 0 = its batched divide and conquer, 1 = rocSOLVER's
 * dstedc; IMAGO_TUNE_GRAPH, 1 = the divide and conquer's launches captured
 * once in a HIP graph and replayed (default; IMAGO_GRAPH=0 turns it off),
 * 0 = launched one by one.  Results do not depend on the choice beyond
 * rounding.  Setting any of them makes the next imago_secular factor the
 * overlap again. */
#define IMAGO_TUNE_POTRF_BLOCK 1
#define IMAGO_TUNE_EIGENSOLVER 2
#define IMAGO_TUNE_CHOLESKY 3
#define IMAGO_TUNE_TRIDIAG 4
#define IMAGO_TUNE_GRAPH 5
int imago_tune(imago_ctx *ctx, int param, int value);

/* The tridiagonal stage of the eigensolver on its own: eigenvalues
 * (ascending, w[b * n + i]) and eigenvectors (column-major n x n per
 * problem, Z[b * n * n + ...]) of nb real symmetric tridiagonals of order
 * n <= 2048, diagonals d[b * n + i], off-diagonals e[b * (n - 1) + i]; host
 * arrays.  No counterpart in the reference's interface: it is the step
 * inside the LAPACK zhegv / dsygv that solveZHEGV / solveDSYGV call
 * (interfaces.F90:105, 185), this library's batched divide and conquer
 * (the default of imago_secular's eigensolve; IMAGO_TUNE_TRIDIAG = 1 or
 * IMAGO_STEDC=rocsolver selects rocSOLVER's dstedc there instead).
 * imago_last_timing stage 17 gives its device time. */
int imago_tridiag_eig(imago_ctx *ctx, int n, int nb, const double *d,
                      const double *e, double *w, double *Z);

/* ---- timing ----------------------------------------------------------- *
 * Milliseconds of the last launch of the named stage measured with HIP
 * events on the library's stream: 0 = last imago_assemble, 1 = H build of
 * the last imago_secular, 2 = its eigensolve, 3 = last density+traces,
 * 4 = the EP panel store of the last IMAGO_EP assembly, 5 = the EP
 * screening and primitive blocks of the last EP assembly that built them
 * (0 when they were already built for the system), 6 = host wall time of
 * the last imago_set_system, 7 = its lattice-image / candidate part,
 * 8 = the EP sum of the last direct H build, 9 = the EP traces of the last
 * direct imago_density_traces, 10 = the cooperative tridiagonal reduction
 * of the last eigensolve (0 when rocSOLVER's eigensolver ran), 11 = the
 * kernels of the last imago_elecstat, 12 = those of the last
 * imago_residual_q (both without the copies to the host), 13 = the
 * reciprocal-space sum kernel of the last imago_residual_q, 14 = the
 * device work of the last imago_xc_overlap, 15 = its GEMMs, 16 = the
 * device work of the last imago_xc_potential (without the first call's
 * operator evaluation), 17 = the device time of the last
 * imago_tridiag_eig; of the last eigensolve: 18 = the Cholesky transform
 * U^-H H U^-1, 19 = the tridiagonal eigensolve (this library's divide and
 * conquer or rocSOLVER's dstedc); of the last eigenvector columns finished
 * (by the traces or imago_get_eigenvectors): 20 = the back-transform,
 * 21 = the solve X = U^-1 Y; 22 = the compact-WY factors of the last
 * eigensolve's reflectors. */
int imago_last_timing(imago_ctx *ctx, int stage, double *ms);

/* Sizes of the EP screening of the current system (after an IMAGO_EP or
 * IMAGO_EP_DIRECT assembly): screened (image, alpha pair, candidate)
 * tuples, (tuple, exponent) primitive items, doubles per stored primitive
 * block, and (IMAGO_EP_DIRECT, else 0) doubles of the H-sum run partial
 * blocks written per H build.  Any output may be NULL.  For measurement
 * (bytes moved per launch). */
int imago_ep_counts(imago_ctx *ctx, long *tuples, long *items, int *block_doubles,
                    long *partial_doubles);

/* ---- electrostatics of the potential-side Gaussians -------------------- *
 * Set-up stage of makeSCFPot's electrostatics (src/imago/elecStat.f90),
 * for the system of the last imago_set_system.  log_elec_thresh is
 * -ln(electroConvg) (logElecThresh, lattice.f90:294); the real-space cells
 * are the system's.  potDim terms are ordered as the EP terms (potential
 * type in the order of its first site, then alpha).
 *
 * neutralAndNuclearQPot (elecStat.f90:59-664): pot_alpha_overlap,
 * non_local_neut_q_pot and local_neut_q_pot are potDim x potDim,
 * column-major, exactly the reference's arrays (element (r, c) sums the
 * pairs whose site-j term is r and site-i term is c; not symmetric);
 * non_local_nuc_q_pot and local_nuc_q_pot are [potDim].  Every element is
 * written. */
int imago_elecstat(imago_ctx *ctx, double log_elec_thresh, double *pot_alpha_overlap,
                   double *non_local_neut_q_pot, double *local_neut_q_pot,
                   double *non_local_nuc_q_pot, double *local_nuc_q_pot);

/* residualQ (elecStat.f90:665-1043): non_local_residual_q is numPotTypes x
 * potDim, column-major (row = the ordinal of site j's type in site order,
 * column = the term of site i), as the reference's nonLocalResidualQ.
 * The reciprocal cells are the caller's (initializeLattice's
 * cellDimsRecip [num_cells_recip][3] and cellSizesRecip, |G|^2 ascending
 * with cell 0 the origin; lattice.f90:386-420).  IMAGO_E_ARG when a
 * potential type has more alphas than the first type of its element run
 * (the reference's tables would hold stale rows). */
int imago_residual_q(imago_ctx *ctx, double log_elec_thresh, int num_cells_recip,
                     const double *cell_dims_recip, const double *cell_sizes_recip,
                     double *non_local_residual_q);

/* ---- exchange-correlation mesh (src/imago/exchCorr.f90) ----------------- *
 * imago_xc_mesh is getECMeshParameters (exchCorr.f90:63-342) with
 * makeSampleVectors (:920-980): the angular sample vectors, the largest
 * radius of every (site, vector) ray and the ray points of every
 * potential site, kept on the device for the two calls below.
 * covalent_radius[numPotTypes] is potTypes(:)%covalentRadius; a site whose
 * type has a radius below 1e-8 gets no mesh, as in the reference.  Returns
 * maxNumRayPoints. */
int imago_xc_mesh(imago_ctx *ctx, const imago_xc_mesh_params *params, double log_elec_thresh,
                  const double *covalent_radius, int *max_num_ray_points);

/* The mesh of one potential site (0-based) as makeECMeshAndOverlap
 * (exchCorr.f90:345-875) writes it: *num_ray_points, radial_weight
 * [maxNumRayPoints] and exch_rho_op, potDim x maxNumRayPoints x numOp
 * column-major (numOp = 1 for LDA, 10 for GGA: the reference's
 * exchRhoOp(potDim, maxNumRayPoints, numOpValues)); entries beyond the
 * site's points are 0. */
int imago_xc_rho_op(imago_ctx *ctx, int site, int gga, int *num_ray_points,
                    double *radial_weight, double *exch_rho_op);

/* exchCorrOverlap (exchCorr.f90:805-810): potDim x potDim column-major,
 * the sum over every site's ray points of radialWeight * op op^T of the
 * LDA operator, upper triangle (row <= column) filled as the reference
 * fills it, lower triangle 0. */
int imago_xc_overlap(imago_ctx *ctx, double *exch_corr_overlap);

/* The exchange-correlation terms of makeSCFPot on the mesh, before its
 * dposvx solve (potentialUpdate.F90:833-1008), for the Wigner functional
 * (xcCode 100 / 300), spin 1, LDA: at every ray point the total and core
 * densities op . rho (negative values set to 0); then exch_corr_pot
 * [3][potDim] = sum over the points of radialWeight * (potential, energy
 * density of the total, energy density of the core) * op.  rho_total and
 * rho_core (may be NULL: 0) are the fitted charge coefficients
 * generalRho(:, 1) and (:, 3) [potDim].  The first call after
 * imago_xc_mesh evaluates the operator at every ray point and keeps it on
 * the device (about potDim/6 x ray points doubles); later calls (one per
 * SCF iteration) reuse it. */
int imago_xc_potential(imago_ctx *ctx, const double *rho_total, const double *rho_core,
                       double *exch_corr_pot);

/* ---- potential-side solves ---------------------------------------------- *
 * solveDPOSVX (interfaces.F90:236-293) with one of the potential-side
 * overlaps, which stay on the device after they are built:
 * IMAGO_SOLVE_POT_OVERLAP, potAlphaOverlap of the last imago_elecstat (the
 * charge fit and the electrostatic potential of makeSCFPot,
 * potentialUpdate.F90:236, 786), or IMAGO_SOLVE_XC_OVERLAP, exchCorrOverlap of
 * the last imago_xc_overlap (:1010).  Upper triangle, equilibrated as
 * dposvx('E') does; on the first solve the Cholesky factor and from it the
 * inverse (dpotri) are made and kept (the matrices do not change between SCF
 * iterations), so a solve is a product with the inverse and two steps of
 * iterative refinement against the equilibrated matrix.  b is potDim x nrhs column-major, overwritten by the
 * solution; *info as dposvx (i > 0: not positive definite, with
 * IMAGO_E_NOTPD). */
#define IMAGO_SOLVE_POT_OVERLAP 1
#define IMAGO_SOLVE_XC_OVERLAP 2
int imago_pot_solve(imago_ctx *ctx, int matrix, int nrhs, double *b, int *info);

/* Sizes of the exchange-correlation mesh (after imago_xc_mesh): all ray
 * points of all sites, and the flops of the GEMMs of the last
 * imago_xc_overlap (0 before one).  Either output may be NULL.  For
 * measurement. */
int imago_xc_counts(imago_ctx *ctx, long *ray_points, double *gemm_flops);

/* ---- the SCF iteration (mainSCF, src/imago/imago.F90:488-523) ----------- *
 * One iteration is imago_secular -> imago_populate -> imago_density_traces
 * -> imago_scf_update, as the reference's secularEqnSCF -> populateStates
 * -> makeValenceRho -> makeSCFPot. */

/* populateStandard (src/imago/populate.F90:190-485) with its merge-sort
 * order (sortSubs.f90:18-177): spin 1, no thermal smearing, no excited core
 * state.  eigenvalues and populations are [num_kpoints][num_states]
 * (k-major, as imago_secular returns them); kpoint_weights sum to 2.
 * Returns IMAGO_E_ARG when the states cannot hold num_electrons.  Any of
 * the three scalar outputs may be NULL.  Host code; no context. */
int imago_populate(int num_kpoints, int num_states, const double *eigenvalues,
                   const double *kpoint_weights, double num_electrons, double *populations,
                   double *occupied_energy, double *band_gap, int *gap_kind);

/* Input of makeSCFPot: the MAIN_INPUT_DATA block of imago.dat as makeinput
 * writes it (makeinput.py:5340-5352) and NUM_ELECTRONS. */
typedef struct imago_scf_params {
    double num_electrons;     /* NUM_ELECTRONS                              */
    int feedback_level;       /* FEEDBACK_LEVEL (makeinput writes 2)        */
    double relax_factor;      /* RELAXATION_FACTOR (makeinput writes 0.2)   */
    double convg_test;        /* CONVERGENCE_TEST (makeinputrc: 1e-4)       */
    int xc_code;              /* XC_CODE: 100 = Wigner, spin 1, LDA, no rel. */
    double log_elec_thresh;   /* -ln(electroConvg) for the electrostatics   */
} imago_scf_params;

/* One row of the reference's per-iteration records: fort.7 (iteration,
 * electron fit error, convergence, total energy; potentialUpdate.F90:
 * 1806-1812) and fort.14 (kinetic, electrostatic, exchange-correlation,
 * total; :1825-1827). */
typedef struct imago_scf_record {
    int iteration;            /* currIteration of this update (1-based)     */
    int converged;            /* testableDelta < convgTest (:1839)          */
    double electron_diff;     /* numElectrons - fitted valence charge       */
    double testable_delta;    /* largest site RMS of the potential change   */
    double kinetic_energy;
    double elec_stat_energy;
    double exch_corr_energy;
    double total_energy;      /* Hartree                                    */
} imago_scf_record;

/* Set-up of makeSCFPot for the system of the last imago_set_system: the
 * electrostatic matrices of neutralAndNuclearQPot and residualQ
 * (imago_elecstat, imago_residual_q; reciprocal cells as for
 * imago_residual_q), the exchange-correlation mesh, its overlap and
 * operator (imago_xc_mesh with mesh and covalent_radius[numPotTypes],
 * imago_xc_overlap), all kept on the device; intgConsts (potential.f90:
 * 150-160).  Resets the iteration count to 1 and the mixing history.
 * IMAGO_E_ARG for an xc_code other than 100. */
int imago_scf_setup(imago_ctx *ctx, const imago_scf_params *params,
                    const imago_xc_mesh_params *mesh, const double *covalent_radius,
                    int num_cells_recip, const double *cell_dims_recip,
                    const double *cell_sizes_recip);

/* makeSCFPot (src/imago/potentialUpdate.F90:53-1866) for spin 1, no core
 * states, non-relativistic LDA (Wigner): the charge fit of pot_rho
 * (potRho of imago_density_traces) against potAlphaOverlap, the
 * electron-count correction, the neutral / residual / nuclear charge split,
 * the electrostatic potential and energy, the exchange-correlation
 * potential and energy on the ray mesh, the total energy, the RMS
 * real-space potential change per site, and the mixed potential
 * (relaxation with the two-step feedback of :1604-1666).  traces3 =
 * (Tr rho S, Tr rho NP, Tr rho KE) of imago_density_traces.  pot_coeffs
 * [potDim] holds the coefficients used in this iteration on entry and the
 * next iteration's on return.  rho_coeffs [2][potDim] (may be NULL) gets
 * the fitted total and valence charge coefficients (generalRho(:, 1:2),
 * the scfV.dat columns rhoT and rhoV).  The iteration count advances
 * unless the record says converged, as currIteration does.  IMAGO_E_SOLVE
 * when a dposvx solve fails and when the total energy is positive (the
 * reference stops in both cases). */
int imago_scf_update(imago_ctx *ctx, const double *traces3, const double *pot_rho,
                     double *pot_coeffs, double *rho_coeffs, imago_scf_record *record);

/* ---- the learned initial potential --------------------------------------- *
 * A per-site MLP predicts, from the site's environment, the correction of
 * the atomic-superposition coefficients that scfV.dat carries (the input of
 * initPotCoeffs, src/imago/potential.f90:433-564; makeinput.py:5805-5941
 * writes the superposition).  The reference has no learned guess; these
 * entry points produce the coefficient vector it reads. */

/* Environment descriptors of num_sites sites: pos [num_sites][3] (bohr),
 * weight [num_sites] (species weight of each site as a neighbour), cells
 * [num_cells][3] the lattice translations to sum over (the origin
 * included), 32 Gaussian radial channels exp(-eta (r - mu_k)^2) with mu_k
 * from r_min to cutoff, the cosine cut-off at cutoff; per channel the l = 0,
 * 1, 2 rotation invariants.  desc [num_sites][96]: channel k at k (l = 0),
 * 32 + k (l = 1), 64 + k (l = 2). */
int imago_site_descriptors(imago_ctx *ctx, int num_sites, const double *pos,
                           const double *weight, int num_cells, const double *cells,
                           double cutoff, double r_min, double eta, double *desc);

typedef struct imago_mlp imago_mlp;
/* An MLP of num_layers layers, widths[0..num_layers] (tanh on the hidden
 * layers, linear output), Glorot-uniform weights from seed, zero biases; on
 * ctx's device and stream.  Parameters are packed per layer as W_l
 * [w_{l+1}][w_l] row-major, then b_l. */
int imago_mlp_create(imago_ctx *ctx, int num_layers, const int *widths, unsigned seed,
                     imago_mlp **mlp);
int imago_mlp_destroy(imago_mlp *mlp);
int imago_mlp_num_params(imago_mlp *mlp, long *num_params);
/* Copy the packed parameters out (get may be NULL) and/or in (set may be
 * NULL; setting clears Adam's moments). */
int imago_mlp_params(imago_mlp *mlp, float *get, const float *set);
/* y [batch][w_L] = MLP(x [batch][w_0]), FP32. */
int imago_mlp_forward(imago_mlp *mlp, int batch, const float *x, float *y);
/* Mean squared error against target [batch][w_L] and its gradient with
 * respect to the packed parameters (grad may be NULL; it stays on the
 * device for imago_mlp_adam_step). */
int imago_mlp_gradient(imago_mlp *mlp, int batch, const float *x, const float *target,
                       double *loss, float *grad);
/* One Adam step (Kingma and Ba) with the last gradient. */
int imago_mlp_adam_step(imago_mlp *mlp, float lr, float beta1, float beta2, float eps);
/* epochs of Adam over shuffled minibatches of the num_samples pairs (x,
 * target), resident on the device; epoch_loss[epochs] (may be NULL) gets
 * each epoch's mean loss. */
int imago_mlp_train(imago_mlp *mlp, int num_samples, const float *x, const float *target,
                    int epochs, int minibatch, float lr, float beta1, float beta2, float eps,
                    unsigned seed, double *epoch_loss);

/* imago_mlp_forward without any allocation: x_dev [batch][w_0] and y_dev
 * [batch][w_L] are device pointers on the MLP's device; the hidden
 * activations use a 4 MiB device array of the library (2 x the widest hidden
 * width x batch floats must fit: batch <= 4096 for 128-wide layers), no
 * hipMalloc, hipFree or memset.  The work is enqueued on the context's
 * stream and the call returns without waiting; calls that overlap on one
 * device share the array and race.  No counterpart in the reference: it is
 * for callers that cannot allocate (HIP graph capture, serving loops). */
int imago_mlp_forward_noalloc(imago_mlp *mlp, int batch, const float *x_dev, float *y_dev);

#ifdef __cplusplus
}
#endif
#endif /* IMAGO_HIP_H */
