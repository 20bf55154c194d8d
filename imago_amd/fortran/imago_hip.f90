! imago_hip.f90 - iso_c_binding interface to libimago_hip.so for the
! reference's Fortran host (UMKC-CPG/imago, src/imago).
!
! The bind(C) interfaces mirror include/imago_hip.h one to one.  The module
! procedures below them keep the argument lists of the reference routines
! they replace, and keep the reference's error convention (a message on
! unit 20, then stop):
!   solveZHEGV_hip  <- O_LAPACKZHEGV::solveZHEGV  (interfaces.F90:73-120)
!   solveDSYGV_hip  <- O_LAPACKDSYGV::solveDSYGV  (interfaces.F90:152-200)
!   hip_check       <- the "stop" after each failing call
module O_ImagoHIP
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: IMAGO_OK = 0
   integer(c_int), parameter :: IMAGO_SOLVE_POT_OVERLAP = 1, IMAGO_SOLVE_XC_OVERLAP = 2
   integer(c_int), parameter :: IMAGO_OL = 1, IMAGO_KE = 2, IMAGO_NP = 3, &
         & IMAGO_EP = 4, IMAGO_EP_DIRECT = 5

   ! imago_system_desc (include/imago_hip.h), field for field
   type, bind(C) :: imago_system_desc
      integer(c_int) :: num_atom_types, max_alphas, max_states
      type(c_ptr) :: type_element, type_num_orb_alphas, type_alphas
      type(c_ptr) :: type_num_states, type_basis, type_lm_index
      integer(c_int) :: num_atoms
      type(c_ptr) :: atom_type, atom_pos
      integer(c_int) :: num_pot_types, max_pot_alphas
      type(c_ptr) :: pot_type_element, pot_type_num_alphas, pot_type_alphas
      type(c_ptr) :: pot_type_nuc_charge, pot_type_nuc_alpha
      type(c_ptr) :: pot_type_first_site, pot_type_multiplicity
      integer(c_int) :: num_pot_sites
      type(c_ptr) :: pot_site_type, pot_site_pos
      real(c_double) :: real_vectors(9)
      integer(c_int) :: num_cells
      type(c_ptr) :: cell_dims, cell_sizes
      integer(c_int) :: num_kpoints
      type(c_ptr) :: kpoints, kpoint_weights
      integer(c_int) :: gamma
      real(c_double) :: log_basis_fn_thresh
      type(c_ptr) :: type_num_core_states
   end type imago_system_desc

   ! readExchCorrMeshParameters input block (exchCorr.f90:876-913)
   type, bind(C) :: imago_xc_mesh_params
      integer(c_int) :: num_sample_vectors
      real(c_double) :: weight_in, weight_out
      real(c_double) :: r_sample_in, r_sample_out, r_sample_space
   end type imago_xc_mesh_params

   ! imago_scf_params: the MAIN_INPUT_DATA block of imago.dat
   type, bind(C) :: imago_scf_params
      real(c_double) :: num_electrons
      integer(c_int) :: feedback_level
      real(c_double) :: relax_factor, convg_test
      integer(c_int) :: xc_code
      real(c_double) :: log_elec_thresh
   end type imago_scf_params

   ! imago_scf_record: one fort.7 / fort.14 row
   type, bind(C) :: imago_scf_record
      integer(c_int) :: iteration, converged
      real(c_double) :: electron_diff, testable_delta
      real(c_double) :: kinetic_energy, elec_stat_energy, exch_corr_energy, total_energy
   end type imago_scf_record

   type(c_ptr), save :: imagoCtx = c_null_ptr

   interface
      integer(c_int) function imago_ctx_create(device, ctx) &
            & bind(C, name="imago_ctx_create")
         import :: c_int, c_ptr
         integer(c_int), value :: device
         type(c_ptr) :: ctx
      end function imago_ctx_create

      integer(c_int) function imago_ctx_destroy(ctx) &
            & bind(C, name="imago_ctx_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function imago_ctx_destroy

      type(c_ptr) function imago_last_error() bind(C, name="imago_last_error")
         import :: c_ptr
      end function imago_last_error

      integer(c_int) function imago_set_system(ctx, sys, valeDim, potDim) &
            & bind(C, name="imago_set_system")
         import :: c_int, c_ptr, imago_system_desc
         type(c_ptr), value :: ctx
         type(imago_system_desc), intent(in) :: sys
         integer(c_int), intent(out) :: valeDim, potDim
      end function imago_set_system

      integer(c_int) function imago_primitive_batch(ctx, kind, n, a, xyz, &
            & l1l2switch, sh) bind(C, name="imago_primitive_batch")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: kind, n
         real(c_double), intent(in) :: a(*), xyz(*)
         integer(c_int), intent(in) :: l1l2switch(*)
         real(c_double), intent(out) :: sh(*)
      end function imago_primitive_batch

      integer(c_int) function imago_boys_batch(ctx, n, T, F7) &
            & bind(C, name="imago_boys_batch")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n
         real(c_double), intent(in) :: T(*)
         real(c_double), intent(out) :: F7(*)
      end function imago_boys_batch

      integer(c_int) function imago_assemble(ctx, kind) &
            & bind(C, name="imago_assemble")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: kind
      end function imago_assemble

      integer(c_int) function imago_get_packed(ctx, kind, term, kpoint, out) &
            & bind(C, name="imago_get_packed")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: kind, term, kpoint
         real(c_double), intent(out) :: out(*)
      end function imago_get_packed

      ! the matrix a resumed run reads from its integral file (readPackedMatrix)
      integer(c_int) function imago_set_packed(ctx, kind, term, kpoint, in) &
            & bind(C, name="imago_set_packed")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: kind, term, kpoint
         real(c_double), intent(in) :: in(*)
      end function imago_set_packed

      integer(c_int) function imago_secular(ctx, potCoeffs, numStates, &
            & eigenValues, info) bind(C, name="imago_secular")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: potCoeffs(*)
         integer(c_int), value :: numStates
         real(c_double), intent(out) :: eigenValues(*)
         integer(c_int), intent(out) :: info(*)
      end function imago_secular

      integer(c_int) function imago_get_eigenvectors(ctx, kpoint, out) &
            & bind(C, name="imago_get_eigenvectors")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: kpoint
         real(c_double), intent(out) :: out(*)
      end function imago_get_eigenvectors

      integer(c_int) function imago_density_traces(ctx, occupations, traces3, &
            & potRho) bind(C, name="imago_density_traces")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: occupations(*)
         real(c_double), intent(out) :: traces3(3), potRho(*)
      end function imago_density_traces

      integer(c_int) function imago_zhegv(ctx, n, A, B, w, info) &
            & bind(C, name="imago_zhegv")
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: n
         complex(c_double_complex), intent(inout) :: A(n,n), B(n,n)
         real(c_double), intent(out) :: w(n)
         integer(c_int), intent(out) :: info
      end function imago_zhegv

      integer(c_int) function imago_dsygv(ctx, n, A, B, w, info) &
            & bind(C, name="imago_dsygv")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n
         real(c_double), intent(inout) :: A(n,n), B(n,n)
         real(c_double), intent(out) :: w(n)
         integer(c_int), intent(out) :: info
      end function imago_dsygv

      integer(c_int) function imago_last_timing(ctx, stage, ms) &
            & bind(C, name="imago_last_timing")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: stage
         real(c_double), intent(out) :: ms
      end function imago_last_timing

      integer(c_int) function imago_ep_counts(ctx, tuples, items, block_doubles, &
            & partial_doubles) bind(C, name="imago_ep_counts")
         import :: c_int, c_ptr, c_long
         type(c_ptr), value :: ctx
         integer(c_long), intent(out) :: tuples, items, partial_doubles
         integer(c_int), intent(out) :: block_doubles
      end function imago_ep_counts

      ! neutralAndNuclearQPot (elecStat.f90:59-664): the arrays the
      ! reference writes to its electrostatics HDF5 datasets
      integer(c_int) function imago_elecstat(ctx, logElecThresh, potAlphaOverlap, &
            & nonLocalNeutQPot, localNeutQPot, nonLocalNucQPot, localNucQPot) &
            & bind(C, name="imago_elecstat")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: logElecThresh
         real(c_double), intent(out) :: potAlphaOverlap(*), nonLocalNeutQPot(*), &
               & localNeutQPot(*), nonLocalNucQPot(*), localNucQPot(*)
      end function imago_elecstat

      ! residualQ (elecStat.f90:665-1043), with O_Lattice's reciprocal cells
      integer(c_int) function imago_residual_q(ctx, logElecThresh, numCellsRecip, &
            & cellDimsRecip, cellSizesRecip, nonLocalResidualQ) &
            & bind(C, name="imago_residual_q")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: logElecThresh
         integer(c_int), value :: numCellsRecip
         real(c_double), intent(in) :: cellDimsRecip(3,*), cellSizesRecip(*)
         real(c_double), intent(out) :: nonLocalResidualQ(*)
      end function imago_residual_q

      ! getECMeshParameters + makeSampleVectors (exchCorr.f90:63-342, 920-980)
      integer(c_int) function imago_xc_mesh(ctx, params, logElecThresh, covalentRadius, &
            & maxNumRayPoints) bind(C, name="imago_xc_mesh")
         import :: c_int, c_ptr, c_double, imago_xc_mesh_params
         type(c_ptr), value :: ctx
         type(imago_xc_mesh_params), intent(in) :: params
         real(c_double), value :: logElecThresh
         real(c_double), intent(in) :: covalentRadius(*)
         integer(c_int), intent(out) :: maxNumRayPoints
      end function imago_xc_mesh

      ! one site's numRayPoints, radialWeight and exchRhoOp (exchCorr.f90:492-790)
      integer(c_int) function imago_xc_rho_op(ctx, site, gga, numRayPoints, radialWeight, &
            & exchRhoOp) bind(C, name="imago_xc_rho_op")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: site, gga
         integer(c_int), intent(out) :: numRayPoints
         real(c_double), intent(out) :: radialWeight(*), exchRhoOp(*)
      end function imago_xc_rho_op

      ! exchCorrOverlap (exchCorr.f90:805-810)
      integer(c_int) function imago_xc_overlap(ctx, exchCorrOverlap) &
            & bind(C, name="imago_xc_overlap")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: exchCorrOverlap(*)
      end function imago_xc_overlap

      ! exchCorrPot(:, 1:3) before the solve (potentialUpdate.F90:833-1008)
      integer(c_int) function imago_xc_potential(ctx, rhoTotal, rhoCore, exchCorrPot) &
            & bind(C, name="imago_xc_potential")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: rhoTotal(*), rhoCore(*)
         real(c_double), intent(out) :: exchCorrPot(*)
      end function imago_xc_potential

      ! solveDPOSVX with potAlphaOverlap (1) or exchCorrOverlap (2) on the device
      integer(c_int) function imago_pot_solve(ctx, matrix, nrhs, b, info) &
            & bind(C, name="imago_pot_solve")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: matrix, nrhs
         real(c_double), intent(inout) :: b(*)
         integer(c_int), intent(out) :: info
      end function imago_pot_solve

      ! populateStandard (populate.F90:190-485)
      integer(c_int) function imago_populate(numKPoints, numStates, eigenValues, &
            & kPointWeight, numElectrons, populations, occupiedEnergy, bandGap, gapKind) &
            & bind(C, name="imago_populate")
         import :: c_int, c_double
         integer(c_int), value :: numKPoints, numStates
         real(c_double), intent(in) :: eigenValues(*), kPointWeight(*)
         real(c_double), value :: numElectrons
         real(c_double), intent(out) :: populations(*), occupiedEnergy, bandGap
         integer(c_int), intent(out) :: gapKind
      end function imago_populate

      ! makeSCFPot's set-up (electrostatics, XC mesh, overlaps on the device)
      integer(c_int) function imago_scf_setup(ctx, params, mesh, covalentRadius, &
            & numCellsRecip, cellDimsRecip, cellSizesRecip) bind(C, name="imago_scf_setup")
         import :: c_int, c_ptr, c_double, imago_scf_params, imago_xc_mesh_params
         type(c_ptr), value :: ctx
         type(imago_scf_params), intent(in) :: params
         type(imago_xc_mesh_params), intent(in) :: mesh
         real(c_double), intent(in) :: covalentRadius(*)
         integer(c_int), value :: numCellsRecip
         real(c_double), intent(in) :: cellDimsRecip(3,*), cellSizesRecip(*)
      end function imago_scf_setup

      ! makeSCFPot (potentialUpdate.F90:53-1866): potCoeffs in and out
      integer(c_int) function imago_scf_update(ctx, traces3, potRho, potCoeffs, &
            & rhoCoeffs, record) bind(C, name="imago_scf_update")
         import :: c_int, c_ptr, c_double, imago_scf_record
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: traces3(3), potRho(*)
         real(c_double), intent(inout) :: potCoeffs(*)
         real(c_double), intent(out) :: rhoCoeffs(*)
         type(imago_scf_record), intent(out) :: record
      end function imago_scf_update

      integer(c_int) function imago_xc_counts(ctx, rayPoints, gemmFlops) &
            & bind(C, name="imago_xc_counts")
         import :: c_int, c_ptr, c_double, c_long
         type(c_ptr), value :: ctx
         integer(c_long), intent(out) :: rayPoints
         real(c_double), intent(out) :: gemmFlops
      end function imago_xc_counts
   end interface

contains

   ! Stop with the library's message on unit 20, the reference's convention.
   subroutine hip_check(status, what)
      integer(c_int), intent(in) :: status
      character(len=*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      integer :: i
      if (status == IMAGO_OK) return
      call c_f_pointer(imago_last_error(), msg, [512])
      write (20,*) what, ' failed, status=', status
      do i = 1, 512
         if (msg(i) == c_null_char) exit
         write (20,fmt="(a1)",advance="no") msg(i)
      enddo
      write (20,*)
      stop
   end subroutine hip_check

   subroutine initImagoHIP(device)
      integer, intent(in) :: device
      if (c_associated(imagoCtx)) return
      call hip_check(imago_ctx_create(int(device,c_int), imagoCtx), &
            & 'imago_ctx_create')
   end subroutine initImagoHIP

   ! Same arguments as solveZHEGV (interfaces.F90:73-120).
   subroutine solveZHEGV_hip(N, resultDim, A, B, eigenValues)
      integer, intent(in) :: N, resultDim
      complex(c_double_complex), dimension(N,N), intent(inout) :: A
      complex(c_double_complex), dimension(N,N), intent(inout) :: B
      real(c_double), dimension(resultDim), intent(out) :: eigenValues
      real(c_double), allocatable :: w(:)
      integer(c_int) :: info, status
      call initImagoHIP(0)
      allocate (w(N))
      info = 0
      status = imago_zhegv(imagoCtx, int(N,c_int), A, B, w, info)
      ! a failure of the library itself leaves info unset: report it first
      if (status /= IMAGO_OK .and. info == 0) call hip_check(status, 'imago_zhegv')
      if (info /= 0) then
         write (20,*) 'ZHEGV error code=', info
         stop
      endif
      eigenValues(:resultDim) = w(:resultDim)
      deallocate (w)
   end subroutine solveZHEGV_hip

   ! Same arguments as solveDSYGV (interfaces.F90:152-200).
   subroutine solveDSYGV_hip(valeDim, numStates, valeValeGamma, &
         & valeValeOLGamma, eigenValues)
      integer, intent(in) :: valeDim, numStates
      real(c_double), dimension(valeDim,valeDim), intent(inout) :: valeValeGamma
      real(c_double), dimension(valeDim,valeDim), intent(inout) :: &
            & valeValeOLGamma
      real(c_double), dimension(valeDim), intent(out) :: eigenValues
      real(c_double), allocatable :: w(:)
      integer(c_int) :: info, status
      call initImagoHIP(0)
      allocate (w(valeDim))
      info = 0
      status = imago_dsygv(imagoCtx, int(valeDim,c_int), valeValeGamma, &
            & valeValeOLGamma, w, info)
      ! a failure of the library itself leaves info unset: report it first
      if (status /= IMAGO_OK .and. info == 0) call hip_check(status, 'imago_dsygv')
      if (info /= 0) then
         write (20,*) 'DSYGV error code=', info
         stop
      endif
      eigenValues(:numStates) = w(:numStates)
      deallocate (w)
   end subroutine solveDSYGV_hip

end module O_ImagoHIP
