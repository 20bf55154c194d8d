! scf_driver.f90 - a Fortran host that runs the reference's setupSCF
! integral stage and its mainSCF loop through O_ImagoHIP, the way a patched
! src/imago would:
!   setupSCF    (imago.F90:254-299)   -> imago_set_system, imago_assemble
!                                        (OL, KE, NP, EP), imago_get_packed
!   mainSCF     (imago.F90:488-523)   -> imago_secular, imago_populate,
!                                        imago_density_traces, imago_scf_update
!   fort.7 row  (potentialUpdate.F90:1806-1808), fort.14 row (:1825-1827)
!
! The module-state arrays the reference keeps (atomic and potential types,
! sites, the lattice, k-points) are read from a stream file (argument 1)
! into allocatable module-style arrays and handed to the library through
! imago_system_desc with c_loc, as a host would hand its own arrays.
! Writes matrices.bin (packed OL, KE, NP and the first potential term at
! the first k-point, then the first iteration's eigenvalues), fort.7 and
! fort.14 in the working directory; errors go to unit 20, then stop.
program scf_driver
   use iso_c_binding
   use O_ImagoHIP
   implicit none
   integer(c_int) :: hdr(16)
   real(c_double) :: dhdr(20)
   integer(c_int), allocatable, target :: typeElement(:), typeNumOrbAlphas(:), &
         & typeNumStates(:), typeLmIndex(:), atomType(:), potTypeElement(:), &
         & potTypeNumAlphas(:), potTypeFirstSite(:), potTypeMultiplicity(:), potSiteType(:)
   real(c_double), allocatable, target :: typeAlphas(:), typeBasis(:), atomPos(:), &
         & potTypeAlphas(:), potTypeNucCharge(:), potTypeNucAlpha(:), potSitePos(:), &
         & cellDims(:), cellSizes(:), kPoints(:), kPointWeight(:), cellDimsRecip(:), &
         & cellSizesRecip(:), covalentRadius(:), potCoeffs(:)
   real(c_double), allocatable :: packed(:), eigenValues(:,:), populations(:,:), &
         & potRho(:), rhoCoeffs(:)
   type(imago_system_desc) :: sys
   type(imago_scf_params) :: params
   type(imago_xc_mesh_params) :: mesh
   type(imago_scf_record) :: rec
   integer(c_int) :: nt, maxa, maxs, na, npt, maxpa, nps, nc, nk, gam, ncr, maxIter
   integer(c_int) :: valeDim, potDim, numStates, gapKind, status, k, iter, dim1
   integer(c_int), allocatable :: info(:)
   real(c_double) :: traces(3), occupiedEnergy, bandGap, energyTol, lastEnergy
   character(len=512) :: path
   logical :: haveLast

   call get_command_argument(1, path)
   open (unit=30, file=trim(path), access='stream', form='unformatted', status='old')
   read (30) hdr
   read (30) dhdr
   nt = hdr(1); maxa = hdr(2); maxs = hdr(3); na = hdr(4); npt = hdr(5); maxpa = hdr(6)
   nps = hdr(7); nc = hdr(8); nk = hdr(9); gam = hdr(10); ncr = hdr(11); maxIter = hdr(12)
   allocate (typeElement(nt), typeNumOrbAlphas(4*nt), typeAlphas(nt*maxa), &
         & typeNumStates(nt), typeBasis(nt*maxs*maxa), typeLmIndex(nt*maxs), &
         & atomType(na), atomPos(3*na), potTypeElement(npt), potTypeNumAlphas(npt), &
         & potTypeAlphas(npt*maxpa), potTypeNucCharge(npt), potTypeNucAlpha(npt), &
         & potTypeFirstSite(npt), potTypeMultiplicity(npt), potSiteType(nps), &
         & potSitePos(3*nps), cellDims(3*nc), cellSizes(nc), kPoints(3*nk), &
         & kPointWeight(nk), cellDimsRecip(3*ncr), cellSizesRecip(ncr), covalentRadius(npt))
   read (30) typeElement, typeNumOrbAlphas, typeAlphas, typeNumStates, typeBasis, typeLmIndex
   read (30) atomType, atomPos
   read (30) potTypeElement, potTypeNumAlphas, potTypeAlphas, potTypeNucCharge, &
         & potTypeNucAlpha, potTypeFirstSite, potTypeMultiplicity
   read (30) potSiteType, potSitePos, cellDims, cellSizes, kPoints, kPointWeight
   read (30) cellDimsRecip, cellSizesRecip, covalentRadius
   potDim = sum(potTypeNumAlphas)
   allocate (potCoeffs(potDim))
   read (30) potCoeffs
   close (30)

   sys%num_atom_types = nt
   sys%max_alphas = maxa
   sys%max_states = maxs
   sys%type_element = c_loc(typeElement)
   sys%type_num_orb_alphas = c_loc(typeNumOrbAlphas)
   sys%type_alphas = c_loc(typeAlphas)
   sys%type_num_states = c_loc(typeNumStates)
   sys%type_basis = c_loc(typeBasis)
   sys%type_lm_index = c_loc(typeLmIndex)
   sys%num_atoms = na
   sys%atom_type = c_loc(atomType)
   sys%atom_pos = c_loc(atomPos)
   sys%num_pot_types = npt
   sys%max_pot_alphas = maxpa
   sys%pot_type_element = c_loc(potTypeElement)
   sys%pot_type_num_alphas = c_loc(potTypeNumAlphas)
   sys%pot_type_alphas = c_loc(potTypeAlphas)
   sys%pot_type_nuc_charge = c_loc(potTypeNucCharge)
   sys%pot_type_nuc_alpha = c_loc(potTypeNucAlpha)
   sys%pot_type_first_site = c_loc(potTypeFirstSite)
   sys%pot_type_multiplicity = c_loc(potTypeMultiplicity)
   sys%num_pot_sites = nps
   sys%pot_site_type = c_loc(potSiteType)
   sys%pot_site_pos = c_loc(potSitePos)
   sys%real_vectors = dhdr(12:20)
   sys%num_cells = nc
   sys%cell_dims = c_loc(cellDims)
   sys%cell_sizes = c_loc(cellSizes)
   sys%num_kpoints = nk
   sys%kpoints = c_loc(kPoints)
   sys%kpoint_weights = c_loc(kPointWeight)
   sys%gamma = gam
   sys%log_basis_fn_thresh = dhdr(1)
   sys%type_num_core_states = c_null_ptr

   ! setupSCF: the integrals
   call initImagoHIP(0)
   call hip_check(imago_set_system(imagoCtx, sys, valeDim, potDim), 'imago_set_system')
   call hip_check(imago_assemble(imagoCtx, IMAGO_OL), 'gaussOverlapOL')
   call hip_check(imago_assemble(imagoCtx, IMAGO_KE), 'gaussOverlapKE')
   call hip_check(imago_assemble(imagoCtx, IMAGO_NP), 'gaussOverlapNP')
   call hip_check(imago_assemble(imagoCtx, IMAGO_EP), 'elecPotGaussOverlap')
   dim1 = 2
   if (gam == 1) dim1 = 1
   allocate (packed(dim1 * valeDim * (valeDim + 1) / 2))
   open (unit=31, file='matrices.bin', access='stream', form='unformatted', status='replace')
   call hip_check(imago_get_packed(imagoCtx, IMAGO_OL, 0, 0, packed), 'get OL')
   write (31) packed
   call hip_check(imago_get_packed(imagoCtx, IMAGO_KE, 0, 0, packed), 'get KE')
   write (31) packed
   call hip_check(imago_get_packed(imagoCtx, IMAGO_NP, 0, 0, packed), 'get NP')
   write (31) packed
   call hip_check(imago_get_packed(imagoCtx, IMAGO_EP, 0, 0, packed), 'get EP')
   write (31) packed

   ! makeSCFPot's set-up
   params%num_electrons = dhdr(3)
   params%feedback_level = hdr(15)
   params%relax_factor = dhdr(4)
   params%convg_test = dhdr(5)
   params%xc_code = hdr(14)
   params%log_elec_thresh = dhdr(2)
   energyTol = dhdr(6)
   mesh%num_sample_vectors = hdr(13)
   mesh%weight_in = dhdr(7)
   mesh%weight_out = dhdr(8)
   mesh%r_sample_in = dhdr(9)
   mesh%r_sample_out = dhdr(10)
   mesh%r_sample_space = dhdr(11)
   call hip_check(imago_scf_setup(imagoCtx, params, mesh, covalentRadius, ncr, &
         & cellDimsRecip, cellSizesRecip), 'imago_scf_setup')

   ! mainSCF
   numStates = valeDim
   allocate (eigenValues(numStates, nk), populations(numStates, nk), info(nk), &
         & potRho(potDim), rhoCoeffs(2 * potDim))
   open (unit=7, file='fort.7', status='replace', form='formatted')
   open (unit=14, file='fort.14', status='replace', form='formatted')
   ! the reference's headers (imago.F90:456-468) on one record each
   write (7,'(a)') " Iter.  Occ._Energy     El._Error   " // &
         & "Convergence    Total_Energy(au)      Mag._Mom.          " // &
         & "Gap(au)  Gap_Kind"
   write (14,'(a)') " Iter. Kinetic        ElecStat       ExchCorr       Total"
   haveLast = .false.
   lastEnergy = 0.0d0
   do iter = 1, maxIter
      status = imago_secular(imagoCtx, potCoeffs, numStates, eigenValues, info)
      if (status /= IMAGO_OK .and. all(info == 0)) call hip_check(status, 'secularEqnSCF')
      do k = 1, nk
         if (info(k) /= 0) then
            write (20,*) 'ZHEGV error code=', info(k), ' at k-point', k
            stop
         endif
      enddo
      if (iter == 1) write (31) eigenValues
      call hip_check(imago_populate(nk, numStates, eigenValues, kPointWeight, &
            & params%num_electrons, populations, occupiedEnergy, bandGap, gapKind), &
            & 'populateStates')
      call hip_check(imago_density_traces(imagoCtx, populations, traces, potRho), &
            & 'makeValenceRho')
      call hip_check(imago_scf_update(imagoCtx, traces, potRho, potCoeffs, rhoCoeffs, rec), &
            & 'makeSCFPot')
      write (7,fmt="(i5,3(1x,f13.8),3(1x,f17.8),1x,i3)") rec%iteration, &
            & occupiedEnergy, rec%electron_diff, rec%testable_delta, &
            & rec%total_energy, 0.0d0, bandGap, gapKind
      write (14,fmt="(i5,4e15.7)") rec%iteration, rec%kinetic_energy, &
            & rec%elec_stat_energy, rec%exch_corr_energy, rec%total_energy
      if (rec%converged == 1 .and. haveLast) then
         if (abs(rec%total_energy - lastEnergy) < energyTol) exit
      endif
      lastEnergy = rec%total_energy
      haveLast = .true.
   enddo
   close (7)
   close (14)
   close (31)
   status = imago_ctx_destroy(imagoCtx)
end program scf_driver
