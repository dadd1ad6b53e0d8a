! cmf2_base_dropin.f90 -- replacement source for Macaw's module cmf2_base
! (cmf2_base.f90:10-345): SetupCMF2 and cmf2 keep their names and argument
! lists, so propagate_wf (propagator.f90:264,271) compiles against this
! module unchanged and its CMF2 steps run on the MI355X engine.
!
! Macaw declares its module arrays without the TARGET attribute (psidef.f90
! targets only psi, nSPF, Aux1 and Aux2; dvr.f90 and hamilton.f90 none), so
! c_loc may not be applied to them.  SetupCMF2 copies every array the
! descriptor points to into a module-local TARGET allocatable of the C type
! (int64 / double / double complex) and hands those to mg_create; the engine
! copies the tables into its own storage inside mg_create.
!
! Macaw's grid and DVR matrices are indexed by DVR-dof (indDVR); the
! descriptor's by DOF, so they are scattered (Gaussian DOFs get zeros).
!
! Gaussian (G-MCTDH) modes: dof_gwp(k) = 1 for a DOF whose grid type is
! 'gwp' (dvr.f90:88), and the width parameters GWPa(k, j, i) of the i-th
! Gaussian mode (psidef.f90:42) are scattered to the descriptor's layout
! [maxdof, nSPFmax, nmodes] of logical modes (imodGWP(i)).  The D matrices
! DD0 are not exchanged: cmf2 recomputes D0 from the SPF overlaps at the
! start of every step (cmf2_base.f90:62-72), and the engine does the same.
!
! The environment variable MACAW_GPU_PRECISION = 1 selects the mixed form
! (complex FP32 H.A / Krylov / mean-field chains, FP64 state); default 0.
module cmf2_base
   use iso_c_binding
   use macaw_gpu
   use psidef
   use dvr, only : dvrType, ngp, ngpMax, grid, d1DVR, d2DVR, xd1DVR, gridOper
   use hamilton, only : nHam, ae, lhs_el, rhs_el, iOpDof, nOpParam
   use globals
   implicit none
   private
   public :: SetupCMF2, cmf2
   type(c_ptr), save :: ctx = c_null_ptr
   type(mg_problem), save :: prob
   integer(c_int64_t), allocatable, target, save :: nmd_t(:), idof_t(:,:), ms_t(:), nspf_t(:,:), ngp_t(:), &
                                                    lhs_t(:), rhs_t(:), iop_t(:,:), gwp_t(:)
   real(c_double), allocatable, target, save :: grid_t(:,:), d1_t(:,:,:), d2_t(:,:,:), xd1_t(:,:,:), gwpa_t(:,:,:), &
                                                gop_t(:,:,:)
   complex(c_double_complex), allocatable, target, save :: ae_t(:)
   integer(c_int64_t), save :: ngmax_t = 0

contains

   subroutine SetupCMF2
      integer :: i, im, n2, nd
      integer(c_int64_t) :: precision
      character(len=8) :: env
      integer :: env_len, env_stat
      if (c_associated(ctx)) call mg_check(ctx, mg_destroy(ctx), 'SetupCMF2')
      ctx = c_null_ptr
      call copy_tables()
      prob%nstates = nstates
      prob%nmodes = nmodes
      prob%ndof = ndof
      prob%maxdof = size(iDOF, 1)
      prob%nmodeDOF = c_loc(nmd_t)
      prob%iDOF = c_loc(idof_t)
      prob%MultiSet = c_loc(ms_t)
      prob%nSPF = c_loc(nspf_t)
      prob%ngp = c_loc(ngp_t)
      prob%ngpmax = ngmax_t
      prob%grid = c_loc(grid_t)
      prob%d1 = c_loc(d1_t)
      prob%d2 = c_loc(d2_t)
      prob%xd1 = c_loc(xd1_t)
      prob%nham = nHam
      prob%ae = c_loc(ae_t)
      prob%lhs_el = c_loc(lhs_t)
      prob%rhs_el = c_loc(rhs_t)
      prob%iOpDof = c_loc(iop_t)
      prob%eps_int = eps_int
      prob%eps_psi = eps_psi
      prob%eps_rho = eps_rho
      prob%writetime = writetime
      prob%min_stepsize = min_stepsize
      prob%int_type_psi = int_type_psi
      prob%eps_C = eps_C
      prob%eps_ovl = eps_ovl
      prob%dof_gwp = c_null_ptr
      prob%gwp_a = c_null_ptr
      prob%nopparam = 0
      prob%gridoper = c_null_ptr
      if (nOpParam > 0) then
         prob%nopparam = nOpParam
         prob%gridoper = c_loc(gop_t)
      end if
      if (nmodGWP > 0) then
         allocate(gwp_t(ndof), gwpa_t(size(iDOF, 1), nSPFmax, nmodes))
         gwp_t = merge(1_c_int64_t, 0_c_int64_t, dvrType == 'gwp')
         gwpa_t = 0.0_c_double
         n2 = min(size(GWPa, 2), nSPFmax)
         do i = 1, nmodGWP
            im = imodGWP(i)
            nd = nmodeDOF(im)
            gwpa_t(1:nd, 1:n2, im) = GWPa(1:nd, 1:n2, i)
         end do
         prob%dof_gwp = c_loc(gwp_t)
         prob%gwp_a = c_loc(gwpa_t)
      end if
      precision = 0
      call get_environment_variable('MACAW_GPU_PRECISION', env, env_len, env_stat)
      if (env_stat == 0 .and. env_len > 0) then
         if (env(1:1) == '1') precision = 1
      end if
      call mg_check(ctx, mg_create(prob, precision, ctx), 'SetupCMF2')
      call mg_check(ctx, mg_set_start(ctx, AVector_Start, psi_Start), 'SetupCMF2')
   end subroutine SetupCMF2

   ! TARGET copies of Macaw's module arrays in the C types of the descriptor.
   subroutine copy_tables()
      integer :: k, n
      if (allocated(nmd_t)) deallocate(nmd_t, idof_t, ms_t, nspf_t, ngp_t, lhs_t, rhs_t, iop_t, &
                                       grid_t, d1_t, d2_t, xd1_t, ae_t)
      if (allocated(gwp_t)) deallocate(gwp_t)
      if (allocated(gwpa_t)) deallocate(gwpa_t)
      if (allocated(gop_t)) deallocate(gop_t)
      allocate(nmd_t(size(nmodeDOF)), idof_t(size(iDOF, 1), size(iDOF, 2)), ms_t(size(MultiSet)), &
               nspf_t(size(nSPF, 1), size(nSPF, 2)), ngp_t(size(ngp)))
      nmd_t = nmodeDOF
      idof_t = iDOF
      ms_t = merge(1_c_int64_t, 0_c_int64_t, MultiSet)
      nspf_t = nSPF
      ngp_t = ngp
      allocate(lhs_t(size(lhs_el)), rhs_t(size(rhs_el)), iop_t(size(iOpDof, 1), size(iOpDof, 2)), ae_t(size(ae)))
      lhs_t = lhs_el
      rhs_t = rhs_el
      iop_t = iOpDof
      ae_t = ae
      ! Macaw's grid arrays are indexed by DVR-dof (indDVR, psidef.f90:33),
      ! the descriptor's by DOF; Gaussian DOFs have no grid (zeros)
      ngmax_t = max(ngpMax, maxval(ngp))
      allocate(grid_t(ngmax_t, ndof), d1_t(ngmax_t, ngmax_t, ndof), d2_t(ngmax_t, ngmax_t, ndof), &
               xd1_t(ngmax_t, ngmax_t, ndof))
      grid_t = 0.0_c_double
      d1_t = 0.0_c_double
      d2_t = 0.0_c_double
      xd1_t = 0.0_c_double
      do k = 1, ndof
         if (dvrType(k) == 'gwp') cycle
         n = ngp(k)
         grid_t(1:n, k) = grid(1:n, indDVR(k))
         d1_t(1:n, 1:n, k) = d1DVR(1:n, 1:n, indDVR(k))
         d2_t(1:n, 1:n, k) = d2DVR(1:n, 1:n, indDVR(k))
         xd1_t(1:n, 1:n, k) = xd1DVR(1:n, 1:n, indDVR(k))
      end do
      ! parametrised operators' grid values, DVR-dof indexed in Macaw
      ! (hamilton.f90:212), DOF-indexed in the descriptor
      if (nOpParam > 0) then
         allocate(gop_t(ngmax_t, nOpParam, ndof))
         gop_t = 0.0_c_double
         do k = 1, ndof
            if (dvrType(k) == 'gwp') cycle
            n = ngp(k)
            gop_t(1:n, 1:nOpParam, k) = gridOper(1:n, 1:nOpParam, indDVR(k))
         end do
      end if
   end subroutine copy_tables

   ! Same signature as cmf2_base.f90:47.  The state is uploaded, stepped on
   ! the GPU and downloaded, so WriteOutput and propagate_wf's own
   ! renormalisation (propagator.f90:288-305) see the updated AVector / psi.
   subroutine cmf2(A0, psi0, DD0, tt, htry, hdid, hnext)
      double complex, dimension(:) :: A0(nAconftot), psi0(dimpsi)
      double complex, dimension(:,:,:,:) :: DD0
      double precision, intent(in) :: tt, htry
      double precision, intent(out) :: hdid, hnext
      call mg_check(ctx, mg_set_state(ctx, A0, psi0), 'cmf2')
      call mg_check(ctx, mg_cmf2_step(ctx, tt, htry, hdid, hnext), 'cmf2')
      call mg_check(ctx, mg_get_state(ctx, A0, psi0), 'cmf2')
   end subroutine cmf2

end module cmf2_base
