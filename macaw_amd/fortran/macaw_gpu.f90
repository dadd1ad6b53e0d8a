! macaw_gpu.f90 — ISO_C_BINDING interface to the MI355X engine's C-ABI
! (include/macaw_gpu.h).  Integer kinds are int64 because Macaw is compiled
! with -fdefault-integer-8 (Makefile:18); complex arrays are passed as the
! Fortran `double complex` storage they already are.
module macaw_gpu
   use iso_c_binding
   implicit none
   private
   public :: mg_problem, mg_output, mg_attempt
   public :: mg_create, mg_destroy, mg_last_error, mg_nconf, mg_dimpsi, mg_dims
   public :: mg_set_state, mg_get_state, mg_set_start, mg_calc_h1mode, mg_calc_rho
   public :: mg_calc_meanfield, mg_apply_h, mg_der_spf, mg_cmf2_step, mg_observables
   public :: mg_propagate, mg_check

   ! Problem descriptor (psidef.f90:11-55, dvr.f90:14-27, hamilton.f90:24-36)
   type, bind(C) :: mg_problem
      integer(c_int64_t) :: nstates, nmodes, ndof, maxdof
      type(c_ptr) :: nmodeDOF, iDOF, MultiSet, nSPF, ngp
      integer(c_int64_t) :: ngpmax
      type(c_ptr) :: grid, d1, d2, xd1
      integer(c_int64_t) :: nham
      type(c_ptr) :: ae, lhs_el, rhs_el, iOpDof
      real(c_double) :: eps_int, eps_psi, eps_rho, writetime, min_stepsize
      integer(c_int64_t) :: int_type_psi
      type(c_ptr) :: dof_gwp = c_null_ptr, gwp_a = c_null_ptr
      real(c_double) :: eps_C = 1.d-8, eps_ovl = 1.d-8
   end type mg_problem

   type, bind(C) :: mg_output
      real(c_double) :: time, auto_re, auto_im, norm, energy
   end type mg_output

   type, bind(C) :: mg_attempt
      real(c_double) :: time, h, err_A, err_psi
      integer(c_int64_t) :: accepted, nkrylov1, nkrylov2
   end type mg_attempt

   interface
      integer(c_int) function mg_create(prob, precision, ctx) bind(C, name='mg_create')
         import :: c_int, c_int64_t, c_ptr, mg_problem
         type(mg_problem), intent(in) :: prob
         integer(c_int64_t), value :: precision
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function mg_destroy(ctx) bind(C, name='mg_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function mg_last_error(ctx) bind(C, name='mg_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int64_t) function mg_nconf(ctx) bind(C, name='mg_nconf')
         import :: c_int64_t, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int64_t) function mg_dimpsi(ctx) bind(C, name='mg_dimpsi')
         import :: c_int64_t, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mg_dims(ctx, nspfmax, nopmax) bind(C, name='mg_dims')
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: nspfmax, nopmax
      end function
      integer(c_int) function mg_set_state(ctx, A, psi) bind(C, name='mg_set_state')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: A(*), psi(*)
      end function
      integer(c_int) function mg_get_state(ctx, A, psi) bind(C, name='mg_get_state')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(out) :: A(*), psi(*)
      end function
      integer(c_int) function mg_set_start(ctx, A, psi) bind(C, name='mg_set_start')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: A(*), psi(*)
      end function
      integer(c_int) function mg_calc_h1mode(ctx, h1mode) bind(C, name='mg_calc_h1mode')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(out) :: h1mode(*)
      end function
      integer(c_int) function mg_calc_rho(ctx, rho) bind(C, name='mg_calc_rho')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(out) :: rho(*)
      end function
      integer(c_int) function mg_calc_meanfield(ctx, mf, rhom1mf) bind(C, name='mg_calc_meanfield')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(out) :: mf(*), rhom1mf(*)
      end function
      integer(c_int) function mg_apply_h(ctx, B, HB) bind(C, name='mg_apply_h')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: B(*)
         complex(c_double_complex), intent(out) :: HB(*)
      end function
      integer(c_int) function mg_der_spf(ctx, im, psi, dpsi) bind(C, name='mg_der_spf')
         import :: c_int, c_int64_t, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: im
         complex(c_double_complex), intent(in) :: psi(*)
         complex(c_double_complex), intent(out) :: dpsi(*)
      end function
      integer(c_int) function mg_cmf2_step(ctx, t, htry, hdid, hnext) bind(C, name='mg_cmf2_step')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: t, htry
         real(c_double), intent(out) :: hdid, hnext
      end function
      integer(c_int) function mg_observables(ctx, t, out, pops) bind(C, name='mg_observables')
         import :: c_int, c_ptr, c_double, mg_output
         type(c_ptr), value :: ctx
         real(c_double), value :: t
         type(mg_output), intent(out) :: out
         real(c_double), intent(out) :: pops(*)
      end function
      integer(c_int) function mg_propagate(ctx, t0, tfinal, outs, pops, max_out, n_out, &
                                           atts, max_att, n_att) bind(C, name='mg_propagate')
         import :: c_int, c_int64_t, c_ptr, c_double, mg_output, mg_attempt
         type(c_ptr), value :: ctx
         real(c_double), value :: t0, tfinal
         type(mg_output), intent(out) :: outs(*)
         real(c_double), intent(out) :: pops(*)
         integer(c_int64_t), value :: max_out
         integer(c_int64_t), intent(out) :: n_out
         type(mg_attempt), intent(out) :: atts(*)
         integer(c_int64_t), value :: max_att
         integer(c_int64_t), intent(out) :: n_att
      end function
      ! Learned CMF2 step proposal (no counterpart in the reference): params =
      ! the network's MG_MLP_NPARAM reals; pass c_null_ptr for the classical rule.
      integer(c_int) function mg_set_step_model(ctx, params) bind(C, name='mg_set_step_model')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, params
      end function
      ! HIP-graph replay of the launch plans on (1) or off (0); on by default.
      integer(c_int) function mg_set_graphs(ctx, enable) bind(C, name='mg_set_graphs')
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: enable
      end function
   end interface

contains

   ! Turn a non-zero status into `error stop` with the library's message,
   ! keeping the reference's abort-on-error behaviour at the Fortran side.
   subroutine mg_check(ctx, status, where)
      type(c_ptr), intent(in) :: ctx
      integer(c_int), intent(in) :: status
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      character(len=256) :: text
      integer :: i
      if (status == 0) return
      text = ''
      call c_f_pointer(mg_last_error(ctx), msg, [256])
      do i = 1, 256
         if (msg(i) == c_null_char) exit
         text(i:i) = msg(i)
      end do
      write(0, '(3a)') where, ': ', trim(text)
      error stop 'macaw_gpu'
   end subroutine mg_check

end module macaw_gpu
