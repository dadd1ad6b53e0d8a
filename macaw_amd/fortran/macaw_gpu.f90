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
   public :: mg_calc_meanfield, mg_apply_h, mg_apply_h_device, mg_der_spf, mg_cmf2_step, mg_observables
   public :: mg_propagate, mg_set_normalize, mg_stepsize, mg_set_stepsize, mg_set_replay, mg_autotune
   public :: mg_calc_derivs, mg_propagate_vmf
   public :: mg_prepare_gwp, mg_gwp_overlaps, mg_der_spf_gwp, mg_a_to_b
   public :: mg_apply_h_stats, mg_apply_h_profile, mg_profile_copy
   public :: mg_set_step_model, mg_set_graphs, mg_log_step_features, mg_step_features, mg_controller_stats
   public :: mg_mlp_train, mg_mlp_forward, mg_mlp_forward_noalloc, mg_mlp_grad, mg_ensemble_propagate
   public :: mg_check

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
      ! parametrised operators: gridoper(ngpmax, nopparam, ndof), DOF-indexed
      integer(c_int64_t) :: nopparam = 0
      type(c_ptr) :: gridoper = c_null_ptr
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
      ! inorm_psi of propagate_wf (globals.f90:23-26): 0 never, 1 first
      ! interval, 2 every output interval, 3 every step.
      integer(c_int) function mg_set_normalize(ctx, mode) bind(C, name='mg_set_normalize')
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: mode
      end function
      integer(c_int) function mg_stepsize(ctx, h) bind(C, name='mg_stepsize')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: h
      end function
      integer(c_int) function mg_set_stepsize(ctx, h) bind(C, name='mg_set_stepsize')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: h
      end function
      integer(c_int) function mg_set_replay(ctx, log, n) bind(C, name='mg_set_replay')
         import :: c_int, c_int64_t, c_ptr, mg_attempt
         type(c_ptr), value :: ctx
         type(mg_attempt), intent(in) :: log(*)
         integer(c_int64_t), value :: n
      end function
      integer(c_int) function mg_autotune(ctx, reps, best_brick, best_score_ms) bind(C, name='mg_autotune')
         import :: c_int, c_int64_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: reps
         integer(c_int64_t), intent(out) :: best_brick
         real(c_double), intent(out) :: best_score_ms
      end function
      integer(c_int) function mg_apply_h_device(ctx, dB, dHB) bind(C, name='mg_apply_h_device')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, dB, dHB
      end function
      integer(c_int) function mg_apply_h_stats(ctx, bytes_alg, flops_alg) bind(C, name='mg_apply_h_stats')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: bytes_alg, flops_alg
      end function
      integer(c_int) function mg_apply_h_profile(ctx, dB, dHB, kernel_ms, nlaunch) bind(C, name='mg_apply_h_profile')
         import :: c_int, c_int64_t, c_ptr, c_double
         type(c_ptr), value :: ctx, dB, dHB
         real(c_double), intent(out) :: kernel_ms
         integer(c_int64_t), intent(out) :: nlaunch
      end function
      integer(c_int) function mg_profile_copy(ctx, n, kernel_ms) bind(C, name='mg_profile_copy')
         import :: c_int, c_int64_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         real(c_double), intent(out) :: kernel_ms
      end function
      ! Calc_Derivs (derivatives.f90:39-243) and propagate_wf's Runge-Kutta
      ! branch (propagator.f90:108-140); method 1 Cash-Karp, 2 Dormand-Prince.
      integer(c_int) function mg_calc_derivs(ctx, dA, dpsi) bind(C, name='mg_calc_derivs')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(out) :: dA(*), dpsi(*)
      end function
      integer(c_int) function mg_propagate_vmf(ctx, method, t0, tfinal, outs, pops, max_out, n_out, &
                                               atts, max_att, n_att) bind(C, name='mg_propagate_vmf')
         import :: c_int, c_int64_t, c_ptr, c_double, mg_output, mg_attempt
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: method
         real(c_double), value :: t0, tfinal
         type(mg_output), intent(out) :: outs(*)
         real(c_double), intent(out) :: pops(*)
         integer(c_int64_t), value :: max_out
         integer(c_int64_t), intent(out) :: n_out
         type(mg_attempt), intent(out) :: atts(*)
         integer(c_int64_t), value :: max_att
         integer(c_int64_t), intent(out) :: n_att
      end function
      ! Gaussian (G-MCTDH) modes: Calc_S00 / InverseSquareRootU / Calc_DHD
      ! (mg_prepare_gwp), the overlaps and D0, Calc_Der_SPF of a Gaussian
      ! mode, A_to_B (cmf2_base.f90:264-305).
      integer(c_int) function mg_prepare_gwp(ctx) bind(C, name='mg_prepare_gwp')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mg_gwp_overlaps(ctx, s00, dd) bind(C, name='mg_gwp_overlaps')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(out) :: s00(*), dd(*)
      end function
      integer(c_int) function mg_der_spf_gwp(ctx, im, psi, dd, dpsi, ddd) bind(C, name='mg_der_spf_gwp')
         import :: c_int, c_int64_t, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: im
         complex(c_double_complex), intent(in) :: psi(*), dd(*)
         complex(c_double_complex), intent(out) :: dpsi(*), ddd(*)
      end function
      integer(c_int) function mg_a_to_b(ctx, A, dd, B) bind(C, name='mg_a_to_b')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: A(*), dd(*)
         complex(c_double_complex), intent(out) :: B(*)
      end function
      ! Learned step controller: training, evaluation and the step logs.
      integer(c_int) function mg_log_step_features(ctx, enable) bind(C, name='mg_log_step_features')
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: enable
      end function
      integer(c_int) function mg_step_features(ctx, features, h, max, n) bind(C, name='mg_step_features')
         import :: c_int, c_int64_t, c_ptr, c_float, c_double
         type(c_ptr), value :: ctx
         real(c_float), intent(out) :: features(*)
         real(c_double), intent(out) :: h(*)
         integer(c_int64_t), value :: max
         integer(c_int64_t), intent(out) :: n
      end function
      integer(c_int) function mg_controller_stats(ctx, proposals, fallbacks) bind(C, name='mg_controller_stats')
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: proposals, fallbacks
      end function
      integer(c_int) function mg_mlp_train(ctx, X, t, n, epochs, batch, lr, seed, params, loss) &
            bind(C, name='mg_mlp_train')
         import :: c_int, c_int64_t, c_ptr, c_float, c_double
         type(c_ptr), value :: ctx
         real(c_float), intent(in) :: X(*), t(*)
         integer(c_int64_t), value :: n, epochs, batch
         real(c_double), value :: lr
         integer(c_int64_t), value :: seed
         real(c_float), intent(inout) :: params(*)
         real(c_double), intent(out) :: loss
      end function
      integer(c_int) function mg_mlp_forward(ctx, params, X, n, y) bind(C, name='mg_mlp_forward')
         import :: c_int, c_int64_t, c_ptr, c_float
         type(c_ptr), value :: ctx
         real(c_float), intent(in) :: params(*), X(*)
         integer(c_int64_t), value :: n
         real(c_float), intent(out) :: y(*)
      end function
      ! mg_mlp_forward without device allocation, n <= 4096 rows (for callers
      ! that cannot allocate device memory; calls overlapping on one GPU race)
      integer(c_int) function mg_mlp_forward_noalloc(ctx, params, X, n, y) bind(C, name='mg_mlp_forward_noalloc')
         import :: c_int, c_int64_t, c_ptr, c_float
         type(c_ptr), value :: ctx
         real(c_float), intent(in) :: params(*), X(*)
         integer(c_int64_t), value :: n
         real(c_float), intent(out) :: y(*)
      end function
      integer(c_int) function mg_mlp_grad(ctx, params, X, t, n, grad, loss) bind(C, name='mg_mlp_grad')
         import :: c_int, c_int64_t, c_ptr, c_float, c_double
         type(c_ptr), value :: ctx
         real(c_float), intent(in) :: params(*), X(*), t(*)
         integer(c_int64_t), value :: n
         real(c_float), intent(out) :: grad(*)
         real(c_double), intent(out) :: loss
      end function
      ! Ensemble of independent wavepackets: probs, A0, psi0 are arrays of
      ! n C pointers (c_loc of each member's descriptor and state).
      integer(c_int) function mg_ensemble_propagate(probs, A0, psi0, n, precision, nthreads, mem_fraction, &
            flags, t0, tfinal, outs, pops, max_out, n_out, status, member_wall, errbuf, errlen) &
            bind(C, name='mg_ensemble_propagate')
         import :: c_int, c_int64_t, c_ptr, c_double, c_char, mg_output
         type(c_ptr), intent(in) :: probs(*), A0(*), psi0(*)
         integer(c_int64_t), value :: n, precision, nthreads
         real(c_double), value :: mem_fraction
         integer(c_int64_t), value :: flags
         real(c_double), value :: t0, tfinal
         type(mg_output), intent(out) :: outs(*)
         real(c_double), intent(out) :: pops(*)
         integer(c_int64_t), value :: max_out
         integer(c_int64_t), intent(out) :: n_out(*), status(*)
         real(c_double), intent(out) :: member_wall(*)
         character(kind=c_char), intent(out) :: errbuf(*)
         integer(c_int64_t), value :: errlen
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
