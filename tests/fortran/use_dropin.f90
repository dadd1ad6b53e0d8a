! A driver fragment in the shape of propagate_wf's CMF2 branch
! (propagator.f90:264-271): it only has to compile and link against the
! drop-in module; it is not run.
program use_dropin
   use psidef
   use globals
   use cmf2_base
   implicit none
   double precision :: step_done, step_next
   if (command_argument_count() > 99) then
      call SetupCMF2
      call cmf2(AVector, psi, DD, time, stepsize, step_done, step_next)
   end if
end program use_dropin
