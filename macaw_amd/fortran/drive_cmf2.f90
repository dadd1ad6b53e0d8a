! drive_cmf2.f90 — a Fortran driver that runs Macaw's CMF2 propagation loop on
! the MI355X engine through the ISO_C_BINDING layer (macaw_gpu.f90).  It reads
! a problem descriptor written by macaw_amd.fortran_io.write_descriptor and
! writes `auto` and `spop` in Macaw's formats (output.f90:228-229, 243-246).
program drive_cmf2
   use iso_c_binding
   use macaw_gpu
   implicit none
   integer(c_int64_t) :: hdr(9), nstates, nmodes, ndof, maxdof, ngpmax, nham, nconf, dimpsi
   real(c_double) :: par(6), tfinal
   integer(c_int64_t), allocatable, target :: nmodeDOF(:), iDOF(:), MultiSet(:), nSPF(:), ngp(:), &
                                              lhs(:), rhs(:), iOpDof(:)
   real(c_double), allocatable, target :: grid(:), d1(:), d2(:), xd1(:), ae(:)
   complex(c_double_complex), allocatable :: A0(:), psi0(:)
   type(mg_problem) :: prob
   type(c_ptr) :: ctx
   type(mg_output), allocatable :: outs(:)
   type(mg_attempt), allocatable :: atts(:)
   real(c_double), allocatable :: pops(:)
   integer(c_int64_t) :: nout, natt, i, maxout
   integer(c_int) :: rc
   character(len=512) :: path
   real(c_double), parameter :: au2fs = 0.02418884326505d0
   integer :: u

   call get_command_argument(1, path)
   open(newunit=u, file=trim(path), access='stream', form='unformatted', status='old', action='read')
   read(u) hdr
   nstates = hdr(1); nmodes = hdr(2); ndof = hdr(3); maxdof = hdr(4); ngpmax = hdr(5)
   nham = hdr(6); nconf = hdr(7); dimpsi = hdr(8)
   read(u) par
   tfinal = par(6)
   allocate(nmodeDOF(nmodes), iDOF(maxdof * nmodes), MultiSet(nmodes), nSPF(nmodes * nstates), &
            ngp(ndof), lhs(nham), rhs(nham), iOpDof(ndof * nham))
   allocate(grid(ngpmax * ndof), d1(ngpmax * ngpmax * ndof), d2(ngpmax * ngpmax * ndof), &
            xd1(ngpmax * ngpmax * ndof), ae(2 * nham), A0(nconf), psi0(dimpsi))
   read(u) nmodeDOF, iDOF, MultiSet, nSPF, ngp, lhs, rhs, iOpDof
   read(u) grid, d1, d2, xd1, ae, A0, psi0
   close(u)

   prob%nstates = nstates; prob%nmodes = nmodes; prob%ndof = ndof; prob%maxdof = maxdof
   prob%nmodeDOF = c_loc(nmodeDOF); prob%iDOF = c_loc(iDOF); prob%MultiSet = c_loc(MultiSet)
   prob%nSPF = c_loc(nSPF); prob%ngp = c_loc(ngp); prob%ngpmax = ngpmax
   prob%grid = c_loc(grid); prob%d1 = c_loc(d1); prob%d2 = c_loc(d2); prob%xd1 = c_loc(xd1)
   prob%nham = nham; prob%ae = c_loc(ae); prob%lhs_el = c_loc(lhs); prob%rhs_el = c_loc(rhs)
   prob%iOpDof = c_loc(iOpDof)
   prob%eps_int = par(1); prob%eps_psi = par(2); prob%eps_rho = par(3)
   prob%writetime = par(4); prob%min_stepsize = par(5); prob%int_type_psi = hdr(9)

   rc = mg_create(prob, 0_c_int64_t, ctx)
   call mg_check(ctx, rc, 'mg_create')
   call mg_check(ctx, mg_set_state(ctx, A0, psi0), 'mg_set_state')
   call mg_check(ctx, mg_set_start(ctx, A0, psi0), 'mg_set_start')
   maxout = int(tfinal / prob%writetime, c_int64_t) + 4
   allocate(outs(maxout), pops(nstates * maxout), atts(1))
   rc = mg_propagate(ctx, 0.0_c_double, tfinal, outs, pops, maxout, nout, atts, 0_c_int64_t, natt)
   call mg_check(ctx, rc, 'mg_propagate')
   open(newunit=u, file='auto', status='replace', action='write')
   write(u, '(a)') '#    time[fs]         Re(autocorrel)     Im(autocorrel)     Abs(autocorrel)'
   do i = 1, nout
      write(u, '(2x,f10.4,5x,3(1x,f18.14))') outs(i)%time * au2fs, outs(i)%auto_re, outs(i)%auto_im, &
           sqrt(outs(i)%auto_re**2 + outs(i)%auto_im**2)
   end do
   close(u)
   open(newunit=u, file='spop', status='replace', action='write')
   write(u, '(a)') '#  time[fs]      State populations'
   do i = 1, nout
      write(u, '(f10.4,2x,30(2x,f10.7))') outs(i)%time * au2fs, pops((i - 1) * nstates + 1 : i * nstates)
   end do
   close(u)
   rc = mg_destroy(ctx)
end program drive_cmf2
