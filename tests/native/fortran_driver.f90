! Fortran driver of the HIP time step through the ISO_C_BINDING module sem2d_hip
! (sem2dpack_amd/fortran/sem2d_hip.f90): the calls a SEM2DPACK host program makes in
! place of its time loop.  It reads a problem from case.bin (written by
! tests/test_gpu_fortran.py from the builder's arrays), creates the solver, registers the
! absorbing boundaries, the point forces with their amplitude tables and the receivers,
! steps, fetches the fields and the seismograms, and writes them to out.bin.
!
! case.bin (stream, native byte order):
!   15 int32  ndof nelem npoin nelast coef_mode nx nz ezflt nsteps nabso nsrc nrec isamp
!             field nsamp
!   3 real64  dt beta gamma
!   ibool(nelem*25) int32, hprime(25), rmass(ndof*npoin), coefs(ncoef), elem_of_cell(nx*nz)
!   per absorbing boundary: n, node(n), C(ndof*n)
!   per point force: iglob, dir(2), nt, ampl(nt)
!   receivers: iglob(nrec)
! out.bin: d, v, a (ndof*npoin each), seismograms (nsamp*ndof*nrec)
program fortran_driver
  use iso_c_binding
  use sem2d_hip
  implicit none

  integer(c_int) :: hdr(15), ndof, nelem, npoin, nelast, coef_mode, nx, nz, ezflt, nsteps, nabso, nsrc
  integer(c_int) :: nrec, isamp, field, nsamp, n, iglob, nt, isrc, taken, rc, k, u
  real(c_double) :: tparm(3), dir(2)
  integer(c_int), allocatable, target :: ibool(:), cells(:), node(:), rnode(:)
  real(c_double), allocatable, target :: hprime(:), rmass(:), coefs(:), C(:), ampl(:), d(:), v(:), a(:), seis(:)
  type(sem2d_desc), target :: desc
  type(c_ptr) :: h

  open(newunit=u, file='case.bin', access='stream', form='unformatted', status='old')
  read(u) hdr
  ndof = hdr(1); nelem = hdr(2); npoin = hdr(3); nelast = hdr(4); coef_mode = hdr(5)
  nx = hdr(6); nz = hdr(7); ezflt = hdr(8); nsteps = hdr(9); nabso = hdr(10); nsrc = hdr(11)
  nrec = hdr(12); isamp = hdr(13); field = hdr(14); nsamp = hdr(15)
  read(u) tparm
  allocate(ibool(nelem*25), hprime(25), rmass(ndof*npoin))
  read(u) ibool, hprime, rmass
  if (coef_mode == 1) then
    allocate(coefs(nelast*25))
  else
    allocate(coefs(nelem*nelast*25))
  endif
  read(u) coefs
  allocate(cells(max(nx*nz, 1)))
  if (nx > 0) read(u) cells

  desc%ndof = ndof; desc%ngll = 5; desc%nelem = nelem; desc%npoin = npoin
  desc%nelast = nelast; desc%coef_mode = coef_mode
  desc%ibool = c_loc(ibool); desc%coefs = c_loc(coefs); desc%hprime = c_loc(hprime)
  desc%rmass = c_loc(rmass)
  desc%dt = tparm(1); desc%beta = tparm(2); desc%gamma = tparm(3)
  desc%nx = nx; desc%nz = nz; desc%ezflt = ezflt
  if (nx > 0) desc%elem_of_cell = c_loc(cells)
  call check(sem2d_create(desc, h), 'sem2d_create')

  do k = 1, nabso
    read(u) n
    allocate(node(n), C(ndof*n))
    read(u) node, C
    call check(sem2d_set_abso(h, n, node, C), 'sem2d_set_abso')
    deallocate(node, C)
  enddo
  do k = 1, nsrc
    read(u) iglob, dir, nt
    allocate(ampl(nt))
    read(u) ampl
    call check(sem2d_set_force(h, iglob, dir, isrc), 'sem2d_set_force')
    call check(sem2d_set_stf(h, isrc, nt, ampl), 'sem2d_set_stf')
    deallocate(ampl)
  enddo
  allocate(rnode(max(nrec, 1)))
  if (nrec > 0) then
    read(u) rnode
    call check(sem2d_set_receivers(h, nrec, rnode, isamp, field, nsamp), 'sem2d_set_receivers')
  endif
  close(u)

  call check(sem2d_step(h, nsteps), 'sem2d_step')
  allocate(d(ndof*npoin), v(ndof*npoin), a(ndof*npoin), seis(max(nsamp*ndof*nrec, 1)))
  call check(sem2d_fetch_fields(h, d, v, a), 'sem2d_fetch_fields')
  taken = 0
  if (nrec > 0) call check(sem2d_fetch_seismo(h, seis, taken), 'sem2d_fetch_seismo')
  open(newunit=u, file='out.bin', access='stream', form='unformatted', status='replace')
  write(u) d, v, a
  if (nrec > 0) write(u) seis(1:taken*ndof*nrec)
  close(u)
  call sem2d_destroy(h)
  write(*, '(a,i0,a,i0)') 'fortran_driver: done, steps ', nsteps, ', samples ', taken

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(*), intent(in) :: what
    if (code /= 0) then
      write(*, '(a,a,i0,a,a)') what, ' failed (', code, '): ', trim(sem2d_error_message(h))
      stop 1
    endif
  end subroutine check

end program fortran_driver
