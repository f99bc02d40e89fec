! sem2d_hip.f90 -- ISO_C_BINDING interface to libsem2d_hip.so (include/sem2d.h).
!
! This is the thin Fortran layer through which the SEM2DPACK host code calls the
! MI355X time step.  It only declares interoperable types and interfaces; it
! depends on no other module, so it compiles on its own and can be dropped into the
! reference's SRC/ directory unchanged.  The derived types mirror sem2d_desc and
! sem2d_fault_desc field by field (same order, C-interoperable kinds).
module sem2d_hip

  use iso_c_binding
  implicit none
  private

  type, bind(C), public :: sem2d_desc
    integer(c_int) :: ndof = 0, ngll = 5, nelem = 0, npoin = 0
    integer(c_int) :: nelast = 0, coef_mode = 0
    type(c_ptr) :: ibool = c_null_ptr, coefs = c_null_ptr, lam_mu = c_null_ptr, &
                   geom = c_null_ptr, wgll = c_null_ptr, hprime = c_null_ptr, rmass = c_null_ptr
    real(c_double) :: dt = 0d0, beta = 0d0, gamma = 0.5d0
    integer(c_int) :: bulk_fp32 = 0
    integer(c_int) :: nx = 0, nz = 0, ezflt = 0
    type(c_ptr) :: elem_of_cell = c_null_ptr
  end type sem2d_desc

  type, bind(C), public :: sem2d_fault_desc
    integer(c_int) :: npoin = 0
    type(c_ptr) :: node1 = c_null_ptr, node2 = c_null_ptr, n1 = c_null_ptr
    type(c_ptr) :: B = c_null_ptr, invM1 = c_null_ptr, invM2 = c_null_ptr, Z = c_null_ptr
    type(c_ptr) :: T0 = c_null_ptr, cohesion = c_null_ptr
    real(c_double) :: coefA2V = 0d0, coefA2D = 0d0
    integer(c_int) :: allow_opening = 1, law = 1, kind = 1, healing = 0
    type(c_ptr) :: p(6) = c_null_ptr
    type(c_ptr) :: theta0 = c_null_ptr, V0 = c_null_ptr
    integer(c_int) :: oit = 0, oitd = 1, oix1 = 1, oixn = 1, oixd = 1
  end type sem2d_fault_desc

  public :: sem2d_create, sem2d_set_kelvin_voigt, sem2d_set_abso, sem2d_set_abso_stacey, &
            sem2d_set_abso_normals, &
            sem2d_set_force, sem2d_set_moment, sem2d_set_stf, &
            sem2d_set_receivers, sem2d_set_receivers_interp, sem2d_set_fault, sem2d_set_dirichlet, &
            sem2d_set_25d, sem2d_set_time_scheme, sem2d_set_fault_sides, sem2d_fetch_fault_sides, &
            sem2d_set_fault_twf, sem2d_set_fault_normal, sem2d_set_periodic, sem2d_set_hht, &
            sem2d_set_abso_wave, sem2d_set_lsf, sem2d_set_abso_periodic, &
            sem2d_step, &
            sem2d_fetch_fields, &
            sem2d_set_fields, sem2d_fetch_seismo, sem2d_fetch_fault, sem2d_fetch_fault_state, &
            sem2d_fetch_potency, &
            sem2d_compute_fint, sem2d_compute_fint_noalloc, sem2d_set_itinfo, sem2d_fetch_info, &
            sem2d_field_info, sem2d_set_tile_config, sem2d_get_step, sem2d_last_step_ms, &
            sem2d_last_error, sem2d_destroy, sem2d_error_message, &
            sem2d_adjoint_init, sem2d_adjoint_forward, sem2d_adjoint_backward, sem2d_adjoint_fetch, &
            sem2d_adjoint_timing, sem2d_adjoint_plan, sem2d_adjoint_release, &
            sem2d_batch_create, sem2d_batch_step, sem2d_batch_tune, sem2d_batch_set_config, &
            sem2d_batch_config, sem2d_batch_fetch_fields, sem2d_batch_fetch_seismo, &
            sem2d_batch_fetch_fault, sem2d_batch_fetch_fault_state, sem2d_batch_fetch_potency, &
            sem2d_batch_get_step, &
            sem2d_batch_last_step_ms, sem2d_batch_last_error, sem2d_batch_destroy

  interface
    integer(c_int) function sem2d_create(desc, h) bind(C, name='sem2d_create')
      import :: c_int, c_ptr, sem2d_desc
      type(sem2d_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function sem2d_set_kelvin_voigt(h, eta) bind(C, name='sem2d_set_kelvin_voigt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: eta(*)
    end function

    integer(c_int) function sem2d_set_abso(h, n, node, C) bind(C, name='sem2d_set_abso')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      integer(c_int), intent(in) :: node(*)
      real(c_double), intent(in) :: C(*)
    end function
    integer(c_int) function sem2d_set_abso_stacey(h, nbe, enode, K) bind(C, name='sem2d_set_abso_stacey')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nbe
      integer(c_int), intent(in) :: enode(*)
      real(c_double), intent(in) :: K(*)
    end function
    integer(c_int) function sem2d_set_abso_normals(h, n, nrm) bind(C, name='sem2d_set_abso_normals')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_double), intent(in) :: nrm(*)
    end function

    integer(c_int) function sem2d_set_force(h, iglob, dir, isrc) bind(C, name='sem2d_set_force')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: iglob
      real(c_double), intent(in) :: dir(*)
      integer(c_int), intent(out) :: isrc
    end function

    integer(c_int) function sem2d_set_moment(h, nel, iglob_xi, iglob_eta, coef_xi, coef_eta, isrc) &
        bind(C, name='sem2d_set_moment')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nel
      integer(c_int), intent(in) :: iglob_xi(*), iglob_eta(*)
      real(c_double), intent(in) :: coef_xi(*), coef_eta(*)
      integer(c_int), intent(out) :: isrc
    end function

    integer(c_int) function sem2d_set_stf(h, isrc, nt, ampl) bind(C, name='sem2d_set_stf')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: isrc, nt
      real(c_double), intent(in) :: ampl(*)
    end function

    integer(c_int) function sem2d_set_receivers(h, nrec, iglob, isamp, field, nsamp) &
        bind(C, name='sem2d_set_receivers')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nrec, isamp, field, nsamp
      integer(c_int), intent(in) :: iglob(*)
    end function

    integer(c_int) function sem2d_set_receivers_interp(h, nrec, elem, interp, isamp, field, nsamp) &
        bind(C, name='sem2d_set_receivers_interp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nrec, isamp, field, nsamp
      integer(c_int), intent(in) :: elem(*)
      real(c_double), intent(in) :: interp(*)
    end function

    integer(c_int) function sem2d_set_fault(h, f, nsamp) bind(C, name='sem2d_set_fault')
      import :: c_int, c_ptr, sem2d_fault_desc
      type(c_ptr), value :: h
      type(sem2d_fault_desc), intent(in) :: f
      integer(c_int), value :: nsamp
    end function

    integer(c_int) function sem2d_step(h, nsteps) bind(C, name='sem2d_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function sem2d_fetch_fields(h, d, v, a) bind(C, name='sem2d_fetch_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: d(*), v(*), a(*)
    end function

    integer(c_int) function sem2d_set_fields(h, d, v, a) bind(C, name='sem2d_set_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: d(*), v(*), a(*)
    end function

    integer(c_int) function sem2d_fetch_seismo(h, out, nsamp_taken) bind(C, name='sem2d_fetch_seismo')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nsamp_taken
    end function

    integer(c_int) function sem2d_fetch_fault(h, out, nsamp_taken) bind(C, name='sem2d_fetch_fault')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nsamp_taken
    end function

    ! time-weakening friction (BC_DYNFLT_TWF): par = MuS, MuD, Mu0, X, Z, V, L, T, Dc,
    ! coord(npoin, 2) the fault node coordinates; after sem2d_set_fault
    integer(c_int) function sem2d_set_fault_twf(h, kind, par, coord) bind(C, name='sem2d_set_fault_twf')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: kind
      real(c_double), intent(in) :: par(9), coord(*)
    end function

    ! normal stress response (BC_DYNFLT_NOR): kind 0..3, coef = exp(-dt/T) (kind 2) or
    ! dt/L (kind 3), v = V (kind 3); after sem2d_set_fault
    integer(c_int) function sem2d_set_fault_normal(h, kind, coef, v) bind(C, name='sem2d_set_fault_normal')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: kind
      real(c_double), value :: coef, v
    end function

    ! periodic boundaries (BC_DEF kind = 'PERIOD'): n node pairs (1-based), master and
    ! slave; right after sem2d_create
    integer(c_int) function sem2d_set_periodic(h, n, master, slave) bind(C, name='sem2d_set_periodic')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: n
      integer(c_int), intent(in) :: master(*), slave(*)
    end function

    ! incident plane wave (mechanism = 'WAVE') through the absorbing boundary set last:
    ! xz(n, 2), vdir(n, ndof), tdir(n, ndof), B(n), par = s1, s2, x0, z0, tdelay, f0, t0, ampli
    integer(c_int) function sem2d_set_abso_wave(h, n, xz, vdir, tdir, B, par) bind(C, name='sem2d_set_abso_wave')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_double), intent(in) :: xz(*), vdir(*), tdir(*), B(*), par(8)
    end function

    ! the absorbing boundary set last ends on periodic boundaries: Stacey end terms summed
    integer(c_int) function sem2d_set_abso_periodic(h) bind(C, name='sem2d_set_abso_periodic')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    ! linear slip fault (BC_DEF kind = 'LISFLT'): node pairs, normals(n, 2), B, W1, W2 (n, ndof)
    integer(c_int) function sem2d_set_lsf(h, n, node1, node2, nrm, B, W1, W2, kt, kn) bind(C, name='sem2d_set_lsf')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      integer(c_int), intent(in) :: node1(*), node2(*)
      real(c_double), intent(in) :: nrm(*), B(*), W1(*), W2(*)
      real(c_double), value :: kt, kn
    end function

    ! HHT-alpha time scheme (TIME kind = 'HHT-alpha'): right after sem2d_create
    integer(c_int) function sem2d_set_hht(h, alpha) bind(C, name='sem2d_set_hht')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: alpha
    end function

    ! per-side fault records (osides): set after sem2d_set_fault; out(nout, 4*ndof, nsamp)
    integer(c_int) function sem2d_set_fault_sides(h, on) bind(C, name='sem2d_set_fault_sides')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function

    integer(c_int) function sem2d_fetch_fault_sides(h, out, nsamp_taken) bind(C, name='sem2d_fetch_fault_sides')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nsamp_taken
    end function

    integer(c_int) function sem2d_fetch_fault_state(h, D, V, T, MU, Tstick) &
        bind(C, name='sem2d_fetch_fault_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: D(*), V(*), T(*), MU(*), Tstick(*)
    end function

    integer(c_int) function sem2d_fetch_potency(h, out, nrows) bind(C, name='sem2d_fetch_potency')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nrows
    end function

    ! 2.5D crustal-plane coefficients beta(ngll,ngll,nelem) (MAT_ELAST_init_25D)
    integer(c_int) function sem2d_set_25d(h, beta) bind(C, name='sem2d_set_25d')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: beta(*)
    end function

    ! BC_DIRNEU Dirichlet components (comp: 1, 2 or 3), in BC order relative to the fault
    integer(c_int) function sem2d_set_dirichlet(h, n, node, comp) bind(C, name='sem2d_set_dirichlet')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: n, comp
      integer(c_int), intent(in) :: node(*)
    end function

    integer(c_int) function sem2d_compute_fint(h, d, f) bind(C, name='sem2d_compute_fint')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: d(*)
      real(c_double), intent(out) :: f(*)
    end function

    ! as sem2d_compute_fint, with staging buffers in the library's static device scratch
    ! array (no device allocation; ndof*npoin <= 4194304)
    integer(c_int) function sem2d_compute_fint_noalloc(h, d, f) bind(C, name='sem2d_compute_fint_noalloc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: d(*)
      real(c_double), intent(out) :: f(*)
    end function

    ! time scheme: 0 Newmark (default), 1 leapfrog (solve_leapfrog)
    integer(c_int) function sem2d_set_time_scheme(h, kind) bind(C, name='sem2d_set_time_scheme')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kind
    end function

    ! progress rows (it, it*dt, max|v|, max|d|) every itinfo steps and the non-finite
    ! watchdog (sem2d_step returns 21 once a row finds a value that is not finite)
    integer(c_int) function sem2d_set_itinfo(h, itinfo) bind(C, name='sem2d_set_itinfo')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: itinfo
    end function

    integer(c_int) function sem2d_fetch_info(h, rows, nrows) bind(C, name='sem2d_fetch_info')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: rows(*)
      integer(c_int), intent(out) :: nrows
    end function

    integer(c_int) function sem2d_field_info(h, vmax, dmax, nonfinite) bind(C, name='sem2d_field_info')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: vmax, dmax
      integer(c_int), intent(out) :: nonfinite
    end function

    integer(c_int) function sem2d_set_tile_config(h, threads, wg_per_cu) bind(C, name='sem2d_set_tile_config')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: threads, wg_per_cu
    end function

    integer(c_int) function sem2d_get_step(h, it) bind(C, name='sem2d_get_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: it
    end function

    integer(c_int) function sem2d_last_step_ms(h, ms) bind(C, name='sem2d_last_step_ms')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ms
    end function

    type(c_ptr) function sem2d_last_error(h) bind(C, name='sem2d_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function

    subroutine sem2d_destroy(h) bind(C, name='sem2d_destroy')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine

    ! adjoint of the time loop
    integer(c_int) function sem2d_adjoint_init(h, nsteps, seg) bind(C, name='sem2d_adjoint_init')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps, seg
    end function

    integer(c_int) function sem2d_adjoint_forward(h) bind(C, name='sem2d_adjoint_forward')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function sem2d_adjoint_backward(h, gseis) bind(C, name='sem2d_adjoint_backward')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: gseis(*)
    end function

    integer(c_int) function sem2d_adjoint_fetch(h, glam, gmu, gmass, gabso) bind(C, name='sem2d_adjoint_fetch')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: glam(*), gmu(*), gmass(*), gabso(*)
    end function

    integer(c_int) function sem2d_adjoint_timing(h, fwd_ms, bwd_ms) bind(C, name='sem2d_adjoint_timing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: fwd_ms, bwd_ms
    end function
    integer(c_int) function sem2d_adjoint_plan(h, history, seg) bind(C, name='sem2d_adjoint_plan')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: history, seg
    end function
    integer(c_int) function sem2d_adjoint_release() bind(C, name='sem2d_adjoint_release')
      import :: c_int
    end function

    ! batched ensemble
    integer(c_int) function sem2d_batch_create(n, members, b) bind(C, name='sem2d_batch_create')
      import :: c_int, c_ptr
      integer(c_int), value :: n
      type(c_ptr), intent(in) :: members(*)
      type(c_ptr), intent(out) :: b
    end function

    integer(c_int) function sem2d_batch_step(b, nsteps) bind(C, name='sem2d_batch_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function sem2d_batch_tune(b, trial_steps, choice) bind(C, name='sem2d_batch_tune')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: trial_steps
      integer(c_int), intent(out) :: choice
    end function

    integer(c_int) function sem2d_batch_set_config(b, threads, wg_per_cu) bind(C, name='sem2d_batch_set_config')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: threads, wg_per_cu
    end function

    integer(c_int) function sem2d_batch_config(b, threads, workgroups, ntiles) bind(C, name='sem2d_batch_config')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), intent(out) :: threads, workgroups, ntiles
    end function

    integer(c_int) function sem2d_batch_fetch_fields(b, m, d, v, a) bind(C, name='sem2d_batch_fetch_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: m
      real(c_double), intent(out) :: d(*), v(*), a(*)
    end function

    integer(c_int) function sem2d_batch_fetch_seismo(b, m, out, nsamp_taken) bind(C, name='sem2d_batch_fetch_seismo')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: m
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nsamp_taken
    end function

    integer(c_int) function sem2d_batch_fetch_fault(b, m, out, nsamp_taken) bind(C, name='sem2d_batch_fetch_fault')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: m
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nsamp_taken
    end function

    integer(c_int) function sem2d_batch_fetch_fault_state(b, m, D, V, T, MU, Tstick) &
        bind(C, name='sem2d_batch_fetch_fault_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: m
      real(c_double), intent(out) :: D(*), V(*), T(*), MU(*), Tstick(*)
    end function

    integer(c_int) function sem2d_batch_fetch_potency(b, m, out, nrows) bind(C, name='sem2d_batch_fetch_potency')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: m
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: nrows
    end function

    integer(c_int) function sem2d_batch_get_step(b, it) bind(C, name='sem2d_batch_get_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), intent(out) :: it
    end function

    integer(c_int) function sem2d_batch_last_step_ms(b, ms) bind(C, name='sem2d_batch_last_step_ms')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      real(c_double), intent(out) :: ms
    end function

    type(c_ptr) function sem2d_batch_last_error(b) bind(C, name='sem2d_batch_last_error')
      import :: c_ptr
      type(c_ptr), value :: b
    end function

    subroutine sem2d_batch_destroy(b) bind(C, name='sem2d_batch_destroy')
      import :: c_ptr
      type(c_ptr), value :: b
    end subroutine
  end interface

contains

  ! Message of the last failed call, as a Fortran string (for IO_abort).
  function sem2d_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: k
    msg = ''
    p = sem2d_last_error(h)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do k = 1, 512
      if (cs(k) == c_null_char) exit
      msg(k:k) = cs(k)
    enddo
  end function sem2d_error_message

end module sem2d_hip
