! ref_harness -- FP64 dump driver linked against the reference modules.
!
! Test infrastructure only (the oracle): it is compiled by oracle/build_ref.py
! together with the reference sources and never shipped.  It follows the
! reference program SRC/main.f90:25-127 (read_main, init_main, the time loop
! calling solve, REC_store and BC_write, then REC_write) and additionally
! writes double precision binary dumps that the reference itself only writes
! in single precision:
!
!   h_grid.bin   numbering, coordinates, inverse mass, GLL data, boundaries
!   h_fint.bin   a deterministic displacement d and f = -K d assembled as in
!                compute_Fint (SRC/solver.f90:273-308), element by element
!                in the reference element order
!   h_nodes.bin  d, v, a at the recorded nodes (receiver nodes and the nodes of
!                boundaries tagged 5 and 6, i.e. split-fault sides) after every
!                step, step 0 included
!   h_final.bin  full d, v, a after the last step
!
! Usage (in a directory holding Par.inp):  ref_harness [nsteps]
!   nsteps overrides the number of steps of the time loop (default: TIME nt).
program ref_harness

  use problem_class
  use input, only : read_main
  use init, only : init_main
  use solver, only : solve
  use receivers, only : REC_store, REC_write, REC_inquire
  use bc_gen, only : BC_write
  use fields_class, only : FIELD_get_elem_sub, FIELD_add_elem
  use mat_gen, only : MAT_Fint
  use echo, only : iout

  implicit none

  type(problem_type) :: pb
  integer :: iexec, it, nt, ngll, nelem, npoin, ndof, e, k, c, i, n, nrec, nb, u
  double precision, allocatable :: dd(:,:), vv(:,:), ff(:,:), dloc(:,:,:), vloc(:,:,:), floc(:,:,:)
  double precision :: E_ep, E_el, sg(3), sgp(3)
  double precision, pointer :: rcoord(:,:)
  integer, allocatable :: recnodes(:)
  character(len=32) :: arg

  call read_main(pb, iexec, input_file='Par.inp')
  call init_main(pb)
  pb%time%time = 0d0

  ngll  = pb%grid%ngll
  nelem = pb%grid%nelem
  npoin = pb%grid%npoin
  ndof  = pb%fields%ndof
  nt    = pb%time%nt
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    read(arg, *) nt
  endif

  ! ---- grid dump ------------------------------------------------------
  open(newunit=u, file='h_grid.bin', access='stream', form='unformatted', status='replace')
  write(u) ngll, nelem, npoin, ndof, nt
  write(u) pb%time%dt
  write(u) pb%grid%ibool
  write(u) pb%grid%coord
  write(u) pb%rmass
  write(u) pb%grid%hprime
  write(u) pb%grid%xgll
  write(u) pb%grid%wgll
  nb = 0
  if (associated(pb%grid%bounds)) nb = size(pb%grid%bounds)
  write(u) nb
  do i = 1, nb
    write(u) pb%grid%bounds(i)%tag, pb%grid%bounds(i)%npoin, pb%grid%bounds(i)%nelem
    write(u) pb%grid%bounds(i)%node
    write(u) pb%grid%bounds(i)%elem
    write(u) pb%grid%bounds(i)%edge
  enddo
  close(u)

  ! ---- f = -K d for a deterministic d ---------------------------------
  allocate(dd(npoin,ndof), vv(npoin,ndof), ff(npoin,ndof))
  allocate(dloc(ngll,ngll,ndof), vloc(ngll,ngll,ndof), floc(ngll,ngll,ndof))
  do c = 1, ndof
    do k = 1, npoin
      dd(k,c) = 1d-3*sin(1d-3*pb%grid%coord(1,k)*dble(c) + 7d-4*pb%grid%coord(2,k)) &
              + 1d-4*cos(0.37d0*dble(k) + 1.1d0*dble(c))
    enddo
  enddo
  vv = 0d0
  ff = 0d0
  do e = 1, nelem
    call FIELD_get_elem_sub(dd, pb%grid%ibool(:,:,e), dloc)
    call FIELD_get_elem_sub(vv, pb%grid%ibool(:,:,e), vloc)
    call MAT_Fint(floc, dloc, vloc, pb%matpro(e), pb%matwrk(e), ngll, ndof, &
                  pb%time%dt, pb%grid, E_ep, E_el, sg, sgp)
    call FIELD_add_elem(floc, ff, pb%grid%ibool(:,:,e))
  enddo
  open(newunit=u, file='h_fint.bin', access='stream', form='unformatted', status='replace')
  write(u) dd
  write(u) ff
  close(u)
  deallocate(dd, vv, ff, dloc, vloc, floc)

  ! ---- recorded nodes: receivers, then fault sides (tags 5, 6) ---------
  nrec = 0
  if (associated(pb%rec)) then
    call REC_inquire(pb%rec, coord=rcoord)
    nrec = size(rcoord, 2)
  endif
  n = nrec
  do i = 1, nb
    if (pb%grid%bounds(i)%tag == 5 .or. pb%grid%bounds(i)%tag == 6) n = n + pb%grid%bounds(i)%npoin
  enddo
  allocate(recnodes(n))
  do k = 1, nrec
    ! receivers inside elements (AtNode=F) sit on no node: dump node 1 for them (their
    ! series are taken from the reference's own seismogram files instead).  Of the nodes
    ! at the receiver's coordinates (the two sides of a split fault) the last one, the
    ! node SE_find_nearest_node keeps (d2 <= d2min)
    recnodes(k) = 1
    do i = 1, npoin
      if (pb%grid%coord(1,i) == rcoord(1,k) .and. pb%grid%coord(2,i) == rcoord(2,k)) recnodes(k) = i
    enddo
  enddo
  n = nrec
  do i = 1, nb
    if (pb%grid%bounds(i)%tag == 5 .or. pb%grid%bounds(i)%tag == 6) then
      recnodes(n+1:n+pb%grid%bounds(i)%npoin) = pb%grid%bounds(i)%node
      n = n + pb%grid%bounds(i)%npoin
    endif
  enddo

  open(newunit=u, file='h_nodes.bin', access='stream', form='unformatted', status='replace')
  write(u) nrec, n
  write(u) recnodes
  call dump_nodes()

  ! ---- time loop, as SRC/main.f90:51-99 -------------------------------
  if (associated(pb%rec)) call REC_store(pb%rec, 0, pb%grid)
  do it = 1, nt
    pb%time%time = it*pb%time%dt
    call solve(pb)
    if (associated(pb%rec)) call REC_store(pb%rec, it, pb%grid)
    call BC_write(pb%bc, it, pb%fields%displ, pb%fields%veloc)
    call dump_nodes()
  enddo
  close(u)

  open(newunit=u, file='h_final.bin', access='stream', form='unformatted', status='replace')
  write(u) pb%fields%displ
  write(u) pb%fields%veloc
  write(u) pb%fields%accel
  close(u)

  if (associated(pb%rec) .and. nt == pb%time%nt) call REC_write(pb%rec, iout)
  write(*,'(a)') 'ref_harness: done'

contains

  subroutine dump_nodes()
    if (n > 0) then
      write(u) pb%fields%displ(recnodes,:)
      write(u) pb%fields%veloc(recnodes,:)
      write(u) pb%fields%accel(recnodes,:)
    endif
  end subroutine dump_nodes

end program ref_harness
