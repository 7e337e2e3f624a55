! isph_main.f90 -- Fortran driver of the device ISPH pipeline.
!
!   isph_driver <case.bin> <state_out.bin> [fused] [snapshots=DIR]
!               [wave=STROKE,PERIOD,RAMP,T0]
!
! Reads a case (run parameters, particles, porous solids, end time and
! display interval) from a binary stream file, runs the production loop of
! the reference (main.f90:199-261) through the C ABI, one call per reference
! subroutine in the reference's order, prints one line per display interval
! and writes the final particle state to <state_out.bin>.  With the argument
! "fused" each iteration is one isph_step call instead.
!
! With wave=STROKE,PERIOD,RAMP,T0 the walls flagged as paddle (flag bit 8)
! are a piston wavemaker (isph_set_wavemaker, after the upload); the
! per-subroutine loop moves the paddle after projection (isph_wavemaker), as
! isph_step does.  The reference has no wavemaker.
!
! With snapshots=DIR the driver writes the reference's output files under DIR
! (whose subdirectories fluid/, buffer/, boundary/ and ghost/ must exist, as
! the reference expects): boundary/boundary_1000.txt, ghost/ghost_1000.txt
! and porous.txt once at the start (output.f90:11-136), and at every display
! time and after the loop fluid/fluid_<1000+iter>.txt and
! buffer/buffer_<1000+iter>.txt (output.f90:188-222, main.f90:263-274), with
! the same edit descriptors.  The library keeps particles in the reference's
! cell order, so the lines come out in the order the reference writes them.
!
! Case file (stream access, native endianness):
!   int32 magic 'ISPH' (0x48505349), int32 version 1
!   isph_params (the C struct, 24 + 23*8 bytes)
!   int64 n, int64 ns
!   float64 x(n), y(n), u(n), v(n), p(n), m(n), rho(n), rho0(n), vol0(n), c(n)
!   int32 tid(n), int8 flags(n)
!   float64 sx(ns), sy(ns), sm(ns), srho(ns)
!   float64 time, displaytime
! State file: int32 magic, int32 version, int64 n, the ten float64 arrays,
! tid and flags in the same layout, then float64 t, dt.
program isph_main
  use iso_c_binding
  use iso_fortran_env, only: output_unit
  use isph_gpu
  implicit none

  integer(c_int32_t), parameter :: MAGIC = int(z'48505349', c_int32_t)
  type(isph_params) :: prm
  type(isph_scalars) :: sc
  type(isph_solve_stats) :: st
  type(c_ptr) :: ctx = c_null_ptr
  character(len=4096) :: inpath, outpath, mode
  integer :: u, j1, iter, ios
  integer(c_int32_t) :: magic_in, version
  integer(c_int64_t) :: n, ns, nout
  real(c_double) :: time, displaytime, told
  real(c_double), allocatable :: x(:), y(:), vx(:), vy(:), p(:), m(:), rho(:), rho0(:)
  real(c_double), allocatable :: vol0(:), con(:), sx(:), sy(:), sm(:), srho(:)
  integer(c_int32_t), allocatable :: tid(:)
  integer(c_int8_t), allocatable :: flags(:)
  logical :: fused
  integer(c_int64_t) :: nsteps
  character(len=4096) :: snapdir, arg
  logical :: snaps, wave
  real(c_double) :: wv(4)
  integer :: ia

  if (command_argument_count() < 2) then
    write(*, '(a)') 'usage: isph_driver <case.bin> <state_out.bin> [fused] [snapshots=DIR] ' // &
                    '[wave=STROKE,PERIOD,RAMP,T0]'
    stop 2
  end if
  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  fused = .false.
  snaps = .false.
  snapdir = ''
  wave = .false.
  do ia = 3, command_argument_count()
    call get_command_argument(ia, arg)
    if (trim(arg) == 'fused') then
      fused = .true.
    else if (arg(1:10) == 'snapshots=') then
      snaps = .true.
      snapdir = arg(11:)
    else if (arg(1:5) == 'wave=') then
      read(arg(6:), *, iostat=ios) wv
      if (ios /= 0) then
        write(*, '(2a)') 'cannot read four numbers from ', trim(arg)
        stop 2
      end if
      wave = .true.
    else
      write(*, '(2a)') 'unknown argument ', trim(arg)
      stop 2
    end if
  end do
  mode = ''

  open(newunit=u, file=trim(inpath), access='stream', form='unformatted', status='old', &
       action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(2a)') 'cannot open ', trim(inpath)
    stop 2
  end if
  read(u) magic_in, version
  if (magic_in /= MAGIC .or. version /= 1) then
    write(*, '(a)') 'not an ISPH case file (version 1)'
    stop 2
  end if
  read(u) prm
  read(u) n, ns
  allocate(x(n), y(n), vx(n), vy(n), p(n), m(n), rho(n), rho0(n), vol0(n), con(n))
  allocate(tid(n), flags(n), sx(max(ns, 1_c_int64_t)), sy(max(ns, 1_c_int64_t)))
  allocate(sm(max(ns, 1_c_int64_t)), srho(max(ns, 1_c_int64_t)))
  read(u) x, y, vx, vy, p, m, rho, rho0, vol0, con
  read(u) tid, flags
  if (ns > 0) read(u) sx(1:ns), sy(1:ns), sm(1:ns), srho(1:ns)
  read(u) time, displaytime
  close(u)

  write(*, '(a, i0, a, i0)') '# case: particles ', n, ', porous solids ', ns
  flush(output_unit)

  ! the state the reference holds when its production loop starts
  call check(isph_create(prm, n, max(ns, 1_c_int64_t), ctx), 'isph_create')
  call check(isph_upload(ctx, n, x, y, vx, vy, p, m, rho, rho0, vol0, con, tid, flags), &
             'isph_upload')
  if (ns > 0) call check(isph_upload_solids(ctx, ns, sx, sy, sm, srho), 'isph_upload_solids')
  if (wave) call check(isph_set_wavemaker(ctx, wv(1), wv(2), wv(3), wv(4)), 'isph_set_wavemaker')
  call check(isph_init_state(ctx), 'isph_init_state')

  if (snaps) then
    call print_walls(1_c_int32_t, 'boundary/boundary_')
    call print_walls(2_c_int32_t, 'ghost/ghost_')
    call print_porous()
  end if

  write(*, '(a)') '#        t            dt      solsteps  iterations    rel_residual        n'
  iter = 1
  nsteps = 0
  call check(isph_get_scalars(ctx, sc), 'isph_get_scalars')
  do while (sc%t < time)
    told = sc%t
    if (fused) then
      call check(isph_step(ctx, 1_c_int32_t, st), 'isph_step')
    else
      sc%t = sc%t + sc%dt
      call check(isph_set_scalars(ctx, sc), 'isph_set_scalars')
      do j1 = 1, sc%solsteps
        call check(isph_scalart_anisotropic(ctx), 'scalart_anisotropic')
        call check(isph_scalarupdate(ctx, sc%dtsol), 'scalarupdate')
      end do
      call check(isph_projection(ctx), 'projection')
      if (wave) call check(isph_wavemaker(ctx), 'wavemaker')
      call check(isph_cellshift(ctx), 'cellshift')
      call check(isph_resetid(ctx), 'resetid')
      call check(isph_neighbour(ctx), 'neighbour')
      call check(isph_effpor(ctx), 'effpor')
      call check(isph_eddyvis(ctx), 'eddyvis')
      call check(isph_int_vel(ctx), 'int_vel')
      call check(isph_ppesolve(ctx, st), 'ppesolve')
      call check(isph_comp_vel(ctx), 'comp_vel')
      call check(isph_comp_pos(ctx), 'comp_pos')
      call check(isph_cellshiftalt(ctx), 'cellshiftalt')
      call check(isph_neighbour(ctx), 'neighbour')
      call check(isph_freesurf(ctx), 'freesurf')
      call check(isph_effpor(ctx), 'effpor')
      call check(isph_opt2_shift(ctx), 'opt2_shift')
      call check(isph_densityupdate(ctx), 'densityupdate')
      call check(isph_timestep(ctx), 'timestep')
      call check(isph_eddyvis(ctx), 'eddyvis')
    end if
    nsteps = nsteps + 1
    call check(isph_get_scalars(ctx, sc), 'isph_get_scalars')
    if (told < iter * displaytime .and. sc%t >= iter * displaytime) then
      write(*, '(es16.8, es14.6, i8, i12, es16.6, i10)') sc%t, sc%dt, sc%solsteps, &
          st%iterations, st%rel_residual, sc%n
      flush(output_unit)
      if (snaps) call combined(iter)
      iter = iter + 1
    end if
  end do
  write(*, '(a, i0, a)') '# ', nsteps, ' steps'
  if (snaps) call combined(iter)

  call check(isph_download(ctx, nout, x, y, vx, vy, p, m, rho, rho0, vol0, con, tid, flags), &
             'isph_download')
  open(newunit=u, file=trim(outpath), access='stream', form='unformatted', status='replace', &
       action='write')
  write(u) MAGIC, 1_c_int32_t, nout
  write(u) x(1:nout), y(1:nout), vx(1:nout), vy(1:nout), p(1:nout), m(1:nout), rho(1:nout)
  write(u) rho0(1:nout), vol0(1:nout), con(1:nout), tid(1:nout), flags(1:nout)
  write(u) sc%t, sc%dt
  close(u)
  call check(isph_destroy(ctx), 'isph_destroy')

contains

  ! fluid/fluid_<1000+it>.txt and buffer/buffer_<1000+it>.txt (output.f90:188-222):
  ! x, y, pressure, vx, vy and the concentration con + 0.5 of every fluid
  ! particle, buffer particles in the second file
  subroutine combined(it)
    integer, intent(in) :: it
    integer :: uf, ub, k
    character(len=4200) :: f1, f2
    call check(isph_download(ctx, nout, x, y, vx, vy, p, m, rho, rho0, vol0, con, tid, flags), &
               'isph_download')
    write(f1, '(a, "/fluid/fluid_", i0, ".txt")') trim(snapdir), 1000 + it
    write(f2, '(a, "/buffer/buffer_", i0, ".txt")') trim(snapdir), 1000 + it
    open(newunit=uf, file=trim(f1), status='replace', action='write')
    open(newunit=ub, file=trim(f2), status='replace', action='write')
    do k = 1, int(nout)
      if (tid(k) /= ISPH_TID_FLUID) cycle
      if (iand(flags(k), ISPH_F_BUFFER) == 0) then
        write(uf, '(F10.3,1X,F10.3,1X,F10.3,1X,ES10.3,1X,ES10.3,1X,F10.5)') x(k), y(k), p(k), &
            vx(k), vy(k), con(k) + 0.50_c_double
      else
        write(ub, '(F10.3,1X,F10.3,1X,F10.3,1X,ES10.3,1X,ES10.3,1X,F10.5)') x(k), y(k), p(k), &
            vx(k), vy(k), con(k) + 0.50_c_double
      end if
    end do
    close(uf)
    close(ub)
  end subroutine combined

  ! boundary/boundary_1000.txt (tid 1) or ghost/ghost_1000.txt (tid 2):
  ! x, y, pressure (output.f90:11-61)
  subroutine print_walls(which, stem)
    integer(c_int32_t), intent(in) :: which
    character(len=*), intent(in) :: stem
    integer :: uf, k
    character(len=4200) :: f1
    call check(isph_download(ctx, nout, x, y, vx, vy, p, m, rho, rho0, vol0, con, tid, flags), &
               'isph_download')
    write(f1, '(a, "/", a, i0, ".txt")') trim(snapdir), stem, 1000
    open(newunit=uf, file=trim(f1), status='replace', action='write')
    do k = 1, int(nout)
      if (tid(k) == which) write(uf, '(F10.3,1X,F10.3,1X,F10.3)') x(k), y(k), p(k)
    end do
    close(uf)
  end subroutine print_walls

  ! porous.txt: x, y of the porous solids (output.f90:114-136), in upload order
  subroutine print_porous()
    integer :: uf, k
    open(newunit=uf, file=trim(snapdir)//'/porous.txt', status='replace', action='write')
    do k = 1, int(ns)
      write(uf, '(F10.3,1X,F10.3)') sx(k), sy(k)
    end do
    close(uf)
  end subroutine print_porous

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= ISPH_OK) then
      ! with a null handle (a failed isph_create) the message is the create's
      write(*, '(3a, i0, 2a)') 'error in ', what, ': code ', rc, ': ', isph_error_message(ctx)
      flush(output_unit)
      error stop 1
    end if
  end subroutine check

end program isph_main
