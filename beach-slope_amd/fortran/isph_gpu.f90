! isph_gpu.f90 -- iso_c_binding interface to the C ABI of the device ISPH
! pipeline (include/isph_abi.h).
!
! A Fortran driver that used the module subroutines of the reference
! (projection, cellshift, ..., timestep; called from main.f90:210-260) calls
! the functions below instead, passing the context handle that owns the
! device-resident state.  Every function returns 0 or a negative ISPH_E* code.
! The derived types are interoperable with the C structs of the same name,
! so they can be filled here and passed by reference.
module isph_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: ISPH_OK = 0, ISPH_EHIP = -1, ISPH_EARG = -2, &
      ISPH_ECAP = -3, ISPH_ENONFINITE = -4, ISPH_ENOCONV = -5, ISPH_ESTATE = -6
  integer(c_int32_t), parameter, public :: ISPH_TID_FIXED = 1, ISPH_TID_GHOST = 2, &
      ISPH_TID_FLUID = 3
  integer(c_int8_t), parameter, public :: ISPH_F_BUFFER = 1_c_int8_t, &
      ISPH_F_FREE = 2_c_int8_t, ISPH_F_VICINITY = 4_c_int8_t
  integer(c_int32_t), parameter, public :: ISPH_PPE_FP64 = 0, ISPH_PPE_FP32_REFINE = 1
  integer, parameter, public :: ISPH_NBR_CAP = 64

  ! run parameters (the module variables of initialize.f90:21-83)
  type, bind(C), public :: isph_params
    integer(c_int32_t) :: cellx = 0, celly = 0
    integer(c_int32_t) :: fixed_dt = 1
    integer(c_int32_t) :: precision = 0
    integer(c_int32_t) :: max_iter = 20000
    integer(c_int32_t) :: max_outer = 20
    real(c_double) :: h = 0, dl = 0, lam = 0, rho = 1000, rel_den = 1.025d0, rhomin = 1000
    real(c_double) :: g = -9.81d0, mu = 1.0d-3, por = 0.4d0, d50 = 0.01d0, lamk = 0
    real(c_double) :: csh = 0, maxshift = 0, fs_thresh = 0, sig1 = 0, tschmidt = 0
    real(c_double) :: xlcutoff = 0, ytcutoff = 0, buf_xmin = 0, buf_xmax = 0
    real(c_double) :: dt = 1.0d-4, tol = 1.0d-10, inner_tol = 0
  end type isph_params

  type, bind(C), public :: isph_solve_stats
    integer(c_int32_t) :: iterations = 0, outer = 0, converged = 0, breakdown = 0
    real(c_double) :: rel_residual = 0, bnorm = 0, nnz = 0
  end type isph_solve_stats

  ! module variables t, dt, dtsol, umax, numax, solsteps plus the live count
  type, bind(C), public :: isph_scalars
    real(c_double) :: t = 0, dt = 0, dtsol = 0, umax = 0, numax = 0
    integer(c_int32_t) :: solsteps = 0
    integer(c_int32_t) :: n = 0
  end type isph_scalars

  public :: isph_create, isph_destroy, isph_last_error, isph_set_params
  public :: isph_upload, isph_upload_solids, isph_download
  public :: isph_get_scalars, isph_set_scalars
  public :: isph_projection, isph_cellshift, isph_resetid, isph_neighbour, isph_effpor
  public :: isph_eddyvis, isph_int_vel, isph_ppesolve, isph_comp_vel, isph_comp_pos
  public :: isph_cellshiftalt, isph_freesurf, isph_opt2_shift, isph_densityupdate
  public :: isph_timestep, isph_scalart_anisotropic, isph_scalarupdate
  public :: isph_init_state, isph_step, isph_footprint_bytes
  public :: isph_set_buffer, isph_buffer_update
  public :: isph_set_wcsph, isph_presstoden, isph_dentopress, isph_calmasscons, isph_updateden
  public :: isph_wavemaker_motion, isph_set_wavemaker, isph_wavemaker
  public :: isph_error_message

  interface
    integer(c_int) function isph_create(params, n_cap, n_solid_cap, ctx) bind(C, name="isph_create")
      import :: c_int, c_int64_t, c_ptr, isph_params
      type(isph_params), intent(in) :: params
      integer(c_int64_t), value :: n_cap, n_solid_cap
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function isph_destroy(ctx) bind(C, name="isph_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function isph_last_error(ctx) bind(C, name="isph_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function isph_set_params(ctx, params) bind(C, name="isph_set_params")
      import :: c_int, c_ptr, isph_params
      type(c_ptr), value :: ctx
      type(isph_params), intent(in) :: params
    end function

    integer(c_int) function isph_upload(ctx, n, x, y, u, v, p, m, rho, rho0, vol0, c, tid, flags) &
        bind(C, name="isph_upload")
      import :: c_int, c_int64_t, c_ptr, c_double, c_int32_t, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: x(*), y(*), u(*), v(*), p(*), m(*), rho(*), rho0(*)
      real(c_double), intent(in) :: vol0(*), c(*)
      integer(c_int32_t), intent(in) :: tid(*)
      integer(c_int8_t), intent(in) :: flags(*)
    end function

    integer(c_int) function isph_upload_solids(ctx, ns, x, y, m, rho) bind(C, name="isph_upload_solids")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ns
      real(c_double), intent(in) :: x(*), y(*), m(*), rho(*)
    end function

    integer(c_int) function isph_download(ctx, n, x, y, u, v, p, m, rho, rho0, vol0, c, tid, flags) &
        bind(C, name="isph_download")
      import :: c_int, c_int64_t, c_ptr, c_double, c_int32_t, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: n
      real(c_double), intent(out) :: x(*), y(*), u(*), v(*), p(*), m(*), rho(*), rho0(*)
      real(c_double), intent(out) :: vol0(*), c(*)
      integer(c_int32_t), intent(out) :: tid(*)
      integer(c_int8_t), intent(out) :: flags(*)
    end function

    integer(c_int) function isph_get_scalars(ctx, s) bind(C, name="isph_get_scalars")
      import :: c_int, c_ptr, isph_scalars
      type(c_ptr), value :: ctx
      type(isph_scalars), intent(out) :: s
    end function

    integer(c_int) function isph_set_scalars(ctx, s) bind(C, name="isph_set_scalars")
      import :: c_int, c_ptr, isph_scalars
      type(c_ptr), value :: ctx
      type(isph_scalars), intent(in) :: s
    end function

    ! one entry per reference subroutine
    integer(c_int) function isph_projection(ctx) bind(C, name="isph_projection")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_cellshift(ctx) bind(C, name="isph_cellshift")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_resetid(ctx) bind(C, name="isph_resetid")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_neighbour(ctx) bind(C, name="isph_neighbour")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_effpor(ctx) bind(C, name="isph_effpor")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_eddyvis(ctx) bind(C, name="isph_eddyvis")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_int_vel(ctx) bind(C, name="isph_int_vel")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_ppesolve(ctx, stats) bind(C, name="isph_ppesolve")
      import :: c_int, c_ptr, isph_solve_stats
      type(c_ptr), value :: ctx
      type(isph_solve_stats), intent(out) :: stats
    end function
    integer(c_int) function isph_comp_vel(ctx) bind(C, name="isph_comp_vel")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_comp_pos(ctx) bind(C, name="isph_comp_pos")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_cellshiftalt(ctx) bind(C, name="isph_cellshiftalt")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_freesurf(ctx) bind(C, name="isph_freesurf")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_opt2_shift(ctx) bind(C, name="isph_opt2_shift")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_densityupdate(ctx) bind(C, name="isph_densityupdate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_timestep(ctx) bind(C, name="isph_timestep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_scalart_anisotropic(ctx) bind(C, name="isph_scalart_anisotropic")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_scalarupdate(ctx, stime) bind(C, name="isph_scalarupdate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: stime
    end function

    ! inflow/outflow buffer zones (insert_buffer1/2, remove_buffer1/2)
    integer(c_int) function isph_set_buffer(ctx, zone, nt, x, y, u, v, p, mass, rho, vol0, c, &
        rv, interval, time_shift) bind(C, name="isph_set_buffer")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: zone
      integer(c_int64_t), value :: nt
      real(c_double), intent(in) :: x(*), y(*), u(*), v(*), p(*)
      real(c_double), value :: mass, rho, vol0, c, rv, interval, time_shift
    end function
    integer(c_int) function isph_buffer_update(ctx, zone) bind(C, name="isph_buffer_update")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: zone
    end function

    ! weakly compressible variant (wcsph.f90)
    integer(c_int) function isph_set_wcsph(ctx, co, t_gam, delt, swl, in_step) &
        bind(C, name="isph_set_wcsph")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: co, t_gam, delt, swl
      integer(c_int32_t), value :: in_step
    end function
    integer(c_int) function isph_presstoden(ctx) bind(C, name="isph_presstoden")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_dentopress(ctx) bind(C, name="isph_dentopress")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_calmasscons(ctx) bind(C, name="isph_calmasscons")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_updateden(ctx) bind(C, name="isph_updateden")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! piston wavemaker (no counterpart in the reference): the ISPH_F_PADDLE
    ! (flag bit 8) walls move by X(t); isph_step moves them after projection
    integer(c_int) function isph_wavemaker_motion(stroke, period, ramp, t0, t, x, u) &
        bind(C, name="isph_wavemaker_motion")
      import :: c_int, c_double
      real(c_double), value :: stroke, period, ramp, t0, t
      real(c_double), intent(out) :: x, u
    end function
    integer(c_int) function isph_set_wavemaker(ctx, stroke, period, ramp, t0) &
        bind(C, name="isph_set_wavemaker")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: stroke, period, ramp, t0
    end function
    integer(c_int) function isph_wavemaker(ctx) bind(C, name="isph_wavemaker")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! fused entries
    integer(c_int) function isph_init_state(ctx) bind(C, name="isph_init_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function isph_step(ctx, nsteps, stats) bind(C, name="isph_step")
      import :: c_int, c_int32_t, c_ptr, isph_solve_stats
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      type(isph_solve_stats), intent(out) :: stats
    end function

    integer(c_int64_t) function isph_footprint_bytes(n_cap, n_solid_cap, cellx, celly, precision) &
        bind(C, name="isph_footprint_bytes")
      import :: c_int64_t, c_int32_t
      integer(c_int64_t), value :: n_cap, n_solid_cap
      integer(c_int32_t), value :: cellx, celly, precision
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! message of the context's last failure as a Fortran string; for a null
  ! handle, the message of the calling thread's last failed isph_create
  function isph_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: cs
    character(kind=c_char), pointer :: chars(:)
    integer :: k, l
    cs = isph_last_error(ctx)
    if (.not. c_associated(cs)) then
      msg = ''
      return
    end if
    l = int(c_strlen(cs))
    call c_f_pointer(cs, chars, [l])
    allocate(character(len=l) :: msg)
    do k = 1, l
      msg(k:k) = chars(k)
    end do
  end function isph_error_message

end module isph_gpu
