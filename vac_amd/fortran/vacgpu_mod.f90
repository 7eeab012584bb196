!==============================================================================
! vacgpu_mod -- Fortran 2003 ISO_C_BINDING interface to libvacgpu (the MI355X
! advect path), for a Fortran 90 host such as VAC.
!
! The derived types are interoperable with the structs of include/vacgpu.h.
! C's  int typeB[6][8]  is  typeB(8,6)  here, i.e. typeB(iw,iB) exactly as
! VAC's own typeB array (src/vacdef.t:110-112).  Pointers that C allows to be
! NULL are passed as type(c_ptr), value (c_loc(x) or c_null_ptr).
!==============================================================================
module vacgpu_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: VACGPU_OK = 0, VACGPU_E_ARG = -1, VACGPU_E_HIP = -2, &
                               VACGPU_E_STATE = -3, VACGPU_E_NOMEM = -4
  integer(c_int), parameter :: VACGPU_HD = 0, VACGPU_MHD = 1
  integer(c_int), parameter :: VACGPU_TVDLF = 0, VACGPU_TVD = 1, VACGPU_TVDMU = 2
  integer(c_int), parameter :: VACGPU_MINMOD = 0, VACGPU_WOODWARD = 1, VACGPU_SUPERBEE = 2
  integer(c_int), parameter :: VACGPU_CONT = 0, VACGPU_PERIODIC = 1, VACGPU_SYMM = 2, &
       VACGPU_ASYMM = 3, VACGPU_FIXED = 4, VACGPU_CONT1 = 5, VACGPU_FIXED1 = 6, VACGPU_GRAD1 = 7, &
       VACGPU_SYMM0 = 8
  integer(c_int), parameter :: VACGPU_FP64 = 0, VACGPU_FP32BULK = 1
  integer(c_int), parameter :: VACGPU_LIM_ORIGINAL = 0, VACGPU_LIM_PREVIOUS = 1, &
       VACGPU_LIM_PREDICTOR = 2
  integer(c_int), parameter :: VACGPU_ENT_NUL = 0, VACGPU_ENT_HARTEN = 1, VACGPU_ENT_POWELL = 2, &
       VACGPU_ENT_RATIO = 3, VACGPU_ENT_YEE = 4

  ! the drop-in's device grid (made by the first advect or projection), and
  ! whether a process() call comes from inside advect
  type(c_ptr), save :: vacgpu_dropin_grid = c_null_ptr
  logical, save :: vacgpu_in_advect = .false.
  ! where the current state is: the device state is current once a step has
  ! uploaded VAC's w and stays current across steps; VAC's w is current
  ! until the first sweep and again after a fetch (before a save, or every
  ! step when a host routine reads w between steps)
  logical, save :: vacgpu_dev_current = .false.
  logical, save :: vacgpu_host_current = .true.
  ! host-side transfers of the state so far (reported by vacgpu_dropin_report)
  integer, save :: vacgpu_n_put = 0, vacgpu_n_get = 0
  ! print a line per transfer (environment variable VACGPU_DROPIN_TRACE)
  logical, save :: vacgpu_trace = .false.

  type, bind(C) :: vacgpu_config
     integer(c_int) :: ndim
     integer(c_int) :: physics
     integer(c_int) :: nx(3)
     real(c_double) :: eqpar(2)
     integer(c_int) :: typeB(8, 6)
     integer(c_int) :: precision
  end type vacgpu_config

  type, bind(C) :: vacgpu_method
     integer(c_int) :: typefull
     integer(c_int) :: typelimiter(8)
     integer(c_int) :: divbfix
     integer(c_int) :: divbwave
     integer(c_int) :: sourceunsplit
     integer(c_int) :: dimsplit_xy
     real(c_double) :: courantpar
     real(c_double) :: smallpcoeff
     integer(c_int) :: typelimited
     integer(c_int) :: typeentropy(8)
     real(c_double) :: entropycoef(8)
     integer(c_int) :: useprimitive
     integer(c_int) :: unsplit
     integer(c_int) :: dtcantgrow
     integer(c_int) :: slowsteps
  end type vacgpu_method

  interface
     integer(c_int) function vacgpu_create(cfg, dx0, dx1, dx2, grid) bind(C, name='vacgpu_create')
       import :: c_int, c_ptr, vacgpu_config
       type(vacgpu_config), intent(in) :: cfg
       type(c_ptr), value :: dx0, dx1, dx2
       type(c_ptr), intent(out) :: grid
     end function vacgpu_create

     integer(c_int) function vacgpu_set_stream(grid, stream) bind(C, name='vacgpu_set_stream')
       import :: c_int, c_ptr
       type(c_ptr), value :: grid, stream
     end function vacgpu_set_stream

     integer(c_int) function vacgpu_put_state(grid, w, ld) bind(C, name='vacgpu_put_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: grid
       real(c_double), intent(in) :: w(*)
       integer(c_int), intent(in) :: ld(*)
     end function vacgpu_put_state

     integer(c_int) function vacgpu_put_state_ghosts(grid, w, ld) bind(C, name='vacgpu_put_state_ghosts')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: grid
       real(c_double), intent(in) :: w(*)
       integer(c_int), intent(in) :: ld(*)
     end function vacgpu_put_state_ghosts

     integer(c_int) function vacgpu_get_state(grid, w, ld) bind(C, name='vacgpu_get_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: grid
       real(c_double), intent(inout) :: w(*)
       integer(c_int), intent(in) :: ld(*)
     end function vacgpu_get_state

     integer(c_int) function vacgpu_getdt_courant(grid, m, dtcourant, smallp, dt) &
          bind(C, name='vacgpu_getdt_courant')
       import :: c_int, c_ptr, c_double, vacgpu_method
       type(c_ptr), value :: grid
       type(vacgpu_method), intent(in) :: m
       type(c_ptr), value :: dtcourant, smallp
       real(c_double), intent(out) :: dt
     end function vacgpu_getdt_courant

     integer(c_int) function vacgpu_advect(grid, m, idimmin, idimmax, dt, firstsweep, lastsweep, &
          dtcourant, nerror, smallp) bind(C, name='vacgpu_advect')
       import :: c_int, c_ptr, c_double, vacgpu_method
       type(c_ptr), value :: grid
       type(vacgpu_method), intent(in) :: m
       integer(c_int), value :: idimmin, idimmax, firstsweep, lastsweep
       real(c_double), value :: dt
       type(c_ptr), value :: dtcourant, nerror, smallp
     end function vacgpu_advect

     integer(c_int) function vacgpu_run(grid, m, nsteps, itmax, tmax, tmaxexact, dtpar, t_hist) &
          bind(C, name='vacgpu_run')
       import :: c_int, c_ptr, c_double, vacgpu_method
       type(c_ptr), value :: grid
       type(vacgpu_method), intent(in) :: m
       integer(c_int), value :: nsteps, itmax, tmaxexact
       real(c_double), value :: tmax, dtpar
       type(c_ptr), value :: t_hist
     end function vacgpu_run

     integer(c_int) function vacgpu_set_time(grid, it, t) bind(C, name='vacgpu_set_time')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: grid
       integer(c_int), value :: it
       real(c_double), value :: t
     end function vacgpu_set_time

     integer(c_int) function vacgpu_get_time(grid, it, t) bind(C, name='vacgpu_get_time')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: grid
       integer(c_int), intent(out) :: it
       real(c_double), intent(out) :: t
     end function vacgpu_get_time

     integer(c_int) function vacgpu_get_scalars(grid, dtcourant, nerror, smallp, dt) &
          bind(C, name='vacgpu_get_scalars')
       import :: c_int, c_ptr
       type(c_ptr), value :: grid, dtcourant, nerror, smallp, dt
     end function vacgpu_get_scalars

     integer(c_int) function vacgpu_profile(grid, enable) bind(C, name='vacgpu_profile')
       import :: c_int, c_ptr
       type(c_ptr), value :: grid
       integer(c_int), value :: enable
     end function vacgpu_profile

     integer(c_int) function vacgpu_kernel_time(grid, k, avg_ms, count) &
          bind(C, name='vacgpu_kernel_time')
       import :: c_int, c_ptr, c_double, c_long
       type(c_ptr), value :: grid
       integer(c_int), value :: k
       real(c_double), intent(out) :: avg_ms
       integer(c_long), intent(out) :: count
     end function vacgpu_kernel_time

     integer(c_int) function vacgpu_set_kernels(grid, fused, seglen) &
          bind(C, name='vacgpu_set_kernels')
       import :: c_int, c_ptr
       type(c_ptr), value :: grid
       integer(c_int), value :: fused, seglen
     end function vacgpu_set_kernels

     integer(c_int) function vacgpu_snapshot_start(grid, slot) &
          bind(C, name='vacgpu_snapshot_start')
       import :: c_int, c_ptr
       type(c_ptr), value :: grid
       integer(c_int), value :: slot
     end function vacgpu_snapshot_start

     integer(c_int) function vacgpu_snapshot_fetch(grid, slot, w, ld, it, t) &
          bind(C, name='vacgpu_snapshot_fetch')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: grid, ld
       integer(c_int), value :: slot
       real(c_double), intent(out) :: w(*)
       integer(c_int), intent(out) :: it
       real(c_double), intent(out) :: t
     end function vacgpu_snapshot_fetch

     integer(c_int) function vacgpu_snapshot_fetch_mesh(grid, slot, w, it, t) &
          bind(C, name='vacgpu_snapshot_fetch_mesh')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: grid
       integer(c_int), value :: slot
       real(c_double), intent(out) :: w(*)
       integer(c_int), intent(out) :: it
       real(c_double), intent(out) :: t
     end function vacgpu_snapshot_fetch_mesh

     integer(c_int) function vacgpu_set_process(grid, nproc1, procpar, itmin) &
          bind(C, name='vacgpu_set_process')
       import :: c_int, c_ptr
       type(c_ptr), value :: grid
       integer(c_int), value :: nproc1, itmin
       type(c_ptr), value :: procpar
     end function vacgpu_set_process

     integer(c_int) function vacgpu_process(grid, count, procpar, info, iter, resid) &
          bind(C, name='vacgpu_process')
       import :: c_int, c_ptr
       type(c_ptr), value :: grid
       integer(c_int), value :: count
       type(c_ptr), value :: procpar, info, iter, resid
     end function vacgpu_process

     integer(c_int) function vacgpu_check_quotients(a, b, n, q_kernel, q_ieee) &
          bind(C, name='vacgpu_check_quotients')
       import :: c_int, c_double, c_long_long
       real(c_double), intent(in) :: a(*), b(*)
       integer(c_long_long), value :: n
       real(c_double), intent(out) :: q_kernel(*), q_ieee(*)
     end function vacgpu_check_quotients

     integer(c_int) function vacgpu_supported(cfg, m) bind(C, name='vacgpu_supported')
       import :: c_int, vacgpu_config, vacgpu_method
       type(vacgpu_config), intent(in) :: cfg
       type(vacgpu_method), intent(in) :: m
     end function vacgpu_supported

     type(c_ptr) function vacgpu_last_error(grid) bind(C, name='vacgpu_last_error')
       import :: c_ptr
       type(c_ptr), value :: grid
     end function vacgpu_last_error

     subroutine vacgpu_destroy(grid) bind(C, name='vacgpu_destroy')
       import :: c_ptr
       type(c_ptr), value :: grid
     end subroutine vacgpu_destroy

     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! The message of the last failed call on grid, as a Fortran string.
  function vacgpu_errmsg(grid) result(msg)
    type(c_ptr), intent(in) :: grid
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = vacgpu_last_error(grid)
    if (.not. c_associated(p)) then
       msg = 'unknown error'
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function vacgpu_errmsg

  ! VAC limiter name -> code (src/vacgrid.t:938-996); -1 if not on the GPU path
  integer(c_int) function vacgpu_limiter_code(name)
    character(len=*), intent(in) :: name
    select case (trim(name))
    case ('minmod')
       vacgpu_limiter_code = VACGPU_MINMOD
    case ('woodward')
       vacgpu_limiter_code = VACGPU_WOODWARD
    case ('superbee')
       vacgpu_limiter_code = VACGPU_SUPERBEE
    case default
       vacgpu_limiter_code = -1
    end select
  end function vacgpu_limiter_code

  ! VAC boundary type -> code (src/vacgrid.t:195-413); -1 if not on the GPU path
  integer(c_int) function vacgpu_bc_code(name)
    character(len=*), intent(in) :: name
    select case (trim(name))
    case ('cont')
       vacgpu_bc_code = VACGPU_CONT
    case ('periodic')
       vacgpu_bc_code = VACGPU_PERIODIC
    case ('symm')
       vacgpu_bc_code = VACGPU_SYMM
    case ('asymm')
       vacgpu_bc_code = VACGPU_ASYMM
    case ('fixed')
       vacgpu_bc_code = VACGPU_FIXED
    case ('cont1')
       vacgpu_bc_code = VACGPU_CONT1
    case ('fixed1')
       vacgpu_bc_code = VACGPU_FIXED1
    case ('grad1')
       vacgpu_bc_code = VACGPU_GRAD1
    case ('symm0')
       vacgpu_bc_code = VACGPU_SYMM0
    case default
       vacgpu_bc_code = -1
    end select
  end function vacgpu_bc_code

end module vacgpu_mod
