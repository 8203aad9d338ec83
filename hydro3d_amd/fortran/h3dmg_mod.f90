! h3dmg_mod.f90 -- ISO_C_BINDING interface of the MI355X pressure-Poisson
! multigrid (include/h3dmg.h) for Fortran hosts such as Hydro3D.
!
! Block indices are 0-based at the C boundary (block ib of a Fortran loop is
! ib-1).  Host arrays are passed as they are stored in the reference
! (dom(ib)%p(ttc_i,ttc_j,ttc_k), src_v70/localparameters.for:140-143): the
! leading dimensions ld1 = ttc_i, ld2 = ttc_j and the 0-based offset of the
! first interior cell off = isp-1 (= pl).
module h3dmg_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  ! status codes (h3dmg_status)
  integer(c_int), parameter, public :: H3DMG_OK = 0, H3DMG_ERR_ARG = 1, &
       H3DMG_ERR_LEVELS = 2, H3DMG_ERR_DEVICE = 3, H3DMG_ERR_NOMEM = 4, &
       H3DMG_ERR_NOT_CONVERGED = 5, H3DMG_ERR_DIVERGED = 6, &
       H3DMG_ERR_UNSUPPORTED = 7, H3DMG_ERR_BUSY = 8
  integer(c_int), parameter, public :: H3DMG_BC_WALL = 0, H3DMG_BC_PERIODIC = 1, &
       H3DMG_BC_NEIGHBOUR = 2
  integer(c_int), parameter, public :: H3DMG_SMOOTH_GS = 1, H3DMG_SMOOTH_TDMA = 2, &
       H3DMG_SMOOTH_RBGS = 3
  integer(c_int), parameter, public :: H3DMG_CRIT_REL_L2 = 0, H3DMG_CRIT_REF_RMAX = 1
  integer(c_int), parameter, public :: H3DMG_FP64 = 0, H3DMG_MIXED = 1
  integer(c_int), parameter, public :: H3DMG_COARSE_SMOOTH = 0, H3DMG_COARSE_EXACT = 1

  type, bind(C), public :: h3dmg_block
     integer(c_int) :: ni, nj, nk
     real(c_double) :: dx, dy, dz
     integer(c_int) :: nlevels
     integer(c_int) :: bc(6)
     integer(c_int) :: nbr(6)   ! 0-based block index across a NEIGHBOUR face
  end type h3dmg_block

  type, bind(C), public :: h3dmg_params
     integer(c_int) :: smoother, nu1, nu2, maxcy, criterion
     real(c_double) :: tol, dt
     integer(c_int) :: precision, coarse, min_cycles, use_graph
  end type h3dmg_params

  public :: h3dmg_default_params, h3dmg_create, h3dmg_destroy, h3dmg_device_bytes, &
       h3dmg_set_rhs, h3dmg_set_p, h3dmg_get_p, h3dmg_vcycle, h3dmg_get_pp, &
       h3dmg_solve, h3dmg_block_stats, h3dmg_remove_mean, h3dmg_set_stream, &
       h3dmg_set_march, h3dmg_get_march, h3dmg_set_lanes, h3dmg_get_lanes, &
       h3dmg_set_tail, h3dmg_get_tail, &
       h3dmg_set_block_tol, h3dmg_set_block_enabled, h3dmg_project_device, &
       h3dmg_velocity_begin, h3dmg_velocity_step, h3dmg_solve_velocity, &
       h3dmg_get_su_device, h3dmg_set_velocity, h3dmg_get_velocity, h3dmg_get_su, &
       h3dmg_get_velocity_calvel, h3dmg_set_differencing, &
       h3dmg_set_dt, h3dmg_set_fused, h3dmg_get_fused, h3dmg_set_density, &
       h3dmg_error_message

  interface
     subroutine h3dmg_default_params(p) bind(C, name="h3dmg_default_params")
       import :: h3dmg_params
       type(h3dmg_params), intent(out) :: p
     end subroutine h3dmg_default_params

     integer(c_int) function h3dmg_create(blocks, nblocks, prm, handle) &
          bind(C, name="h3dmg_create")
       import :: c_int, c_ptr, h3dmg_block, h3dmg_params
       type(h3dmg_block), intent(in) :: blocks(*)
       integer(c_int), value :: nblocks
       type(h3dmg_params), intent(in) :: prm
       type(c_ptr), intent(out) :: handle
     end function h3dmg_create

     subroutine h3dmg_destroy(handle) bind(C, name="h3dmg_destroy")
       import :: c_ptr
       type(c_ptr), value :: handle
     end subroutine h3dmg_destroy

     type(c_ptr) function h3dmg_last_error_c() bind(C, name="h3dmg_last_error")
       import :: c_ptr
     end function h3dmg_last_error_c

     integer(c_long_long) function h3dmg_device_bytes(handle) &
          bind(C, name="h3dmg_device_bytes")
       import :: c_ptr, c_long_long
       type(c_ptr), value :: handle
     end function h3dmg_device_bytes

     integer(c_int) function h3dmg_set_rhs(handle, b, su, ld1, ld2, off) &
          bind(C, name="h3dmg_set_rhs")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(in) :: su(*)
     end function h3dmg_set_rhs

     integer(c_int) function h3dmg_set_p(handle, b, p, ld1, ld2, off) &
          bind(C, name="h3dmg_set_p")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(in) :: p(*)
     end function h3dmg_set_p

     integer(c_int) function h3dmg_get_p(handle, b, p, ld1, ld2, off) &
          bind(C, name="h3dmg_get_p")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(inout) :: p(*)
     end function h3dmg_get_p

     integer(c_int) function h3dmg_vcycle(handle) bind(C, name="h3dmg_vcycle")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
     end function h3dmg_vcycle

     integer(c_int) function h3dmg_get_pp(handle, b, pp, ld1, ld2, off) &
          bind(C, name="h3dmg_get_pp")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(inout) :: pp(*)
     end function h3dmg_get_pp

     integer(c_int) function h3dmg_solve(handle, hist, hist_cap, ncycles, final_res) &
          bind(C, name="h3dmg_solve")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       real(c_double), intent(inout) :: hist(*)
       integer(c_int), value :: hist_cap
       integer(c_int), intent(out) :: ncycles
       real(c_double), intent(out) :: final_res
     end function h3dmg_solve

     integer(c_int) function h3dmg_block_stats(handle, cycles, rel_res) &
          bind(C, name="h3dmg_block_stats")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), intent(out) :: cycles(*)
       real(c_double), intent(out) :: rel_res(*)
     end function h3dmg_block_stats

     integer(c_int) function h3dmg_set_block_tol(handle, b, tol) &
          bind(C, name="h3dmg_set_block_tol")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b
       real(c_double), value :: tol
     end function h3dmg_set_block_tol

     integer(c_int) function h3dmg_set_block_enabled(handle, b, enabled) &
          bind(C, name="h3dmg_set_block_enabled")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
       integer(c_int), value :: b, enabled
     end function h3dmg_set_block_enabled

     integer(c_int) function h3dmg_project_device(handle, b, ustar, vstar, wstar, u, v, w, &
          su, alfapr, rmax) bind(C, name="h3dmg_project_device")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b
       type(c_ptr), value :: ustar, vstar, wstar, u, v, w, su
       real(c_double), value :: alfapr
       real(c_double), intent(out) :: rmax
     end function h3dmg_project_device

     ! velocity-form loop (newsolv_mg.for:17-59) on the device; the six list
     ! arguments are c_null_ptr to use the arrays of h3dmg_set_velocity
     integer(c_int) function h3dmg_velocity_begin(handle, ustar, vstar, wstar, u, v, w, alfapr, &
          rmax) bind(C, name="h3dmg_velocity_begin")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle, ustar, vstar, wstar, u, v, w
       real(c_double), value :: alfapr
       real(c_double), intent(out) :: rmax
     end function h3dmg_velocity_begin

     integer(c_int) function h3dmg_velocity_step(handle, rmax) bind(C, name="h3dmg_velocity_step")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       real(c_double), intent(out) :: rmax
     end function h3dmg_velocity_step

     integer(c_int) function h3dmg_solve_velocity(handle, ustar, vstar, wstar, u, v, w, alfapr, &
          hist, hist_cap, ncycles, rmax) bind(C, name="h3dmg_solve_velocity")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle, ustar, vstar, wstar, u, v, w
       real(c_double), value :: alfapr
       real(c_double), intent(inout) :: hist(*)
       integer(c_int), value :: hist_cap
       integer(c_int), intent(out) :: ncycles
       real(c_double), intent(out) :: rmax
     end function h3dmg_solve_velocity

     integer(c_int) function h3dmg_get_su_device(handle, b, su) bind(C, name="h3dmg_get_su_device")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle, su
       integer(c_int), value :: b
     end function h3dmg_get_su_device

     integer(c_int) function h3dmg_set_velocity(handle, b, ustar, vstar, wstar, u, v, w, ld1, &
          ld2, off) bind(C, name="h3dmg_set_velocity")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(in) :: ustar(*), vstar(*), wstar(*), u(*), v(*), w(*)
     end function h3dmg_set_velocity

     integer(c_int) function h3dmg_get_velocity(handle, b, u, v, w, ld1, ld2, off) &
          bind(C, name="h3dmg_get_velocity")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(inout) :: u(*), v(*), w(*)
     end function h3dmg_get_velocity

     ! calmas' branch: the reference's differencing (1: 2nd order, 2: 4th)
     integer(c_int) function h3dmg_set_differencing(handle, differencing) &
          bind(C, name="h3dmg_set_differencing")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
       integer(c_int), value :: differencing
     end function h3dmg_set_differencing

     ! the faces calvel computes (isu..ieu, jsv..jev, ksw..kew); the host's
     ! exchange(1..3) fills the others, as after the reference's calvel
     integer(c_int) function h3dmg_get_velocity_calvel(handle, b, u, v, w, ld1, ld2, off) &
          bind(C, name="h3dmg_get_velocity_calvel")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(inout) :: u(*), v(*), w(*)
     end function h3dmg_get_velocity_calvel

     integer(c_int) function h3dmg_get_su(handle, b, su, ld1, ld2, off) bind(C, name="h3dmg_get_su")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(inout) :: su(*)
     end function h3dmg_get_su

     integer(c_int) function h3dmg_set_fused(handle, max_cells) bind(C, name="h3dmg_set_fused")
       import :: c_ptr, c_int, c_long_long
       type(c_ptr), value :: handle
       integer(c_long_long), value :: max_cells
     end function h3dmg_set_fused

     integer(c_long_long) function h3dmg_get_fused(handle) bind(C, name="h3dmg_get_fused")
       import :: c_ptr, c_long_long
       type(c_ptr), value :: handle
     end function h3dmg_get_fused

     integer(c_int) function h3dmg_set_density(handle, b, dens, ld1, ld2, off) &
          bind(C, name="h3dmg_set_density")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: b, ld1, ld2, off
       real(c_double), intent(in) :: dens(*)
     end function h3dmg_set_density

     integer(c_int) function h3dmg_set_dt(handle, dt) bind(C, name="h3dmg_set_dt")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: handle
       real(c_double), value :: dt
     end function h3dmg_set_dt

     integer(c_int) function h3dmg_set_tail(handle, level) bind(C, name="h3dmg_set_tail")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
       integer(c_int), value :: level
     end function h3dmg_set_tail

     integer(c_int) function h3dmg_get_tail(handle) bind(C, name="h3dmg_get_tail")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
     end function h3dmg_get_tail

     integer(c_int) function h3dmg_remove_mean(handle) bind(C, name="h3dmg_remove_mean")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
     end function h3dmg_remove_mean

     integer(c_int) function h3dmg_set_stream(handle, stream) bind(C, name="h3dmg_set_stream")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle, stream
     end function h3dmg_set_stream

     integer(c_int) function h3dmg_set_march(handle, level, planes) &
          bind(C, name="h3dmg_set_march")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
       integer(c_int), value :: level, planes
     end function h3dmg_set_march

     integer(c_int) function h3dmg_get_march(handle, level) bind(C, name="h3dmg_get_march")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
       integer(c_int), value :: level
     end function h3dmg_get_march

     integer(c_int) function h3dmg_set_lanes(handle, level, lanes) &
          bind(C, name="h3dmg_set_lanes")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
       integer(c_int), value :: level, lanes
     end function h3dmg_set_lanes

     integer(c_int) function h3dmg_get_lanes(handle, level) bind(C, name="h3dmg_get_lanes")
       import :: c_ptr, c_int
       type(c_ptr), value :: handle
       integer(c_int), value :: level
     end function h3dmg_get_lanes
  end interface

contains

  ! The library's last error message as a Fortran string.
  function h3dmg_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: cp
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    cp = h3dmg_last_error_c()
    if (.not. c_associated(cp)) then
       msg = ''
       return
    end if
    call c_f_pointer(cp, chars, [4096])
    n = 0
    do while (n < 4096)
       if (chars(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function h3dmg_error_message

end module h3dmg_mod
