! newsolv_mg_h3d.f90 -- drop-in replacement of Hydro3D's pressure solve
! (src_v70/newsolv_mg.for:2-93) that runs the multigrid on the GPU.
!
! Use: compile this file and h3dmg_mod.f90 with the reference's modules, link
! libh3dmg.so, and call newsolv_mg_h3d where src_v70/flosol.for:215 calls
! newsolv_mg.  Nothing else in the LES driver changes.
!
! The reference's loop itself runs on the device (h3dmg_velocity_begin /
! h3dmg_velocity_step): calvel, calmas, then per iteration mgkcyc from su,
! p = p + pp, calvel, calmas, in the reference's arithmetic.  u*, v*, w*,
! u, v, w and p go to the device once per call; each iteration returns this
! rank's rmax, which is reduced over the ranks here as calmas reduces it
! (press.for:72-73).  After every device calvel the faces it computed come
! back and the host runs calvel's own exchange(1..3) (press.for:139-141), so
! the velocity ghost layers evolve exactly as in the reference's loop (its
! exchanges carry data from call to call in their send/receive buffers);
! su and p come back once, followed by exchange(4) and pbound.  The host
! keeps the not-converged / BIG RMAX handling, boundu/v/w and the log
! lines.
!
! Scope: the multigrid with any of the reference's smoothers (mg_itrsch 1
! lexicographic GS, 2 line TDMA, 3 red-black GS), constant-coefficient or
! variable-density (L_LSM: the density goes to the device every call, as the
! reference rebuilds its coefficients every solve, newsolv_mg.for:13-15), on
! independent blocks; and the red-black multigrid with constant coefficients
! on the blocks of a rank that share faces (same-level neighbours and
! periodic partners; the halo exchange is the device reading the
! neighbour's cells), one MPI rank per GPU; calmas in either branch
! (differencing 1 or 2, press.for:17-63).  Anything else -- a neighbour on
! another rank, LMR refinement, the 4th-order branch under L_LSM -- stops
! with a message rather than running a different algorithm.  The precision is FP64 (bitwise the reference loop)
! unless the environment variable H3DMG_PRECISION=mixed selects the FP32
! V-cycle with the exact coarsest solve inside the FP64 velocity loop
! (constant coefficients and red-black GS only).
subroutine newsolv_mg_h3d
  use vars
  use mpi
  use multidata
  use h3dmg_mod
  use, intrinsic :: iso_c_binding
  implicit none
  type(c_ptr), save :: h = c_null_ptr
  type(h3dmg_block), allocatable :: blk(:)   ! the block table of setup()
  integer :: ib
  integer(c_int) :: st
  real(c_double) :: rloc
  logical :: converged
  double precision :: wtimedum
  ! the host's velocity halo exchange after every device iteration (the
  ! default: ghost layers bitwise the reference loop's), or once after the
  ! loop (H3DMG_VELOCITY_HALO=final: one download instead of one per
  ! iteration; every face of the interior rows is still the reference's)
  logical, save :: every_halo = .true.

  if (.not. c_associated(h)) call setup()
  st = h3dmg_set_dt(h, dt)
  call chk(st, 'set_dt')

  do ib = 1, nbp
     associate (d => dom(ib))
       st = h3dmg_set_velocity(h, ib - 1, d%ustar, d%vstar, d%wstar, d%u, d%v, d%w, &
                               d%ttc_i, d%ttc_j, d%isp - 1)
       call chk(st, 'set_velocity')
       st = h3dmg_set_p(h, ib - 1, d%p, d%ttc_i, d%ttc_j, d%isp - 1)
       call chk(st, 'set_p')
       if (L_LSM) then
          st = h3dmg_set_density(h, ib - 1, d%dens, d%ttc_i, d%ttc_j, d%isp - 1)
          call chk(st, 'set_density')
       end if
       d%pp = 0.0d0
     end associate
  end do

  ! calvel; calmas (newsolv_mg.for:17-18)
  st = h3dmg_velocity_begin(h, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                            c_null_ptr, alfapr, rloc)
  call chk(st, 'velocity_begin')
  if (every_halo) call velocity_halo()
  call MPI_ALLREDUCE(rloc, rmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, MPI_COMM_WORLD, ierr)

  ! newsolv_mg.for:24-59
  converged = .false.
  do iter = 1, maxcy
     st = h3dmg_velocity_step(h, rloc)
     call chk(st, 'velocity_step')
     if (every_halo) call velocity_halo()
     call MPI_ALLREDUCE(rloc, rmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, MPI_COMM_WORLD, ierr)
     if (rmax < eps .and. iter > 1) then
        converged = .true.
        exit
     end if
  end do

  do ib = 1, nbp
     associate (d => dom(ib))
       st = h3dmg_get_p(h, ib - 1, d%p, d%ttc_i, d%ttc_j, d%isp - 1)
       call chk(st, 'get_p')
       if (L_LSM) then
          st = h3dmg_get_su(h, ib - 1, d%sup, d%ttc_i, d%ttc_j, d%isp - 1)
       else
          st = h3dmg_get_su(h, ib - 1, d%su, d%ttc_i, d%ttc_j, d%isp - 1)
       end if
       call chk(st, 'get_su')
     end associate
  end do
  call exchange(4)
  call pbound
  if (.not. every_halo) call velocity_halo()

  if (.not. converged) then
     if (myrank == 0) print *, 'not converged!! ', maxcy, rmax
     if (myrank == 0) write (numfile, *) 'not converged!! ', maxcy, rmax
  end if

  if (rmax > 10.0d0) then
     if (myrank == 0) then
        write (numfile, *) 'BIG RMAX!! STOP!!!!!', rmax
        open (unit=101, file='final_ctime.dat')
        write (101, '(i8,3F15.6)') ntime, ctime, forcn, qstpn
        close (101)
     end if
     call MPI_BARRIER(MPI_COMM_WORLD, ierr)
     stop
  end if

  call boundu
  call boundv
  call boundw

  if (myrank == 0) then
     wtimedum = MPI_WTIME() - wtime
     write (6, '(1x,a,i8,a,e13.6,a,i8)') ' ntime:', ntime, ' rmax:', rmax, '    iter', iter
     write (numfile, '(1x,a,i8,a,e13.6,a,i8)') ' ntime:', ntime, ' rmax:', rmax, '    iter', iter
     write (numfile2, '(i8,f15.6,3e20.6)') ntime, wtimedum, rmax, dt, Mdef
  end if

contains

  subroutine setup()
    type(h3dmg_params) :: prm
    character(len=16) :: pv
    integer :: b, plen, pst
    allocate (blk(nbp))
    do b = 1, nbp
       associate (d => dom(b))
         ! every block coarsens to the global level count (no LMR levels)
         if (d%ngrid /= ngrd_gl) &
              call h3d_stop('blocks with their own level count (LMR, rdiv > 1) are not on the device path')
         blk(b)%ni = d%iep - d%isp + 1
         blk(b)%nj = d%jep - d%jsp + 1
         blk(b)%nk = d%kep - d%ksp + 1
         blk(b)%dx = d%dx
         blk(b)%dy = d%dy
         blk(b)%dz = d%dz
         blk(b)%nlevels = d%ngrid
         ! faces shared with another block of this rank (alloc_dom.for:524-690)
         ! or, on a periodic axis spanning several blocks, with the periodic
         ! partner (alloc_dom.for:1040-1215); the device reads the neighbour's
         ! cells where exchangepp would copy them
         call face(b, 1, d%iprev, d%inext, d%per_ip, d%bc_west)
         call face(b, 2, d%inext, d%iprev, d%per_in, d%bc_east)
         call face(b, 3, d%jprev, d%jnext, d%per_jp, d%bc_south)
         call face(b, 4, d%jnext, d%jprev, d%per_jn, d%bc_north)
         call face(b, 5, d%kprev, d%knext, d%per_kp, d%bc_bottom)
         call face(b, 6, d%knext, d%kprev, d%per_kn, d%bc_top)
       end associate
    end do
    call h3dmg_default_params(prm)
    prm%smoother = mg_itrsch
    prm%nu1 = irestr
    prm%nu2 = iproln
    prm%maxcy = maxcy
    prm%criterion = H3DMG_CRIT_REF_RMAX
    prm%tol = eps
    prm%dt = dt
    prm%min_cycles = 2
    call get_environment_variable('H3DMG_VELOCITY_HALO', pv, plen, pst)
    every_halo = .not. (pst == 0 .and. pv(1:5) == 'final')
    call get_environment_variable('H3DMG_PRECISION', pv, plen, pst)
    if (pst == 0 .and. pv(1:5) == 'mixed' .and. .not. L_LSM .and. mg_itrsch == 3) then
       prm%precision = H3DMG_MIXED
       prm%coarse = H3DMG_COARSE_EXACT
    end if
    st = h3dmg_create(blk, int(nbp, c_int), prm, h)
    call chk(st, 'create')
    ! calmas' branch (press.for:17-37 or :48-63); the 4th-order stencil's
    ! second ghost faces at walls come with h3dmg_set_velocity (pl >= 2)
    st = h3dmg_set_differencing(h, int(differencing, c_int))
    call chk(st, 'set_differencing')
  end subroutine setup

  ! u, v, w of the last device calvel on the faces calvel computes, then the
  ! exchange(1..3) that ends the reference's calvel (press.for:139-141): the
  ! ghost layers evolve as in the reference's loop, whose exchanges also
  ! carry data between calls in their send/receive buffers
  subroutine velocity_halo()
    integer :: b
    do b = 1, nbp
       associate (d => dom(b))
         st = h3dmg_get_velocity_calvel(h, b - 1, d%u, d%v, d%w, d%ttc_i, d%ttc_j, d%isp - 1)
         call chk(st, 'get_velocity_calvel')
       end associate
    end do
    call exchange(1)
    call exchange(2)
    call exchange(3)
  end subroutine velocity_halo

  ! face f (1..6) of block b: `across` is the neighbour's domain number
  ! (-1: none), `opposite` the neighbour on the other face of the axis,
  ! `partner` the periodic partner, `code` the boundary code (5 periodic)
  subroutine face(b, f, across, opposite, partner, code)
    integer, intent(in) :: b, f, across, opposite, partner, code
    integer :: other
    blk(b)%nbr(f) = -1
    other = across
    if (other < 0 .and. code == 5) then
       if (opposite < 0) then
          blk(b)%bc(f) = H3DMG_BC_PERIODIC   ! the block spans the axis
          return
       end if
       other = partner
    end if
    if (other < 0) then
       blk(b)%bc(f) = H3DMG_BC_WALL
       return
    end if
    if (dom_ad(other) /= myrank) &
         call h3d_stop('a block''s neighbour on another rank: the exchange between ranks is not on the device path')
    blk(b)%bc(f) = H3DMG_BC_NEIGHBOUR
    blk(b)%nbr(f) = dom_indid(other) - 1
  end subroutine face

  ! bc code 5 is periodic, every other code a Neumann wall for the pressure
  ! (mgsolver.for:194-205)
  integer(c_int) function bcmap(code)
    integer, intent(in) :: code
    if (code == 5) then
       bcmap = H3DMG_BC_PERIODIC
    else
       bcmap = H3DMG_BC_WALL
    end if
  end function bcmap

  subroutine chk(s, what)
    integer(c_int), intent(in) :: s
    character(len=*), intent(in) :: what
    if (s /= H3DMG_OK) call h3d_stop('h3dmg_' // what // ': ' // h3dmg_error_message())
  end subroutine chk

  subroutine h3d_stop(msg)
    character(len=*), intent(in) :: msg
    if (myrank == 0) then
       print *, 'newsolv_mg_h3d: ', msg
       write (numfile, *) 'newsolv_mg_h3d: ', msg
    end if
    call MPI_ABORT(MPI_COMM_WORLD, 1, ierr)
    stop 1
  end subroutine h3d_stop

end subroutine newsolv_mg_h3d
