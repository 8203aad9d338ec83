! newsolv_mg_h3d.f90 -- drop-in replacement of Hydro3D's pressure solve
! (src_v70/newsolv_mg.for:2-93) that runs the multigrid on the GPU.
!
! Use: compile this file and h3dmg_mod.f90 with the reference's modules, link
! libh3dmg.so, and call newsolv_mg_h3d where src_v70/flosol.for:215 calls
! newsolv_mg.  Nothing else in the LES driver changes: the host steps of the
! reference routine stay on the host -- calvel / calmas (press.for:2-144),
! exchange(4), pbound, boundu/v/w and the log lines -- and only the cycle
! loop (newsolv_mg.for:24-59) moves to the device.
!
! Mapping.  The reference iterates p <- p + mgkcyc(su) with su = div(u)/dt
! recomputed from the corrected velocity after every cycle
! (newsolv_mg.for:26-55); div(u* - dt grad p)/dt = f - A p, with A the
! 7-point operator of mgsolver.for (calvel excludes the wall-normal faces,
! localparameters.for:187-192, which makes div(grad) equal to A with Neumann
! walls).  So one device solve of A pp = su0 from pp = 0, with
! su0 = calmas(calvel(p_old)), under the reference criterion (rmax < eps and
! iter > 1, rmax = max |r| dt dx dy dz) gives p = p_old + pp, after which
! the host recomputes u and rmax exactly as the reference ends its loop.
!
! Scope: the constant-coefficient RBGS multigrid on blocks without
! neighbours (every face a wall or periodic onto itself), one MPI rank per
! GPU.  Anything else stops with a message rather than running a different
! algorithm.  The precision is FP64 unless the environment variable
! H3DMG_PRECISION=mixed selects the FP32 V-cycle inside FP64 refinement.
subroutine newsolv_mg_h3d
  use vars
  use mpi
  use multidata
  use h3dmg_mod
  use, intrinsic :: iso_c_binding
  implicit none
  type(c_ptr), save :: h = c_null_ptr
  real(c_double), allocatable, save :: hist(:)
  integer :: ib
  integer(c_int) :: st, ncyc
  real(c_double) :: fin
  logical :: converged
  double precision :: wtimedum

  if (L_LSM) call h3d_stop('the variable-density (LSM) operator is not on the device path')
  if (.not. c_associated(h)) call setup()

  ! su0 = f - A p_old from the current pressure (newsolv_mg.for:17-18)
  call calvel
  call calmas
  do ib = 1, nbp
     dom(ib)%pp = 0.0d0
     st = h3dmg_set_rhs(h, ib - 1, dom(ib)%su, dom(ib)%ttc_i, dom(ib)%ttc_j, dom(ib)%isp - 1)
     call chk(st, 'set_rhs')
     st = h3dmg_set_p(h, ib - 1, dom(ib)%pp, dom(ib)%ttc_i, dom(ib)%ttc_j, dom(ib)%isp - 1)
     call chk(st, 'set_p')
  end do

  st = h3dmg_solve(h, hist, size(hist, kind=c_int), ncyc, fin)
  converged = st == H3DMG_OK
  if (st /= H3DMG_OK .and. st /= H3DMG_ERR_NOT_CONVERGED .and. st /= H3DMG_ERR_DIVERGED) &
       call chk(st, 'solve')

  ! p = p + pp over the interior and one ghost layer (newsolv_mg.for:39-49,
  ! ppref = 0)
  do ib = 1, nbp
     st = h3dmg_get_p(h, ib - 1, dom(ib)%pp, dom(ib)%ttc_i, dom(ib)%ttc_j, dom(ib)%isp - 1)
     call chk(st, 'get_p')
     associate (d => dom(ib))
       d%p(d%isp - 1:d%iep + 1, d%jsp - 1:d%jep + 1, d%ksp - 1:d%kep + 1) = &
            d%p(d%isp - 1:d%iep + 1, d%jsp - 1:d%jep + 1, d%ksp - 1:d%kep + 1) + &
            d%pp(d%isp - 1:d%iep + 1, d%jsp - 1:d%jep + 1, d%ksp - 1:d%kep + 1)
       d%pp = 0.0d0
     end associate
  end do
  call exchange(4)
  call pbound
  call calvel
  call calmas

  if (converged) then
     iter = ncyc
  else
     iter = maxcy + 1
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
    type(h3dmg_block), allocatable :: blk(:)
    type(h3dmg_params) :: prm
    character(len=16) :: pv
    integer :: b, plen, pst
    allocate (blk(nbp))
    do b = 1, nbp
       associate (d => dom(b))
         if (d%iprev >= 0 .or. d%inext >= 0 .or. d%jprev >= 0 .or. d%jnext >= 0 .or. &
             d%kprev >= 0 .or. d%knext >= 0) then
            if (.not. self_periodic(d)) &
                 call h3d_stop('blocks with neighbours need the halo exchange, not on the device path')
         end if
         blk(b)%ni = d%iep - d%isp + 1
         blk(b)%nj = d%jep - d%jsp + 1
         blk(b)%nk = d%kep - d%ksp + 1
         blk(b)%dx = d%dx
         blk(b)%dy = d%dy
         blk(b)%dz = d%dz
         blk(b)%nlevels = ngrd_gl
         blk(b)%bc = [bcmap(d%bc_west), bcmap(d%bc_east), bcmap(d%bc_south), &
                      bcmap(d%bc_north), bcmap(d%bc_bottom), bcmap(d%bc_top)]
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
    call get_environment_variable('H3DMG_PRECISION', pv, plen, pst)
    if (pst == 0 .and. pv(1:5) == 'mixed') then
       prm%precision = H3DMG_MIXED
       prm%coarse = H3DMG_COARSE_EXACT
    end if
    st = h3dmg_create(blk, int(nbp, c_int), prm, h)
    call chk(st, 'create')
    allocate (hist(0:maxcy))
  end subroutine setup

  ! a block whose only neighbour on a periodic axis is itself
  logical function self_periodic(d)
    type(multidom), intent(in) :: d
    self_periodic = (d%iprev < 0 .or. d%bc_west == 5) .and. (d%inext < 0 .or. d%bc_east == 5) .and. &
         (d%jprev < 0 .or. d%bc_south == 5) .and. (d%jnext < 0 .or. d%bc_north == 5) .and. &
         (d%kprev < 0 .or. d%bc_bottom == 5) .and. (d%knext < 0 .or. d%bc_top == 5) .and. &
         nprocs == 1 .and. nbp == 1
  end function self_periodic

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
