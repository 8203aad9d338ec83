! ref_multiblock.f90 -- test-only driver around the reference Fortran
! multigrid for SEVERAL blocks on one rank, with the reference's own layout
! files and neighbour search.
!
! This program is oracle infrastructure: it is linked against the reference
! objects built from src_v70 by oracle/Makefile (the multigrid, the halo
! exchanges and alloc_dom.for) and is used only to write golden fixtures and
! as the live reference in tests.  The module state comes from mb_setup.f90.
!
! Job description: namelist /job/ in the file named by argv(1); raw input
! doubles in argv(2); raw output doubles to argv(3).  Per-block arrays are
! read and written block after block in the order of the rank's mdmap list.
! Padded level arrays are (n_i+2, n_j+2, n_k+2), interior arrays
! (n_i, n_j, n_k), Fortran order.
!
! Modes:
!   nbr     : per block: dom_id, the cell counts, iprev..knext, corprev1..4,
!             cornext1..4, edgprev1..6, edgnext1..6, per_ip..per_kn
!   relax   : per block phi_g (interior), rhs_g -> exchangepp(g) (the ghost
!             state of a level after its last sweep) -> mgrelax(g, nsw)
!             -> per block phi_g (padded)
!   restr   : per block phi_{g-1} (interior), rhs_{g-1} -> exchangepp(g-1)
!             -> mgrestr(g) -> per block phi_g (padded), rhs_g
!   corr    : per block phi_g (interior), phi_{g+1} (interior) ->
!             exchangepp(g), exchangepp(g+1) -> mgcorr(g) -> phi_g (padded)
!   vcycle  : per block su -> mgkcyc -> per block pp (padded)
!   newsolv : per block u*, v*, w* face arrays and p0 -> the newsolv_mg loop
!             (newsolv_mg.for:17-59, statement by statement) -> rmax per
!             iteration, then per block p (padded) and su
!   timeloop: the same loop for exactly ncycles iterations on every rank of
!             an MPI run (mpiexec -n <ranks of mdmap.cin>), each rank reading
!             its own blocks' inputs at their place in the input (blocks of
!             one size, in domain order) -> rank 0 writes the seconds between
!             barriers around the loop (first calvel + calmas included),
!             ncycles and the last rmax
program ref_multiblock
  use mb_setup
  implicit none

  call start()
  select case (trim(mode))
  case ('nbr')
     call run_nbr()
  case ('relax')
     call run_relax()
  case ('restr')
     call run_restr()
  case ('corr')
     call run_corr()
  case ('vcycle')
     call run_vcycle()
  case ('newsolv')
     call run_newsolv()
  case ('timeloop')
     call run_timeloop()
  case default
     print *, 'ref_multiblock: unknown mode ', trim(mode)
     call MPI_ABORT(MPI_COMM_WORLD, 2, ierr)
  end select
  call finish()

contains

  ! ------------------------------------------------------------------
  subroutine run_nbr()
    integer :: b
    do b = 1, nbp
       associate (d => dom(b))
         write(uout) dble((/ dom_id(b), nc(1, b), nc(2, b), nc(3, b), &
              d%iprev, d%inext, d%jprev, d%jnext, d%kprev, d%knext, &
              d%corprev1, d%corprev2, d%corprev3, d%corprev4, &
              d%cornext1, d%cornext2, d%cornext3, d%cornext4, &
              d%edgprev1, d%edgprev2, d%edgprev3, d%edgprev4, d%edgprev5, d%edgprev6, &
              d%edgnext1, d%edgnext2, d%edgnext3, d%edgnext4, d%edgnext5, d%edgnext6, &
              d%per_ip, d%per_in, d%per_jp, d%per_jn, d%per_kp, d%per_kn /))
       end associate
    end do
  end subroutine run_nbr

  subroutine run_relax()
    integer :: b
    double precision, allocatable :: a3(:,:,:), b3(:,:,:)
    do b = 1, nbp
       call read_level_interior(b, level, a3)
       call read_level_interior(b, level, b3)
       call put_phi_interior(b, level, a3)
       call put_rhs(b, level, b3)
    end do
    call exchangepp(level)
    call mgrelax(level, nsw)
    do b = 1, nbp
       call write_phi(b, level)
    end do
  end subroutine run_relax

  subroutine run_restr()
    integer :: b, mi, mj, mk
    double precision, allocatable :: a3(:,:,:), b3(:,:,:)
    do b = 1, nbp
       call read_level_interior(b, level-1, a3)
       call read_level_interior(b, level-1, b3)
       call put_phi_interior(b, level-1, a3)
       call put_rhs(b, level-1, b3)
       ! poison the coarse level so that what mgrestr does not write shows up
       call lev_dims(b, level, mi, mj, mk)
       deallocate(a3)
       allocate(a3(mi-2, mj-2, mk-2)); a3 = 7.0d0
       call put_phi_interior(b, level, a3)
       call put_rhs(b, level, a3)
    end do
    call exchangepp(level-1)
    call mgrestr(level)
    do b = 1, nbp
       call write_phi(b, level)
       call write_rhs(b, level)
    end do
  end subroutine run_restr

  subroutine run_corr()
    integer :: b
    double precision, allocatable :: a3(:,:,:), b3(:,:,:)
    do b = 1, nbp
       call read_level_interior(b, level, a3)
       call read_level_interior(b, level+1, b3)
       call put_phi_interior(b, level, a3)
       call put_phi_interior(b, level+1, b3)
    end do
    call exchangepp(level)
    call exchangepp(level+1)
    call mgcorr(level)
    do b = 1, nbp
       call write_phi(b, level)
    end do
  end subroutine run_corr

  subroutine run_vcycle()
    integer :: b, i, j, k
    double precision, allocatable :: a3(:,:,:), pad(:,:,:)
    do b = 1, nbp
       call read_level_interior(b, 1, a3)
       dom(b)%pp = 0.0d0
       do k = 1, nc(3, b); do j = 1, nc(2, b); do i = 1, nc(1, b)
          dom(b)%su(i+pl, j+pl, k+pl) = a3(i, j, k)
       end do; end do; end do
    end do
    call mgkcyc
    do b = 1, nbp
       ! copy-out range of src_v70/mgsolver.for:354-367
       allocate(pad(nc(1, b)+2, nc(2, b)+2, nc(3, b)+2))
       do k = 1, nc(3, b)+2; do j = 1, nc(2, b)+2; do i = 1, nc(1, b)+2
          pad(i, j, k) = dom(b)%pp(i+pl-1, j+pl-1, k+pl-1)
       end do; end do; end do
       write(uout) pad
       deallocate(pad)
    end do
  end subroutine run_vcycle

  ! newsolv_mg.for:17-59 replayed statement by statement (newsolv_mg.o pulls
  ! boundu/v/w -> wall_function.for, which flang rejects), with the
  ! reference's calvel/calmas/pbound/exchange over all blocks of the rank.
  ! early = 0 runs all ncycles iterations; early = 1 leaves the loop where
  ! newsolv_mg does (rmax < eps .and. iter > 1, newsolv_mg.for:57) and also
  ! writes iter, rmax and the whole u, v, w arrays.
  subroutine run_newsolv()
    double precision, allocatable :: hist(:)
    integer :: b, i, j, k, it, stop_at
    double precision :: ppref

    call read_velocity_state()
    allocate(hist(0:ncycles))
    hist = 0.0d0
    stop_at = -1
    call calvel
    call calmas
    hist(0) = rmax
    do b = 1, nbp
       dom(b)%pp = 0.0d0
    end do
    do it = 1, ncycles
       iter = it
       call mgkcyc
       ppref = 0.0d0
       do b = 1, nbp
          do k = dom(b)%ksp-1, dom(b)%kep+1
             do i = dom(b)%isp-1, dom(b)%iep+1
                do j = dom(b)%jsp-1, dom(b)%jep+1
                   dom(b)%p(i,j,k) = (dom(b)%p(i,j,k) + (dom(b)%pp(i,j,k) - ppref))
                   dom(b)%pp(i,j,k) = 0.0
                end do
             end do
          end do
       end do
       call exchange(4)
       call pbound
       call calvel
       call calmas
       hist(it) = rmax
       if (stop_at < 0 .and. rmax < eps .and. it > 1) then
          stop_at = it
          if (early /= 0) exit
       end if
    end do
    if (early /= 0) then
       ! a DO loop that runs out leaves its variable one past maxcy
       if (stop_at < 0) iter = ncycles + 1
       write(uout) dble(iter), rmax
       call write_pressure_state()
       call write_velocity_arrays()
       return
    end if
    write(uout) dble(stop_at)
    write(uout) hist
    call write_pressure_state()
  end subroutine run_newsolv

  subroutine run_timeloop()
    double precision, allocatable :: uf(:,:,:), vf(:,:,:), wf(:,:,:), p0(:,:,:)
    integer :: b, i, j, k, it, isp, jsp, ksp, ni, nj, nk
    integer(kind=8) :: per_block, pos
    double precision :: ppref, t0, t1
    do b = 1, nbp
       ni = nc(1, b); nj = nc(2, b); nk = nc(3, b)
       per_block = 8_8 * (int(ni + 1, 8) * nj * nk + int(ni, 8) * (nj + 1) * nk + &
                          int(ni, 8) * nj * (nk + 1) + int(ni, 8) * nj * nk)
       pos = 1_8 + int(dom_id(b), 8) * per_block
       isp = dom(b)%isp; jsp = dom(b)%jsp; ksp = dom(b)%ksp
       allocate(uf(0:ni, nj, nk), vf(ni, 0:nj, nk), wf(ni, nj, 0:nk), p0(ni, nj, nk))
       read(uin, pos=pos) uf
       read(uin) vf
       read(uin) wf
       read(uin) p0
       do k = 1, nk; do j = 1, nj; do i = 0, ni
          dom(b)%ustar(isp-1+i, jsp-1+j, ksp-1+k) = uf(i, j, k)
       end do; end do; end do
       do k = 1, nk; do j = 0, nj; do i = 1, ni
          dom(b)%vstar(isp-1+i, jsp-1+j, ksp-1+k) = vf(i, j, k)
       end do; end do; end do
       do k = 0, nk; do j = 1, nj; do i = 1, ni
          dom(b)%wstar(isp-1+i, jsp-1+j, ksp-1+k) = wf(i, j, k)
       end do; end do; end do
       dom(b)%u = dom(b)%ustar; dom(b)%v = dom(b)%vstar; dom(b)%w = dom(b)%wstar
       do k = 1, nk; do j = 1, nj; do i = 1, ni
          dom(b)%p(isp-1+i, jsp-1+j, ksp-1+k) = p0(i, j, k)
       end do; end do; end do
       deallocate(uf, vf, wf, p0)
    end do
    call exchange(4)
    call pbound
    call MPI_BARRIER(MPI_COMM_WORLD, ierr)
    t0 = MPI_WTIME()
    call calvel
    call calmas
    do b = 1, nbp
       dom(b)%pp = 0.0d0
    end do
    do it = 1, ncycles
       iter = it
       call mgkcyc
       ppref = 0.0d0
       do b = 1, nbp
          do k = dom(b)%ksp-1, dom(b)%kep+1
             do i = dom(b)%isp-1, dom(b)%iep+1
                do j = dom(b)%jsp-1, dom(b)%jep+1
                   dom(b)%p(i,j,k) = (dom(b)%p(i,j,k) + (dom(b)%pp(i,j,k) - ppref))
                   dom(b)%pp(i,j,k) = 0.0
                end do
             end do
          end do
       end do
       call exchange(4)
       call pbound
       call calvel
       call calmas
    end do
    call MPI_BARRIER(MPI_COMM_WORLD, ierr)
    t1 = MPI_WTIME()
    if (myrank == 0) write(uout) t1 - t0, dble(ncycles), rmax
  end subroutine run_timeloop

end program ref_multiblock
