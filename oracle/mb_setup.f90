! mb_setup.f90 -- module state of several blocks on one rank for the
! reference-linked test drivers (ref_multiblock.f90, dropin_driver.f90).
!
! TEST INFRASTRUCTURE: linked only with the reference objects that
! oracle/Makefile builds from src_v70.  Set-up follows src_v70/fdstag.for:
! 10-22: read_mdmap and read_infodom (alloc_dom.for:284-1458) on the
! mdmap.cin / infodom.cin of the working directory, the block boundary codes
! of read_control (initial.for:98-100), then the sizes, index ranges, level
! storage and message buffers of localparameters (localparameters.for:
! 19-265) and initial (initial.for:250-268) for LMR = 1 and rdiv = 1, and
! coeff (mgsolver.for:2-224).
module mb_setup
  use vars
  use multidata
  use mpi
  implicit none

  character(len=16)  :: mode
  integer :: nlev, nu1, nu2, level, nsw, ncycles, bcs(6), scheme, pl_extra, early, cds4
  double precision :: gdx(3), dtin, epsin
  ! cds4 = 1: differencing = 2 (calmas' 4th-order divergence, press.for:17-37);
  ! the velocity input then holds faces -1..n+1 of every axis per block
  namelist /job/ mode, nlev, nu1, nu2, level, nsw, ncycles, bcs, gdx, dtin, epsin, scheme, &
       pl_extra, early, cds4

  integer :: uin, uout
  integer, allocatable :: nc(:,:)   ! cells per axis of every block

contains

  ! namelist defaults, job file, MPI, module state; opens the input and
  ! output streams
  subroutine start()
    character(len=512) :: fjob, fin, fout
    mode = 'nbr'; nlev = 1; nu1 = 2; nu2 = 2; level = 1; nsw = 1; ncycles = 1
    bcs = 1; gdx = 1.0d0; dtin = 1.0d-3; epsin = 1.0d-6; scheme = 3; pl_extra = 1; early = 0
    cds4 = 0
    call get_command_argument(1, fjob)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    open(newunit=uin, file=trim(fjob), status='old', action='read')
    read(uin, nml=job)
    close(uin)
    call MPI_INIT(ierr)
    call MPI_COMM_RANK(MPI_COMM_WORLD, myrank, ierr)
    call MPI_COMM_SIZE(MPI_COMM_WORLD, nprocs, ierr)
    MPI_FLT = MPI_DOUBLE_PRECISION
    call setup()
    open(newunit=uin, file=trim(fin), access='stream', form='unformatted', &
         status='old', action='read')
    open(newunit=uout, file=trim(fout), access='stream', form='unformatted', &
         status='replace', action='write')
  end subroutine start

  subroutine finish()
    close(uin)
    close(uout)
    call MPI_FINALIZE(ierr)
  end subroutine finish

  subroutine setup()
    integer :: b, g, gl, mi, mj, mk, tt(3), nmax, nemax
    double precision :: n1, n2, n3

    call read_mdmap
    LMR = 1; pl_ex = pl_extra; pl = 1 + pl_ex
    ngrid_input = nlev; ngrd_gl = nlev
    maxcy = ncycles; irestr = nu1; iproln = nu2; mg_itrsch = scheme
    solver = 2; L_LSM = .false.; L_LSMbase = .false.
    differencing = 1; dt = dtin; alfapr = 1.0d0; eps = epsin
    if (cds4 /= 0) differencing = 2
    g_dx = gdx(1); g_dy = gdx(2); g_dz = gdx(3)
    bc_w = bcs(1); bc_e = bcs(2); bc_s = bcs(3); bc_n = bcs(4)
    bc_b = bcs(5); bc_t = bcs(6)
    do b = 1, nbp
       ! read_control, initial.for:98-100
       dom(b)%bc_west = bc_w; dom(b)%bc_east = bc_e
       dom(b)%bc_south = bc_s; dom(b)%bc_north = bc_n
       dom(b)%bc_bottom = bc_b; dom(b)%bc_top = bc_t
    end do
    call read_infodom

    allocate(rdv(0:num_domains-1, ngrd_gl)); rdv = 1
    allocate(nc(3, nbp))
    nmax = 0; nemax = 0
    do b = 1, nbp
       n1 = (dom(b)%xel - dom(b)%xsl)/g_dx
       n2 = (dom(b)%yel - dom(b)%ysl)/g_dy
       n3 = (dom(b)%zel - dom(b)%zsl)/g_dz
       nc(:, b) = (/ nint(n1), nint(n2), nint(n3) /)
       dom(b)%ngrid = nlev
       dom(b)%dx = g_dx; dom(b)%dy = g_dy; dom(b)%dz = g_dz
       dom(b)%isp = 1 + pl; dom(b)%iep = nc(1, b) + pl
       dom(b)%jsp = 1 + pl; dom(b)%jep = nc(2, b) + pl
       dom(b)%ksp = 1 + pl; dom(b)%kep = nc(3, b) + pl
       dom(b)%isu = dom(b)%isp; dom(b)%isv = dom(b)%isp; dom(b)%isw = dom(b)%isp
       dom(b)%ieu = dom(b)%iep; dom(b)%iev = dom(b)%iep; dom(b)%iew = dom(b)%iep
       dom(b)%jsu = dom(b)%jsp; dom(b)%jsv = dom(b)%jsp; dom(b)%jsw = dom(b)%jsp
       dom(b)%jeu = dom(b)%jep; dom(b)%jev = dom(b)%jep; dom(b)%jew = dom(b)%jep
       dom(b)%ksu = dom(b)%ksp; dom(b)%ksv = dom(b)%ksp; dom(b)%ksw = dom(b)%ksp
       dom(b)%keu = dom(b)%kep; dom(b)%kev = dom(b)%kep; dom(b)%kew = dom(b)%kep
       ! localparameters.for:187-192
       if (dom(b)%inext < 0 .and. dom(b)%bc_east /= 5) dom(b)%ieu = dom(b)%ieu - 1
       if (dom(b)%jnext < 0 .and. dom(b)%bc_north /= 5) dom(b)%jev = dom(b)%jev - 1
       if (dom(b)%knext < 0 .and. dom(b)%bc_top /= 5) dom(b)%kew = dom(b)%kew - 1
       tt = nc(:, b) + 2*pl
       dom(b)%ttc_i = tt(1); dom(b)%ttc_j = tt(2); dom(b)%ttc_k = tt(3)
       dom(b)%ttc_ijk = tt(1)*tt(2)*tt(3)
       nmax = max(tt(1)*tt(2), tt(1)*tt(3), tt(2)*tt(3), nmax)
       nemax = max(tt(1), tt(2), tt(3), nemax)
       dom(b)%nwork = 0; dom(b)%nvars = 0
       do g = 1, ngrd_gl
          gl = min(g, dom(b)%ngrid)
          mi = nc(1, b)/2**(gl-1) + 2; mj = nc(2, b)/2**(gl-1) + 2
          mk = nc(3, b)/2**(gl-1) + 2
          dom(b)%nwork = dom(b)%nwork + mi*mj*mk + 8*(mi-2)*(mj-2)*(mk-2)
          dom(b)%nvars = dom(b)%nvars + mi*mj*mk
       end do
       allocate(dom(b)%cof(dom(b)%nwork)); dom(b)%cof = 0.0d0
       allocate(dom(b)%faz(ngrd_gl))
       allocate(dom(b)%p(tt(1),tt(2),tt(3)), dom(b)%pp(tt(1),tt(2),tt(3)))
       allocate(dom(b)%su(tt(1),tt(2),tt(3)), dom(b)%sup(tt(1),tt(2),tt(3)))
       allocate(dom(b)%u(tt(1),tt(2),tt(3)), dom(b)%v(tt(1),tt(2),tt(3)))
       allocate(dom(b)%w(tt(1),tt(2),tt(3)))
       allocate(dom(b)%ustar(tt(1),tt(2),tt(3)), dom(b)%vstar(tt(1),tt(2),tt(3)))
       allocate(dom(b)%wstar(tt(1),tt(2),tt(3)))
       dom(b)%p = 0; dom(b)%pp = 0; dom(b)%su = 0; dom(b)%sup = 0
       dom(b)%u = 0; dom(b)%v = 0; dom(b)%w = 0
       dom(b)%ustar = 0; dom(b)%vstar = 0; dom(b)%wstar = 0
    end do
    ! localparameters.for:259-265
    ngg = nmax
    nge = nemax*(pl+1)**2
    ngc = (pl+1)**3
    do b = 1, nbp
       ! MPI tags of the block's receives (initial.for:139-141)
       do g = 1, 26
          dom(b)%tg(g) = g*tag_offset1 + dom_id(b)
       end do
       allocate(dom(b)%stfcinf(6, pl, ngg))
       allocate(dom(b)%sendb_m1(ngg), dom(b)%sendb_p1(ngg))
       allocate(dom(b)%recvb_m1(ngg), dom(b)%recvb_p1(ngg))
       allocate(dom(b)%sendb_m2(ngg), dom(b)%sendb_p2(ngg))
       allocate(dom(b)%recvb_m2(ngg), dom(b)%recvb_p2(ngg))
       allocate(dom(b)%sendb_m3(ngg), dom(b)%sendb_p3(ngg))
       allocate(dom(b)%recvb_m3(ngg), dom(b)%recvb_p3(ngg))
       allocate(dom(b)%sc1m(ngc), dom(b)%sc1p(ngc), dom(b)%rc1m(ngc), dom(b)%rc1p(ngc))
       allocate(dom(b)%sc2m(ngc), dom(b)%sc2p(ngc), dom(b)%rc2m(ngc), dom(b)%rc2p(ngc))
       allocate(dom(b)%sc3m(ngc), dom(b)%sc3p(ngc), dom(b)%rc3m(ngc), dom(b)%rc3p(ngc))
       allocate(dom(b)%sc4m(ngc), dom(b)%sc4p(ngc), dom(b)%rc4m(ngc), dom(b)%rc4p(ngc))
       allocate(dom(b)%se1m(nge), dom(b)%se1p(nge), dom(b)%re1m(nge), dom(b)%re1p(nge))
       allocate(dom(b)%se2m(nge), dom(b)%se2p(nge), dom(b)%re2m(nge), dom(b)%re2p(nge))
       allocate(dom(b)%se3m(nge), dom(b)%se3p(nge), dom(b)%re3m(nge), dom(b)%re3p(nge))
       allocate(dom(b)%se4m(nge), dom(b)%se4p(nge), dom(b)%re4m(nge), dom(b)%re4p(nge))
       allocate(dom(b)%se5m(nge), dom(b)%se5p(nge), dom(b)%re5m(nge), dom(b)%re5p(nge))
       allocate(dom(b)%se6m(nge), dom(b)%se6p(nge), dom(b)%re6m(nge), dom(b)%re6p(nge))
    end do
    call coeff
  end subroutine setup

  ! ------------------------------------------------------------------
  subroutine lev_dims(b, g, mi, mj, mk)
    integer, intent(in) :: b, g
    integer, intent(out) :: mi, mj, mk
    integer :: gl
    gl = min(g, dom(b)%ngrid)
    mi = nc(1, b)/2**(gl-1) + 2; mj = nc(2, b)/2**(gl-1) + 2; mk = nc(3, b)/2**(gl-1) + 2
  end subroutine lev_dims

  ! level storage of mgsolver.for:23-41,412-420 (j fastest) <-> i-fastest
  subroutine put_phi_interior(b, g, arr)
    integer, intent(in) :: b, g
    double precision, intent(in) :: arr(:,:,:)
    integer :: mi, mj, mk, i, j, k, base
    call lev_dims(b, g, mi, mj, mk)
    base = dom(b)%faz(g) - mi*mj*mk
    do k = 2, mk-1; do i = 2, mi-1; do j = 2, mj-1
       dom(b)%cof(base + (k-1)*mi*mj + (i-1)*mj + j) = arr(i-1, j-1, k-1)
    end do; end do; end do
  end subroutine put_phi_interior

  subroutine write_phi(b, g)
    integer, intent(in) :: b, g
    integer :: mi, mj, mk, i, j, k, base
    double precision, allocatable :: arr(:,:,:)
    call lev_dims(b, g, mi, mj, mk)
    allocate(arr(mi, mj, mk))
    base = dom(b)%faz(g) - mi*mj*mk
    do k = 1, mk; do i = 1, mi; do j = 1, mj
       arr(i, j, k) = dom(b)%cof(base + (k-1)*mi*mj + (i-1)*mj + j)
    end do; end do; end do
    write(uout) arr
  end subroutine write_phi

  subroutine put_rhs(b, g, arr)
    integer, intent(in) :: b, g
    double precision, intent(in) :: arr(:,:,:)
    integer :: mi, mj, mk, i, j, k, incr
    call lev_dims(b, g, mi, mj, mk)
    incr = (mi-2)*(mj-2)*(mk-2)
    do k = 2, mk-1; do i = 2, mi-1; do j = 2, mj-1
       dom(b)%cof(dom(b)%faz(g) + 7*incr + (k-2)*(mi-2)*(mj-2) + (i-2)*(mj-2) + (j-1)) = &
            arr(i-1, j-1, k-1)
    end do; end do; end do
  end subroutine put_rhs

  subroutine write_rhs(b, g)
    integer, intent(in) :: b, g
    integer :: mi, mj, mk, i, j, k, incr
    double precision, allocatable :: arr(:,:,:)
    call lev_dims(b, g, mi, mj, mk)
    allocate(arr(mi-2, mj-2, mk-2))
    incr = (mi-2)*(mj-2)*(mk-2)
    do k = 2, mk-1; do i = 2, mi-1; do j = 2, mj-1
       arr(i-1, j-1, k-1) = dom(b)%cof(dom(b)%faz(g) + 7*incr + (k-2)*(mi-2)*(mj-2) + &
            (i-2)*(mj-2) + (j-1))
    end do; end do; end do
    write(uout) arr
  end subroutine write_rhs

  subroutine read_level_interior(b, g, arr)
    integer, intent(in) :: b, g
    double precision, allocatable, intent(out) :: arr(:,:,:)
    integer :: mi, mj, mk
    call lev_dims(b, g, mi, mj, mk)
    allocate(arr(mi-2, mj-2, mk-2))
    read(uin) arr
  end subroutine read_level_interior


  ! per block u*, v*, w* face arrays and p0 from the input stream (face f of
  ! cell i, its east face, lives at index i+pl; face 0 at isp-1); u = u*;
  ! then exchange(4) and pbound, the state newsolv_mg starts from
  subroutine read_velocity_state()
    double precision, allocatable :: uf(:,:,:), vf(:,:,:), wf(:,:,:), p0(:,:,:)
    integer :: b, i, j, k, isp, jsp, ksp, ni, nj, nk, fo
    fo = 0
    if (cds4 /= 0) fo = 1
    do b = 1, nbp
       ni = nc(1, b); nj = nc(2, b); nk = nc(3, b)
       isp = dom(b)%isp; jsp = dom(b)%jsp; ksp = dom(b)%ksp
       allocate(uf(-fo:ni+fo, nj, nk), vf(ni, -fo:nj+fo, nk), wf(ni, nj, -fo:nk+fo))
       allocate(p0(ni, nj, nk))
       read(uin) uf
       read(uin) vf
       read(uin) wf
       read(uin) p0
       do k = 1, nk; do j = 1, nj; do i = -fo, ni + fo
          dom(b)%ustar(isp-1+i, jsp-1+j, ksp-1+k) = uf(i, j, k)
       end do; end do; end do
       do k = 1, nk; do j = -fo, nj + fo; do i = 1, ni
          dom(b)%vstar(isp-1+i, jsp-1+j, ksp-1+k) = vf(i, j, k)
       end do; end do; end do
       do k = -fo, nk + fo; do j = 1, nj; do i = 1, ni
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
  end subroutine read_velocity_state

  ! per block p with one ghost layer and su over the interior
  subroutine write_pressure_state()
    double precision, allocatable :: pout(:,:,:), sout(:,:,:)
    integer :: b, i, j, k, ni, nj, nk
    do b = 1, nbp
       ni = nc(1, b); nj = nc(2, b); nk = nc(3, b)
       allocate(pout(ni+2, nj+2, nk+2), sout(ni, nj, nk))
       do k = 1, nk+2; do j = 1, nj+2; do i = 1, ni+2
          pout(i, j, k) = dom(b)%p(dom(b)%isp-2+i, dom(b)%jsp-2+j, dom(b)%ksp-2+k)
       end do; end do; end do
       do k = 1, nk; do j = 1, nj; do i = 1, ni
          sout(i, j, k) = dom(b)%su(dom(b)%isp-1+i, dom(b)%jsp-1+j, dom(b)%ksp-1+k)
       end do; end do; end do
       write(uout) pout
       write(uout) sout
       deallocate(pout, sout)
    end do
  end subroutine write_pressure_state

  ! per block the whole u, v, w arrays, ghost layers included
  subroutine write_velocity_arrays()
    integer :: b
    do b = 1, nbp
       write(uout) dom(b)%u
       write(uout) dom(b)%v
       write(uout) dom(b)%w
    end do
  end subroutine write_velocity_arrays

end module mb_setup
