! ref_driver.f90 -- test-only driver around the reference Fortran multigrid.
!
! This program is oracle infrastructure: it is linked against the reference
! objects built from src_v70/{module_vars,module_mpi,module_multidata,mgsolver,
! exchangepp,exchangep,press,exchange,exchange_bc,exchange_phi,exchangesca,
! exchangeu,exchangev,exchangew,sipsol}.for by oracle/Makefile and is used only
! to write golden fixtures and to time the reference CPU path.  It fills the
! reference module state for ONE block with no same-level neighbours (the
! single-block layout of src_v70/alloc_dom.for:524-690 when no face matches),
! calls the reference routines, and dumps results as raw little-endian doubles.
!
! Job description: namelist /job/ in the file named by argv(1); raw input
! doubles in argv(2); raw output doubles to argv(3).
!
! Array convention at this boundary: every 3-D array is Fortran column-major
! (i fastest), padded level arrays are (n_i+2, n_j+2, n_k+2) with the ghost
! layer at index 1 and n+2, interior arrays are (n_i, n_j, n_k).
!
! Modes:
!   vcycle  : su -> one mgkcyc (src_v70/mgsolver.for:226-372) from pp = 0 -> pp
!   relax   : phi_g, rhs_g -> mgrelax(g, nsw) (mgsolver.for:374-621) -> phi_g
!   restr   : phi_{g-1}, rhs_{g-1} -> mgrestr(g) (mgsolver.for:623-766)
!             -> phi_g (zeroed), rhs_g
!   corr    : phi_g, phi_{g+1} -> mgcorr(g) (mgsolver.for:768-910) -> phi_g
!   newsolv : u*, v*, w* face arrays, p0 -> the newsolv_mg outer loop
!             (newsolv_mg.for:17-59, replayed statement by statement because
!             newsolv_mg.o pulls boundu/v/w -> wall_function.for, which flang
!             rejects) -> stop cycle, history(rmax, ||su||_2), final p
!             (with ghosts), final su, su after the first calmas (the f of
!             the pressure form).
! cds4 = 1 sets differencing = 2 (calmas' 4th-order divergence,
! press.for:17-37); newsolv then reads faces -1..n+1 of every axis (the
! second ghost faces the 4th-order stencil reaches).
! lsm = 1 switches the reference's variable-density operator on (L_LSM):
! the density (interior n_i*n_j*n_k, ghosts set to images) is read first,
! coeff builds the density-weighted coefficients of every level
! (src_v70/mgsolver.for:72-184), calvel divides by the face density and
! calmas writes sup (press.for:53-55) into the V-cycle's rhs.
program ref_driver
  use vars
  use multidata
  use mpi
  implicit none

  character(len=512) :: fjob, fin, fout
  character(len=16)  :: mode
  integer :: ni, nj, nk, nlev, nu1, nu2, level, nsw, ncycles, bcs(6)
  integer :: scheme, pl_extra, lsm, cds4
  double precision :: h(3), dtin, epsin
  namelist /job/ mode, ni, nj, nk, nlev, nu1, nu2, level, nsw, ncycles, &
       bcs, h, dtin, epsin, scheme, pl_extra, lsm, cds4

  integer :: uin, uout, ib, g
  double precision, allocatable :: a3(:,:,:), b3(:,:,:)

  mode = 'vcycle'; ni = 8; nj = 8; nk = 8; nlev = 1; nu1 = 2; nu2 = 2
  level = 1; nsw = 1; ncycles = 1; bcs = 1; h = 1.0d0; dtin = 1.0d-3
  epsin = 1.0d-6; scheme = 3; pl_extra = 1; lsm = 0; cds4 = 0

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

  call setup_block()

  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', &
       status='old', action='read')
  if (lsm /= 0) call read_density()
  call coeff
  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', &
       status='replace', action='write')

  ib = 1
  select case (trim(mode))
  case ('vcycle')
     call run_vcycle()
  case ('relax')
     call run_relax()
  case ('restr')
     call run_restr()
  case ('corr')
     call run_corr()
  case ('newsolv')
     call run_newsolv()
  case default
     print *, 'ref_driver: unknown mode ', trim(mode)
     call MPI_ABORT(MPI_COMM_WORLD, 2, ierr)
  end select

  close(uin)
  close(uout)
  call MPI_FINALIZE(ierr)

contains

  ! ------------------------------------------------------------------
  ! Module state for one block, following the sizes that
  ! src_v70/localparameters.for:19-26,106-143,187-192,225-239 and
  ! src_v70/initial.for:250-268 derive from the input decks.
  subroutine setup_block()
    integer :: b, gl, mi, mj, mk, tt(3)

    nbp = 1; num_domains = 1; rdivmax = 1; nbpmax = 1
    allocate(rdiv(0:0), dom_id(1), dom_ad(0:0), dom_indid(0:0))
    rdiv(0) = 1; dom_id(1) = 0; dom_ad(0) = 0; dom_indid(0) = 1

    LMR = 1; pl_ex = pl_extra; pl = 1 + pl_ex
    ngrid_input = nlev; ngrd_gl = nlev
    allocate(rdv(0:0, ngrd_gl)); rdv = 1
    maxcy = ncycles; irestr = nu1; iproln = nu2; mg_itrsch = scheme
    solver = 2; L_LSM = lsm /= 0; L_LSMbase = .false.
    differencing = 1; dt = dtin; alfapr = 1.0d0; eps = epsin
    if (cds4 /= 0) differencing = 2
    bc_w = bcs(1); bc_e = bcs(2); bc_s = bcs(3); bc_n = bcs(4)
    bc_b = bcs(5); bc_t = bcs(6)
    PERIODIC = any(bcs == 5)

    allocate(dom(1))
    b = 1
    dom(b)%iprev = -1; dom(b)%inext = -1; dom(b)%jprev = -1
    dom(b)%jnext = -1; dom(b)%kprev = -1; dom(b)%knext = -1
    dom(b)%corprev1 = -1; dom(b)%corprev2 = -1; dom(b)%corprev3 = -1
    dom(b)%corprev4 = -1; dom(b)%cornext1 = -1; dom(b)%cornext2 = -1
    dom(b)%cornext3 = -1; dom(b)%cornext4 = -1
    dom(b)%edgprev1 = -1; dom(b)%edgprev2 = -1; dom(b)%edgprev3 = -1
    dom(b)%edgprev4 = -1; dom(b)%edgprev5 = -1; dom(b)%edgprev6 = -1
    dom(b)%edgnext1 = -1; dom(b)%edgnext2 = -1; dom(b)%edgnext3 = -1
    dom(b)%edgnext4 = -1; dom(b)%edgnext5 = -1; dom(b)%edgnext6 = -1
    dom(b)%per_in = -1; dom(b)%per_ip = -1; dom(b)%per_jn = -1
    dom(b)%per_jp = -1; dom(b)%per_kn = -1; dom(b)%per_kp = -1
    dom(b)%bc_west = bcs(1); dom(b)%bc_east = bcs(2)
    dom(b)%bc_south = bcs(3); dom(b)%bc_north = bcs(4)
    dom(b)%bc_bottom = bcs(5); dom(b)%bc_top = bcs(6)
    dom(b)%ngrid = nlev
    dom(b)%dx = h(1); dom(b)%dy = h(2); dom(b)%dz = h(3)

    dom(b)%isp = 1 + pl; dom(b)%iep = ni + pl
    dom(b)%jsp = 1 + pl; dom(b)%jep = nj + pl
    dom(b)%ksp = 1 + pl; dom(b)%kep = nk + pl
    dom(b)%isu = dom(b)%isp; dom(b)%isv = dom(b)%isp; dom(b)%isw = dom(b)%isp
    dom(b)%ieu = dom(b)%iep; dom(b)%iev = dom(b)%iep; dom(b)%iew = dom(b)%iep
    dom(b)%jsu = dom(b)%jsp; dom(b)%jsv = dom(b)%jsp; dom(b)%jsw = dom(b)%jsp
    dom(b)%jeu = dom(b)%jep; dom(b)%jev = dom(b)%jep; dom(b)%jew = dom(b)%jep
    dom(b)%ksu = dom(b)%ksp; dom(b)%ksv = dom(b)%ksp; dom(b)%ksw = dom(b)%ksp
    dom(b)%keu = dom(b)%kep; dom(b)%kev = dom(b)%kep; dom(b)%kew = dom(b)%kep
    ! wall-normal faces at non-periodic ends are not corrected
    ! (src_v70/localparameters.for:187-192)
    if (dom(b)%bc_east /= 5) dom(b)%ieu = dom(b)%ieu - 1
    if (dom(b)%bc_north /= 5) dom(b)%jev = dom(b)%jev - 1
    if (dom(b)%bc_top /= 5) dom(b)%kew = dom(b)%kew - 1

    tt = (/ ni + 2*pl, nj + 2*pl, nk + 2*pl /)
    dom(b)%ttc_i = tt(1); dom(b)%ttc_j = tt(2); dom(b)%ttc_k = tt(3)
    dom(b)%ttc_ijk = tt(1)*tt(2)*tt(3)

    dom(b)%nwork = 0; dom(b)%nvars = 0
    do g = 1, ngrd_gl
       gl = min(g, dom(b)%ngrid)
       mi = ni/2**(gl-1) + 2; mj = nj/2**(gl-1) + 2; mk = nk/2**(gl-1) + 2
       dom(b)%nwork = dom(b)%nwork + mi*mj*mk + 8*(mi-2)*(mj-2)*(mk-2)
       dom(b)%nvars = dom(b)%nvars + mi*mj*mk
    end do
    allocate(dom(b)%cof(dom(b)%nwork)); dom(b)%cof = 0.0d0
    allocate(dom(b)%faz(ngrd_gl))

    ngg = max(tt(1)*tt(2), tt(1)*tt(3), tt(2)*tt(3))
    nge = max(tt(1), tt(2), tt(3))*(pl+1)**2
    ngc = (pl+1)**3

    allocate(dom(b)%p(tt(1),tt(2),tt(3)), dom(b)%pp(tt(1),tt(2),tt(3)))
    allocate(dom(b)%su(tt(1),tt(2),tt(3)))
    allocate(dom(b)%u(tt(1),tt(2),tt(3)), dom(b)%v(tt(1),tt(2),tt(3)))
    allocate(dom(b)%w(tt(1),tt(2),tt(3)))
    allocate(dom(b)%ustar(tt(1),tt(2),tt(3)), dom(b)%vstar(tt(1),tt(2),tt(3)))
    allocate(dom(b)%wstar(tt(1),tt(2),tt(3)))
    dom(b)%p = 0; dom(b)%pp = 0; dom(b)%su = 0
    dom(b)%u = 0; dom(b)%v = 0; dom(b)%w = 0
    dom(b)%ustar = 0; dom(b)%vstar = 0; dom(b)%wstar = 0
    allocate(dom(b)%sup(tt(1),tt(2),tt(3))); dom(b)%sup = 0
    if (L_LSM) then
       ! as initial.for:196-214 allocates them
       allocate(dom(b)%dens(tt(1),tt(2),tt(3)), dom(b)%ijkp_lsm(0:dom(b)%ngrid))
       dom(b)%ijkp_lsm = 0
       dom(b)%ijkp_lsm(1) = ni*nj*nk
       do g = 2, dom(b)%ngrid
          dom(b)%ijkp_lsm(g) = dom(b)%ijkp_lsm(g-1) + &
               (ni/2**(g-1))*(nj/2**(g-1))*(nk/2**(g-1))
       end do
       dom(b)%tot = dom(b)%ijkp_lsm(dom(b)%ngrid)
       allocate(dom(b)%dens_mg(dom(b)%tot))
    end if

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
  end subroutine setup_block

  ! density of the interior cells; the ghost layers hold images (mirror at
  ! walls, wrap at periodic faces), as the LES's exchange leaves them
  subroutine read_density()
    double precision, allocatable :: d(:,:,:)
    integer :: i, j, k, ii, jj, kk, isp, jsp, ksp
    allocate(d(ni, nj, nk))
    read(uin) d
    isp = dom(1)%isp; jsp = dom(1)%jsp; ksp = dom(1)%ksp
    do k = ksp - pl, ksp + nk - 1 + pl
       do j = jsp - pl, jsp + nj - 1 + pl
          do i = isp - pl, isp + ni - 1 + pl
             ii = img(i - isp + 1, ni, bcs(1) == 5)
             jj = img(j - jsp + 1, nj, bcs(3) == 5)
             kk = img(k - ksp + 1, nk, bcs(5) == 5)
             dom(1)%dens(i, j, k) = d(ii, jj, kk)
          end do
       end do
    end do
  end subroutine read_density

  ! 1-based image of x in [1-pl, n+pl]
  integer function img(x, n, periodic)
    integer, intent(in) :: x, n
    logical, intent(in) :: periodic
    if (x < 1) then
       if (periodic) then
          img = x + n
       else
          img = 1 - x
       end if
    else if (x > n) then
       if (periodic) then
          img = x - n
       else
          img = 2*n + 1 - x
       end if
    else
       img = x
    end if
  end function img

  ! ------------------------------------------------------------------
  ! Level geometry and packing between the reference's j-fastest packed
  ! level storage (src_v70/mgsolver.for:23-41,412-420) and i-fastest arrays.
  subroutine lev_dims(g, mi, mj, mk)
    integer, intent(in) :: g
    integer, intent(out) :: mi, mj, mk
    integer :: gl
    gl = min(g, dom(1)%ngrid)
    mi = ni/2**(gl-1) + 2; mj = nj/2**(gl-1) + 2; mk = nk/2**(gl-1) + 2
  end subroutine lev_dims

  subroutine get_phi(g, arr)
    integer, intent(in) :: g
    double precision, intent(out) :: arr(:,:,:)
    integer :: mi, mj, mk, i, j, k, base
    call lev_dims(g, mi, mj, mk)
    base = dom(1)%faz(g) - mi*mj*mk
    do k = 1, mk; do i = 1, mi; do j = 1, mj
       arr(i, j, k) = dom(1)%cof(base + (k-1)*mi*mj + (i-1)*mj + j)
    end do; end do; end do
  end subroutine get_phi

  subroutine set_phi(g, arr)
    integer, intent(in) :: g
    double precision, intent(in) :: arr(:,:,:)
    integer :: mi, mj, mk, i, j, k, base
    call lev_dims(g, mi, mj, mk)
    base = dom(1)%faz(g) - mi*mj*mk
    do k = 1, mk; do i = 1, mi; do j = 1, mj
       dom(1)%cof(base + (k-1)*mi*mj + (i-1)*mj + j) = arr(i, j, k)
    end do; end do; end do
  end subroutine set_phi

  subroutine get_rhs(g, arr)
    integer, intent(in) :: g
    double precision, intent(out) :: arr(:,:,:)
    integer :: mi, mj, mk, i, j, k, incr
    call lev_dims(g, mi, mj, mk)
    incr = (mi-2)*(mj-2)*(mk-2)
    do k = 2, mk-1; do i = 2, mi-1; do j = 2, mj-1
       arr(i-1, j-1, k-1) = dom(1)%cof(dom(1)%faz(g) + 7*incr + &
            (k-2)*(mi-2)*(mj-2) + (i-2)*(mj-2) + (j-1))
    end do; end do; end do
  end subroutine get_rhs

  subroutine set_rhs(g, arr)
    integer, intent(in) :: g
    double precision, intent(in) :: arr(:,:,:)
    integer :: mi, mj, mk, i, j, k, incr
    call lev_dims(g, mi, mj, mk)
    incr = (mi-2)*(mj-2)*(mk-2)
    do k = 2, mk-1; do i = 2, mi-1; do j = 2, mj-1
       dom(1)%cof(dom(1)%faz(g) + 7*incr + (k-2)*(mi-2)*(mj-2) + &
            (i-2)*(mj-2) + (j-1)) = arr(i-1, j-1, k-1)
    end do; end do; end do
  end subroutine set_rhs

  subroutine read_padded(g, arr)
    integer, intent(in) :: g
    double precision, allocatable, intent(out) :: arr(:,:,:)
    integer :: mi, mj, mk
    call lev_dims(g, mi, mj, mk)
    allocate(arr(mi, mj, mk))
    read(uin) arr
  end subroutine read_padded

  subroutine read_interior(g, arr)
    integer, intent(in) :: g
    double precision, allocatable, intent(out) :: arr(:,:,:)
    integer :: mi, mj, mk
    call lev_dims(g, mi, mj, mk)
    allocate(arr(mi-2, mj-2, mk-2))
    read(uin) arr
  end subroutine read_interior

  ! ------------------------------------------------------------------
  subroutine run_vcycle()
    integer :: i, j, k
    double precision, allocatable :: pad(:,:,:)
    call read_interior(1, a3)
    dom(1)%pp = 0.0d0
    do k = 1, nk; do j = 1, nj; do i = 1, ni
       dom(1)%su(i+pl, j+pl, k+pl) = a3(i, j, k)
       dom(1)%sup(i+pl, j+pl, k+pl) = a3(i, j, k)
    end do; end do; end do
    call mgkcyc
    allocate(pad(ni+2, nj+2, nk+2))
    ! copy-out range of src_v70/mgsolver.for:354-367
    do k = 1, nk+2; do j = 1, nj+2; do i = 1, ni+2
       pad(i, j, k) = dom(1)%pp(i+pl-1, j+pl-1, k+pl-1)
    end do; end do; end do
    write(uout) pad
  end subroutine run_vcycle

  subroutine run_relax()
    call read_padded(level, a3)
    call read_interior(level, b3)
    call set_phi(level, a3)
    call set_rhs(level, b3)
    call mgrelax(level, nsw)
    call get_phi(level, a3)
    write(uout) a3
  end subroutine run_relax

  subroutine run_restr()
    double precision, allocatable :: c3(:,:,:), d3(:,:,:)
    integer :: mi, mj, mk
    call read_padded(level-1, a3)
    call read_interior(level-1, b3)
    call set_phi(level-1, a3)
    call set_rhs(level-1, b3)
    ! poison the coarse level so that what mgrestr does not write shows up
    call lev_dims(level, mi, mj, mk)
    allocate(c3(mi, mj, mk), d3(mi-2, mj-2, mk-2))
    c3 = 7.0d0; d3 = 7.0d0
    call set_phi(level, c3); call set_rhs(level, d3)
    call mgrestr(level)
    call get_phi(level, c3)
    call get_rhs(level, d3)
    write(uout) c3
    write(uout) d3
  end subroutine run_restr

  subroutine run_corr()
    call read_padded(level, a3)
    call read_padded(level+1, b3)
    call set_phi(level, a3)
    call set_phi(level+1, b3)
    call mgcorr(level)
    call get_phi(level, a3)
    write(uout) a3
  end subroutine run_corr

  ! newsolv_mg.for:17-59 with the reference's calvel/calmas/pbound/exchange.
  ! Records, for it = 0..ncycles: rmax (press.for:66-73) and the L2 norm of
  ! su = div(u)/dt after calmas.  Runs all ncycles cycles (no early exit) and
  ! writes the cycle at which `rmax < eps .and. iter > 1` first held.
  subroutine run_newsolv()
    double precision, allocatable :: uf(:,:,:), vf(:,:,:), wf(:,:,:), p0(:,:,:)
    double precision, allocatable :: hist(:,:), pout(:,:,:), sout(:,:,:), su0(:,:,:)
    integer :: i, j, k, it, stop_at, isp, jsp, ksp, fo
    double precision :: ppref

    isp = dom(1)%isp; jsp = dom(1)%jsp; ksp = dom(1)%ksp
    ! cds4: faces -1..n+1 along the axis (the second ghost faces calmas'
    ! 4th-order branch reads, press.for:17-37); else faces 0..n
    fo = 0
    if (cds4 /= 0) fo = 1
    allocate(uf(-fo:ni+fo, nj, nk), vf(ni, -fo:nj+fo, nk), wf(ni, nj, -fo:nk+fo))
    allocate(p0(ni, nj, nk))
    read(uin) uf
    read(uin) vf
    read(uin) wf
    read(uin) p0
    ! face f of cell i (east face) lives at index i+pl; face 0 at isp-1
    do k = 1, nk; do j = 1, nj; do i = -fo, ni + fo
       dom(1)%ustar(isp-1+i, jsp-1+j, ksp-1+k) = uf(i, j, k)
    end do; end do; end do
    do k = 1, nk; do j = -fo, nj + fo; do i = 1, ni
       dom(1)%vstar(isp-1+i, jsp-1+j, ksp-1+k) = vf(i, j, k)
    end do; end do; end do
    do k = -fo, nk + fo; do j = 1, nj; do i = 1, ni
       dom(1)%wstar(isp-1+i, jsp-1+j, ksp-1+k) = wf(i, j, k)
    end do; end do; end do
    dom(1)%u = dom(1)%ustar; dom(1)%v = dom(1)%vstar; dom(1)%w = dom(1)%wstar
    do k = 1, nk; do j = 1, nj; do i = 1, ni
       dom(1)%p(isp-1+i, jsp-1+j, ksp-1+k) = p0(i, j, k)
    end do; end do; end do
    call exchange(4)
    call pbound

    allocate(hist(2, 0:ncycles))
    hist = 0.0d0
    stop_at = -1

    call calvel
    call calmas
    hist(1, 0) = rmax; hist(2, 0) = su_norm()
    allocate(su0(ni, nj, nk))
    do k = 1, nk; do j = 1, nj; do i = 1, ni
       su0(i, j, k) = rhs_field(isp-1+i, jsp-1+j, ksp-1+k)
    end do; end do; end do
    dom(1)%pp = 0.0d0
    do it = 1, ncycles
       iter = it
       call mgkcyc
       ppref = 0.0d0
       do k = dom(1)%ksp-1, dom(1)%kep+1
          do i = dom(1)%isp-1, dom(1)%iep+1
             do j = dom(1)%jsp-1, dom(1)%jep+1
                dom(1)%p(i,j,k) = (dom(1)%p(i,j,k) + (dom(1)%pp(i,j,k) - ppref))
                dom(1)%pp(i,j,k) = 0.0
             end do
          end do
       end do
       call exchange(4)
       call pbound
       call calvel
       call calmas
       hist(1, it) = rmax; hist(2, it) = su_norm()
       if (stop_at < 0 .and. rmax < eps .and. it > 1) stop_at = it
    end do

    allocate(pout(ni+2, nj+2, nk+2), sout(ni, nj, nk))
    do k = 1, nk+2; do j = 1, nj+2; do i = 1, ni+2
       pout(i, j, k) = dom(1)%p(isp-2+i, jsp-2+j, ksp-2+k)
    end do; end do; end do
    do k = 1, nk; do j = 1, nj; do i = 1, ni
       sout(i, j, k) = rhs_field(isp-1+i, jsp-1+j, ksp-1+k)
    end do; end do; end do
    write(uout) dble(stop_at)
    write(uout) hist
    write(uout) pout
    write(uout) sout
    write(uout) su0
  end subroutine run_newsolv

  double precision function su_norm()
    integer :: i, j, k
    double precision :: s
    s = 0.0d0
    do k = dom(1)%ksp, dom(1)%kep
       do j = dom(1)%jsp, dom(1)%jep
          do i = dom(1)%isp, dom(1)%iep
             s = s + rhs_field(i,j,k)**2
          end do
       end do
    end do
    su_norm = sqrt(s)
  end function su_norm

  ! the V-cycle's right-hand side as calmas leaves it: su, or sup under L_LSM
  double precision function rhs_field(i, j, k)
    integer, intent(in) :: i, j, k
    if (L_LSM) then
       rhs_field = dom(1)%sup(i, j, k)
    else
       rhs_field = dom(1)%su(i, j, k)
    end if
  end function rhs_field

end program ref_driver
