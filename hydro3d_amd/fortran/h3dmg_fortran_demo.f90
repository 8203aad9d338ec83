! h3dmg_fortran_demo.f90 -- a Fortran host driving the device multigrid
! through h3dmg_mod, with arrays laid out as Hydro3D lays out p and su (two
! ghost layers, i fastest).
!
! The right-hand side is one discrete eigenmode of the 7-point operator on a
! 32^3 block (periodic in x and y, Neumann walls in z):
!     f(i,j,k) = cos(2 pi (i+1/2)/n) cos(2 pi (j+1/2)/n) cos(pi (k+1/2)/n)
! whose exact discrete solution is f / lambda with
!     lambda = sum_a 2 (cos(theta_a) - 1) / h^2,
!     theta = (2 pi/n, 2 pi/n, pi/n).
! The demo solves it in FP64 and in mixed precision and checks the pressure
! against that closed form and the residual it computes itself.
!
! It then runs the velocity-form loop of newsolv_mg on the device from face
! arrays in Hydro3D's layout (u(i,j,k) on the face between cells i and i+1,
! two ghost layers): a smooth u*, v*, w* (w* = 0 on the z walls), p = 0,
! until rmax < 1e-10.  On the host it recomputes calvel from the returned p
! and u* and calmas from the returned u, v, w, and checks them against the
! returned u, v, w, su and rmax.  Prints "h3dmg fortran demo: OK" and exits 0
! on success, stops with code 1 otherwise.
program h3dmg_fortran_demo
  use, intrinsic :: iso_c_binding
  use h3dmg_mod
  implicit none
  integer, parameter :: n = 32, pl = 2, ld = n + 2 * pl
  real(c_double), allocatable :: f(:, :, :), p(:, :, :), pex(:, :, :)
  type(h3dmg_block) :: blk(1)
  type(h3dmg_params) :: prm
  type(c_ptr) :: h
  real(c_double) :: hh, lam, pi, hist(0:60), fin, err, res, fn
  integer(c_int) :: st, ncyc, prec
  integer :: i, j, k, run
  logical :: ok

  pi = acos(-1.0d0)
  hh = 1.0d0 / n
  allocate (f(ld, ld, ld), p(ld, ld, ld), pex(ld, ld, ld))
  f = 0.0d0
  lam = 2.0d0 * (cos(2.0d0 * pi / n) - 1.0d0) / hh**2 * 2.0d0 + &
        2.0d0 * (cos(pi / n) - 1.0d0) / hh**2
  do k = 1, n
     do j = 1, n
        do i = 1, n
           f(pl + i, pl + j, pl + k) = cos(2.0d0 * pi * (i - 0.5d0) / n) * &
                cos(2.0d0 * pi * (j - 0.5d0) / n) * cos(pi * (k - 0.5d0) / n)
        end do
     end do
  end do
  pex = f / lam

  blk(1)%ni = n; blk(1)%nj = n; blk(1)%nk = n
  blk(1)%dx = hh; blk(1)%dy = hh; blk(1)%dz = hh
  blk(1)%nlevels = 4
  blk(1)%nbr = -1
  blk(1)%bc = [H3DMG_BC_PERIODIC, H3DMG_BC_PERIODIC, H3DMG_BC_PERIODIC, &
               H3DMG_BC_PERIODIC, H3DMG_BC_WALL, H3DMG_BC_WALL]

  ok = .true.
  ! red-black in FP64 and mixed, then the reference's other two smoothers
  ! (mg_itrsch 1 lexicographic GS, 2 line TDMA) in FP64
  do run = 1, 4
     call h3dmg_default_params(prm)
     prm%tol = 1.0d-10
     prec = merge(H3DMG_MIXED, H3DMG_FP64, run == 2)
     prm%precision = prec
     if (prec == H3DMG_MIXED) prm%coarse = H3DMG_COARSE_EXACT
     if (run == 3) prm%smoother = H3DMG_SMOOTH_GS
     if (run == 4) prm%smoother = H3DMG_SMOOTH_TDMA
     st = h3dmg_create(blk, 1, prm, h)
     call check(st, 'create')
     st = h3dmg_set_rhs(h, 0, f, ld, ld, pl)
     call check(st, 'set_rhs')
     p = 0.0d0
     st = h3dmg_set_p(h, 0, p, ld, ld, pl)
     call check(st, 'set_p')
     st = h3dmg_solve(h, hist, size(hist, kind=c_int), ncyc, fin)
     call check(st, 'solve')
     st = h3dmg_get_p(h, 0, p, ld, ld, pl)
     call check(st, 'get_p')
     call h3dmg_destroy(h)
     call residual(p, f, res, fn)
     ! pressure is defined up to a constant: compare mean-free fields
     p(pl + 1:pl + n, pl + 1:pl + n, pl + 1:pl + n) = &
          p(pl + 1:pl + n, pl + 1:pl + n, pl + 1:pl + n) - &
          sum(p(pl + 1:pl + n, pl + 1:pl + n, pl + 1:pl + n)) / n**3
     err = sqrt(sum((p(pl + 1:pl + n, pl + 1:pl + n, pl + 1:pl + n) - &
                     pex(pl + 1:pl + n, pl + 1:pl + n, pl + 1:pl + n))**2)) / &
           sqrt(sum(pex(pl + 1:pl + n, pl + 1:pl + n, pl + 1:pl + n)**2))
     write (*, '(a,i1,a,i1,a,i3,a,es10.3,a,es10.3,a,es10.3)') ' precision ', prec, &
          ' smoother ', prm%smoother, ': cycles', ncyc, ' final ', fin, ' |p-pex|/|pex| ', err, &
          ' |f-Ap|/|f| ', res / fn
     if (err > 1.0d-8 .or. res / fn > 1.0d-9) ok = .false.
  end do
  call velocity_form(ok)
  if (.not. ok) then
     write (*, '(a)') 'h3dmg fortran demo: FAILED'
     stop 1
  end if
  write (*, '(a)') 'h3dmg fortran demo: OK'

contains

  subroutine velocity_form(ok)
    logical, intent(inout) :: ok
    real(c_double), allocatable :: us(:, :, :), vs(:, :, :), ws(:, :, :)
    real(c_double), allocatable :: u(:, :, :), v(:, :, :), w(:, :, :), pv(:, :, :), su(:, :, :)
    real(c_double), parameter :: dt = 1.0d-3, eps = 1.0d-10
    real(c_double) :: r, x, y, z, eu, es, smax, rhost, fact, umax
    integer(c_int) :: it
    integer :: ii, jj, kk
    allocate (us(ld, ld, ld), vs(ld, ld, ld), ws(ld, ld, ld), u(ld, ld, ld), v(ld, ld, ld), &
              w(ld, ld, ld), pv(ld, ld, ld), su(ld, ld, ld))
    us = 0.0d0; vs = 0.0d0; ws = 0.0d0
    ! faces pl..pl+n on each axis: index pl+f is the face between cells f and
    ! f+1 (cells pl+1..pl+n); periodic x, y; z walls at faces pl and pl+n
    do k = 1, n
       do j = 1, n
          do i = 0, n
             x = i * hh; y = (j - 0.5d0) * hh; z = (k - 0.5d0) * hh
             us(pl + i, pl + j, pl + k) = sin(2.0d0 * pi * x) * cos(2.0d0 * pi * y) + z
             vs(pl + j, pl + i, pl + k) = cos(2.0d0 * pi * x) * z * z
          end do
       end do
    end do
    do k = 1, n - 1
       do j = 1, n
          do i = 1, n
             ws(pl + i, pl + j, pl + k) = sin(pi * k * hh) * cos(2.0d0 * pi * (i - 0.5d0) * hh)
          end do
       end do
    end do
    u = us; v = vs; w = ws
    call h3dmg_default_params(prm)
    prm%dt = dt
    prm%tol = eps
    prm%maxcy = 60
    st = h3dmg_create(blk, 1, prm, h)
    call check(st, 'create')
    pv = 0.0d0
    st = h3dmg_set_p(h, 0, pv, ld, ld, pl)
    call check(st, 'set_p')
    st = h3dmg_set_velocity(h, 0, us, vs, ws, u, v, w, ld, ld, pl)
    call check(st, 'set_velocity')
    st = h3dmg_velocity_begin(h, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                              c_null_ptr, 1.0d0, r)
    call check(st, 'velocity_begin')
    it = 0
    do while (it < prm%maxcy)
       st = h3dmg_velocity_step(h, r)
       call check(st, 'velocity_step')
       it = it + 1
       if (r < eps .and. it > 1) exit
    end do
    st = h3dmg_get_p(h, 0, pv, ld, ld, pl)
    call check(st, 'get_p')
    st = h3dmg_get_velocity(h, 0, u, v, w, ld, ld, pl)
    call check(st, 'get_velocity')
    st = h3dmg_get_su(h, 0, su, ld, ld, pl)
    call check(st, 'get_su')
    call h3dmg_destroy(h)
    ! calvel on the host: corrected faces of x and y (all, periodic) and the
    ! interior z faces; the z-wall faces keep w*
    eu = 0.0d0
    umax = 0.0d0
    do kk = pl + 1, pl + n
       do jj = pl + 1, pl + n
          do ii = pl, pl + n
             if (ii > pl) eu = max(eu, abs(u(ii, jj, kk) - (us(ii, jj, kk) - dt * &
                  (pv(ii + 1, jj, kk) - pv(ii, jj, kk)) / hh)))
             umax = max(umax, abs(u(ii, jj, kk)))
          end do
       end do
    end do
    do kk = pl, pl + n
       do jj = pl + 1, pl + n
          do ii = pl + 1, pl + n
             if (kk == pl .or. kk == pl + n) then
                eu = max(eu, abs(w(ii, jj, kk) - ws(ii, jj, kk)))
             else
                eu = max(eu, abs(w(ii, jj, kk) - (ws(ii, jj, kk) - dt * &
                     (pv(ii, jj, kk + 1) - pv(ii, jj, kk)) / hh)))
             end if
          end do
       end do
    end do
    ! calmas on the host
    es = 0.0d0
    smax = 0.0d0
    rhost = 0.0d0
    do kk = pl + 1, pl + n
       do jj = pl + 1, pl + n
          do ii = pl + 1, pl + n
             fact = (u(ii, jj, kk) - u(ii - 1, jj, kk)) / hh + (v(ii, jj, kk) - v(ii, jj - 1, kk)) / hh + &
                    (w(ii, jj, kk) - w(ii, jj, kk - 1)) / hh
             es = max(es, abs(su(ii, jj, kk) - fact / dt))
             smax = max(smax, abs(fact / dt))
             rhost = max(rhost, abs(fact * hh**3))
          end do
       end do
    end do
    write (*, '(a,i3,a,es10.3,a,es10.3,a,es10.3,a,es10.3)') ' velocity form: cycles', it, &
         ' rmax ', r, ' host rmax ', rhost, ' |u-calvel| ', eu, ' |su-calmas| ', es
    if (r >= eps .or. abs(rhost - r) > 1.0d-6 * r + 1.0d-20 .or. eu > 1.0d-12 * umax .or. &
        es > 1.0d-9 * max(smax, 1.0d0)) ok = .false.
  end subroutine velocity_form

  subroutine check(st, what)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: what
    if (st /= H3DMG_OK) then
       write (*, '(a,a,a,i2,a,a)') 'h3dmg_', what, ' failed: status ', st, ' ', &
            h3dmg_error_message()
       stop 1
    end if
  end subroutine check

  ! ||f - A p||_2 and ||f||_2 on the host, ghosts from get_p (images)
  subroutine residual(p, f, res, fn)
    real(c_double), intent(in) :: p(ld, ld, ld), f(ld, ld, ld)
    real(c_double), intent(out) :: res, fn
    real(c_double) :: c, a, aw, ae, ab, at, r
    integer :: i, j, k, ii, jj, kk
    c = 1.0d0 / hh**2
    res = 0.0d0
    fn = 0.0d0
    do k = 1, n
       do j = 1, n
          do i = 1, n
             ii = pl + i; jj = pl + j; kk = pl + k
             aw = c; ae = c; ab = c; at = c
             if (k == 1) ab = 0.0d0
             if (k == n) at = 0.0d0
             a = -(aw + ae + c + c + ab + at)
             r = f(ii, jj, kk) - (aw * p(ii - 1, jj, kk) + ae * p(ii + 1, jj, kk) + &
                  c * p(ii, jj - 1, kk) + c * p(ii, jj + 1, kk) + ab * p(ii, jj, kk - 1) + &
                  at * p(ii, jj, kk + 1) + a * p(ii, jj, kk))
             res = res + r * r
             fn = fn + f(ii, jj, kk)**2
          end do
       end do
    end do
    res = sqrt(res)
    fn = sqrt(fn)
  end subroutine residual

end program h3dmg_fortran_demo
