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
! against that closed form and the residual it computes itself.  Prints
! "h3dmg fortran demo: OK" and exits 0 on success, stops with code 1 otherwise.
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
  integer :: i, j, k
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
  blk(1)%bc = [H3DMG_BC_PERIODIC, H3DMG_BC_PERIODIC, H3DMG_BC_PERIODIC, &
               H3DMG_BC_PERIODIC, H3DMG_BC_WALL, H3DMG_BC_WALL]

  ok = .true.
  do prec = H3DMG_FP64, H3DMG_MIXED
     call h3dmg_default_params(prm)
     prm%tol = 1.0d-10
     prm%precision = prec
     if (prec == H3DMG_MIXED) prm%coarse = H3DMG_COARSE_EXACT
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
     write (*, '(a,i1,a,i3,a,es10.3,a,es10.3,a,es10.3)') ' precision ', prec, ': cycles', &
          ncyc, ' final ', fin, ' |p-pex|/|pex| ', err, ' |f-Ap|/|f| ', res / fn
     if (err > 1.0d-8 .or. res / fn > 1.0d-9) ok = .false.
  end do
  if (.not. ok) then
     write (*, '(a)') 'h3dmg fortran demo: FAILED'
     stop 1
  end if
  write (*, '(a)') 'h3dmg fortran demo: OK'

contains

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
