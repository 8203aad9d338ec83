! tecbin_io -- test infrastructure: the Fortran runtime's own unformatted
! sequential I/O of Hydro3D's tecbin restart records, to pin the library's
! C reader and writer (hydro3d_amd/csrc/h3dmg_tecbin.cpp) to it.
!
! The record structure is that of the reference's WRITE statements in
! tecbin (src_v70/post.for:346-369: dimensions, pl, the three wall flags,
! then one record per cell of x, y, z and 20 fields) and of the restart
! READ statements (src_v70/initial.for:719-733).
!
!   tecbin_io write RAW BIN   RAW (stream) -> tecbin file BIN
!   tecbin_io read  BIN RAW   tecbin file BIN -> RAW (stream)
!
! RAW (stream access, no record markers): tt(3), pl, ind(3) as 4-byte
! integers; x(tt(1)), y(tt(2)), z(tt(3)); then the 20 fields, each a
! Fortran (tt(1), tt(2), tt(3)) array of doubles.
program tecbin_io
  implicit none
  character(len=512) :: mode, fin, fout
  integer :: tt(3), pl, ind(3), i, j, k, q
  double precision, allocatable :: x(:), y(:), z(:), fld(:,:,:,:)
  double precision :: dm(23)

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  if (trim(mode) == 'write') then
     open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
     read(10) tt, pl, ind
     allocate(x(tt(1)), y(tt(2)), z(tt(3)), fld(tt(1), tt(2), tt(3), 20))
     read(10) x, y, z
     read(10) fld
     close(10)
     open(unit=88, file=trim(fout), form='unformatted')
     write(88) tt(1), tt(2), tt(3)
     write(88) pl
     write(88) ind(1), ind(2), ind(3)
     do k = 1, tt(3)
        do j = 1, tt(2)
           do i = 1, tt(1)
              write(88) x(i), y(j), z(k), (fld(i, j, k, q), q = 1, 20)
           end do
        end do
     end do
     close(88)
  else if (trim(mode) == 'read') then
     open(unit=700, file=trim(fin), form='unformatted', status='old')
     read(700) tt(1), tt(2), tt(3)
     read(700) pl
     read(700) ind(1), ind(2), ind(3)
     allocate(x(tt(1)), y(tt(2)), z(tt(3)), fld(tt(1), tt(2), tt(3), 20))
     do k = 1, tt(3)
        do j = 1, tt(2)
           do i = 1, tt(1)
              read(700) dm(1), dm(2), dm(3), dm(4), dm(5), dm(6), dm(7), dm(8), dm(9), &
                   dm(10), dm(11), dm(12), dm(13), dm(14), dm(15), dm(16), dm(17), dm(18), &
                   dm(19), dm(20), dm(21), dm(22), dm(23)
              x(i) = dm(1); y(j) = dm(2); z(k) = dm(3)
              do q = 1, 20
                 fld(i, j, k, q) = dm(3 + q)
              end do
           end do
        end do
     end do
     close(700)
     open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')
     write(11) tt, pl, ind
     write(11) x, y, z
     write(11) fld
     close(11)
  else
     print *, 'usage: tecbin_io write|read IN OUT'
     stop 2
  end if
end program tecbin_io
