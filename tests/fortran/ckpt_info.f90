! Writes FN.info records in the sequence MPISAVEX uses and an MSAVE-style
! file, with the compiler's own unformatted I/O, then reads a file written
! by the Python side and prints its values.  usage: ckpt_info <dir>
program ckpt_info
  implicit none
  character(len=512) :: d
  integer :: u, ndimr, ndimth, ndimx, nrc, ntc, nxc, sp, minc, mklink, s1, s2, s3
  real(8) :: ln, zlen, ell
  complex(8) :: e(3, 2, 2)
  integer :: i, j, k
  call get_command_argument(1, d)
  open (newunit=u, file=trim(d)//'/f.info', form='unformatted', status='replace')
  write (u) 387, 257, 256
  write (u) 262, 86, 171
  write (u) 1, -0.0795774715459476d0
  write (u) 12.566370614359172d0, 2.0d0
  write (u) 1, 0
  close (u)
  do k = 1, 2
    do j = 1, 2
      do i = 1, 3
        e(i, j, k) = cmplx(i + 10 * j + 100 * k, -i, kind=8)
      end do
    end do
  end do
  open (newunit=u, file=trim(d)//'/f.msave', form='unformatted', status='replace')
  write (u) 5, 4, 3
  write (u) 3, 2, 2
  write (u) 1, 0.5d0
  write (u) 6.0d0, 1.0d0
  write (u) 1, 0
  write (u) 3, 2, 2
  write (u) e
  close (u)
  open (newunit=u, file=trim(d)//'/p.info', form='unformatted', status='old')
  read (u) ndimr, ndimth, ndimx
  read (u) nrc, ntc, nxc
  read (u) sp, ln
  read (u) zlen, ell
  read (u) minc, mklink
  close (u)
  write (*, '(6i6,i3,3es24.16,2i3)') ndimr, ndimth, ndimx, nrc, ntc, nxc, sp, ln, zlen, ell, minc, mklink
end program ckpt_info
