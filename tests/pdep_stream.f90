! Fortran stream I/O of one PDEP eigenvector file, written with the
! statements the reference uses (IO_kernel/pdep_io.f90:128-139 to write,
! :261-280 to read): a header of 32 default integers at POS=1 (format
! version 210405, little-endian flag, dimension), the complex(8) data at
! POS=1+SIZEOF(header).  Used by tests/test_pdep_db.py to check that the
! Python writer and reader agree with a Fortran compiler byte for byte.
!
!   pdep_stream w <file> <n>   writes c(k) = (k, -k/2 + 1/3), k = 1..n
!   pdep_stream r <file> <n>   reads and prints n, sum Re c, sum Im c, c(n)
program pdep_stream
  implicit none
  integer, parameter :: dp = kind(1.d0)
  integer :: header(32), n, k, iun, ierr
  integer(8) :: offset
  complex(dp), allocatable :: c(:)
  character(len=256) :: mode, fname, arg
  call get_command_argument(1, mode)
  call get_command_argument(2, fname)
  call get_command_argument(3, arg)
  read(arg, *) n
  allocate(c(n))
  if (mode(1:1) == 'w') then
    do k = 1, n
      c(k) = cmplx(real(k, dp), -real(k, dp) / 2._dp + 1._dp / 3._dp, kind=dp)
    end do
    header = 0
    header(1) = 210405
    header(3) = n
    if (ichar(transfer(1, 'a')) /= 0) header(2) = 1
    open(newunit=iun, file=trim(fname), access='STREAM', form='UNFORMATTED', status='REPLACE')
    offset = 1
    write(iun, pos=offset) header
    offset = offset + sizeof(header)
    write(iun, pos=offset) c(1:n)
    close(iun)
  else
    open(newunit=iun, file=trim(fname), access='STREAM', form='UNFORMATTED', status='OLD', iostat=ierr)
    if (ierr /= 0) stop 2
    offset = 1
    read(iun, pos=offset) header
    if (header(1) /= 210405) stop 3
    if (header(3) /= n) stop 4
    offset = offset + sizeof(header)
    read(iun, pos=offset) c(1:n)
    close(iun)
    write(*, '(I0,3(1X,ES24.16))') header(2), sum(real(c)), sum(aimag(c)), real(c(n))
  end if
end program pdep_stream
