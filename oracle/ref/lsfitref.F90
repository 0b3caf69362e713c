! Checker driver (test infrastructure only): the reference least-squares
! pseudo-inverse of getlvl (LSfitting.F90 pseudoInverse, single-precision
! LAPACK sgesvd from prep/lsfit/svd.f) on point sets read from a file.
!
! usage: lsfitref <in> <out>
!   in:  int32 ncase, then per case int32 nobs (5 or 6) and nobs (x, y)
!        float32 pairs
!   out: per case the nobs x 5 pseudo-inverse, float32, inv(ibf, iobs) with
!        ibf fastest (the reference's array order)
program lsfitref
  use LSfitting, only: pseudoInverse
  implicit none
  integer :: ncase, icase, nobs, uin, uout
  real :: xy(2, 6), inv(5, 6)
  character(len=1024) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  read(uin) ncase
  do icase = 1, ncase
    read(uin) nobs
    if (nobs < 5 .or. nobs > 6) stop 2
    read(uin) xy(:, 1:nobs)
    inv = 0.
    call solve_one(nobs)
    write(uout) inv(:, 1:nobs)
  end do
  close(uin)
  close(uout)
contains
  subroutine solve_one(n)
    integer, intent(in) :: n
    real :: a(2, n), p(5, n)
    a = xy(:, 1:n)
    call pseudoInverse(a, n, 5, p)
    inv(:, 1:n) = p
  end subroutine
end program
