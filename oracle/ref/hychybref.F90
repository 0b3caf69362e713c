! hychybref: driver that runs the reference HYCOM hybrid grid generator
! (hycom_hybgn1.F90 hybgn1 with hybgn1_1d, ppmad3, pcwise, equilb and
! hycom_sigetc.F90's equation of state), compiled from the reference
! sources by oracle/build_ref.sh.  Test infrastructure only.
!
! Usage: hychybref <state.bin> <out.bin>
!
! State file (stream, little endian):
!   int32  magic(=1213809480), kdm, nip, leap, numtr, with_tracers, 0, 0
!   real64 batrop
!   int32  wet(nip)
!   real64 targt(kdm), salmin(kdm), dpmin(kdm), latij(nip), dens(kdm,nip),
!          dp(kdm,nip,2), pres(kdm+1,nip), temp(kdm,nip), saln(kdm,nip),
!          uvel(kdm,nip,2), vvel(kdm,nip,2), passv_tr(kdm,nip,numtr)
! Output: records  character(16) tag, int64 n, real64 vals(n): 'dens', 'dp',
! 'pres', 'temp', 'saln', 'uvel', 'vvel', 'passv'.

program hychybref
  use module_control,   only: nip
  use fimnamelist,      only: kdm, itest, diag_intvl
  use module_constants, only: perm
  use hycom_constants,  only: wet, salmin, dpmin, batrop
  use hycom_control,    only: numtr
  use hycom_hybgn1,     only: hybgn1
  implicit none

  character(len=512) :: fstate, fout
  integer(4) :: hdr(8)
  real(8)    :: r1(1)
  integer    :: leap, withtr, i
  integer(4), allocatable :: wet4(:)
  real(8), allocatable :: r8k(:), r8n(:), r8c(:,:), r8d(:,:,:), r8p(:,:), r8t(:,:,:)
  real, allocatable :: targt(:), latij(:), dens(:,:), dp(:,:,:), pres(:,:), temp(:,:), saln(:,:)
  real, allocatable :: uvel(:,:,:), vvel(:,:,:), passv(:,:,:)

  call get_command_argument(1, fstate)
  call get_command_argument(2, fout)
  open (21, file=trim(fstate), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1213809480) stop 'hychybref: bad state magic'
  kdm = hdr(2); nip = hdr(3); leap = hdr(4); numtr = hdr(5); withtr = hdr(6)
  read (21) r1
  batrop = r1(1)
  itest = -1
  diag_intvl = 1000000000           ! no diagnostics within the run

  allocate (perm(nip), wet(nip), wet4(nip), salmin(kdm), dpmin(kdm), targt(kdm), latij(nip))
  do i = 1, nip
    perm(i) = i
  end do
  read (21) wet4
  wet = wet4
  allocate (r8k(kdm), r8n(nip), r8c(kdm,nip), r8d(kdm,nip,2), r8p(kdm+1,nip), r8t(kdm,nip,max(numtr,1)))
  read (21) r8k; targt = r8k
  read (21) r8k; salmin = r8k
  read (21) r8k; dpmin = r8k
  read (21) r8n; latij = r8n
  allocate (dens(kdm,nip), dp(kdm,nip,2), pres(kdm+1,nip), temp(kdm,nip), saln(kdm,nip))
  allocate (uvel(kdm,nip,2), vvel(kdm,nip,2), passv(kdm,nip,max(numtr,1)))
  read (21) r8c; dens = r8c
  read (21) r8d; dp = r8d
  read (21) r8p; pres = r8p
  read (21) r8c; temp = r8c
  read (21) r8c; saln = r8c
  read (21) r8d; uvel = r8d
  read (21) r8d; vvel = r8d
  passv = 0.
  if (numtr > 0) then
    read (21) r8t(:,:,1:numtr); passv(:,:,1:numtr) = r8t(:,:,1:numtr)
  end if
  close (21)

  if (withtr /= 0) then
    call hybgn1(1, leap, targt, dens, dp, pres, temp, saln, uvel, vvel, latij, passv)
  else
    call hybgn1(1, leap, targt, dens, dp, pres, temp, saln, uvel, vvel, latij)
  end if

  open (23, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  call out('dens', reshape(real(dens,8), (/ size(dens) /)))
  call out('dp', reshape(real(dp,8), (/ size(dp) /)))
  call out('pres', reshape(real(pres,8), (/ size(pres) /)))
  call out('temp', reshape(real(temp,8), (/ size(temp) /)))
  call out('saln', reshape(real(saln,8), (/ size(saln) /)))
  call out('uvel', reshape(real(uvel,8), (/ size(uvel) /)))
  call out('vvel', reshape(real(vvel,8), (/ size(vvel) /)))
  if (numtr > 0) call out('passv', reshape(real(passv(:,:,1:numtr),8), (/ kdm*nip*numtr /)))
  close (23)

contains
  subroutine out(tag, v)
    character(*), intent(in) :: tag
    real(8), intent(in) :: v(:)
    character(16) :: t16
    t16 = tag
    write (23) t16, int(size(v), 8), v
  end subroutine
end program
