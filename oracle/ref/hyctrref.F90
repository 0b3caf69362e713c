! hyctrref: driver that runs the reference HYCOM tracer transport
! (hycom_transp3d.F90 transp0, transp1, transp2 with hycom_fct3d.F90's fct3d
! and hycom_dffusn.F90's dffusn_lyr), compiled from the reference sources by
! oracle/build_ref.sh.  Test infrastructure only.
!
! Usage: hyctrref <glvl.dat> <state.bin> <out.bin>
!
! State file (stream, little endian):
!   int32  magic(=1213809490), kdm, nip, leapn, numtr, 0, 0, 0
!   real64 batrop, temdff
!   int32  wet(nip)
!   real64 dp(kdm,nip,2), massflx(kdm,npp,nip), cumuflx(kdm,npp,nip),
!          dpinit(kdm,nip), dpfinl(kdm,nip), temp(kdm,nip), saln(kdm,nip),
!          passv_tr(kdm,nip,numtr)
! Runs transp0 on copies of cumuflx and dpinit (records 'cumuflx0',
! 'dpinit0'); then transp1 with massflx on cumuflx ('cumuflx1'); then
! transp2 with that cumuflx, dpinit and dpfinl ('temp', 'saln', 'passv').
! Output: records  character(16) tag, int64 n, real64 vals(n).

program hyctrref
  use module_control,   only: nip, npp
  use fimnamelist,      only: kdm, itest, diag_intvl, temdff
  use module_constants, only: prox, proxs, nprox, area, rarea, sidevec_e, sideln, rsideln, rprox_ln, &
                              nedge, permedge, perm, inv_perm, cs, sn
  use hycom_constants,  only: wet, batrop
  use hycom_control,    only: numtr
  use hycom_variables,  only: totmass
  use hycom_transp3d,   only: transp0, transp1, transp2
  implicit none

  character(len=512) :: fgrid, fstate, fout
  integer(4) :: hdr(8)
  real(8)    :: r2(2)
  integer    :: leapn, ipn, isn, idx
  real(4), allocatable :: work4(:)
  integer, allocatable :: iwork(:)
  integer(4), allocatable :: wet4(:)
  character(16) :: header
  real(8), allocatable :: r8c(:,:), r8d(:,:,:), r8e(:,:,:), r8p(:,:,:)
  real, allocatable :: dp(:,:,:), massflx(:,:,:), cumuflx(:,:,:), cumu0(:,:,:), dpinit(:,:), dpini0(:,:)
  real, allocatable :: dpfinl(:,:), temp(:,:), saln(:,:), passv(:,:,:)

  call get_command_argument(1, fgrid)
  call get_command_argument(2, fstate)
  call get_command_argument(3, fout)
  open (21, file=trim(fstate), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1213809490) stop 'hyctrref: bad state magic'
  kdm = hdr(2); nip = hdr(3); leapn = hdr(4); numtr = hdr(5)
  read (21) r2
  batrop = r2(1); temdff = r2(2)
  itest = -1
  diag_intvl = 1000000000           ! no diagnostics within the run
  totmass = 1.

  ! ---- grid: the glvl.dat records these routines use (getlvl.F90:296-346) ----
  allocate (nprox(nip), proxs(npp,nip), prox(npp,nip), area(nip), rarea(nip))
  allocate (sidevec_e(2,npp,nip), sideln(npp,nip), rsideln(npp,nip), rprox_ln(npp,nip))
  allocate (cs(4,npp,nip), sn(4,npp,nip))
  allocate (perm(nip), inv_perm(nip), nedge(nip), permedge(npp,nip), work4(nip), iwork(nip))
  open (22, file=trim(fgrid), form='unformatted', status='old', action='read')
  read (22) header
  read (22) header
  read (22) work4                   ! lat
  read (22) work4                   ! lon
  read (22) nprox
  do isn = 1, npp
    read (22) iwork; proxs(isn,:) = iwork
  end do
  do isn = 1, npp
    read (22) iwork; prox(isn,:) = iwork
  end do
  read (22) work4; area = work4
  do isn = 1, npp
    do idx = 1, 4
      read (22) work4; cs(idx,isn,:) = work4
    end do
  end do
  do isn = 1, npp
    do idx = 1, 4
      read (22) work4; sn(idx,isn,:) = work4
    end do
  end do
  do isn = 1, npp
    do idx = 1, 2
      read (22) work4               ! sidevec_c
    end do
  end do
  do isn = 1, npp
    do idx = 1, 2
      read (22) work4; sidevec_e(idx,isn,:) = work4
    end do
  end do
  do isn = 1, npp
    read (22) work4; sideln(isn,:) = work4
  end do
  do isn = 1, npp
    read (22) work4; rprox_ln(isn,:) = work4
  end do
  read (22) inv_perm
  close (22)
  do ipn = 1, nip
    perm(inv_perm(ipn)) = ipn
  end do
  do ipn = 1, nip
    if (nprox(ipn) == 5) prox(6,ipn) = prox(5,ipn)
    rarea(ipn) = 1./area(ipn)
    nedge(ipn) = nprox(ipn)
    do isn = 1, nprox(ipn)
      permedge(isn,ipn) = isn
      rsideln(isn,ipn) = 1./sideln(isn,ipn)
    end do
  end do

  ! ---- state ----
  allocate (wet(nip), wet4(nip))
  read (21) wet4
  wet = wet4
  allocate (dp(kdm,nip,2), massflx(kdm,npp,nip), cumuflx(kdm,npp,nip), cumu0(kdm,npp,nip))
  allocate (dpinit(kdm,nip), dpini0(kdm,nip), dpfinl(kdm,nip), temp(kdm,nip), saln(kdm,nip))
  allocate (passv(kdm,nip,numtr))
  allocate (r8c(kdm,nip), r8d(kdm,nip,2), r8e(kdm,npp,nip), r8p(kdm,nip,numtr))
  read (21) r8d; dp = r8d
  read (21) r8e; massflx = r8e
  read (21) r8e; cumuflx = r8e
  read (21) r8c; dpinit = r8c
  read (21) r8c; dpfinl = r8c
  read (21) r8c; temp = r8c
  read (21) r8c; saln = r8c
  read (21) r8p; passv = r8p
  close (21)

  open (23, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  cumu0 = cumuflx
  dpini0 = dpinit
  call transp0(0, leapn, cumu0, dp, dpini0)
  call out('cumuflx0', reshape(real(cumu0,8), (/ size(cumu0) /)))
  call out('dpinit0', reshape(real(dpini0,8), (/ size(dpini0) /)))
  call transp1(0, cumuflx, massflx)
  call out('cumuflx1', reshape(real(cumuflx,8), (/ size(cumuflx) /)))
  call transp2(0, temp, saln, passv, cumuflx, dpinit, dpfinl)
  call out('temp', reshape(real(temp,8), (/ size(temp) /)))
  call out('saln', reshape(real(saln,8), (/ size(saln) /)))
  call out('passv', reshape(real(passv,8), (/ size(passv) /)))
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
