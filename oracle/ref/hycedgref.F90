! hycedgref: driver that runs the reference HYCOM edge interpolation
! (hycom_edgvar.F90, subroutine bclin_edgvar) and, on request, the continuity
! equation after it (hycom_cnuity.F90 cnuity with dp_edg as its dpm_edg, as
! hycom_run.F90:364-377 calls them), compiled from the reference sources by
! oracle/build_ref.sh.  Test infrastructure only.
!
! Usage: hycedgref <glvl.dat> <state.bin> <out.bin>
!
! State file (stream, little endian):
!   int32  magic(=1213809477), kdm, nip, nstep, leapm, janjic_ocn, chain, 0
!   real64 dtleap, thkdff, biharm, totmass, massglb0, masscor
!   int32  wet(nip)
!   real64 utrop(3,nip), vtrop(3,nip), uclin(kdm,nip,2), vclin(kdm,nip,2),
!          dp(kdm,nip,2), pres(kdm+1,nip), geop(kdm+1,nip), montg(kdm,nip),
!          spcifv(kdm,nip), massflx(kdm,npp,nip), mssfx_avg(kdm,npp,nip)
! The edge arrays bclin_edgvar updates start at 7.0 everywhere (it leaves
! some of them untouched: geop_edg without janjic_ocn, a pentagon's sixth
! slot).  leapn = 3 - leapm.
!
! Output: records  character(16) tag, int64 n, real64 vals(n): 'pres',
! 'um_edg', 'vm_edg', 'un_edg', 'vn_edg', 'dp_edg', 'pr_edg', 'geop_edg',
! 'mont_edg', 'uvsq_edg', 'spcv_edg'; with chain also cnuity's 'dp', 'pres2',
! 'massflx', 'mssfx_avg' and 'mass'.

program hycedgref
  use module_control,   only: nip, npp
  use fimnamelist,      only: kdm, thkdff, biharm, itest, diag_intvl, janjic_ocn
  use module_constants, only: prox, proxs, nprox, area, rarea, sidevec_e, sideln, rsideln, rprox_ln, &
                              nedge, permedge, perm, inv_perm, cs, sn
  use hycom_constants,  only: wet, dtleap
  use hycom_variables,  only: totmass, massglb0, masscor
  use hycom_edgvar,     only: bclin_edgvar
  use hycom_cnuity,     only: cnuity
  implicit none

  character(len=512) :: fgrid, fstate, fout
  integer(4) :: hdr(8)
  real(8)    :: r6(6)
  integer    :: nstep, leapm, leapn, chain, ipn, isn, idx
  real(4), allocatable :: work4(:)
  integer, allocatable :: iwork(:)
  integer(4), allocatable :: wet4(:)
  character(16) :: header
  real(8), allocatable :: r8t(:,:), r8c(:,:), r8k(:,:), r8d(:,:,:), r8e(:,:,:)
  real, allocatable :: utrop(:,:), vtrop(:,:), uclin(:,:,:), vclin(:,:,:), dp(:,:,:), pres(:,:), geop(:,:)
  real, allocatable :: montg(:,:), spcifv(:,:), massflx(:,:,:), mssfx_avg(:,:,:)
  real, allocatable :: um_edg(:,:,:), vm_edg(:,:,:), un_edg(:,:,:), vn_edg(:,:,:), dp_edg(:,:,:)
  real, allocatable :: pr_edg(:,:,:), geop_edg(:,:,:), mont_edg(:,:,:), uvsq_edg(:,:,:), spcv_edg(:,:,:)

  call get_command_argument(1, fgrid)
  call get_command_argument(2, fstate)
  call get_command_argument(3, fout)
  open (21, file=trim(fstate), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1213809477) stop 'hycedgref: bad state magic'
  kdm = hdr(2); nip = hdr(3); nstep = hdr(4); leapm = hdr(5)
  janjic_ocn = hdr(6) /= 0
  chain = hdr(7)
  leapn = 3 - leapm
  read (21) r6
  dtleap = r6(1); thkdff = r6(2); biharm = r6(3)
  itest = -1
  diag_intvl = 1000000000           ! no diagnostics within the run

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
  allocate (utrop(3,nip), vtrop(3,nip), uclin(kdm,nip,2), vclin(kdm,nip,2), dp(kdm,nip,2))
  allocate (pres(kdm+1,nip), geop(kdm+1,nip), montg(kdm,nip), spcifv(kdm,nip))
  allocate (massflx(kdm,npp,nip), mssfx_avg(kdm,npp,nip))
  allocate (r8t(3,nip), r8d(kdm,nip,2), r8c(kdm+1,nip), r8k(kdm,nip), r8e(kdm,npp,nip))
  read (21) r8t; utrop = r8t
  read (21) r8t; vtrop = r8t
  read (21) r8d; uclin = r8d
  read (21) r8d; vclin = r8d
  read (21) r8d; dp = r8d
  read (21) r8c; pres = r8c
  read (21) r8c; geop = r8c
  read (21) r8k; montg = r8k
  read (21) r8k; spcifv = r8k
  read (21) r8e; massflx = r8e
  read (21) r8e; mssfx_avg = r8e
  close (21)
  totmass = r6(4); massglb0 = r6(5); masscor = r6(6)

  allocate (um_edg(kdm,npp,nip), vm_edg(kdm,npp,nip), un_edg(kdm,npp,nip), vn_edg(kdm,npp,nip))
  allocate (dp_edg(kdm,npp,nip), pr_edg(kdm+1,npp,nip), geop_edg(kdm,npp,nip), mont_edg(kdm,npp,nip))
  allocate (uvsq_edg(kdm,npp,nip), spcv_edg(kdm,npp,nip))
  um_edg = 7.; vm_edg = 7.; un_edg = 7.; vn_edg = 7.; dp_edg = 7.; pr_edg = 7.
  geop_edg = 7.; mont_edg = 7.; uvsq_edg = 7.; spcv_edg = 7.

  call bclin_edgvar(nstep, leapm, leapn, uclin, vclin, utrop, vtrop, dp, pres, geop, montg, spcifv, &
                    um_edg, vm_edg, un_edg, vn_edg, dp_edg, pr_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg)

  open (23, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  call out('pres', reshape(real(pres,8), (/ size(pres) /)))
  call out('um_edg', reshape(real(um_edg,8), (/ size(um_edg) /)))
  call out('vm_edg', reshape(real(vm_edg,8), (/ size(vm_edg) /)))
  call out('un_edg', reshape(real(un_edg,8), (/ size(un_edg) /)))
  call out('vn_edg', reshape(real(vn_edg,8), (/ size(vn_edg) /)))
  call out('dp_edg', reshape(real(dp_edg,8), (/ size(dp_edg) /)))
  call out('pr_edg', reshape(real(pr_edg,8), (/ size(pr_edg) /)))
  call out('geop_edg', reshape(real(geop_edg,8), (/ size(geop_edg) /)))
  call out('mont_edg', reshape(real(mont_edg,8), (/ size(mont_edg) /)))
  call out('uvsq_edg', reshape(real(uvsq_edg,8), (/ size(uvsq_edg) /)))
  call out('spcv_edg', reshape(real(spcv_edg,8), (/ size(spcv_edg) /)))
  if (chain /= 0) then
    call cnuity(nstep, leapm, leapn, um_edg, vm_edg, un_edg, vn_edg, dp_edg, dp, pres, massflx, mssfx_avg)
    call out('dp', reshape(real(dp,8), (/ size(dp) /)))
    call out('pres2', reshape(real(pres,8), (/ size(pres) /)))
    call out('massflx', reshape(real(massflx,8), (/ size(massflx) /)))
    call out('mssfx_avg', reshape(real(mssfx_avg,8), (/ size(mssfx_avg) /)))
    call out('mass', (/ real(totmass,8), real(massglb0,8), real(masscor,8) /))
  end if
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
