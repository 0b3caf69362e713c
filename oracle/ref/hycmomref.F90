! hycmomref: driver that runs the reference HYCOM momentum equation
! (hycom_momtum.F90, subroutine momtum, with hycom_dissip.F90's del4prep and
! dissip), compiled from the reference sources by oracle/build_ref.sh.  Test
! infrastructure only.
!
! Usage: hycmomref <glvl.dat> <state.bin> <out.bin>
!
! State file (stream, little endian):
!   int32  magic(=1213809485), kdm, nip, nstep, leapm, janjic_ocn, 0, 0
!   real64 dtleap, batrop, veldff, smagcf
!   int32  wet(nip)
!   real64 utrop(3,nip), vtrop(3,nip), uclin(kdm,nip,2), vclin(kdm,nip,2),
!          um_edg, vm_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg (kdm,npp,nip),
!          taux(nip), tauy(nip), dp(kdm,nip,2), pres(kdm+1,nip), ustarb(nip)
! leapn = 3 - leapm; corio = omegx2*sin(lat) as dyn_init.F90:753 sets it and
! meshsz = sqrt(area) as hycom_init.F90:327 does.
!
! Output: records  character(16) tag, int64 n, real64 vals(n): 'uclin',
! 'vclin', 'pres', 'ustarb'.

program hycmomref
  use module_control,   only: nip, npp
  use fimnamelist,      only: kdm, itest, diag_intvl, janjic_ocn, veldff, smagcf
  use module_constants, only: prox, proxs, nprox, area, rarea, sidevec_e, sidevec_c, sideln, rsideln, &
                              rprox_ln, nedge, permedge, perm, inv_perm, cs, sn, corio, omegx2
  use hycom_constants,  only: wet, dtleap, batrop, meshsz
  use hycom_momtum,     only: momtum
  implicit none

  character(len=512) :: fgrid, fstate, fout
  integer(4) :: hdr(8)
  real(8)    :: r4v(4)
  integer    :: nstep, leapm, leapn, ipn, isn, idx
  real(4), allocatable :: work4(:)
  integer, allocatable :: iwork(:)
  integer(4), allocatable :: wet4(:)
  character(16) :: header
  real, allocatable :: lat(:)
  real(8), allocatable :: r8t(:,:), r8c(:,:), r8d(:,:,:), r8e(:,:,:), r8n(:)
  real, allocatable :: utrop(:,:), vtrop(:,:), uclin(:,:,:), vclin(:,:,:), dp(:,:,:), pres(:,:)
  real, allocatable :: um_edg(:,:,:), vm_edg(:,:,:), geop_edg(:,:,:), mont_edg(:,:,:), uvsq_edg(:,:,:)
  real, allocatable :: spcv_edg(:,:,:), taux(:), tauy(:), ubforc(:), vbforc(:), ustarb(:)

  call get_command_argument(1, fgrid)
  call get_command_argument(2, fstate)
  call get_command_argument(3, fout)
  open (21, file=trim(fstate), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1213809485) stop 'hycmomref: bad state magic'
  kdm = hdr(2); nip = hdr(3); nstep = hdr(4); leapm = hdr(5)
  janjic_ocn = hdr(6) /= 0
  leapn = 3 - leapm
  read (21) r4v
  dtleap = r4v(1); batrop = r4v(2); veldff = r4v(3); smagcf = r4v(4)
  itest = -1
  diag_intvl = 1000000000           ! no diagnostics within the run

  ! ---- grid: the glvl.dat records these routines use (getlvl.F90:296-346) ----
  allocate (nprox(nip), proxs(npp,nip), prox(npp,nip), area(nip), rarea(nip))
  allocate (sidevec_e(2,npp,nip), sideln(npp,nip), rsideln(npp,nip), rprox_ln(npp,nip))
  allocate (cs(4,npp,nip), sn(4,npp,nip), sidevec_c(2,npp,nip), corio(nip), meshsz(nip), lat(nip))
  allocate (perm(nip), inv_perm(nip), nedge(nip), permedge(npp,nip), work4(nip), iwork(nip))
  open (22, file=trim(fgrid), form='unformatted', status='old', action='read')
  read (22) header
  read (22) header
  read (22) work4; lat = work4
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
      read (22) work4; sidevec_c(idx,isn,:) = work4
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
    corio(ipn) = omegx2*sin(lat(ipn))              ! dyn_init.F90:753
    meshsz(ipn) = sqrt(area(ipn))                  ! hycom_init.F90:327
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
  allocate (utrop(3,nip), vtrop(3,nip), uclin(kdm,nip,2), vclin(kdm,nip,2), dp(kdm,nip,2), pres(kdm+1,nip))
  allocate (um_edg(kdm,npp,nip), vm_edg(kdm,npp,nip), geop_edg(kdm,npp,nip), mont_edg(kdm,npp,nip))
  allocate (uvsq_edg(kdm,npp,nip), spcv_edg(kdm,npp,nip), taux(nip), tauy(nip), ubforc(nip), vbforc(nip))
  allocate (ustarb(nip))
  allocate (r8t(3,nip), r8d(kdm,nip,2), r8c(kdm+1,nip), r8e(kdm,npp,nip), r8n(nip))
  read (21) r8t; utrop = r8t
  read (21) r8t; vtrop = r8t
  read (21) r8d; uclin = r8d
  read (21) r8d; vclin = r8d
  read (21) r8e; um_edg = r8e
  read (21) r8e; vm_edg = r8e
  read (21) r8e; geop_edg = r8e
  read (21) r8e; mont_edg = r8e
  read (21) r8e; uvsq_edg = r8e
  read (21) r8e; spcv_edg = r8e
  read (21) r8n; taux = r8n
  read (21) r8n; tauy = r8n
  read (21) r8d; dp = r8d
  read (21) r8c; pres = r8c
  read (21) r8n; ustarb = r8n
  close (21)
  ubforc = 0.; vbforc = 0.

  call momtum(nstep, leapm, leapn, utrop, vtrop, uclin, vclin, um_edg, vm_edg, geop_edg, mont_edg, uvsq_edg, &
              spcv_edg, taux, tauy, ubforc, vbforc, dp, pres, ustarb)

  open (23, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  call out('uclin', reshape(real(uclin,8), (/ size(uclin) /)))
  call out('vclin', reshape(real(vclin,8), (/ size(vclin) /)))
  call out('pres', reshape(real(pres,8), (/ size(pres) /)))
  call out('ustarb', reshape(real(ustarb,8), (/ size(ustarb) /)))
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
