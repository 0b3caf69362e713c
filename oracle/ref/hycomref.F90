! hycomref: driver that runs the reference HYCOM ocean continuity equation
! (hycom_cnuity.F90, subroutine cnuity), compiled from the reference sources
! by oracle/build_ref.sh, on a grid file (glvl.dat) and a state file.  Test
! infrastructure only.
!
! Usage: hycomref <glvl.dat> <state.bin> <out.bin>
!
! State file (stream, little endian):
!   int32  magic(=1213809491), kdm, nip, nsteps, nstep0, 0, 0, 0
!   real64 dtleap, thkdff, biharm, totmass, massglb0, masscor
!          (the run's first step is nstep0, and hycom_variables' mass state
!          starts from the three values given)
!   int32  wet(nip)
!   real64 dp(kdm,nip,2), pres(kdm+1,nip)
!   real64 um_edg, vm_edg, un_edg, vn_edg, dpm_edg   (kdm,npp,nip) each
!   real64 massflx, mssfx_avg                        (kdm,npp,nip) each
! Each step calls cnuity(nstep, leapm, leapn, ...), nstep = nstep0 .. nstep0 +
! nsteps - 1, with leapm = 1 + mod(nstep,2)
! and leapn the other slot (the new thickness replaces the old one in leapn),
! the edge velocities and thicknesses held fixed.
!
! Output: records  character(16) tag, int64 n, real64 vals(n): after every
! step 'dp', 'pres', 'massflx', 'mssfx_avg' and 'mass' (totmass, massglb0,
! masscor).

program hycomref
  use module_control,   only: nip, npp
  use fimnamelist,      only: kdm, thkdff, biharm, itest, diag_intvl
  use module_constants, only: prox, proxs, nprox, area, rarea, sidevec_e, sideln, rsideln, rprox_ln, &
                              nedge, permedge, perm, inv_perm
  use hycom_constants,  only: wet, dtleap
  use hycom_variables,  only: totmass, massglb0, masscor
  use hycom_cnuity,     only: cnuity
  implicit none

  character(len=512) :: fgrid, fstate, fout
  integer(4) :: hdr(8)
  real(8)    :: r6(6)
  integer    :: nsteps, nstep0, nstep, leapm, leapn, ipn, isn, idx
  real(4), allocatable :: work4(:)
  integer, allocatable :: iwork(:)
  integer(4), allocatable :: wet4(:)
  character(16) :: header
  real(8), allocatable :: r8c(:,:), r8e(:,:,:), r8d(:,:,:)
  real, allocatable :: dp(:,:,:), pres(:,:), um_edg(:,:,:), vm_edg(:,:,:), un_edg(:,:,:), vn_edg(:,:,:)
  real, allocatable :: dpm_edg(:,:,:), massflx(:,:,:), mssfx_avg(:,:,:)

  call get_command_argument(1, fgrid)
  call get_command_argument(2, fstate)
  call get_command_argument(3, fout)
  open (21, file=trim(fstate), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1213809491) stop 'hycomref: bad state magic'
  kdm = hdr(2); nip = hdr(3); nsteps = hdr(4)     ! npp = 6 (module_control)
  nstep0 = max(1, hdr(5))
  read (21) r6
  dtleap = r6(1); thkdff = r6(2); biharm = r6(3)
  itest = -1
  diag_intvl = 1000000000           ! no diagnostics within the run

  ! ---- grid: the glvl.dat records cnuity uses (getlvl.F90:296-346) ----
  allocate (nprox(nip), proxs(npp,nip), prox(npp,nip), area(nip), rarea(nip))
  allocate (sidevec_e(2,npp,nip), sideln(npp,nip), rsideln(npp,nip), rprox_ln(npp,nip))
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
      read (22) work4               ! cs
    end do
  end do
  do isn = 1, npp
    do idx = 1, 4
      read (22) work4               ! sn
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
  allocate (dp(kdm,nip,2), pres(kdm+1,nip), r8d(kdm,nip,2), r8c(kdm+1,nip))
  allocate (um_edg(kdm,npp,nip), vm_edg(kdm,npp,nip), un_edg(kdm,npp,nip), vn_edg(kdm,npp,nip))
  allocate (dpm_edg(kdm,npp,nip), massflx(kdm,npp,nip), mssfx_avg(kdm,npp,nip), r8e(kdm,npp,nip))
  read (21) r8d; dp = r8d
  read (21) r8c; pres = r8c
  read (21) r8e; um_edg = r8e
  read (21) r8e; vm_edg = r8e
  read (21) r8e; un_edg = r8e
  read (21) r8e; vn_edg = r8e
  read (21) r8e; dpm_edg = r8e
  read (21) r8e; massflx = r8e
  read (21) r8e; mssfx_avg = r8e
  close (21)
  totmass = r6(4); massglb0 = r6(5); masscor = r6(6)

  open (23, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  do nstep = nstep0, nstep0 + nsteps - 1
    leapm = 1 + mod(nstep, 2)
    leapn = 3 - leapm
    call cnuity (nstep, leapm, leapn, um_edg, vm_edg, un_edg, vn_edg, dpm_edg, dp, pres, massflx, mssfx_avg)
    call out('dp', reshape(real(dp,8), (/ size(dp) /)))
    call out('pres', reshape(real(pres,8), (/ size(pres) /)))
    call out('massflx', reshape(real(massflx,8), (/ size(massflx) /)))
    call out('mssfx_avg', reshape(real(mssfx_avg,8), (/ size(mssfx_avg) /)))
    call out('mass', (/ real(totmass,8), real(massglb0,8), real(masscor,8) /))
  end do
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
