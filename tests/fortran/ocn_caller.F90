! ocn_caller: drives the library's HYCOM ocean from Fortran through
! fimhip_bind the way hycom_run would call the replaced routines (test driver
! for the Fortran binding; built by the package build with amdflang).
!
!   ocn_caller <case.bin> <out.bin>   transp0 at the start; two steps of
!        bclin_edgvar, cnuity and momtum (hycom_run.F90:364-385), transp1
!        after the second; then transp2 with dpfinl = dp(:,:,leapn), hybgn1
!        (no tracers) and transp0 (hycom_run.F90:386-556); every routine
!        through its _r8 wrapper with host arrays
!
! case.bin (stream, little endian): int32 magic 1329808716, nip, kdm, numtr,
! 0, 0; real64 dtleap, thkdff, biharm, batrop, veldff, smagcf, temdff; int32
! nprox(nip), prox(6,nip), proxs(6,nip) (1-based), wet(nip); real64
! cs(4,6,nip), sn(4,6,nip), sidevec_c(2,6,nip), sidevec_e(2,6,nip),
! sideln(6,nip), area(nip), lat(nip), rprox_ln(6,nip), uclin(kdm,nip,2),
! vclin(kdm,nip,2), utrop(3,nip), vtrop(3,nip), dp(kdm,nip,2),
! pres(kdm+1,nip), geop(kdm+1,nip), montg(kdm,nip), spcifv(kdm,nip),
! um_edg, vm_edg, un_edg, vn_edg, dp_edg (kdm,6,nip), pr_edg(kdm+1,6,nip),
! geop_edg, mont_edg, uvsq_edg, spcv_edg, massflx, mssfx_avg (kdm,6,nip),
! taux(nip), tauy(nip), ustarb(nip), temp(kdm,nip), saln(kdm,nip),
! dens(kdm,nip), passv(kdm,nip,numtr), targt(kdm), salmin(kdm), dpmin(kdm),
! latij(nip).
! out.bin: real64 dp, pres, massflx, mssfx_avg, uclin, vclin, ustarb, temp,
! saln, dens, cumuflx, dpinit, passv.
program ocn_caller
  use iso_c_binding
  use fimhip_bind
  implicit none
  integer, parameter :: npp = 6
  character(len=512) :: a1, a2
  integer(c_int) :: hdr(6), nip, kdm, numtr, st, leapm, leapn
  real(c_double) :: r7(7)
  type(fimhip_grid_desc), target :: gd
  type(c_ptr) :: ocn
  integer(c_int), allocatable, target :: nprox(:), prox(:,:), proxs(:,:), wet(:)
  real(c_double), allocatable, target :: cs(:,:,:), sn(:,:,:), svc(:,:,:), sve(:,:,:), sideln(:,:), area(:), &
      lat(:), rprox_ln(:,:)
  real(c_double), allocatable :: uclin(:,:,:), vclin(:,:,:), utrop(:,:), vtrop(:,:), dp(:,:,:), pres(:,:), &
      geop(:,:), montg(:,:), spcifv(:,:), um_edg(:,:,:), vm_edg(:,:,:), un_edg(:,:,:), vn_edg(:,:,:), &
      dp_edg(:,:,:), pr_edg(:,:,:), geop_edg(:,:,:), mont_edg(:,:,:), uvsq_edg(:,:,:), spcv_edg(:,:,:), &
      massflx(:,:,:), mssfx_avg(:,:,:), taux(:), tauy(:), ustarb(:), ubforc(:), vbforc(:), dens(:,:), &
      targt(:), salmin(:), dpmin(:), latij(:), cumuflx(:,:,:), dpinit(:,:), dpfinl(:,:)
  real(c_double), allocatable, target :: temp(:,:), saln(:,:), passv(:,:,:)

  call get_command_argument(1, a1)
  call get_command_argument(2, a2)
  open (21, file=trim(a1), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1329808716) stop 'ocn_caller: bad case magic'
  nip = hdr(2); kdm = hdr(3); numtr = hdr(4)
  read (21) r7
  allocate (nprox(nip), prox(npp,nip), proxs(npp,nip), wet(nip), cs(4,npp,nip), sn(4,npp,nip), &
            svc(2,npp,nip), sve(2,npp,nip), sideln(npp,nip), area(nip), lat(nip), rprox_ln(npp,nip))
  read (21) nprox, prox, proxs, wet, cs, sn, svc, sve, sideln, area, lat, rprox_ln
  allocate (uclin(kdm,nip,2), vclin(kdm,nip,2), utrop(3,nip), vtrop(3,nip), dp(kdm,nip,2), pres(kdm+1,nip), &
            geop(kdm+1,nip), montg(kdm,nip), spcifv(kdm,nip), um_edg(kdm,npp,nip), vm_edg(kdm,npp,nip), &
            un_edg(kdm,npp,nip), vn_edg(kdm,npp,nip), dp_edg(kdm,npp,nip), pr_edg(kdm+1,npp,nip), &
            geop_edg(kdm,npp,nip), mont_edg(kdm,npp,nip), uvsq_edg(kdm,npp,nip), spcv_edg(kdm,npp,nip), &
            massflx(kdm,npp,nip), mssfx_avg(kdm,npp,nip), taux(nip), tauy(nip), ustarb(nip), ubforc(nip), &
            vbforc(nip), temp(kdm,nip), saln(kdm,nip), dens(kdm,nip), passv(kdm,nip,numtr), targt(kdm), &
            salmin(kdm), dpmin(kdm), latij(nip), cumuflx(kdm,npp,nip), dpinit(kdm,nip), dpfinl(kdm,nip))
  read (21) uclin, vclin, utrop, vtrop, dp, pres, geop, montg, spcifv, um_edg, vm_edg, un_edg, vn_edg, &
            dp_edg, pr_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg, massflx, mssfx_avg, taux, tauy, ustarb, &
            temp, saln, dens, passv, targt, salmin, dpmin, latij
  close (21)
  cumuflx = 0.d0; dpinit = 0.d0; ubforc = 0.d0; vbforc = 0.d0

  ! the instance (hycom_init): grid, wet, dtleap, thkdff, biharm; batrop,
  ! veldff, smagcf; numtr, temdff; theta, salmin, dpmin, deg_lat
  gd%nip = nip; gd%nvl = kdm; gd%ntra = 0; gd%dt = r7(1)
  gd%nprox = c_loc(nprox); gd%prox = c_loc(prox); gd%proxs = c_loc(proxs)
  gd%cs = c_loc(cs); gd%sn = c_loc(sn); gd%sidevec_c = c_loc(svc); gd%sidevec_e = c_loc(sve)
  gd%sideln = c_loc(sideln); gd%area = c_loc(area); gd%lat = c_loc(lat)
  gd%hfcoef = c_null_ptr; gd%lon = c_null_ptr; gd%thetac = c_null_ptr; gd%dpsig = c_null_ptr
  call fimhip_check(fimhip_ocn_create(gd, rprox_ln, kdm, wet, r7(1), r7(2), r7(3), 0, ocn), 'ocn_create')
  call fimhip_check(fimhip_ocn_set_momtum(ocn, r7(4), r7(5), r7(6)), 'ocn_set_momtum')
  call fimhip_check(fimhip_ocn_set_transport(ocn, numtr, r7(7)), 'ocn_set_transport')
  call fimhip_check(fimhip_ocn_set_hybgen(ocn, targt, salmin, dpmin, latij), 'ocn_set_hybgen')

  ! hycom_run's order
  call fimhip_check(fimhip_ocn_transp0_r8(ocn, 0, 1, cumuflx, dp, dpinit), 'ocn_transp0')
  do st = 1, 2
    leapm = 1 + mod(st, 2)
    leapn = 3 - leapm
    call fimhip_check(fimhip_ocn_edgvar_r8(ocn, st, leapm, leapn, uclin, vclin, utrop, vtrop, dp, pres, geop, &
                      montg, spcifv, um_edg, vm_edg, un_edg, vn_edg, dp_edg, pr_edg, geop_edg, mont_edg, &
                      uvsq_edg, spcv_edg), 'ocn_edgvar')
    call fimhip_check(fimhip_ocn_cnuity_r8(ocn, st, leapm, leapn, um_edg, vm_edg, un_edg, vn_edg, dp_edg, dp, &
                      pres, massflx, mssfx_avg), 'ocn_cnuity')
    call fimhip_check(fimhip_ocn_momtum_r8(ocn, st, leapm, leapn, utrop, vtrop, uclin, vclin, um_edg, vm_edg, &
                      geop_edg, mont_edg, uvsq_edg, spcv_edg, taux, tauy, ubforc, vbforc, dp, pres, ustarb), &
                      'ocn_momtum')
    if (mod(st, 2) == 0) call fimhip_check(fimhip_ocn_transp1_r8(ocn, st, cumuflx, massflx), 'ocn_transp1')
  end do
  dpfinl(:,:) = dp(:,:,leapn)
  call fimhip_check(fimhip_ocn_transp2_r8(ocn, 2, temp, saln, passv, cumuflx, dpinit, dpfinl), 'ocn_transp2')
  call fimhip_check(fimhip_ocn_hybgn1_r8(ocn, 2, leapn, targt, dens, dp, pres, temp, saln, uclin, vclin, &
                    latij, c_null_ptr), 'ocn_hybgn1')
  call fimhip_check(fimhip_ocn_transp0_r8(ocn, 2, leapn, cumuflx, dp, dpinit), 'ocn_transp0')
  call fimhip_check(fimhip_ocn_destroy(ocn), 'ocn_destroy')

  open (22, file=trim(a2), access='stream', form='unformatted', status='replace', action='write')
  write (22) dp, pres, massflx, mssfx_avg, uclin, vclin, ustarb, temp, saln, dens, cumuflx, dpinit, passv
  close (22)
end program
