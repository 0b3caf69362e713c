! fimhip_caller: drives the library from Fortran through fimhip_bind the way
! a serial FIM build's dyn_init and dyn_run would call the replaced routines
! (test driver for the Fortran binding; built by the package build with
! amdflang).
!
!   fimhip_caller <case.bin> <out.bin>   the start-up (dyn_init.F90:880-893,
!        1037: hybgen(its=0), pressure from the Exner function, hystat(0))
!        and nsteps full steps in dyn_run's order (dyn_run.F90:329-423:
!        abstart, edgvar1, edgvar3, cnuity, trcadv, momtum, hybgen, hystat),
!        every routine through its _r8 wrapper with host arrays
!   fimhip_caller --reject               bad sizes and null handles; prints
!        the status codes and exits 0 when every call returned -1
!
! case.bin (stream, little endian): int32 magic 1179208498, nip, nvl, ntra,
! nsteps, init_hybgen; real64 dt; int32 nprox(nip), prox(6,nip),
! proxs(6,nip) (1-based); real64 cs(4,6,nip), sn(4,6,nip), sidevec_c(2,6,nip),
! sidevec_e(2,6,nip), sideln(6,nip), hfcoef(6,6,nip), area(nip), lat(nip),
! lon(nip), thetac(nvl), dpsig(nvl), us(nvl,nip), vs(nvl,nip), dp(nvl,nip),
! pr(nvl+1,nip), ex(nvl+1,nip), ph(nvl+1,nip), tr(nvl,nip,ntra).
! out.bin: real64 us, vs, dp, pr, ex, ph, mp, tr, trdp after the last step.
program fimhip_caller
  use iso_c_binding
  use fimhip_bind
  implicit none
  integer, parameter :: npp = 6, nabl = 3
  real(c_double), parameter :: p1000 = 100000.d0, cp = 1004.6855d0, rd = 287.0586d0
  character(len=512) :: a1, a2
  integer(c_int) :: hdr(6), nip, nvl, ntra, nsteps, nvlp1, its, rc, k, ipn
  integer(c_int) :: nf, of, vof
  real(c_double) :: dt, ab1, ab2, ab3
  type(fimhip_grid_desc), target :: gd
  type(c_ptr) :: ctx
  integer(c_int), allocatable, target :: nprox(:), prox(:,:), proxs(:,:)
  real(c_double), allocatable, target :: cs(:,:,:), sn(:,:,:), svc(:,:,:), sve(:,:,:), sideln(:,:), &
      hfcoef(:,:,:), area(:), lat(:), lon(:), thetac(:), dpsig(:)
  real(c_double), allocatable :: us(:,:), vs(:,:), dp(:,:), pr(:,:), ex(:,:), ph(:,:), mp(:,:), &
      tr(:,:,:), trdp(:,:,:), sdot(:,:), ws(:,:), relvor(:,:), psrf(:), ptdcy(:,:), &
      u_tdcy(:,:,:), v_tdcy(:,:,:), dp_tdcy(:,:,:), dpl_tdcy(:,:,:), trc_tdcy(:,:,:,:), trl_tdcy(:,:,:,:), &
      massfx(:,:,:,:), u_edg(:,:,:), v_edg(:,:,:), dp_edg(:,:,:), lp_edg(:,:,:), geop_edg(:,:,:), &
      mont_edg(:,:,:), uvsq_edg(:,:,:), trc_edg(:,:,:,:)

  call get_command_argument(1, a1)
  if (trim(a1) == '--reject') then
    call reject()
    stop
  end if
  call get_command_argument(2, a2)
  open (21, file=trim(a1), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1179208498) stop 'fimhip_caller: bad case magic'
  nip = hdr(2); nvl = hdr(3); ntra = hdr(4); nsteps = hdr(5); nvlp1 = nvl + 1
  read (21) dt
  allocate (nprox(nip), prox(npp,nip), proxs(npp,nip), cs(4,npp,nip), sn(4,npp,nip), svc(2,npp,nip), &
            sve(2,npp,nip), sideln(npp,nip), hfcoef(6,npp,nip), area(nip), lat(nip), lon(nip), &
            thetac(nvl), dpsig(nvl))
  read (21) nprox, prox, proxs, cs, sn, svc, sve, sideln, hfcoef, area, lat, lon, thetac, dpsig
  allocate (us(nvl,nip), vs(nvl,nip), dp(nvl,nip), pr(nvlp1,nip), ex(nvlp1,nip), ph(nvlp1,nip), &
            mp(nvl,nip), tr(nvl,nip,ntra), trdp(nvl,nip,ntra), sdot(nvlp1,nip), ws(nvl,nip), &
            relvor(nvl,nip), psrf(nip), ptdcy(nip,2), u_tdcy(nvl,nip,nabl), v_tdcy(nvl,nip,nabl), &
            dp_tdcy(nvl,nip,nabl), dpl_tdcy(nvl,nip,nabl), trc_tdcy(nvl,nip,nabl,ntra), &
            trl_tdcy(nvl,nip,nabl,ntra), massfx(nvl,npp,nip,nabl), u_edg(nvl,npp,nip), v_edg(nvl,npp,nip), &
            dp_edg(nvl,npp,nip), lp_edg(nvl,npp,nip), geop_edg(nvl,npp,nip), mont_edg(nvl,npp,nip), &
            uvsq_edg(nvl,npp,nip), trc_edg(nvl,npp,nip,ntra))
  read (21) us, vs, dp, pr, ex, ph, tr
  close (21)

  ! the instance (dyn_init.F90:199, 1198)
  gd%nip = nip; gd%nvl = nvl; gd%ntra = ntra; gd%dt = dt
  gd%nprox = c_loc(nprox); gd%prox = c_loc(prox); gd%proxs = c_loc(proxs)
  gd%cs = c_loc(cs); gd%sn = c_loc(sn); gd%sidevec_c = c_loc(svc); gd%sidevec_e = c_loc(sve)
  gd%sideln = c_loc(sideln); gd%hfcoef = c_loc(hfcoef); gd%area = c_loc(area)
  gd%lat = c_loc(lat); gd%lon = c_loc(lon); gd%thetac = c_loc(thetac); gd%dpsig = c_loc(dpsig)
  call fimhip_check(fimhip_create(gd, 0, ctx), 'create')

  ! start-up (dyn_init.F90:880-893, 1037-1094)
  mp = 0.d0; sdot = 0.d0; ws = 0.d0; relvor = 0.d0; psrf = 0.d0; ptdcy = 0.d0
  u_tdcy = 0.d0; v_tdcy = 0.d0; dp_tdcy = 0.d0; dpl_tdcy = 0.d0; trc_tdcy = 0.d0; trl_tdcy = 0.d0
  massfx = 0.d0; u_edg = 0.d0; v_edg = 0.d0; dp_edg = 0.d0; lp_edg = 0.d0; geop_edg = 0.d0
  mont_edg = 0.d0; uvsq_edg = 0.d0; trc_edg = 0.d0
  do k = 1, ntra
    trdp(:,:,k) = tr(:,:,k) * dp
  end do
  if (hdr(6) /= 0) then
    call fimhip_check(fimhip_hybgen_r8(ctx, 0, thetac, us, vs, tr, sdot, ex, dp, pr, 0), 'hybgen(0)')
    do ipn = 1, nip
      do k = 1, nvlp1
        pr(k,ipn) = p1000 * fimhip_pow(ex(k,ipn) / cp, cp / rd)
      end do
    end do
  end if
  call fimhip_check(fimhip_hystat_r8(ctx, 0, ph, ex, mp, dp, tr, trdp, psrf, ptdcy), 'hystat(0)')
  nf = 0; of = 2; vof = 1

  ! the steps (dyn_run.F90:329-423)
  do its = 1, nsteps
    call fimhip_check(fimhip_abstart(ctx, its, 1, nf, of, vof, ab1, ab2, ab3), 'abstart')
    call fimhip_check(fimhip_edgvar1_r8(ctx, its, us, vs, dp, pr, ex, ph, mp, tr, u_edg, v_edg, dp_edg, &
                      lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg), 'edgvar1')
    call fimhip_check(fimhip_edgvar3_r8(ctx, its, us, vs, dp, pr, ex, ph, mp, tr, u_edg, v_edg, dp_edg, &
                      lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg), 'edgvar3')
    call fimhip_check(fimhip_cnuity_r8(ctx, its, us, vs, u_edg, v_edg, dp_edg, lp_edg, dp, pr, ex, &
                      dp_tdcy, dpl_tdcy, massfx, ws, 0), 'cnuity')
    call fimhip_check(fimhip_trcadv_r8(ctx, its, trc_edg, tr, trdp, trc_tdcy, trl_tdcy, massfx, dp, 0), &
                      'trcadv')
    call fimhip_check(fimhip_momtum_r8(ctx, its, us, vs, ex, relvor, u_edg, v_edg, trc_edg, geop_edg, &
                      mont_edg, uvsq_edg, u_tdcy, v_tdcy, dp), 'momtum')
    call fimhip_check(fimhip_hybgen_r8(ctx, its, thetac, us, vs, tr, sdot, ex, dp, pr, 0), 'hybgen')
    call fimhip_check(fimhip_hystat_r8(ctx, its, ph, ex, mp, dp, tr, trdp, psrf, ptdcy), 'hystat')
  end do
  call fimhip_check(fimhip_destroy(ctx), 'destroy')

  open (22, file=trim(a2), access='stream', form='unformatted', status='replace', action='write')
  write (22) us, vs, dp, pr, ex, ph, mp, tr, trdp
  close (22)

contains

  ! bad arguments are rejected with -1 (FIMHIP_EINVAL) before any GPU call
  subroutine reject()
    integer(c_int) :: r(5)
    real(c_double), target :: x(64)
    integer(c_int), target :: ix(64)
    integer(c_int) :: i1, i2, i3
    real(c_double) :: d1, d2, d3
    x = 0.d0; ix = 0
    gd%nip = 42; gd%nvl = 65; gd%ntra = 4; gd%dt = 240.d0          ! nvl > 64
    gd%nprox = c_loc(ix); gd%prox = c_loc(ix); gd%proxs = c_loc(ix)
    gd%cs = c_loc(x); gd%sn = c_loc(x); gd%sidevec_c = c_loc(x); gd%sidevec_e = c_loc(x)
    gd%sideln = c_loc(x); gd%hfcoef = c_loc(x); gd%area = c_loc(x); gd%lat = c_loc(x); gd%lon = c_loc(x)
    gd%thetac = c_loc(x); gd%dpsig = c_loc(x)
    r(1) = fimhip_create(gd, 0, ctx)
    gd%nvl = 64; gd%nip = 10                                         ! too few cells
    r(2) = fimhip_create(gd, 0, ctx)
    r(3) = fimhip_cnuity_r8(c_null_ptr, 1, x, x, x, x, x, x, x, x, x, x, x, x, x, 0)
    r(4) = fimhip_hystat_r8(c_null_ptr, 1, x, x, x, x, x, x, x, x)
    r(5) = fimhip_abstart(c_null_ptr, 1, 1, i1, i2, i3, d1, d2, d3)
    write (*,'(a,5i4)') 'reject:', r
    if (any(r /= -1)) stop 1
  end subroutine reject

end program fimhip_caller
