! fimhip_bind: the Fortran side of the C ABI in include/fimhip.h, for a serial
! FIM build (-DSERIAL, one GPU).  The bind(C) interfaces keep the argument
! lists of the reference subroutines they replace:
!   fimhip_edgvar1_r8 / fimhip_edgvar3_r8   edgvar1.F90:27, edgvar3.F90:25
!   fimhip_cnuity_r8                        cnuity.F90:23
!   fimhip_mass_adv_r8                      massadv.F90:21
!   fimhip_trcadv_r8                        trcadv.F90:19
!   fimhip_momtum_r8                        momtum.F90:16
!   fimhip_hybgen_r8                        hybgen.F90:20
!   fimhip_hystat_r8                        hystat.F90:23
!   fimhip_transp0/1/2_r8                   transp3d.F90:22, 44, 81
!   fimhip_abstart                          abstart.F90:14
! plus instance creation (in place of GPUinit + copytoGPUinit,
! dyn_init.F90:199, 1198) and the device-resident fast path.  Arrays are
! passed as the reference passes them (explicit-shape, by reference); real*8
! arrays bind to real(c_double).  Every function returns 0 or a negative
! status; fimhip_check stops the program on an error, as dyn_init does on a
! failed GPUinit (dyn_init.F90:199-203).
module fimhip_bind
  use iso_c_binding
  implicit none
  type(c_ptr), save :: fimhip_ctx = c_null_ptr

  type, bind(C) :: fimhip_grid_desc
    integer(c_int) :: nip, nvl, ntra
    real(c_double) :: dt
    type(c_ptr) :: nprox, prox, proxs, cs, sn, sidevec_c, sidevec_e, sideln, &
                   hfcoef, area, lat, lon, thetac, dpsig
  end type

  interface
    integer(c_int) function fimhip_create(grid, device, ctx) bind(C, name="fimhip_create")
      import; type(fimhip_grid_desc), intent(in) :: grid
      integer(c_int), value :: device; type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function fimhip_destroy(ctx) bind(C, name="fimhip_destroy")
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function fimhip_abstart(ctx, its, itsDFI, nf, of, vof, ab1, ab2, ab3) &
        bind(C, name="fimhip_abstart")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, itsDFI
      integer(c_int) :: nf, of, vof; real(c_double) :: ab1, ab2, ab3
    end function
    integer(c_int) function fimhip_edgvar1_r8(ctx, its, u, v, delp, pres, exner, geop, montg, tracr, &
        u_edg, v_edg, dp_edg, lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg) &
        bind(C, name="fimhip_edgvar1_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its
      real(c_double), intent(in) :: u(*), v(*), delp(*), pres(*), exner(*), geop(*), montg(*), tracr(*)
      real(c_double) :: u_edg(*), v_edg(*), dp_edg(*), lp_edg(*), geop_edg(*), mont_edg(*), &
                        uvsq_edg(*), trc_edg(*)
    end function
    integer(c_int) function fimhip_edgvar3_r8(ctx, its, u, v, delp, pres, exner, geop, montg, tracr, &
        u_edg, v_edg, dp_edg, lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg) &
        bind(C, name="fimhip_edgvar3_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its
      real(c_double), intent(in) :: u(*), v(*), delp(*), pres(*), exner(*), geop(*), montg(*), tracr(*)
      real(c_double) :: u_edg(*), v_edg(*), dp_edg(*), lp_edg(*), geop_edg(*), mont_edg(*), &
                        uvsq_edg(*), trc_edg(*)
    end function
    integer(c_int) function fimhip_cnuity_r8(ctx, its, u_velo, v_velo, u_edg, v_edg, dp_edg, lp_edg, &
        delp, pres, exner, dp_tndcy, dplo_tndcy, massfx, omega, TimingBarriers) &
        bind(C, name="fimhip_cnuity_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, TimingBarriers
      real(c_double), intent(in) :: u_velo(*), v_velo(*), u_edg(*), v_edg(*), dp_edg(*), lp_edg(*)
      real(c_double) :: delp(*), pres(*), exner(*), dp_tndcy(*), dplo_tndcy(*), massfx(*), omega(*)
    end function
    integer(c_int) function fimhip_mass_adv_r8(ctx, its, u_velo, v_velo, u_edg, v_edg, dp_edg, lp_edg, &
        delp, pres, exner, dp_tndcy, massfx, omega, TimingBarriers) bind(C, name="fimhip_mass_adv_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, TimingBarriers
      real(c_double), intent(in) :: u_velo(*), v_velo(*), u_edg(*), v_edg(*), lp_edg(*)
      real(c_double) :: dp_edg(*), delp(*), pres(*), exner(*), dp_tndcy(*), massfx(*), omega(*)
    end function
    integer(c_int) function fimhip_set_flux_scheme(ctx, flux_conserv_schm) &
        bind(C, name="fimhip_set_flux_scheme")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: flux_conserv_schm
    end function
    integer(c_int) function fimhip_set_class_b(ctx, ntrb, dtratio) bind(C, name="fimhip_set_class_b")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: ntrb, dtratio
    end function
    integer(c_int) function fimhip_transp0_r8(ctx, its, cumufx, dp3d, dpinit) bind(C, name="fimhip_transp0_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its
      real(c_double), intent(in) :: dp3d(*); real(c_double) :: cumufx(*), dpinit(*)
    end function
    integer(c_int) function fimhip_transp1_r8(ctx, its, nf, of, vof, adbash1, adbash2, adbash3, &
        cumufx, massfx) bind(C, name="fimhip_transp1_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, nf, of, vof
      real(c_double), value :: adbash1, adbash2, adbash3
      real(c_double) :: cumufx(*); real(c_double), intent(in) :: massfx(*)
    end function
    integer(c_int) function fimhip_transp2_r8(ctx, its, tracr, cumufx, dpinit, dpfinl, TimingBarriers) &
        bind(C, name="fimhip_transp2_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, TimingBarriers
      real(c_double) :: tracr(*); real(c_double), intent(in) :: cumufx(*), dpinit(*), dpfinl(*)
    end function
    integer(c_int) function fimhip_trcadv_r8(ctx, its, trcr_edg, tracr, trcdp, trc_tdcy, trclo_tdcy, &
        massfx, delp, TimingBarriers) bind(C, name="fimhip_trcadv_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, TimingBarriers
      real(c_double), intent(in) :: trcr_edg(*), massfx(*), delp(*)
      real(c_double) :: tracr(*), trcdp(*), trc_tdcy(*), trclo_tdcy(*)
    end function
    integer(c_int) function fimhip_momtum_r8(ctx, its, u_velo, v_velo, exner, relvor, u_edg, v_edg, &
        trcr_edg, geop_edg, mont_edg, uvsq_edg, u_tndcy, v_tndcy, dp3d) bind(C, name="fimhip_momtum_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its
      real(c_double), intent(in) :: exner(*), u_edg(*), v_edg(*), trcr_edg(*), geop_edg(*), &
                                    mont_edg(*), uvsq_edg(*), dp3d(*)
      real(c_double) :: u_velo(*), v_velo(*), relvor(*), u_tndcy(*), v_tndcy(*)
    end function
    integer(c_int) function fimhip_hybgen_r8(ctx, its, targt, us3d, vs3d, tr3d, sdot, ex3d, dp3d, pr3d, &
        TimingBarriers) bind(C, name="fimhip_hybgen_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, TimingBarriers
      real(c_double), intent(in) :: targt(*)
      real(c_double) :: us3d(*), vs3d(*), tr3d(*), sdot(*), ex3d(*), dp3d(*), pr3d(*)
    end function
    integer(c_int) function fimhip_hystat_r8(ctx, its, ph3d, ex3d, mp3d, dp3d, tr3d, trdp, psrf, ptdcy) &
        bind(C, name="fimhip_hystat_r8")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its
      real(c_double), intent(in) :: ex3d(*), dp3d(*)
      real(c_double) :: ph3d(*), mp3d(*), tr3d(*), trdp(*), psrf(*), ptdcy(*)
    end function
    ! device-resident fast path
    integer(c_int) function fimhip_upload(ctx, field, host) bind(C, name="fimhip_upload")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: field
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function fimhip_download(ctx, field, host) bind(C, name="fimhip_download")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: field
      real(c_double) :: host(*)
    end function
    integer(c_int) function fimhip_init_state(ctx, init_hybgen) bind(C, name="fimhip_init_state")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: init_hybgen
    end function
    integer(c_int) function fimhip_step(ctx, its, mode) bind(C, name="fimhip_step")
      import; type(c_ptr), value :: ctx; integer(c_int), value :: its, mode
    end function
    integer(c_int) function fimhip_sync(ctx) bind(C, name="fimhip_sync")
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function fimhip_globsum(ctx, mass, trtot) bind(C, name="fimhip_globsum")
      import; type(c_ptr), value :: ctx; real(c_double) :: mass, trtot(*)
    end function
    ! x**y as the device computes it
    real(c_double) function fimhip_pow(x, y) bind(C, name="fimhip_pow")
      import; real(c_double), value :: x, y
    end function
    ! the HYCOM ocean's continuity equation (hycom_cnuity.F90:17): an instance
    ! made in hycom_init from the grid descriptor, glvl.dat's rprox_ln, kdm, the
    ! land/sea mask wet, dtleap, thkdff and biharm
    integer(c_int) function fimhip_ocn_create(grid, rprox_ln, kdm, wet, dtleap, thkdff, biharm, device, ocn) &
        bind(C, name="fimhip_ocn_create")
      import; type(fimhip_grid_desc), intent(in) :: grid
      real(c_double), intent(in) :: rprox_ln(*); integer(c_int), value :: kdm, device
      integer(c_int), intent(in) :: wet(*); real(c_double), value :: dtleap, thkdff, biharm
      type(c_ptr), intent(out) :: ocn
    end function
    integer(c_int) function fimhip_ocn_destroy(ocn) bind(C, name="fimhip_ocn_destroy")
      import; type(c_ptr), value :: ocn
    end function
    integer(c_int) function fimhip_ocn_cnuity_r8(ocn, nstep, leapm, leapn, um_edg, vm_edg, un_edg, vn_edg, &
        dpm_edg, dp, pres, massflx, mssfx_avg) bind(C, name="fimhip_ocn_cnuity_r8")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: nstep, leapm, leapn
      real(c_double), intent(in) :: um_edg(*), vm_edg(*), un_edg(*), vn_edg(*), dpm_edg(*)
      real(c_double) :: dp(*), pres(*), massflx(*), mssfx_avg(*)
    end function
    integer(c_int) function fimhip_ocn_edgvar_r8(ocn, nstep, leapm, leapn, uclin, vclin, utrop, vtrop, dp, &
        pres, geop, montg, spcifv, um_edg, vm_edg, un_edg, vn_edg, dp_edg, pr_edg, geop_edg, mont_edg, &
        uvsq_edg, spcv_edg) bind(C, name="fimhip_ocn_edgvar_r8")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: nstep, leapm, leapn
      real(c_double), intent(in) :: uclin(*), vclin(*), utrop(*), vtrop(*), dp(*), geop(*), montg(*), spcifv(*)
      real(c_double) :: pres(*), um_edg(*), vm_edg(*), un_edg(*), vn_edg(*), dp_edg(*), pr_edg(*), &
                        geop_edg(*), mont_edg(*), uvsq_edg(*), spcv_edg(*)
    end function
    integer(c_int) function fimhip_ocn_momtum_r8(ocn, nstep, leapm, leapn, utrop, vtrop, uclin, vclin, um_edg, &
        vm_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg, taux, tauy, ubforc, vbforc, dp, pres, ustarb) &
        bind(C, name="fimhip_ocn_momtum_r8")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: nstep, leapm, leapn
      real(c_double), intent(in) :: utrop(*), vtrop(*), um_edg(*), vm_edg(*), geop_edg(*), mont_edg(*), &
                                    uvsq_edg(*), spcv_edg(*), taux(*), tauy(*), dp(*)
      real(c_double) :: uclin(*), vclin(*), ubforc(*), vbforc(*), pres(*), ustarb(*)
    end function
    integer(c_int) function fimhip_ocn_set_momtum(ocn, batrop, veldff, smagcf) bind(C, name="fimhip_ocn_set_momtum")
      import; type(c_ptr), value :: ocn; real(c_double), value :: batrop, veldff, smagcf
    end function
    integer(c_int) function fimhip_ocn_set_transport(ocn, numtr, temdff) bind(C, name="fimhip_ocn_set_transport")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: numtr; real(c_double), value :: temdff
    end function
    ! hycom_transp3d.F90:18 transp0, :54 transp1, :88 transp2 (passv_tr is
    ! not read when numtr = 0)
    integer(c_int) function fimhip_ocn_transp0_r8(ocn, nstep, leapn, cumuflx, dp, dpinit) &
        bind(C, name="fimhip_ocn_transp0_r8")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: nstep, leapn
      real(c_double), intent(in) :: dp(*)
      real(c_double) :: cumuflx(*), dpinit(*)
    end function
    integer(c_int) function fimhip_ocn_transp1_r8(ocn, nstep, cumuflx, massflx) bind(C, name="fimhip_ocn_transp1_r8")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: nstep
      real(c_double) :: cumuflx(*)
      real(c_double), intent(in) :: massflx(*)
    end function
    integer(c_int) function fimhip_ocn_transp2_r8(ocn, nstep, temp, saln, passv_tr, cumuflx, dpinit, dpfinl) &
        bind(C, name="fimhip_ocn_transp2_r8")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: nstep
      real(c_double) :: temp(*), saln(*), passv_tr(*)
      real(c_double), intent(in) :: cumuflx(*), dpinit(*), dpfinl(*)
    end function
    ! hycom_hybgn1.F90:9 hybgn1 (passv_tr: pass c_null_ptr for none, as
    ! hycom_run.F90:516 calls it)
    integer(c_int) function fimhip_ocn_set_hybgen(ocn, targt, salmin, dpmin, latij) &
        bind(C, name="fimhip_ocn_set_hybgen")
      import; type(c_ptr), value :: ocn
      real(c_double), intent(in) :: targt(*), salmin(*), dpmin(*), latij(*)
    end function
    integer(c_int) function fimhip_ocn_hybgn1_r8(ocn, nstep, leap, targt, dens, dp, pres, temp, saln, uvel, &
        vvel, latij, passv_tr) bind(C, name="fimhip_ocn_hybgn1_r8")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: nstep, leap
      real(c_double), intent(in) :: targt(*), latij(*)
      real(c_double) :: dens(*), dp(*), pres(*), temp(*), saln(*), uvel(*), vvel(*)
      type(c_ptr), value :: passv_tr
    end function
    integer(c_int) function fimhip_ocn_sync(ocn) bind(C, name="fimhip_ocn_sync")
      import; type(c_ptr), value :: ocn
    end function
    integer(c_int) function fimhip_ocn_set_janjic(ocn, on) bind(C, name="fimhip_ocn_set_janjic")
      import; type(c_ptr), value :: ocn; integer(c_int), value :: on
    end function
    integer(c_int) function fimhip_ocn_set_dtleap(ocn, dtleap) bind(C, name="fimhip_ocn_set_dtleap")
      import; type(c_ptr), value :: ocn; real(c_double), value :: dtleap
    end function
    integer(c_int) function fimhip_ocn_mass(ocn, totmass, massglb0, masscor) bind(C, name="fimhip_ocn_mass")
      import; type(c_ptr), value :: ocn; real(c_double) :: totmass, massglb0, masscor
    end function
    integer(c_int) function fimhip_ocn_set_mass(ocn, totmass, massglb0, masscor) bind(C, name="fimhip_ocn_set_mass")
      import; type(c_ptr), value :: ocn; real(c_double), value :: totmass, massglb0, masscor
    end function
  end interface

  ! field ids of fimhip_upload / fimhip_download (the FIMHIP_* enum)
  integer(c_int), parameter :: FIMHIP_US = 0, FIMHIP_VS = 1, FIMHIP_DP = 2, FIMHIP_PR = 3, &
    FIMHIP_EX = 4, FIMHIP_PH = 5, FIMHIP_MP = 6, FIMHIP_TR = 7, FIMHIP_TRDP = 8

contains

  subroutine fimhip_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= 0) then
      write(*,'(a,a,a,i0)') 'fimhip: ', where, ' failed with status ', rc
      stop 1
    end if
  end subroutine

end module fimhip_bind
