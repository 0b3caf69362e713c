! dynref: driver that runs the reference FIM dynamics hot path, compiled from
! the reference sources by oracle/build_ref.sh, on a grid file (glvl.dat) and a
! seeded state file.  Test infrastructure only.
!
! Usage: dynref <glvl.dat> <state.bin> <out.bin>
!
! State file (stream, little endian):
!   int32  magic(=1179208497), nvl, nip, ntra, nsteps, mode, dumpstep, flags
!   real64 dt
!   real64 thetac(nvl), dpsig(nvl)
!   real64 us(nvl,nip), vs(nvl,nip), dp(nvl,nip), pr(nvl+1,nip), ex(nvl+1,nip),
!          ph(nvl+1,nip), tr(nvl,nip,ntra)
! mode 0: full step = DYN_1(its) then DYN_2(its)   (dyn_run.F90:99-425 order)
! mode 1: transport = edgvar1, edgvar3, cnuity, trcadv with fixed winds
! flags bit 0: run the dyn_init start-up hybgen(its=0) + pressure refresh
!              (dyn_init.F90:880-893) before the start-up hystat.
! flags bit 1: continuity by mass_adv (flux_conserv_schm = 2,
!              dyn_run.F90:377-386) instead of cnuity.
! flags bits 8-15: ntrb, class-B tracers (tr in the state file then holds
!              ntra+ntrb fields), transported by transp1/transp2/transp0
!              (dyn_run.F90:388-405, dyn_init.F90:1096-1098);
!              bits 16-23: dtratio.
!
! Output: a sequence of records  character(16) tag, int64 n, real64 vals(n).
! Stage records are written for step 'dumpstep' after every sub-call.

program dynref
  use module_control,   only: nip, nvlp1, dt, ntra, ntrb, npp, nabl, kbl
  use fimnamelist,      only: nvl, intfc_smooth, veldff_bkgnd, veldff_boost, dtratio
  use module_constants
  use module_variables, only: nf, of, vof, adbash1, adbash2, adbash3, worka, workb, q_min_out, q_max_out
  use global_bounds,    only: set_global_bounds, set_global_bounds_halo
  use module_abstart,   only: abstart
  use module_edgvar1,   only: edgvar1
  use module_edgvar3,   only: edgvar3
  use cnuity_driver,    only: cnuity
  use massadv_driver,   only: mass_adv
  use trcadv_driver,    only: trcadv
  use momtum_driver,    only: momtum
  use module_hybgen,    only: hybgen
  use module_hystat,    only: hystat
  use module_transp3d,  only: transp0, transp1, transp2
  use module_globsum,   only: globsum, qmass
  implicit none

  character(len=512) :: fgrid, fstate, fout
  integer(4) :: hdr(8)
  real(8)    :: dt8
  integer    :: nsteps, mode, dumpstep, flags, its, k, ipn, isn, edg, idx, t, n
  integer    :: ndamp
  real       :: wgt
  real(8), allocatable :: r8a(:,:), r8b(:,:,:), r81(:)
  real(4), allocatable :: work4(:)
  integer, allocatable :: iwork(:)
  character(16) :: header

  real, allocatable :: us3d(:,:), vs3d(:,:), dp3d(:,:), pr3d(:,:), ex3d(:,:), ph3d(:,:)
  real, allocatable :: mp3d(:,:), tr3d(:,:,:), trdp(:,:,:), sdot(:,:), ws3d(:,:), relvor(:,:)
  real, allocatable :: psrf(:), ptdcy(:,:)
  real, allocatable :: u_tdcy(:,:,:), v_tdcy(:,:,:), dp_tdcy(:,:,:), dpl_tdcy(:,:,:)
  real, allocatable :: trc_tdcy(:,:,:,:), trl_tdcy(:,:,:,:), massfx(:,:,:,:)
  real, allocatable :: u_edg(:,:,:), v_edg(:,:,:), dp_edg(:,:,:), lp_edg(:,:,:)
  real, allocatable :: geop_edg(:,:,:), mont_edg(:,:,:), uvsq_edg(:,:,:), trc_edg(:,:,:,:)
  real, allocatable :: cumufx(:,:,:), dpinit(:,:)
  real, allocatable :: rn2d(:), rc2d(:), qf2d(:), qtrcr(:)

  call get_command_argument(1, fgrid)
  call get_command_argument(2, fstate)
  call get_command_argument(3, fout)

  open (21, file=trim(fstate), access='stream', form='unformatted', status='old', action='read')
  read (21) hdr
  if (hdr(1) /= 1179208497) stop 'dynref: bad state magic'
  nvl = hdr(2); nip = hdr(3); ntra = hdr(4); nsteps = hdr(5)
  mode = hdr(6); dumpstep = hdr(7); flags = hdr(8)
  ntrb = iand(ishft(flags, -8), 255)
  dtratio = iand(ishft(flags, -16), 255)
  read (21) dt8
  dt = dt8
  nvlp1 = nvl + 1
  call set_global_bounds ()
  call set_global_bounds_halo ()

  ! ---- grid (dyn_init.F90:373-548) ----
  allocate (lat(nip), lon(nip), nprox(nip), proxs(npp,nip), prox(npp,nip), area(nip))
  allocate (cs(4,npp,nip), sn(4,npp,nip), sidevec_c(2,npp,nip), sidevec_e(2,npp,nip))
  allocate (sideln(npp,nip), rsideln(npp,nip), rprox_ln(npp,nip), corner_xy(npp,2,nip))
  allocate (hfcoef(6,npp,nip), perm(nip), inv_perm(nip), nedge(nip), permedge(npp,nip))
  allocate (rarea(nip), corio(nip), grndspd(nip), deg_lat(nip), deg_lon(nip))
  allocate (idxtrl(npp,3,nip), idxldg(npp,3,nip), wgttrl(npp,3,nip), wgtldg(npp,3,nip))
  allocate (work4(nip), iwork(nip))
  open (22, file=trim(fgrid), form='unformatted', status='old', action='read')
  read (22) header
  read (22) header
  read (22) work4; lat = work4
  read (22) work4; lon = work4
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
  do ipn = 1, nip
    perm(inv_perm(ipn)) = ipn
  end do
  do idx = 1, 2
    do isn = 1, npp
      read (22) work4; corner_xy(isn,idx,:) = work4
    end do
  end do
  do isn = 1, npp
    do idx = 1, 6
      read (22) work4; hfcoef(idx,isn,:) = work4
    end do
  end do
  close (22)
  do ipn = 1, nip
    if (nprox(ipn) == 5) prox(6,ipn) = prox(5,ipn)
    nedge(ipn) = nprox(ipn)
    do isn = 1, nprox(ipn)
      permedge(isn,ipn) = isn
    end do
  end do
  ! derived geometry (dyn_init.F90:753-757)
  do ipn = 1, nip
    corio(ipn) = omegx2*sin(lat(ipn))
    deg_lat(ipn) = raddeg*lat(ipn)
    deg_lon(ipn) = raddeg*lon(ipn)
    rarea(ipn) = 1./area(ipn)
    grndspd(ipn) = earthrad*2.*pi*cos(lat(ipn))/86164.
    do isn = 1, nprox(ipn)
      rsideln(isn,ipn) = 1./sideln(isn,ipn)
    end do
  end do
  ! edge stencils with equal weights (cornwgt.F90:199-201,247-250,268-270,313-316,331-333)
  idxtrl = -999; idxldg = -999
  do ipn = 1, nip
    do edg = 1, nprox(ipn)
      idxtrl(edg,1,ipn) = ipn
      idxtrl(edg,2,ipn) = prox(edg,ipn)
      idxtrl(edg,3,ipn) = prox(mod(edg-2+nprox(ipn),nprox(ipn))+1,ipn)
      idxldg(edg,1,ipn) = ipn
      idxldg(edg,2,ipn) = prox(edg,ipn)
      idxldg(edg,3,ipn) = prox(mod(edg,nprox(ipn))+1,ipn)
    end do
  end do
  wgttrl = 1./3.; wgtldg = 1./3.

  ! ---- vertical parameters (dyn_init.F90:309-338) ----
  allocate (dfflen(nvl), biharw(nvl), smoo_coeff(nvl), thetac(nvl), dpsig(nvl), sigak(nvlp1), sigbk(nvlp1))
  ndamp = 0.1*nvl
  do k = 1, nvl
    wgt = max(0., real(k-nvl+ndamp)/real(ndamp))
    dfflen(k) = dt*(veldff_bkgnd*(1.-wgt) + max(veldff_bkgnd, veldff_boost)*wgt)
  end do
  dfflen_max = maxval(dfflen)
  smoo_coeff(:) = 0.
  do k = kbl+2, nvl
    smoo_coeff(k) = dt*intfc_smooth*taper(k,kbl+1)*taper(k,nvlp1)
  end do
  biharw(:) = 0.
  sigak = 0.; sigbk = 0.

  ! ---- state ----
  allocate (r81(nvl))
  read (21) r81; thetac = r81
  read (21) r81; dpsig = r81
  allocate (us3d(nvl,nip), vs3d(nvl,nip), dp3d(nvl,nip), pr3d(nvlp1,nip), ex3d(nvlp1,nip))
  allocate (ph3d(nvlp1,nip), mp3d(nvl,nip), tr3d(nvl,nip,ntra+ntrb), trdp(nvl,nip,ntra))
  allocate (cumufx(nvl,npp,nip), dpinit(nvl,nip))
  allocate (sdot(nvlp1,nip), ws3d(nvl,nip), relvor(nvl,nip), psrf(nip), ptdcy(nip,2))
  allocate (u_tdcy(nvl,nip,nabl), v_tdcy(nvl,nip,nabl), dp_tdcy(nvl,nip,nabl), dpl_tdcy(nvl,nip,nabl))
  allocate (trc_tdcy(nvl,nip,nabl,ntra), trl_tdcy(nvl,nip,nabl,ntra), massfx(nvl,npp,nip,nabl))
  allocate (u_edg(nvl,npp,nip), v_edg(nvl,npp,nip), dp_edg(nvl,npp,nip), lp_edg(nvl,npp,nip))
  allocate (geop_edg(nvl,npp,nip), mont_edg(nvl,npp,nip), uvsq_edg(nvl,npp,nip), trc_edg(nvl,npp,nip,ntra))
  allocate (worka(nvl,nip), workb(nvl,nip))
  allocate (q_min_out(nvl,nip), q_max_out(nvl,nip))   ! dyn_alloc.F90:284-302
  allocate (rn2d(nip), rc2d(nip), qf2d(nip), qtrcr(ntra+ntrb))   ! physics inputs of globsum: none here
  rn2d = 0.; rc2d = 0.; qf2d = 0.; qtrcr = 0.
  q_min_out = 0.
  q_max_out = 0.
  allocate (r8a(nvl,nip))
  read (21) r8a; us3d = r8a
  read (21) r8a; vs3d = r8a
  read (21) r8a; dp3d = r8a
  deallocate (r8a); allocate (r8a(nvlp1,nip))
  read (21) r8a; pr3d = r8a
  read (21) r8a; ex3d = r8a
  read (21) r8a; ph3d = r8a
  allocate (r8b(nvl,nip,ntra+ntrb))
  read (21) r8b; tr3d = r8b
  close (21)

  ! zeroed start-up fields (dyn_init.F90:1042-1094)
  u_tdcy = 0.; v_tdcy = 0.; dp_tdcy = 0.; dpl_tdcy = 0.
  trc_tdcy = 0.; trl_tdcy = 0.; massfx = 0.
  ws3d = 0.; relvor = 0.; sdot = 0.; psrf = 0.; ptdcy = 0.; mp3d = 0.
  u_edg = 0.; v_edg = 0.; dp_edg = 0.; lp_edg = 0.; geop_edg = 0.; mont_edg = 0.
  uvsq_edg = 0.; trc_edg = 0.
  do t = 1, ntra
    do ipn = 1, nip
      do k = 1, nvl
        trdp(k,ipn,t) = tr3d(k,ipn,t)*dp3d(k,ipn)
      end do
    end do
  end do

  open (23, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  call out1('grid.corio', real(corio,8))
  call out1('grid.rarea', real(rarea,8))
  call out1('grid.grndspd', real(grndspd,8))
  call out1('vert.dfflen', real(dfflen,8))
  call out1('vert.smoo', real(smoo_coeff,8))

  its = 0
  if (btest(flags,0)) then
    call hybgen (its, thetac, us3d, vs3d, tr3d, sdot, ex3d, dp3d, pr3d, .false.)
    do ipn = 1, nip
      pr3d(1,ipn) = p1000*(ex3d(1,ipn)/cp)**(cp/rd)
      do k = 1, nvl
        pr3d(k+1,ipn) = p1000*(ex3d(k+1,ipn)/cp)**(cp/rd)
      end do
    end do
    call dumpstate('init.hybgen')
  end if
  call hystat (its, ph3d, ex3d, mp3d, dp3d, tr3d, trdp, psrf, ptdcy)
  call dumpstate('init')
  call dumpsums('init')
  nf = 0; of = 2; vof = 1
  if (ntrb > 0) call transp0 (its, cumufx, dp3d, dpinit)

  do its = 1, nsteps
    call abstart (its, 1, nf, of, vof, adbash1, adbash2, adbash3)
    call edgvar1 (its, us3d, vs3d, dp3d, pr3d, ex3d, ph3d, mp3d, tr3d, u_edg, v_edg, dp_edg, &
                  lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg)
    if (its == dumpstep) call dumpedges('edgvar1')
    call edgvar3 (its, us3d, vs3d, dp3d, pr3d, ex3d, ph3d, mp3d, tr3d, u_edg, v_edg, dp_edg, &
                  lp_edg, geop_edg, mont_edg, uvsq_edg, trc_edg)
    if (its == dumpstep) call dumpedges('edgvar3')
    if (btest(flags,1)) then
      call mass_adv (its, us3d, vs3d, u_edg, v_edg, dp_edg, lp_edg, dp3d, pr3d, ex3d, &
                     dp_tdcy, massfx, ws3d, .false.)
    else
      call cnuity (its, us3d, vs3d, u_edg, v_edg, dp_edg, lp_edg, dp3d, pr3d, ex3d, &
                   dp_tdcy, dpl_tdcy, massfx, ws3d, .false.)
    end if
    if (ntrb > 0) then
      call transp1 (its, nf, of, vof, adbash1, adbash2, adbash3, cumufx, massfx)
      if (its == dumpstep) call out3('transp1.cumufx', real(cumufx,8))
      if (mod(its,dtratio) == 0) then
        call transp2 (its, tr3d, cumufx, dpinit, dp3d, .false.)
        if (its == dumpstep) then
          do t = ntra+1, ntra+ntrb
            call out2('transp2.trb', real(tr3d(:,:,t),8))
          end do
        end if
        call transp0 (its, cumufx, dp3d, dpinit)
      end if
    end if
    if (its == dumpstep) then
      call dumpstate('cnuity')
      call out3('cnuity.massfx', real(massfx(:,:,:,nf),8))
      call out2('cnuity.omega', real(ws3d,8))
      call out2('cnuity.dptdcy', real(dp_tdcy(:,:,nf),8))
      call out2('cnuity.dpltdcy', real(dpl_tdcy(:,:,nf),8))
    end if
    call trcadv (its, trc_edg, tr3d, trdp, trc_tdcy, trl_tdcy, massfx, dp3d, .false.)
    if (its == dumpstep) then
      call dumpstate('trcadv')
      do t = 1, ntra
        call out2('trcadv.trdp', real(trdp(:,:,t),8))
        call out2('trcadv.tdcy', real(trc_tdcy(:,:,nf,t),8))
      end do
    end if
    if (mode == 0) then
      call momtum (its, us3d, vs3d, ex3d, relvor, u_edg, v_edg, trc_edg, geop_edg, mont_edg, &
                   uvsq_edg, u_tdcy, v_tdcy, dp3d)
      if (its == dumpstep) then
        call dumpstate('momtum')
        call out2('momtum.relvor', real(relvor,8))
        call out2('momtum.utdcy', real(u_tdcy(:,:,nf),8))
        call out2('momtum.vtdcy', real(v_tdcy(:,:,nf),8))
      end if
      do ipn = 1, nip
        worka(:,ipn) = tr3d(:,ipn,1)
      end do
      call hybgen (its, thetac, us3d, vs3d, tr3d, sdot, ex3d, dp3d, pr3d, .false.)
      if (its == dumpstep) then
        call dumpstate('hybgen')
        call out2('hybgen.sdot', real(sdot,8))
      end if
      call hystat (its, ph3d, ex3d, mp3d, dp3d, tr3d, trdp, psrf, ptdcy)
      if (its == dumpstep) call dumpstate('hystat')
    end if
  end do
  call dumpstate('final')
  call dumpsums('final')
  call out1('final.psrf', real(psrf,8))
  do t = 1, ntra
    call out2('final.trdp', real(trdp(:,:,t),8))
  end do
  call out2('final.dptdcy', real(dp_tdcy(:,:,nf),8))
  call out2('final.utdcy', real(u_tdcy(:,:,nf),8))
  if (ntrb > 0) then
    do t = ntra+1, ntra+ntrb
      call out2('final.trb', real(tr3d(:,:,t),8))
    end do
    call out3('final.cumufx', real(cumufx,8))
    call out2('final.dpinit', real(dpinit,8))
  end if
  close (23)

contains

  real function taper(k1, k2)   ! dyn_init.F90:125
    integer, intent(in) :: k1, k2
    taper = float(k2-k1)**2/(9.+float(k2-k1)**2)
  end function taper

  subroutine outrec(tag, n, v)
    character(*), intent(in) :: tag
    integer, intent(in) :: n
    real(8), intent(in) :: v(n)
    character(16) :: t16
    t16 = tag
    write (23) t16, int(n,8), v
  end subroutine outrec

  subroutine out1(tag, v)
    character(*), intent(in) :: tag
    real(8), intent(in) :: v(:)
    call outrec(tag, size(v), v)
  end subroutine out1

  subroutine out2(tag, v)
    character(*), intent(in) :: tag
    real(8), intent(in) :: v(:,:)
    call outrec(tag, size(v), reshape(v, [size(v)]))
  end subroutine out2

  subroutine out3(tag, v)
    character(*), intent(in) :: tag
    real(8), intent(in) :: v(:,:,:)
    call outrec(tag, size(v), reshape(v, [size(v)]))
  end subroutine out3

  subroutine dumpstate(stage)
    character(*), intent(in) :: stage
    integer :: tt
    call out2(stage//'.us', real(us3d,8))
    call out2(stage//'.vs', real(vs3d,8))
    call out2(stage//'.dp', real(dp3d,8))
    call out2(stage//'.pr', real(pr3d,8))
    call out2(stage//'.ex', real(ex3d,8))
    call out2(stage//'.ph', real(ph3d,8))
    call out2(stage//'.mp', real(mp3d,8))
    do tt = 1, ntra
      call out2(stage//'.tr', real(tr3d(:,:,tt),8))
    end do
  end subroutine dumpstate

  subroutine dumpsums(stage)     ! global mass and tracer totals (globsum.F90:39-130)
    character(*), intent(in) :: stage
    call globsum (its, dp3d, tr3d, rn2d, rc2d, pr3d, ex3d, qf2d, qtrcr)
    call out1(stage//'.qmass', [real(qmass,8)])
    call out1(stage//'.qtrcr', real(qtrcr,8))
  end subroutine dumpsums

  subroutine dumpedges(stage)
    character(*), intent(in) :: stage
    integer :: tt
    call out3(stage//'.u', real(u_edg,8))
    call out3(stage//'.v', real(v_edg,8))
    call out3(stage//'.dp', real(dp_edg,8))
    call out3(stage//'.lp', real(lp_edg,8))
    call out3(stage//'.mont', real(mont_edg,8))
    call out3(stage//'.uvsq', real(uvsq_edg,8))
    do tt = 1, ntra
      call out3(stage//'.trc', real(trc_edg(:,:,:,tt),8))
    end do
  end subroutine dumpedges

end program dynref
