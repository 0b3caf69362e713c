! restref: driver for the reference restart writer and reader of the
! dynamics section (restart_dyn.F90, with writearr32_r.F90 / readarr32_r.F90),
! compiled from the reference sources by oracle/build_ref.sh in default real
! 4, as FIM runs.  Test infrastructure only.
!
! Usage: restref write <values.bin> <restart.out>   fill every field the dynamics
!                                                     section holds from a raw stream,
!                                                     then write_restart_dyn
!        restref read  <restart.in> <values.out>    read_restart_dyn, then dump every
!                                                     field as a raw stream
! values.bin (stream, little endian): int32 nvl, nip, ntra, ntrb, then the
! items of the six header records in the order write_restart_dyn writes them,
! then the whole arrays in the order of the last two read statements below
! (int32, real32 or logical(4) each).  The read mode needs <restart.in>.dims
! holding the four int32 sizes.

program restref
  use module_control,   only: nip, nvlp1, ntra, ntrb, npp, nabl, next_date, have_next_sst
  use fimnamelist,      only: nvl
  use module_constants, only: dpsig, thetac, lat, lon, nprox, proxs, area, cs, sn, sidevec_c, sidevec_e, &
                              sideln, rsideln, rprox_ln, rarea, corio, deg_lat, deg_lon, sigak, sigbk
  use module_variables, only: curr_write_time, nf, of, vof, adbash1, adbash2, adbash3, psrf, ptdcy, pw2d, &
                              u_tdcy, v_tdcy, dp_tdcy, dpl_tdcy, tr3d, trdp, trc_tdcy, trl_tdcy, us3d, vs3d, &
                              ws3d, mp3d, tk3d, dp3d, rh3d, relvor, dpinit, pr3d, ex3d, ph3d, sdot, massfx, cumufx
  use module_sfc_variables, only: rn2d, rc2d, u10m, v10m, sst_prev, sst_next, fice2d_prev, fice2d_next, &
                                  fice2d, runoff
  use module_globsum,   only: qmstr, qmstrc, qmstrn, qdtr_set
  use restart_dyn,      only: read_restart_dyn, write_restart_dyn
  implicit none
  character(len=512) :: mode, fin, fout
  integer(4) :: dims(4)

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  if (trim(mode) == 'write') then
    open (21, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
    read (21) dims
    call setup(dims)
    read (21) curr_write_time, nf, of, vof, adbash1, adbash2, adbash3, thetac, dpsig, sigak, sigbk
    read (21) lat, lon, nprox, proxs, area, cs, sn, psrf, ptdcy
    read (21) qmstr, qmstrc, qmstrn, qdtr_set
    read (21) sidevec_c, sidevec_e, sideln, rprox_ln
    read (21) sst_prev, sst_next, fice2d_prev, fice2d_next, fice2d, next_date, have_next_sst
    read (21) rarea, rsideln, corio, deg_lat, deg_lon, rn2d, pw2d, rc2d, u10m, v10m, runoff
    read (21) u_tdcy, v_tdcy, dp_tdcy, dpl_tdcy, tr3d, trdp, trc_tdcy, trl_tdcy
    read (21) us3d, vs3d, ws3d, mp3d, tk3d, dp3d, rh3d, relvor, dpinit, pr3d, ex3d, ph3d, sdot, massfx, cumufx
    close (21)
    open (31, file=trim(fout), form='unformatted', status='replace', action='write')
    call write_restart_dyn (31)
    close (31)
  else
    open (21, file=trim(fin), form='unformatted', status='old', action='read')
    open (22, file=trim(fin)//'.dims', access='stream', form='unformatted', status='old', action='read')
    read (22) dims
    close (22)
    call setup(dims)
    call read_restart_dyn (21)
    close (21)
    open (31, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write (31) dims
    write (31) curr_write_time, nf, of, vof, adbash1, adbash2, adbash3, thetac, dpsig, sigak, sigbk
    write (31) lat, lon, nprox, proxs, area, cs, sn, psrf, ptdcy
    write (31) qmstr, qmstrc, qmstrn, qdtr_set
    write (31) sidevec_c, sidevec_e, sideln, rprox_ln
    write (31) sst_prev, sst_next, fice2d_prev, fice2d_next, fice2d, next_date, have_next_sst
    write (31) rarea, rsideln, corio, deg_lat, deg_lon, rn2d, pw2d, rc2d, u10m, v10m, runoff
    write (31) u_tdcy, v_tdcy, dp_tdcy, dpl_tdcy, tr3d, trdp, trc_tdcy, trl_tdcy
    write (31) us3d, vs3d, ws3d, mp3d, tk3d, dp3d, rh3d, relvor, dpinit, pr3d, ex3d, ph3d, sdot, massfx, cumufx
    close (31)
  end if

contains

  subroutine setup(d)
    integer(4), intent(in) :: d(4)
    nvl = d(1); nip = d(2); ntra = d(3); ntrb = d(4)
    nvlp1 = nvl + 1
    allocate (thetac(nvl), dpsig(nvl), sigak(nvlp1), sigbk(nvlp1))
    allocate (lat(nip), lon(nip), nprox(nip), proxs(npp,nip), area(nip), cs(4,npp,nip), sn(4,npp,nip))
    allocate (psrf(nip), ptdcy(nip,2))
    allocate (sidevec_c(2,npp,nip), sidevec_e(2,npp,nip), sideln(npp,nip), rprox_ln(npp,nip))
    allocate (sst_prev(nip), sst_next(nip), fice2d_prev(nip), fice2d_next(nip), fice2d(nip))
    allocate (rarea(nip), rsideln(npp,nip), corio(nip), deg_lat(nip), deg_lon(nip), rn2d(nip), pw2d(nip))
    allocate (rc2d(nip), u10m(nip), v10m(nip), runoff(nip))
    allocate (u_tdcy(nvl,nip,nabl), v_tdcy(nvl,nip,nabl), dp_tdcy(nvl,nip,nabl), dpl_tdcy(nvl,nip,nabl))
    allocate (tr3d(nvl,nip,ntra+ntrb), trdp(nvl,nip,ntra+ntrb))
    allocate (trc_tdcy(nvl,nip,nabl,ntra+ntrb), trl_tdcy(nvl,nip,nabl,ntra+ntrb))
    allocate (us3d(nvl,nip), vs3d(nvl,nip), ws3d(nvl,nip), mp3d(nvl,nip), tk3d(nvl,nip), dp3d(nvl,nip))
    allocate (rh3d(nvl,nip), relvor(nvl,nip), dpinit(nvl,nip))
    allocate (pr3d(nvlp1,nip), ex3d(nvlp1,nip), ph3d(nvlp1,nip), sdot(nvlp1,nip))
    allocate (massfx(nvl,npp,nip,nabl), cumufx(nvl,npp,nip))
  end subroutine setup

end program restref
