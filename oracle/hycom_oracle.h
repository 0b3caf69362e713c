/* hycom_oracle.h - the CPU checker of the HYCOM ocean continuity equation
 * and edge interpolation (test infrastructure only).  See hycom_oracle.c for
 * the array layouts. */
#pragma once

typedef struct {
  int nip, kdm;
  const int *prox, *proxs, *nprox, *wet;     /* 0-based neighbour tables; wet > 0: ocean */
  const double *area, *rarea, *sidevec_e, *sideln, *rsideln, *rprox_ln;
  double dtleap, thkdff, biharm;
} fo_hyc_grid;

/* hycom_variables.F90:128-135: the total mass, its value at the first step
 * and the thickness factor that corrects its drift */
typedef struct {
  double totmass, massglb0, masscor;
} fo_hyc_mass;

/* hycom_cnuity.F90 cnuity(nstep, leapm, leapn, ...); returns 1 where the
 * reference's duplicate-flux check would stop the program */
int fo_hycom_cnuity(const fo_hyc_grid *g, fo_hyc_mass *ms, int nstep, int leapm, int leapn,
                    const double *um_edg, const double *vm_edg, const double *un_edg, const double *vn_edg,
                    const double *dpm_edg, double *dp, double *pres, double *massflx, double *mssfx_avg);

/* hycom_edgvar.F90 bclin_edgvar(nstep, leapm, leapn, uclin, vclin, utrop,
 * vtrop, dp, pres, geop, montg, spcifv, um_edg, vm_edg, un_edg, vn_edg,
 * dp_edg, pr_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg) with uvedge and
 * blokf; cs/sn (4,6,nip) as [(i*6+e)*4 + j].  Only the grid fields nip, kdm,
 * prox, nprox and wet are read. */
void fo_hycom_edgvar(const fo_hyc_grid *g, const double *cs, const double *sn, int janjic, int leapm, int leapn,
                     const double *uclin, const double *vclin, const double *utrop, const double *vtrop,
                     const double *dp, double *pres, const double *geop, const double *montg,
                     const double *spcifv, double *um_edg, double *vm_edg, double *un_edg, double *vn_edg,
                     double *dp_edg, double *pr_edg, double *geop_edg, double *mont_edg, double *uvsq_edg,
                     double *spcv_edg);

/* the grid and namelist values hycom_momtum reads beyond fo_hyc_grid:
 * cs/sn (4,npp,nip) as [(i*6+e)*4 + j], sidevec_c (2,npp,nip) as
 * [(i*6+e)*2 + c], the latitude (radians; corio = omegx2*sin(lat) as
 * dyn_init.F90:753 forms it, and meshsz = sqrt(area) as hycom_init.F90:327
 * does, with the C library's sin and sqrt), the barotropic time step batrop,
 * veldff and smagcf (fimnamelist.F90:253-254) and janjic_ocn */
typedef struct {
  const double *cs, *sn, *sidevec_c, *lat;
  double batrop, veldff, smagcf;
  int janjic;
} fo_hyc_mom;

/* hycom_momtum.F90 momtum(nstep, leapm, leapn, utrop, vtrop, uclin, vclin,
 * um_edg, vm_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg, taux, tauy,
 * ubforc, vbforc, dp, pres, ustarb) with hycom_dissip's del4prep and dissip;
 * modesplit is .false. (hycom_control.F90:5), so ubforc and vbforc are not
 * set and are left out.  Reads nip, kdm, prox, nprox, wet, rarea, sideln,
 * sidevec_e and dtleap of g. */
void fo_hycom_momtum(const fo_hyc_grid *g, const fo_hyc_mom *m, int leapm, int leapn, const double *utrop,
                     const double *vtrop, double *uclin, double *vclin, const double *um_edg, const double *vm_edg,
                     const double *geop_edg, const double *mont_edg, const double *uvsq_edg,
                     const double *spcv_edg, const double *taux, const double *tauy, const double *dp,
                     double *pres, double *ustarb);

/* hycom_transp3d.F90: transp0 (start of a transport interval), transp1 (the
 * mass flux added to its time integral, cumuflx += massflx*2*batrop) and
 * transp2 (the tracers moved by hycom_fct3d.F90's fct3d with the integral,
 * then temperature and salinity diffused by hycom_dffusn.F90's dffusn_lyr
 * when temdff > 0).  temp, saln (kdm,nip); passv (kdm,nip,numtr) as
 * [(t*nip + i)*kdm + k]; cumuflx, massflx edge layout; dpinit, dpfinl
 * (kdm,nip).  Reads nip, kdm, prox, proxs, nprox, wet, area, rarea and
 * sideln of g. */
void fo_hycom_transp0(const fo_hyc_grid *g, int leapn, const double *dp, double *dpinit, double *cumuflx);
void fo_hycom_transp1(const fo_hyc_grid *g, double batrop, const double *massflx, double *cumuflx);
void fo_hycom_transp2(const fo_hyc_grid *g, int numtr, double batrop, double temdff, double *temp, double *saln,
                      double *passv, const double *cumuflx, const double *dpinit, const double *dpfinl);

/* hycom_hybgn1.F90 hybgn1(nstep, leap, targt, dens, dp, pres, temp, saln,
 * uvel, vvel, latij[, passv_tr]) with hybgn1_1d, ppmad3, pcwise and
 * equilb, and hycom_sigetc's sigocn, dsigdt, dsigds: the hybrid vertical
 * coordinate restored column by column on the ocean cells.  targt, salmin,
 * dpmin (kdm) are hycom_constants' theta (target densities), salmin and
 * dpmin (m); batrop the barotropic step; latij (nip) latitude in degrees;
 * dens, temp, saln (kdm,nip); dp, uvel, vvel (kdm,nip,2), level leap
 * updated; pres (kdm+1,nip); passv (kdm,nip,numtr) moved when not NULL
 * (hycom_run.F90:516 passes none). */
void fo_hycom_hybgn1(int nip, int kdm, const int *wet, int leap, const double *targt, const double *salmin,
                     const double *dpmin, double batrop, const double *latij, double *dens, double *dp, double *pres,
                     double *temp, double *saln, double *uvel, double *vvel, int numtr, double *passv);
