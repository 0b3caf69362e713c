/* fim_oracle.h - the CPU checker's interface (test infrastructure only).
 * See fim_oracle.c for the array layouts. */
#pragma once
#include <stddef.h>

typedef struct {
  int nip, nvl, ntra;
  int pow_kind;              /* 0: C library pow (as the Fortran), 1: fim_pow (as the GPU) */
  int flux_scheme;           /* flux_conserv_schm: 1 cnuity (FCT), 2 mass_adv (flux limiter) */
  int ntrb;                  /* class-B tracers (transported every dtratio steps by fct3d) */
  int dtratio;               /* class-B / class-A time step ratio */
  double dt;
  const int *prox, *proxs, *nprox;
  const double *cs, *sn, *sidevec_c, *sidevec_e, *sideln, *hfcoef;
  const double *area, *rarea, *corio, *grndspd;
  const double *dfflen, *smoo_coeff, *thetac, *dpsig;
  /* land/sea mask (nip, 0 = land) or NULL: no counterpart in the reference's
   * atmosphere.  An edge carries mass and tracer flux only if both cells are
   * wet (the gating of hycom_cnuity.F90:100-125); momentum diffusion through a
   * land edge is zero; land columns keep thickness, winds and tracers and
   * get no hybgen (sdot = 0). */
  const int *wet;
} fo_grid;

typedef struct {
  double *us, *vs, *dp, *pr, *ex, *ph, *mp, *tr, *trdp;
  double *u_tdcy, *v_tdcy, *dp_tdcy, *dpl_tdcy, *trc_tdcy, *trl_tdcy, *massfx;
  double *ws, *relvor, *sdot, *psrf, *ptdcy, *psurf;
  double *u_edg, *v_edg, *dp_edg, *lp_edg, *mont_edg, *uvsq_edg, *trc_edg;
  double *work_e1, *work_e2;                               /* edge-sized scratch */
  double *work_c1, *work_c2, *work_c3, *work_c4, *work_c5; /* cell-sized scratch */
  double *work_l1, *work_l2;                               /* level-sized scratch */
  int nf, of, vof;
  double ab1, ab2, ab3;
  double *trb;      /* class-B tracers (nvl,nip,ntrb): tr3d(:,:,ntra+1:ntra+ntrb) */
  double *cumufx;   /* time-integrated mass flux (nvl,6,nip) */
  double *dpinit;   /* thickness at the start of the integration (nvl,nip) */
} fo_state;

void fo_vertical_params(int nvl, double dt, double *dfflen, double *smoo_coeff);
void fo_cell_params(int nip, const double *lat, const double *area, double *corio, double *rarea, double *grndspd);
void fo_abstart(fo_state *s, int its, double dt);
void fo_edgvar1(const fo_grid *g, fo_state *s);
void fo_edgvar3(const fo_grid *g, fo_state *s);
void fo_cnuity(const fo_grid *g, fo_state *s);
void fo_mass_adv(const fo_grid *g, fo_state *s);
void fo_trcadv(const fo_grid *g, fo_state *s);
void fo_momtum(const fo_grid *g, fo_state *s);
void fo_hybgen(const fo_grid *g, fo_state *s, int its);
void fo_hystat(const fo_grid *g, fo_state *s, int its);
void fo_transp0(const fo_grid *g, fo_state *s);
void fo_transp1(const fo_grid *g, fo_state *s);
void fo_transp2(const fo_grid *g, fo_state *s);
void fo_fct3d(const fo_grid *g, double *fields, int nfld, const double *u, const double *w,
              const double *thko, const double *thkn);
void fo_init(const fo_grid *g, fo_state *s, int init_hybgen);
void fo_step(const fo_grid *g, fo_state *s, int its, int mode);
void fo_globsum(const fo_grid *g, const fo_state *s, double *mass, double *trtot);
double fo_pow(double x, double y);
double fo_pow_fast(double x, double y);   /* the mixed-precision form's pow */
