/* fim_oracle.c - CPU restatement, in plain C and FP64, of the FIM dynamics
 * hot path.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP path is
 * compared with, and the "port" CPU baseline; the product never links it.
 *
 * Every routine follows the reference operation for operation, in the
 * reference loop order, so that with fp contraction off it reproduces the
 * reference Fortran built with -fdefault-real-8 (oracle/_ref/bin/dynref_r8)
 * bit for bit when pow_kind selects the C library pow (the one the Fortran
 * calls).  pow_kind = 1 selects fim_pow, the project pow that the HIP kernels
 * use, so that the two can be compared bit for bit as well.
 *
 * Layout (C index, level k fastest, cells ipn, 0-based):
 *   cell field  (nvl,nip)          a[ipn*nvl + k]
 *   level field (nvl+1,nip)        a[ipn*(nvl+1) + k]
 *   edge field  (nvl,6,nip)        a[(ipn*6 + e)*nvl + k]
 *   tracers     (nvl,nip,ntra)     a[(t*nip + ipn)*nvl + k]
 *   AB history  (nvl,nip,3)        a[(s*nip + ipn)*nvl + k]      s = slot-1
 *   tracer AB   (nvl,nip,3,ntra)   a[((t*3 + s)*nip + ipn)*nvl + k]
 *   massfx      (nvl,6,nip,3)      a[((s*nip + ipn)*6 + e)*nvl + k]
 * Grid tables: prox/proxs 0-based [ipn*6+e] (pentagon slot 6 = slot 5),
 * cs/sn [(ipn*6+e)*4+i], sidevec [(ipn*6+e)*2+i], hfcoef [(ipn*6+i)*6+j]
 * holding Fortran hfcoef(j,i,ipn).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fim_oracle.h"
#include "fim_pow.h"

#define NPP 6
#define KBL 2

/* module_constants.F90:14-24 */
static const double PI_T = 3.14159265;
static const double OMEGX2 = 1.4584e-4;
static const double EARTHRAD = 6371220.;
static const double GRVITY = 9.80665;
static const double P1000 = 100000.;
static const double CP = 1004.6855;
static const double RD = 287.0586;
static const double QVMIN = 1.e-10, QWMIN = 1.e-10;
/* fimnamelist.F90:100-117 */
static const double INTFC_SMOOTH = 30., PTOP = 30., THKTOP = 30.;
static const double RLEIGH = 2., RLTHRESH = 90., RLTAPER = 0.4, SLAK = 0.5;
static const double VELDFF_BKGND = 0.3, VELDFF_BOOST = 3.;

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

static inline double xpow(const fo_grid *g, double x, double y) {
  return g->pow_kind ? fim_pow(x, y) : pow(x, y);
}

#define C2(a, ipn, k) (a)[(size_t)(ipn) * nvl + (k)]
#define L2(a, ipn, k) (a)[(size_t)(ipn) * nvlp1 + (k)]
#define E3(a, ipn, e, k) (a)[((size_t)(ipn) * NPP + (e)) * nvl + (k)]
#define AB(a, s, ipn, k) (a)[((size_t)(s) * nip + (ipn)) * nvl + (k)]
#define TR(a, t, ipn, k) (a)[((size_t)(t) * nip + (ipn)) * nvl + (k)]
#define TAB(a, t, s, ipn, k) (a)[(((size_t)(t) * 3 + (s)) * nip + (ipn)) * nvl + (k)]
#define MF(a, s, ipn, e, k) (a)[(((size_t)(s) * nip + (ipn)) * NPP + (e)) * nvl + (k)]
#define PROX(ipn, e) g->prox[(ipn) * NPP + (e)]
#define PROXS(ipn, e) g->proxs[(ipn) * NPP + (e)]
#define CS(ipn, e, i) g->cs[((ipn) * NPP + (e)) * 4 + (i)]
#define SN(ipn, e, i) g->sn[((ipn) * NPP + (e)) * 4 + (i)]
#define SVC(ipn, e, i) g->sidevec_c[((ipn) * NPP + (e)) * 2 + (i)]
#define SVE(ipn, e, i) g->sidevec_e[((ipn) * NPP + (e)) * 2 + (i)]
#define HF(ipn, i, j) g->hfcoef[((ipn) * NPP + (i)) * NPP + (j)]

/* ------------------------------------------------------------------------
 * start-up derived arrays: dyn_init.F90:309-321 (dfflen, smoo_coeff) */
static double taper(int k1, int k2) {   /* dyn_init.F90:125 */
  double d = (double)(k2 - k1);
  return d * d / (9. + d * d);
}

void fo_vertical_params(int nvl, double dt, double *dfflen, double *smoo_coeff) {
  int ndamp = (int)(0.1 * nvl);
  for (int k = 1; k <= nvl; k++) {
    double wgt = dmax(0., (double)(k - nvl + ndamp) / (double)ndamp);
    dfflen[k - 1] = dt * (VELDFF_BKGND * (1. - wgt) + dmax(VELDFF_BKGND, VELDFF_BOOST) * wgt);
  }
  for (int k = 1; k <= nvl; k++) smoo_coeff[k - 1] = 0.;
  for (int k = KBL + 2; k <= nvl; k++)
    smoo_coeff[k - 1] = dt * INTFC_SMOOTH * taper(k, KBL + 1) * taper(k, nvl + 1);
}

/* dyn_init.F90:753-757 */
void fo_cell_params(int nip, const double *lat, const double *area,
                    double *corio, double *rarea, double *grndspd) {
  for (int i = 0; i < nip; i++) {
    corio[i] = OMEGX2 * sin(lat[i]);
    rarea[i] = 1. / area[i];
    grndspd[i] = EARTHRAD * 2. * PI_T * cos(lat[i]) / 86164.;
  }
}

/* abstart.F90:14-68 */
void fo_abstart(fo_state *s, int its, double dt) {
  s->nf = s->nf + 1; if (s->nf == 4) s->nf = 1;
  s->of = s->of + 1; if (s->of == 4) s->of = 1;
  s->vof = s->vof + 1; if (s->vof == 4) s->vof = 1;
  if (its == 1) { s->ab1 = dt; s->ab2 = 0.; s->ab3 = 0.; }
  if (its == 2) { s->ab1 = dt * 1.5; s->ab2 = -dt * .5; s->ab3 = 0.; }
  if (its == 3) { s->ab1 = dt * 23. / 12.; s->ab2 = -dt * 16. / 12.; s->ab3 = dt * 5. / 12.; }
}

/* the land/sea mask (NULL: every cell wet) */
static inline int fo_wet(const fo_grid *g, int i) { return !g->wet || g->wet[i]; }
static inline int fo_open(const fo_grid *g, int i, int x) { return fo_wet(g, i) && fo_wet(g, x); }

/* ------------------------------------------------------------------------
 * edgvar1.F90:160-421 (eqwgt branch; janjic = 0; hiOrderFluxComp: no dp_edg)
 * and the tracer loop 425-511 for tracer 1 only */
void fo_edgvar1(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl, nvlp1 = nvl + 1;
  const double divby6 = 1. / 6., eps = 0.001;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    int np = g->nprox[ipn];
    for (int edg = 0; edg < np; edg++) {
      int i1 = ipn, i2 = PROX(ipn, edg);
      int i3t = PROX(ipn, (edg - 1 + np) % np);
      int i3l = PROX(ipn, (edg + 1) % np);
      for (int k = 0; k < nvl; k++) {
        double u3trl = CS(ipn, edg, 2) * C2(s->us, i3t, k) + SN(ipn, edg, 2) * C2(s->vs, i3t, k);
        double v3trl = -SN(ipn, edg, 2) * C2(s->us, i3t, k) + CS(ipn, edg, 2) * C2(s->vs, i3t, k);
        double u1ldg = CS(ipn, edg, 0) * C2(s->us, i1, k) + SN(ipn, edg, 0) * C2(s->vs, i1, k);
        double v1ldg = -SN(ipn, edg, 0) * C2(s->us, i1, k) + CS(ipn, edg, 0) * C2(s->vs, i1, k);
        double u2ldg = CS(ipn, edg, 1) * C2(s->us, i2, k) + SN(ipn, edg, 1) * C2(s->vs, i2, k);
        double v2ldg = -SN(ipn, edg, 1) * C2(s->us, i2, k) + CS(ipn, edg, 1) * C2(s->vs, i2, k);
        double u3ldg = CS(ipn, edg, 3) * C2(s->us, i3l, k) + SN(ipn, edg, 3) * C2(s->vs, i3l, k);
        double v3ldg = -SN(ipn, edg, 3) * C2(s->us, i3l, k) + CS(ipn, edg, 3) * C2(s->vs, i3l, k);
        double ue = (2. * (u1ldg + u2ldg) + u3trl + u3ldg) * divby6;
        double ve = (2. * (v1ldg + v2ldg) + v3trl + v3ldg) * divby6;
        double m_1ldg = sqrt(u1ldg * u1ldg + v1ldg * v1ldg);
        double m_2ldg = sqrt(u2ldg * u2ldg + v2ldg * v2ldg);
        double m_3trl = sqrt(u3trl * u3trl + v3trl * v3trl);
        double m_3ldg = sqrt(u3ldg * u3ldg + v3ldg * v3ldg);
        double m_1 = sqrt(ue * ue + ve * ve);
        double m_2 = (2. * (m_1ldg + m_2ldg) + m_3trl + m_3ldg) * divby6;
        ue = ue * m_2 / dmax(eps, m_1);
        ve = ve * m_2 / dmax(eps, m_1);
        E3(s->u_edg, ipn, edg, k) = ue;
        E3(s->v_edg, ipn, edg, k) = ve;
        E3(s->lp_edg, ipn, edg, k) =
            (L2(s->pr, i1, k) + L2(s->pr, i1, k + 1) + L2(s->pr, i2, k) + L2(s->pr, i2, k + 1) +
             .5 * (L2(s->pr, i3t, k) + L2(s->pr, i3t, k + 1) + L2(s->pr, i3l, k) + L2(s->pr, i3l, k + 1))) * divby6;
        E3(s->mont_edg, ipn, edg, k) =
            (2. * (C2(s->mp, i1, k) + C2(s->mp, i2, k)) + C2(s->mp, i3t, k) + C2(s->mp, i3l, k)) * divby6;
        E3(s->uvsq_edg, ipn, edg, k) = .5 * (ue * ue + ve * ve);
      }
    }
  }
  /* tracer 1 (theta) */
  double *trc = s->trc_edg;  /* tracer 0 block */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    int np = g->nprox[ipn];
    for (int k = 0; k < nvl; k++) E3(trc, ipn, NPP - 1, k) = 0.;
    for (int edg = 0; edg < np; edg++) {
      int i1 = ipn, i2 = PROX(ipn, edg);
      int i3t = PROX(ipn, (edg - 1 + np) % np);
      int i3l = PROX(ipn, (edg + 1) % np);
      for (int k = 0; k < nvl; k++)
        E3(trc, ipn, edg, k) = (2. * (TR(s->tr, 0, i1, k) + TR(s->tr, 0, i2, k)) + TR(s->tr, 0, i3t, k) +
                                TR(s->tr, 0, i3l, k)) * divby6;
    }
  }
}

/* edgvar3.F90:65-163: LS edge reconstruction of dp and tracers 2..ntra, then
 * the average of both sides of each edge */
static void edgval_eval(const fo_grid *g, int ipn, const double *fld, double *edgval) {
  const int nvl = g->nvl;
  int n = g->nprox[ipn];
  for (int i = n; i < NPP; i++)
    for (int k = 0; k < nvl; k++) edgval[i * nvl + k] = 0.;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < nvl; k++) edgval[i * nvl + k] = fld[(size_t)ipn * nvl + k];
    for (int j = 0; j < n; j++) {
      int nb = PROX(ipn, j);
      double w = HF(ipn, i, j);
      for (int k = 0; k < nvl; k++)
        edgval[i * nvl + k] = edgval[i * nvl + k] + (fld[(size_t)nb * nvl + k] - fld[(size_t)ipn * nvl + k]) * w;
    }
  }
}

void fo_edgvar3(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl, ntra = g->ntra;
  const size_t esz = (size_t)nip * NPP * nvl;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    edgval_eval(g, ipn, s->dp, &s->dp_edg[(size_t)ipn * NPP * nvl]);
    for (int t = 1; t < ntra; t++)
      edgval_eval(g, ipn, &s->tr[(size_t)t * nip * nvl], &s->trc_edg[t * esz + (size_t)ipn * NPP * nvl]);
  }
  /* loop 2: serial, as each edge pair is written once (ipn < nb) */
  for (int ipn = 0; ipn < nip; ipn++) {
    int np = g->nprox[ipn];
    for (int edg = 0; edg < np; edg++) {
      int nb = PROX(ipn, edg), eo = PROXS(ipn, edg);
      if (ipn < nb) {
        for (int k = 0; k < nvl; k++) {
          double w = 0.5 * (E3(s->dp_edg, ipn, edg, k) + E3(s->dp_edg, nb, eo, k));
          E3(s->dp_edg, ipn, edg, k) = w;
          E3(s->dp_edg, nb, eo, k) = w;
        }
        for (int t = 1; t < ntra; t++) {
          double *tb = &s->trc_edg[t * esz];
          for (int k = 0; k < nvl; k++) {
            double w = 0.5 * (E3(tb, ipn, edg, k) + E3(tb, nb, eo, k));
            E3(tb, ipn, edg, k) = w;
            E3(tb, nb, eo, k) = w;
          }
        }
      }
    }
  }
}

static void update_pres(const fo_grid *g, fo_state *s);

/* ------------------------------------------------------------------------
 * cnuity.F90:23-249 with cnuity_compute.F90:17-451 */
void fo_cnuity(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl, nvlp1 = nvl + 1;
  const int nf = s->nf - 1, of = s->of - 1, vof = s->vof - 1;
  const double ab1 = s->ab1, ab2 = s->ab2, ab3 = s->ab3;
  const double thshld = 1.e-11;
  double *vnorm = s->work_e1, *antifx = s->work_e2;
  double *delp_lo = s->work_c1, *r_plus = s->work_c2, *r_mnus = s->work_c3;
  double *massfx = s->massfx;
  /* cnuity1 */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++)
    for (int edg = 0; edg < g->nprox[ipn]; edg++)
      for (int k = 0; k < nvl; k++)
        E3(vnorm, ipn, edg, k) = !fo_open(g, ipn, PROX(ipn, edg)) ? 0. :
            SVE(ipn, edg, 1) * E3(s->u_edg, ipn, edg, k) - SVE(ipn, edg, 0) * E3(s->v_edg, ipn, edg, k);
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    for (int k = 0; k < nvl; k++) { E3(antifx, ipn, NPP - 1, k) = 0.; MF(massfx, nf, ipn, NPP - 1, k) = 0.; }
    for (int edg = 0; edg < g->nprox[ipn]; edg++) {
      int ipx = PROX(ipn, edg), edx = PROXS(ipn, edg);
      for (int k = 0; k < nvl; k++) {
        double vn = E3(vnorm, ipn, edg, k), vx = E3(vnorm, ipx, edx, k);
        double flxhi = vn * E3(s->dp_edg, ipn, edg, k);
        double m = 0.5 * ((vn + fabs(vn)) * C2(s->dp, ipn, k) - (vx + fabs(vx)) * C2(s->dp, ipx, k));
        MF(massfx, nf, ipn, edg, k) = m;
        E3(antifx, ipn, edg, k) = flxhi - m;
      }
    }
  }
  /* cnuity2 */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    for (int k = 0; k < nvl; k++) AB(s->dpl_tdcy, nf, ipn, k) = 0.;
    for (int edg = 0; edg < g->nprox[ipn]; edg++)
      for (int k = 0; k < nvl; k++) AB(s->dpl_tdcy, nf, ipn, k) = AB(s->dpl_tdcy, nf, ipn, k) + MF(massfx, nf, ipn, edg, k);
    for (int k = 0; k < nvl; k++) {
      AB(s->dpl_tdcy, nf, ipn, k) = -AB(s->dpl_tdcy, nf, ipn, k) * g->rarea[ipn];
      C2(delp_lo, ipn, k) = C2(s->dp, ipn, k) + ab1 * AB(s->dpl_tdcy, nf, ipn, k) +
                            ab2 * AB(s->dpl_tdcy, of, ipn, k) + ab3 * AB(s->dpl_tdcy, vof, ipn, k);
    }
  }
  /* cnuity3 */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    for (int k = 0; k < nvl; k++) {
      double pp = 0., pm = 0.;
      for (int edg = 0; edg < g->nprox[ipn]; edg++) {
        double a = E3(antifx, ipn, edg, k);
        if (a <= 0.) pp = pp - a; else pm = pm + a;
      }
      double dpmax = dmax(C2(delp_lo, ipn, k), C2(s->dp, ipn, k));
      double dpmin = dmin(C2(delp_lo, ipn, k), C2(s->dp, ipn, k));
      for (int e = 0; e < NPP; e++) {
        int x = PROX(ipn, e);
        dpmax = dmax(dpmax, dmax(C2(delp_lo, x, k), C2(s->dp, x, k)));
        dpmin = dmin(dpmin, dmin(C2(delp_lo, x, k), C2(s->dp, x, k)));
      }
      dpmax = dmax(0., dpmax);
      dpmin = dmax(0., dpmin);
      double q_plus = (dpmax - C2(delp_lo, ipn, k) - (ab2 * dmin(0., AB(s->dp_tdcy, of, ipn, k)) +
                       ab3 * dmax(0., AB(s->dp_tdcy, vof, ipn, k)))) / ab1 * g->area[ipn];
      double q_mnus = (C2(delp_lo, ipn, k) - dpmin + (ab2 * dmax(0., AB(s->dp_tdcy, of, ipn, k)) +
                       ab3 * dmin(0., AB(s->dp_tdcy, vof, ipn, k)))) / ab1 * g->area[ipn];
      C2(r_plus, ipn, k) = dmax(0., dmin(1., q_plus / dmax(pp, thshld)));
      C2(r_mnus, ipn, k) = dmax(0., dmin(1., q_mnus / dmax(pm, thshld)));
    }
  }
  /* cnuity4 */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    double ps = 0.;
    for (int k = nvl - 1; k >= 0; k--) ps = ps + C2(s->dp, ipn, k);
    s->psurf[ipn] = ps;
    for (int edg = 0; edg < g->nprox[ipn]; edg++) {
      int x = PROX(ipn, edg);
      for (int k = 0; k < nvl; k++) {
        double a = E3(antifx, ipn, edg, k), clip;
        if (a >= 0.) clip = dmin(C2(r_mnus, ipn, k), C2(r_plus, x, k));
        else clip = dmin(C2(r_plus, ipn, k), C2(r_mnus, x, k));
        a = a * clip;
        E3(antifx, ipn, edg, k) = a;
        MF(massfx, nf, ipn, edg, k) = MF(massfx, nf, ipn, edg, k) + a;
      }
    }
  }
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    for (int k = 0; k < nvl; k++) {
      double at = 0.;
      for (int edg = 0; edg < g->nprox[ipn]; edg++) at = at + E3(antifx, ipn, edg, k);
      at = -at * g->rarea[ipn];
      AB(s->dp_tdcy, nf, ipn, k) = AB(s->dpl_tdcy, nf, ipn, k) + at;
      C2(s->dp, ipn, k) = C2(s->dp, ipn, k) + ab1 * AB(s->dp_tdcy, nf, ipn, k) +
                          ab2 * AB(s->dp_tdcy, of, ipn, k) + ab3 * AB(s->dp_tdcy, vof, ipn, k);
    }
  }
  update_pres(g, s);
}

/* cnuity5 (cnuity_compute.F90:376-451), the same as update_pres of mass_adv
 * (massadv_compute.F90:73-171): top-down pressure, Exner function and omega */
static void update_pres(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl, nvlp1 = nvl + 1;
  const int nf = s->nf - 1, of = s->of - 1, vof = s->vof - 1;
  const double ab1 = s->ab1, ab2 = s->ab2, ab3 = s->ab3;
  const double rdcp = RD / CP;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    double coltend = 0.;
    for (int k = nvl - 1; k >= 0; k--) {
      L2(s->pr, ipn, k) = L2(s->pr, ipn, k + 1) + C2(s->dp, ipn, k);
      L2(s->ex, ipn, k) = CP * xpow(g, L2(s->pr, ipn, k) / P1000, rdcp);
      double lyrtend = (ab1 * AB(s->dp_tdcy, nf, ipn, k) + ab2 * AB(s->dp_tdcy, of, ipn, k) +
                        ab3 * AB(s->dp_tdcy, vof, ipn, k)) / g->dt;
      double om = coltend + 0.5 * lyrtend;
      coltend = coltend + lyrtend;
      double dpdx = 0., dpdy = 0.;
      for (int edg = 0; edg < g->nprox[ipn]; edg++) {
        dpdx = dpdx + E3(s->lp_edg, ipn, edg, k) * SVC(ipn, edg, 1);
        dpdy = dpdy - E3(s->lp_edg, ipn, edg, k) * SVC(ipn, edg, 0);
      }
      C2(s->ws, ipn, k) = om + (C2(s->us, ipn, k) * dpdx + C2(s->vs, ipn, k) * dpdy) * g->rarea[ipn];
    }
  }
}

/* ------------------------------------------------------------------------
 * mass_adv (flux_conserv_schm = 2, dyn_run.F90:377-386): massadv.F90:21-160
 * with flux_limiter.F90:35-262 (init_mass_limiter, edgeNormVel, limiter
 * with M1 defined) and massadv_compute.F90:9-171 (massflux_newdepth,
 * update_pres).  The limited edge depth replaces dp_edg, as edgeVal does. */
void fo_mass_adv(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl;
  const int nf = s->nf - 1, of = s->of - 1, vof = s->vof - 1;
  const double ab1 = s->ab1, ab2 = s->ab2, ab3 = s->ab3, dt = g->dt;
  const double max_out_bound = 100000.0;
  double *vnorm = s->work_e1, *c_num = s->work_e2;
  double *q_min_out = s->work_c1, *q_max_out = s->work_c2;
  double *massfx = s->massfx, *dp = s->dp, *dpe = s->dp_edg;
  /* edgeNormVel (flux_limiter.F90:78-103) */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++)
    for (int edg = 0; edg < g->nprox[ipn]; edg++)
      for (int k = 0; k < nvl; k++) {
        const double vn = !fo_open(g, ipn, PROX(ipn, edg)) ? 0. :
            SVE(ipn, edg, 1) * E3(s->u_edg, ipn, edg, k) - SVE(ipn, edg, 0) * E3(s->v_edg, ipn, edg, k);
        E3(vnorm, ipn, edg, k) = vn;
        E3(c_num, ipn, edg, k) = fabs(vn) * dt * g->rarea[ipn];
      }
  /* limiter (flux_limiter.F90:132-260): the edge value clamped to the four
   * cells around the edge (q_star, kept in dp_edg), then the in/out bounds */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    const int np = g->nprox[ipn];
    for (int edg = 0; edg < np; edg++) {
      const int nbm1 = PROX(ipn, (edg + np - 1) % np), nbp1 = PROX(ipn, (edg + 1) % np), ipx = PROX(ipn, edg);
      for (int k = 0; k < nvl; k++) {
        const double a = C2(dp, ipn, k), b = C2(dp, ipx, k), c = C2(dp, nbm1, k), d = C2(dp, nbp1, k);
        const double q_min = dmin(dmin(a, b), dmin(c, d)), q_max = dmax(dmax(a, b), dmax(c, d));
        E3(dpe, ipn, edg, k) = dmax(q_min, dmin(q_max, E3(dpe, ipn, edg, k)));
      }
    }
  }
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    const int np = g->nprox[ipn];
    for (int k = 0; k < nvl; k++) {
      double q_min_in_min = 1.7976931348623157e308, q_max_in_max = 0.;   /* huge(), 0 */
      double cq_min_in_sum = 0., cq_max_in_sum = 0., c_num_in_sum = 0., c_num_out_sum = 0.;
      const double gv = C2(dp, ipn, k);
      for (int edg = 0; edg < np; edg++) {
        const int ipx = PROX(ipn, edg);
        const double qs = E3(dpe, ipn, edg, k), cn = E3(c_num, ipn, edg, k);
        if (E3(vnorm, ipn, edg, k) > 0.) {
          c_num_out_sum = c_num_out_sum + cn;
        } else {
          const double q_min = dmin(qs, C2(dp, ipx, k)), q_max = dmax(qs, C2(dp, ipx, k));
          c_num_in_sum = c_num_in_sum + cn;
          q_min_in_min = dmin(q_min_in_min, q_min);
          q_max_in_max = dmax(q_max_in_max, q_max);
          cq_min_in_sum = cq_min_in_sum + cn * q_min;
          cq_max_in_sum = cq_max_in_sum + cn * q_max;
        }
      }
      const double q_min_next = dmin(gv, q_min_in_min), q_max_next = dmax(gv, q_max_in_max);
      if (c_num_out_sum != 0.) {
        const double at = ab2 * AB(s->dp_tdcy, of, ipn, k), av = ab3 * AB(s->dp_tdcy, vof, ipn, k);
        C2(q_min_out, ipn, k) = (gv + ab1 / dt * cq_max_in_sum - q_max_next * (1.0 + c_num_in_sum - c_num_out_sum) +
                                 at + av) / c_num_out_sum / ab1 * dt;
        C2(q_max_out, ipn, k) = (gv + ab1 / dt * cq_min_in_sum - q_min_next * (1.0 + c_num_in_sum - c_num_out_sum) +
                                 at + av) / c_num_out_sum / ab1 * dt;
      } else {
        C2(q_min_out, ipn, k) = -max_out_bound;
        C2(q_max_out, ipn, k) = max_out_bound;
      }
    }
  }
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++)
    for (int edg = 0; edg < g->nprox[ipn]; edg++) {
      const int ipx = PROX(ipn, edg);
      for (int k = 0; k < nvl; k++) {
        const int up = E3(vnorm, ipn, edg, k) > 0. ? ipn : ipx;
        E3(dpe, ipn, edg, k) = dmax(C2(q_min_out, up, k), dmin(C2(q_max_out, up, k), E3(dpe, ipn, edg, k)));
      }
    }
  /* massflux_newdepth (massadv_compute.F90:9-70) */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    double ps = 0.;
    for (int k = nvl - 1; k >= 0; k--) ps = ps + C2(dp, ipn, k);
    s->psurf[ipn] = ps;
    for (int k = 0; k < nvl; k++) MF(massfx, nf, ipn, NPP - 1, k) = 0.;
    for (int edg = 0; edg < g->nprox[ipn]; edg++)
      for (int k = 0; k < nvl; k++) MF(massfx, nf, ipn, edg, k) = E3(vnorm, ipn, edg, k) * E3(dpe, ipn, edg, k);
  }
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    for (int k = 0; k < nvl; k++) AB(s->dp_tdcy, nf, ipn, k) = 0.;
    for (int edg = 0; edg < g->nprox[ipn]; edg++)
      for (int k = 0; k < nvl; k++) AB(s->dp_tdcy, nf, ipn, k) = AB(s->dp_tdcy, nf, ipn, k) + MF(massfx, nf, ipn, edg, k);
    for (int k = 0; k < nvl; k++) AB(s->dp_tdcy, nf, ipn, k) = -AB(s->dp_tdcy, nf, ipn, k) * g->rarea[ipn];
    for (int k = 0; k < nvl; k++) {
      double d = C2(dp, ipn, k) + ab1 * AB(s->dp_tdcy, nf, ipn, k) + ab2 * AB(s->dp_tdcy, of, ipn, k) +
                 ab3 * AB(s->dp_tdcy, vof, ipn, k);
      C2(dp, ipn, k) = dmax(d, 0.0);
    }
  }
  update_pres(g, s);
}

/* ------------------------------------------------------------------------
 * trcadv.F90:19-214 with trcadv_compute.F90:19-315 */
void fo_trcadv(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl, ntra = g->ntra;
  const int nf = s->nf - 1, of = s->of - 1, vof = s->vof - 1;
  const double ab1 = s->ab1, ab2 = s->ab2, ab3 = s->ab3;
  const double thshld = 1.e-11;
  const size_t esz = (size_t)nip * NPP * nvl;
  double *flx_lo = s->work_e1, *antiflx = s->work_e2;
  double *trcr_lo = s->work_c1, *r_plus = s->work_c2, *r_mnus = s->work_c3;
  double *trmax = s->work_c4, *trmin = s->work_c5;
  const double *massfx = s->massfx;
  for (int t = 0; t < ntra; t++) {
    const double *trc_edg = &s->trc_edg[t * esz];
    double *tracr = &s->tr[(size_t)t * nip * nvl];
    double *trcdp = &s->trdp[(size_t)t * nip * nvl];
    /* trcadv1 */
    #pragma omp parallel for schedule(static)
    for (int ipn = 0; ipn < nip; ipn++) {
      for (int k = 0; k < nvl; k++) {
        TAB(s->trl_tdcy, t, nf, ipn, k) = 0.;
        TAB(s->trc_tdcy, t, nf, ipn, k) = 0.;
        E3(antiflx, ipn, NPP - 1, k) = 0.;
      }
      for (int edg = 0; edg < g->nprox[ipn]; edg++) {
        int ipx = PROX(ipn, edg), edx = PROXS(ipn, edg);
        for (int k = 0; k < nvl; k++) {
          double m = MF(massfx, nf, ipn, edg, k), mx = MF(massfx, nf, ipx, edx, k);
          double flx_hi = m * E3(trc_edg, ipn, edg, k);
          double fl = 0.5 * ((m + fabs(m)) * C2(tracr, ipn, k) - (mx + fabs(mx)) * C2(tracr, ipx, k));
          E3(flx_lo, ipn, edg, k) = fl;
          E3(antiflx, ipn, edg, k) = flx_hi - fl;
        }
      }
      for (int edg = 0; edg < g->nprox[ipn]; edg++)
        for (int k = 0; k < nvl; k++)
          TAB(s->trl_tdcy, t, nf, ipn, k) = TAB(s->trl_tdcy, t, nf, ipn, k) + E3(flx_lo, ipn, edg, k);
      for (int k = 0; k < nvl; k++) {
        TAB(s->trl_tdcy, t, nf, ipn, k) = -TAB(s->trl_tdcy, t, nf, ipn, k) * g->rarea[ipn];
        double trcdp_lo = C2(trcdp, ipn, k) + ab1 * TAB(s->trl_tdcy, t, nf, ipn, k) +
                          ab2 * TAB(s->trl_tdcy, t, of, ipn, k) + ab3 * TAB(s->trl_tdcy, t, vof, ipn, k);
        double mx = C2(tracr, ipn, k), mn = C2(tracr, ipn, k);
        for (int e = 0; e < NPP; e++) {
          mx = dmax(mx, C2(tracr, PROX(ipn, e), k));
          mn = dmin(mn, C2(tracr, PROX(ipn, e), k));
        }
        mx = dmax(0., mx);
        mn = dmax(0., mn);
        C2(trmax, ipn, k) = mx;
        C2(trmin, ipn, k) = mn;
        C2(trcr_lo, ipn, k) = dmax(mn, dmin(mx, trcdp_lo / dmax(thshld, C2(s->dp, ipn, k))));
      }
    }
    /* trcadv2 */
    #pragma omp parallel for schedule(static)
    for (int ipn = 0; ipn < nip; ipn++) {
      for (int k = 0; k < nvl; k++) {
        double sp = 0., sm = 0.;
        for (int edg = 0; edg < g->nprox[ipn]; edg++) {
          sp = sp - dmin(0., E3(antiflx, ipn, edg, k));
          sm = sm + dmax(0., E3(antiflx, ipn, edg, k));
        }
        double mx = dmax(C2(trmax, ipn, k), C2(trcr_lo, ipn, k));
        double mn = dmin(C2(trmin, ipn, k), C2(trcr_lo, ipn, k));
        for (int e = 0; e < NPP; e++) {
          mx = dmax(mx, C2(trcr_lo, PROX(ipn, e), k));
          mn = dmin(mn, C2(trcr_lo, PROX(ipn, e), k));
        }
        mx = dmax(0., mx);
        mn = dmax(0., mn);
        C2(trmax, ipn, k) = mx;
        C2(trmin, ipn, k) = mn;
        double lo = C2(trcr_lo, ipn, k), dp = C2(s->dp, ipn, k);
        double q_plus = ((mx - lo) * dp - (ab2 * dmin(0., TAB(s->trc_tdcy, t, of, ipn, k)) +
                         ab3 * dmax(0., TAB(s->trc_tdcy, t, vof, ipn, k)))) / ab1 * g->area[ipn];
        double q_mnus = ((lo - mn) * dp + (ab2 * dmax(0., TAB(s->trc_tdcy, t, of, ipn, k)) +
                         ab3 * dmin(0., TAB(s->trc_tdcy, t, vof, ipn, k)))) / ab1 * g->area[ipn];
        C2(r_plus, ipn, k) = dmax(0., dmin(1., q_plus / dmax(sp, thshld)));
        C2(r_mnus, ipn, k) = dmax(0., dmin(1., q_mnus / dmax(sm, thshld)));
      }
    }
    /* trcadv3 */
    #pragma omp parallel for schedule(static)
    for (int ipn = 0; ipn < nip; ipn++) {
      for (int k = 0; k < nvl; k++) {
        double at = 0.;
        for (int edg = 0; edg < g->nprox[ipn]; edg++) {
          int x = PROX(ipn, edg);
          double a = E3(antiflx, ipn, edg, k);
          if (a >= 0.) a = a * dmin(C2(r_mnus, ipn, k), C2(r_plus, x, k));
          else a = a * dmin(C2(r_plus, ipn, k), C2(r_mnus, x, k));
          E3(antiflx, ipn, edg, k) = a;
          at = at + a;
        }
        at = -at * g->rarea[ipn];
        TAB(s->trc_tdcy, t, nf, ipn, k) = TAB(s->trl_tdcy, t, nf, ipn, k) + at;
        C2(trcdp, ipn, k) = C2(trcdp, ipn, k) + ab1 * TAB(s->trc_tdcy, t, nf, ipn, k) +
                            ab2 * TAB(s->trc_tdcy, t, of, ipn, k) + ab3 * TAB(s->trc_tdcy, t, vof, ipn, k);
        if (fo_wet(g, ipn))      /* a land column keeps its tracers */
          C2(tracr, ipn, k) = dmax(C2(trmin, ipn, k), dmin(C2(trmax, ipn, k),
                                   C2(trcdp, ipn, k) / dmax(thshld, C2(s->dp, ipn, k))));
      }
    }
  }
}

/* ------------------------------------------------------------------------
 * momtum.F90:16-158 with momtum_compute.F90:102-450 (dissip, momtum1) */
void fo_momtum(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl, nvlp1 = nvl + 1;
  const int nf = s->nf - 1, of = s->of - 1, vof = s->vof - 1;
  const double ab1 = s->ab1, ab2 = s->ab2, ab3 = s->ab3;
  const double thshld = 1.e-11;
  double *worku = s->work_c1, *workv = s->work_c2;
  double *ufx = s->work_e1, *vfx = s->work_e2;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++)
    for (int k = 0; k < nvl; k++) {
      C2(worku, ipn, k) = C2(s->us, ipn, k) + g->grndspd[ipn];
      C2(workv, ipn, k) = C2(s->vs, ipn, k);
    }
  /* dissip loop 1 */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++)
    for (int edg = 0; edg < g->nprox[ipn]; edg++) {
      int ipx = PROX(ipn, edg);
      for (int k = 0; k < nvl; k++) {
        double uxy1 = CS(ipn, edg, 0) * C2(worku, ipn, k) + SN(ipn, edg, 0) * C2(workv, ipn, k);
        double vxy1 = -SN(ipn, edg, 0) * C2(worku, ipn, k) + CS(ipn, edg, 0) * C2(workv, ipn, k);
        double uxy2 = CS(ipn, edg, 1) * C2(worku, ipx, k) + SN(ipn, edg, 1) * C2(workv, ipx, k);
        double vxy2 = -SN(ipn, edg, 1) * C2(worku, ipx, k) + CS(ipn, edg, 1) * C2(workv, ipx, k);
        double a = dmax(C2(s->dp, ipn, k), thshld), b = dmax(C2(s->dp, ipx, k), thshld);
        double hh = a * b / (a + b);
        const int open = fo_open(g, ipn, ipx);     /* no diffusion through a land edge */
        E3(ufx, ipn, edg, k) = open ? (uxy2 - uxy1) * g->dfflen[k] * g->sideln[ipn * NPP + edg] * 2. * hh : 0.;
        E3(vfx, ipn, edg, k) = open ? (vxy2 - vxy1) * g->dfflen[k] * g->sideln[ipn * NPP + edg] * 2. * hh : 0.;
      }
    }
  /* dissip loop 2, then the increment */
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    for (int k = 0; k < nvl; k++) {
      double wu = 0., wv = 0.;
      double factor = g->rarea[ipn] / dmax(C2(s->dp, ipn, k), thshld);
      for (int edg = 0; edg < g->nprox[ipn]; edg++) {
        double uflux = CS(ipn, edg, 0) * E3(ufx, ipn, edg, k) - SN(ipn, edg, 0) * E3(vfx, ipn, edg, k);
        double vflux = SN(ipn, edg, 0) * E3(ufx, ipn, edg, k) + CS(ipn, edg, 0) * E3(vfx, ipn, edg, k);
        wu = wu + uflux * factor;
        wv = wv + vflux * factor;
      }
      if (!fo_wet(g, ipn)) continue;      /* a land column keeps its winds */
      C2(s->us, ipn, k) = C2(s->us, ipn, k) + wu;
      C2(s->vs, ipn, k) = C2(s->vs, ipn, k) + wv;
    }
  }
  /* momtum1 */
  const int ndamp = (int)(0.25 * nvl);
  const double *trc1 = s->trc_edg;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    if (!fo_wet(g, ipn)) {               /* a land column keeps its winds */
      for (int k = 0; k < nvl; k++) {
        C2(s->relvor, ipn, k) = 0.;
        AB(s->u_tdcy, nf, ipn, k) = 0.;
        AB(s->v_tdcy, nf, ipn, k) = 0.;
      }
      continue;
    }
    for (int k = 0; k < nvl; k++) {
      double px = 0., py = 0., rv = 0.;
      double exm = .5 * (L2(s->ex, ipn, k) + L2(s->ex, ipn, k + 1));
      for (int edg = 0; edg < g->nprox[ipn]; edg++) {
        double b = E3(s->mont_edg, ipn, edg, k) + E3(s->uvsq_edg, ipn, edg, k);
        px = px - b * SVC(ipn, edg, 1) + exm * E3(trc1, ipn, edg, k) * SVC(ipn, edg, 1);
        py = py + b * SVC(ipn, edg, 0) - exm * E3(trc1, ipn, edg, k) * SVC(ipn, edg, 0);
        rv = rv + (SVE(ipn, edg, 0) * E3(s->u_edg, ipn, edg, k) + SVE(ipn, edg, 1) * E3(s->v_edg, ipn, edg, k));
      }
      px = px * g->rarea[ipn];
      py = py * g->rarea[ipn];
      rv = rv * g->rarea[ipn];
      C2(s->relvor, ipn, k) = rv;
      double uzeta = (g->corio[ipn] + rv) * C2(s->us, ipn, k);
      double vzeta = (g->corio[ipn] + rv) * C2(s->vs, ipn, k);
      AB(s->u_tdcy, nf, ipn, k) = px + vzeta;
      AB(s->v_tdcy, nf, ipn, k) = py - uzeta;
      C2(s->us, ipn, k) = C2(s->us, ipn, k) + ab1 * AB(s->u_tdcy, nf, ipn, k) + ab2 * AB(s->u_tdcy, of, ipn, k) +
                          ab3 * AB(s->u_tdcy, vof, ipn, k);
      C2(s->vs, ipn, k) = C2(s->vs, ipn, k) + ab1 * AB(s->v_tdcy, nf, ipn, k) + ab2 * AB(s->v_tdcy, of, ipn, k) +
                          ab3 * AB(s->v_tdcy, vof, ipn, k);
    }
    /* Rayleigh damping near the model top (momtum_compute.F90:406-416) */
    double u0 = C2(s->us, ipn, nvl - 1), v0 = C2(s->vs, ipn, nvl - 1);
    double u1 = C2(s->us, ipn, nvl - 2), v1 = C2(s->vs, ipn, nvl - 2);
    double u2 = C2(s->us, ipn, nvl - 3), v2 = C2(s->vs, ipn, nvl - 3);
    double qq = dmax(dmax(u0 * u0 + v0 * v0, u1 * u1 + v1 * v1), u2 * u2 + v2 * v2) / (RLTHRESH * RLTHRESH);
    if (qq > 1.) {
      qq = sqrt((qq - 1.) * qq);
      for (int k1 = nvl - ndamp; k1 <= nvl; k1++) {
        double factor = 1. - (g->dt / 86400.) * RLEIGH * qq * xpow(g, 2., RLTAPER * (double)(k1 - nvl));
        C2(s->us, ipn, k1 - 1) = C2(s->us, ipn, k1 - 1) * factor;
        C2(s->vs, ipn, k1 - 1) = C2(s->vs, ipn, k1 - 1) * factor;
      }
    }
  }
}

/* ------------------------------------------------------------------------
 * hybgen.F90: regrid_1d (207-512), restp_1d (792-890), inflate (893-1007),
 * equilb (1282-1437), remap_1d (515-789), pcwise (1010-1279) and the driver
 * (20-201) with dffusn_lev (dffusn.F90:37-58, dffusn_lev.exe). */
static const double THIN = 287.0586 / 100000.;   /* rd/p1000 */

static void restp_1d(const double *thold, const double *pkold, int kold, double *thnew,
                     double *pknew, const double *targt, int knew, double *oldth) {
  oldth[kold - 1] = thold[kold - 1];
  for (int k = kold; k >= 2; k--) oldth[k - 2] = dmin(oldth[k - 1], thold[k - 2]);
  for (int k = 0; k < knew; k++) thnew[k] = targt[k];
  thnew[0] = dmin(dmin(oldth[0], oldth[kold - 1]), targt[0]);
  thnew[knew - 1] = dmax(dmax(oldth[0], oldth[kold - 1]), targt[knew - 1]);
  pknew[0] = pkold[0];
  pknew[knew] = pkold[kold];
  int ko = 1;
  for (int k = 1; k <= knew - 1; k++) {
    double pinteg = 0., tha, thb = thnew[k - 1];
    for (;;) {
      tha = thb;
      thb = dmin(thnew[k], dmax(thnew[k - 1], oldth[ko - 1]));
      pinteg = pinteg + pkold[ko - 1] * (thb - tha);
      if (oldth[ko - 1] < thnew[k]) {
        if (ko < kold) { ko = ko + 1; continue; }
        tha = thb;
        thb = thnew[k];
        pinteg = pinteg + pkold[kold] * (thb - tha);
      }
      break;
    }
    pknew[k] = pknew[k - 1];
    if (thnew[k] > thnew[k - 1]) pknew[k] = pinteg / (thnew[k] - thnew[k - 1]);
  }
}

static inline double cushn(double arg, double arg1) {
  double m = dmin(arg, 2. * arg1);
  return .25 * (m * m) / arg1 + dmax(arg, 2. * arg1) - arg1;
}

static void inflate(const fo_grid *g, double psurf, double *exner) {
  const int nvl = g->nvl;
  double dpsum = 0.;
  double extop = CP * xpow(g, PTOP / P1000, RD / CP);
  const double top = .2;
  double shrink = dmin(1., (psurf / P1000 - top) / (1. - top));
  int last = 1;
  for (int k = 1; k <= nvl - 1; k++) {
    double dp0 = RD * g->dpsig[k - 1] / P1000 * shrink;
    dpsum = dpsum + dp0;
    double exold = exner[k];
    if (k <= KBL) {
      double dp1 = dp0;
      last = k + 1;
      exner[k] = exner[k - 1] - dp1;
    } else {
      if (k > last) {
        double p2 = 1.;
        for (int i = 0; i < k - last; i++) p2 = p2 * 2.;
        dp0 = dp0 / dmin(10., 1.25 * p2);
        dp0 = dmax(dp0, RD * 100. / P1000);
      }
      double dp1 = cushn(dmax(0., exner[k - 1] - dmin(exold, exner[0] - dpsum)), dp0);
      if (exner[k - 1] - dp1 < exold - THIN) {
        if (k == last) last = k + 1;
        exner[k] = dmax(extop, exner[k - 1] - dp1);
      }
    }
  }
  /* THKTOP > 0 */
  for (int k = nvl; k >= 2; k--) {
    double q = exner[k - 1] / CP;
    q = q * q * sqrt(q);
    double tq = .25 * (double)(nvl - k);
    double tap = 1. / (1. + tq * tq);
    double dp0 = RD * THKTOP / P1000 * tap / q;
    double exold = exner[k - 1];
    double dp1 = cushn(dmax(0., exold - exner[k]), dp0);
    if (exner[k] + dp1 > exold + THIN) exner[k - 1] = exner[k] + dp1;
    if (exner[k - 1] < exner[k - 2] - dp0) break;
  }
}

static void equilb(const double *thet, double *pres, int kk, double slak, double *pnew) {
  for (int k = 0; k <= kk; k++) pnew[k] = pres[k];
  for (int k = 3; k <= kk - 2; k++) {
    if (pnew[k - 3] > pnew[0] - THIN) continue;
    double dp1 = pnew[k - 3] - pnew[k - 2];
    double dp2 = pnew[k - 2] - pnew[k - 1];
    double dp3 = pnew[k - 1] - pnew[k];
    double dp4 = pnew[k] - pnew[k + 1];
    double dp5 = pnew[k + 1] - pnew[k + 2];
    double th2 = thet[k - 2], th3 = thet[k - 1], th4 = thet[k];
    if (dp2 > dp1 && dp4 > dp5) {
      double goal = .5 * (dp3 + dmin(dp2, dp4));
      if (dp2 > goal && dp4 > goal) {
        double dis3 = dmin(dp2 - goal, goal - dp3) * slak;
        double dis4 = dmin(dp4 - goal, goal - dp3) * slak;
        if (th3 > th2 && th4 > th3) {
          double ratio = (th4 - th3) / (th3 - th2);
          if (ratio > 1.) dis4 = dis3 / ratio;
          else dis3 = dis4 * ratio;
        }
        pnew[k - 1] = pnew[k - 1] + dis3;
        pnew[k] = pnew[k] - dis4;
      }
    }
  }
  for (int k = 0; k <= kk; k++) pres[k] = pnew[k];
}

/* work: at least 6*(nvl+1) doubles */
static void regrid_1d(const fo_grid *g, int its, const double *targt, double *thcol, double *excol,
                      double psurf, double *work) {
  const int nvl = g->nvl, nvlp1 = nvl + 1;
  double *exwrk = work, *exnew = work + nvlp1, *thnew = work + 2 * nvlp1;
  double *heatfx = work + 3 * nvlp1, *tmp = work + 4 * nvlp1;
  const double dffudt = .1, tolrnce = 0.001;
  double eqlb_slak = .1 * (double)its / (10. + (double)its);
  for (int k = 1; k <= nvl; k++)
    if (excol[k - 1] < excol[k] - THIN) excol[k] = excol[k - 1];
  for (int k = nvl - 1; k >= 1; k--) thcol[k - 1] = dmin(thcol[k - 1], thcol[k]);
  for (int k = 0; k < nvlp1; k++) exwrk[k] = excol[k];
  for (int k = nvl; k >= 2; k--) {
    if (thcol[k - 1] <= targt[nvl - 1] + tolrnce) break;
    double wgt1 = dmax(0., exwrk[k - 2] - exwrk[k - 1]);
    double wgt2 = dmax(THIN, exwrk[k - 1] - exwrk[k]);
    double try_ = (thcol[k - 2] * wgt1 + thcol[k - 1] * wgt2) / (wgt1 + wgt2);
    if (try_ > targt[nvl - 1] - tolrnce) {
      exwrk[k - 1] = exwrk[nvlp1 - 1];
      for (int k1 = k - 1; k1 <= nvl; k1++) thcol[k1 - 1] = try_;
    } else {
      exwrk[k - 1] = dmin(exwrk[k - 2], dmax(exwrk[k], (exwrk[k - 1] * (thcol[k - 1] - thcol[k - 2]) +
                                                         exwrk[k] * (targt[nvl - 1] - thcol[k - 1])) /
                                                            (targt[nvl - 1] - thcol[k - 2])));
      for (int k1 = k; k1 <= nvl; k1++) thcol[k1 - 1] = targt[nvl - 1];
      break;
    }
  }
  restp_1d(thcol, exwrk, nvl, thnew, exnew, targt, nvl, tmp);
  heatfx[nvl - 1] = 0.;
  heatfx[0] = 0.;
  for (int iter = 1; iter <= 3; iter++) {
    for (int k = 0; k < nvlp1; k++) exwrk[k] = exnew[k];
    for (int k = 2; k <= nvl - 1; k++) {
      heatfx[k - 1] = 0.;
      if (exwrk[k - 1] < exwrk[0] - .01 && exwrk[k] > exwrk[nvlp1 - 1] + .01) {
        double cnv = GRVITY / targt[k - 1];
        heatfx[k - 1] = dffudt * cnv * cnv * .5 * (targt[k] - targt[k - 2]) / dmax(.03 * cnv, exwrk[k - 1] - exwrk[k]);
      }
    }
    for (int k = 1; k <= nvl - 1; k++) {
      if (exwrk[k] < exwrk[0] - .01 && exwrk[k] > exwrk[nvlp1 - 1] + .01)
        exnew[k] = dmax(exwrk[k + 1], dmin(exwrk[k - 1], exwrk[k] + (heatfx[k] - heatfx[k - 1]) / (targt[k] - targt[k - 1])));
    }
    int event = 0;
    for (int k = 2; k <= nvlp1; k++)
      if (exnew[k - 1] > exnew[k - 2] + THIN) event = 1;
    if (!event) break;
  }
  equilb(thnew, exnew, nvl, eqlb_slak, tmp);
  inflate(g, psurf, exnew);
  for (int k = 0; k < nvlp1; k++) excol[k] = exnew[k];
}

/* work: at least 10*(nvl+1) doubles */
static void pcwise(int kk, const double *xold, const double *yold, const double *xnew, double *ynew,
                   double *work) {
  double *zold = work, *znew = work + (kk + 1), *delz = work + 2 * (kk + 1), *fco = work + 3 * (kk + 1);
  double *fcn = work + 4 * (kk + 1), *vertfx = work + 5 * (kk + 1), *fld = work + 6 * (kk + 1);
  double *a = work + 7 * (kk + 1), *b = work + 8 * (kk + 1);
  const double small = 1.e-22;
  if (xold[0] < xold[kk]) {
    for (int k = 0; k <= kk; k++) { zold[k] = xold[k]; znew[k] = xnew[k]; }
  } else {
    for (int k = 0; k <= kk; k++) { zold[k] = -xold[k]; znew[k] = -xnew[k]; }
  }
  for (int k = 0; k <= kk; k++) delz[k] = znew[k] - zold[k];
  for (int k = 0; k < kk; k++) {
    fco[k] = dmax(0., zold[k + 1] - zold[k]);
    fcn[k] = dmax(0., znew[k + 1] - znew[k]);
  }
  for (int k = 0; k < kk; k++) fld[k] = yold[k];
  for (int k = kk - 1; k >= 1; k--)
    fld[k - 1] = (fld[k - 1] * fco[k - 1] + fld[k] * small) / (fco[k - 1] + small);
  for (int k = 2; k <= kk; k++)
    fld[k - 1] = (fld[k - 1] * fco[k - 1] + fld[k - 2] * small) / (fco[k - 1] + small);
  a[0] = fld[0]; b[0] = 0.;
  a[kk - 1] = fld[kk - 1]; b[kk - 1] = 0.;
  for (int k = 2; k <= kk - 1; k++) {   /* PLM (remap_optn = 2) */
    double fm = fld[k - 2], f0 = fld[k - 1], fp = fld[k];
    a[k - 1] = f0;
    b[k - 1] = 0.;
    if (f0 <= dmin(fm, fp) || f0 >= dmax(fm, fp)) b[k - 1] = 0.;
    else if ((fp - fm) * (fm + fp - 2. * f0) > 0.) b[k - 1] = f0 - fm;
    else b[k - 1] = fp - f0;
  }
  /* miglim = 1; c = 0 for PLM */
  for (int k = 2; k <= kk; k++) {
    double slab = 0., amount = 0.;
    vertfx[k - 1] = 0.;
    double dz = delz[k - 1];
    if (dz > 0.) {
      int lyr = k;
      if (fco[lyr - 1] > 0.) {
        double dslab = dmin(slab + fco[lyr - 1], dz) - dmin(slab, dz);
        double dx = dslab / fco[lyr - 1];
        double fcdx = a[lyr - 1] + b[lyr - 1] * .5 * (dx - 1.) + 0. * (.25 - dx * (.5 - dx * (1. / 3.)));
        amount = amount + fcdx * dslab;
        slab = slab + dslab;
      }
    } else if (dz < 0.) {
      int lyr = k - 1;
      if (fco[lyr - 1] > 0.) {
        double dslab = dmin(slab + fco[lyr - 1], -dz) - dmin(slab, -dz);
        double dx = dslab / fco[lyr - 1];
        double fcdx = a[lyr - 1] + b[lyr - 1] * .5 * (1. - dx) + 0. * (.25 - dx * (.5 - dx * (1. / 3.)));
        amount = amount + fcdx * dslab;
        slab = slab + dslab;
      }
    }
    if (slab != 0.) vertfx[k - 1] = -dz * amount / slab;
  }
  vertfx[0] = 0.;
  vertfx[kk] = 0.;
  for (int k = 0; k < kk; k++) {
    double vertdv = vertfx[k + 1] - vertfx[k];
    double amount = fld[k] * fco[k] - vertdv;
    ynew[k] = (amount + fld[k] * small) / (fcn[k] + small);
  }
}

/* work: at least (14 + 2*ntra)*(nvl+1) doubles */
static void remap_1d(const fo_grid *g, int its, const double *targt, int ntra, double *trcol,
                     size_t trstride, double *ucol, double *vcol, const double *exold, double *excol,
                     double *dpcol, double *prcol, double *work) {
  const int nvl = g->nvl, nvlp1 = nvl + 1;
  double *pk1col = work, *exnew = work + nvlp1, *prnew = work + 2 * nvlp1, *pk1new = work + 3 * nvlp1;
  double *unew = work + 4 * nvlp1, *vnew = work + 5 * nvlp1, *trnew = work + 6 * nvlp1;
  double *pw = work + (6 + ntra) * nvlp1;
  const double small = 1.e-6;
  const double cprd = CP / RD;
  for (int k = 0; k < nvlp1; k++) pk1col[k] = exold[k] * prcol[k] / CP;
  for (int k = 0; k < nvlp1; k++) exnew[k] = excol[k];
  inflate(g, prcol[0], exnew);   /* intfc_smooth > 0 */
  if (its > 0)
    for (int k = 0; k < nvlp1; k++) exnew[k] = SLAK * exnew[k] + (1. - SLAK) * exold[k];
  /* miglim = 1 */
  for (int k = 2; k <= nvlp1 - 1; k++)
    exnew[k - 1] = dmax(.0001 * exold[k - 1] + .9999 * exold[k], dmin(.0001 * exold[k - 1] + .9999 * exold[k - 2], exnew[k - 1]));
  for (int k = nvlp1; k >= 1; k--) prnew[k - 1] = P1000 * xpow(g, exnew[k - 1] / CP, cprd);
  for (int k = 0; k < nvl; k++) dpcol[k] = prnew[k] - prnew[k + 1];
  for (int k = 0; k < nvlp1; k++) pk1new[k] = exnew[k] * prnew[k] / CP;
  prnew[nvlp1 - 1] = prcol[nvlp1 - 1];
  pk1new[nvlp1 - 1] = pk1col[nvlp1 - 1];
  prnew[0] = prcol[0];
  pk1new[0] = pk1col[0];
  pcwise(nvl, prcol, ucol, prnew, unew, pw);
  pcwise(nvl, prcol, vcol, prnew, vnew, pw);
  for (int n = 1; n < ntra; n++) pcwise(nvl, prcol, trcol + n * trstride, prnew, trnew + n * nvlp1, pw);
  pcwise(nvl, pk1col, trcol, pk1new, trnew, pw);
  for (int k = nvl; k >= 3; k--) {
    double dplo = dmax(pk1new[k - 2] - pk1new[k - 1], pk1new[0] * small);
    double dpup = dmax(pk1new[k - 1] - pk1new[k], pk1new[0] * small);
    double tha = trnew[k - 2], thb = trnew[k - 1];
    double devilo = (tha - targt[k - 2]) * dplo;
    double deviup = (thb - targt[k - 1]) * dpup;
    if (deviup > 0. && devilo < 0.) {
      trnew[k - 2] = tha + dmin(deviup, -devilo) / dplo;
      trnew[k - 1] = thb - dmin(deviup, -devilo) / dpup;
    } else if (deviup < 0. && devilo > 0.) {
      trnew[k - 2] = tha - dmin(devilo, -deviup) / dplo;
      trnew[k - 1] = thb + dmin(devilo, -deviup) / dpup;
    }
  }
  for (int k = 0; k < nvl; k++) { ucol[k] = unew[k]; vcol[k] = vnew[k]; }
  for (int n = 0; n < ntra; n++)
    for (int k = 0; k < nvl; k++) trcol[n * trstride + k] = trnew[n * nvlp1 + k];
  for (int k = 0; k < nvlp1; k++) { excol[k] = exnew[k]; prcol[k] = prnew[k]; }
  for (int k = 1; k <= nvl; k++)
    if (excol[k - 1] < excol[k]) excol[k] = excol[k - 1];
}

void fo_hybgen(const fo_grid *g, fo_state *s, int its) {
  const int nip = g->nip, nvl = g->nvl, nvlp1 = nvl + 1, ntra = g->ntra;
  double *exsmo = s->work_l1;
  /* regrid every column (hybgen.F90:126-139) */
  #pragma omp parallel
  {
    double *work = (double *)malloc(sizeof(double) * 6 * nvlp1);
    double *thcol = (double *)malloc(sizeof(double) * nvl);
    #pragma omp for schedule(static)
    for (int ipn = 0; ipn < nip; ipn++) {
      for (int k = 0; k < nvl; k++) thcol[k] = TR(s->tr, 0, ipn, k);   /* worka = theta */
      for (int k = 0; k < nvlp1; k++) L2(exsmo, ipn, k) = L2(s->ex, ipn, k);
      if (fo_wet(g, ipn)) regrid_1d(g, its, g->thetac, thcol, &L2(exsmo, ipn, 0), L2(s->pr, ipn, 0), work);
    }
    free(work);
    free(thcol);
  }
  /* lateral interface smoothing dffusn_lev(exsmo, smoo_coeff, nvlp1, kbl+2, nvl) */
  double *flxdv = s->work_l2;
  int k1 = KBL + 2, k2 = nvl;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    for (int k = 0; k < nvlp1; k++) L2(flxdv, ipn, k) = 0.;
    for (int edg = 0; edg < g->nprox[ipn]; edg++) {
      int ipx = PROX(ipn, edg);
      for (int k = k1; k <= k2; k++)
        L2(flxdv, ipn, k - 1) = L2(flxdv, ipn, k - 1) + (L2(exsmo, ipn, k - 1) - L2(exsmo, ipx, k - 1)) * g->sideln[ipn * NPP + edg];
    }
  }
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++)
    for (int k = k1; k <= k2; k++) {
      double factor = -g->smoo_coeff[k - 1] * g->rarea[ipn];
      L2(exsmo, ipn, k - 1) = L2(exsmo, ipn, k - 1) + L2(flxdv, ipn, k - 1) * factor;
    }
  /* remap every column (hybgen.F90:177-192) */
  #pragma omp parallel
  {
    double *work = (double *)malloc(sizeof(double) * (14 + 2 * ntra) * nvlp1);
    #pragma omp for schedule(static)
    for (int ipn = 0; ipn < nip; ipn++) {
      if (!fo_wet(g, ipn)) {             /* a land column: no remap */
        for (int k = 0; k < nvlp1; k++) L2(s->sdot, ipn, k) = 0.;
        continue;
      }
      remap_1d(g, its, g->thetac, ntra, &TR(s->tr, 0, ipn, 0), (size_t)nip * nvl, &C2(s->us, ipn, 0),
               &C2(s->vs, ipn, 0), &L2(s->ex, ipn, 0), &L2(exsmo, ipn, 0), &C2(s->dp, ipn, 0),
               &L2(s->pr, ipn, 0), work);
      for (int k = 0; k < nvlp1; k++) {
        L2(s->sdot, ipn, k) = L2(exsmo, ipn, k) - L2(s->ex, ipn, k);
        L2(s->ex, ipn, k) = L2(exsmo, ipn, k);
      }
    }
    free(work);
  }
}

/* hystat.F90:23-115 */
void fo_hystat(const fo_grid *g, fo_state *s, int its) {
  const int nip = g->nip, nvl = g->nvl, nvlp1 = nvl + 1, ntra = g->ntra;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++) {
    double work = s->psrf[ipn];
    double ps = C2(s->dp, ipn, nvl - 1);
    for (int k = nvl - 1; k >= 1; k--) ps = ps + C2(s->dp, ipn, k - 1);
    s->psrf[ipn] = ps;
    int ns = its % 2;
    if (its > 0) s->ptdcy[ns * nip + ipn] = ps - work;
    C2(s->mp, ipn, 0) = L2(s->ex, ipn, 0) * TR(s->tr, 0, ipn, 0) + L2(s->ph, ipn, 0);
    L2(s->ph, ipn, 1) = C2(s->mp, ipn, 0) - TR(s->tr, 0, ipn, 0) * L2(s->ex, ipn, 1);
    for (int k = 2; k <= nvl; k++) {
      C2(s->mp, ipn, k - 1) = C2(s->mp, ipn, k - 2) + L2(s->ex, ipn, k - 1) * (TR(s->tr, 0, ipn, k - 1) - TR(s->tr, 0, ipn, k - 2));
      L2(s->ph, ipn, k) = C2(s->mp, ipn, k - 1) - TR(s->tr, 0, ipn, k - 1) * L2(s->ex, ipn, k);
    }
    for (int k = 0; k < nvl; k++) {
      if (ntra > 1) TR(s->tr, 1, ipn, k) = dmax(QVMIN, TR(s->tr, 1, ipn, k));
      if (ntra > 2) TR(s->tr, 2, ipn, k) = dmax(QWMIN, TR(s->tr, 2, ipn, k));
    }
    for (int t = 0; t < ntra; t++)
      for (int k = 0; k < nvl; k++) TR(s->trdp, t, ipn, k) = TR(s->tr, t, ipn, k) * C2(s->dp, ipn, k);
  }
}

/* start-up: hybgen(its=0), pressure refresh (dyn_init.F90:884-893), hystat;
 * zero tendencies and set the AB slots (dyn_init.F90:1042-1094) */
/* ------------------------------------------------------------------------
 * Class-B tracer transport on the long time step (module_transp3d,
 * transp3d.F90:22-241; dyn_run.F90:388-405; dyn_init.F90:1096-1098).
 * transp0: reset the time-integrated mass flux and keep the thickness. */
void fo_transp0(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl;
  memset(s->cumufx, 0, sizeof(double) * (size_t)nip * NPP * nvl);
  memcpy(s->dpinit, s->dp, sizeof(double) * (size_t)nip * nvl);
}

/* transp1 (transp3d.F90:44-78): add this step's AB-weighted mass flux */
void fo_transp1(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl;
  const int nf = s->nf - 1, of = s->of - 1, vof = s->vof - 1;
  const double ab1 = s->ab1, ab2 = s->ab2, ab3 = s->ab3;
  #pragma omp parallel for schedule(static)
  for (int ipn = 0; ipn < nip; ipn++)
    for (int edg = 0; edg < g->nprox[ipn]; edg++)
      for (int k = 0; k < nvl; k++)
        E3(s->cumufx, ipn, edg, k) = E3(s->cumufx, ipn, edg, k) + ab1 * MF(s->massfx, nf, ipn, edg, k) +
                                     ab2 * MF(s->massfx, of, ipn, edg, k) + ab3 * MF(s->massfx, vof, ipn, edg, k);
}

/* fct3d (fct3d.F90:4-483, diagno = .false., recovr = .false.): 3-D FCT of
 * nfld fields (nvl,nip) with time-integrated horizontal mass fluxes u
 * (nvl,6,nip), vertical mass fluxes w (nvl,nip), thickness thko -> thkn */
void fo_fct3d(const fo_grid *g, double *fields, int nfld, const double *u, const double *w,
              const double *thko, const double *thkn) {
  const int nip = g->nip, nvl = g->nvl;
  const double athird = 1. / 3., epsil = 1.e-11, onemu = 1.e-6;
  const size_t cz = (size_t)nip * nvl, ez = cz * NPP;
  double *fld = malloc(sizeof(double) * cz), *vfx = malloc(sizeof(double) * cz);
  double *vdv = malloc(sizeof(double) * cz), *fmx = malloc(sizeof(double) * cz);
  double *fmn = malloc(sizeof(double) * cz), *flp = malloc(sizeof(double) * cz);
  double *fln = malloc(sizeof(double) * cz);
  double *flx = malloc(sizeof(double) * ez), *uan = malloc(sizeof(double) * ez);
  for (int n = 0; n < nfld; n++) {
    double *f = fields + (size_t)n * cz;
    /* massless cells take their values from the layers above and below */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nip; i++) {
      for (int k = 0; k < nvl; k++) C2(fld, i, k) = C2(f, i, k);
      for (int k = nvl - 2; k >= 0; k--)
        C2(fld, i, k) = (C2(fld, i, k) * C2(thko, i, k) + C2(fld, i, k + 1) * onemu) / (C2(thko, i, k) + onemu);
      for (int k = 1; k < nvl; k++)
        C2(fld, i, k) = (C2(fld, i, k) * C2(thko, i, k) + C2(fld, i, k - 1) * onemu) / (C2(thko, i, k) + onemu);
    }
    /* piecewise parabolic fit per layer, vertical flux over the upstream slab */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nip; i++) {
      double a[64 + 2], b[64 + 2], c[64 + 2];
      const double *fc = fld + (size_t)i * nvl, *th = thko + (size_t)i * nvl, *wc = w + (size_t)i * nvl;
      a[0] = fc[0]; b[0] = 0.; c[0] = 0.;
      a[nvl - 1] = fc[nvl - 1]; b[nvl - 1] = 0.; c[nvl - 1] = 0.;
      for (int k = 1; k < nvl - 1; k++) {
        const double dxlft = dmax(epsil, th[k - 1]), dxmid = dmax(epsil, th[k]), dxrgt = dmax(epsil, th[k + 1]);
        const double yl = (dxlft * fc[k] + dxmid * fc[k - 1]) / (dxlft + dxmid);
        const double yr = (dxrgt * fc[k] + dxmid * fc[k + 1]) / (dxrgt + dxmid);
        a[k] = 1.5 * fc[k] - .25 * (yl + yr);
        b[k] = yr - yl;
        c[k] = 6. * (.5 * (yl + yr) - fc[k]);
        if (fabs(yr - yl) < 6. * fabs(.5 * (yl + yr) - fc[k])) {
          if (fabs(yr - yl) > 2. * fabs(.5 * (yl + yr) - fc[k])) {
            if ((yr - yl) * (.5 * (yl + yr) - fc[k]) > 0.) {
              a[k] = .25 * (3. * fc[k] + yl);
              c[k] = 3. * (fc[k] - yl);
              b[k] = c[k];
            } else {
              a[k] = .25 * (3. * fc[k] + yr);
              c[k] = 3. * (fc[k] - yr);
              b[k] = -c[k];
            }
          } else {
            a[k] = fc[k]; b[k] = 0.; c[k] = 0.;
          }
        }
      }
      double amount = 0.;   /* a zero w keeps the previous layer's amount; its flux is 0 */
      for (int k = 0; k < nvl - 1; k++) {
        double slab = onemu;
        if (wc[k] < 0.) {                 /* interface moves up: slab from above */
          amount = slab * fc[k + 1];
          int kp = k;
          for (;;) {
            kp = kp + 1;
            if (slab >= -wc[k]) break;
            if (th[kp] > 0.) {
              const double dslab = dmin(slab + th[kp], -wc[k]) - dmin(slab, -wc[k]);
              const double dx = dslab / th[kp];
              const double fcdx = a[kp] + b[kp] * .5 * (dx - 1.) + c[kp] * (.25 - dx * (.5 - dx * athird));
              amount = amount + fcdx * dslab;
              slab = slab + dslab;
            }
            if (!(kp < nvl - 1)) break;
          }
        } else if (wc[k] > 0.) {          /* interface moves down: slab from below */
          amount = slab * fc[k];
          int kp = k + 1;
          for (;;) {
            kp = kp - 1;
            if (slab >= wc[k]) break;
            if (th[kp] > 0.) {
              const double dslab = dmin(slab + th[kp], wc[k]) - dmin(slab, wc[k]);
              const double dx = dslab / th[kp];
              const double fcdx = a[kp] + b[kp] * .5 * (1. - dx) + c[kp] * (.25 - dx * (.5 - dx * athird));
              amount = amount + fcdx * dslab;
              slab = slab + dslab;
            }
            if (!(kp > 1)) break;     /* the Fortran continues while kp > 2 (1-based) */
          }
        }
        C2(vfx, i, k) = wc[k] * amount / slab;
      }
      C2(vfx, i, nvl - 1) = 0.;
      C2(vdv, i, 0) = C2(vfx, i, 0);
      for (int k = 1; k < nvl; k++) C2(vdv, i, k) = C2(vfx, i, k) - C2(vfx, i, k - 1);
    }
    /* low-order and antidiffusive fluxes, local bounds */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nip; i++)
      for (int k = 0; k < nvl; k++) {
        double mx = C2(fld, i, k), mn = C2(fld, i, k);
        for (int edg = 0; edg < g->nprox[i]; edg++) {
          const int ix = PROX(i, edg), edx = PROXS(i, edg);
          const double ui = E3(u, i, edg, k), ux = E3(u, ix, edx, k);
          E3(flx, i, edg, k) = 0.5 * ((ui + fabs(ui)) * C2(fld, i, k) - (ux + fabs(ux)) * C2(fld, ix, k));
          const double q = .5 * (C2(fld, i, k) + C2(fld, ix, k));
          E3(uan, i, edg, k) = q * ui - E3(flx, i, edg, k);
          mx = dmax(mx, C2(fld, ix, k));
          mn = dmin(mn, C2(fld, ix, k));
        }
        if (k < nvl - 1 && C2(w, i, k) < 0.) {
          mx = dmax(mx, C2(vfx, i, k) / C2(w, i, k));
          mn = dmin(mn, C2(vfx, i, k) / C2(w, i, k));
        }
        if (k > 0 && C2(w, i, k - 1) > 0.) {
          mx = dmax(mx, C2(vfx, i, k - 1) / C2(w, i, k - 1));
          mn = dmin(mn, C2(vfx, i, k - 1) / C2(w, i, k - 1));
        }
        C2(fmx, i, k) = mx;
        C2(fmn, i, k) = mn;
      }
    /* low-order update, then the ratios of permissible to incoming and
     * outgoing antidiffusive fluxes */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nip; i++)
      for (int k = 0; k < nvl; k++) {
        double hd = 0.;
        for (int edg = 0; edg < g->nprox[i]; edg++) hd = hd + E3(flx, i, edg, k);
        hd = hd * g->rarea[i];
        const double q = C2(fld, i, k) * C2(thko, i, k) - hd - C2(vdv, i, k);
        const double amount = dmax(dmax(0., C2(fmn, i, k) * C2(thkn, i, k)), dmin(q, C2(fmx, i, k) * C2(thkn, i, k)));
        C2(fld, i, k) = (C2(fld, i, k) * onemu + amount) / (onemu + C2(thkn, i, k));
      }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nip; i++)
      for (int k = 0; k < nvl; k++) {
        double totin = 0., totou = 0.;
        for (int edg = 0; edg < g->nprox[i]; edg++) {
          totin = totin - dmin(0., E3(uan, i, edg, k));
          totou = totou + dmax(0., E3(uan, i, edg, k));
        }
        C2(flp, i, k) = (C2(fmx, i, k) - C2(fld, i, k)) * C2(thkn, i, k) / (totin + epsil) * g->rarea[i];
        C2(fln, i, k) = (C2(fld, i, k) - C2(fmn, i, k)) * C2(thkn, i, k) / (totou + epsil) * g->rarea[i];
      }
    /* limited antidiffusive fluxes and the final update */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nip; i++)
      for (int k = 0; k < nvl; k++) {
        double hd = 0.;
        for (int edg = 0; edg < g->nprox[i]; edg++) {
          const int ix = PROX(i, edg);
          const double an = E3(uan, i, edg, k);
          const double clip = an >= 0. ? dmin(dmin(1., C2(fln, i, k)), C2(flp, ix, k))
                                       : dmin(dmin(1., C2(flp, i, k)), C2(fln, ix, k));
          hd = hd + an * clip;
        }
        hd = hd * g->rarea[i];
        const double q = C2(fld, i, k) * C2(thkn, i, k) - hd;
        const double amount = dmax(dmax(0., C2(fmn, i, k) * C2(thkn, i, k)), dmin(q, C2(fmx, i, k) * C2(thkn, i, k)));
        C2(f, i, k) = (C2(fld, i, k) * onemu + amount) / (onemu + C2(thkn, i, k));
      }
  }
  free(fld); free(vfx); free(vdv); free(fmx); free(fmn); free(flp); free(fln); free(flx); free(uan);
}

/* transp2 (transp3d.F90:81-241): the vertical mass flux that closes the
 * time-integrated continuity equation, then fct3d of the class-B tracers
 * from dpinit to the current thickness */
void fo_transp2(const fo_grid *g, fo_state *s) {
  const int nip = g->nip, nvl = g->nvl;
  double *w = malloc(sizeof(double) * (size_t)nip * nvl);
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < nip; i++) {
    double hordiv[64];
    for (int k = 0; k < nvl; k++) hordiv[k] = 0.;
    for (int edg = 0; edg < g->nprox[i]; edg++)
      for (int k = 0; k < nvl; k++) hordiv[k] = hordiv[k] + E3(s->cumufx, i, edg, k);
    for (int k = 0; k < nvl; k++) {
      const double vertdv = (C2(s->dpinit, i, k) - C2(s->dp, i, k)) - hordiv[k] * g->rarea[i];
      C2(w, i, k) = k == 0 ? vertdv : C2(w, i, k - 1) + vertdv;
    }
  }
  if (g->ntrb > 0) fo_fct3d(g, s->trb, g->ntrb, s->cumufx, w, s->dpinit, s->dp);
  free(w);
}

void fo_init(const fo_grid *g, fo_state *s, int init_hybgen) {
  const int nip = g->nip, nvl = g->nvl, nvlp1 = nvl + 1, ntra = g->ntra;
  for (int t = 0; t < ntra; t++)
    for (int ipn = 0; ipn < nip; ipn++)
      for (int k = 0; k < nvl; k++) TR(s->trdp, t, ipn, k) = TR(s->tr, t, ipn, k) * C2(s->dp, ipn, k);
  if (init_hybgen) {
    fo_hybgen(g, s, 0);
    const double cprd = CP / RD;
    for (int ipn = 0; ipn < nip; ipn++) {
      L2(s->pr, ipn, 0) = P1000 * xpow(g, L2(s->ex, ipn, 0) / CP, cprd);
      for (int k = 1; k <= nvl; k++) L2(s->pr, ipn, k) = P1000 * xpow(g, L2(s->ex, ipn, k) / CP, cprd);
    }
  }
  fo_hystat(g, s, 0);
  memset(s->u_tdcy, 0, sizeof(double) * 3 * nip * nvl);
  memset(s->v_tdcy, 0, sizeof(double) * 3 * nip * nvl);
  memset(s->dp_tdcy, 0, sizeof(double) * 3 * nip * nvl);
  memset(s->dpl_tdcy, 0, sizeof(double) * 3 * nip * nvl);
  memset(s->trc_tdcy, 0, sizeof(double) * 3 * ntra * nip * nvl);
  memset(s->trl_tdcy, 0, sizeof(double) * 3 * ntra * nip * nvl);
  memset(s->massfx, 0, sizeof(double) * 3 * nip * NPP * nvl);
  s->nf = 0; s->of = 2; s->vof = 1;
  if (g->ntrb > 0) fo_transp0(g, s);
}

/* one step in dyn_run order: DYN_1(its) then DYN_2(its) (dyn_run.F90:119-425);
 * mode 1 is transport only (edgvar1, edgvar3, cnuity, trcadv) */
void fo_step(const fo_grid *g, fo_state *s, int its, int mode) {
  fo_abstart(s, its, g->dt);
  fo_edgvar1(g, s);
  fo_edgvar3(g, s);
  if (g->flux_scheme == 2) fo_mass_adv(g, s);
  else fo_cnuity(g, s);
  if (g->ntrb > 0) {                 /* dyn_run.F90:388-405 */
    fo_transp1(g, s);
    if (its % g->dtratio == 0) {
      fo_transp2(g, s);
      fo_transp0(g, s);
    }
  }
  fo_trcadv(g, s);
  if (mode == 0) {
    fo_momtum(g, s);
    fo_hybgen(g, s, its);
    fo_hystat(g, s, its);
  }
}

/* globsum.F90:70-101: area-weighted column mass and tracer totals
 * (ntra + ntrb of them) */
void fo_globsum(const fo_grid *g, const fo_state *s, double *mass, double *trtot) {
  const int nip = g->nip, nvl = g->nvl, ntra = g->ntra;
  double summ = 0.;
  for (int ipn = 0; ipn < nip; ipn++)
    for (int k = 0; k < nvl; k++) summ = summ + g->area[ipn] * C2(s->dp, ipn, k);
  *mass = summ / GRVITY;
  for (int t = 0; t < ntra; t++) {
    double st = 0.;
    for (int ipn = 0; ipn < nip; ipn++)
      for (int k = 0; k < nvl; k++) st = st + g->area[ipn] * C2(s->dp, ipn, k) * TR(s->tr, t, ipn, k);
    trtot[t] = st;
  }
  for (int t = 0; t < g->ntrb; t++) {     /* class-B tracers follow (tr3d(:,:,ntra+1:)) */
    double st = 0.;
    for (int ipn = 0; ipn < nip; ipn++)
      for (int k = 0; k < nvl; k++) st = st + g->area[ipn] * C2(s->dp, ipn, k) * TR(s->trb, t, ipn, k);
    trtot[ntra + t] = st;
  }
}

double fo_pow(double x, double y) { return fim_pow(x, y); }
double fo_pow_fast(double x, double y) { return fim_pow_fast(x, y); }
