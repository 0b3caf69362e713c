/* hycom_oracle.c - CPU restatement, in plain C and FP64, of the HYCOM ocean
 * continuity equation (FIMsrc/fim/horizontal/hycom_cnuity.F90, subroutine
 * cnuity), of its edge interpolation (hycom_edgvar.F90, bclin_edgvar with
 * uvedge and blokf), of its momentum equation (hycom_momtum.F90 momtum
 * with hycom_dissip.F90's del4prep and dissip) and of its tracer transport
 * (hycom_transp3d.F90 transp0/1/2 with hycom_fct3d.F90's fct3d and
 * hycom_dffusn.F90's dffusn_lyr).  TEST INFRASTRUCTURE ONLY: the checker the HIP kernels
 * are compared with; the product never links it.
 *
 * Every loop follows the reference operation for operation, in the
 * reference loop order, so that with fp contraction off it reproduces the
 * reference built with -fdefault-real-8 (oracle/_ref/bin/hycomref) bit for
 * bit.  Diagnostics (prints, stencil dumps) are left out; the reference's
 * duplicate-flux check (chkdup, hycom_cnuity.F90:1030-1075), which stops the
 * program, returns 1 here instead.
 *
 * Layout (C index, level k fastest, cells i, 0-based):
 *   dp      (kdm,nip,2)     dp[(s*nip + i)*kdm + k]      s = leapfrog slot-1
 *   pres    (kdm+1,nip)     pres[i*(kdm+1) + k]
 *   edge    (kdm,npp,nip)   a[(i*6 + e)*kdm + k]
 * Grid: prox/proxs 0-based [i*6+e] (a pentagon's slot 6 repeats slot 5),
 * sidevec_e [(i*6+e)*2 + c], sideln/rsideln/rprox_ln [i*6+e].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hycom_oracle.h"

#define NPP 6
/* hycom_constants.F90:30-77, module_constants.F90:19 */
static const double GRVITY = 9.80665;
static const double ONEM = 1000. * 9.80665;
#define ONECM (0.01 * ONEM)
static const double ASSLDP = .25;
static const double LAND_SPVAL = -1.e33;
static const double EPSIL = 1.e-11;

static double dmax2(double a, double b) { return a > b ? a : b; }
static double dmin2(double a, double b) { return a < b ? a : b; }

int fo_hycom_cnuity(const fo_hyc_grid *g, fo_hyc_mass *ms, int nstep, int leapm, int leapn,
                    const double *um_edg, const double *vm_edg, const double *un_edg, const double *vn_edg,
                    const double *dpm_edg, double *dp, double *pres, double *massflx, double *mssfx_avg) {
  const int nip = g->nip, kdm = g->kdm, kp = kdm + 1;
  const double dtleap = g->dtleap, dtinv = 1. / dtleap;
  const size_t cz = (size_t)nip * kdm, ez = cz * NPP;
  double *antiflx = malloc(ez * sizeof(double)), *bolusfx = malloc(ez * sizeof(double));
  double *flxdv = malloc(cz * sizeof(double)), *dp_lo = malloc(cz * sizeof(double));
  double *dpold = malloc(cz * sizeof(double)), *flxplus = malloc(cz * sizeof(double));
  double *flxmnus = malloc(cz * sizeof(double)), *rplus = malloc(cz * sizeof(double));
  double *rmnus = malloc(cz * sizeof(double)), *del2p = malloc(cz * sizeof(double));
  double *dptmp = malloc(cz * sizeof(double)), *clippd = malloc((size_t)nip * NPP * sizeof(double));
  double *rpbot = malloc((size_t)nip * sizeof(double)), *aux1 = malloc((size_t)nip * sizeof(double));
  double *sum = malloc((size_t)(kdm + 1) * sizeof(double));
  memset(dpold, 0, cz * sizeof(double));
  memset(rpbot, 0, (size_t)nip * sizeof(double));
  double *dpm = dp + (size_t)(leapm - 1) * cz, *dpn = dp + (size_t)(leapn - 1) * cz;
#define C(a, i, k) (a)[(size_t)(i) * kdm + (k)]
#define P(i, k) pres[(size_t)(i) * kp + (k)]
#define E(a, i, e, k) (a)[((size_t)(i) * NPP + (e)) * kdm + (k)]
#define SVE(i, e, c) g->sidevec_e[((size_t)(i) * NPP + (e)) * 2 + (c)]
#define WET(i) (g->wet[i] > 0)
  /* low- and high-order fluxes (hycom_cnuity.F90:98-159) */
  for (int i = 0; i < nip; i++) {
    for (int e = 0; e < NPP; e++)
      for (int k = 0; k < kdm; k++) {
        E(antiflx, i, e, k) = LAND_SPVAL;
        E(massflx, i, e, k) = LAND_SPVAL;
      }
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++) {
      const int ix = g->prox[i * NPP + e];
      if (!WET(ix)) {
        for (int k = 0; k < kdm; k++) {
          E(massflx, i, e, k) = 0.;
          E(antiflx, i, e, k) = 0.;
        }
        continue;
      }
      const int ex = g->proxs[i * NPP + e];
      for (int k = 0; k < kdm; k++) {
        double valu1 = SVE(i, e, 1) * E(um_edg, i, e, k) - SVE(i, e, 0) * E(vm_edg, i, e, k);
        double valu2 = SVE(ix, ex, 1) * E(um_edg, ix, ex, k) - SVE(ix, ex, 0) * E(vm_edg, ix, ex, k);
        const double outgom = .5 * (valu1 - valu2);
        valu1 = g->rsideln[i * NPP + e] * dtleap * g->rprox_ln[i * NPP + e];
        valu2 = g->rsideln[ix * NPP + ex] * dtleap * g->rprox_ln[ix * NPP + ex];
        const double q = dmin2(fabs(outgom) * .5 * (valu1 + valu2), 1.);
        const double dpedg = .5 * (E(dpm_edg, i, e, k) + E(dpm_edg, ix, ex, k));
        double hi;
        if (outgom >= 0.) hi = dpedg * (1. - q) + C(dpm, i, k) * q;
        else hi = dpedg * (1. - q) + C(dpm, ix, k) * q;
        hi = hi * outgom;
        valu1 = SVE(i, e, 1) * E(un_edg, i, e, k) - SVE(i, e, 0) * E(vn_edg, i, e, k);
        valu2 = SVE(ix, ex, 1) * E(un_edg, ix, ex, k) - SVE(ix, ex, 0) * E(vn_edg, ix, ex, k);
        const double outgon = .5 * (valu1 - valu2);
        const double lo = outgon >= 0. ? outgon * C(dpn, i, k) : outgon * C(dpn, ix, k);
        E(antiflx, i, e, k) = hi - lo;
        E(massflx, i, e, k) = lo;
      }
    }
  }
  /* low-order thickness (hycom_cnuity.F90:169-205) */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) C(dp_lo, i, k) = C(flxdv, i, k) = 0.;
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++)
      for (int k = 0; k < kdm; k++) C(flxdv, i, k) = C(flxdv, i, k) + E(massflx, i, e, k);
    for (int k = 0; k < kdm; k++) {
      C(dpold, i, k) = C(dpn, i, k);
      P(i, k + 1) = P(i, k) + C(dpn, i, k);
      C(dp_lo, i, k) = C(dpn, i, k) - C(flxdv, i, k) * g->rarea[i] * dtleap;
      if (C(dp_lo, i, k) < 0. && C(dp_lo, i, k) + ONECM > 0.) C(dp_lo, i, k) = 0.;
    }
  }
  /* in- and outgoing antidiffusive fluxes (246-265) */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) {
      C(flxplus, i, k) = EPSIL;
      C(flxmnus, i, k) = -EPSIL;
    }
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++)
      for (int k = 0; k < kdm; k++) {
        if (E(antiflx, i, e, k) >= 0.) C(flxplus, i, k) = C(flxplus, i, k) + E(antiflx, i, e, k);
        else C(flxmnus, i, k) = C(flxmnus, i, k) + E(antiflx, i, e, k);
      }
  }
  /* limiter ratios from the local bounds (269-300) */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) C(rplus, i, k) = C(rmnus, i, k) = 0.;
    if (!WET(i)) continue;
    double dpmax[kdm], dpmin[kdm];
    for (int k = 0; k < kdm; k++) {
      dpmax[k] = dmax2(C(dp_lo, i, k), C(dpn, i, k));
      dpmin[k] = dmin2(C(dp_lo, i, k), C(dpn, i, k));
    }
    for (int e = 0; e < g->nprox[i]; e++) {
      const int ix = g->prox[i * NPP + e];
      if (!WET(ix)) continue;
      for (int k = 0; k < kdm; k++) {
        dpmax[k] = dmax2(dmax2(dpmax[k], C(dp_lo, ix, k)), C(dpn, ix, k));
        dpmin[k] = dmin2(dmin2(dpmin[k], C(dp_lo, ix, k)), C(dpn, ix, k));
      }
    }
    for (int k = 0; k < kdm; k++) {
      dpmin[k] = dmax2(0., dpmin[k]);
      C(rmnus, i, k) = (C(dp_lo, i, k) - dpmin[k]) * g->area[i] / (dtleap * C(flxplus, i, k));
      C(rplus, i, k) = (C(dp_lo, i, k) - dpmax[k]) * g->area[i] / (dtleap * C(flxmnus, i, k));
    }
  }
  /* limit the antidiffusive fluxes, keeping the clipped part (315-352) */
  for (int i = 0; i < nip; i++) {
    for (int e = 0; e < NPP; e++) clippd[i * NPP + e] = 0.;
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++) {
      const int ix = g->prox[i * NPP + e];
      if (!WET(ix)) continue;
      for (int k = 0; k < kdm; k++) {
        const double a = E(antiflx, i, e, k);
        const double clip = a >= 0. ? dmin2(dmin2(1., C(rmnus, i, k)), C(rplus, ix, k))
                                    : dmin2(dmin2(1., C(rplus, i, k)), C(rmnus, ix, k));
        clippd[i * NPP + e] = clippd[i * NPP + e] + a * (1. - clip);
        E(antiflx, i, e, k) = a * clip;
        E(massflx, i, e, k) = E(massflx, i, e, k) + E(antiflx, i, e, k);
      }
    }
  }
  /* the antidiffusive update (357-393) */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) C(flxdv, i, k) = 0.;
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++)
      for (int k = 0; k < kdm; k++) C(flxdv, i, k) = C(flxdv, i, k) + E(antiflx, i, e, k);
    for (int k = 0; k < kdm; k++) {
      C(dpn, i, k) = C(dp_lo, i, k) - C(flxdv, i, k) * g->rarea[i] * dtleap;
      if (C(dpn, i, k) < 0. && C(dpn, i, k) + ONECM > 0.) C(dpn, i, k) = 0.;
      P(i, k + 1) = P(i, k) + C(dpn, i, k);
    }
    rpbot[i] = 1. / P(i, kdm);
  }
  /* the clipped fluxes back as an upstream barotropic correction (434-491) */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) C(flxdv, i, k) = 0.;
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++) {
      const int ix = g->prox[i * NPP + e];
      if (!WET(ix)) continue;
      const double cl = clippd[i * NPP + e];
      for (int k = 0; k < kdm; k++) {
        double q = 0.;
        if (cl > 0.) q = cl * (C(dpn, i, k) * rpbot[i]);
        else if (cl < 0.) q = cl * (C(dpn, ix, k) * rpbot[ix]);
        C(flxdv, i, k) = C(flxdv, i, k) + q;
        E(massflx, i, e, k) = E(massflx, i, e, k) + q;
      }
    }
  }
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) {
      C(dpn, i, k) = C(dpn, i, k) - C(flxdv, i, k) * g->rarea[i] * dtleap;
      if (C(dpn, i, k) < 0. && C(dpn, i, k) + ONECM > 0.) C(dpn, i, k) = 0.;
      P(i, k + 1) = P(i, k) + C(dpn, i, k);
    }
  }
  if (g->thkdff != 0.) {
    /* interface smoothing (bolus fluxes, 508-835) */
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      for (int k = 0; k < kdm; k++) C(dpn, i, k) = P(i, k + 1) - P(i, k);
    }
    for (int i = 0; i < nip; i++) {
      for (int k = 0; k < kdm; k++) C(del2p, i, k) = LAND_SPVAL;
      if (g->biharm == 0.)
        for (int k = 1; k < kdm; k++) C(del2p, i, k) = P(i, k);
    }
    if (g->biharm > 0.) {
      for (int i = 0; i < nip; i++) {
        if (!WET(i)) continue;
        for (int k = 0; k < kdm; k++) sum[k] = 0.;
        int ncount = 0;
        for (int e = 0; e < g->nprox[i]; e++) {
          const int ix = g->prox[i * NPP + e];
          if (!WET(ix)) continue;
          ncount = ncount + 1;
          for (int k = 1; k < kdm; k++) {
            double deltap;
            if (P(ix, kdm) < P(i, k)) deltap = P(ix, kdm) - P(ix, k);
            else if (P(i, kdm) < P(ix, k)) deltap = P(i, k) - P(i, kdm);
            else deltap = P(i, k) - P(ix, k);
            sum[k] = sum[k] + deltap;
          }
        }
        for (int k = 1; k < kdm; k++)
          C(del2p, i, k) = g->biharm * sum[k] / (double)ncount + (1. - g->biharm) * P(i, k);
      }
    }
    for (int i = 0; i < nip; i++) {
      for (int e = 0; e < NPP; e++)
        for (int k = 0; k < kdm; k++) E(bolusfx, i, e, k) = WET(i) ? 0. : LAND_SPVAL;
      if (!WET(i)) continue;
      for (int e = 0; e < g->nprox[i]; e++) {
        const int ix = g->prox[i * NPP + e];
        if (!WET(ix)) continue;
        for (int k = 1; k < kdm; k++) {
          const double q = dtleap * g->thkdff * (C(del2p, i, k) - C(del2p, ix, k)) * g->sideln[i * NPP + e];
          const double flxmx = (P(ix, kdm) - P(ix, k)) * g->area[ix] / (double)g->nprox[ix];
          const double flxmn = -((P(i, kdm) - P(i, k)) * g->area[i] / (double)g->nprox[i]);
          E(bolusfx, i, e, k) = dmin2(flxmx * .999, dmax2(flxmn * .999, q));
          E(bolusfx, i, e, k - 1) = E(bolusfx, i, e, k) - E(bolusfx, i, e, k - 1);
        }
        E(bolusfx, i, e, kdm - 1) = -E(bolusfx, i, e, kdm - 1);
      }
    }
    if (g->biharm > 0.) {
      for (int i = 0; i < nip; i++) {
        for (int k = 0; k < kdm; k++) C(flxplus, i, k) = EPSIL;
        if (!WET(i)) continue;
        for (int e = 0; e < g->nprox[i]; e++)
          for (int k = 0; k < kdm; k++) C(flxplus, i, k) = C(flxplus, i, k) + dmax2(0., E(bolusfx, i, e, k));
        for (int k = 0; k < kdm; k++) C(rmnus, i, k) = C(dpn, i, k) * g->area[i] / C(flxplus, i, k);
      }
      for (int i = 0; i < nip; i++) {
        for (int e = 0; e < NPP; e++) clippd[i * NPP + e] = 0.;
        if (!WET(i)) continue;
        for (int e = 0; e < g->nprox[i]; e++) {
          const int ix = g->prox[i * NPP + e];
          if (!WET(ix)) continue;
          for (int k = 0; k < kdm; k++) {
            const double b = E(bolusfx, i, e, k);
            if (b != 0.) {
              const double clip = b > 0. ? dmin2(1., C(rmnus, i, k)) : dmin2(1., C(rmnus, ix, k));
              clippd[i * NPP + e] = clippd[i * NPP + e] + b * (1. - clip);
              E(bolusfx, i, e, k) = b * clip;
            }
          }
        }
      }
    }
    for (int i = 0; i < nip; i++) {
      for (int k = 0; k < kdm; k++) C(flxdv, i, k) = 0.;
      if (!WET(i)) continue;
      for (int e = 0; e < g->nprox[i]; e++)
        for (int k = 0; k < kdm; k++) C(flxdv, i, k) = C(flxdv, i, k) + E(bolusfx, i, e, k);
      for (int k = 0; k < kdm; k++) {
        C(dpn, i, k) = C(dpn, i, k) - C(flxdv, i, k) * g->rarea[i];
        if (C(dpn, i, k) < 0. && C(dpn, i, k) + ONECM > 0.) C(dpn, i, k) = 0.;
        C(dpn, i, k) = C(dpn, i, k) * (1. + ms->masscor);
        P(i, k + 1) = P(i, k) + C(dpn, i, k);
        C(dptmp, i, k) = C(dpn, i, k);
      }
      rpbot[i] = 1. / P(i, kdm);
    }
    if (g->biharm > 0.) {
      for (int i = 0; i < nip; i++) {
        for (int k = 0; k < kdm; k++) C(flxdv, i, k) = 0.;
        if (!WET(i)) continue;
        for (int e = 0; e < g->nprox[i]; e++) {
          const int ix = g->prox[i * NPP + e];
          const double cl = clippd[i * NPP + e];
          double q = 0.;
          for (int k = 0; k < kdm; k++) {
            if (cl > 0.) q = cl * C(dptmp, i, k) * rpbot[i];
            else if (cl < 0.) q = cl * C(dptmp, ix, k) * rpbot[ix];
            C(flxdv, i, k) = C(flxdv, i, k) + q;
            E(massflx, i, e, k) = E(massflx, i, e, k) + q * dtinv;
          }
        }
        for (int k = 0; k < kdm; k++) {
          C(dpn, i, k) = C(dpn, i, k) - C(flxdv, i, k) * g->rarea[i];
          P(i, k + 1) = P(i, k) + C(dpn, i, k);
        }
      }
    }
    /* the instantaneous mass flux into its time average (896-905) */
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      for (int e = 0; e < g->nprox[i]; e++)
        for (int k = 0; k < kdm; k++) E(mssfx_avg, i, e, k) = E(mssfx_avg, i, e, k) + E(massflx, i, e, k);
    }
  }
  /* time smoothing of the mid-time thickness (935-954) */
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++)
      C(dpm, i, k) = ASSLDP * (C(dpold, i, k) + C(dpn, i, k)) + (1. - ASSLDP - ASSLDP) * C(dpm, i, k);
  }
  /* duplicate edge fluxes must be bit-identical up to sign (chkdup) */
  int bad = 0;
  for (int i = 0; i < nip && !bad; i++) {
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++) {
      const int ix = g->prox[i * NPP + e], ex = g->proxs[i * NPP + e];
      if (!WET(ix)) continue;
      for (int k = 0; k < kdm; k++)
        if (E(massflx, i, e, k) + E(massflx, ix, ex, k) != 0. && E(massflx, i, e, k) - E(massflx, ix, ex, k) != 0.)
          bad = 1;
    }
  }
  /* drift correction factor for the total mass (977-998), glob2d in cell order */
  const double day = (double)nstep * .5 * dtleap / 86400.;
  if (nstep == 1 || fmod(day + .0001, 10.) < .0002) {
    double tot = 0.;
    for (int i = 0; i < nip; i++)
      if (WET(i)) tot = tot + P(i, kdm) * g->area[i];
    ms->totmass = tot / GRVITY;
    if (nstep == 1) ms->massglb0 = ms->totmass;
    ms->masscor = dtleap * (ms->massglb0 - ms->totmass) / (ms->massglb0 * 10. * 86400.);
  }
  free(antiflx); free(bolusfx); free(flxdv); free(dp_lo); free(dpold); free(flxplus); free(flxmnus);
  free(rplus); free(rmnus); free(del2p); free(dptmp); free(clippd); free(rpbot); free(aux1); free(sum);
  return bad;
}

/* ------------------------------------------------------------------------
 * hycom_edgvar.F90.  Layouts as above, plus pr_edg (kdm+1,npp,nip) as
 * [(i*6+e)*(kdm+1) + k], utrop/vtrop (3,nip) as [i*3 + slot-1] and
 * cs/sn (4,npp,nip) as [(i*6+e)*4 + j].  The reference's nedge/permedge are
 * nprox and the identity in a serial run. */
static const double SLIP = -1.;               /* hycom_constants.F90:79 */
#define TENM (10. * ONEM)                     /* hycom_constants.F90:36 */
static const double DIVBY18 = 1. / 18.;       /* hycom_edgvar.F90:26 */

/* hycom_edgvar.F90:6-124: rotate the velocities of the edge's two cells (double
 * weight) and of the two joint neighbours into the edge's local frame and
 * interpolate, walls by the slip condition, joint land neighbours as 0 */
static void hyc_uvedge(const fo_hyc_grid *g, const double *cs, const double *sn, const double *uvel,
                       const double *vvel, double *u_edg, double *v_edg) {
  const int nip = g->nip, kdm = g->kdm;
  for (int i = 0; i < nip; i++) {
    for (int e = 0; e < NPP; e++)
      for (int k = 0; k < kdm; k++) E(u_edg, i, e, k) = E(v_edg, i, e, k) = 0.;
    const int np = g->nprox[i];
    for (int edg = 0; edg < np; edg++) {
      const int ipx = g->prox[i * NPP + edg];
      const int im1 = g->prox[i * NPP + (edg - 1 + np) % np];
      const int ip1 = g->prox[i * NPP + (edg + 1) % np];
      const double *c = cs + (size_t)(i * NPP + edg) * 4, *s = sn + (size_t)(i * NPP + edg) * 4;
      for (int k = 0; k < kdm; k++) {
        double u1 = c[0] * C(uvel, i, k) + s[0] * C(vvel, i, k);
        double v1 = -s[0] * C(uvel, i, k) + c[0] * C(vvel, i, k);
        double u2, v2, u3, v3, u4, v4;
        if (WET(ipx)) {
          u2 = c[1] * C(uvel, ipx, k) + s[1] * C(vvel, ipx, k);
          v2 = -s[1] * C(uvel, ipx, k) + c[1] * C(vvel, ipx, k);
          if (g->wet[i] == 0) {
            u1 = u2 * SLIP;
            v1 = v2 * SLIP;
          }
        } else {
          u2 = u1 * SLIP;
          v2 = v1 * SLIP;
        }
        if (WET(im1)) {
          u3 = c[2] * C(uvel, im1, k) + s[2] * C(vvel, im1, k);
          v3 = -s[2] * C(uvel, im1, k) + c[2] * C(vvel, im1, k);
        } else {
          u3 = v3 = 0.;
        }
        if (WET(ip1)) {
          u4 = c[3] * C(uvel, ip1, k) + s[3] * C(vvel, ip1, k);
          v4 = -s[3] * C(uvel, ip1, k) + c[3] * C(vvel, ip1, k);
        } else {
          u4 = v4 = 0.;
        }
        E(u_edg, i, edg, k) = (8. * (u1 + u2) + u3 + u4) * DIVBY18;
        E(v_edg, i, edg, k) = (8. * (v1 + v2) + v3 + v4) * DIVBY18;
      }
    }
  }
}

/* hycom_edgvar.F90:479-493: the extent to which a sidewall is present */
static double hyc_blokf(double dp, double plo, double pbx, double pupx) {
  (void)pupx;
  return dmax2(0., dmin2(1., (plo - pbx) / dmax2(EPSIL, dp)));
}

void fo_hycom_edgvar(const fo_hyc_grid *g, const double *cs, const double *sn, int janjic, int leapm, int leapn,
                     const double *uclin, const double *vclin, const double *utrop, const double *vtrop,
                     const double *dp, double *pres, const double *geop, const double *montg,
                     const double *spcifv, double *um_edg, double *vm_edg, double *un_edg, double *vn_edg,
                     double *dp_edg, double *pr_edg, double *geop_edg, double *mont_edg, double *uvsq_edg,
                     double *spcv_edg) {
  const int nip = g->nip, kdm = g->kdm, kp = kdm + 1;
  const size_t cz = (size_t)nip * kdm;
  const double *dpm = dp + (size_t)(leapm - 1) * cz;
  double *utotm = malloc(cz * sizeof(double)), *utotn = malloc(cz * sizeof(double));
  double *vtotm = malloc(cz * sizeof(double)), *vtotn = malloc(cz * sizeof(double));
  double *geop_lyr = malloc(cz * sizeof(double)), *geopx = malloc(cz * NPP * sizeof(double));
  double *p_lyr = malloc((size_t)kdm * sizeof(double));
#define GEO(i, k) geop[(size_t)(i) * kp + (k)]
#define PR(i, e, k) pr_edg[((size_t)(i) * NPP + (e)) * kp + (k)]
  if (janjic) {
    /* hydrostatic integration to the layer's mid pressure in the neighbouring
     * columns, down to the shallowest neighbour at most (200-265) */
    for (int i = 0; i < nip; i++) {
      for (int k = 0; k < kdm; k++) {
        C(geop_lyr, i, k) = LAND_SPVAL;
        for (int e = 0; e < NPP; e++) E(geopx, i, e, k) = LAND_SPVAL;
      }
      if (!WET(i)) continue;
      const int np = g->nprox[i];
      for (int k = 0; k < kdm; k++) {
        p_lyr[k] = .5 * (P(i, k) + P(i, k + 1));
        C(geop_lyr, i, k) = .5 * (GEO(i, k) + GEO(i, k + 1));
      }
      double silhou = 1.e33;
      for (int e = 0; e < np; e++) {
        for (int k = 0; k < kdm; k++) E(geopx, i, e, k) = C(geop_lyr, i, k);
        const int ipx = g->prox[i * NPP + e];
        if (WET(ipx)) silhou = dmin2(silhou, P(ipx, kdm));
      }
      for (int e = 0; e < np; e++) {
        const int ipx = g->prox[i * NPP + e];
        if (!WET(ipx)) continue;
        for (int k = 0; k < kdm; k++) {
          const double plyrk = dmin2(p_lyr[k], silhou);
          int kx = 0;
          while (kx < kdm - 1 && P(ipx, kx + 1) < plyrk) kx++;
          if (P(ipx, kx + 1) >= plyrk) {
            const double thav = C(spcifv, ipx, kx);
            E(geopx, i, e, k) = GEO(ipx, kx) - thav * (plyrk - P(ipx, kx));
          }
        }
      }
      /* nudge the pressure force to zero below the silhouette depth */
      for (int k = 0; k < kdm; k++) {
        const double tapr = dmax2(0., dmin2(1., (p_lyr[k] - silhou) / (100. * ONEM)));
        for (int e = 0; e < np; e++) E(geopx, i, e, k) = (1. - tapr) * E(geopx, i, e, k) + tapr * C(geop_lyr, i, k);
      }
    }
  }
  /* total velocities and the interface pressure at mid-time (273-291) */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) C(utotm, i, k) = C(utotn, i, k) = C(vtotm, i, k) = C(vtotn, i, k) = LAND_SPVAL;
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) {
      C(utotm, i, k) = C(uclin + (size_t)(leapm - 1) * cz, i, k) + utrop[i * 3 + leapm - 1];
      C(utotn, i, k) = C(uclin + (size_t)(leapn - 1) * cz, i, k) + utrop[i * 3 + leapn - 1];
      C(vtotm, i, k) = C(vclin + (size_t)(leapm - 1) * cz, i, k) + vtrop[i * 3 + leapm - 1];
      C(vtotn, i, k) = C(vclin + (size_t)(leapn - 1) * cz, i, k) + vtrop[i * 3 + leapn - 1];
      P(i, k + 1) = P(i, k) + C(dpm, i, k);
    }
  }
  hyc_uvedge(g, cs, sn, utotm, vtotm, um_edg, vm_edg);
  hyc_uvedge(g, cs, sn, utotn, vtotn, un_edg, vn_edg);
  /* layer thickness, interface pressure, Montgomery potential, kinetic energy
   * and specific volume on the edges (300-418) */
  for (int i = 0; i < nip; i++) {
    for (int e = 0; e < NPP; e++) {
      for (int k = 0; k < kdm; k++)
        E(dp_edg, i, e, k) = E(mont_edg, i, e, k) = E(uvsq_edg, i, e, k) = E(spcv_edg, i, e, k) = 0.;
      for (int k = 0; k < kp; k++) PR(i, e, k) = 0.;
    }
    if (!WET(i)) continue;
    const int np = g->nprox[i];
    for (int edg = 0; edg < np; edg++) {
      int ipx = g->prox[i * NPP + edg];
      if (WET(ipx)) PR(i, edg, kdm) = dmin2(P(i, kdm), P(ipx, kdm));
      else ipx = i;                               /* avoid stepping on land */
      int im1 = g->prox[i * NPP + (edg - 1 + np) % np];
      if (g->wet[im1] == 0) im1 = i;
      int ip1 = g->prox[i * NPP + (edg + 1) % np];
      if (g->wet[ip1] == 0) ip1 = i;
      for (int k = kdm - 1; k >= 1; k--)
        PR(i, edg, k) = PR(i, edg, k + 1) - .5 * dmin2(dmin2(TENM, C(dpm, i, k)), C(dpm, ipx, k));
      for (int k = 0; k < kdm; k++) {
        if (k < kdm - 1)
          PR(i, edg, k + 1) = dmin2(PR(i, edg, k + 1), PR(i, edg, k) + (8. * (C(dpm, i, k) + C(dpm, ipx, k)) +
                                                                      C(dpm, im1, k) + C(dpm, ip1, k)) * DIVBY18);
        E(dp_edg, i, edg, k) = dmax2(0., PR(i, edg, k + 1) - PR(i, edg, k));
      }
      ipx = g->prox[i * NPP + edg];
      im1 = g->prox[i * NPP + (edg - 1 + np) % np];
      ip1 = g->prox[i * NPP + (edg + 1) % np];
      for (int k = 0; k < kdm; k++) {
        double blok = WET(ipx) ? hyc_blokf(C(dpm, i, k), P(i, k + 1), P(ipx, kdm), P(ipx, k)) : 1.;
        const double ngborm1 = C(montg, ipx, k) * (1. - blok) + C(montg, i, k) * blok;
        const double ngborv1 = C(spcifv, ipx, k) * (1. - blok) + C(spcifv, i, k) * blok;
        blok = WET(im1) ? hyc_blokf(C(dpm, i, k), P(i, k + 1), P(im1, kdm), P(im1, k)) : 1.;
        const double ngborm2 = C(montg, im1, k) * (1. - blok) + C(montg, i, k) * blok;
        const double ngborv2 = C(spcifv, im1, k) * (1. - blok) + C(spcifv, i, k) * blok;
        blok = WET(ip1) ? hyc_blokf(C(dpm, i, k), P(i, k + 1), P(ip1, kdm), P(ip1, k)) : 1.;
        const double ngborm3 = C(montg, ip1, k) * (1. - blok) + C(montg, i, k) * blok;
        const double ngborv3 = C(spcifv, ip1, k) * (1. - blok) + C(spcifv, i, k) * blok;
        E(mont_edg, i, edg, k) = (8. * (C(montg, i, k) + ngborm1) + ngborm2 + ngborm3) * DIVBY18;
        E(uvsq_edg, i, edg, k) = .5 * (E(um_edg, i, edg, k) * E(um_edg, i, edg, k) +
                                       E(vm_edg, i, edg, k) * E(vm_edg, i, edg, k));
        E(spcv_edg, i, edg, k) = (8. * (C(spcifv, i, k) + ngborv1) + ngborv2 + ngborv3) * DIVBY18;
      }
    }
  }
  if (janjic) {
    /* geopotential on the edges (437-462) */
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      const int np = g->nprox[i];
      for (int edg = 0; edg < np; edg++) {
        const int em1 = (edg - 1 + np) % np, ep1 = (edg + 1) % np;
        for (int k = 0; k < kdm; k++)
          E(geop_edg, i, edg, k) =
              (8. * (C(geop_lyr, i, k) + E(geopx, i, edg, k)) + E(geopx, i, em1, k) + E(geopx, i, ep1, k)) * DIVBY18;
      }
    }
  }
#undef GEO
#undef PR
  free(utotm); free(utotn); free(vtotm); free(vtotn); free(geop_lyr); free(geopx); free(p_lyr);
}

/* ------------------------------------------------------------------------
 * hycom_momtum.F90 with hycom_dissip.F90.  Layouts as above; uclin/vclin
 * (kdm,nip,2) as [(s*nip + i)*kdm + k], sidevec_c like sidevec_e. */
#define MGRVITY 9.80665                        /* module_constants grvity */
#define ONEMM (0.001 * ONEM)                   /* hycom_constants.F90:39 */
#define TENCM (0.1 * ONEM)                     /* hycom_constants.F90:37 */
static const double DRCOEF = 3.e-3;           /* hycom_constants.F90:44 */
static const double ASSLUV = .25;             /* hycom_constants.F90:48 */
static const double THSHLD = 1.e-11;          /* hycom_dissip.F90:145 */
static const double OMEGX2 = 1.4584e-4;       /* module_constants.F90:18 */

void fo_hycom_momtum(const fo_hyc_grid *g, const fo_hyc_mom *m, int leapm, int leapn, const double *utrop,
                     const double *vtrop, double *uclin, double *vclin, const double *um_edg, const double *vm_edg,
                     const double *geop_edg, const double *mont_edg, const double *uvsq_edg,
                     const double *spcv_edg, const double *taux, const double *tauy, const double *dp,
                     double *pres, double *ustarb) {
  const int nip = g->nip, kdm = g->kdm, kp = kdm + 1;
  const size_t cz = (size_t)nip * kdm;
  const double dtleap = g->dtleap;
  const double thktop = 50. * ONEM, thkbot = 10. * ONEM, cbar = .5;       /* hycom_momtum.F90:87-91 */
  const double *dpm = dp + (size_t)(leapm - 1) * cz, *dpn = dp + (size_t)(leapn - 1) * cz;
  double *ucm = uclin + (size_t)(leapm - 1) * cz, *ucn = uclin + (size_t)(leapn - 1) * cz;
  double *vcm = vclin + (size_t)(leapm - 1) * cz, *vcn = vclin + (size_t)(leapn - 1) * cz;
  double *gradx = calloc(cz, sizeof(double)), *grady = calloc(cz, sizeof(double));
  double *utotn = calloc(cz, sizeof(double)), *vtotn = calloc(cz, sizeof(double));
  double *worku = malloc(cz * sizeof(double)), *workv = malloc(cz * sizeof(double));
  double *d4u = malloc(cz * sizeof(double)), *d4v = malloc(cz * sizeof(double));
  double *ufx = malloc(cz * NPP * sizeof(double)), *vfx = malloc(cz * NPP * sizeof(double));
  double *dpres = malloc((size_t)nip * kp * sizeof(double)), *drag = calloc((size_t)nip, sizeof(double));
  double *pgfx = malloc((size_t)kdm * sizeof(double)), *pgfy = malloc((size_t)kdm * sizeof(double));
  double *avgu = malloc((size_t)kdm * sizeof(double)), *avgv = malloc((size_t)kdm * sizeof(double));
  double *dudx = malloc((size_t)kdm * sizeof(double)), *dudy = malloc((size_t)kdm * sizeof(double));
  double *dvdx = malloc((size_t)kdm * sizeof(double)), *dvdy = malloc((size_t)kdm * sizeof(double));
  double *uold = malloc((size_t)kdm * sizeof(double)), *vold = malloc((size_t)kdm * sizeof(double));
#define CSN(a, i, e, j) (a)[((size_t)(i) * NPP + (e)) * 4 + (j)]
#define SVC(i, e, c) m->sidevec_c[((size_t)(i) * NPP + (e)) * 2 + (c)]
#define DPR(i, k) dpres[(size_t)(i) * kp + (k)]
  /* bottom drag, pressure force (hycom_momtum.F90:104-183) */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) C(worku, i, k) = C(workv, i, k) = LAND_SPVAL;
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) P(i, k + 1) = P(i, k) + C(dpn, i, k);
    double ubot = 0., vbot = 0.;
    for (int k = 0; k < kdm; k++) {
      C(utotn, i, k) = C(ucn, i, k) + utrop[i * 3 + leapn - 1];
      C(vtotn, i, k) = C(vcn, i, k) + vtrop[i * 3 + leapn - 1];
      const double q = dmin2(thkbot, P(i, kdm));
      ubot = ubot + C(utotn, i, k) * (dmin2(P(i, k + 1), P(i, kdm) - q) - dmin2(P(i, k), P(i, kdm) - q));
      vbot = vbot + C(vtotn, i, k) * (dmin2(P(i, k + 1), P(i, kdm) - q) - dmin2(P(i, k), P(i, kdm) - q));
    }
    const double botvel = sqrt(ubot * ubot + vbot * vbot) / thkbot;
    drag[i] = dmin2(DRCOEF * (botvel + cbar) * ONEM / thkbot, 1. / m->batrop);
    ustarb[i] = sqrt(DRCOEF) * botvel;
    for (int k = 0; k < kdm; k++) pgfx[k] = pgfy[k] = 0.;
    for (int e = 0; e < g->nprox[i]; e++)
      for (int k = 0; k < kdm; k++) {
        if (m->janjic) {
          pgfx[k] = pgfx[k] - (E(geop_edg, i, e, k) + E(uvsq_edg, i, e, k)) * SVC(i, e, 1);
          pgfy[k] = pgfy[k] + (E(geop_edg, i, e, k) + E(uvsq_edg, i, e, k)) * SVC(i, e, 0);
        } else {
          P(i, k + 1) = P(i, k) + C(dpm, i, k);
          pgfx[k] = pgfx[k] - (E(mont_edg, i, e, k) + E(uvsq_edg, i, e, k)) * SVC(i, e, 1);
          pgfy[k] = pgfy[k] + (E(mont_edg, i, e, k) + E(uvsq_edg, i, e, k)) * SVC(i, e, 0);
          pgfx[k] = pgfx[k] + .5 * (P(i, k) + P(i, k + 1)) * E(spcv_edg, i, e, k) * SVC(i, e, 1);
          pgfy[k] = pgfy[k] - .5 * (P(i, k) + P(i, k + 1)) * E(spcv_edg, i, e, k) * SVC(i, e, 0);
        }
      }
    for (int k = 0; k < kdm; k++) {
      C(gradx, i, k) = pgfx[k] * g->rarea[i];
      C(grady, i, k) = pgfy[k] * g->rarea[i];
      C(worku, i, k) = C(ucn, i, k);
      C(workv, i, k) = C(vcn, i, k);
    }
  }
  /* del4prep (hycom_dissip.F90:9-110): minus the mean of the neighbours,
   * rotated into the cell's frame through the edge's, walls by the slip
   * condition */
  memcpy(d4u, worku, cz * sizeof(double));
  memcpy(d4v, workv, cz * sizeof(double));
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    const int np = g->nprox[i];
    for (int k = 0; k < kdm; k++) avgu[k] = avgv[k] = 0.;
    for (int e = 0; e < np; e++) {
      const int ix = g->prox[i * NPP + e];
      for (int k = 0; k < kdm; k++) {
        double ungbor, vngbor;
        if (WET(ix)) {
          const double uedge = CSN(m->cs, i, e, 1) * C(worku, ix, k) + CSN(m->sn, i, e, 1) * C(workv, ix, k);
          const double vedge = -CSN(m->sn, i, e, 1) * C(worku, ix, k) + CSN(m->cs, i, e, 1) * C(workv, ix, k);
          ungbor = CSN(m->cs, i, e, 0) * uedge - CSN(m->sn, i, e, 0) * vedge;
          vngbor = CSN(m->sn, i, e, 0) * uedge + CSN(m->cs, i, e, 0) * vedge;
        } else {
          ungbor = SLIP * C(worku, i, k);
          vngbor = SLIP * C(workv, i, k);
        }
        avgu[k] = avgu[k] + ungbor;
        avgv[k] = avgv[k] + vngbor;
      }
    }
    for (int k = 0; k < kdm; k++) {
      const double factor = 1. / np;
      C(d4u, i, k) = C(worku, i, k) - avgu[k] * factor;
      C(d4v, i, k) = C(workv, i, k) - avgv[k] * factor;
    }
  }
  /* dissip (hycom_dissip.F90:113-272): thickness-weighted momentum fluxes
   * through the edges, walls blended in by blokf, summed into a tendency */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kp; k++) DPR(i, k) = LAND_SPVAL;
    if (!WET(i)) continue;
    DPR(i, 0) = 0.;
    for (int k = 0; k < kdm; k++) DPR(i, k + 1) = DPR(i, k) + C(dpn, i, k);
  }
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++) {
      const int ix = g->prox[i * NPP + e];
      for (int k = 0; k < kdm; k++) {
        const double urot = CSN(m->cs, i, e, 0) * C(d4u, i, k) + CSN(m->sn, i, e, 0) * C(d4v, i, k);
        const double vrot = -CSN(m->sn, i, e, 0) * C(d4u, i, k) + CSN(m->cs, i, e, 0) * C(d4v, i, k);
        double ungbor, vngbor;
        if (WET(ix)) {
          ungbor = CSN(m->cs, i, e, 1) * C(d4u, ix, k) + CSN(m->sn, i, e, 1) * C(d4v, ix, k);
          vngbor = -CSN(m->sn, i, e, 1) * C(d4u, ix, k) + CSN(m->cs, i, e, 1) * C(d4v, ix, k);
          const double blokd = hyc_blokf(C(dpn, i, k), DPR(i, k + 1), DPR(ix, kdm), DPR(ix, k));
          ungbor = (1. - blokd) * ungbor + blokd * urot * SLIP;
          vngbor = (1. - blokd) * vngbor + blokd * vrot * SLIP;
        } else {
          ungbor = urot * SLIP;
          vngbor = vrot * SLIP;
        }
        const double a = dmax2(C(dpn, i, k), THSHLD), b = dmax2(C(dpn, ix, k), THSHLD);
        const double hh = a * b / (a + b);
        E(ufx, i, e, k) = (ungbor - urot) * g->sideln[i * NPP + e] * 2. * hh;
        E(vfx, i, e, k) = (vngbor - vrot) * g->sideln[i * NPP + e] * 2. * hh;
      }
    }
  }
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) C(worku, i, k) = C(workv, i, k) = 0.;
    for (int e = 0; e < g->nprox[i]; e++)
      for (int k = 0; k < kdm; k++) {
        const double factor = g->rarea[i] / dmax2(C(dpn, i, k), THSHLD);
        const double uflux = CSN(m->cs, i, e, 0) * E(ufx, i, e, k) - CSN(m->sn, i, e, 0) * E(vfx, i, e, k);
        const double vflux = CSN(m->sn, i, e, 0) * E(ufx, i, e, k) + CSN(m->cs, i, e, 0) * E(vfx, i, e, k);
        C(worku, i, k) = C(worku, i, k) + uflux * factor;
        C(workv, i, k) = C(workv, i, k) + vflux * factor;
      }
  }
  /* vorticity, deformation, dissipation, Coriolis force, stresses, the
   * leapfrog update, downward mixing and time smoothing (211-389) */
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) dudx[k] = dudy[k] = dvdx[k] = dvdy[k] = 0.;
    for (int e = 0; e < g->nprox[i]; e++)
      for (int k = 0; k < kdm; k++) {
        dudx[k] = dudx[k] + SVE(i, e, 1) * E(um_edg, i, e, k);
        dudy[k] = dudy[k] - SVE(i, e, 0) * E(um_edg, i, e, k);
        dvdx[k] = dvdx[k] + SVE(i, e, 1) * E(vm_edg, i, e, k);
        dvdy[k] = dvdy[k] - SVE(i, e, 0) * E(vm_edg, i, e, k);
      }
    for (int k = 0; k < kdm; k++) {
      dudx[k] = dudx[k] * g->rarea[i];
      dudy[k] = dudy[k] * g->rarea[i];
      dvdx[k] = dvdx[k] * g->rarea[i];
      dvdy[k] = dvdy[k] * g->rarea[i];
    }
    for (int k = 0; k < kdm; k++) {
      const double relvor = dvdx[k] - dudy[k];
      const double a = dudx[k] - dvdy[k], b = dudy[k] + dvdx[k];
      const double deform = sqrt(a * a + b * b);
      const double smag = deform * sqrt(g->area[i]) * m->smagcf;
      const double udissp = C(worku, i, k) * dmax2(m->veldff, smag);
      const double vdissp = C(workv, i, k) * dmax2(m->veldff, smag);
      const double corio = OMEGX2 * sin(m->lat[i]);
      const double uzeta = C(ucm, i, k) * (relvor + corio) + utrop[i * 3 + leapm - 1] * relvor;
      const double vzeta = C(vcm, i, k) * (relvor + corio) + vtrop[i * 3 + leapm - 1] * relvor;
      double ustres = taux[i] * MGRVITY / thktop * (dmin2(P(i, k + 1) + ONEMM, thktop) - dmin2(P(i, k), thktop)) /
                      dmax2(C(dpn, i, k), ONEMM);
      double vstres = tauy[i] * MGRVITY / thktop * (dmin2(P(i, k + 1) + ONEMM, thktop) - dmin2(P(i, k), thktop)) /
                      dmax2(C(dpn, i, k), ONEMM);
      const double q = (dmax2(P(i, kdm) - thkbot, P(i, k + 1)) -
                        dmax2(P(i, kdm) - thkbot, dmin2(P(i, k + 1) - ONEMM, P(i, k)))) / dmax2(C(dpn, i, k), ONEMM);
      ustres = ustres - drag[i] * C(utotn, i, k) * q;
      vstres = vstres - drag[i] * C(vtotn, i, k) * q;
      uold[k] = C(ucn, i, k);
      vold[k] = C(vcn, i, k);
      C(ucn, i, k) = C(ucn, i, k) + dtleap * (C(gradx, i, k) + vzeta + udissp + ustres);
      C(vcn, i, k) = C(vcn, i, k) + dtleap * (C(grady, i, k) - uzeta + vdissp + vstres);
      if (k > 0) {            /* downward momentum mixing into thin bottom layers */
        const double slab = dmax2(TENCM, P(i, k + 1) - P(i, kdm) + TENM);
        const double qq = dmin2(C(dpn, i, k), slab);
        C(ucn, i, k) = (C(ucn, i, k) * qq + C(ucn, i, k - 1) * (slab - qq)) / slab;
        C(vcn, i, k) = (C(vcn, i, k) * qq + C(vcn, i, k - 1) * (slab - qq)) / slab;
      }
    }
    for (int k = 0; k < kdm; k++) {
      C(ucm, i, k) = ASSLUV * (uold[k] + C(ucn, i, k)) + (1. - ASSLUV - ASSLUV) * C(ucm, i, k);
      C(vcm, i, k) = ASSLUV * (vold[k] + C(vcn, i, k)) + (1. - ASSLUV - ASSLUV) * C(vcm, i, k);
    }
  }
#undef CSN
#undef SVC
#undef DPR
  free(gradx); free(grady); free(utotn); free(vtotn); free(worku); free(workv); free(d4u); free(d4v);
  free(ufx); free(vfx); free(dpres); free(drag); free(pgfx); free(pgfy); free(avgu); free(avgv);
  free(dudx); free(dudy); free(dvdx); free(dvdy); free(uold); free(vold);
}

/* ------------------------------------------------------------------------
 * hycom_transp3d.F90, hycom_fct3d.F90, hycom_dffusn.F90 */
static const int BCLIN_FRQ = 30;              /* hycom_control.F90:9 */
static const double TOFFSET = 10.;            /* hycom_transp3d.F90: keeps temperature positive */

/* hycom_transp3d.F90 transp0 */
void fo_hycom_transp0(const fo_hyc_grid *g, int leapn, const double *dp, double *dpinit, double *cumuflx) {
  const int nip = g->nip, kdm = g->kdm;
  const double *dpn = dp + (size_t)(leapn - 1) * nip * kdm;
  for (int i = 0; i < nip; i++) {
    for (int e = 0; e < NPP; e++)
      for (int k = 0; k < kdm; k++) E(cumuflx, i, e, k) = LAND_SPVAL;
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) {
      for (int e = 0; e < NPP; e++) E(cumuflx, i, e, k) = 0.;
      C(dpinit, i, k) = C(dpn, i, k);
    }
  }
}

/* hycom_transp3d.F90 transp1 */
void fo_hycom_transp1(const fo_hyc_grid *g, double batrop, const double *massflx, double *cumuflx) {
  const int nip = g->nip, kdm = g->kdm;
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++)
      for (int k = 0; k < kdm; k++) E(cumuflx, i, e, k) = E(cumuflx, i, e, k) + E(massflx, i, e, k) * 2. * batrop;
  }
}

/* hycom_fct3d.F90 fct3d(nstep, fld3d, numfld, u, w, area, rarea, thko, thkn,
 * diagno) with recovr = .true.: each field advected by the time-integrated
 * mass fluxes u (horizontal) and w (vertical, positive up), low order plus
 * limited antidiffusive fluxes, then shifted so that its mass-weighted mean
 * over the ocean is what it was (the sums serial over cells, as written) */
static void hyc_fct3d(const fo_hyc_grid *g, double *fld3d, int numfld, const double *u, const double *w,
                      const double *thko, const double *thkn) {
  const int nip = g->nip, kdm = g->kdm;
  const size_t cz = (size_t)nip * kdm;
  const double athird = 1. / 3., epsil = 1.e-11, onemu = 1.e-6;
  double *fld = malloc(cz * sizeof(double)), *vertfx = malloc(cz * sizeof(double));
  double *vertdv = malloc(cz * sizeof(double)), *fmx = malloc(cz * sizeof(double));
  double *fmn = malloc(cz * sizeof(double)), *flp = malloc(cz * sizeof(double));
  double *fln = malloc(cz * sizeof(double)), *flx = malloc(cz * NPP * sizeof(double));
  double *uan = malloc(cz * NPP * sizeof(double));
  double *a = malloc((size_t)kdm * sizeof(double)), *b = malloc((size_t)kdm * sizeof(double));
  double *c = malloc((size_t)kdm * sizeof(double));
  for (int type = 0; type < numfld; type++) {
    double *f3 = fld3d + (size_t)type * cz;
    for (int i = 0; i < nip; i++)
      for (int k = 0; k < kdm; k++) C(fld, i, k) = WET(i) ? C(f3, i, k) : LAND_SPVAL;
    /* piecewise-parabolic fits in each layer and the vertical fluxes they give
     * (hycom_fct3d.F90:135-288) */
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      a[0] = C(fld, i, 0);
      b[0] = c[0] = 0.;
      a[kdm - 1] = C(fld, i, kdm - 1);
      b[kdm - 1] = c[kdm - 1] = 0.;
      for (int k = 1; k < kdm - 1; k++) {
        const double dxlft = dmax2(epsil, C(thko, i, k - 1)), dxmid = dmax2(epsil, C(thko, i, k));
        const double dxrgt = dmax2(epsil, C(thko, i, k + 1));
        const double fk = C(fld, i, k);
        const double yl = (dxlft * fk + dxmid * C(fld, i, k - 1)) / (dxlft + dxmid);
        const double yr = (dxrgt * fk + dxmid * C(fld, i, k + 1)) / (dxrgt + dxmid);
        a[k] = 1.5 * fk - .25 * (yl + yr);
        b[k] = yr - yl;
        c[k] = 6. * (.5 * (yl + yr) - fk);
        if (fabs(yr - yl) < 6. * fabs(.5 * (yl + yr) - fk)) {
          if (fabs(yr - yl) > 2. * fabs(.5 * (yl + yr) - fk)) {
            if ((yr - yl) * (.5 * (yl + yr) - fk) > 0.) {
              a[k] = .25 * (3. * fk + yl);
              c[k] = 3. * (fk - yl);
              b[k] = c[k];
            } else {
              a[k] = .25 * (3. * fk + yr);
              c[k] = 3. * (fk - yr);
              b[k] = -c[k];
            }
          } else {                 /* -1/6 < x < +1/6 */
            a[k] = fk;
            b[k] = c[k] = 0.;
          }
        }
      }
      double amount = 0.;          /* (kept from the interface above where w = 0, as written) */
      for (int k = 0; k < kdm - 1; k++) {
        double slab = onemu;
        const double wk = C(w, i, k);
        if (wk < 0.) {             /* the interface moves down: take from the layers below */
          amount = slab * C(fld, i, k + 1);
          for (int kp = k + 1; kp < kdm; kp++) {
            if (slab >= -wk) break;
            if (C(thko, i, kp) > 0.) {
              const double dslab = dmin2(slab + C(thko, i, kp), -wk) - dmin2(slab, -wk);
              const double dx = dslab / C(thko, i, kp);
              const double fcdx = a[kp] + b[kp] * .5 * (dx - 1.) + c[kp] * (.25 - dx * (.5 - dx * athird));
              amount = amount + fcdx * dslab;
              slab = slab + dslab;
            }
          }
        } else if (wk > 0.) {      /* the interface moves up: take from the layers above */
          amount = slab * C(fld, i, k);
          /* the reference's loop tests its end after each layer (goto 25 while
           * kp > 2, 1-based): it takes layer k first, whatever k is, and goes on
           * up to the second layer */
          int kp = k;
          for (;;) {
            if (slab >= wk) break;
            if (C(thko, i, kp) > 0.) {
              const double dslab = dmin2(slab + C(thko, i, kp), wk) - dmin2(slab, wk);
              const double dx = dslab / C(thko, i, kp);
              const double fcdx = a[kp] + b[kp] * .5 * (1. - dx) + c[kp] * (.25 - dx * (.5 - dx * athird));
              amount = amount + fcdx * dslab;
              slab = slab + dslab;
            }
            if (kp > 1) kp--;
            else break;
          }
        }
        C(vertfx, i, k) = wk * amount / slab;
      }
      C(vertfx, i, kdm - 1) = 0.;  /* no flux through the bottom */
      C(vertdv, i, 0) = C(vertfx, i, 0);
      for (int k = 1; k < kdm; k++) C(vertdv, i, k) = C(vertfx, i, k) - C(vertfx, i, k - 1);
    }
    /* the mass-weighted mean before (recovr) */
    double bfore = 0., wtold = 0.;
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      double col1 = 0., col2 = 0.;
      for (int k = 0; k < kdm; k++) {
        col1 = col1 + C(thko, i, k) * g->area[i];
        col2 = col2 + C(fld, i, k) * C(thko, i, k) * g->area[i];
      }
      wtold = wtold + col1;
      bfore = bfore + col2;
    }
    /* low-order and antidiffusive fluxes, local bounds (hycom_fct3d.F90:311-379) */
    for (int i = 0; i < nip; i++) {
      for (int e = 0; e < NPP; e++)
        for (int k = 0; k < kdm; k++) E(flx, i, e, k) = E(uan, i, e, k) = LAND_SPVAL;
      if (!WET(i)) continue;
      for (int k = 0; k < kdm; k++) {
        C(fmx, i, k) = C(fmn, i, k) = C(fld, i, k);
        for (int e = 0; e < g->nprox[i]; e++) {
          E(flx, i, e, k) = E(uan, i, e, k) = 0.;
          const int ix = g->prox[i * NPP + e];
          if (!WET(ix)) continue;
          const int edx = g->proxs[i * NPP + e];
          E(flx, i, e, k) = 0.5 * ((E(u, i, e, k) + fabs(E(u, i, e, k))) * C(fld, i, k) -
                                   (E(u, ix, edx, k) + fabs(E(u, ix, edx, k))) * C(fld, ix, k));
          const double q = .5 * (C(fld, i, k) + C(fld, ix, k));
          E(uan, i, e, k) = q * E(u, i, e, k) - E(flx, i, e, k);
          C(fmx, i, k) = dmax2(C(fmx, i, k), C(fld, ix, k));
          C(fmn, i, k) = dmin2(C(fmn, i, k), C(fld, ix, k));
        }
        if (k < kdm - 1 && C(w, i, k) < 0.) {
          C(fmx, i, k) = dmax2(C(fmx, i, k), C(vertfx, i, k) / C(w, i, k));
          C(fmn, i, k) = dmin2(C(fmn, i, k), C(vertfx, i, k) / C(w, i, k));
        }
        if (k > 0 && C(w, i, k - 1) > 0.) {
          C(fmx, i, k) = dmax2(C(fmx, i, k), C(vertfx, i, k - 1) / C(w, i, k - 1));
          C(fmn, i, k) = dmin2(C(fmn, i, k), C(vertfx, i, k - 1) / C(w, i, k - 1));
        }
      }
    }
    /* the low-order update (380-388) */
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      for (int k = 0; k < kdm; k++) {
        double hordiv = 0.;
        for (int e = 0; e < g->nprox[i]; e++) hordiv = hordiv + E(flx, i, e, k);
        hordiv = hordiv * g->rarea[i];
        const double unbound = C(fld, i, k) * C(thko, i, k) - hordiv - C(vertdv, i, k);
        const double bounded =
            dmax2(dmax2(0., C(fmn, i, k) * C(thkn, i, k)), dmin2(unbound, C(fmx, i, k) * C(thkn, i, k)));
        C(fld, i, k) = (C(fld, i, k) * onemu + bounded) / (onemu + C(thkn, i, k));
      }
    }
    /* limiter ratios with the bounds widened by the low-order neighbours (389-436) */
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      for (int k = 0; k < kdm; k++) {
        double totin = 0., totou = 0.;
        for (int e = 0; e < g->nprox[i]; e++) {
          totin = totin - dmin2(0., E(uan, i, e, k));
          totou = totou + dmax2(0., E(uan, i, e, k));
          const int ix = g->prox[i * NPP + e];
          if (WET(ix)) {
            C(fmx, i, k) = dmax2(C(fmx, i, k), C(fld, ix, k));
            C(fmn, i, k) = dmin2(C(fmn, i, k), C(fld, ix, k));
          }
        }
        C(flp, i, k) = (C(fmx, i, k) - C(fld, i, k)) * C(thkn, i, k) / (totin + epsil) * g->rarea[i];
        C(fln, i, k) = (C(fld, i, k) - C(fmn, i, k)) * C(thkn, i, k) / (totou + epsil) * g->rarea[i];
      }
    }
    /* the limited antidiffusive update (437-481) */
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      for (int k = 0; k < kdm; k++) {
        double hordiv = 0.;
        for (int e = 0; e < g->nprox[i]; e++) {
          const int ix = g->prox[i * NPP + e];
          if (!WET(ix)) continue;
          const double clip = E(uan, i, e, k) >= 0. ? dmin2(dmin2(1., C(fln, i, k)), C(flp, ix, k))
                                                    : dmin2(dmin2(1., C(flp, i, k)), C(fln, ix, k));
          E(flx, i, e, k) = E(uan, i, e, k) * clip;
          hordiv = hordiv + E(flx, i, e, k);
        }
        hordiv = hordiv * g->rarea[i];
        const double unbound = C(fld, i, k) * C(thkn, i, k) - hordiv;
        const double bounded =
            dmax2(dmax2(0., C(fmn, i, k) * C(thkn, i, k)), dmin2(unbound, C(fmx, i, k) * C(thkn, i, k)));
        C(fld, i, k) = (C(fld, i, k) * onemu + bounded) / (onemu + C(thkn, i, k));
      }
    }
    /* global conservation (recovr, 482-524): the shift of the mean */
    double after = 0., wtnew = 0.;
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      double col1 = 0., col2 = 0.;
      for (int k = 0; k < kdm; k++) {
        col1 = col1 + C(thkn, i, k) * g->area[i];
        col2 = col2 + C(fld, i, k) * C(thkn, i, k) * g->area[i];
      }
      wtnew = wtnew + col1;
      after = after + col2;
    }
    bfore = bfore / wtold;
    after = after / wtnew;
    const double q = bfore - after;
    for (int i = 0; i < nip; i++) {
      if (!WET(i)) continue;
      for (int k = 0; k < kdm; k++) C(fld, i, k) = C(fld, i, k) + q;
    }
    for (int i = 0; i < nip; i++)
      if (WET(i))
        for (int k = 0; k < kdm; k++) C(f3, i, k) = C(fld, i, k);
  }
  free(fld); free(vertfx); free(vertdv); free(fmx); free(fmn); free(flp); free(fln); free(flx); free(uan);
  free(a); free(b); free(c);
}

/* hycom_dffusn.F90 dffusn_lyr(fld, delp, kdm, 1, kdm, dffveldt): thickness-
 * weighted Laplacian smoothing through the ocean edges */
static void hyc_dffusn_lyr(const fo_hyc_grid *g, double *fld, const double *delp, double dffveldt) {
  const int nip = g->nip, kdm = g->kdm;
  const double thshld = 1.e-11;
  double *flxdv = calloc((size_t)nip * kdm, sizeof(double));
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int e = 0; e < g->nprox[i]; e++) {
      const int ix = g->prox[i * NPP + e];
      if (!WET(ix)) continue;
      for (int k = 0; k < kdm; k++) {
        const double a = dmax2(C(delp, i, k), thshld), b = dmax2(C(delp, ix, k), thshld);
        C(flxdv, i, k) = C(flxdv, i, k) + (C(fld, i, k) - C(fld, ix, k)) * g->sideln[i * NPP + e] * 2. * (a * b / (a + b));
      }
    }
  }
  for (int i = 0; i < nip; i++) {
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) {
      const double factor = -dffveldt * g->rarea[i] / dmax2(C(delp, i, k), thshld);
      C(fld, i, k) = C(fld, i, k) + C(flxdv, i, k) * factor;
    }
  }
  free(flxdv);
}

/* hycom_transp3d.F90 transp2 */
void fo_hycom_transp2(const fo_hyc_grid *g, int numtr, double batrop, double temdff, double *temp, double *saln,
                      double *passv, const double *cumuflx, const double *dpinit, const double *dpfinl) {
  const int nip = g->nip, kdm = g->kdm;
  const size_t cz = (size_t)nip * kdm;
  const int numfld = 2 + numtr;
  double *vertfx = malloc(cz * sizeof(double)), *all = malloc(cz * numfld * sizeof(double));
  /* the vertical mass flux the time-integrated horizontal fluxes imply */
  for (int i = 0; i < nip; i++) {
    for (int k = 0; k < kdm; k++) C(vertfx, i, k) = LAND_SPVAL;
    if (!WET(i)) continue;
    for (int k = 0; k < kdm; k++) {
      double hordiv = 0.;
      for (int e = 0; e < g->nprox[i]; e++) hordiv = hordiv + E(cumuflx, i, e, k);
      const double vertdv = (C(dpinit, i, k) - C(dpfinl, i, k)) - hordiv * g->rarea[i];
      C(vertfx, i, k) = k == 0 ? vertdv : C(vertfx, i, k - 1) + vertdv;
    }
  }
  for (int t = 0; t < numfld; t++) {
    const double *src = t == 0 ? temp : t == 1 ? saln : passv + (size_t)(t - 2) * cz;
    double *dst = all + (size_t)t * cz;
    for (int i = 0; i < nip; i++)
      for (int k = 0; k < kdm; k++) C(dst, i, k) = WET(i) ? C(src, i, k) : LAND_SPVAL;
  }
  for (int i = 0; i < nip; i++)
    if (WET(i))
      for (int k = 0; k < kdm; k++) C(all, i, k) = C(all, i, k) + TOFFSET;
  hyc_fct3d(g, all, numfld, cumuflx, vertfx, dpinit, dpfinl);
  for (int i = 0; i < nip; i++)
    if (WET(i))
      for (int k = 0; k < kdm; k++) C(all, i, k) = C(all, i, k) - TOFFSET;
  for (int t = 0; t < numfld; t++) {
    double *dst = t == 0 ? temp : t == 1 ? saln : passv + (size_t)(t - 2) * cz;
    const double *src = all + (size_t)t * cz;
    for (int i = 0; i < nip; i++)
      if (WET(i))
        for (int k = 0; k < kdm; k++) C(dst, i, k) = C(src, i, k);
  }
  if (temdff > 0.) {
    hyc_dffusn_lyr(g, temp, dpfinl, temdff * batrop * BCLIN_FRQ);
    hyc_dffusn_lyr(g, saln, dpfinl, temdff * batrop * BCLIN_FRQ);
  }
  free(vertfx); free(all);
}
