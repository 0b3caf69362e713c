/* hycom_oracle.c - CPU restatement, in plain C and FP64, of the HYCOM ocean
 * continuity equation (FIMsrc/fim/horizontal/hycom_cnuity.F90, subroutine
 * cnuity), of its edge interpolation (hycom_edgvar.F90, bclin_edgvar with
 * uvedge and blokf), of its momentum equation (hycom_momtum.F90 momtum
 * with hycom_dissip.F90's del4prep and dissip), of its tracer transport
 * (hycom_transp3d.F90 transp0/1/2 with hycom_fct3d.F90's fct3d and
 * hycom_dffusn.F90's dffusn_lyr) and of its hybrid grid generator
 * (hycom_hybgn1.F90 hybgn1 with hycom_sigetc.F90's equation of state).
 * TEST INFRASTRUCTURE ONLY: the checker the HIP kernels are compared with;
 * the product never links it.
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

/* ------------------------------------------------------------------------
 * hycom_hybgn1.F90 (hybgn1, hybgn1_1d, ppmad3, pcwise, equilb) with
 * hycom_sigetc.F90's sigocn, dsigdt and dsigds.  The column routines use
 * 1-based arrays, as the reference does (element 0 unused). */
static const double HYB_PREF = 1.e7;          /* hycom_constants.F90:56 (sigma_1) */
static const double HYB_C2 = 1.470290E-02, HYB_C3 = 7.904674E-01, HYB_C4 = -6.861022E-03;
static const double HYB_C5 = -3.051459E-03, HYB_C6 = 3.099320E-05, HYB_C7 = 3.050299E-05;
static const double HYB_C8 = 1.046036E-04, HYB_C9 = 5.287567E-06;

/* hycom_sigetc.F90 sigocn: potential density (sigma_1) of t, s */
static double hyb_sigocn(double t, double s) {
  const double k01 = 9.9984085444849347e2, a1 = 7.3471625860981584, a2 = -5.3211231792841769e-2,
               a3 = 3.6492439109814549e-4, b1 = 2.588057102399139, b2 = -6.7168282786692355e-3,
               b3 = 1.9203202055760151e-3, cc1 = 1.1798263740430364e-2, cc2 = 9.8920219266399117e-8,
               cc3 = 4.699664277175473e-6, cc4 = -2.5862187075154352e-8, cc5 = -3.2921414007960662e-12;
  const double k02 = 1.0, d1 = 7.2815210113327091e-3, d2 = -4.4787265461983921e-5, d3 = 3.3851002965802430e-7,
               d4 = 1.3651202389758572e-10, e1 = 1.7632126669040377e-3, e2 = -8.8066583251206474e-6,
               e3 = -1.8832689434804897e-10, e4 = 5.7463776745432097e-6, e5 = 1.4716275472242334e-9,
               f1 = 6.7103246285651894e-6, f2 = -2.4461698007024582e-17, f3 = -9.1534417604289062e-18;
  const double p = HYB_PREF * 1.e-4;
  const double pn = k01 + t * (a1 + t * (a2 + a3 * t)) + s * (b1 + b2 * t + b3 * s) +
                    p * (cc1 + cc2 * t * t + cc3 * s + p * (cc4 + cc5 * t * t));
  const double pd = k02 + t * (d1 + d2 * t + t * t * (d3 + d4 * t)) +
                    s * (e1 + e2 * t + e3 * t * t * t + sqrt(dmax2(0., s)) * (e4 + e5 * t * t)) +
                    p * (f1 + p * t * (f2 * t * t + f3 * p));
  return pn / pd - 1000.;
}
static double hyb_dsigdt(double t, double s) {
  return HYB_C2 + s * (HYB_C5 + HYB_C9 * s) + 2. * t * (HYB_C4 + HYB_C7 * s + 1.5 * HYB_C6 * t);
}
static double hyb_dsigds(double t, double s) {
  return HYB_C3 + 2. * s * (HYB_C8 + HYB_C9 * t) + t * (HYB_C5 + t * HYB_C7);
}

/* the parabola of a cell (value y, neighbours yl/yr at its faces), with the
 * over/undershoot fix the reference applies in ppmad3, pcwise and fct3d */
static void hyb_parab(double y, double yl, double yr, double *a, double *b, double *c) {
  *a = 1.5 * y - .25 * (yl + yr);
  *b = yr - yl;
  *c = 6. * (.5 * (yl + yr) - y);
  if (fabs(yr - yl) < 6. * fabs(.5 * (yl + yr) - y)) {
    if (fabs(yr - yl) > 2. * fabs(.5 * (yl + yr) - y)) {
      if ((yr - yl) * (.5 * (yl + yr) - y) > 0.) {
        *a = .25 * (3. * y + yl);
        *c = 3. * (y - yl);
        *b = *c;
      } else {
        *a = .25 * (3. * y + yr);
        *c = 3. * (y - yr);
        *b = -*c;
      }
    } else {
      *a = y;
      *b = 0.;
      *c = 0.;
    }
  }
}

/* hycom_hybgn1.F90 ppmad3(x, dx, y): three cells x(1..4), displacements
 * dx(1..4), values y(1..3) updated */
static void hyb_ppmad3(const double *x, const double *dx, double *y) {
  const double onemu = .098, athird = 1. / 3.;
  double yold[4], ytmp[4];
  for (int n = 1; n <= 3; n++) {
    yold[n] = y[n];
    ytmp[n] = y[n] * (x[n + 1] - x[n]);
  }
  const double dxlft = dmax2(onemu, x[2] - x[1]), dxmid = dmax2(onemu, x[3] - x[2]);
  const double dxrgt = dmax2(onemu, x[4] - x[3]);
  const double yl = (dxlft * yold[2] + dxmid * yold[1]) / (dxlft + dxmid);
  const double yr = (dxrgt * yold[2] + dxmid * yold[3]) / (dxrgt + dxmid);
  double a, b, c;
  hyb_parab(yold[2], yl, yr, &a, &b, &c);
  if (dx[2] > 0.) {
    const double wdth = dx[2] / (x[3] - x[2]);
    const double slab = (x[3] - x[2]) * wdth * (a + b * .5 * (wdth - 1.) + c * (.25 - wdth * (.5 - wdth * athird)));
    ytmp[2] = ytmp[2] - slab;
    ytmp[1] = ytmp[1] + slab;
  }
  if (dx[3] < 0.) {
    const double wdth = -dx[3] / (x[3] - x[2]);
    const double slab = (x[3] - x[2]) * wdth * (a + b * .5 * (1. - wdth) + c * (.25 - wdth * (.5 - wdth * athird)));
    ytmp[2] = ytmp[2] - slab;
    ytmp[3] = ytmp[3] + slab;
  }
  for (int n = 1; n <= 3; n++) {
    const double dxnew = x[n + 1] - x[n] + dx[n + 1] - dx[n];
    if (dxnew > 0.) y[n] = ytmp[n] / dxnew;
  }
}

/* hycom_hybgn1.F90 pcwise(remap_optn, kk, xold, fldold, xnew, fldnew,
 * miglim) with miglim = -1 (no migration limit: hybgn1_1d's only use),
 * remap_optn 1 (PCM) or 3 (PPM); arrays 1-based, w a work area of
 * 11*(kk+2) */
static void hyb_pcwise(int optn, int kk, const double *xold, const double *fldold, const double *xnew,
                       double *fldnew, double *w) {
  const double athird = 1. / 3., small = 1.e-11;
  const int n = kk + 2;
  double *zold = w, *znew = w + n, *delz = w + 2 * n, *fco = w + 3 * n, *fcn = w + 4 * n;
  double *vertfx = w + 5 * n, *vertdv = w + 6 * n, *a = w + 7 * n, *b = w + 8 * n, *c = w + 9 * n;
  if (xold[1] < xold[kk + 1]) {
    for (int k = 1; k <= kk + 1; k++) {
      zold[k] = xold[k];
      znew[k] = xnew[k];
    }
  } else {
    for (int k = 1; k <= kk + 1; k++) {
      zold[k] = -xold[k];
      znew[k] = -xnew[k];
    }
  }
  for (int k = 1; k <= kk + 1; k++) delz[k] = znew[k] - zold[k];
  for (int k = 1; k <= kk; k++) {
    fco[k] = dmax2(0., zold[k + 1] - zold[k]);
    fcn[k] = dmax2(0., znew[k + 1] - znew[k]);
  }
  for (int k = 1; k <= kk; k++) fldnew[k] = fldold[k];
  /* zero-width cells take their neighbours' values */
  for (int k = kk - 1; k >= 1; k--) fldnew[k] = (fldnew[k] * fco[k] + fldnew[k + 1] * small) / (fco[k] + small);
  for (int k = 2; k <= kk; k++) fldnew[k] = (fldnew[k] * fco[k] + fldnew[k - 1] * small) / (fco[k] + small);
  a[1] = fldnew[1];
  b[1] = c[1] = 0.;
  a[kk] = fldnew[kk];
  b[kk] = c[kk] = 0.;
  for (int k = 2; k <= kk - 1; k++) {
    if (optn == 1) {
      a[k] = fldnew[k];
      b[k] = c[k] = 0.;
    } else {
      const double yl = (dmax2(small, fco[k - 1]) * fldnew[k] + dmax2(small, fco[k]) * fldnew[k - 1]) /
                        (dmax2(small, fco[k - 1]) + dmax2(small, fco[k]));
      const double yr = (dmax2(small, fco[k + 1]) * fldnew[k] + dmax2(small, fco[k]) * fldnew[k + 1]) /
                        (dmax2(small, fco[k + 1]) + dmax2(small, fco[k]));
      hyb_parab(fldnew[k], yl, yr, &a[k], &b[k], &c[k]);
    }
  }
  /* the flux through each interface from the upstream slab it sweeps */
  for (int k = 2; k <= kk; k++) {
    double slab = 0., amount = 0.;
    vertfx[k] = 0.;
    if (delz[k] > 0.) {
      int lyr = k - 1;
      for (;;) {
        lyr = lyr + 1;
        if (slab >= delz[k]) break;
        if (fco[lyr] > 0.) {
          const double dslab = dmin2(slab + fco[lyr], delz[k]) - dmin2(slab, delz[k]);
          const double dx = dslab / fco[lyr];
          const double fcdx = a[lyr] + b[lyr] * .5 * (dx - 1.) + c[lyr] * (.25 - dx * (.5 - dx * athird));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
        if (!(lyr < kk)) break;
      }
    } else if (delz[k] < 0.) {
      /* the reference's walk stops after layer 2 (goto 25 while lyr > 2):
       * layer 1 is swept only from interface 2 */
      int lyr = k;
      for (;;) {
        lyr = lyr - 1;
        if (slab >= -delz[k]) break;
        if (fco[lyr] > 0.) {
          const double dslab = dmin2(slab + fco[lyr], -delz[k]) - dmin2(slab, -delz[k]);
          const double dx = dslab / fco[lyr];
          const double fcdx = a[lyr] + b[lyr] * .5 * (1. - dx) + c[lyr] * (.25 - dx * (.5 - dx * athird));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
        if (!(lyr > 2)) break;
      }
    }
    if (slab != 0.) vertfx[k] = -delz[k] * amount / slab;
  }
  vertfx[1] = 0.;
  vertfx[kk + 1] = 0.;
  for (int k = 1; k <= kk; k++) vertdv[k] = vertfx[k + 1] - vertfx[k];
  for (int k = 1; k <= kk; k++) {
    const double amount = fldnew[k] * fco[k] - vertdv[k];
    fldnew[k] = (fldnew[k] * small + amount) / (small + fcn[k]);
  }
}

/* hycom_hybgn1.F90 equilb(nstep, pres, thet, temp, saln, kdm): heat
 * diffusion (3 passes) to inflate thin layers, then scenario 4 (the others
 * are commented out in the reference); thet is read only; w a work area of
 * 7*(kdm+2) */
static void hyb_equilb(int kdm, double *pres, const double *thet, double *temp, double *saln, double *w) {
  const double tenm = 10. * ONEM, tencm = 0.1 * ONEM, onemu = 1.e-6 * ONEM;
  const double slak = .25, dffudt = .1;
  const int n = kdm + 2;
  double *pnew = w, *pwrk = w + n, *heatfx = w + 2 * n, *tdz = w + 3 * n, *sdz = w + 4 * n;
  double *twrk = w + 5 * n, *swrk = w + 6 * n;
  for (int k = 1; k <= kdm + 1; k++) pnew[k] = pwrk[k] = pres[k];
  for (int k = 1; k <= kdm; k++) {
    twrk[k] = temp[k];
    swrk[k] = saln[k];
    tdz[k] = dmax2(EPSIL, pres[k + 1] - pres[k]) * temp[k];
    sdz[k] = dmax2(EPSIL, pres[k + 1] - pres[k]) * saln[k];
  }
  heatfx[kdm] = 0.;
  heatfx[1] = 0.;
  for (int iter = 1; iter <= 3; iter++) {
    for (int k = 2; k <= kdm - 1; k++) {
      heatfx[k] = 0.;
      if (pwrk[k] > pwrk[1] + ONEM && pwrk[k + 1] < pwrk[kdm + 1] - ONEM) {
        const double cnv = ONEM;
        heatfx[k] = dffudt * cnv * cnv * dmin2(thet[k + 1] - thet[k], thet[k] - thet[k - 1]) /
                    dmax2(.01 * cnv, pwrk[k + 1] - pwrk[k]);
        const double plimit1 = dmax2(pwrk[k] - tenm, .5 * (pwrk[k] + pwrk[k - 1]));
        const double plimit2 = dmin2(pwrk[k + 1] + tenm, .5 * (pwrk[k + 1] + pwrk[k + 2]));
        heatfx[k] = dmin2(dmin2(dmax2(0., heatfx[k]), dmax2(0., (plimit2 - pwrk[k + 1]) * (thet[k + 1] - thet[k]))),
                          dmax2(0., (pwrk[k] - plimit1) * (thet[k] - thet[k - 1])));
      }
    }
    for (int k = 1; k <= kdm - 1; k++) {
      if (thet[k + 1] > thet[k]) {
        pnew[k + 1] = pwrk[k + 1] + (heatfx[k] - heatfx[k + 1]) / (thet[k + 1] - thet[k]);
        const double delp = pnew[k + 1] - pwrk[k + 1];
        double dtdz, dsdz;
        if (delp > 0.) {
          dtdz = twrk[k + 1] * delp;
          dsdz = swrk[k + 1] * delp;
        } else {
          dtdz = twrk[k] * delp;
          dsdz = swrk[k] * delp;
        }
        tdz[k] = tdz[k] + dtdz;
        tdz[k + 1] = tdz[k + 1] - dtdz;
        sdz[k] = sdz[k] + dsdz;
        sdz[k + 1] = sdz[k + 1] - dsdz;
      }
    }
    for (int k = 1; k <= kdm; k++) {
      if (pnew[k + 1] > pnew[k]) {
        twrk[k] = tdz[k] / (pnew[k + 1] - pnew[k]);
        swrk[k] = sdz[k] / (pnew[k + 1] - pnew[k]);
      }
    }
    for (int k = 1; k <= kdm + 1; k++) pwrk[k] = pnew[k];
  }
  /* scenario 4: thick-(thin-thin)-thick */
  for (int k = 4; k <= kdm - 2; k++) {
    if (pnew[k - 2] < pnew[1] + tencm || pnew[k + 2] > pnew[kdm + 1] - tencm) continue;
    const double dp1 = pnew[k - 1] - pnew[k - 2], dp2 = pnew[k] - pnew[k - 1];
    const double dp3 = pnew[k + 1] - pnew[k], dp4 = pnew[k + 2] - pnew[k + 1];
    const double th1 = thet[k - 2], th2 = thet[k - 1], th3 = thet[k], th4 = thet[k + 1];
    if (dp1 > dp2 + dp3 && dp4 > dp2 + dp3) {
      const double goal = .5 * (dp2 + dp3 + dmin2(dp1, dp4));
      double dis2 = (goal - dp2 - dp3) * slak, dis4 = dis2;
      if (th2 > th1 && th3 > th2 && th4 > th3) {
        const double ratio = (th4 - th3) / (th2 - th1);
        if (ratio > 1.) dis4 = dis2 / ratio;
        else dis2 = dis4 * ratio;
      }
      pnew[k - 1] = pnew[k - 1] - dis2;
      pnew[k + 1] = pnew[k + 1] + dis4;
      tdz[k - 2] = tdz[k - 2] - dis2 * twrk[k - 2];
      tdz[k - 1] = tdz[k - 1] + dis2 * twrk[k - 2];
      tdz[k] = tdz[k] + dis4 * twrk[k + 1];
      tdz[k + 1] = tdz[k + 1] - dis4 * twrk[k + 1];
      sdz[k - 2] = sdz[k - 2] - dis2 * swrk[k - 2];
      sdz[k - 1] = sdz[k - 1] + dis2 * swrk[k - 2];
      sdz[k] = sdz[k] + dis4 * swrk[k + 1];
      sdz[k + 1] = sdz[k + 1] - dis4 * swrk[k + 1];
    }
  }
  for (int k = 1; k <= kdm; k++) {
    if (pnew[k + 1] - pnew[k] > onemu) {
      temp[k] = tdz[k] / (pnew[k + 1] - pnew[k]);
      saln[k] = sdz[k] / (pnew[k + 1] - pnew[k]);
    }
    pres[k] = pnew[k];
  }
}

/* hycom_hybgn1.F90 hybgn1_1d: one column (1-based arrays); trc (numtr
 * columns of kdm+2) moved only with dotrcr; w a work area of
 * (20 + 2*numtr)*(kdm+2) */
static void hyb_1d(int kdm, const double *targt, const double *salmin, const double *dpmin, double batrop,
                   double lati, double *pr, double *tem, double *sal, double *dns, double *u, double *v,
                   double *trc, int numtr, int dotrcr, double *w) {
  const double tencm = 0.1 * ONEM, onemm = 0.001 * ONEM, r_onem = 1. / ONEM, huuge = 1.e33;
  const double tfreez = -2.2, sigjmp = .01, scalt = -30., scals = 10.;
  const double slak = 1. / 86400.;
  const double dlt = batrop * 30;                   /* batrop*bclin_frq (hycom_control.F90:9) */
  const int n = kdm + 2;
  double *uold = w, *vold = w + n, *pold = w + 2 * n, *unew = w + 3 * n, *vnew = w + 4 * n;
  double *trold = w + 5 * n, *trnew = w + (5 + numtr) * n, *wk = w + (5 + 2 * numtr) * n;
  double p_hat = 0., q, displ[5];
#define SLAKF() ((1. + 9. * dmax2(0., 1. - .01 * r_onem * p_hat)) * slak * dlt)
  for (int k = 1; k <= kdm; k++) {
    uold[k] = u[k];
    vold[k] = v[k];
    if (dotrcr)
      for (int nt = 0; nt < numtr; nt++) trold[nt * n + k] = trc[nt * n + k];
  }
  /* absorb near-massless layers on the sea floor in the layer above */
  int kp = 1;
  for (int k = 2; k <= kdm; k++) {
    if (pr[k] < pr[kdm + 1] - tencm) {
      kp = k;
    } else {
      const double q1 = dmax2(EPSIL, pr[k] - pr[kp]), q2 = dmax2(0., pr[k + 1] - pr[k]);
      q = q1 / (q1 + q2);
      tem[kp] = tem[kp] * q + tem[k] * (1. - q);
      sal[kp] = sal[kp] * q + sal[k] * (1. - q);
      tem[k] = tem[kp];
      sal[k] = dmax2(sal[kp], salmin[k]);
      if (dotrcr)
        for (int nt = 0; nt < numtr; nt++) {
          double *t = trold + nt * n;
          t[kp] = t[kp] * q + t[k] * (1. - q);
          t[k] = t[kp];
        }
    }
  }
  dns[kp] = hyb_sigocn(tem[kp], sal[kp]);
  for (int k = kp + 1; k <= kdm; k++) {
    dns[k] = dns[kp];
    pr[k] = pr[kdm + 1];
  }
  /* a layer touching the sea floor too light: split it, the lower part at
   * its target density, the upper part joined to the layer above */
  if (kp != 1) {
    const int k = kp;
    if (!(dns[k] <= dmax2(dns[k - 1], targt[k - 1]) || dns[k] >= targt[k])) {
      const double tem_ = tem[k], sal_ = sal[k];
      const double dsgdt = hyb_dsigdt(tem_, sal_) * scalt, dsgds = hyb_dsigds(tem_, sal_) * scals;
      q = 1. / (dsgdt + dsgds);
      const double sal_lo = dmin2(sal_ + (targt[k] - dns[k]) * q * scals, dmax2(sal[k - 1], sal_ + .01 * scals));
      const double tem_lo = dmax2(tem_ + (targt[k] - dns[k]) * q * scalt, tfreez);
      const double rho_lo = hyb_sigocn(tem_lo, sal_lo);
      const double rho_up = dmax2(dns[k - 1], dmin2(targt[k - 1], dns[k]));
      q = (rho_lo - dns[k]) / dmax2(EPSIL, rho_lo - rho_up);
      if (!(q < 0. || q > 1.)) {
        p_hat = pr[k] * (1. - q) + pr[k + 1] * q;
        q = (pr[k + 1] - p_hat) / dmax2(EPSIL, p_hat - pr[k]);
        const double sal_up = sal_ * (1. + q) - sal_lo * q;
        const double tem_up = tem_ * (1. + q) - tem_lo * q;
        q = (p_hat - pr[k]) / dmax2(p_hat - pr[k - 1], EPSIL);
        sal[k - 1] = sal_up * q + sal[k - 1] * (1. - q);
        tem[k - 1] = tem_up * q + tem[k - 1] * (1. - q);
        if (dotrcr)
          for (int nt = 0; nt < numtr; nt++) {
            double *t = trold + nt * n;
            t[k - 1] = t[k] * q + t[k - 1] * (1. - q);
          }
        dns[k - 1] = hyb_sigocn(tem[k - 1], sal[k - 1]);
        tem[k] = tem_lo;
        sal[k] = sal_lo;
        dns[k] = rho_lo;
        pr[k] = p_hat;
      }
    }
  }
  for (int k = 1; k <= kdm + 1; k++) pold[k] = pr[k];
  /* restore isopycnic conditions by moving interfaces */
  double dpsum = 0., dp0 = huuge;
  for (int k = 1; k <= kdm; k++) {
    const double dp0abv = dp0;
    dp0 = dpmin[k] * ONEM;
    if (k > 1) dp0 = dp0 * dmax2(.6, dmin2(1., (fabs(lati) + 5.) * .04));
    if (4 * k < kdm) dp0 = dp0 * dmin2(1., pr[kdm + 1] / (200. * ONEM) + .4);
    dpsum = dpsum + dp0;
    int lower = 0;                   /* 1: go to the lower-interface move (label 5) */
    if (k == 1) {
      p_hat = dp0;
      if (p_hat > pr[2]) {           /* layer 1 too thin */
        p_hat = dmin2(p_hat, pr[2] + dmax2(ONECM, SLAKF() * (p_hat - pr[2])));
        lower = 1;
      } else if (p_hat < pr[2]) {    /* too thick: layer 1 water into layer 2 */
        p_hat = dmax2(p_hat, pr[2] + dmin2(-ONECM, SLAKF() * (p_hat - pr[2])));
        q = (pr[2] - p_hat) / dmax2(pr[3] - p_hat, EPSIL);
        sal[2] = sal[2] * (1. - q) + sal[1] * q;
        tem[2] = tem[2] * (1. - q) + tem[1] * q;
        dns[2] = hyb_sigocn(tem[2], sal[2]);
        pr[2] = p_hat;
      }
      if (!lower) continue;
    } else {
      if (pr[k] == pr[kdm + 1]) dns[k] = dmax2(targt[k], dns[k]);
      if (pr[k] > pr[kdm + 1] - ONECM) continue;
      p_hat = dpsum;
      if (k < kdm && p_hat > pr[k + 1]) {          /* lower interface too close to the surface */
        p_hat = dmin2(p_hat, pr[k + 1] + dmax2(ONECM, SLAKF() * (p_hat - pr[k + 1])));
        lower = 1;
      } else {
        if (fabs(dns[k] - targt[k]) < .1 * sigjmp) continue;
        if (dns[k] <= targt[k]) {                  /* too light (label 7) */
          if (k >= kdm || pr[k + 1] > pr[kdm + 1] - onemm) continue;
          q = (dns[k] - targt[k]) / dmax2(dns[k + 1] - targt[k], sigjmp * 10.);
          p_hat = pr[k + 1] * (1. - q) + pr[k] * q;
          p_hat = dmin2(p_hat, dmax2(.5 * (pr[k] + pr[k + 2]), 2. * pr[k + 1] - pr[k]));
          p_hat = dmin2(p_hat, pr[k + 1] + dmax2(ONECM, SLAKF() * (p_hat - pr[k + 1])));
          lower = 1;
        } else {                                   /* too dense */
          if (k != 2) {
            q = (targt[k] - dns[k]) / dmax2(targt[k] - dns[k - 1], sigjmp * 10.);
            p_hat = pr[k] * (1. - q) + pr[k + 1] * q;
            p_hat = dmax2(p_hat, pr[k - 1] + dp0abv);
            p_hat = dmin2(p_hat, .5 * (pr[k - 1] + pr[k + 1]));
            if (p_hat < pr[k]) {                   /* upper interface up: entrain layer k-1 */
              p_hat = dmax2(dmax2(p_hat, pr[k - 1]), pr[k] + dmin2(-ONECM, SLAKF() * (p_hat - pr[k])));
              if (fabs(dns[k - 1] - targt[k - 1]) > .1 * sigjmp) {
                displ[1] = 0.;
                displ[2] = 0.;
                displ[3] = p_hat - pr[k];
                displ[4] = 0.;
                hyb_ppmad3(pr + k - 3, displ, sal + k - 3);
                hyb_ppmad3(pr + k - 3, displ, tem + k - 3);
                dns[k - 1] = hyb_sigocn(tem[k - 1], sal[k - 1]);
                dns[k] = hyb_sigocn(tem[k], sal[k]);
              } else {
                q = (pr[k] - p_hat) / dmax2(pr[k + 1] - p_hat, EPSIL);
                sal[k] = sal[k] * (1. - q) + sal[k - 1] * q;
                tem[k] = tem[k] * (1. - q) + tem[k - 1] * q;
                dns[k] = hyb_sigocn(tem[k], sal[k]);
              }
              pr[k] = p_hat;
            } else if (p_hat > pr[k]) {            /* upper interface down: layer k into k-1 */
              p_hat = dmin2(dmin2(p_hat, pr[k + 1]), pr[k] + dmax2(ONECM, SLAKF() * (p_hat - pr[k])));
              if (k < kdm) {
                displ[1] = 0.;
                displ[2] = p_hat - pr[k];
                displ[3] = 0.;
                displ[4] = 0.;
                hyb_ppmad3(pr + k - 2, displ, sal + k - 2);
                hyb_ppmad3(pr + k - 2, displ, tem + k - 2);
                dns[k - 1] = hyb_sigocn(tem[k - 1], sal[k - 1]);
                dns[k] = hyb_sigocn(tem[k], sal[k]);
              } else {
                q = (p_hat - pr[k]) / dmax2(p_hat - pr[k - 1], EPSIL);
                sal[k - 1] = sal[k - 1] * (1. - q) + sal[k] * q;
                tem[k - 1] = tem[k - 1] * (1. - q) + tem[k] * q;
                dns[k - 1] = hyb_sigocn(tem[k - 1], sal[k - 1]);
              }
              pr[k] = p_hat;
            }
          }
          /* label 6: inflate layer k by lowering its lower interface? */
          p_hat = pr[k] + dpmin[2] * ONEM;
          if (k < kdm && pr[k + 1] < pr[kdm + 1] - onemm && pr[k + 1] < p_hat) {
            p_hat = dmin2(p_hat, pr[k + 1] + dmax2(ONECM, SLAKF() * (p_hat - pr[k + 1])));
            lower = 1;
          }
          if (!lower) continue;
        }
      }
    }
    /* label 5: move the lower interface down to p_hat */
    p_hat = dmin2(p_hat, pr[k + 2]);
    if (p_hat > pr[k + 1] + onemm) {
      if (k < kdm - 1 && fabs(dns[k + 1] - targt[k + 1]) > .1 * sigjmp) {
        displ[1] = 0.;
        displ[2] = p_hat - pr[k + 1];
        displ[3] = 0.;
        displ[4] = 0.;
        hyb_ppmad3(pr + k - 1, displ, sal + k - 1);
        hyb_ppmad3(pr + k - 1, displ, tem + k - 1);
        dns[k] = hyb_sigocn(tem[k], sal[k]);
        dns[k + 1] = hyb_sigocn(tem[k + 1], sal[k + 1]);
      } else {
        q = (p_hat - pr[k + 1]) / dmax2(p_hat - pr[k], EPSIL);
        sal[k] = sal[k] * (1. - q) + sal[k + 1] * q;
        tem[k] = tem[k] * (1. - q) + tem[k + 1] * q;
        dns[k] = hyb_sigocn(tem[k], sal[k]);
      }
      pr[k + 1] = p_hat;
    }
  }
#undef SLAKF
  hyb_equilb(kdm, pr, dns, tem, sal, wk);
  for (int k = 1; k <= kdm - 2; k++)
    if (pr[k + 1] < pr[k] || pr[k + 1] > pr[k + 2]) pr[k + 1] = dmax2(pr[k], dmin2(pr[k + 1], pold[k + 2]));
  hyb_pcwise(1, kdm, pold, uold, pr, unew, wk);
  hyb_pcwise(1, kdm, pold, vold, pr, vnew, wk);
  if (dotrcr)
    for (int nt = 0; nt < numtr; nt++) hyb_pcwise(3, kdm, pold, trold + nt * n, pr, trnew + nt * n, wk);
  for (int k = 1; k <= kdm; k++) {
    u[k] = unew[k];
    v[k] = vnew[k];
    if (dotrcr)
      for (int nt = 0; nt < numtr; nt++) trc[nt * n + k] = trnew[nt * n + k];
  }
}

void fo_hycom_hybgn1(int nip, int kdm, const int *wet, int leap, const double *targt, const double *salmin,
                     const double *dpmin, double batrop, const double *latij, double *dens, double *dp, double *pres,
                     double *temp, double *saln, double *uvel, double *vvel, int numtr, double *passv) {
  const int n = kdm + 2, dotrcr = passv != NULL && numtr > 0;
  const size_t cz = (size_t)nip * kdm;
  double *col = malloc((size_t)(7 + numtr) * n * sizeof(double));
  double *w = malloc((size_t)(20 + 2 * numtr) * n * sizeof(double));
  double *tg = malloc((size_t)3 * n * sizeof(double));
  double *pr = col, *tem = col + n, *sal = col + 2 * n, *dns = col + 3 * n, *u = col + 4 * n, *v = col + 5 * n;
  double *trc = col + 7 * n;
  for (int k = 1; k <= kdm; k++) {
    tg[k] = targt[k - 1];
    tg[n + k] = salmin[k - 1];
    tg[2 * n + k] = dpmin[k - 1];
  }
  double *dpl = dp + (size_t)(leap - 1) * cz, *ul = uvel + (size_t)(leap - 1) * cz, *vl = vvel + (size_t)(leap - 1) * cz;
  for (int i = 0; i < nip; i++) {
    if (wet[i] <= 0) continue;
    pr[1] = pres[(size_t)i * (kdm + 1)];
    for (int k = 1; k <= kdm; k++) {
      const size_t ik = (size_t)i * kdm + k - 1;
      dns[k] = dens[ik];
      tem[k] = temp[ik];
      sal[k] = saln[ik];
      u[k] = ul[ik];
      v[k] = vl[ik];
      if (dotrcr)
        for (int nt = 0; nt < numtr; nt++) trc[nt * n + k] = passv[nt * cz + ik];
      pr[k + 1] = pr[k] + dpl[ik];
    }
    hyb_1d(kdm, tg, tg + n, tg + 2 * n, batrop, latij[i], pr, tem, sal, dns, u, v, trc, numtr, dotrcr, w);
    for (int k = 1; k <= kdm; k++) {
      const size_t ik = (size_t)i * kdm + k - 1;
      dens[ik] = dns[k];
      temp[ik] = tem[k];
      saln[ik] = sal[k];
      ul[ik] = u[k];
      vl[ik] = v[k];
      dpl[ik] = pr[k + 1] - pr[k];
      if (dotrcr)
        for (int nt = 0; nt < numtr; nt++) passv[nt * cz + ik] = trc[nt * n + k];
      pres[(size_t)i * (kdm + 1) + k] = pr[k + 1];
    }
  }
  free(col);
  free(w);
  free(tg);
}
