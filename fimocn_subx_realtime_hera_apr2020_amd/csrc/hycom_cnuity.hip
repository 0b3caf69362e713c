// The HYCOM ocean's continuity equation (FIMsrc/fim/horizontal/hycom_cnuity.F90,
// subroutine cnuity) on the GPU, in FP64, for a grid whose cells are ocean
// or land (wet).
//
// Each column runs on LPC lanes of a wavefront (LPC = 32 for up to 32
// layers, two columns per wavefront; 64 otherwise), lane k holding layer k.
// Every operation is the reference's, in its order: the sums over a cell's
// edges run in edge order per lane, and the two sums the reference forms
// along the column (interface pressure pres(k+1) = pres(k) + dp(k), and the
// clipped part of each edge's flux) are exact sequential chains that every
// lane of the column evaluates in layer order from broadcast values.  A
// kernel boundary stands wherever a stage reads a neighbour's value that the
// previous stage writes:
//   hyc_flux     low/high-order edge fluxes, low-order thickness, pressure,
//                in/outflow sums of the antidiffusive fluxes    (98-265)
//   hyc_limits   local bounds and limiter ratios              (269-300)
//   hyc_clip     limited antidiffusive fluxes and update      (315-393)
//   hyc_recover  the clipped flux as a barotropic correction  (434-462)
//   hyc_update   its thickness update; dp from pres, del2p    (465-548)
//   hyc_del2p    the biharmonic del2p from the neighbours     (553-603)
//   hyc_bolus    interface-slope bolus fluxes, outflow ratios (605-683)
//   hyc_bclip    limited bolus fluxes and their update        (685-776)
//   hyc_final    the clipped bolus correction, the mass-flux average, the
//                leapfrog time smoothing, the bottom pressure (778-954)
// The duplicate-flux check (chkdup) counts differing pairs into a flag; the
// global mass and its drift correction (glob2d) are summed on the host in
// cell order, as the reference does serially.
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "fimhip.h"
#include "host_common.h"

namespace fimhip {
namespace {

constexpr double H_GRVITY = 9.80665;
constexpr double H_ONECM = 0.01 * (1000. * 9.80665);
constexpr double H_ASSLDP = .25;
constexpr double H_LAND = -1.e33;
constexpr double H_EPSIL = 1.e-11;

struct HycGrid {
  int nip, kdm;
  const int *prox, *proxs, *nprox, *wet;
  const double *area, *rarea, *sve, *sideln, *rsideln, *rprox_ln;
};
struct HycState {
  const double *um, *vm, *un, *vn, *dpm_edg;
  double *dp, *pres, *massflx, *mssfx_avg;
  // work arrays
  double *antiflx, *dp_lo, *dpold, *flxplus, *flxmnus, *rplus, *rmnus, *del2p, *dptmp, *flxdv;
  double *clippd, *rpbot, *aux;
  int *flag;
};

__device__ __forceinline__ double hmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double hmax(double a, double b) { return a > b ? a : b; }

// column c of the block and the lane's layer
template <int LPC>
struct Col {
  int i, k;
  bool ok;
};
template <int LPC>
__device__ __forceinline__ Col<LPC> hcol(int ncol) {
  constexpr int CPB = 256 / LPC;
  Col<LPC> c;
  c.i = blockIdx.x * CPB + threadIdx.x / LPC;
  c.k = threadIdx.x % LPC;
  c.ok = c.i < ncol;
  return c;
}
// value of lane j of this column's LPC lanes
template <int LPC>
__device__ __forceinline__ double bcast(double v, int j) { return __shfl(v, j, LPC); }

// pres(k+1) = pres(k) + d(k), k = 0..kdm-1, from pres(0) = p0: every lane forms
// the whole chain in layer order and keeps the interface below its layer
template <int LPC>
__device__ __forceinline__ double col_pres(double d, double p0, int k, int kdm) {
  double prev = p0, mine = 0.;
  for (int j = 0; j < kdm; j++) {
    prev = prev + bcast<LPC>(d, j);
    if (k == j) mine = prev;
  }
  return mine;
}
// the sum over the column's layers of v, in layer order from 0.
template <int LPC>
__device__ __forceinline__ double col_sum(double v, int kdm) {
  double acc = 0.;
  for (int j = 0; j < kdm; j++) acc = acc + bcast<LPC>(v, j);
  return acc;
}

#define HC(a, i, k) (a)[(size_t)(i) * kdm + (k)]
#define HP(i, k) s.pres[(size_t)(i) * (kdm + 1) + (k)]
#define HE(a, i, e, k) (a)[((size_t)(i) * NPP + (e)) * kdm + (k)]
#define HSVE(i, e, c) g.sve[((size_t)(i) * NPP + (e)) * 2 + (c)]

template <int LPC>
__global__ void __launch_bounds__(256) hyc_flux(HycGrid g, HycState s, int leapm, int leapn, double dtleap) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const double *dpm = s.dp + (size_t)(leapm - 1) * nip * kdm, *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  if (live)
    for (int e = 0; e < NPP; e++) HE(s.antiflx, i, e, k) = HE(s.massflx, i, e, k) = H_LAND;
  if (!g.wet[i]) {
    if (live) {
      HC(s.dp_lo, i, k) = 0.;
      HC(s.flxplus, i, k) = H_EPSIL;
      HC(s.flxmnus, i, k) = -H_EPSIL;
    }
    return;
  }
  const int np = g.nprox[i];
  double mf[NPP], af[NPP];
  for (int e = 0; e < NPP; e++) mf[e] = af[e] = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    const int ix = g.prox[i * NPP + e];
    if (!live) continue;
    if (!g.wet[ix]) {
      HE(s.massflx, i, e, k) = HE(s.antiflx, i, e, k) = 0.;
      continue;
    }
    const int ex = g.proxs[i * NPP + e];
    double valu1 = HSVE(i, e, 1) * HE(s.um, i, e, k) - HSVE(i, e, 0) * HE(s.vm, i, e, k);
    double valu2 = HSVE(ix, ex, 1) * HE(s.um, ix, ex, k) - HSVE(ix, ex, 0) * HE(s.vm, ix, ex, k);
    const double outgom = .5 * (valu1 - valu2);
    valu1 = g.rsideln[i * NPP + e] * dtleap * g.rprox_ln[i * NPP + e];
    valu2 = g.rsideln[ix * NPP + ex] * dtleap * g.rprox_ln[ix * NPP + ex];
    const double q = hmin(fabs(outgom) * .5 * (valu1 + valu2), 1.);
    const double dpedg = .5 * (HE(s.dpm_edg, i, e, k) + HE(s.dpm_edg, ix, ex, k));
    double hi = outgom >= 0. ? dpedg * (1. - q) + HC(dpm, i, k) * q : dpedg * (1. - q) + HC(dpm, ix, k) * q;
    hi = hi * outgom;
    valu1 = HSVE(i, e, 1) * HE(s.un, i, e, k) - HSVE(i, e, 0) * HE(s.vn, i, e, k);
    valu2 = HSVE(ix, ex, 1) * HE(s.un, ix, ex, k) - HSVE(ix, ex, 0) * HE(s.vn, ix, ex, k);
    const double outgon = .5 * (valu1 - valu2);
    const double lo = outgon >= 0. ? outgon * HC(dpn, i, k) : outgon * HC(dpn, ix, k);
    af[e] = hi - lo;
    mf[e] = lo;
    HE(s.antiflx, i, e, k) = af[e];
    HE(s.massflx, i, e, k) = mf[e];
  }
  double flx = 0., d = 0.;
  if (live) {
#pragma unroll
    for (int e = 0; e < NPP; e++)
      if (e < np) flx = flx + mf[e];
    d = HC(dpn, i, k);
    HC(s.dpold, i, k) = d;
  }
  const double p = col_pres<LPC>(d, HP(i, 0), k, kdm);
  if (!live) return;
  HP(i, k + 1) = p;
  double lo = d - flx * g.rarea[i] * dtleap;
  if (lo < 0. && lo + H_ONECM > 0.) lo = 0.;
  HC(s.dp_lo, i, k) = lo;
  double fp = H_EPSIL, fm = -H_EPSIL;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    if (af[e] >= 0.) fp = fp + af[e];
    else fm = fm + af[e];
  }
  HC(s.flxplus, i, k) = fp;
  HC(s.flxmnus, i, k) = fm;
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_limits(HycGrid g, HycState s, int leapn, double dtleap) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  if (!g.wet[i]) {
    HC(s.rplus, i, k) = HC(s.rmnus, i, k) = 0.;
    return;
  }
  const double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  const double lo = HC(s.dp_lo, i, k);
  double dpmax = hmax(lo, HC(dpn, i, k)), dpmin = hmin(lo, HC(dpn, i, k));
  const int np = g.nprox[i];
  for (int e = 0; e < np; e++) {
    const int ix = g.prox[i * NPP + e];
    if (!g.wet[ix]) continue;
    dpmax = hmax(hmax(dpmax, HC(s.dp_lo, ix, k)), HC(dpn, ix, k));
    dpmin = hmin(hmin(dpmin, HC(s.dp_lo, ix, k)), HC(dpn, ix, k));
  }
  dpmin = hmax(0., dpmin);
  HC(s.rmnus, i, k) = (lo - dpmin) * g.area[i] / (dtleap * HC(s.flxplus, i, k));
  HC(s.rplus, i, k) = (lo - dpmax) * g.area[i] / (dtleap * HC(s.flxmnus, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_clip(HycGrid g, HycState s, int leapn, double dtleap) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  double cl[NPP];
  for (int e = 0; e < NPP; e++) cl[e] = 0.;
  if (g.wet[i]) {
    const int np = g.nprox[i];
    double flx = 0.;
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      if (e >= np) break;
      const int ix = g.prox[i * NPP + e];
      double part = 0., a = 0.;         // an edge to land carries no flux
      if (g.wet[ix]) {
        if (live) {
          a = HE(s.antiflx, i, e, k);
          const double clip = a >= 0. ? hmin(hmin(1., HC(s.rmnus, i, k)), HC(s.rplus, ix, k))
                                      : hmin(hmin(1., HC(s.rplus, i, k)), HC(s.rmnus, ix, k));
          part = a * (1. - clip);
          a = a * clip;
          HE(s.massflx, i, e, k) = HE(s.massflx, i, e, k) + a;
        }
        cl[e] = col_sum<LPC>(part, kdm);
      }
      flx = flx + a;
    }
    double d = 0.;
    double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
    if (live) {
      d = HC(s.dp_lo, i, k) - flx * g.rarea[i] * dtleap;
      if (d < 0. && d + H_ONECM > 0.) d = 0.;
      HC(dpn, i, k) = d;
    }
    const double p = col_pres<LPC>(d, HP(i, 0), k, kdm);
    if (live) HP(i, k + 1) = p;
    if (k == kdm - 1) s.rpbot[i] = 1. / p;
  }
  if (k == 0)
    for (int e = 0; e < NPP; e++) s.clippd[i * NPP + e] = cl[e];
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_recover(HycGrid g, HycState s, int leapn) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  double flx = 0.;
  if (g.wet[i]) {
    const double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
    const int np = g.nprox[i];
    for (int e = 0; e < np; e++) {
      const int ix = g.prox[i * NPP + e];
      if (!g.wet[ix]) continue;
      const double cl = s.clippd[i * NPP + e];
      double q = 0.;
      if (cl > 0.) q = cl * (HC(dpn, i, k) * s.rpbot[i]);
      else if (cl < 0.) q = cl * (HC(dpn, ix, k) * s.rpbot[ix]);
      flx = flx + q;
      HE(s.massflx, i, e, k) = HE(s.massflx, i, e, k) + q;
    }
  }
  HC(s.flxdv, i, k) = flx;
}

// the recovered flux's update (465-491); with interface smoothing, dp again
// from the interface pressures and del2p's own part (522-548)
template <int LPC>
__global__ void __launch_bounds__(256) hyc_update(HycGrid g, HycState s, int leapn, double dtleap, int smooth,
                                                  double biharm) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  // pk: the pressure of interface k (the top of layer k) after the update
  double pk = 0.;
  if (g.wet[i]) {
    double d = 0.;
    if (live) {
      d = HC(dpn, i, k) - HC(s.flxdv, i, k) * g.rarea[i] * dtleap;
      if (d < 0. && d + H_ONECM > 0.) d = 0.;
      HC(dpn, i, k) = d;
    }
    const double p0 = HP(i, 0);
    const double p = col_pres<LPC>(d, p0, k, kdm);
    if (live) HP(i, k + 1) = p;
    const double pa = __shfl_up(p, 1, LPC);
    pk = k > 0 ? pa : p0;
    if (smooth && live) HC(dpn, i, k) = p - pk;
  } else if (live) {
    pk = HP(i, k);
  }
  if (smooth && live) HC(s.del2p, i, k) = biharm == 0. && k >= 1 ? pk : H_LAND;
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_del2p(HycGrid g, HycState s, double biharm) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm || c.k < 1) return;
  const int i = c.i, k = c.k;
  if (!g.wet[i]) return;
  const int np = g.nprox[i];
  double sum = 0.;
  int ncount = 0;
  const double pi = HP(i, k), pbi = HP(i, kdm);
  for (int e = 0; e < np; e++) {
    const int ix = g.prox[i * NPP + e];
    if (!g.wet[ix]) continue;
    ncount = ncount + 1;
    const double px = HP(ix, k), pbx = HP(ix, kdm);
    double deltap;
    if (pbx < pi) deltap = pbx - px;
    else if (pbi < px) deltap = pi - pbi;
    else deltap = pi - px;
    sum = sum + deltap;
  }
  HC(s.del2p, i, k) = biharm * sum / (double)ncount + (1. - biharm) * pi;
}

// bolus fluxes (605-646; the antidiffusive-flux array holds them) and, for
// the biharmonic form, the ratio of each layer's mass to its outflow (664-683)
template <int LPC>
__global__ void __launch_bounds__(256) hyc_bolus(HycGrid g, HycState s, int leapn, double dtleap, double thkdff,
                                                 double biharm) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const bool wi = g.wet[i] != 0;
  if (live)
    for (int e = 0; e < NPP; e++) HE(s.antiflx, i, e, k) = wi ? 0. : H_LAND;
  if (!wi) return;
  const int np = g.nprox[i];
  double fp = H_EPSIL;
  for (int e = 0; e < np; e++) {
    const int ix = g.prox[i * NPP + e];
    double f = 0.;
    if (g.wet[ix]) {
      double b = 0.;      // the limited interface flux of interface k (layers k-1|k), lanes 1..kdm-1
      if (live && k >= 1) {
        const double q = dtleap * thkdff * (HC(s.del2p, i, k) - HC(s.del2p, ix, k)) * g.sideln[i * NPP + e];
        const double flxmx = (HP(ix, kdm) - HP(ix, k)) * g.area[ix] / (double)g.nprox[ix];
        const double flxmn = -((HP(i, kdm) - HP(i, k)) * g.area[i] / (double)np);
        b = hmin(flxmx * .999, hmax(flxmn * .999, q));
      }
      // layer k's flux: interface k+1's minus interface k's (0 above the
      // top), the bottom layer minus interface kdm-1's
      const double bn = __shfl_down(b, 1, LPC);
      if (live) f = k < kdm - 1 ? bn - b : -b;
      if (live) HE(s.antiflx, i, e, k) = f;
    }
    if (biharm > 0.) fp = fp + hmax(0., f);
  }
  if (biharm > 0. && live) {
    const double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
    HC(s.rmnus, i, k) = HC(dpn, i, k) * g.area[i] / fp;
  }
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_bclip(HycGrid g, HycState s, int leapn, double biharm, double masscor) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  double cl[NPP];
  for (int e = 0; e < NPP; e++) cl[e] = 0.;
  if (g.wet[i]) {
    const int np = g.nprox[i];
    double flx = 0.;
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      if (e >= np) break;
      const int ix = g.prox[i * NPP + e];
      double b = live ? HE(s.antiflx, i, e, k) : 0.;
      if (biharm > 0. && g.wet[ix]) {
        double part = 0.;
        if (live && b != 0.) {
          const double clip = b > 0. ? hmin(1., HC(s.rmnus, i, k)) : hmin(1., HC(s.rmnus, ix, k));
          part = b * (1. - clip);
          b = b * clip;
        }
        cl[e] = col_sum<LPC>(part, kdm);
      }
      flx = flx + b;
    }
    double d = 0.;
    double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
    if (live) {
      d = HC(dpn, i, k) - flx * g.rarea[i];
      if (d < 0. && d + H_ONECM > 0.) d = 0.;
      d = d * (1. + masscor);
      HC(dpn, i, k) = d;
      HC(s.dptmp, i, k) = d;
    }
    const double p = col_pres<LPC>(d, HP(i, 0), k, kdm);
    if (live) HP(i, k + 1) = p;
    if (k == kdm - 1) s.rpbot[i] = 1. / p;
  }
  if (biharm > 0. && k == 0)
    for (int e = 0; e < NPP; e++) s.clippd[i * NPP + e] = cl[e];
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_final(HycGrid g, HycState s, int leapm, int leapn, double dtleap,
                                                 int smooth, double biharm) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  if (!g.wet[i]) {
    if (k == 0) s.aux[i] = 0.;
    return;
  }
  double *dpm = s.dp + (size_t)(leapm - 1) * nip * kdm, *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  const int np = g.nprox[i];
  if (smooth && biharm > 0.) {
    const double dtinv = 1. / dtleap;
    double flx = 0., d = 0.;
    if (live) {
      for (int e = 0; e < np; e++) {
        const int ix = g.prox[i * NPP + e];
        const double cl = s.clippd[i * NPP + e];
        double q = 0.;
        if (cl > 0.) q = cl * HC(s.dptmp, i, k) * s.rpbot[i];
        else if (cl < 0.) q = cl * HC(s.dptmp, ix, k) * s.rpbot[ix];
        flx = flx + q;
        HE(s.massflx, i, e, k) = HE(s.massflx, i, e, k) + q * dtinv;
      }
      d = HC(dpn, i, k) - flx * g.rarea[i];
      HC(dpn, i, k) = d;
    }
    const double p = col_pres<LPC>(d, HP(i, 0), k, kdm);
    if (live) HP(i, k + 1) = p;
  }
  if (!live) return;
  if (smooth)
    for (int e = 0; e < np; e++) HE(s.mssfx_avg, i, e, k) = HE(s.mssfx_avg, i, e, k) + HE(s.massflx, i, e, k);
  HC(dpm, i, k) = H_ASSLDP * (HC(s.dpold, i, k) + HC(dpn, i, k)) + (1. - H_ASSLDP - H_ASSLDP) * HC(dpm, i, k);
  if (k == kdm - 1) s.aux[i] = HP(i, kdm);
}

// duplicate edge fluxes must be equal and opposite bit for bit (chkdup)
template <int LPC>
__global__ void __launch_bounds__(256) hyc_check(HycGrid g, HycState s) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm || !g.wet[c.i]) return;
  const int i = c.i, k = c.k;
  const int np = g.nprox[i];
  for (int e = 0; e < np; e++) {
    const int ix = g.prox[i * NPP + e], ex = g.proxs[i * NPP + e];
    if (!g.wet[ix]) continue;
    const double a = HE(s.massflx, i, e, k), b = HE(s.massflx, ix, ex, k);
    if (a + b != 0. && a - b != 0.) atomicOr(s.flag, 1);
  }
}

}  // namespace
}  // namespace fimhip

using namespace fimhip;

struct fimhip_ocn {
  int device = 0, nip = 0, kdm = 0;
  double dtleap = 0., thkdff = 0., biharm = 0.;
  double totmass = 0., massglb0 = 0., masscor = 0.;
  hipStream_t stream = nullptr;
  Arena arena;
  HycGrid g{};
  HycState s{};
  std::vector<int> wet;
  std::vector<double> area;
  KTimer *kt = nullptr;              // non-null while profiling
};

static int ocn_launch(fimhip_ocn *o, int nstep, int leapm, int leapn);

extern "C" int fimhip_ocn_create(const fimhip_grid_desc *gd, const double *rprox_ln, int kdm, const int *wet,
                                 double dtleap, double thkdff, double biharm, int device, fimhip_ocn **out) {
  if (!gd || !rprox_ln || !wet || !out || kdm < 2 || kdm > 64 || gd->nip < 12 || !gd->nprox || !gd->prox ||
      !gd->proxs || !gd->sidevec_e || !gd->sideln || !gd->area || !(dtleap > 0.))
    return FIMHIP_EINVAL;
  *out = nullptr;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return FIMHIP_ENODEV;
  HIPCHK(hipSetDevice(device));
  fimhip_ocn *o = new fimhip_ocn();
  o->device = device;
  o->nip = gd->nip;
  o->kdm = kdm;
  o->dtleap = dtleap;
  o->thkdff = thkdff;
  o->biharm = biharm;
  if (hipStreamCreate(&o->stream) != hipSuccess) { delete o; return FIMHIP_EHIP; }
  o->arena.stream = o->stream;
  const int nip = gd->nip;
  std::vector<int> prox(nip * NPP), proxs(nip * NPP), nprox(nip);
  std::vector<double> rarea(nip), rsideln(nip * NPP, 0.), sve(nip * NPP * 2), sideln(nip * NPP), rpl(nip * NPP);
  o->wet.assign(wet, wet + nip);
  o->area.assign(gd->area, gd->area + nip);
  for (int i = 0; i < nip; i++) {
    const int np = gd->nprox[i];
    if (np < 5 || np > 6) { fimhip_ocn_destroy(o); return FIMHIP_EINVAL; }
    nprox[i] = np;
    o->wet[i] = wet[i] > 0 ? 1 : 0;
    rarea[i] = 1. / gd->area[i];
    for (int e = 0; e < NPP; e++) {
      const int ee = e < np ? e : np - 1;            // a pentagon's slot 6 repeats slot 5
      const int x = gd->prox[i * NPP + ee] - 1, xs = gd->proxs[i * NPP + ee] - 1;
      if (x < 0 || x >= nip || xs < 0 || xs >= NPP) { fimhip_ocn_destroy(o); return FIMHIP_EINVAL; }
      prox[i * NPP + e] = x;
      proxs[i * NPP + e] = xs;
      sideln[i * NPP + e] = gd->sideln[i * NPP + e];
      rpl[i * NPP + e] = rprox_ln[i * NPP + e];
      if (e < np) rsideln[i * NPP + e] = 1. / gd->sideln[i * NPP + e];
      sve[(i * NPP + e) * 2] = gd->sidevec_e[(i * NPP + e) * 2];
      sve[(i * NPP + e) * 2 + 1] = gd->sidevec_e[(i * NPP + e) * 2 + 1];
    }
  }
  int r = 0;
  HycGrid &g = o->g;
  g.nip = nip;
  g.kdm = kdm;
  int *pp, *ps, *pn, *pw;
  double *pa, *pr, *pv, *psl, *prs, *prl;
  if ((r = o->arena.upload(&pp, prox)) || (r = o->arena.upload(&ps, proxs)) || (r = o->arena.upload(&pn, nprox)) ||
      (r = o->arena.upload(&pw, o->wet)) || (r = o->arena.upload(&pa, o->area)) || (r = o->arena.upload(&pr, rarea)) ||
      (r = o->arena.upload(&pv, sve)) || (r = o->arena.upload(&psl, sideln)) ||
      (r = o->arena.upload(&prs, rsideln)) || (r = o->arena.upload(&prl, rpl))) {
    fimhip_ocn_destroy(o);
    return r;
  }
  g.prox = pp; g.proxs = ps; g.nprox = pn; g.wet = pw;
  g.area = pa; g.rarea = pr; g.sve = pv; g.sideln = psl; g.rsideln = prs; g.rprox_ln = prl;
  const size_t cz = (size_t)nip * kdm, ez = cz * NPP;
  HycState &s = o->s;
  double *um, *vm, *un, *vn, *dpe;
  struct { double **p; size_t n; } arrs[] = {
      {&um, ez}, {&vm, ez}, {&un, ez}, {&vn, ez}, {&dpe, ez}, {&s.dp, 2 * cz}, {&s.pres, (size_t)nip * (kdm + 1)},
      {&s.massflx, ez}, {&s.mssfx_avg, ez}, {&s.antiflx, ez}, {&s.dp_lo, cz}, {&s.dpold, cz}, {&s.flxplus, cz},
      {&s.flxmnus, cz}, {&s.rplus, cz}, {&s.rmnus, cz}, {&s.del2p, cz}, {&s.dptmp, cz}, {&s.flxdv, cz},
      {&s.clippd, (size_t)nip * NPP}, {&s.rpbot, (size_t)nip}, {&s.aux, (size_t)nip}};
  for (auto &a : arrs)
    if ((r = o->arena.alloc(a.p, a.n))) { fimhip_ocn_destroy(o); return r; }
  s.um = um; s.vm = vm; s.un = un; s.vn = vn; s.dpm_edg = dpe;
  if ((r = o->arena.alloc(&s.flag, 1))) { fimhip_ocn_destroy(o); return r; }
  if (hipStreamSynchronize(o->stream) != hipSuccess) { fimhip_ocn_destroy(o); return FIMHIP_EHIP; }
  *out = o;
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_destroy(fimhip_ocn *o) {
  if (!o) return FIMHIP_EINVAL;
  (void)hipSetDevice(o->device);
  if (o->stream) (void)hipStreamSynchronize(o->stream);
  o->arena.release();
  delete o->kt;
  if (o->stream) (void)hipStreamDestroy(o->stream);
  delete o;
  return FIMHIP_OK;
}

static double *ocn_field(fimhip_ocn *o, int f, size_t *n) {
  const size_t cz = (size_t)o->nip * o->kdm, ez = cz * NPP;
  HycState &s = o->s;
  switch (f) {
    case FIMHIP_OCN_DP: *n = 2 * cz; return s.dp;
    case FIMHIP_OCN_PRES: *n = (size_t)o->nip * (o->kdm + 1); return s.pres;
    case FIMHIP_OCN_UM_EDG: *n = ez; return const_cast<double *>(s.um);
    case FIMHIP_OCN_VM_EDG: *n = ez; return const_cast<double *>(s.vm);
    case FIMHIP_OCN_UN_EDG: *n = ez; return const_cast<double *>(s.un);
    case FIMHIP_OCN_VN_EDG: *n = ez; return const_cast<double *>(s.vn);
    case FIMHIP_OCN_DPM_EDG: *n = ez; return const_cast<double *>(s.dpm_edg);
    case FIMHIP_OCN_MASSFLX: *n = ez; return s.massflx;
    case FIMHIP_OCN_MSSFX_AVG: *n = ez; return s.mssfx_avg;
  }
  return nullptr;
}

extern "C" long long fimhip_ocn_field_size(const fimhip_ocn *o, int f) {
  if (!o) return FIMHIP_EINVAL;
  size_t n = 0;
  if (!ocn_field(const_cast<fimhip_ocn *>(o), f, &n)) return FIMHIP_EINVAL;
  return (long long)n;
}

extern "C" int fimhip_ocn_upload(fimhip_ocn *o, int f, const double *h) {
  if (!o || !h) return FIMHIP_EINVAL;
  size_t n = 0;
  double *d = ocn_field(o, f, &n);
  if (!d) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  HIPCHK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, o->stream));
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_download(fimhip_ocn *o, int f, double *h) {
  if (!o || !h) return FIMHIP_EINVAL;
  size_t n = 0;
  double *d = ocn_field(o, f, &n);
  if (!d) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  HIPCHK(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, o->stream));
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}

template <int LPC>
static void ocn_kernels(fimhip_ocn *o, int leapm, int leapn) {
  const HycGrid &g = o->g;
  const HycState &s = o->s;
  hipStream_t st = o->stream;
  const dim3 grid((o->nip + 256 / LPC - 1) / (256 / LPC)), blk(256);
  const double dt = o->dtleap;
  const int smooth = o->thkdff != 0.;
  KT_LAUNCH("hyc_flux", st, hyc_flux<LPC><<<grid, blk, 0, st>>>(g, s, leapm, leapn, dt));
  KT_LAUNCH("hyc_limits", st, hyc_limits<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt));
  KT_LAUNCH("hyc_clip", st, hyc_clip<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt));
  KT_LAUNCH("hyc_recover", st, hyc_recover<LPC><<<grid, blk, 0, st>>>(g, s, leapn));
  KT_LAUNCH("hyc_update", st, hyc_update<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt, smooth, o->biharm));
  if (smooth) {
    if (o->biharm > 0.) KT_LAUNCH("hyc_del2p", st, hyc_del2p<LPC><<<grid, blk, 0, st>>>(g, s, o->biharm));
    KT_LAUNCH("hyc_bolus", st, hyc_bolus<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt, o->thkdff, o->biharm));
    KT_LAUNCH("hyc_bclip", st, hyc_bclip<LPC><<<grid, blk, 0, st>>>(g, s, leapn, o->biharm, o->masscor));
  }
  KT_LAUNCH("hyc_final", st, hyc_final<LPC><<<grid, blk, 0, st>>>(g, s, leapm, leapn, dt, smooth, o->biharm));
  hyc_check<LPC><<<grid, blk, 0, st>>>(g, s);
}

// one call of cnuity on the device-resident arrays (hycom_cnuity.F90:21-361)
static int ocn_launch(fimhip_ocn *o, int nstep, int leapm, int leapn) {
  if (nstep < 1 || leapm < 1 || leapm > 2 || leapn < 1 || leapn > 2 || leapm == leapn) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  HIPCHK(hipMemsetAsync(o->s.flag, 0, sizeof(int), o->stream));
  g_ktimer = o->kt;
  if (o->kdm <= 32) ocn_kernels<32>(o, leapm, leapn);
  else ocn_kernels<64>(o, leapm, leapn);
  g_ktimer = nullptr;
  HIPCHK(hipGetLastError());
  // the drift correction of the total mass, at the first step and every 10
  // days (977-998): glob2d sums bottom pressure times area over the ocean in
  // cell order
  const double day = (double)nstep * .5 * o->dtleap / 86400.;
  int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, o->s.flag, sizeof(int), hipMemcpyDeviceToHost, o->stream));
  if (nstep == 1 || fmod(day + .0001, 10.) < .0002) {
    std::vector<double> pb(o->nip);
    HIPCHK(hipMemcpyAsync(pb.data(), o->s.aux, pb.size() * sizeof(double), hipMemcpyDeviceToHost, o->stream));
    HIPCHK(hipStreamSynchronize(o->stream));
    double tot = 0.;
    for (int i = 0; i < o->nip; i++)
      if (o->wet[i]) tot = tot + pb[i] * o->area[i];
    o->totmass = tot / H_GRVITY;
    if (nstep == 1) o->massglb0 = o->totmass;
    o->masscor = o->dtleap * (o->massglb0 - o->totmass) / (o->massglb0 * 10. * 86400.);
  } else {
    HIPCHK(hipStreamSynchronize(o->stream));
  }
  // the reference stops the program here (chkdup, hycom_cnuity.F90:1030-1075)
  return flag ? FIMHIP_EDUP : FIMHIP_OK;
}

extern "C" int fimhip_ocn_cnuity(fimhip_ocn *o, int nstep, int leapm, int leapn) {
  if (!o) return FIMHIP_EINVAL;
  return ocn_launch(o, nstep, leapm, leapn);
}

extern "C" int fimhip_ocn_cnuity_r8(fimhip_ocn *o, int nstep, int leapm, int leapn, const double *um_edg,
                                   const double *vm_edg, const double *un_edg, const double *vn_edg,
                                   const double *dpm_edg, double *dp, double *pres, double *massflx,
                                   double *mssfx_avg) {
  if (!o || !um_edg || !vm_edg || !un_edg || !vn_edg || !dpm_edg || !dp || !pres || !massflx || !mssfx_avg)
    return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  const double *in[] = {um_edg, vm_edg, un_edg, vn_edg, dpm_edg, dp, pres, massflx, mssfx_avg};
  const int fi[] = {FIMHIP_OCN_UM_EDG, FIMHIP_OCN_VM_EDG, FIMHIP_OCN_UN_EDG, FIMHIP_OCN_VN_EDG, FIMHIP_OCN_DPM_EDG,
                    FIMHIP_OCN_DP, FIMHIP_OCN_PRES, FIMHIP_OCN_MASSFLX, FIMHIP_OCN_MSSFX_AVG};
  for (int j = 0; j < 9; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fi[j], &n);
    HIPCHK(hipMemcpyAsync(d, in[j], n * sizeof(double), hipMemcpyHostToDevice, o->stream));
  }
  const int r = ocn_launch(o, nstep, leapm, leapn);
  if (r && r != FIMHIP_EDUP) return r;
  double *outp[] = {dp, pres, massflx, mssfx_avg};
  const int fo[] = {FIMHIP_OCN_DP, FIMHIP_OCN_PRES, FIMHIP_OCN_MASSFLX, FIMHIP_OCN_MSSFX_AVG};
  for (int j = 0; j < 4; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fo[j], &n);
    HIPCHK(hipMemcpyAsync(outp[j], d, n * sizeof(double), hipMemcpyDeviceToHost, o->stream));
  }
  HIPCHK(hipStreamSynchronize(o->stream));
  return r;
}

extern "C" int fimhip_ocn_mass(const fimhip_ocn *o, double *totmass, double *massglb0, double *masscor) {
  if (!o || !totmass || !massglb0 || !masscor) return FIMHIP_EINVAL;
  *totmass = o->totmass;
  *massglb0 = o->massglb0;
  *masscor = o->masscor;
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_set_mass(fimhip_ocn *o, double totmass, double massglb0, double masscor) {
  if (!o) return FIMHIP_EINVAL;
  o->totmass = totmass;
  o->massglb0 = massglb0;
  o->masscor = masscor;
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_profile(fimhip_ocn *o, int on) {
  if (!o) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  HIPCHK(hipStreamSynchronize(o->stream));
  delete o->kt;
  o->kt = on ? new KTimer() : nullptr;
  return FIMHIP_OK;
}

// "name count total_ms" lines of the kernels timed since profiling began
extern "C" int fimhip_ocn_profile_report(fimhip_ocn *o, char *buf, int len) {
  if (!o || !buf || len < 1) return FIMHIP_EINVAL;
  buf[0] = 0;
  if (!o->kt) return FIMHIP_OK;
  HIPCHK(hipSetDevice(o->device));
  const int r = o->kt->flush(o->stream);
  if (r) return r;
  std::string out;
  char line[256];
  for (auto &kv : o->kt->acc) {
    snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.second, kv.second.first);
    out += line;
  }
  if ((int)out.size() >= len) return FIMHIP_EINVAL;
  memcpy(buf, out.c_str(), out.size() + 1);
  return FIMHIP_OK;
}
