// The HYCOM ocean's continuity equation (FIMsrc/fim/horizontal/hycom_cnuity.F90,
// subroutine cnuity) on the GPU, in FP64, for a grid whose cells are ocean
// or land (wet).
//
// Each column runs on LPC lanes of a wavefront (LPC = 32 for up to 32
// layers, two columns per wavefront; 64 otherwise), lane k holding layer k.
// Every operation is the reference's, in its order: the sums over a cell's
// edges run in edge order per lane, and the two sums the reference forms
// along the column (interface pressure pres(k+1) = pres(k) + dp(k), and the
// clipped part of each edge's flux) are exact sequential chains over values
// the column's lanes exchange through LDS (a column never leaves its
// wavefront, so the exchange needs no workgroup barrier).  The neighbour
// values of all six edges are loaded unconditionally (every neighbour index
// is a valid cell) and the land/sea mask selects afterwards, so each lane has
// all its gathers in flight at once; a per-cell word packs the edge count,
// the cell's own mask and the mask of its six neighbours.  Blocks are ordered
// so that each XCD works on a contiguous eighth of the cells and neighbour
// rows are found in its L2.  A kernel boundary stands wherever a stage reads
// a neighbour's value that the previous stage writes:
//   hyc_flux     low/high-order edge fluxes, each edge's pair once (98-159)
//   hyc_lowsum   low-order thickness, pressure, in/outflow sums of the
//                antidiffusive fluxes                           (169-265)
//   hyc_limits   local bounds and limiter ratios              (269-300)
//   hyc_clip     limited antidiffusive fluxes and update      (315-393)
//   hyc_recover  the clipped flux as a barotropic correction and its
//                thickness update; dp from pres, del2p's own part (434-548)
//   hyc_del2p    the biharmonic del2p from the neighbours     (553-603)
//   hyc_bolus    interface-slope bolus fluxes, outflow ratios (605-683)
//   hyc_bclip    limited bolus fluxes and their update        (685-776)
//   hyc_final    the clipped bolus correction, the mass-flux average, the
//                leapfrog time smoothing, the bottom pressure (778-954)
// The thickness after the antidiffusive update goes to its own array (dpc),
// which hyc_recover reads at the neighbours while it writes the new level.
// The duplicate-flux check (chkdup) compares each pair of edge fluxes once
// and sets a flag; the global mass and its drift correction (glob2d) are
// summed on the host in cell order, as the reference does serially.
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "hycom_internal.h"

namespace fimhip {
namespace hyc {
namespace {


// the high- and low-order fluxes of one edge seen from the cell whose values
// are 'a' (its edge ea) towards the cell 'b' (its edge eb)  (98-159)
struct EdgeFlux {
  double anti, low;
};
__device__ __forceinline__ EdgeFlux edge_flux(const HycGrid &g, const HycState &s, const double *dpm,
                                              const double *dpn, int kdm, int a, int ea, int b, int eb, int k,
                                              double dtleap) {
  double valu1 = HSVE(a, ea, 1) * HE(s.um, a, ea, k) - HSVE(a, ea, 0) * HE(s.vm, a, ea, k);
  double valu2 = HSVE(b, eb, 1) * HE(s.um, b, eb, k) - HSVE(b, eb, 0) * HE(s.vm, b, eb, k);
  const double outgom = .5 * (valu1 - valu2);
  valu1 = g.rsideln[a * NPP + ea] * dtleap * g.rprox_ln[a * NPP + ea];
  valu2 = g.rsideln[b * NPP + eb] * dtleap * g.rprox_ln[b * NPP + eb];
  const double q = hmin(fabs(outgom) * .5 * (valu1 + valu2), 1.);
  const double dpedg = .5 * (HE(s.dpm_edg, a, ea, k) + HE(s.dpm_edg, b, eb, k));
  double hi = outgom >= 0. ? dpedg * (1. - q) + HC(dpm, a, k) * q : dpedg * (1. - q) + HC(dpm, b, k) * q;
  hi = hi * outgom;
  valu1 = HSVE(a, ea, 1) * HE(s.un, a, ea, k) - HSVE(a, ea, 0) * HE(s.vn, a, ea, k);
  valu2 = HSVE(b, eb, 1) * HE(s.un, b, eb, k) - HSVE(b, eb, 0) * HE(s.vn, b, eb, k);
  const double outgon = .5 * (valu1 - valu2);
  const double lo = outgon >= 0. ? outgon * HC(dpn, a, k) : outgon * HC(dpn, b, k);
  return EdgeFlux{hi - lo, lo};
}

// the edge fluxes, each pair of an edge's two sides once: the cell of lower
// index evaluates the reference's expression from both sides (the same
// operations with the two cells' roles exchanged) and writes both
template <int LPC>
__global__ void __launch_bounds__(256) hyc_flux(HycGrid g, HycState s, int leapm, int leapn, double dtleap) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  const int ci = g.cinfo[i];
  int ix[NPP], ex[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    ix[e] = g.prox[i * NPP + e];
    ex[e] = g.proxs[i * NPP + e];
  }
  if (!ci_wet(ci)) {      // (the work array of antidiffusive fluxes is never read on land)
#pragma unroll
    for (int e = 0; e < NPP; e++) HE(s.massflx, i, e, k) = H_LAND;
    return;
  }
  const double *dpm = s.dp + (size_t)(leapm - 1) * nip * kdm, *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  const int np = ci_np(ci);
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const bool open = ci_open(ci, e);
    if (open && i < ix[e]) {
      const EdgeFlux fa = edge_flux(g, s, dpm, dpn, kdm, i, e, ix[e], ex[e], k, dtleap);
      const EdgeFlux fb = edge_flux(g, s, dpm, dpn, kdm, ix[e], ex[e], i, e, k, dtleap);
      __builtin_nontemporal_store(fa.anti, &HE(s.antiflx, i, e, k));
      __builtin_nontemporal_store(fa.low, &HE(s.massflx, i, e, k));
      __builtin_nontemporal_store(fb.anti, &HE(s.antiflx, ix[e], ex[e], k));
      __builtin_nontemporal_store(fb.low, &HE(s.massflx, ix[e], ex[e], k));
    } else if (!open) {      // an edge to land carries no flux; a pentagon's sixth slot is land-valued
      HE(s.antiflx, i, e, k) = HE(s.massflx, i, e, k) = e < np ? 0. : H_LAND;
    }
  }
}

// low-order thickness, pressure and the in/outflow sums of the antidiffusive
// fluxes (169-265)
template <int LPC>
__global__ void __launch_bounds__(256) hyc_lowsum(HycGrid g, HycState s, int leapn, double dtleap) {
  __shared__ double sh[256 / LPC][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  if (!ci_wet(ci)) {
    if (live) {
      __builtin_nontemporal_store(0., &HC(s.dp_lo, i, k));
      __builtin_nontemporal_store(H_EPSIL, &HC(s.flxplus, i, k));
      __builtin_nontemporal_store(-H_EPSIL, &HC(s.flxmnus, i, k));
    }
    return;
  }
  const double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  const int np = ci_np(ci);
  double mf[NPP], af[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    mf[e] = HE(s.massflx, i, e, kk);
    af[e] = HE(s.antiflx, i, e, kk);
  }
  const double d = HC(dpn, i, kk);
  double flx = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++)
    if (e < np) flx = flx + mf[e];
  if (live) __builtin_nontemporal_store(d, &HC(s.dpold, i, k));
  const double p = col_pres(sh[c.cc], live ? d : 0., HP(i, 0), k, kdm);
  if (!live) return;
  __builtin_nontemporal_store(p, &HP(i, k + 1));
  double lo = d - flx * g.rarea[i] * dtleap;
  if (lo < 0. && lo + H_ONECM > 0.) lo = 0.;
  __builtin_nontemporal_store(lo, &HC(s.dp_lo, i, k));
  double fp = H_EPSIL, fm = -H_EPSIL;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    if (af[e] >= 0.) fp = fp + af[e];
    else fm = fm + af[e];
  }
  __builtin_nontemporal_store(fp, &HC(s.flxplus, i, k));
  __builtin_nontemporal_store(fm, &HC(s.flxmnus, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_limits(HycGrid g, HycState s, int leapn, double dtleap) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) {
    HC(s.rplus, i, k) = HC(s.rmnus, i, k) = 0.;
    return;
  }
  const double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  const double lo = HC(s.dp_lo, i, k), dn = HC(dpn, i, k);
  double dpmax = hmax(lo, dn), dpmin = hmin(lo, dn);
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int ix = nb[e];
    const double a = HC(s.dp_lo, ix, k), b = HC(dpn, ix, k);
    const double nmax = hmax(hmax(dpmax, a), b), nmin = hmin(hmin(dpmin, a), b);
    const bool open = ci_open(ci, e);
    dpmax = open ? nmax : dpmax;
    dpmin = open ? nmin : dpmin;
  }
  dpmin = hmax(0., dpmin);
  __builtin_nontemporal_store((lo - dpmin) * g.area[i] / (dtleap * HC(s.flxplus, i, k)), &HC(s.rmnus, i, k));
  __builtin_nontemporal_store((lo - dpmax) * g.area[i] / (dtleap * HC(s.flxmnus, i, k)), &HC(s.rplus, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_clip(HycGrid g, HycState s, double dtleap) {
  __shared__ double sh[256 / LPC][NPP][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) {
    if (k < NPP) s.clippd[i * NPP + k] = 0.;
    return;
  }
  const int np = ci_np(ci);
  const double rmi = HC(s.rmnus, i, kk), rpi = HC(s.rplus, i, kk);
  double part[NPP], flx = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int ix = nb[e];
    const double a = HE(s.antiflx, i, e, kk), mfv = HE(s.massflx, i, e, kk);
    const double rpx = HC(s.rplus, ix, kk), rmx = HC(s.rmnus, ix, kk);
    const double clip = a >= 0. ? hmin(hmin(1., rmi), rpx) : hmin(hmin(1., rpi), rmx);
    const bool open = ci_open(ci, e);
    const double ac = a * clip;
    part[e] = open && live ? a * (1. - clip) : 0.;
    if (open && live) __builtin_nontemporal_store(mfv + ac, &HE(s.massflx, i, e, k));
    if (e < np) flx = flx + (open ? ac : a);    // a closed edge's antidiffusive flux is 0
  }
  const double cl = col_sums6<LPC>(sh[c.cc], part, k, kdm);
  if (k < NPP) s.clippd[i * NPP + k] = cl;
  double d = 0.;
  if (live) {
    d = HC(s.dp_lo, i, k) - flx * g.rarea[i] * dtleap;
    if (d < 0. && d + H_ONECM > 0.) d = 0.;
    __builtin_nontemporal_store(d, &HC(s.dpc, i, k));
  }
  const double p = col_pres(sh[c.cc][0], d, HP(i, 0), k, kdm);
  if (live) __builtin_nontemporal_store(p, &HP(i, k + 1));
  if (k == kdm - 1) s.rpbot[i] = 1. / p;
}

// the clipped fluxes back as an upstream barotropic correction and its
// thickness update (434-491); with interface smoothing, dp again from the
// interface pressures and del2p's own part (508-548)
template <int LPC>
__global__ void __launch_bounds__(256) hyc_recover(HycGrid g, HycState s, int leapn, double dtleap, int smooth,
                                                   double biharm) {
  __shared__ double sh[256 / LPC][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) {
    if (smooth && live) __builtin_nontemporal_store(biharm == 0. && k >= 1 ? HP(i, k) : H_LAND, &HC(s.del2p, i, k));
    return;
  }
  double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  const double dci = HC(s.dpc, i, kk), rpi = s.rpbot[i];
  double flx = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int ix = nb[e];
    const double cl = s.clippd[i * NPP + e], dcx = HC(s.dpc, ix, kk), rpx = s.rpbot[ix];
    const double mfv = HE(s.massflx, i, e, kk);
    const double q = cl > 0. ? cl * (dci * rpi) : (cl < 0. ? cl * (dcx * rpx) : 0.);
    if (ci_open(ci, e)) {
      flx = flx + q;
      if (live) __builtin_nontemporal_store(mfv + q, &HE(s.massflx, i, e, k));
    }
  }
  double d = 0.;
  if (live) {
    d = dci - flx * g.rarea[i] * dtleap;
    if (d < 0. && d + H_ONECM > 0.) d = 0.;
  }
  const double p0 = HP(i, 0);
  const double p = col_pres(sh[c.cc], d, p0, k, kdm);
  if (live) __builtin_nontemporal_store(p, &HP(i, k + 1));
  if (smooth) {
    const double pa = __shfl_up(p, 1, LPC);
    const double pk = k > 0 ? pa : p0;         // the interface above layer k
    if (live) {
      __builtin_nontemporal_store(p - pk, &HC(dpn, i, k));
      __builtin_nontemporal_store(biharm == 0. && k >= 1 ? pk : H_LAND, &HC(s.del2p, i, k));
    }
  } else if (live) {
    __builtin_nontemporal_store(d, &HC(dpn, i, k));
  }
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_del2p(HycGrid g, HycState s, double biharm) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm || c.k < 1) return;
  const int i = c.i, k = c.k;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  double sum = 0.;
  const double pi = HP(i, k), pbi = HP(i, kdm);
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int ix = nb[e];
    const double px = HP(ix, k), pbx = HP(ix, kdm);
    double deltap;
    if (pbx < pi) deltap = pbx - px;
    else if (pbi < px) deltap = pi - pbi;
    else deltap = pi - px;
    sum = ci_open(ci, e) ? sum + deltap : sum;
  }
  const int ncount = __builtin_popcount((ci >> 4) & 63);
  __builtin_nontemporal_store(biharm * sum / (double)ncount + (1. - biharm) * pi, &HC(s.del2p, i, k));
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
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;      // the bolus fluxes of land are never read
  const int np = ci_np(ci);
  const double pik = HP(i, kk), pbi = HP(i, kdm), di = HC(s.del2p, i, kk), ai = g.area[i];
  double fp = H_EPSIL;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int ix = nb[e];
    const double dx = HC(s.del2p, ix, kk), pxk = HP(ix, kk), pbx = HP(ix, kdm), ax = g.area[ix];
    const int npx = ci_np(g.cinfo[ix]);
    const double sl = g.sideln[i * NPP + e];
    // the limited flux of interface k (between layers k-1 and k), lanes 1..kdm-1
    double b = 0.;
    if (live && k >= 1) {
      const double q = dtleap * thkdff * (di - dx) * sl;
      const double flxmx = (pbx - pxk) * ax / (double)npx;
      const double flxmn = -((pbi - pik) * ai / (double)np);
      b = hmin(flxmx * .999, hmax(flxmn * .999, q));
    }
    // layer k's flux: interface k+1's minus interface k's (0 above the top),
    // the bottom layer minus interface kdm-1's
    const double bn = __shfl_down(b, 1, LPC);
    double f = k < kdm - 1 ? bn - b : -b;
    f = ci_open(ci, e) ? f : 0.;
    if (live) __builtin_nontemporal_store(f, &HE(s.antiflx, i, e, k));
    if (biharm > 0. && e < np) fp = fp + hmax(0., f);
  }
  if (biharm > 0. && live) {
    const double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
    __builtin_nontemporal_store(HC(dpn, i, k) * ai / fp, &HC(s.rmnus, i, k));
  }
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_bclip(HycGrid g, HycState s, int leapn, double biharm, double masscor) {
  __shared__ double sh[256 / LPC][NPP][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) {
    if (biharm > 0. && k < NPP) s.clippd[i * NPP + k] = 0.;
    return;
  }
  const int np = ci_np(ci);
  const double rmi = HC(s.rmnus, i, kk);
  double part[NPP], flx = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int ix = nb[e];
    double b = HE(s.antiflx, i, e, kk);
    const double rmx = HC(s.rmnus, ix, kk);
    double pt = 0.;
    if (biharm > 0.) {
      const double clip = b > 0. ? hmin(1., rmi) : hmin(1., rmx);
      const bool doit = ci_open(ci, e) && b != 0.;
      pt = doit ? b * (1. - clip) : 0.;
      b = doit ? b * clip : b;
    }
    part[e] = live ? pt : 0.;
    if (e < np) flx = flx + b;
  }
  if (biharm > 0.) {
    const double cl = col_sums6<LPC>(sh[c.cc], part, k, kdm);
    if (k < NPP) s.clippd[i * NPP + k] = cl;
  }
  double d = 0.;
  double *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  if (live) {
    d = HC(dpn, i, k) - flx * g.rarea[i];
    if (d < 0. && d + H_ONECM > 0.) d = 0.;
    d = d * (1. + masscor);
    __builtin_nontemporal_store(d, &HC(dpn, i, k));
    __builtin_nontemporal_store(d, &HC(s.dptmp, i, k));
  }
  const double p = col_pres(sh[c.cc][0], d, HP(i, 0), k, kdm);
  if (live) __builtin_nontemporal_store(p, &HP(i, k + 1));
  if (k == kdm - 1) s.rpbot[i] = 1. / p;
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_final(HycGrid g, HycState s, int leapm, int leapn, double dtleap,
                                                 int smooth, double biharm) {
  __shared__ double sh[256 / LPC][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) {
    if (k == 0) s.aux[i] = 0.;
    return;
  }
  double *dpm = s.dp + (size_t)(leapm - 1) * nip * kdm, *dpn = s.dp + (size_t)(leapn - 1) * nip * kdm;
  const int np = ci_np(ci);
  double mf[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) mf[e] = HE(s.massflx, i, e, kk);
  double dn = HC(dpn, i, kk);
  if (smooth && biharm > 0.) {
    const double dtinv = 1. / dtleap;
    const double dti = HC(s.dptmp, i, kk), rpi = s.rpbot[i];
    double flx = 0.;
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      const int ix = nb[e];
      const double cl = s.clippd[i * NPP + e], dtx = HC(s.dptmp, ix, kk), rpx = s.rpbot[ix];
      const double q = cl > 0. ? cl * dti * rpi : (cl < 0. ? cl * dtx * rpx : 0.);
      if (e < np) {
        flx = flx + q;
        mf[e] = mf[e] + q * dtinv;
        if (live) __builtin_nontemporal_store(mf[e], &HE(s.massflx, i, e, k));
      }
    }
    dn = dn - flx * g.rarea[i];
    if (live) __builtin_nontemporal_store(dn, &HC(dpn, i, k));
    const double p = col_pres(sh[c.cc], live ? dn : 0., HP(i, 0), k, kdm);
    if (live) __builtin_nontemporal_store(p, &HP(i, k + 1));
  }
  if (!live) return;
  if (smooth)
#pragma unroll
    for (int e = 0; e < NPP; e++)
      if (e < np) __builtin_nontemporal_store(HE(s.mssfx_avg, i, e, k) + mf[e], &HE(s.mssfx_avg, i, e, k));
  __builtin_nontemporal_store(H_ASSLDP * (HC(s.dpold, i, k) + dn) + (1. - H_ASSLDP - H_ASSLDP) * HC(dpm, i, k), &HC(dpm, i, k));
  if (k == kdm - 1) s.aux[i] = HP(i, kdm);
}

// duplicate edge fluxes must be equal and opposite bit for bit (chkdup); the
// cell of lower index checks each pair
template <int LPC>
__global__ void __launch_bounds__(256) hyc_check(HycGrid g, HycState s) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  const int ci = g.cinfo[i];
  int nb[NPP];      // neighbours, loaded before the mask is tested
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  bool bad = false;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int ix = nb[e], ex = g.proxs[i * NPP + e];
    if (!ci_open(ci, e) || ix < i) continue;
    const double a = HE(s.massflx, i, e, k), b = HE(s.massflx, ix, ex, k);
    bad = bad || (a + b != 0. && a - b != 0.);
  }
  if (bad) atomicOr(s.flag, 1);
}

}  // namespace
}  // namespace hyc
}  // namespace fimhip

using namespace fimhip;
using namespace fimhip::hyc;

static int ocn_launch(fimhip_ocn *o, int nstep, int leapm, int leapn);

extern "C" int fimhip_ocn_create(const fimhip_grid_desc *gd, const double *rprox_ln, int kdm, const int *wet,
                                 double dtleap, double thkdff, double biharm, int device, fimhip_ocn **out) {
  if (!gd || !rprox_ln || !wet || !out || kdm < 2 || kdm > 64 || gd->nip < 12 || !gd->nprox || !gd->prox ||
      !gd->proxs || !gd->sidevec_e || !gd->sideln || !gd->area || !gd->cs || !gd->sn || !gd->sidevec_c ||
      !gd->lat || !(dtleap > 0.))
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
    // the reference tests wet > 0 in some places and wet == 0 in others
    if (wet[i] < 0) { fimhip_ocn_destroy(o); return FIMHIP_EINVAL; }
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
  // every edge's twin must point back (each pair of edge fluxes is formed once)
  for (int i = 0; i < nip; i++)
    for (int e = 0; e < nprox[i]; e++) {
      const int x = prox[i * NPP + e], xs = proxs[i * NPP + e];
      if (xs >= nprox[x] || prox[x * NPP + xs] != i) { fimhip_ocn_destroy(o); return FIMHIP_EINVAL; }
    }
  // the packed per-cell word: edge count, own mask, mask of each edge's neighbour
  std::vector<int> cinfo(nip);
  for (int i = 0; i < nip; i++) {
    int w = nprox[i] | (o->wet[i] << 3);
    for (int e = 0; e < nprox[i]; e++)
      if (o->wet[prox[i * NPP + e]]) w |= 1 << (4 + e);
    cinfo[i] = w;
  }
  int r = 0;
  HycGrid &g = o->g;
  g.nip = nip;
  g.kdm = kdm;
  int *pp, *ps, *pc;
  double *pa, *pr, *pv, *psl, *prs, *prl, *pcs, *psn;
  const std::vector<double> csv(gd->cs, gd->cs + (size_t)nip * NPP * 4), snv(gd->sn, gd->sn + (size_t)nip * NPP * 4);
  const std::vector<double> svcv(gd->sidevec_c, gd->sidevec_c + (size_t)nip * NPP * 2);
  // the momentum equation's cell values, with the C library's sin and sqrt as
  // the reference forms them (dyn_init.F90:753, hycom_init.F90:327)
  std::vector<double> corv(nip), meshv(nip);
  for (int i = 0; i < nip; i++) {
    corv[i] = 1.4584e-4 * sin(gd->lat[i]);
    meshv[i] = sqrt(gd->area[i]);
  }
  double *pvc, *pco, *pms;
  if ((r = o->arena.upload(&pp, prox)) || (r = o->arena.upload(&ps, proxs)) || (r = o->arena.upload(&pc, cinfo)) ||
      (r = o->arena.upload(&pa, o->area)) || (r = o->arena.upload(&pr, rarea)) || (r = o->arena.upload(&pv, sve)) ||
      (r = o->arena.upload(&psl, sideln)) || (r = o->arena.upload(&prs, rsideln)) ||
      (r = o->arena.upload(&prl, rpl)) || (r = o->arena.upload(&pcs, csv)) || (r = o->arena.upload(&psn, snv)) ||
      (r = o->arena.upload(&pvc, svcv)) || (r = o->arena.upload(&pco, corv)) || (r = o->arena.upload(&pms, meshv))) {
    fimhip_ocn_destroy(o);
    return r;
  }
  g.prox = pp; g.proxs = ps; g.cinfo = pc;
  g.area = pa; g.rarea = pr; g.sve = pv; g.sideln = psl; g.rsideln = prs; g.rprox_ln = prl;
  g.cs = pcs; g.sn = psn;
  o->mo.svc = pvc; o->mo.corio = pco; o->mo.meshsz = pms;
  o->mo.veldff = 0.1;              // fimnamelist.F90:254
  o->mo.smagcf = 0.3;              // fimnamelist.F90:253
  const size_t cz = (size_t)nip * kdm, ez = cz * NPP;
  HycState &s = o->s;
  double *um, *vm, *un, *vn, *dpe;
  struct { double **p; size_t n; } arrs[] = {
      {&um, ez}, {&vm, ez}, {&un, ez}, {&vn, ez}, {&dpe, ez}, {&s.dp, 2 * cz}, {&s.pres, (size_t)nip * (kdm + 1)},
      {&s.massflx, ez}, {&s.mssfx_avg, ez}, {&s.antiflx, ez}, {&s.dp_lo, cz}, {&s.dpold, cz}, {&s.flxplus, cz},
      {&s.flxmnus, cz}, {&s.rplus, cz}, {&s.rmnus, cz}, {&s.del2p, cz}, {&s.dptmp, cz}, {&s.dpc, cz},
      {&s.clippd, (size_t)nip * NPP}, {&s.rpbot, (size_t)nip}, {&s.aux, (size_t)nip},
      // the edge interpolation's inputs and outputs
      {&o->ed.uclin, 2 * cz}, {&o->ed.vclin, 2 * cz}, {&o->ed.utrop, (size_t)nip * 3}, {&o->ed.vtrop, (size_t)nip * 3},
      {&o->ed.geop, (size_t)nip * (kdm + 1)}, {&o->ed.montg, cz}, {&o->ed.spcifv, cz},
      {&o->ed.pr_edg, (size_t)nip * NPP * (kdm + 1)}, {&o->ed.geop_edg, ez}, {&o->ed.mont_edg, ez},
      {&o->ed.uvsq_edg, ez}, {&o->ed.spcv_edg, ez},
      // the momentum equation's forcing and work arrays
      {&o->mo.taux, (size_t)nip}, {&o->mo.tauy, (size_t)nip}, {&o->mo.ustarb, (size_t)nip},
      {&o->mo.drag, (size_t)nip}, {&o->mo.dpbot, (size_t)nip}, {&o->mo.gradx, cz}, {&o->mo.grady, cz},
      {&o->mo.d4u, cz}, {&o->mo.d4v, cz}, {&o->mo.du, cz}, {&o->mo.dv, cz},
      // the tracers and density the transport and the grid generator share
      {&o->tr.temp, cz}, {&o->tr.saln, cz}, {&o->hy.dens, cz}};
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

double *ocn_field(fimhip_ocn *o, int f, size_t *n) {
  const size_t cz = (size_t)o->nip * o->kdm, ez = cz * NPP;
  HycState &s = o->s;
  switch (f) {
    case FIMHIP_OCN_DP: *n = 2 * cz; return s.dp;
    case FIMHIP_OCN_PRES: *n = (size_t)o->nip * (o->kdm + 1); return s.pres;
    case FIMHIP_OCN_UM_EDG: *n = ez; return s.um;
    case FIMHIP_OCN_VM_EDG: *n = ez; return s.vm;
    case FIMHIP_OCN_UN_EDG: *n = ez; return s.un;
    case FIMHIP_OCN_VN_EDG: *n = ez; return s.vn;
    case FIMHIP_OCN_DPM_EDG: *n = ez; return s.dpm_edg;
    case FIMHIP_OCN_MASSFLX: *n = ez; return s.massflx;
    case FIMHIP_OCN_MSSFX_AVG: *n = ez; return s.mssfx_avg;
    case FIMHIP_OCN_UCLIN: *n = 2 * cz; return o->ed.uclin;
    case FIMHIP_OCN_VCLIN: *n = 2 * cz; return o->ed.vclin;
    case FIMHIP_OCN_UTROP: *n = (size_t)o->nip * 3; return o->ed.utrop;
    case FIMHIP_OCN_VTROP: *n = (size_t)o->nip * 3; return o->ed.vtrop;
    case FIMHIP_OCN_GEOP: *n = (size_t)o->nip * (o->kdm + 1); return o->ed.geop;
    case FIMHIP_OCN_MONTG: *n = cz; return o->ed.montg;
    case FIMHIP_OCN_SPCIFV: *n = cz; return o->ed.spcifv;
    case FIMHIP_OCN_PR_EDG: *n = (size_t)o->nip * NPP * (o->kdm + 1); return o->ed.pr_edg;
    case FIMHIP_OCN_GEOP_EDG: *n = ez; return o->ed.geop_edg;
    case FIMHIP_OCN_MONT_EDG: *n = ez; return o->ed.mont_edg;
    case FIMHIP_OCN_UVSQ_EDG: *n = ez; return o->ed.uvsq_edg;
    case FIMHIP_OCN_SPCV_EDG: *n = ez; return o->ed.spcv_edg;
    case FIMHIP_OCN_TAUX: *n = (size_t)o->nip; return o->mo.taux;
    case FIMHIP_OCN_TAUY: *n = (size_t)o->nip; return o->mo.tauy;
    case FIMHIP_OCN_USTARB: *n = (size_t)o->nip; return o->mo.ustarb;
    case FIMHIP_OCN_CUMUFLX: *n = ez; return o->tr.cumuflx;
    case FIMHIP_OCN_DPINIT: *n = cz; return o->tr.dpinit;
    case FIMHIP_OCN_DPFINL: *n = cz; return o->tr.dpfinl;
    case FIMHIP_OCN_TEMP: *n = cz; return o->tr.temp;
    case FIMHIP_OCN_SALN: *n = cz; return o->tr.saln;
    case FIMHIP_OCN_PASSV: *n = o->tr.numtr * cz; return o->tr.numtr > 0 ? o->tr.passv : nullptr;
    case FIMHIP_OCN_DENS: *n = cz; return o->hy.dens;
  }
  return nullptr;
}

// waits for every asynchronous call queued on the instance
extern "C" int fimhip_ocn_sync(fimhip_ocn *o) {
  if (!o) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
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
  const dim3 grid(ocn_blocks(o, LPC)), blk(256);     // a multiple of 8 blocks: see hcol
  const double dt = o->dtleap;
  const int smooth = o->thkdff != 0.;
  KT_LAUNCH("hyc_flux", st, hyc_flux<LPC><<<grid, blk, 0, st>>>(g, s, leapm, leapn, dt));
  KT_LAUNCH("hyc_lowsum", st, hyc_lowsum<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt));
  KT_LAUNCH("hyc_limits", st, hyc_limits<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt));
  KT_LAUNCH("hyc_clip", st, hyc_clip<LPC><<<grid, blk, 0, st>>>(g, s, dt));
  KT_LAUNCH("hyc_recover", st, hyc_recover<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt, smooth, o->biharm));
  if (smooth) {
    if (o->biharm > 0.) KT_LAUNCH("hyc_del2p", st, hyc_del2p<LPC><<<grid, blk, 0, st>>>(g, s, o->biharm));
    KT_LAUNCH("hyc_bolus", st, hyc_bolus<LPC><<<grid, blk, 0, st>>>(g, s, leapn, dt, o->thkdff, o->biharm));
    KT_LAUNCH("hyc_bclip", st, hyc_bclip<LPC><<<grid, blk, 0, st>>>(g, s, leapn, o->biharm, o->masscor));
  }
  KT_LAUNCH("hyc_final", st, hyc_final<LPC><<<grid, blk, 0, st>>>(g, s, leapm, leapn, dt, smooth, o->biharm));
  KT_LAUNCH("hyc_check", st, hyc_check<LPC><<<grid, blk, 0, st>>>(g, s));
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

extern "C" int fimhip_ocn_set_dtleap(fimhip_ocn *o, double dtleap) {
  if (!o || !(dtleap > 0.)) return FIMHIP_EINVAL;
  o->dtleap = dtleap;
  return FIMHIP_OK;
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
