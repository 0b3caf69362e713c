// HYCOM's momentum equation (FIMsrc/fim/horizontal/hycom_momtum.F90, momtum,
// with hycom_dissip.F90's del4prep and dissip) on the GPU, in FP64, with the
// reference's operations in its order.  Four kernels on the column layout of
// the other ocean kernels (lanes over layers):
//   hyc_mom_pgf   interface pressure of the new level, the near-bottom
//                 velocity, bottom drag and friction velocity; the pressure
//                 force around the cell's edges (Bernoulli function and p grad
//                 alpha on the mid-time interfaces, or the janjic_ocn
//                 geopotential)  (104-183); the bottom of dissip's own
//                 pressure column (hycom_dissip.F90:150-159)
//   hyc_del4      del4prep: minus the mean of the neighbours' velocities
//                 rotated into the cell's frame, walls by the slip condition
//                 (hycom_dissip.F90:9-110)
//   hyc_dissip    dissip: thickness-weighted momentum fluxes through the edges
//                 with the sidewall blend, summed into a tendency
//                 (hycom_dissip.F90:113-272)
//   hyc_mom_upd   vorticity, deformation and Smagorinsky viscosity from the
//                 edge velocities; Coriolis force, wind stress and bottom drag;
//                 the leapfrog update; the downward mixing into thin bottom
//                 layers (a recurrence down the column, on one lane per
//                 velocity component over LDS); the time smoothing (211-389)
// The sums along a column (the interface pressures, the near-bottom
// velocity) are sequential chains through LDS as in the continuity kernels.
// modesplit is .false. in the reference (hycom_control.F90:5): the
// barotropic forcing (ubforc, vbforc) is not formed.
#include "hycom_internal.h"

namespace fimhip {
namespace hyc {
namespace {

constexpr double H_GRV = 9.80665;                 // module_constants grvity
constexpr double H_ONEMM = 0.001 * H_ONEM;
constexpr double H_TENCM = 0.1 * H_ONEM;
constexpr double H_DRCOEF = 3.e-3;                // hycom_constants.F90:44
constexpr double H_ASSLUV = .25;                  // hycom_constants.F90:48
constexpr double H_THSHLD = 1.e-11;               // hycom_dissip.F90:145
constexpr double H_THKTOP = 50. * H_ONEM, H_THKBOT = 10. * H_ONEM, H_CBAR = .5;   // hycom_momtum.F90:87-91

#define HSVC(i, e, c) m.svc[((size_t)(i) * NPP + (e)) * 2 + (c)]
#define HCS(i, e, j) g.cs[((size_t)(i) * NPP + (e)) * 4 + (j)]
#define HSN(i, e, j) g.sn[((size_t)(i) * NPP + (e)) * 4 + (j)]

// the sum over the column's layers of v, in layer order from 0, on every lane
__device__ __forceinline__ double col_total(double *buf, double v, int k, int kdm) {
  buf[k] = v;
  wave_sync();
  double acc = 0.;
  for (int j = 0; j < kdm; j++) acc = acc + buf[j];
  wave_sync();
  return acc;
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_mom_pgf(HycGrid g, HycState s, HycEdg ed, HycMom m, int leapm, int leapn,
                                                   int janjic) {
  __shared__ double sh[256 / LPC][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const size_t cz = (size_t)nip * kdm;
  const double *dpm = s.dp + (size_t)(leapm - 1) * cz, *dpn = s.dp + (size_t)(leapn - 1) * cz;
  const double dn = live ? HC(dpn, i, k) : 0.;
  const double p0 = HP(i, 0);
  // the new level's interfaces: pn1 below the layer, pn0 above, pb the bottom
  // (every shuffle is taken by all the column's lanes: a lane that skips one
  // gives no value to the lane that reads it)
  const double pn1 = col_pres(sh[c.cc], dn, p0, k, kdm);
  const double pn1_up = __shfl_up(pn1, 1, LPC);
  const double pn0 = k > 0 ? pn1_up : p0;
  const double pb = __shfl(pn1, kdm - 1, LPC);
  // the near-bottom velocity (thkbot above the bottom) and the drag
  const double utn = HC(ed.uclin + (size_t)(leapn - 1) * cz, i, kk) + ed.utrop[i * 3 + leapn - 1];
  const double vtn = HC(ed.vclin + (size_t)(leapn - 1) * cz, i, kk) + ed.vtrop[i * 3 + leapn - 1];
  const double q = hmin(H_THKBOT, pb);
  const double wgt = hmin(pn1, pb - q) - hmin(pn0, pb - q);
  const double ubot = col_total(sh[c.cc], live ? utn * wgt : 0., k, kdm);
  const double vbot = col_total(sh[c.cc], live ? vtn * wgt : 0., k, kdm);
  const double botvel = sqrt(ubot * ubot + vbot * vbot) / H_THKBOT;
  if (k == 0) {
    m.drag[i] = hmin(H_DRCOEF * (botvel + H_CBAR) * H_ONEM / H_THKBOT, 1. / m.batrop);
    m.ustarb[i] = sqrt(H_DRCOEF) * botvel;
  }
  // dissip's pressure column starts at 0 (hycom_dissip.F90:150-159)
  const double pz = col_pres(sh[c.cc], dn, 0., k, kdm);
  if (k == kdm - 1) m.dpbot[i] = pz;
  // the pressure force around the edges
  double pgfx = 0., pgfy = 0., pk1 = pn1;
  if (janjic) {
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      if (e >= np) break;
      const double t = HE(ed.geop_edg, i, e, kk) + HE(ed.uvsq_edg, i, e, kk);
      pgfx = pgfx - t * HSVC(i, e, 1);
      pgfy = pgfy + t * HSVC(i, e, 0);
    }
  } else {
    // the mid-time interfaces replace the new level's (recomputed at every
    // edge in the reference, to the same values)
    pk1 = col_pres(sh[c.cc], live ? HC(dpm, i, k) : 0., p0, k, kdm);
    const double pk1_up = __shfl_up(pk1, 1, LPC);
    const double pk0 = k > 0 ? pk1_up : p0;
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      if (e >= np) break;
      const double t = HE(ed.mont_edg, i, e, kk) + HE(ed.uvsq_edg, i, e, kk);
      const double sv = HE(ed.spcv_edg, i, e, kk);
      pgfx = pgfx - t * HSVC(i, e, 1);
      pgfy = pgfy + t * HSVC(i, e, 0);
      pgfx = pgfx + .5 * (pk0 + pk1) * sv * HSVC(i, e, 1);
      pgfy = pgfy - .5 * (pk0 + pk1) * sv * HSVC(i, e, 0);
    }
  }
  if (!live) return;
  __builtin_nontemporal_store(pk1, &HP(i, k + 1));
  __builtin_nontemporal_store(pgfx * g.rarea[i], &HC(m.gradx, i, k));
  __builtin_nontemporal_store(pgfy * g.rarea[i], &HC(m.grady, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_del4(HycGrid g, HycEdg ed, HycMom m, int leapn) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const size_t cz = (size_t)nip * kdm;
  const double *un = ed.uclin + (size_t)(leapn - 1) * cz, *vn = ed.vclin + (size_t)(leapn - 1) * cz;
  const double ui = HC(un, i, k), vi = HC(vn, i, k);
  // every slot's operands are loaded before any is used (a pentagon's sixth
  // slot and a land neighbour are valid cells), so the loads overlap
  double ux[NPP], vx[NPP], c0[NPP], s0[NPP], c1[NPP], s1[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    ux[e] = HC(un, nb[e], k);
    vx[e] = HC(vn, nb[e], k);
    c0[e] = HCS(i, e, 0);
    s0[e] = HSN(i, e, 0);
    c1[e] = HCS(i, e, 1);
    s1[e] = HSN(i, e, 1);
  }
  double avgu = 0., avgv = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    double ungbor, vngbor;
    if (ci_open(ci, e)) {
      const double uedge = c1[e] * ux[e] + s1[e] * vx[e];
      const double vedge = -s1[e] * ux[e] + c1[e] * vx[e];
      ungbor = c0[e] * uedge - s0[e] * vedge;
      vngbor = s0[e] * uedge + c0[e] * vedge;
    } else {
      ungbor = H_SLIP * ui;
      vngbor = H_SLIP * vi;
    }
    if (e < np) {
      avgu = avgu + ungbor;
      avgv = avgv + vngbor;
    }
  }
  const double factor = 1. / np;
  __builtin_nontemporal_store(ui - avgu * factor, &HC(m.d4u, i, k));
  __builtin_nontemporal_store(vi - avgv * factor, &HC(m.d4v, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_dissip(HycGrid g, HycState s, HycMom m, int leapn) {
  __shared__ double sh[256 / LPC][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const double *dpn = s.dp + (size_t)(leapn - 1) * (size_t)nip * kdm;
  const double di = HC(dpn, i, kk);
  const double pz1 = col_pres(sh[c.cc], live ? di : 0., 0., k, kdm);     // dissip's own column
  if (!live) return;
  const double ui = HC(m.d4u, i, k), vi = HC(m.d4v, i, k);
  const double factor = g.rarea[i] / hmax(di, H_THSHLD);
  double du = 0., dv = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    const int ix = nb[e];
    const double c0 = HCS(i, e, 0), s0 = HSN(i, e, 0);
    const double urot = c0 * ui + s0 * vi;
    const double vrot = -s0 * ui + c0 * vi;
    double ungbor, vngbor;
    if (ci_open(ci, e)) {
      const double ux = HC(m.d4u, ix, k), vx = HC(m.d4v, ix, k);
      ungbor = HCS(i, e, 1) * ux + HSN(i, e, 1) * vx;
      vngbor = -HSN(i, e, 1) * ux + HCS(i, e, 1) * vx;
      const double blokd = hmax(0., hmin(1., (pz1 - m.dpbot[ix]) / hmax(H_EPSIL, di)));
      ungbor = (1. - blokd) * ungbor + blokd * urot * H_SLIP;
      vngbor = (1. - blokd) * vngbor + blokd * vrot * H_SLIP;
    } else {
      ungbor = urot * H_SLIP;
      vngbor = vrot * H_SLIP;
    }
    const double a = hmax(di, H_THSHLD), b = hmax(HC(dpn, ix, k), H_THSHLD);
    const double hh = a * b / (a + b);
    const double sl = g.sideln[i * NPP + e];
    const double ufx = (ungbor - urot) * sl * 2. * hh;
    const double vfx = (vngbor - vrot) * sl * 2. * hh;
    const double uflux = c0 * ufx - s0 * vfx;
    const double vflux = s0 * ufx + c0 * vfx;
    du = du + uflux * factor;
    dv = dv + vflux * factor;
  }
  __builtin_nontemporal_store(du, &HC(m.du, i, k));
  __builtin_nontemporal_store(dv, &HC(m.dv, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_mom_upd(HycGrid g, HycState s, HycEdg ed, HycMom m, int leapm,
                                                   int leapn, double dtleap) {
  constexpr int CPB = 256 / LPC;
  __shared__ double sx[CPB][2][LPC + 1], sq[CPB][LPC + 1], ssl[CPB][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k, cc = c.cc;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const size_t cz = (size_t)nip * kdm;
  const double *dpn = s.dp + (size_t)(leapn - 1) * cz;
  double *ucm = ed.uclin + (size_t)(leapm - 1) * cz, *ucn = ed.uclin + (size_t)(leapn - 1) * cz;
  double *vcm = ed.vclin + (size_t)(leapm - 1) * cz, *vcn = ed.vclin + (size_t)(leapn - 1) * cz;
  // velocity derivatives from the tangential velocity along the edges
  double dudx = 0., dudy = 0., dvdx = 0., dvdy = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    const double ue = HE(s.um, i, e, kk), ve = HE(s.vm, i, e, kk);
    const double s0 = HSVE(i, e, 0), s1 = HSVE(i, e, 1);
    dudx = dudx + s1 * ue;
    dudy = dudy - s0 * ue;
    dvdx = dvdx + s1 * ve;
    dvdy = dvdy - s0 * ve;
  }
  const double ra = g.rarea[i];
  dudx = dudx * ra;
  dudy = dudy * ra;
  dvdx = dvdx * ra;
  dvdy = dvdy * ra;
  const double relvor = dvdx - dudy;
  const double da = dudx - dvdy, db = dudy + dvdx;
  const double deform = sqrt(da * da + db * db);
  const double smag = deform * m.meshsz[i] * m.smagcf;
  const double udissp = HC(m.du, i, kk) * hmax(m.veldff, smag);
  const double vdissp = HC(m.dv, i, kk) * hmax(m.veldff, smag);
  const double cor = m.corio[i];
  const double uzeta = HC(ucm, i, kk) * (relvor + cor) + ed.utrop[i * 3 + leapm - 1] * relvor;
  const double vzeta = HC(vcm, i, kk) * (relvor + cor) + ed.vtrop[i * 3 + leapm - 1] * relvor;
  const double pk = HP(i, kk), pk1 = HP(i, kk + 1), pb = HP(i, kdm);
  const double dn = HC(dpn, i, kk);
  double ustres = m.taux[i] * H_GRV / H_THKTOP * (hmin(pk1 + H_ONEMM, H_THKTOP) - hmin(pk, H_THKTOP)) /
                  hmax(dn, H_ONEMM);
  double vstres = m.tauy[i] * H_GRV / H_THKTOP * (hmin(pk1 + H_ONEMM, H_THKTOP) - hmin(pk, H_THKTOP)) /
                  hmax(dn, H_ONEMM);
  const double q = (hmax(pb - H_THKBOT, pk1) - hmax(pb - H_THKBOT, hmin(pk1 - H_ONEMM, pk))) / hmax(dn, H_ONEMM);
  const double uold = HC(ucn, i, kk), vold = HC(vcn, i, kk);
  const double utn = uold + ed.utrop[i * 3 + leapn - 1], vtn = vold + ed.vtrop[i * 3 + leapn - 1];
  ustres = ustres - m.drag[i] * utn * q;
  vstres = vstres - m.drag[i] * vtn * q;
  const double un = uold + dtleap * (HC(m.gradx, i, kk) + vzeta + udissp + ustres);
  const double vn = vold + dtleap * (HC(m.grady, i, kk) - uzeta + vdissp + vstres);
  // downward momentum mixing into thin bottom layers: each layer below the
  // first takes (slab - q) of the final value of the layer above
  const double slab = hmax(H_TENCM, pk1 - pb + H_TENM);
  const double qq = hmin(dn, slab);
  sx[cc][0][k] = un;
  sx[cc][1][k] = vn;
  sq[cc][k] = qq;
  ssl[cc][k] = slab;
  wave_sync();
  if (k < 2) {
    double *x = sx[cc][k];
    double prev = x[0];
    for (int j = 1; j < kdm; j++) {
      const double sj = ssl[cc][j], qj = sq[cc][j];
      prev = (x[j] * qj + prev * (sj - qj)) / sj;
      x[j] = prev;
    }
  }
  wave_sync();
  if (!live) return;
  const double uf = sx[cc][0][k], vf = sx[cc][1][k];
  __builtin_nontemporal_store(uf, &HC(ucn, i, k));
  __builtin_nontemporal_store(vf, &HC(vcn, i, k));
  __builtin_nontemporal_store(H_ASSLUV * (uold + uf) + (1. - H_ASSLUV - H_ASSLUV) * HC(ucm, i, k), &HC(ucm, i, k));
  __builtin_nontemporal_store(H_ASSLUV * (vold + vf) + (1. - H_ASSLUV - H_ASSLUV) * HC(vcm, i, k), &HC(vcm, i, k));
}

}  // namespace
}  // namespace hyc
}  // namespace fimhip

using namespace fimhip;
using namespace fimhip::hyc;

template <int LPC>
static void mom_kernels(fimhip_ocn *o, int leapm, int leapn) {
  const HycGrid &g = o->g;
  const HycState &s = o->s;
  const HycEdg &ed = o->ed;
  const HycMom &m = o->mo;
  hipStream_t st = o->stream;
  const dim3 grid(ocn_blocks(o, LPC)), blk(256);
  KT_LAUNCH("hyc_mom_pgf", st, hyc_mom_pgf<LPC><<<grid, blk, 0, st>>>(g, s, ed, m, leapm, leapn, o->janjic ? 1 : 0));
  KT_LAUNCH("hyc_del4", st, hyc_del4<LPC><<<grid, blk, 0, st>>>(g, ed, m, leapn));
  KT_LAUNCH("hyc_dissip", st, hyc_dissip<LPC><<<grid, blk, 0, st>>>(g, s, m, leapn));
  KT_LAUNCH("hyc_mom_upd", st, hyc_mom_upd<LPC><<<grid, blk, 0, st>>>(g, s, ed, m, leapm, leapn, o->dtleap));
}

// momtum on the device-resident arrays (hycom_momtum.F90:16-389)
static int mom_launch(fimhip_ocn *o, int leapm, int leapn) {
  if (leapm < 1 || leapm > 2 || leapn < 1 || leapn > 2 || leapm == leapn) return FIMHIP_EINVAL;
  if (!(o->mo.batrop > 0.)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  g_ktimer = o->kt;
  if (o->kdm <= 32) mom_kernels<32>(o, leapm, leapn);
  else mom_kernels<64>(o, leapm, leapn);
  g_ktimer = nullptr;
  HIPCHK(hipGetLastError());
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_set_momtum(fimhip_ocn *o, double batrop, double veldff, double smagcf) {
  if (!o || !(batrop > 0.)) return FIMHIP_EINVAL;
  o->mo.batrop = batrop;
  o->mo.veldff = veldff;
  o->mo.smagcf = smagcf;
  return FIMHIP_OK;
}

// asynchronous on the instance's stream, like fimhip_ocn_edgvar
extern "C" int fimhip_ocn_momtum(fimhip_ocn *o, int leapm, int leapn) {
  if (!o) return FIMHIP_EINVAL;
  return mom_launch(o, leapm, leapn);
}

extern "C" int fimhip_ocn_momtum_r8(fimhip_ocn *o, int nstep, int leapm, int leapn, const double *utrop,
                                   const double *vtrop, double *uclin, double *vclin, const double *um_edg,
                                   const double *vm_edg, const double *geop_edg, const double *mont_edg,
                                   const double *uvsq_edg, const double *spcv_edg, const double *taux,
                                   const double *tauy, double *ubforc, double *vbforc, const double *dp,
                                   double *pres, double *ustarb) {
  (void)nstep;
  (void)ubforc;     // formed only under modesplit, which the reference has off
  (void)vbforc;
  if (!o || !utrop || !vtrop || !uclin || !vclin || !um_edg || !vm_edg || !geop_edg || !mont_edg || !uvsq_edg ||
      !spcv_edg || !taux || !tauy || !dp || !pres || !ustarb)
    return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  const double *in[] = {utrop, vtrop, uclin, vclin, um_edg, vm_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg,
                        taux, tauy, dp, pres, ustarb};
  const int fi[] = {FIMHIP_OCN_UTROP, FIMHIP_OCN_VTROP, FIMHIP_OCN_UCLIN, FIMHIP_OCN_VCLIN, FIMHIP_OCN_UM_EDG,
                    FIMHIP_OCN_VM_EDG, FIMHIP_OCN_GEOP_EDG, FIMHIP_OCN_MONT_EDG, FIMHIP_OCN_UVSQ_EDG,
                    FIMHIP_OCN_SPCV_EDG, FIMHIP_OCN_TAUX, FIMHIP_OCN_TAUY, FIMHIP_OCN_DP, FIMHIP_OCN_PRES,
                    FIMHIP_OCN_USTARB};
  for (int j = 0; j < 15; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fi[j], &n);
    HIPCHK(hipMemcpyAsync(d, in[j], n * sizeof(double), hipMemcpyHostToDevice, o->stream));
  }
  const int r = mom_launch(o, leapm, leapn);
  if (r) return r;
  double *outp[] = {uclin, vclin, pres, ustarb};
  const int fo[] = {FIMHIP_OCN_UCLIN, FIMHIP_OCN_VCLIN, FIMHIP_OCN_PRES, FIMHIP_OCN_USTARB};
  for (int j = 0; j < 4; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fo[j], &n);
    HIPCHK(hipMemcpyAsync(outp[j], d, n * sizeof(double), hipMemcpyDeviceToHost, o->stream));
  }
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}
