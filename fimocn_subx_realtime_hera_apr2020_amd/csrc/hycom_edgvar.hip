// HYCOM's edge interpolation (FIMsrc/fim/horizontal/hycom_edgvar.F90,
// bclin_edgvar with uvedge and blokf) on the GPU, in FP64, with the
// reference's operations in its order.  Three kernels, lanes over layers as
// in the continuity kernels:
//   hyc_janjic   (janjic_ocn only) the layer-mean geopotential integrated
//                hydrostatically to each layer's mid pressure in the
//                neighbouring columns, nudged below the silhouette depth, and
//                geop_edg from the six results  (200-265, 437-462).  Each
//                lane searches its own level in the neighbour's column,
//                staged in LDS.
//   hyc_uvedge   total velocities, rotated into each edge's frame from the
//                cell, its neighbour and the two joint neighbours, one launch
//                per leapfrog level (6-124, 273-296); with the first, the
//                interface pressure of the mid-time thickness (the column sum
//                through LDS) and the kinetic energy on the edges (411-412).
//   hyc_edgint   interface pressure and thickness on the edges: each edge's
//                two sequential chains (down from the shallower bottom, then
//                up, limited by the interpolated thickness) run on one lane
//                over increments the column's lanes put in LDS; Montgomery
//                potential and specific volume with the sidewall blend on the
//                edges  (300-418).
// A neighbour's column values are loaded once per lane and shared by the
// three edges that use it (as the edge's cell or as a joint neighbour).
#include "hycom_internal.h"

namespace fimhip {
namespace hyc {
namespace {

#define HG(i, k) ed.geop[(size_t)(i) * (kdm + 1) + (k)]
#define HPR(i, e, k) ed.pr_edg[((size_t)(i) * NPP + (e)) * (kdm + 1) + (k)]

// the edges before and after edge e of a cell with np edges
__device__ __forceinline__ int e_prev(int e, int np) { return e == 0 ? np - 1 : e - 1; }
__device__ __forceinline__ int e_next(int e, int np) { return e == np - 1 ? 0 : e + 1; }
// a register array's entry at edge e's previous and next edge: with e a
// constant of an unrolled loop, every index is a constant and only the
// wrap-around selects on np (a run-time index would put the array in scratch)
__device__ __forceinline__ double at_prev(const double (&a)[NPP], int e, int np) {
  return e == 0 ? (np == 6 ? a[5] : a[4]) : a[e == 0 ? 0 : e - 1];
}
__device__ __forceinline__ double at_next(const double (&a)[NPP], int e, int np) {
  return e < 4 ? a[e < 4 ? e + 1 : 0] : (e == 4 ? (np == 5 ? a[0] : a[5]) : a[0]);
}
// the extent to which a sidewall is present (hycom_edgvar.F90:479-493; its
// fourth argument is not used)
__device__ __forceinline__ double blokf(double dp, double plo, double pbx) {
  return hmax(0., hmin(1., (plo - pbx) / hmax(H_EPSIL, dp)));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_janjic(HycGrid g, HycState s, HycEdg ed) {
  constexpr int CPB = 256 / LPC;
  __shared__ double sp[CPB][LPC + 1], sg[CPB][LPC + 1], sv[CPB][LPC];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k, cc = c.cc;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const double p_lyr = .5 * (HP(i, kk) + HP(i, kk + 1));
  const double glyr = .5 * (HG(i, kk) + HG(i, kk + 1));
  double silhou = 1.e33;        // the shallowest ocean neighbour's bottom
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const double pb = HP(nb[e], kdm);
    if (ci_open(ci, e)) silhou = hmin(silhou, pb);
  }
  const double plyrk = hmin(p_lyr, silhou);
  double gx[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    gx[e] = glyr;
    if (!ci_open(ci, e)) continue;
    const int ipx = nb[e];
    wave_sync();
    for (int j = k; j <= kdm; j += LPC) {
      sp[cc][j] = HP(ipx, j);
      sg[cc][j] = HG(ipx, j);
    }
    if (live) sv[cc][k] = HC(ed.spcifv, ipx, k);
    wave_sync();
    if (live) {
      int kx = 0;
      while (kx < kdm - 1 && sp[cc][kx + 1] < plyrk) kx++;
      if (sp[cc][kx + 1] >= plyrk) gx[e] = sg[cc][kx] - sv[cc][kx] * (plyrk - sp[cc][kx]);
    }
  }
  if (!live) return;
  const double tapr = hmax(0., hmin(1., (p_lyr - silhou) / (100. * H_ONEM)));
#pragma unroll
  for (int e = 0; e < NPP; e++)
    if (e < np) gx[e] = (1. - tapr) * gx[e] + tapr * glyr;
#pragma unroll
  for (int e = 0; e < NPP; e++)
    if (e < np)
      HE(ed.geop_edg, i, e, k) =
          (8. * (glyr + gx[e]) + at_prev(gx, e, np) + at_next(gx, e, np)) * H_DIVBY18;
}

// one velocity pair rotated and interpolated to edge e (uvedge's layer loop)
struct UV {
  double u, v;
};
struct Rot {
  double cs[4], sn[4];
};
__device__ __forceinline__ UV uv_edge(const Rot &rt, bool wi, bool wpx, bool wm1, bool wp1,
                                      double ui, double vi, double ux, double vx, double um1, double vm1, double up1,
                                      double vp1) {
  const double *cs = rt.cs, *sn = rt.sn;
  double u1 = cs[0] * ui + sn[0] * vi;
  double v1 = -sn[0] * ui + cs[0] * vi;
  double u2, v2;
  if (wpx) {
    u2 = cs[1] * ux + sn[1] * vx;
    v2 = -sn[1] * ux + cs[1] * vx;
    if (!wi) {          // the neighbour is ocean and the cell is not
      u1 = u2 * H_SLIP;
      v1 = v2 * H_SLIP;
    }
  } else {
    u2 = u1 * H_SLIP;
    v2 = v1 * H_SLIP;
  }
  const double u3 = wm1 ? cs[2] * um1 + sn[2] * vm1 : 0.;
  const double v3 = wm1 ? -sn[2] * um1 + cs[2] * vm1 : 0.;
  const double u4 = wp1 ? cs[3] * up1 + sn[3] * vp1 : 0.;
  const double v4 = wp1 ? -sn[3] * up1 + cs[3] * vp1 : 0.;
  return UV{(8. * (u1 + u2) + u3 + u4) * H_DIVBY18, (8. * (v1 + v2) + v3 + v4) * H_DIVBY18};
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_uvedge(HycGrid g, HycState s, HycEdg ed, int leapm, int slot,
                                                  double *ue, double *ve, int with_pres, double *uvsq) {
  __shared__ double sh[256 / LPC][LPC + 1];
  __shared__ double rot[256 / LPC][2 * NPP * 4];     // the cell's cs then sn, (4,6) each
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
  const int np = ci_np(ci);
  const bool wi = ci_wet(ci);
  const size_t cz = (size_t)nip * kdm;
  const double *dpm = s.dp + (size_t)(leapm - 1) * cz;
  const double *uc = ed.uclin + (size_t)(slot - 1) * cz, *vc = ed.vclin + (size_t)(slot - 1) * cz;
  // every operand is loaded before the column sum and the LDS staging, so
  // that the loads' latencies overlap each other and the sum: the cell's
  // rotation coefficients (one or two per lane), and the total velocities
  // (baroclinic + barotropic; land: -1e33) of the cell and its neighbours at
  // leapfrog level 'slot' (a land cell's and a pentagon's sixth slot's
  // values are valid memory, replaced or unused)
  constexpr int NROT = 2 * NPP * 4, NRL = (NROT + LPC - 1) / LPC;
  double rv[NRL];
#pragma unroll
  for (int q = 0; q < NRL; q++) {
    const int j = k + q * LPC;
    rv[q] = j < NPP * 4 ? g.cs[(size_t)i * NPP * 4 + j] : j < NROT ? g.sn[(size_t)i * NPP * 4 + j - NPP * 4] : 0.;
  }
  const double uo = HC(uc, i, kk) + ed.utrop[i * 3 + slot - 1];
  const double vo = HC(vc, i, kk) + ed.vtrop[i * 3 + slot - 1];
  double ux[NPP], vx[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int x = nb[e];
    const bool wx = ci_open(ci, e);
    const double a = HC(uc, x, kk) + ed.utrop[x * 3 + slot - 1], b = HC(vc, x, kk) + ed.vtrop[x * 3 + slot - 1];
    ux[e] = wx ? a : H_LAND;
    vx[e] = wx ? b : H_LAND;
  }
  // the interface pressure of the mid-time thickness (273-291)
  if (with_pres && wi) {
    const double p = col_pres(sh[c.cc], live ? HC(dpm, i, k) : 0., HP(i, 0), k, kdm);
    if (live) HP(i, k + 1) = p;
  }
  // the column's lanes stage the cell's rotation coefficients in LDS
#pragma unroll
  for (int q = 0; q < NRL; q++)
    if (k + q * LPC < NROT) rot[c.cc][k + q * LPC] = rv[q];
  wave_sync();
  {
    const double ui = wi ? uo : H_LAND;
    const double vi = wi ? vo : H_LAND;
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      UV r{0., 0.};
      if (e < np) {
        const int em1 = e_prev(e, np), ep1 = e_next(e, np);
        Rot rt;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          rt.cs[j] = rot[c.cc][e * 4 + j];
          rt.sn[j] = rot[c.cc][NPP * 4 + e * 4 + j];
        }
        r = uv_edge(rt, wi, ci_open(ci, e), ci_open(ci, em1), ci_open(ci, ep1), ui, vi, ux[e], vx[e],
                    at_prev(ux, e, np), at_prev(vx, e, np), at_next(ux, e, np), at_next(vx, e, np));
      }
      if (live) {
        __builtin_nontemporal_store(r.u, &HE(ue, i, e, k));
        __builtin_nontemporal_store(r.v, &HE(ve, i, e, k));
        // the kinetic energy on the edges, from the values just stored (0 on
        // land and in a pentagon's sixth slot)
        if (uvsq) __builtin_nontemporal_store(wi && e < np ? .5 * (r.u * r.u + r.v * r.v) : 0., &HE(uvsq, i, e, k));
      }
    }
  }
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_edgint(HycGrid g, HycState s, HycEdg ed, int leapm) {
  constexpr int CPB = 256 / LPC;
  __shared__ double sh_h[CPB][NPP][LPC + 1], sh_t[CPB][NPP][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k, cc = c.cc;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) {
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      if (live) {
        __builtin_nontemporal_store(0., &HE(s.dpm_edg, i, e, k));
        __builtin_nontemporal_store(0., &HE(ed.mont_edg, i, e, k));
        __builtin_nontemporal_store(0., &HE(ed.spcv_edg, i, e, k));
      }
      for (int j = k; j <= kdm; j += LPC) __builtin_nontemporal_store(0., &HPR(i, e, j));
    }
    return;
  }
  const int np = ci_np(ci);
  const double *dpm = s.dp + (size_t)(leapm - 1) * (size_t)nip * kdm;
  const double di = HC(dpm, i, kk), pik1 = HP(i, kk + 1), pbi = HP(i, kdm);
  const double mi = HC(ed.montg, i, kk), svi = HC(ed.spcifv, i, kk);
  double D[NPP], PB[NPP], M[NPP], SV[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int x = nb[e];
    D[e] = HC(dpm, x, kk);
    PB[e] = HP(x, kdm);
    M[e] = HC(ed.montg, x, kk);
    SV[e] = HC(ed.spcifv, x, kk);
  }
  // the increments of each edge's two interface chains; a land neighbour is
  // replaced by the cell itself (300-330)
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    double h = 0., t = 0.;
    if (e < np) {
      const int em1 = e_prev(e, np), ep1 = e_next(e, np);
      const double dx = ci_open(ci, e) ? D[e] : di;
      const double dm1 = ci_open(ci, em1) ? at_prev(D, e, np) : di;
      const double dp1 = ci_open(ci, ep1) ? at_next(D, e, np) : di;
      h = .5 * hmin(hmin(H_TENM, di), dx);
      t = (8. * (di + dx) + dm1 + dp1) * H_DIVBY18;
    }
    sh_h[cc][e][k] = h;
    sh_t[cc][e][k] = t;
  }
  wave_sync();
  if (k < np) {
    // lane e: edge e's interface pressures, down from the shallower of the two
    // bottoms (0 beside land), then up, each interface no deeper than the one
    // above plus the interpolated thickness; they replace the increments
    double *hrow = sh_h[cc][k];
    const double *trow = sh_t[cc][k];
    double pr = ci_open(ci, k) ? hmin(pbi, HP(g.prox[i * NPP + k], kdm)) : 0.;
    hrow[kdm] = pr;
    for (int j = kdm - 1; j >= 1; j--) {
      pr = pr - hrow[j];
      hrow[j] = pr;
    }
    hrow[0] = 0.;
    double prk = 0.;
    for (int j = 0; j < kdm - 1; j++) {
      const double nx = hmin(hrow[j + 1], prk + trow[j]);
      hrow[j + 1] = nx;
      prk = nx;
    }
  }
  wave_sync();
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) {
      if (live) {
        __builtin_nontemporal_store(0., &HE(s.dpm_edg, i, e, k));
        __builtin_nontemporal_store(0., &HE(ed.mont_edg, i, e, k));
        __builtin_nontemporal_store(0., &HE(ed.spcv_edg, i, e, k));
      }
      for (int j = k; j <= kdm; j += LPC) __builtin_nontemporal_store(0., &HPR(i, e, j));
      continue;
    }
    const double *hrow = sh_h[cc][e];
    for (int j = k; j <= kdm; j += LPC) __builtin_nontemporal_store(hrow[j], &HPR(i, e, j));
    const int em1 = e_prev(e, np), ep1 = e_next(e, np);
    double blok = ci_open(ci, e) ? blokf(di, pik1, PB[e]) : 1.;
    const double nm1 = M[e] * (1. - blok) + mi * blok, nv1 = SV[e] * (1. - blok) + svi * blok;
    blok = ci_open(ci, em1) ? blokf(di, pik1, at_prev(PB, e, np)) : 1.;
    const double nm2 = at_prev(M, e, np) * (1. - blok) + mi * blok, nv2 = at_prev(SV, e, np) * (1. - blok) + svi * blok;
    blok = ci_open(ci, ep1) ? blokf(di, pik1, at_next(PB, e, np)) : 1.;
    const double nm3 = at_next(M, e, np) * (1. - blok) + mi * blok, nv3 = at_next(SV, e, np) * (1. - blok) + svi * blok;
    if (live) {
      __builtin_nontemporal_store(hmax(0., hrow[k + 1] - hrow[k]), &HE(s.dpm_edg, i, e, k));
      __builtin_nontemporal_store((8. * (mi + nm1) + nm2 + nm3) * H_DIVBY18, &HE(ed.mont_edg, i, e, k));
      __builtin_nontemporal_store((8. * (svi + nv1) + nv2 + nv3) * H_DIVBY18, &HE(ed.spcv_edg, i, e, k));
    }
  }
}

}  // namespace
}  // namespace hyc
}  // namespace fimhip

using namespace fimhip;
using namespace fimhip::hyc;

template <int LPC>
static void edg_kernels(fimhip_ocn *o, int leapm, int leapn) {
  const HycGrid &g = o->g;
  const HycState &s = o->s;
  const HycEdg &ed = o->ed;
  hipStream_t st = o->stream;
  const dim3 grid(ocn_blocks(o, LPC)), blk(256);
  if (o->janjic) KT_LAUNCH("hyc_janjic", st, hyc_janjic<LPC><<<grid, blk, 0, st>>>(g, s, ed));
  // one launch per leapfrog level (both levels in one kernel would hold every
  // edge's rotation coefficients in registers across them)
  KT_LAUNCH("hyc_uvedge", st,
            hyc_uvedge<LPC><<<grid, blk, 0, st>>>(g, s, ed, leapm, leapm, s.um, s.vm, 1, ed.uvsq_edg));
  KT_LAUNCH("hyc_uvedge", st, hyc_uvedge<LPC><<<grid, blk, 0, st>>>(g, s, ed, leapm, leapn, s.un, s.vn, 0, nullptr));
  KT_LAUNCH("hyc_edgint", st, hyc_edgint<LPC><<<grid, blk, 0, st>>>(g, s, ed, leapm));
}

// bclin_edgvar on the device-resident arrays (hycom_edgvar.F90:128-478)
static int edg_launch(fimhip_ocn *o, int leapm, int leapn) {
  if (leapm < 1 || leapm > 2 || leapn < 1 || leapn > 2 || leapm == leapn) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  g_ktimer = o->kt;
  if (o->kdm <= 32) edg_kernels<32>(o, leapm, leapn);
  else edg_kernels<64>(o, leapm, leapn);
  g_ktimer = nullptr;
  HIPCHK(hipGetLastError());
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_set_janjic(fimhip_ocn *o, int on) {
  if (!o) return FIMHIP_EINVAL;
  o->janjic = on != 0;
  return FIMHIP_OK;
}

// asynchronous on the instance's stream: the continuity step, a download or a
// profile report that follows waits for it
extern "C" int fimhip_ocn_edgvar(fimhip_ocn *o, int leapm, int leapn) {
  if (!o) return FIMHIP_EINVAL;
  return edg_launch(o, leapm, leapn);
}

extern "C" int fimhip_ocn_edgvar_r8(fimhip_ocn *o, int nstep, int leapm, int leapn, const double *uclin,
                                   const double *vclin, const double *utrop, const double *vtrop, const double *dp,
                                   double *pres, const double *geop, const double *montg, const double *spcifv,
                                   double *um_edg, double *vm_edg, double *un_edg, double *vn_edg, double *dp_edg,
                                   double *pr_edg, double *geop_edg, double *mont_edg, double *uvsq_edg,
                                   double *spcv_edg) {
  (void)nstep;      // the reference uses it for diagnostics only
  if (!o || !uclin || !vclin || !utrop || !vtrop || !dp || !pres || !geop || !montg || !spcifv || !um_edg ||
      !vm_edg || !un_edg || !vn_edg || !dp_edg || !pr_edg || !geop_edg || !mont_edg || !uvsq_edg || !spcv_edg)
    return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  // inputs, and the one output bclin_edgvar leaves partly untouched (geop_edg
  // outside the ocean, and entirely without janjic_ocn)
  const double *in[] = {uclin, vclin, utrop, vtrop, dp, pres, geop, montg, spcifv, geop_edg};
  const int fi[] = {FIMHIP_OCN_UCLIN, FIMHIP_OCN_VCLIN, FIMHIP_OCN_UTROP, FIMHIP_OCN_VTROP, FIMHIP_OCN_DP,
                    FIMHIP_OCN_PRES, FIMHIP_OCN_GEOP, FIMHIP_OCN_MONTG, FIMHIP_OCN_SPCIFV, FIMHIP_OCN_GEOP_EDG};
  const int nin = o->janjic ? 10 : 9;
  for (int j = 0; j < nin; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fi[j], &n);
    HIPCHK(hipMemcpyAsync(d, in[j], n * sizeof(double), hipMemcpyHostToDevice, o->stream));
  }
  const int r = edg_launch(o, leapm, leapn);
  if (r) return r;
  double *outp[] = {pres, um_edg, vm_edg, un_edg, vn_edg, dp_edg, pr_edg, mont_edg, uvsq_edg, spcv_edg, geop_edg};
  const int fo[] = {FIMHIP_OCN_PRES, FIMHIP_OCN_UM_EDG, FIMHIP_OCN_VM_EDG, FIMHIP_OCN_UN_EDG, FIMHIP_OCN_VN_EDG,
                    FIMHIP_OCN_DPM_EDG, FIMHIP_OCN_PR_EDG, FIMHIP_OCN_MONT_EDG, FIMHIP_OCN_UVSQ_EDG,
                    FIMHIP_OCN_SPCV_EDG, FIMHIP_OCN_GEOP_EDG};
  const int nout = o->janjic ? 11 : 10;
  for (int j = 0; j < nout; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fo[j], &n);
    HIPCHK(hipMemcpyAsync(outp[j], d, n * sizeof(double), hipMemcpyDeviceToHost, o->stream));
  }
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}
