// HYCOM's tracer transport (FIMsrc/fim/horizontal/hycom_transp3d.F90 transp0,
// transp1, transp2, with hycom_fct3d.F90's fct3d and hycom_dffusn.F90's
// dffusn_lyr) on the GPU, in FP64, with the reference's operations in its
// order.  The tracers of transp2 (temperature + 10, salinity, the passive
// tracers) go through fct3d together, one per grid row (blockIdx.y):
//   hyc_tr_w      the vertical mass flux the time-integrated horizontal
//                 fluxes imply (a column sum) and the tracers' working copies
//                 (hycom_transp3d.F90:149-172, 196-230)
//   hyc_fct_vert  parabolic fits in each layer and the vertical fluxes: each
//                 interface's lane walks the layers its displacement spans,
//                 over fits the column's lanes put in LDS; the
//                 conservation sums' column parts before the update
//                 (hycom_fct3d.F90:135-309)
//   hyc_fct_lo    low-order and antidiffusive edge fluxes, local bounds, the
//                 low-order update (into its own array: the neighbours still
//                 read the old values) (311-388)
//   hyc_fct_ratio the bounds widened by the low-order neighbours, limiter
//                 ratios (389-436)
//   hyc_fct_clip  the limited antidiffusive update; the conservation sums'
//                 column parts after it (437-498)
//   hyc_fct_sums  the conservation sums over the cells, one wave per sum,
//                 in cell order as the reference's serial loops add them
//   hyc_fct_fix   the shift that restores each tracer's ocean mean, the
//                 tracers written back (512-544, hycom_transp3d.F90:250-264)
//   hyc_dff_flux, hyc_dff_apply  dffusn_lyr for temperature and salinity
//                 (hycom_dffusn.F90:11-116)
#include <vector>

#include "hycom_internal.h"

namespace fimhip {
namespace hyc {
namespace {

constexpr double H_TOFFSET = 10.;        // keeps temperature positive (hycom_transp3d.F90)
constexpr double H_ONEMU_F = 1.e-6;      // fct3d's onemu
constexpr double H_ATHIRD = 1. / 3.;
constexpr int H_BCLIN_FRQ = 30;          // hycom_control.F90:9

// the column scan that starts at its first element: v(0), v(0)+v(1), ...
__device__ __forceinline__ double col_scan0(double *buf, double v, int k, int kdm) {
  buf[k] = v;
  wave_sync();
  double prev = buf[0], mine = prev;
  for (int j = 1; j < kdm; j++) {
    prev = prev + buf[j];
    if (j == k) mine = prev;
  }
  wave_sync();
  return mine;
}
// the sum over the column's layers of v, in layer order from 0
__device__ __forceinline__ double col_sum0(double *buf, double v, int k, int kdm) {
  buf[k] = v;
  wave_sync();
  double acc = 0.;
  for (int j = 0; j < kdm; j++) acc = acc + buf[j];
  wave_sync();
  return acc;
}

#define TF(a, t, i, k) (a)[((size_t)(t) * nip + (i)) * kdm + (k)]

// transp0: the integral restarts, dpinit from the new level
template <int LPC>
__global__ void __launch_bounds__(256) hyc_tr_init(HycGrid g, HycState s, HycTr tr, int leapn) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  const bool wi = ci_wet(g.cinfo[i]);
#pragma unroll
  for (int e = 0; e < NPP; e++) HE(tr.cumuflx, i, e, k) = wi ? 0. : H_LAND;
  if (wi) __builtin_nontemporal_store(HC(s.dp + (size_t)(leapn - 1) * nip * kdm, i, k), &HC(tr.dpinit, i, k));
}

// transp1: cumuflx += massflx * 2 * batrop
template <int LPC>
__global__ void __launch_bounds__(256) hyc_tr_accum(HycGrid g, HycTr tr, const double *massflx, double batrop) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k;
  const int ci = g.cinfo[i];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
#pragma unroll
  for (int e = 0; e < NPP; e++)
    if (e < np) __builtin_nontemporal_store(HE(tr.cumuflx, i, e, k) + HE(massflx, i, e, k) * 2. * batrop, &HE(tr.cumuflx, i, e, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_tr_w(HycGrid g, HycTr tr, const double *dpfinl) {
  __shared__ double sh[256 / LPC][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  const bool wi = ci_wet(ci);
  // the tracers' working copies, temperature offset
  if (live)
    for (int t = 0; t < tr.nfld; t++) {
      const double *src = t == 0 ? tr.temp : t == 1 ? tr.saln : tr.passv + (size_t)(t - 2) * nip * kdm;
      __builtin_nontemporal_store(wi ? (t == 0 ? HC(src, i, k) + H_TOFFSET : HC(src, i, k)) : H_LAND, &TF(tr.fld, t, i, k));
    }
  if (!wi) {
    if (live) __builtin_nontemporal_store(H_LAND, &HC(tr.w, i, k));
    return;
  }
  const int np = ci_np(ci);
  double hordiv = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++)
    if (e < np) hordiv = hordiv + HE(tr.cumuflx, i, e, kk);
  const double vertdv = (HC(tr.dpinit, i, kk) - HC(dpfinl, i, kk)) - hordiv * g.rarea[i];
  const double w = col_scan0(sh[c.cc], live ? vertdv : 0., k, kdm);
  if (live) __builtin_nontemporal_store(w, &HC(tr.w, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_fct_vert(HycGrid g, HycTr tr) {
  constexpr int CPB = 256 / LPC;
  __shared__ double sa[CPB][LPC + 1], sb[CPB][LPC + 1], sc[CPB][LPC + 1], st[CPB][LPC + 1];
  __shared__ double sm[CPB][LPC + 1];
  __shared__ int sz[CPB][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k, cc = c.cc, t = blockIdx.y;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  if (!ci_wet(g.cinfo[i])) return;
  const double eps = 1.e-11;
  const double f = TF(tr.fld, t, i, kk), th = HC(tr.dpinit, i, kk);
  // neighbours in the column (taken by every lane)
  const double fu = __shfl_up(f, 1, LPC), fd = __shfl_down(f, 1, LPC);
  const double tu = __shfl_up(th, 1, LPC), td = __shfl_down(th, 1, LPC);
  double a = f, b = 0., cq = 0.;
  if (k > 0 && k < kdm - 1) {
    const double dxlft = hmax(eps, tu), dxmid = hmax(eps, th), dxrgt = hmax(eps, td);
    const double yl = (dxlft * f + dxmid * fu) / (dxlft + dxmid);
    const double yr = (dxrgt * f + dxmid * fd) / (dxrgt + dxmid);
    a = 1.5 * f - .25 * (yl + yr);
    b = yr - yl;
    cq = 6. * (.5 * (yl + yr) - f);
    if (fabs(yr - yl) < 6. * fabs(.5 * (yl + yr) - f)) {
      if (fabs(yr - yl) > 2. * fabs(.5 * (yl + yr) - f)) {
        if ((yr - yl) * (.5 * (yl + yr) - f) > 0.) {
          a = .25 * (3. * f + yl);
          cq = 3. * (f - yl);
          b = cq;
        } else {
          a = .25 * (3. * f + yr);
          cq = 3. * (f - yr);
          b = -cq;
        }
      } else {
        a = f;
        b = cq = 0.;
      }
    }
  }
  sa[cc][k] = a;
  sb[cc][k] = b;
  sc[cc][k] = cq;
  st[cc][k] = th;
  wave_sync();
  // the vertical flux through the interface below layer k (lanes k < kdm-1)
  const double wk = live ? HC(tr.w, i, k) : 0.;
  double amount = 0., slab = H_ONEMU_F;
  const bool moved = wk < 0. || wk > 0.;
  if (k < kdm - 1) {
    if (wk < 0.) {
      amount = slab * fd;
      for (int kp = k + 1; kp < kdm; kp++) {
        if (slab >= -wk) break;
        const double tk = st[cc][kp];
        if (tk > 0.) {
          const double dslab = hmin(slab + tk, -wk) - hmin(slab, -wk);
          const double dx = dslab / tk;
          const double fcdx = sa[cc][kp] + sb[cc][kp] * .5 * (dx - 1.) + sc[cc][kp] * (.25 - dx * (.5 - dx * H_ATHIRD));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
      }
    } else if (wk > 0.) {
      amount = slab * f;
      int kp = k;
      for (;;) {          // layer k first, whatever k, then up to the second layer
        if (slab >= wk) break;
        const double tk = st[cc][kp];
        if (tk > 0.) {
          const double dslab = hmin(slab + tk, wk) - hmin(slab, wk);
          const double dx = dslab / tk;
          const double fcdx = sa[cc][kp] + sb[cc][kp] * .5 * (1. - dx) + sc[cc][kp] * (.25 - dx * (.5 - dx * H_ATHIRD));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
        if (kp > 1) kp--;
        else break;
      }
    }
  }
  // where w is neither < 0 nor > 0 the reference keeps the amount of the
  // nearest interface above where it was (its loop variable), 0 if none
  sm[cc][k] = amount;
  sz[cc][k] = moved ? 0 : 1;
  wave_sync();
  if (k < kdm - 1 && !moved) {
    amount = 0.;
    for (int j = k - 1; j >= 0; j--)
      if (!sz[cc][j]) {
        amount = sm[cc][j];
        break;
      }
  }
  wave_sync();
  const double vfx = k < kdm - 1 ? wk * amount / slab : 0.;
  const double vfu = __shfl_up(vfx, 1, LPC);
  if (live) {
    __builtin_nontemporal_store(vfx, &TF(tr.vfx, t, i, k));
    __builtin_nontemporal_store(k == 0 ? vfx : vfx - vfu, &TF(tr.vdv, t, i, k));
  }
  // the conservation sums before the update: this column's parts
  const double area = g.area[i];
  const double col1 = col_sum0(sm[cc], live ? th * area : 0., k, kdm);
  const double col2 = col_sum0(sm[cc], live ? f * th * area : 0., k, kdm);
  if (k == 0) {
    const int r = tr.rank[i];
    tr.colsum[((size_t)t * 4 + 0) * tr.nwp + r] = col1;
    tr.colsum[((size_t)t * 4 + 1) * tr.nwp + r] = col2;
  }
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_fct_lo(HycGrid g, HycTr tr, const double *thkn) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k, t = blockIdx.y;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const double f = TF(tr.fld, t, i, k);
  double fmx = f, fmn = f, hordiv = 0.;
  // every slot's operands are loaded before any is used (a pentagon's sixth
  // slot repeats its fifth), so the loads overlap
  double uea[NPP], uxa[NPP], fxa[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    uea[e] = HE(tr.cumuflx, i, e, k);
    uxa[e] = HE(tr.cumuflx, nb[e], g.proxs[i * NPP + e], k);
    fxa[e] = TF(tr.fld, t, nb[e], k);
  }
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    const double ue = uea[e], ux = uxa[e], fx = fxa[e];
    double flx = 0., uan = 0.;
    if (ci_open(ci, e)) {
      flx = 0.5 * ((ue + fabs(ue)) * f - (ux + fabs(ux)) * fx);
      const double q = .5 * (f + fx);
      uan = q * ue - flx;
      fmx = hmax(fmx, fx);
      fmn = hmin(fmn, fx);
    }
    tr.uan[((size_t)t * nip * NPP + (size_t)i * NPP + e) * kdm + k] = uan;
    hordiv = hordiv + flx;
  }
  const double wk = HC(tr.w, i, k);
  if (k < kdm - 1 && wk < 0.) {
    const double v = TF(tr.vfx, t, i, k) / wk;
    fmx = hmax(fmx, v);
    fmn = hmin(fmn, v);
  }
  if (k > 0) {
    const double wu = HC(tr.w, i, k - 1);
    if (wu > 0.) {
      const double v = TF(tr.vfx, t, i, k - 1) / wu;
      fmx = hmax(fmx, v);
      fmn = hmin(fmn, v);
    }
  }
  hordiv = hordiv * g.rarea[i];
  const double tn = HC(thkn, i, k);
  const double unbound = f * HC(tr.dpinit, i, k) - hordiv - TF(tr.vdv, t, i, k);
  const double bounded = hmax(hmax(0., fmn * tn), hmin(unbound, fmx * tn));
  __builtin_nontemporal_store((f * H_ONEMU_F + bounded) / (H_ONEMU_F + tn), &TF(tr.flo, t, i, k));
  __builtin_nontemporal_store(fmx, &TF(tr.fmx, t, i, k));
  __builtin_nontemporal_store(fmn, &TF(tr.fmn, t, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_fct_ratio(HycGrid g, HycTr tr, const double *thkn) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k, t = blockIdx.y;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  double fmx = TF(tr.fmx, t, i, k), fmn = TF(tr.fmn, t, i, k);
  const double f = TF(tr.flo, t, i, k);
  double totin = 0., totou = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    const double u = tr.uan[((size_t)t * nip * NPP + (size_t)i * NPP + e) * kdm + k];
    totin = totin - hmin(0., u);
    totou = totou + hmax(0., u);
    const double fx = TF(tr.flo, t, nb[e], k);
    if (ci_open(ci, e)) {
      fmx = hmax(fmx, fx);
      fmn = hmin(fmn, fx);
    }
  }
  const double tn = HC(thkn, i, k), ra = g.rarea[i];
  __builtin_nontemporal_store((fmx - f) * tn / (totin + 1.e-11) * ra, &TF(tr.flp, t, i, k));
  __builtin_nontemporal_store((f - fmn) * tn / (totou + 1.e-11) * ra, &TF(tr.fln, t, i, k));
  __builtin_nontemporal_store(fmx, &TF(tr.fmx, t, i, k));
  __builtin_nontemporal_store(fmn, &TF(tr.fmn, t, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_fct_clip(HycGrid g, HycTr tr, const double *thkn) {
  __shared__ double sh[256 / LPC][LPC + 1];
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok) return;
  const int i = c.i, k = c.k, t = blockIdx.y;
  const bool live = k < kdm;
  const int kk = live ? k : 0;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const double fln = TF(tr.fln, t, i, kk), flp = TF(tr.flp, t, i, kk);
  double hordiv = 0.;
  double ua[NPP], flpa[NPP], flna[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    ua[e] = tr.uan[((size_t)t * nip * NPP + (size_t)i * NPP + e) * kdm + kk];
    flpa[e] = TF(tr.flp, t, nb[e], kk);
    flna[e] = TF(tr.fln, t, nb[e], kk);
  }
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    const double u = ua[e], flpx = flpa[e], flnx = flna[e];
    if (ci_open(ci, e)) {
      const double clip = u >= 0. ? hmin(hmin(1., fln), flpx) : hmin(hmin(1., flp), flnx);
      hordiv = hordiv + u * clip;
    }
  }
  hordiv = hordiv * g.rarea[i];
  const double f = TF(tr.flo, t, i, kk), tn = HC(thkn, i, kk);
  const double unbound = f * tn - hordiv;
  const double bounded = hmax(hmax(0., TF(tr.fmn, t, i, kk) * tn), hmin(unbound, TF(tr.fmx, t, i, kk) * tn));
  const double fnew = (f * H_ONEMU_F + bounded) / (H_ONEMU_F + tn);
  if (live) __builtin_nontemporal_store(fnew, &TF(tr.fld, t, i, k));
  // the conservation sums after the update: this column's parts
  const double area = g.area[i];
  const double col1 = col_sum0(sh[c.cc], live ? tn * area : 0., k, kdm);
  const double col2 = col_sum0(sh[c.cc], live ? fnew * tn * area : 0., k, kdm);
  if (k == 0) {
    const int r = tr.rank[i];
    tr.colsum[((size_t)t * 4 + 2) * tr.nwp + r] = col1;
    tr.colsum[((size_t)t * 4 + 3) * tr.nwp + r] = col2;
  }
}

// sum c of the conservation step: 0 wtold (the thickness before, part 0 of
// the first tracer's column parts), 1 wtnew (after, part 2), then each
// tracer's bfore (part 1) and after (part 3), over the ocean columns in cell
// order (the column parts are stored by ocean rank, each part's row padded
// with +0 to nwp = a multiple of 256 past nwet + 256).  One wave per sum: it
// loads blocks of 256 parts with coalesced loads, the next block in flight
// while the current one, staged in LDS, is added in order into the chain
// every lane carries (a dependent FP64 add takes ~7 ns here, so a sum over
// G8's 458753 ocean columns takes ~3.4 ms).  The padding leaves a sum
// unchanged: no partial sum is -0 (it starts at +0, and a sum is -0 only when
// both terms are).
constexpr int SUMS_BLK = 256;
__global__ void __launch_bounds__(64) hyc_fct_sums(const double *__restrict__ colsum, int nwet, int nwp,
                                                   double *__restrict__ sums) {
  __shared__ double buf[SUMS_BLK];
  const int c = blockIdx.x, l = threadIdx.x;
  const int t = c < 2 ? 0 : (c - 2) >> 1;
  const int part = c == 0 ? 0 : c == 1 ? 2 : (c & 1) ? 3 : 1;
  const double *__restrict__ x = colsum + ((size_t)t * 4 + part) * nwp;
  double r[SUMS_BLK / 64];
#pragma unroll
  for (int j = 0; j < SUMS_BLK / 64; j++) r[j] = x[l + 64 * j];
  double acc = 0.;
  for (int base = 0; base < nwet; base += SUMS_BLK) {
#pragma unroll
    for (int j = 0; j < SUMS_BLK / 64; j++) buf[l + 64 * j] = r[j];
    const int nx = base + SUMS_BLK;
#pragma unroll
    for (int j = 0; j < SUMS_BLK / 64; j++) r[j] = x[nx + l + 64 * j];     // nx + 256 <= nwp
    wave_sync();
    // 32 parts in registers ahead of their adds, so that the LDS reads
    // overlap the chain
#pragma unroll
    for (int h = 0; h < SUMS_BLK; h += 32) {
      double v[32];
#pragma unroll
      for (int j = 0; j < 32; j++) v[j] = buf[h + j];
#pragma unroll
      for (int j = 0; j < 32; j++) acc = acc + v[j];
    }
    wave_sync();
  }
  if (l == 0) sums[c] = acc;
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_fct_fix(HycGrid g, HycTr tr, const double *sums) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k, t = blockIdx.y;
  if (!ci_wet(g.cinfo[i])) return;
  const double bfore = sums[2 + 2 * t] / sums[0], after = sums[3 + 2 * t] / sums[1];
  const double f = TF(tr.fld, t, i, k) + (bfore - after);
  double *dst = t == 0 ? tr.temp : t == 1 ? tr.saln : tr.passv + (size_t)(t - 2) * nip * kdm;
  __builtin_nontemporal_store(t == 0 ? f - H_TOFFSET : f, &HC(dst, i, k));
}

// dffusn_lyr: the fluxes through the ocean edges, then the update (two
// kernels: the update changes what the neighbours read)
template <int LPC>
__global__ void __launch_bounds__(256) hyc_dff_flux(HycGrid g, HycTr tr, const double *delp) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k, t = blockIdx.y;
  const int ci = g.cinfo[i];
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = g.prox[i * NPP + e];
  if (!ci_wet(ci)) return;
  const int np = ci_np(ci);
  const double *fld = t == 0 ? tr.temp : tr.saln;
  const double fi = HC(fld, i, k), a = hmax(HC(delp, i, k), 1.e-11);
  double flxdv = 0.;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    if (e >= np) break;
    const int ix = nb[e];
    const double fx = HC(fld, ix, k), b = hmax(HC(delp, ix, k), 1.e-11);
    if (ci_open(ci, e)) flxdv = flxdv + (fi - fx) * g.sideln[i * NPP + e] * 2. * (a * b / (a + b));
  }
  __builtin_nontemporal_store(flxdv, &TF(tr.flxdv, t, i, k));
}

template <int LPC>
__global__ void __launch_bounds__(256) hyc_dff_apply(HycGrid g, HycTr tr, const double *delp, double dffveldt) {
  const int kdm = g.kdm, nip = g.nip;
  const Col<LPC> c = hcol<LPC>(nip);
  if (!c.ok || c.k >= kdm) return;
  const int i = c.i, k = c.k, t = blockIdx.y;
  if (!ci_wet(g.cinfo[i])) return;
  double *fld = t == 0 ? tr.temp : tr.saln;
  const double factor = -dffveldt * g.rarea[i] / hmax(HC(delp, i, k), 1.e-11);
  __builtin_nontemporal_store(HC(fld, i, k) + TF(tr.flxdv, t, i, k) * factor, &HC(fld, i, k));
}

}  // namespace
}  // namespace hyc
}  // namespace fimhip

using namespace fimhip;
using namespace fimhip::hyc;

// the transport's arrays exist; transp1 and transp2 also need batrop
static int tr_alloc(const fimhip_ocn *o) { return o->tr.nfld > 0; }
static int tr_ready(const fimhip_ocn *o) { return tr_alloc(o) && o->mo.batrop > 0.; }

template <int LPC>
static int tr2_kernels(fimhip_ocn *o, const double *dpfinl) {
  const HycGrid &g = o->g;
  HycTr &tr = o->tr;
  hipStream_t st = o->stream;
  const unsigned nb = ocn_blocks(o, LPC);
  const dim3 g1(nb), gf(nb, tr.nfld), blk(256);
  KT_LAUNCH("hyc_tr_w", st, hyc_tr_w<LPC><<<g1, blk, 0, st>>>(g, tr, dpfinl));
  KT_LAUNCH("hyc_fct_vert", st, hyc_fct_vert<LPC><<<gf, blk, 0, st>>>(g, tr));
  KT_LAUNCH("hyc_fct_lo", st, hyc_fct_lo<LPC><<<gf, blk, 0, st>>>(g, tr, dpfinl));
  KT_LAUNCH("hyc_fct_ratio", st, hyc_fct_ratio<LPC><<<gf, blk, 0, st>>>(g, tr, dpfinl));
  KT_LAUNCH("hyc_fct_clip", st, hyc_fct_clip<LPC><<<gf, blk, 0, st>>>(g, tr, dpfinl));
  // the conservation shift of each tracer (hycom_fct3d.F90:290-309, 482-511:
  // serial sums over the ocean cells, as written; the same thickness sums
  // serve every tracer)
  KT_LAUNCH("hyc_fct_sums", st, hyc_fct_sums<<<dim3(2 + 2 * tr.nfld), dim3(64), 0, st>>>(tr.colsum, tr.nwet, tr.nwp, tr.q));
  KT_LAUNCH("hyc_fct_fix", st, hyc_fct_fix<LPC><<<gf, blk, 0, st>>>(g, tr, tr.q));
  if (tr.temdff > 0.) {
    const dim3 g2(nb, 2);
    const double dffveldt = tr.temdff * o->mo.batrop * H_BCLIN_FRQ;
    KT_LAUNCH("hyc_dff_flux", st, hyc_dff_flux<LPC><<<g2, blk, 0, st>>>(g, tr, dpfinl));
    KT_LAUNCH("hyc_dff_apply", st, hyc_dff_apply<LPC><<<g2, blk, 0, st>>>(g, tr, dpfinl, dffveldt));
  }
  HIPCHK(hipGetLastError());
  return FIMHIP_OK;
}

static int tr2_launch(fimhip_ocn *o, const double *dpfinl) {
  if (!tr_ready(o)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  g_ktimer = o->kt;
  const int r = o->kdm <= 32 ? tr2_kernels<32>(o, dpfinl) : tr2_kernels<64>(o, dpfinl);
  g_ktimer = nullptr;
  return r;
}

extern "C" int fimhip_ocn_set_transport(fimhip_ocn *o, int numtr, double temdff) {
  if (!o || numtr < 0 || numtr > 8) return FIMHIP_EINVAL;
  if (o->tr.nfld > 0) {             // the arrays exist: only temdff may change
    if (o->tr.nfld != 2 + numtr) return FIMHIP_EINVAL;
    o->tr.temdff = temdff;
    return FIMHIP_OK;
  }
  HIPCHK(hipSetDevice(o->device));
  const int nf = 2 + numtr, nip = o->nip, kdm = o->kdm;
  const size_t cz = (size_t)nip * kdm, ez = cz * NPP;
  HycTr &tr = o->tr;
  // each ocean cell's rank among the ocean cells in cell order; the rows of
  // the conservation sums' column parts padded for their summing kernel
  std::vector<int> rank(nip, 0);
  int nwet = 0;
  for (int i = 0; i < nip; i++)
    if (o->wet[i]) rank[i] = nwet++;
  const int nwp = (nwet + 255) / 256 * 256 + 256;
  int r = 0;
  struct { double **p; size_t n; } arrs[] = {
      {&tr.cumuflx, ez}, {&tr.dpinit, cz}, {&tr.dpfinl, cz}, {&tr.passv, numtr > 0 ? numtr * cz : 1}, {&tr.w, cz}, {&tr.fld, nf * cz}, {&tr.flo, nf * cz},
      {&tr.vfx, nf * cz}, {&tr.vdv, nf * cz}, {&tr.fmx, nf * cz}, {&tr.fmn, nf * cz}, {&tr.flp, nf * cz},
      {&tr.fln, nf * cz}, {&tr.uan, nf * ez}, {&tr.colsum, (size_t)nf * 4 * nwp}, {&tr.flxdv, 2 * cz},
      {&tr.q, (size_t)(2 + 2 * nf)}};
  for (auto &a : arrs)
    if ((r = o->arena.alloc(a.p, a.n))) return r;
  int *drank = nullptr;
  if ((r = o->arena.alloc(&drank, (size_t)nip))) return r;
  HIPCHK(hipMemcpyAsync(drank, rank.data(), nip * sizeof(int), hipMemcpyHostToDevice, o->stream));
  HIPCHK(hipMemsetAsync(tr.colsum, 0, (size_t)nf * 4 * nwp * sizeof(double), o->stream));
  tr.rank = drank;
  tr.nwet = nwet;
  tr.nwp = nwp;
  HIPCHK(hipStreamSynchronize(o->stream));
  tr.nfld = nf;
  tr.numtr = numtr;
  tr.temdff = temdff;
  return FIMHIP_OK;
}

// transp0 on the device-resident arrays (hycom_transp3d.F90:18-51)
extern "C" int fimhip_ocn_transp0(fimhip_ocn *o, int leapn) {
  if (!o || !tr_alloc(o) || leapn < 1 || leapn > 2) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  g_ktimer = o->kt;
  const dim3 blk(256);
  if (o->kdm <= 32) KT_LAUNCH("hyc_tr_init", o->stream, hyc_tr_init<32><<<ocn_blocks(o, 32), blk, 0, o->stream>>>(o->g, o->s, o->tr, leapn));
  else KT_LAUNCH("hyc_tr_init", o->stream, hyc_tr_init<64><<<ocn_blocks(o, 64), blk, 0, o->stream>>>(o->g, o->s, o->tr, leapn));
  g_ktimer = nullptr;
  HIPCHK(hipGetLastError());
  return FIMHIP_OK;
}

// transp1 with the instance's mass flux (hycom_transp3d.F90:54-85)
extern "C" int fimhip_ocn_transp1(fimhip_ocn *o) {
  if (!o || !tr_ready(o)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  g_ktimer = o->kt;
  const dim3 blk(256);
  if (o->kdm <= 32) KT_LAUNCH("hyc_tr_accum", o->stream, hyc_tr_accum<32><<<ocn_blocks(o, 32), blk, 0, o->stream>>>(o->g, o->tr, o->s.massflx, o->mo.batrop));
  else KT_LAUNCH("hyc_tr_accum", o->stream, hyc_tr_accum<64><<<ocn_blocks(o, 64), blk, 0, o->stream>>>(o->g, o->tr, o->s.massflx, o->mo.batrop));
  g_ktimer = nullptr;
  HIPCHK(hipGetLastError());
  return FIMHIP_OK;
}

// transp2 (hycom_transp3d.F90:88-321) with dpfinl = dp(:,:,leapn), as
// hycom_run.F90:398-400 calls it
extern "C" int fimhip_ocn_transp2(fimhip_ocn *o, int leapn) {
  if (!o || leapn < 1 || leapn > 2) return FIMHIP_EINVAL;
  return tr2_launch(o, o->s.dp + (size_t)(leapn - 1) * o->nip * o->kdm);
}

static int ocn_put(fimhip_ocn *o, int f, const double *h) {
  size_t n = 0;
  double *d = ocn_field(o, f, &n);
  if (!d) return FIMHIP_EINVAL;
  HIPCHK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, o->stream));
  return FIMHIP_OK;
}
static int ocn_get(fimhip_ocn *o, int f, double *h) {
  size_t n = 0;
  double *d = ocn_field(o, f, &n);
  if (!d) return FIMHIP_EINVAL;
  HIPCHK(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, o->stream));
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_transp0_r8(fimhip_ocn *o, int nstep, int leapn, double *cumuflx, const double *dp,
                                    double *dpinit) {
  (void)nstep;
  if (!o || !cumuflx || !dp || !dpinit || !tr_alloc(o)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  int r;
  if ((r = ocn_put(o, FIMHIP_OCN_DP, dp)) || (r = ocn_put(o, FIMHIP_OCN_DPINIT, dpinit))) return r;
  if ((r = fimhip_ocn_transp0(o, leapn))) return r;
  if ((r = ocn_get(o, FIMHIP_OCN_CUMUFLX, cumuflx)) || (r = ocn_get(o, FIMHIP_OCN_DPINIT, dpinit))) return r;
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_transp1_r8(fimhip_ocn *o, int nstep, double *cumuflx, const double *massflx) {
  (void)nstep;
  if (!o || !cumuflx || !massflx || !tr_ready(o)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  int r;
  if ((r = ocn_put(o, FIMHIP_OCN_CUMUFLX, cumuflx)) || (r = ocn_put(o, FIMHIP_OCN_MASSFLX, massflx))) return r;
  if ((r = fimhip_ocn_transp1(o))) return r;
  if ((r = ocn_get(o, FIMHIP_OCN_CUMUFLX, cumuflx))) return r;
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_transp2_r8(fimhip_ocn *o, int nstep, double *temp, double *saln, double *passv_tr,
                                    const double *cumuflx, const double *dpinit, const double *dpfinl) {
  (void)nstep;
  if (!o || !temp || !saln || !cumuflx || !dpinit || !dpfinl || !tr_ready(o) || (o->tr.numtr > 0 && !passv_tr))
    return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  int r;
  if ((r = ocn_put(o, FIMHIP_OCN_TEMP, temp)) || (r = ocn_put(o, FIMHIP_OCN_SALN, saln)) ||
      (r = ocn_put(o, FIMHIP_OCN_CUMUFLX, cumuflx)) || (r = ocn_put(o, FIMHIP_OCN_DPINIT, dpinit)) ||
      (r = ocn_put(o, FIMHIP_OCN_DPFINL, dpfinl)))
    return r;
  if (o->tr.numtr > 0 && (r = ocn_put(o, FIMHIP_OCN_PASSV, passv_tr))) return r;
  if ((r = tr2_launch(o, o->tr.dpfinl))) return r;
  if ((r = ocn_get(o, FIMHIP_OCN_TEMP, temp)) || (r = ocn_get(o, FIMHIP_OCN_SALN, saln))) return r;
  if (o->tr.numtr > 0 && (r = ocn_get(o, FIMHIP_OCN_PASSV, passv_tr))) return r;
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}
