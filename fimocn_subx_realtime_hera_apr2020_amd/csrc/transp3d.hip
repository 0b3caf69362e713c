// Class-B tracer transport on the long time step, FP64 (module_transp3d,
// transp3d.F90:22-241, and module_fct3d, fct3d.F90:4-483, with
// diagno = recovr = .false.).  Class-B tracers are not touched by the
// class-A tracer kernels or hybgen; every step transp1 adds the AB-weighted
// mass flux to a time integral, and every dtratio steps transp2 derives the
// vertical mass flux that closes the time-integrated continuity equation and
// moves the class-B tracers with the 3-D FCT scheme fct3d, after which
// transp0 restarts the integral (dyn_run.F90:388-405).
//
// Operation order follows the reference exactly (this file is built with
// -ffp-contract=off), so the results have the bytes of the reference
// Fortran built with -fdefault-real-8.  Mapping: one wavefront per cell
// column, lane = level, blockIdx.y = class-B tracer; the column-serial parts
// (the vertical running sum of transp2) run in lane 0 over the column held
// in LDS; the two massless-cell sweeps of fct3d run one thread per column.
#include "fimhip_internal.h"

namespace fimhip {

#define C2(a, ipn, k) (a)[(size_t)(ipn) * nvl + (k)]
#define E3(a, ipn, e, k) (a)[((size_t)(ipn) * NPP + (e)) * nvl + (k)]
#define TB(a, n, ipn, k) (a)[((size_t)(n) * nip + (ipn)) * nvl + (k)]
#define TE(a, n, ipn, e, k) (a)[(((size_t)(n) * nip + (ipn)) * NPP + (e)) * nvl + (k)]
#define MF(a, s, ipn, e, k) (a)[(((size_t)(s) * nip + (ipn)) * NPP + (e)) * nvl + (k)]
#define PROX(ipn, e) ldc(g.prox, (ipn) * NPP + (e))
#define PROXS(ipn, e) ldc(g.proxs, (ipn) * NPP + (e))
#define GNP(ipn) ldc(g.nprox, (ipn))
#define GV(a, ipn) ldc(g.a, (ipn))

namespace {

constexpr int COLS = 4;   // columns per block
constexpr double ATHIRD = 1. / 3., EPSIL = 1.e-11, ONEMU = 1.e-6;   // fct3d.F90:48

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

__device__ __forceinline__ int wave_col() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * COLS + threadIdx.y));
}

// transp0 (transp3d.F90:22-41)
__global__ void __launch_bounds__(256) k_transp0(DevGrid g, DevState s, DevClassB b) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  C2(b.dpinit, ipn, k) = C2(s.dp, ipn, k);
#pragma unroll
  for (int e = 0; e < NPP; e++) E3(b.cumufx, ipn, e, k) = 0.;
}

// transp1 (transp3d.F90:44-78)
__global__ void __launch_bounds__(256) k_transp1(DevGrid g, DevState s, DevClassB b, AB ab) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  const int nf = ab.nf - 1, of = ab.of - 1, vof = ab.vof - 1;
  const int np = GNP(ipn);
  for (int edg = 0; edg < np; edg++)
    E3(b.cumufx, ipn, edg, k) = E3(b.cumufx, ipn, edg, k) + ab.ab1 * MF(s.massfx, nf, ipn, edg, k) +
                                ab.ab2 * MF(s.massfx, of, ipn, edg, k) + ab.ab3 * MF(s.massfx, vof, ipn, edg, k);
}

// transp2 (transp3d.F90:137-169): vertical mass flux w from the integrated
// horizontal divergence and the thickness change, summed upward from the
// bottom layer in lane 0
__global__ void __launch_bounds__(256) k_transp2_w(DevGrid g, DevState s, DevClassB b) {
  __shared__ double sh[COLS][NVL_MAX];
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, c = threadIdx.y, ipn = wave_col();
  const bool live = ipn < nip && k < nvl;
  if (live) {
    const int np = GNP(ipn);
    double hordiv = 0.;
    for (int edg = 0; edg < np; edg++) hordiv = hordiv + E3(b.cumufx, ipn, edg, k);
    sh[c][k] = (C2(b.dpinit, ipn, k) - C2(s.dp, ipn, k)) - hordiv * GV(rarea, ipn);
  }
  __syncthreads();
  if (k == 0 && ipn < nip) {
    double acc = sh[c][0];
    for (int kk = 1; kk < nvl; kk++) {
      acc = acc + sh[c][kk];
      sh[c][kk] = acc;
    }
  }
  __syncthreads();
  if (live) C2(b.w, ipn, k) = sh[c][k];
}

// fct3d (fct3d.F90:127-140): massless cells take their values from the
// layers above and below, in two dependent sweeps (downward, then upward).
// One thread per column and class-B tracer (64 columns per wavefront) with
// the column in registers; the column data pass through LDS in chunks of 16
// levels so that every global load and store instruction covers whole cache
// lines (16 levels of 4 columns).  Needs nvl == NVL.
template <int NVL>
__global__ void __launch_bounds__(64) k_fct3d_fill(DevGrid g, DevClassB b) {
  constexpr int CH = 16, NCH = NVL / CH;
  static_assert(NVL % CH == 0, "whole chunks");
  __shared__ double sf[CH][65], sth[CH][65];
  const int nip = g.nip, nvl = NVL;
  const int lane = threadIdx.x, c0 = blockIdx.x * 64, n = blockIdx.y;
  const double *trb = b.trb + (size_t)n * nip * nvl;
  double *fld = b.fld + (size_t)n * nip * nvl;
  double x[NVL];
#pragma unroll
  for (int ch = NCH - 1; ch >= 0; ch--) {       // downward sweep, top chunk first
    const int k0 = ch * CH;
    __syncthreads();
    for (int i = 0; i < 16; i++) {
      const int cc = i * 4 + (lane >> 4), kk = lane & 15, gc = c0 + cc;
      const bool in = gc < nip;
      sf[kk][cc] = in ? trb[(size_t)gc * nvl + k0 + kk] : 0.;
      sth[kk][cc] = in ? b.dpinit[(size_t)gc * nvl + k0 + kk] : 1.;
    }
    __syncthreads();
#pragma unroll
    for (int kk = CH - 1; kk >= 0; kk--) {
      const int k = k0 + kk;
      if (k == NVL - 1) {
        x[k] = sf[kk][lane];
      } else {
        const double t = sth[kk][lane];
        x[k] = (sf[kk][lane] * t + x[k + 1] * ONEMU) / (t + ONEMU);
      }
    }
  }
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {            // upward sweep, bottom chunk first
    const int k0 = ch * CH;
    __syncthreads();
    for (int i = 0; i < 16; i++) {
      const int cc = i * 4 + (lane >> 4), kk = lane & 15, gc = c0 + cc;
      sth[kk][cc] = gc < nip ? b.dpinit[(size_t)gc * nvl + k0 + kk] : 1.;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < CH; kk++) {
      const int k = k0 + kk;
      if (k > 0) {
        const double t = sth[kk][lane];
        x[k] = (x[k] * t + x[k - 1] * ONEMU) / (t + ONEMU);
      }
      sf[kk][lane] = x[k];
    }
    __syncthreads();
    for (int i = 0; i < 16; i++) {
      const int cc = i * 4 + (lane >> 4), kk = lane & 15, gc = c0 + cc;
      if (gc < nip) fld[(size_t)gc * nvl + k0 + kk] = sf[kk][cc];
    }
  }
}

// fct3d, first part (fct3d.F90:127-260): massless cells filled from the
// layers above and below (two serial sweeps, in lane 0, unless FILLED: then
// k_fct3d_fill has left the filled field in fld), the piecewise parabolic fit
// of each layer, and the vertical tracer flux over the upstream slab of
// thickness |w|.  Writes the filled field and the vertical flux.
template <bool FILLED>
__global__ void __launch_bounds__(256) k_fct3d_vert(DevGrid g, DevClassB b) {
  __shared__ double sf[COLS][NVL_MAX], sth[COLS][NVL_MAX];
  __shared__ double sa[COLS][NVL_MAX], sb[COLS][NVL_MAX], sc[COLS][NVL_MAX];
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, c = threadIdx.y, ipn = wave_col(), n = blockIdx.y;
  const bool live = ipn < nip && k < nvl;
  if (live) {
    sf[c][k] = FILLED ? TB(b.fld, n, ipn, k) : TB(b.trb, n, ipn, k);
    sth[c][k] = C2(b.dpinit, ipn, k);
  }
  __syncthreads();
  double *f = sf[c];
  const double *th = sth[c];
  if (!FILLED && k == 0 && ipn < nip) {
    for (int kk = nvl - 2; kk >= 0; kk--) f[kk] = (f[kk] * th[kk] + f[kk + 1] * ONEMU) / (th[kk] + ONEMU);
    for (int kk = 1; kk < nvl; kk++) f[kk] = (f[kk] * th[kk] + f[kk - 1] * ONEMU) / (th[kk] + ONEMU);
  }
  __syncthreads();
  if (live) {
    double a = f[k], bb = 0., cc = 0.;
    if (k > 0 && k < nvl - 1) {
      const double dxlft = dmax(EPSIL, th[k - 1]), dxmid = dmax(EPSIL, th[k]), dxrgt = dmax(EPSIL, th[k + 1]);
      const double fk = f[k];
      const double yl = (dxlft * fk + dxmid * f[k - 1]) / (dxlft + dxmid);
      const double yr = (dxrgt * fk + dxmid * f[k + 1]) / (dxrgt + dxmid);
      a = 1.5 * fk - .25 * (yl + yr);
      bb = yr - yl;
      cc = 6. * (.5 * (yl + yr) - fk);
      if (fabs(yr - yl) < 6. * fabs(.5 * (yl + yr) - fk)) {
        if (fabs(yr - yl) > 2. * fabs(.5 * (yl + yr) - fk)) {
          if ((yr - yl) * (.5 * (yl + yr) - fk) > 0.) {
            a = .25 * (3. * fk + yl);
            cc = 3. * (fk - yl);
            bb = cc;
          } else {
            a = .25 * (3. * fk + yr);
            cc = 3. * (fk - yr);
            bb = -cc;
          }
        } else {
          a = fk; bb = 0.; cc = 0.;
        }
      }
    }
    sa[c][k] = a; sb[c][k] = bb; sc[c][k] = cc;
  }
  __syncthreads();
  if (!live) return;
  const double *A = sa[c], *B = sb[c], *Cc = sc[c];
  double vfx = 0.;
  if (k < nvl - 1) {
    const double wk = C2(b.w, ipn, k);
    double slab = ONEMU, amount = 0.;   // a zero w gives a zero flux
    if (wk < 0.) {                      // interface moves up: slab from above
      amount = slab * f[k + 1];
      int kp = k;
      for (;;) {
        kp = kp + 1;
        if (slab >= -wk) break;
        if (th[kp] > 0.) {
          const double dslab = dmin(slab + th[kp], -wk) - dmin(slab, -wk);
          const double dx = dslab / th[kp];
          const double fcdx = A[kp] + B[kp] * .5 * (dx - 1.) + Cc[kp] * (.25 - dx * (.5 - dx * ATHIRD));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
        if (!(kp < nvl - 1)) break;
      }
    } else if (wk > 0.) {               // interface moves down: slab from below
      amount = slab * f[k];
      int kp = k + 1;
      for (;;) {
        kp = kp - 1;
        if (slab >= wk) break;
        if (th[kp] > 0.) {
          const double dslab = dmin(slab + th[kp], wk) - dmin(slab, wk);
          const double dx = dslab / th[kp];
          const double fcdx = A[kp] + B[kp] * .5 * (1. - dx) + Cc[kp] * (.25 - dx * (.5 - dx * ATHIRD));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
        if (!(kp > 1)) break;           // the reference continues while kp > 2 (1-based)
      }
    }
    vfx = wk * amount / slab;
  }
  TB(b.vfx, n, ipn, k) = vfx;
  if (!FILLED) TB(b.fld, n, ipn, k) = f[k];
}

// fct3d, second part (fct3d.F90:276-370): low-order and antidiffusive
// horizontal fluxes, the local bounds (including the tracer carried through
// the layer interfaces), the low-order update and the ratios of permissible
// change to incoming / outgoing antidiffusive flux
__global__ void __launch_bounds__(256) k_fct3d_lo(DevGrid g, DevState s, DevClassB b) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, ipn = wave_col(), n = blockIdx.y;
  if (ipn >= nip || k >= nvl) return;
  const int np = GNP(ipn);
  const double fl = TB(b.fld, n, ipn, k);
  double mx = fl, mn = fl, hd = 0., totin = 0., totou = 0.;
  for (int edg = 0; edg < np; edg++) {
    const int ix = PROX(ipn, edg), edx = PROXS(ipn, edg);
    const double ui = E3(b.cumufx, ipn, edg, k), ux = E3(b.cumufx, ix, edx, k);
    const double fx = TB(b.fld, n, ix, k);
    const double flx = 0.5 * ((ui + fabs(ui)) * fl - (ux + fabs(ux)) * fx);
    const double q = .5 * (fl + fx);
    const double an = q * ui - flx;
    TE(b.uan, n, ipn, edg, k) = an;
    mx = dmax(mx, fx);
    mn = dmin(mn, fx);
    hd = hd + flx;
    totin = totin - dmin(0., an);
    totou = totou + dmax(0., an);
  }
  const double vk = TB(b.vfx, n, ipn, k);
  if (k < nvl - 1) {
    const double wk = C2(b.w, ipn, k);
    if (wk < 0.) {
      mx = dmax(mx, vk / wk);
      mn = dmin(mn, vk / wk);
    }
  }
  double vdv = vk;
  if (k > 0) {
    const double vm = TB(b.vfx, n, ipn, k - 1), wm = C2(b.w, ipn, k - 1);
    if (wm > 0.) {
      mx = dmax(mx, vm / wm);
      mn = dmin(mn, vm / wm);
    }
    vdv = vk - vm;
  }
  const double rar = GV(rarea, ipn);
  const double thko = C2(b.dpinit, ipn, k), thkn = C2(s.dp, ipn, k);
  hd = hd * rar;
  const double q = fl * thko - hd - vdv;
  const double amount = dmax(dmax(0., mn * thkn), dmin(q, mx * thkn));
  const double fl1 = (fl * ONEMU + amount) / (ONEMU + thkn);
  TB(b.fmx, n, ipn, k) = mx;
  TB(b.fmn, n, ipn, k) = mn;
  TB(b.fld1, n, ipn, k) = fl1;
  TB(b.flp, n, ipn, k) = (mx - fl1) * thkn / (totin + EPSIL) * rar;
  TB(b.fln, n, ipn, k) = (fl1 - mn) * thkn / (totou + EPSIL) * rar;
}

// fct3d, last part (fct3d.F90:372-405): limited antidiffusive fluxes and the
// final update of the class-B tracer
__global__ void __launch_bounds__(256) k_fct3d_clip(DevGrid g, DevState s, DevClassB b) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, ipn = wave_col(), n = blockIdx.y;
  if (ipn >= nip || k >= nvl) return;
  const int np = GNP(ipn);
  const double lp = TB(b.flp, n, ipn, k), ln = TB(b.fln, n, ipn, k);
  double hd = 0.;
  for (int edg = 0; edg < np; edg++) {
    const int ix = PROX(ipn, edg);
    const double an = TE(b.uan, n, ipn, edg, k);
    const double clip = an >= 0. ? dmin(dmin(1., ln), TB(b.flp, n, ix, k)) : dmin(dmin(1., lp), TB(b.fln, n, ix, k));
    hd = hd + an * clip;
  }
  const double thkn = C2(s.dp, ipn, k);
  hd = hd * GV(rarea, ipn);
  const double fl1 = TB(b.fld1, n, ipn, k);
  const double q = fl1 * thkn - hd;
  const double amount = dmax(dmax(0., TB(b.fmn, n, ipn, k) * thkn), dmin(q, TB(b.fmx, n, ipn, k) * thkn));
  TB(b.trb, n, ipn, k) = (fl1 * ONEMU + amount) / (ONEMU + thkn);
}

inline dim3 col_grid(int nip, int z = 1) { return dim3((nip + COLS - 1) / COLS, z); }
inline dim3 col_block() { return dim3(NVL_MAX, COLS); }

}  // namespace

void launch_transp0(const DevGrid &g, const DevState &s, const DevClassB &b, hipStream_t st) {
  KT_LAUNCH("k_transp0", st, k_transp0<<<col_grid(g.nip), col_block(), 0, st>>>(g, s, b));
}
void launch_transp1(const DevGrid &g, const DevState &s, const DevClassB &b, const AB &ab, hipStream_t st) {
  KT_LAUNCH("k_transp1", st, k_transp1<<<col_grid(g.nip), col_block(), 0, st>>>(g, s, b, ab));
}
void launch_transp2(const DevGrid &g, const DevState &s, const DevClassB &b, hipStream_t st) {
  KT_LAUNCH("k_transp2_w", st, k_transp2_w<<<col_grid(g.nip), col_block(), 0, st>>>(g, s, b));
  if (b.ntrb < 1) return;
  if (g.nvl == NVL_MAX) {
    KT_LAUNCH("k_fct3d_fill", st, k_fct3d_fill<NVL_MAX><<<dim3((g.nip + 63) / 64, b.ntrb), 64, 0, st>>>(g, b));
    KT_LAUNCH("k_fct3d_vert", st, k_fct3d_vert<true><<<col_grid(g.nip, b.ntrb), col_block(), 0, st>>>(g, b));
  } else {
    KT_LAUNCH("k_fct3d_vert", st, k_fct3d_vert<false><<<col_grid(g.nip, b.ntrb), col_block(), 0, st>>>(g, b));
  }
  KT_LAUNCH("k_fct3d_lo", st, k_fct3d_lo<<<col_grid(g.nip, b.ntrb), col_block(), 0, st>>>(g, s, b));
  KT_LAUNCH("k_fct3d_clip", st, k_fct3d_clip<<<col_grid(g.nip, b.ntrb), col_block(), 0, st>>>(g, s, b));
}

}  // namespace fimhip
