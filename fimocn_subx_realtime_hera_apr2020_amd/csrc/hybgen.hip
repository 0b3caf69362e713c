// hybgen (FIMsrc/fim/horizontal/hybgen.F90) on the GPU, in FP64.
//
// regrid_1d (207-512) with restp_1d (792-890), equilb (1282-1437) and
// inflate (893-1007), and remap_1d (515-789) with pcwise (1010-1279), run one
// wavefront per column (lane = level) with the column in LDS; at 64 levels
// the inherently serial sweeps (equilb, inflate, the theta redistribution)
// run in separate kernels one thread per column with the column in
// registers.  The lateral interface smoothing dffusn_lev (dffusn.F90:37-58)
// is a cell-level kernel.  Loop orders and operation orders are those of the
// reference.  Columns a land/sea mask marks as land are left as they are.
#include "fimhip_internal.h"
#include "fim_pow.h"

#include <stdlib.h>
#include <string.h>

namespace fimhip {

// Column indices run over every column of every member (ncol = members * nip,
// column c = member * nip + cell); tracers are (nvl, ncol, ntra).
#define C2(a, col, k) (a)[(size_t)(col) * nvl + (k)]
#define L2(a, col, k) (a)[(size_t)(col) * nvlp1 + (k)]
#define TR(a, t, col, k) (a)[((size_t)(t) * ncol + (col)) * nvl + (k)]

constexpr int LMAX = NVL_MAX + 1;
constexpr double THIN = RD / P1000;

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

__device__ void restp_1d(const double *thold, const double *pkold, int kold, double *thnew, double *pknew,
                         const double *targt, int knew) {
  double oldth[NVL_MAX];
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

__device__ __forceinline__ double cushn(double arg, double arg1) {
  double m = dmin(arg, 2. * arg1);
  return .25 * (m * m) / arg1 + dmax(arg, 2. * arg1) - arg1;
}

__device__ void equilb(const double *thet, double *pnew, int kk, double slak, int k0 = 3) {
  for (int k = k0; k <= kk - 2; k++) {
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
}

// dffusn_lev(exsmo, smoo_coeff, nvlp1, kbl+2, nvl): reads the regridded Exner
// function from sdot and writes the smoothed one to exsmo
template <typename V>
__global__ void __launch_bounds__(256) k_dffusn_lev(HybCols<V> s) {
  const int nvl = s.nvl, nvlp1 = nvl + 1, nip = s.nip;
  const int col = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + threadIdx.y));
  if (col >= s.ncol) return;
  const int m = col / nip, ipn = col - m * nip;
  const int k1 = KBL + 2, k2 = nvl;
  // every neighbour load is issued before the sum (a pentagon's sixth slot
  // repeats its fifth and is not added)
  const int np = ldc(s.nprox, ipn);
  int nb[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) nb[e] = m * nip + ldc(s.prox, ipn * NPP + e);
  // lane kk < 64 smooths interface kk; the levels outside k1..k2 (the
  // bottom ones and the top interface, which lane 0 copies as well) are copied
  for (int kk = threadIdx.x; kk < nvlp1; kk += blockDim.x) {
    const double f = L2(s.sdot, col, kk);
    const bool top = kk == 0 && nvlp1 - 1 >= (int)blockDim.x;
    const double ftop = top ? L2(s.sdot, col, nvlp1 - 1) : 0.;
    double out = f;
    const int k = kk + 1;
    if (k >= k1 && k <= k2) {
      double v[NPP];
#pragma unroll
      for (int e = 0; e < NPP; e++) v[e] = L2(s.sdot, nb[e], kk);
      double flx = 0.;
#pragma unroll
      for (int e = 0; e < NPP; e++)
        if (e < np) flx = flx + (f - v[e]) * ldc(s.sideln, ipn * NPP + e);
      double factor = -s.smoo[kk] * ldc(s.rarea, ipn);
      out = f + flx * factor;
    }
    // nontemporal for the batched FP32 state (measured: 1.15 -> 1.07 ms at
    // 16 x G7; the same in the column kernels, whose threads store strided
    // rows, made them up to 3.7x slower)
    if constexpr (sizeof(V) < sizeof(double)) {
      __builtin_nontemporal_store(out, &L2(s.exsmo, col, kk));
      if (top) __builtin_nontemporal_store(ftop, &L2(s.exsmo, col, nvlp1 - 1));
    } else {
      L2(s.exsmo, col, kk) = out;
      if (top) L2(s.exsmo, col, nvlp1 - 1) = ftop;
    }
    if (kk + (int)blockDim.x == nvlp1 - 1) break;   // the top interface is done
  }
}

// ===========================================================================
// The same column algorithm with one wavefront per column (lane = level) and
// the column in LDS.  Every element is computed by the same operations in the
// same order as above; the loops of the reference are mapped as follows:
//  * element-wise loops (heat diffusion, slak, miglim, pow, the PLM slopes,
//    fluxes and outputs) run one level per lane;
//  * running minima (thcol, oldth, the final Exner monotonicity) are wavefront
//    scans, exact because min is;
//  * restp_1d: each target layer's integration range starts where the
//    previous one stopped, which for non-decreasing targets is found by a
//    binary search, so each lane integrates its own layer in the reference
//    order (a lane-0 sweep takes over for non-monotone targets);
//  * the excol fix and equilb change nothing unless their condition holds on
//    the unmodified column somewhere; lane 0 then runs the sweep from the
//    first such level;
//  * the two fill sweeps of pcwise: a cell's update ignores its neighbour
//    whenever neighbour*1e-22 is provably below half an ulp of the cell's own
//    term (bounded by the running maximum of the column); the remaining
//    chains are resolved neighbour by neighbour in sweep order;
//  * inflate, the top-of-column loop and the theta redistribution run on
//    lane 0.
constexpr int HW = 4;   // columns (wavefronts) per block

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ bool wany(bool p) { return __ballot(p) != 0ull; }
// a / b: exact division in the FP64 form; in the mixed-precision form the
// hardware reciprocal refined by two Newton steps (within an ulp or two)
template <bool FAST>
__device__ __forceinline__ double ddiv(double a, double b) {
  if (!FAST) return a / b;
  double r = __builtin_amdgcn_rcp(b);
  r = fma(r, fma(-b, r, 1.), r);
  r = fma(r, fma(-b, r, 1.), r);
  return a * r;
}
__device__ __forceinline__ double w_suffix_min(double x, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_down(x, d, 64);
    if (lane + d < 64) x = dmin(x, y);
  }
  return x;
}
__device__ __forceinline__ double w_prefix_min(double x, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(x, d, 64);
    if (lane >= d) x = dmin(x, y);
  }
  return x;
}
// Cross-lane moves on the VALU (DPP) instead of the LDS crossbar:
// w_up1(x) on lane j is lane j-1's x, w_dn1(x) lane j+1's (0 where that lane
// does not exist), w_max_all the column-wide maximum of x >= 0 on every lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int old, int x) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_z(double x) {   // lanes without a source get 0
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_d(double old, double x) {
  const unsigned long long o = (unsigned long long)__double_as_longlong(old);
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)dpp_i<CTRL, ROWS>((int)(unsigned)o, (int)(unsigned)b);
  const unsigned hi = (unsigned)dpp_i<CTRL, ROWS>((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_XOR1 = 0xb1, DPP_XOR2 = 0x4e, DPP_ROW_MIRROR = 0x140, DPP_HALF_MIRROR = 0x141;
constexpr int DPP_BCAST15 = 0x142, DPP_BCAST31 = 0x143;
__device__ __forceinline__ double w_up1(double x) { return dpp_z<DPP_WAVE_SHR1>(x); }
__device__ __forceinline__ double w_dn1(double x) { return dpp_z<DPP_WAVE_SHL1>(x); }
__device__ __forceinline__ int w_dn1(int x) { return __builtin_amdgcn_mov_dpp(x, DPP_WAVE_SHL1, 0xf, 0xf, true); }
__device__ __forceinline__ int w_up1(int x) { return __builtin_amdgcn_mov_dpp(x, DPP_WAVE_SHR1, 0xf, 0xf, true); }
__device__ __forceinline__ double w_max_all(double x) {
  x = dmax(x, dpp_d<DPP_XOR1>(x, x));          // quads
  x = dmax(x, dpp_d<DPP_XOR2>(x, x));
  x = dmax(x, dpp_d<DPP_HALF_MIRROR>(x, x));   // half rows
  x = dmax(x, dpp_d<DPP_ROW_MIRROR>(x, x));    // rows of 16
  x = dmax(x, dpp_d<DPP_BCAST15, 0xa>(x, x));  // row 1 <- row 0, row 3 <- row 2
  x = dmax(x, dpp_d<DPP_BCAST31, 0xc>(x, x));  // rows 2, 3 <- row 1
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float w_up1(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), DPP_WAVE_SHR1, 0xf, 0xf, true));
}
__device__ __forceinline__ float w_dn1(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), DPP_WAVE_SHL1, 0xf, 0xf, true));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_fmax(float v) {
  return fmaxf(v, __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(v), __float_as_int(v))));
}
__device__ __forceinline__ float w_max_all(float x) {
  x = dpp_fmax<DPP_XOR1, 0xf>(x);
  x = dpp_fmax<DPP_XOR2, 0xf>(x);
  x = dpp_fmax<DPP_HALF_MIRROR, 0xf>(x);
  x = dpp_fmax<DPP_ROW_MIRROR, 0xf>(x);
  x = dpp_fmax<DPP_BCAST15, 0xa>(x);
  x = dpp_fmax<DPP_BCAST31, 0xc>(x);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}
__device__ __forceinline__ int w_min_int(int x) {
  for (int d = 1; d < 64; d <<= 1) x = min(x, __shfl_xor(x, d, 64));
  return x;
}

// inflate (hybgen.F90:893-1007) for one column on lane 0, the per-level
// constants prepared beforehand by all lanes: dp0b[k-1] = rd*dpsig(k)/p1000*
// shrink (k = 1..nvl-1) and tap[k] = 1/(1+tq**2), tq = (nvl-k)/4 (k = 2..nvl),
// each formed by the same operations as in inflate().
__device__ void inflate_w(const double *dp0b, const double *tap, int nvl, double extop, double *exner) {
  double dpsum = 0.;
  int last = 1;
  for (int k = 1; k <= nvl - 1; k++) {
    double dp0 = dp0b[k - 1];
    dpsum = dpsum + dp0;
    const double exold = exner[k];
    if (k <= KBL) {
      last = k + 1;
      exner[k] = exner[k - 1] - dp0;
    } else {
      if (k > last) {
        const double p2 = ldexp(1., k - last);
        dp0 = dp0 / dmin(10., 1.25 * p2);
        dp0 = dmax(dp0, RD * 100. / P1000);
      }
      const double dp1 = cushn(dmax(0., exner[k - 1] - dmin(exold, exner[0] - dpsum)), dp0);
      if (exner[k - 1] - dp1 < exold - THIN) {
        if (k == last) last = k + 1;
        exner[k] = dmax(extop, exner[k - 1] - dp1);
      }
    }
  }
  for (int k = nvl; k >= 2; k--) {
    double q = exner[k - 1] / CP;
    q = q * q * sqrt(q);
    const double dp0 = RD * THKTOP / P1000 * tap[k] / q;
    const double exold = exner[k - 1];
    const double dp1 = cushn(dmax(0., exold - exner[k]), dp0);
    if (exner[k] + dp1 > exold + THIN) exner[k - 1] = exner[k] + dp1;
    if (exner[k - 1] < exner[k - 2] - dp0) break;
  }
}

// all lanes: the per-level constants of inflate_w for surface pressure psurf
__device__ __forceinline__ void inflate_prep(const double *dpsig, int nvl, double psurf, double *dp0b, double *tap,
                                             int lane) {
  const double top = .2;
  const double shrink = dmin(1., (psurf / P1000 - top) / (1. - top));
  for (int k = 1 + lane; k <= nvl; k += 64) {
    if (k <= nvl - 1) dp0b[k - 1] = RD * ldc(dpsig, k - 1) / P1000 * shrink;
    if (k >= 2) {
      const double tq = .25 * (double)(nvl - k);
      tap[k] = 1. / (1. + tq * tq);
    }
  }
}

// true when y*1e-22 cannot move t, for every |y| <= bound (|y*1e-22| < |t|*2^-55,
// with margin for the growth of the bound along a sweep)
__device__ __forceinline__ bool absorbs(double t, double bound) {
  const double at = fabs(t);
  return at > 1.e-280 && bound * (1. + 1.e-12) * (1.e-22 * 36028797018963968.) < at;
}

// The interface geometry of pcwise for one column, lane j holding layer j:
// the old and new layer widths (fco, fcn), the width of the layer below
// (fco(j-1)) and the displacement of interface j.  xold/xnew (LDS) are the
// interface coordinates; the reference works in the direction in which they
// increase (zold/znew = +-x).
struct PcwGeom {
  double f, fl, fn, dz;
  double rf, rfn, dx;   // FAST only: 1/(f+1e-22), 1/(fn+1e-22), the flux slab fraction
};
template <bool FAST>
__device__ __forceinline__ PcwGeom pcwise_geom(int kk, const double *xold, const double *xnew, int lane) {
  const bool up = xold[0] < xold[kk];
  auto zo = [&](int k) { return up ? xold[k] : -xold[k]; };
  auto zn = [&](int k) { return up ? xnew[k] : -xnew[k]; };
  PcwGeom g{0., 0., 0., 0., 0., 0., 0.};
  if (lane < kk) {
    g.f = dmax(0., zo(lane + 1) - zo(lane));
    g.fn = dmax(0., zn(lane + 1) - zn(lane));
    g.dz = zn(lane) - zo(lane);
  }
  g.fl = w_up1(g.f);
  if (FAST) {
    g.rf = 1. / (g.f + 1.e-22);
    g.rfn = 1. / (g.fn + 1.e-22);
    const double fc = g.dz > 0. ? g.f : g.fl;
    const double adz = fabs(g.dz);
    g.dx = fc > 0. ? dmin(fc, adz) / fc : 0.;
  }
  return g;
}

// pcwise for one column: lane j < kk holds layer j's value a; returns the
// remapped value of layer j.  FAST (FP32 state) replaces the divisions by
// multiplications with reciprocals from the geometry, which changes results by
// rounding only; otherwise every operation is the reference's.
template <bool FAST>
__device__ double pcwise_w(int kk, const PcwGeom &gm, double a, int lane) {
  const double small = 1.e-22;
  const bool live = lane < kk;
  const double f = gm.f;
  // a bound on every value either fill sweep can produce: the largest
  // magnitude in the column (the sweeps form convex combinations)
  double bnd = live ? fabs(a) : 0.;
  bnd = w_max_all(bnd);
  // sweep 1, from the top: layer j (j <= kk-2) takes layer j+1's new value
  double v;
  {
    bool fin;
    if (lane >= kk - 1) {
      v = a;
      fin = true;
    } else {
      const double t = a * f;
      fin = absorbs(t, bnd);
      v = fin ? (FAST ? t * gm.rf : t / (f + small)) : a;
    }
    while (wany(!fin)) {
      const double vu = w_dn1(v);
      const int fu = w_dn1((int)fin);
      if (!fin && fu) {
        v = FAST ? (a * f + vu * small) * gm.rf : (a * f + vu * small) / (f + small);
        fin = true;
      }
    }
  }
  // sweep 2, from the bottom: layer j (j >= 1) takes layer j-1's new value
  double w;
  {
    const double a2 = v;
    bool fin;
    if (lane == 0 || !live) {
      w = a2;
      fin = true;
    } else {
      const double t = a2 * f;
      fin = absorbs(t, bnd);
      w = fin ? (FAST ? t * gm.rf : t / (f + small)) : a2;
    }
    while (wany(!fin)) {
      const double wd = w_up1(w);
      const int fd = w_up1((int)fin);
      if (!fin && fd) {
        w = FAST ? (a2 * f + wd * small) * gm.rf : (a2 * f + wd * small) / (f + small);
        fin = true;
      }
    }
  }
  // PLM slope of layer lane (1-based k = lane + 1)
  const double fm = w_up1(w), fp = w_dn1(w);
  double b = 0.;
  if (live && lane != 0 && lane != kk - 1) {
    if (!(w <= dmin(fm, fp) || w >= dmax(fm, fp))) {
      if ((fp - fm) * (fm + fp - 2. * w) > 0.) b = w - fm;
      else b = fp - w;
    }
  }
  // flux through interface lane (1-based vertfx(lane + 1)), lanes 1..kk-1
  double vf = 0.;
  {
    const double wl = w_up1(w), bl = w_up1(b);
    if (FAST) {
      // the slab is one layer's part, so amount/slab is the mean over it
      if (live && lane >= 1) {
        const double dz = gm.dz, dx = gm.dx;
        if (dz > 0. && gm.f > 0.) vf = -dz * (w + b * .5 * (dx - 1.));
        else if (dz < 0. && gm.fl > 0.) vf = -dz * (wl + bl * .5 * (1. - dx));
      }
    } else if (live && lane >= 1) {
      double slab = 0., amount = 0.;
      const double dz = gm.dz;
      if (dz > 0.) {
        const double fc = f;                       // fco(layer lane)
        if (fc > 0.) {
          double dslab = dmin(slab + fc, dz) - dmin(slab, dz);
          double dx = dslab / fc;
          double fcdx = w + b * .5 * (dx - 1.) + 0. * (.25 - dx * (.5 - dx * (1. / 3.)));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
      } else if (dz < 0.) {
        const double fc = gm.fl;                   // fco(layer lane - 1)
        if (fc > 0.) {
          double dslab = dmin(slab + fc, -dz) - dmin(slab, -dz);
          double dx = dslab / fc;
          double fcdx = wl + bl * .5 * (1. - dx) + 0. * (.25 - dx * (.5 - dx * (1. / 3.)));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
      }
      if (slab != 0.) vf = -dz * amount / slab;
    }
  }
  double vhi = w_dn1(vf);
  if (lane >= kk - 1) vhi = 0.;
  const double vlo = vf;
  double y = 0.;
  if (live) {
    const double amount = w * f - (vhi - vlo);
    y = FAST ? (amount + w * small) * gm.rfn : (amount + w * small) / (gm.fn + small);
  }
  return y;
}

// pcwise in FP32 for a field held in FP32 (the mixed-precision form's winds
// and tracers 2..ntra): the geometry is formed in FP64 and rounded once, and
// the sweeps, slopes and fluxes are those of pcwise_w<true> in FP32.  A
// lane's sweep value is final once the carried-in term y*1e-22 is below a
// quarter ulp of a*f (|y| <= the column maximum), as in absorbs() with
// FP32's precision.
struct PcwGeomF {
  float f, fl, fn, dz, rf, rfn, dx;
};
__device__ __forceinline__ PcwGeomF pcwise_geom_f(const PcwGeom &g) {
  return PcwGeomF{(float)g.f, (float)g.fl, (float)g.fn, (float)g.dz, (float)g.rf, (float)g.rfn, (float)g.dx};
}
__device__ __forceinline__ bool absorbs_f(float t, float bound) {
  const float at = fabsf(t);
  return at > 1.e-30f && bound * (1.f + 1.e-5f) * (1.e-22f * 67108864.f) < at;   // 2^26
}
__device__ float pcwise_wf(int kk, const PcwGeomF &gm, float a, int lane) {
#pragma clang fp contract(fast)   // FP32 state: fused multiply-adds
  const float small = 1.e-22f;
  const bool live = lane < kk;
  const float f = gm.f;
  const float bnd = w_max_all(live ? fabsf(a) : 0.f);
  float v;
  {
    bool fin;
    if (lane >= kk - 1) {
      v = a;
      fin = true;
    } else {
      const float t = a * f;
      fin = absorbs_f(t, bnd);
      v = fin ? t * gm.rf : a;
    }
    while (wany(!fin)) {
      const float vu = w_dn1(v);
      const int fu = w_dn1((int)fin);
      if (!fin && fu) {
        v = (a * f + vu * small) * gm.rf;
        fin = true;
      }
    }
  }
  float w;
  {
    const float a2 = v;
    bool fin;
    if (lane == 0 || !live) {
      w = a2;
      fin = true;
    } else {
      const float t = a2 * f;
      fin = absorbs_f(t, bnd);
      w = fin ? t * gm.rf : a2;
    }
    while (wany(!fin)) {
      const float wd = w_up1(w);
      const int fd = w_up1((int)fin);
      if (!fin && fd) {
        w = (a2 * f + wd * small) * gm.rf;
        fin = true;
      }
    }
  }
  const float fm = w_up1(w), fp = w_dn1(w);
  float b = 0.f;
  if (live && lane != 0 && lane != kk - 1) {
    if (!(w <= fminf(fm, fp) || w >= fmaxf(fm, fp))) {
      if ((fp - fm) * (fm + fp - 2.f * w) > 0.f) b = w - fm;
      else b = fp - w;
    }
  }
  float vf = 0.f;
  {
    const float wl = w_up1(w), bl = w_up1(b);
    if (live && lane >= 1) {
      const float dz = gm.dz, dx = gm.dx;
      if (dz > 0.f && gm.f > 0.f) vf = -dz * (w + b * .5f * (dx - 1.f));
      else if (dz < 0.f && gm.fl > 0.f) vf = -dz * (wl + bl * .5f * (1.f - dx));
    }
  }
  float vhi = w_dn1(vf);
  if (lane >= kk - 1) vhi = 0.f;
  float y = 0.f;
  if (live) y = (w * f - (vhi - vf) + w * small) * gm.rfn;
  return y;
}

// regrid_1d (hybgen.F90:207-512), one wavefront per column
// With SPLIT the column-serial equilb and inflate are left to
// k_regrid_tail, which runs them one thread per column with the column in
// registers; the Exner function after the heat diffusion goes to sdot.
template <typename V, bool SPLIT>
__global__ void __launch_bounds__(64 * HW) k_regrid_w(HybCols<V> s, const StepCtl *ctl, int its) {
  __shared__ double sh[HW][7][LMAX];
  const int ncol = s.ncol, nvl = s.nvl, nvlp1 = nvl + 1;
  const int lane = threadIdx.x, c = threadIdx.y;
  const int col = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * HW + c));
  if (col >= ncol) return;
  if (s.wet && !ldc(s.wet, col)) {     // land: no regrid; its Exner function for the neighbours' smoothing
    for (int k = lane; k < nvlp1; k += 64) L2(s.sdot, col, k) = L2(s.ex, col, k);
    return;
  }
  if (ctl) its = load_ctl(ctl).its;
  double *thcol = sh[c][0], *excol = sh[c][1], *exwrk = sh[c][2], *exnew = sh[c][3], *thnew = sh[c][4],
         *heatfx = sh[c][5], *oldth = sh[c][6];
  const double *targt = s.thetac;
  const double tolrnce = 0.001;
  for (int k = lane; k < nvl; k += 64) thcol[k] = (double)TR(s.tr, 0, col, k);
  for (int k = lane; k < nvlp1; k += 64) excol[k] = L2(s.ex, col, k);
  const double psurf = L2(s.pr, col, 0);
  const double eqlb_slak = .1 * (double)its / (10. + (double)its);
  wsync();
  {
    bool t = false;
    for (int k = 1 + lane; k <= nvl; k += 64) t = t || (excol[k - 1] < excol[k] - THIN);
    if (wany(t)) {
      if (lane == 0)
        for (int k = 1; k <= nvl; k++)
          if (excol[k - 1] < excol[k] - THIN) excol[k] = excol[k - 1];
      wsync();
    }
  }
  {
    const double m = w_suffix_min(lane < nvl ? thcol[lane] : 1.e300, lane);
    if (lane < nvl) thcol[lane] = m;
  }
  for (int k = lane; k < nvlp1; k += 64) exwrk[k] = excol[k];
  wsync();
  const double ttop = ldc(targt, nvl - 1);
  if (lane == 0) {
    for (int k = nvl; k >= 2; k--) {
      if (thcol[k - 1] <= ttop + tolrnce) break;
      double wgt1 = dmax(0., exwrk[k - 2] - exwrk[k - 1]);
      double wgt2 = dmax(THIN, exwrk[k - 1] - exwrk[k]);
      double try_ = (thcol[k - 2] * wgt1 + thcol[k - 1] * wgt2) / (wgt1 + wgt2);
      if (try_ > ttop - tolrnce) {
        exwrk[k - 1] = exwrk[nvlp1 - 1];
        for (int k1 = k - 1; k1 <= nvl; k1++) thcol[k1 - 1] = try_;
      } else {
        exwrk[k - 1] = dmin(exwrk[k - 2], dmax(exwrk[k], (exwrk[k - 1] * (thcol[k - 1] - thcol[k - 2]) +
                                                         exwrk[k] * (ttop - thcol[k - 1])) /
                                                            (ttop - thcol[k - 2])));
        for (int k1 = k; k1 <= nvl; k1++) thcol[k1 - 1] = ttop;
        break;
      }
    }
  }
  wsync();
  // restp_1d(thcol, exwrk, nvl, thnew, exnew, targt, nvl)
  {
    const int kold = nvl, knew = nvl;
    const double m = w_suffix_min(lane < kold ? thcol[lane] : 1.e300, lane);
    if (lane < kold) oldth[lane] = m;
    wsync();
    const double o0 = oldth[0], on = oldth[kold - 1];
    bool nonmono = false;
    for (int k = lane; k < knew; k += 64) {
      double t = ldc(targt, k);
      if (k == 0) t = dmin(dmin(o0, on), ldc(targt, 0));
      if (k == knew - 1) t = dmax(dmax(o0, on), ldc(targt, knew - 1));
      thnew[k] = t;
    }
    wsync();
    for (int k = 1 + lane; k < knew; k += 64) nonmono = nonmono || (thnew[k] < thnew[k - 1]);
    if (wany(nonmono)) {
      if (lane == 0) restp_1d(thcol, exwrk, kold, thnew, exnew, targt, knew);
      wsync();
    } else {
      if (lane == 0) {
        exnew[0] = exwrk[0];
        exnew[knew] = exwrk[kold];
      }
      bool copy = false;
      for (int k = 1 + lane; k <= knew - 1; k += 64) {
        const double tk = thnew[k], tkm = thnew[k - 1];
        int ko = 1;
        if (k >= 2) {            // where layer k-1's integration stopped
          int lo = 0, hi = kold;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (oldth[mid] < tkm) lo = mid + 1;
            else hi = mid;
          }
          ko = lo + 1 < kold ? lo + 1 : kold;
        }
        double pinteg = 0., tha, thb = tkm;
        for (;;) {
          tha = thb;
          thb = dmin(tk, dmax(tkm, oldth[ko - 1]));
          pinteg = pinteg + exwrk[ko - 1] * (thb - tha);
          if (oldth[ko - 1] < tk) {
            if (ko < kold) {
              ko = ko + 1;
              continue;
            }
            tha = thb;
            thb = tk;
            pinteg = pinteg + exwrk[kold] * (thb - tha);
          }
          break;
        }
        if (tk > tkm) exnew[k] = ddiv<sizeof(V) < sizeof(double)>(pinteg, tk - tkm);
        else copy = true;
      }
      if (wany(copy)) {
        wsync();
        if (lane == 0)
          for (int k = 1; k <= knew - 1; k++)
            if (!(thnew[k] > thnew[k - 1])) exnew[k] = exnew[k - 1];
      }
      wsync();
    }
  }
  // heat diffusion, up to 3 iterations
  if (lane == 0) {
    heatfx[nvl - 1] = 0.;
    heatfx[0] = 0.;
  }
  for (int iter = 1; iter <= 3; iter++) {
    for (int k = lane; k < nvlp1; k += 64) exwrk[k] = exnew[k];
    wsync();
    for (int k = 2 + lane; k <= nvl - 1; k += 64) {
      double h = 0.;
      if (exwrk[k - 1] < exwrk[0] - .01 && exwrk[k] > exwrk[nvlp1 - 1] + .01) {
        constexpr bool F = sizeof(V) < sizeof(double);
        double cnv = ddiv<F>(GRVITY, ldc(targt, k - 1));
        h = ddiv<F>(.1 * cnv * cnv * .5 * (ldc(targt, k) - ldc(targt, k - 2)), dmax(.03 * cnv, exwrk[k - 1] - exwrk[k]));
      }
      heatfx[k - 1] = h;
    }
    wsync();
    for (int k = 1 + lane; k <= nvl - 1; k += 64) {
      if (exwrk[k] < exwrk[0] - .01 && exwrk[k] > exwrk[nvlp1 - 1] + .01)
        exnew[k] = dmax(exwrk[k + 1], dmin(exwrk[k - 1], exwrk[k] + ddiv<sizeof(V) < sizeof(double)>(
                                                                       heatfx[k] - heatfx[k - 1],
                                                                       ldc(targt, k) - ldc(targt, k - 1))));
    }
    wsync();
    bool ev = false;
    for (int k = 2 + lane; k <= nvlp1; k += 64) ev = ev || (exnew[k - 1] > exnew[k - 2] + THIN);
    if (!wany(ev)) break;
  }
  if (SPLIT) {
    for (int k = lane; k < nvlp1; k += 64) L2(s.sdot, col, k) = exnew[k];
    return;
  }
  // equilb from the first level where it would move an interface
  {
    int first = 1 << 30;
    for (int k = 3 + lane; k <= nvl - 2; k += 64) {
      if (exnew[k - 3] > exnew[0] - THIN) continue;
      const double dp1 = exnew[k - 3] - exnew[k - 2], dp2 = exnew[k - 2] - exnew[k - 1];
      const double dp3 = exnew[k - 1] - exnew[k], dp4 = exnew[k] - exnew[k + 1], dp5 = exnew[k + 1] - exnew[k + 2];
      if (dp2 > dp1 && dp4 > dp5) {
        const double goal = .5 * (dp3 + dmin(dp2, dp4));
        if (dp2 > goal && dp4 > goal) first = min(first, k);
      }
    }
    first = w_min_int(first);
    if (first <= nvl - 2) {
      if (lane == 0) equilb(thnew, exnew, nvl, eqlb_slak, first);
      wsync();
    }
  }
  inflate_prep(s.dpsig, nvl, psurf, heatfx, oldth, lane);
  wsync();
  if (lane == 0) inflate_w(heatfx, oldth, nvl, CP * fim_pow(PTOP / P1000, RD / CP), exnew);
  wsync();
  for (int k = lane; k < nvlp1; k += 64) L2(s.sdot, col, k) = exnew[k];   // pre-smoothing Exner, kept in sdot
}

// remap_1d (hybgen.F90:515-789), one wavefront per column
// With SPLIT, inflate has been applied to exsmo by k_remap_inflate, and the
// theta redistribution is left to k_remap_theta: the remapped theta goes to
// thbuf and the layer depths in p**(kappa+1), max(pk1new(k)-pk1new(k+1),
// pk1new(1)*1e-6), to exsmo.
template <typename V, bool SPLIT>
__global__ void __launch_bounds__(64 * HW) k_remap_w(HybCols<V> s, const StepCtl *ctl, int its) {
  __shared__ double sh[HW][7][LMAX];
  const int ncol = s.ncol, nvl = s.nvl, nvlp1 = nvl + 1, ntra = s.ntra;
  const int lane = threadIdx.x, c = threadIdx.y;
  const int col = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * HW + c));
  if (col >= ncol) return;
  if (s.wet && !ldc(s.wet, col)) {    // land: no remap
    for (int k = lane; k < nvlp1; k += 64) L2(s.sdot, col, k) = 0.;
    return;
  }
  if (ctl) its = load_ctl(ctl).its;
  double *prcol = sh[c][0], *pk1col = sh[c][1], *exnew = sh[c][2], *prnew = sh[c][3], *pk1new = sh[c][4],
         *th1 = sh[c][5], *exold = sh[c][6];
  const double *targt = s.thetac;
  const double small = 1.e-6;
  for (int k = lane; k < nvlp1; k += 64) {
    const double eo = L2(s.ex, col, k);
    const double p = L2(s.pr, col, k);
    exold[k] = eo;
    prcol[k] = p;
    exnew[k] = L2(s.exsmo, col, k);
    pk1col[k] = eo * p / CP;
  }
  if (!SPLIT) {
    inflate_prep(s.dpsig, nvl, L2(s.pr, col, 0), prnew, pk1new, lane);
    wsync();
    if (lane == 0) inflate_w(prnew, pk1new, nvl, CP * fim_pow(PTOP / P1000, RD / CP), exnew);
  }
  wsync();
  for (int k = lane; k < nvlp1; k += 64) {
    double e = exnew[k];
    if (its > 0) e = SLAK * e + (1. - SLAK) * exold[k];
    if (k >= 1 && k <= nvlp1 - 2) e = dmax(.0001 * exold[k] + .9999 * exold[k + 1], dmin(.0001 * exold[k] + .9999 * exold[k - 1], e));
    exnew[k] = e;
    // FP32 state: the plain-double pow (a few ulp) instead of the correctly rounded one
    const double pn = P1000 * (sizeof(V) < sizeof(double) ? fim_pow_fast(e / CP, CP / RD) : fim_pow(e / CP, CP / RD));
    prnew[k] = pn;
    pk1new[k] = e * pn / CP;
  }
  wsync();
  for (int k = lane; k < nvl; k += 64) C2(s.dp, col, k) = prnew[k] - prnew[k + 1];
  wsync();
  if (lane == 0) {
    prnew[nvlp1 - 1] = prcol[nvlp1 - 1];
    pk1new[nvlp1 - 1] = pk1col[nvlp1 - 1];
    prnew[0] = prcol[0];
    pk1new[0] = pk1col[0];
  }
  wsync();
  // u, v and tracers 2..ntra are remapped in pressure
  {
    const bool live = lane < nvl;
    constexpr bool FAST = sizeof(V) < sizeof(double);
    const PcwGeom gp = pcwise_geom<FAST>(nvl, prcol, prnew, lane);
    const double th0 = live ? (double)TR(s.tr, 0, col, lane) : 0.;   // loaded before the stores below
    double y;
    if (FAST) {   // FP32 state: FP32 remap
      const PcwGeomF gf = pcwise_geom_f(gp);
      // the inputs of four fields are loaded before any result is stored
      const int kl = live ? lane : 0;
      const float u0 = (float)C2(s.us_in, col, kl), v0 = (float)C2(s.vs_in, col, kl);
      float yf = pcwise_wf(nvl, gf, live ? u0 : 0.f, lane);
      const float yv = pcwise_wf(nvl, gf, live ? v0 : 0.f, lane);
      if (live) {
        C2(s.us, col, lane) = (V)yf;
        C2(s.vs, col, lane) = (V)yv;
      }
      for (int n0 = 1; n0 < ntra; n0 += 4) {
        float x[4];
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = (float)TR(s.tr, min(n0 + j, ntra - 1), col, kl);
        float y4[4];
#pragma unroll
        for (int j = 0; j < 4; j++) y4[j] = n0 + j < ntra ? pcwise_wf(nvl, gf, live ? x[j] : 0.f, lane) : 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (live && n0 + j < ntra) TR(s.tr, n0 + j, col, lane) = (V)y4[j];
      }
    } else {
      y = pcwise_w<FAST>(nvl, gp, live ? (double)C2(s.us_in, col, lane) : 0., lane);
      if (live) C2(s.us, col, lane) = (V)y;
      y = pcwise_w<FAST>(nvl, gp, live ? (double)C2(s.vs_in, col, lane) : 0., lane);
      if (live) C2(s.vs, col, lane) = (V)y;
      for (int n = 1; n < ntra; n++) {
        y = pcwise_w<FAST>(nvl, gp, live ? (double)TR(s.tr, n, col, lane) : 0., lane);
        if (live) TR(s.tr, n, col, lane) = (V)y;
      }
    }
    // theta is remapped in p^(kappa+1)
    y = pcwise_w<FAST>(nvl, pcwise_geom<FAST>(nvl, pk1col, pk1new, lane), th0, lane);
    if (live) th1[lane] = y;
  }
  wsync();
  if (SPLIT) {
    for (int k = lane; k < nvl; k += 64) {
      C2(s.thbuf, col, k) = th1[k];
      L2(s.exsmo, col, k) = dmax(pk1new[k] - pk1new[k + 1], pk1new[0] * small);
    }
  } else if (lane == 0) {
    for (int k = nvl; k >= 3; k--) {
      double dplo = dmax(pk1new[k - 2] - pk1new[k - 1], pk1new[0] * small);
      double dpup = dmax(pk1new[k - 1] - pk1new[k], pk1new[0] * small);
      double tha = th1[k - 2], thb = th1[k - 1];
      double devilo = (tha - ldc(targt, k - 2)) * dplo;
      double deviup = (thb - ldc(targt, k - 1)) * dpup;
      if (deviup > 0. && devilo < 0.) {
        th1[k - 2] = tha + dmin(deviup, -devilo) / dplo;
        th1[k - 1] = thb - dmin(deviup, -devilo) / dpup;
      } else if (deviup < 0. && devilo > 0.) {
        th1[k - 2] = tha - dmin(devilo, -deviup) / dplo;
        th1[k - 1] = thb + dmin(devilo, -deviup) / dpup;
      }
    }
  }
  wsync();
  if (!SPLIT && lane < nvl) TR(s.tr, 0, col, lane) = (V)th1[lane];
  // exnew(k) = min over exnew(0..k): the reference's top-down monotonicity fix
  {
    const double m = w_prefix_min(lane < nvlp1 ? exnew[lane] : 1.e300, lane);
    double mtop = 0.;
    if (nvlp1 > 64) mtop = dmin(__shfl(m, 63, 64), exnew[64]);
    wsync();
    if (lane < nvlp1) exnew[lane] = m;
    if (nvlp1 > 64 && lane == 0) exnew[64] = mtop;
  }
  wsync();
  for (int k = lane; k < nvlp1; k += 64) {
    L2(s.pr, col, k) = prnew[k];
    L2(s.sdot, col, k) = exnew[k] - exold[k];
    L2(s.ex, col, k) = exnew[k];
  }
}


// ===========================================================================
// The column-serial sweeps of hybgen, one thread per column with the column
// in registers (NVL levels, every loop unrolled so each array index is a
// compile-time constant), for the split form of the wave kernels above.  The
// operations are those of equilb, inflate and the theta redistribution of
// remap_1d, in the same order.
template <int NVL>
__device__ __forceinline__ void equilb_r(const double *targt, double *pnew, double slak) {
#pragma unroll
  for (int k = 3; k <= NVL - 2; k++) {
    if (pnew[k - 3] > pnew[0] - THIN) continue;
    const double dp1 = pnew[k - 3] - pnew[k - 2];
    const double dp2 = pnew[k - 2] - pnew[k - 1];
    const double dp3 = pnew[k - 1] - pnew[k];
    const double dp4 = pnew[k] - pnew[k + 1];
    const double dp5 = pnew[k + 1] - pnew[k + 2];
    const double th2 = ldc(targt, k - 2), th3 = ldc(targt, k - 1), th4 = ldc(targt, k);
    if (dp2 > dp1 && dp4 > dp5) {
      const double goal = .5 * (dp3 + dmin(dp2, dp4));
      if (dp2 > goal && dp4 > goal) {
        double dis3 = dmin(dp2 - goal, goal - dp3) * slak;
        double dis4 = dmin(dp4 - goal, goal - dp3) * slak;
        if (th3 > th2 && th4 > th3) {
          const double ratio = (th4 - th3) / (th3 - th2);
          if (ratio > 1.) dis4 = dis3 / ratio;
          else dis3 = dis4 * ratio;
        }
        pnew[k - 1] = pnew[k - 1] + dis3;
        pnew[k] = pnew[k] - dis4;
      }
    }
  }
}

template <int NVL>
__device__ __forceinline__ void inflate_r(const double *dpsig, double psurf, double extop, double *exner) {
  double dpsum = 0.;
  const double top = .2;
  const double shrink = dmin(1., (psurf / P1000 - top) / (1. - top));
  int last = 1;
#pragma unroll
  for (int k = 1; k <= NVL - 1; k++) {
    double dp0 = RD * ldc(dpsig, k - 1) / P1000 * shrink;
    dpsum = dpsum + dp0;
    const double exold = exner[k];
    if (k <= KBL) {
      last = k + 1;
      exner[k] = exner[k - 1] - dp0;
    } else {
      if (k > last) {
        const double p2 = ldexp(1., k - last);
        dp0 = dp0 / dmin(10., 1.25 * p2);
        dp0 = dmax(dp0, RD * 100. / P1000);
      }
      const double dp1 = cushn(dmax(0., exner[k - 1] - dmin(exold, exner[0] - dpsum)), dp0);
      if (exner[k - 1] - dp1 < exold - THIN) {
        if (k == last) last = k + 1;
        exner[k] = dmax(extop, exner[k - 1] - dp1);
      }
    }
  }
#pragma unroll
  for (int k = NVL; k >= 2; k--) {
    double q = exner[k - 1] / CP;
    q = q * q * sqrt(q);
    const double tq = .25 * (double)(NVL - k);
    const double tap = 1. / (1. + tq * tq);
    const double dp0 = RD * THKTOP / P1000 * tap / q;
    const double exold = exner[k - 1];
    const double dp1 = cushn(dmax(0., exold - exner[k]), dp0);
    if (exner[k] + dp1 > exold + THIN) exner[k - 1] = exner[k] + dp1;
    if (exner[k - 1] < exner[k - 2] - dp0) break;
  }
}

// equilb + inflate after the wave part of regrid_1d (Exner function in sdot)
template <int NVL, typename V>
__global__ void __launch_bounds__(64) k_regrid_tail(HybCols<V> s, const StepCtl *ctl, int its, double extop) {
  const int nvl = NVL, nvlp1 = NVL + 1;
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= s.ncol || (s.wet && !s.wet[col])) return;
  if (ctl) its = load_ctl(ctl).its;
  double ex[NVL + 1];
#pragma unroll
  for (int k = 0; k <= NVL; k++) ex[k] = L2(s.sdot, col, k);
  const double psurf = L2(s.pr, col, 0);
  equilb_r<NVL>(s.thetac, ex, .1 * (double)its / (10. + (double)its));
  inflate_r<NVL>(s.dpsig, psurf, extop, ex);
#pragma unroll
  for (int k = 0; k <= NVL; k++) L2(s.sdot, col, k) = ex[k];
  (void)nvl;
}

// inflate of the smoothed Exner function at the start of remap_1d (in place)
template <int NVL, typename V>
__global__ void __launch_bounds__(64) k_remap_inflate(HybCols<V> s, double extop) {
  const int nvlp1 = NVL + 1;
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= s.ncol || (s.wet && !s.wet[col])) return;
  double ex[NVL + 1];
#pragma unroll
  for (int k = 0; k <= NVL; k++) ex[k] = L2(s.exsmo, col, k);
  inflate_r<NVL>(s.dpsig, L2(s.pr, col, 0), extop, ex);
#pragma unroll
  for (int k = 0; k <= NVL; k++) L2(s.exsmo, col, k) = ex[k];
}

// one exchange of the theta redistribution of remap_1d (hybgen.F90:515-789,
// final loop): layers p (tha, depth dplo) and p+1 (thb, dpup)
__device__ __forceinline__ void theta_pair(double &tha, double &thb, double dplo, double dpup, double tlo,
                                           double tup) {
  const double devilo = (tha - tlo) * dplo;
  const double deviup = (thb - tup) * dpup;
  if (deviup > 0. && devilo < 0.) {
    tha = tha + dmin(deviup, -devilo) / dplo;
    thb = thb - dmin(deviup, -devilo) / dpup;
  } else if (deviup < 0. && devilo > 0.) {
    tha = tha - dmin(devilo, -deviup) / dplo;
    thb = thb + dmin(devilo, -deviup) / dpup;
  }
}

// theta redistribution of remap_1d (final loop over k = nvl..3, i.e. pairs
// p = k-2 = nvl-2 .. 1 in that order), one thread per column for 64 columns,
// from the remapped theta in thbuf and the layer depths left in exsmo.  The
// inputs are staged through LDS in chunks of 16 levels (each load instruction
// reads 16 levels of 4 columns), the sweep works on one chunk in registers and
// carries the upper layer of the pair that straddles two chunks, and the
// result goes back to column-major tracer 1 through an LDS tile, so every
// global access is coalesced.
template <int NVL, typename V>
__global__ void __launch_bounds__(64) k_remap_theta(HybCols<V> s) {
  constexpr int CH = 16, NCH = NVL / CH;
  static_assert(NVL % CH == 0, "whole chunks");
  __shared__ double sth[CH][65], sdl[CH][65];
  __shared__ V sout[CH + 1][65];   // levels k0 .. k0+CH of the chunk being finished
  const int nvl = NVL, nvlp1 = NVL + 1, ncol = s.ncol;
  const int lane = threadIdx.x, c0 = blockIdx.x * 64;
  const int sl = lane & 15, sj = lane >> 4;
  double cth = 0., cdl = 0.;   // level k0+CH (the chunk above): theta still to be finalised, layer depth
#pragma unroll 1
  for (int ch = NCH - 1; ch >= 0; ch--) {
    const int k0 = ch * CH;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int c = i * 4 + sj;
      const int cc = min(c0 + c, ncol - 1);   // a partial last tile repeats its last column
      sth[sl][c] = C2(s.thbuf, cc, k0 + sl);
      sdl[sl][c] = L2(s.exsmo, cc, k0 + sl);
    }
    wsync();
    double th[CH], dl[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) {
      th[j] = sth[j][lane];
      dl[j] = sdl[j][lane];
    }
    if (ch < NCH - 1) {   // the pair (k0+CH-1, k0+CH) finishes level k0+CH
      theta_pair(th[CH - 1], cth, dl[CH - 1], cdl, ldc(s.thetac, k0 + CH - 1), ldc(s.thetac, k0 + CH));
      sout[CH][lane] = (V)cth;
    }
#pragma unroll
    for (int j = CH - 2; j >= 0; j--) {
      if (k0 + j < 1) break;
      theta_pair(th[j], th[j + 1], dl[j], dl[j + 1], ldc(s.thetac, k0 + j), ldc(s.thetac, k0 + j + 1));
    }
    // levels k0+1 .. k0+CH-1 are final; k0 is carried (final in the bottom chunk)
#pragma unroll
    for (int j = 0; j < CH; j++) sout[j][lane] = (V)th[j];
    cth = th[0];
    cdl = dl[0];
    wsync();
    // store levels k0+1 .. k0+CH (16 levels of 4 columns per instruction), and level 0 at the bottom
    const int lev = k0 + 1 + sl;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int c = i * 4 + sj;
      if (c0 + c < ncol && lev < NVL && (!s.wet || s.wet[c0 + c])) TR(s.tr, 0, c0 + c, lev) = sout[1 + sl][c];
    }
    if (ch == 0 && c0 + lane < ncol && (!s.wet || s.wet[c0 + lane])) TR(s.tr, 0, c0 + lane, 0) = sout[0][lane];
    wsync();
  }
  (void)nvl;
  (void)nvlp1;
}

template <typename V>
void launch_hybgen_cols(const HybCols<V> &h, const StepCtl *ctl, int its, hipStream_t st, hipEvent_t winds_ready) {
  const dim3 wg((h.ncol + HW - 1) / HW), wb(64, HW);
  if (h.nvl == NVL_MAX && h.thbuf) {
    const dim3 tg((h.ncol + 63) / 64), tb(64);
    const double extop = CP * fim_pow(PTOP / P1000, RD / CP);
    KT_LAUNCH("k_regrid", st, (k_regrid_w<V, true><<<wg, wb, 0, st>>>(h, ctl, its)));
    KT_LAUNCH("k_regrid_tail", st, (k_regrid_tail<NVL_MAX, V><<<tg, tb, 0, st>>>(h, ctl, its, extop)));
    KT_LAUNCH("k_dffusn_lev", st, k_dffusn_lev<V><<<(h.ncol + 3) / 4, dim3(64, 4), 0, st>>>(h));
    KT_LAUNCH("k_remap_inflate", st, (k_remap_inflate<NVL_MAX, V><<<tg, tb, 0, st>>>(h, extop)));
    if (winds_ready) (void)hipStreamWaitEvent(st, winds_ready, 0);
    KT_LAUNCH("k_remap", st, (k_remap_w<V, true><<<wg, wb, 0, st>>>(h, ctl, its)));
    KT_LAUNCH("k_remap_theta", st, (k_remap_theta<NVL_MAX, V><<<tg, tb, 0, st>>>(h)));
    return;
  }
  KT_LAUNCH("k_regrid", st, (k_regrid_w<V, false><<<wg, wb, 0, st>>>(h, ctl, its)));
  KT_LAUNCH("k_dffusn_lev", st, k_dffusn_lev<V><<<(h.ncol + 3) / 4, dim3(64, 4), 0, st>>>(h));
  if (winds_ready) (void)hipStreamWaitEvent(st, winds_ready, 0);
  KT_LAUNCH("k_remap", st, (k_remap_w<V, false><<<wg, wb, 0, st>>>(h, ctl, its)));
}
template void launch_hybgen_cols<float>(const HybCols<float> &, const StepCtl *, int, hipStream_t, hipEvent_t);

void launch_hybgen(const DevGrid &g, const DevState &s, int its, hipStream_t st) {
  HybCols<double> h{g.nip, g.nip, g.nvl, g.ntra, s.us, s.vs, s.us, s.vs, s.tr, s.dp, s.pr, s.ex, s.sdot, s.exsmo,
                    s.thbuf, g.thetac, g.dpsig, g.smoo, g.sideln, g.rarea, g.prox, g.nprox};
  launch_hybgen_cols<double>(h, nullptr, its, st);
}

}  // namespace fimhip
