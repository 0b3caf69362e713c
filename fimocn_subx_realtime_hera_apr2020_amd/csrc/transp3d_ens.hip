// Class-B tracer transport (transp3d.F90:22-241, fct3d.F90:4-483) in the
// batched mixed-precision form: the class-B tracers, the antidiffusive fluxes
// and the fct3d work fields in FP32, the thickness at the start of the
// integration, the vertical mass flux and every tracer-mass sum in FP64.
//
// The batched form keeps only the current step's mass flux, so transp1 keeps
// its own ring of the last three steps' fluxes for the AB-weighted time
// integral.  The mass fluxes of the two cells of an edge are exact negatives
// of each other, so the integrated fluxes, fct3d's low-order and
// antidiffusive fluxes and the clipped fluxes are too, and the class-B tracer
// mass moves between cells without loss (fct3d's own clipping aside).
//
// The step is captured once as a graph: whether a step moves the class-B
// tracers (its % dtratio == 0, dyn_run.F90:396) is decided on the device from
// the step counter, and the transport kernels of other steps return at once.
// Mapping: one wavefront per column (lane = level), blockIdx.y = member,
// blockIdx.z = class-B tracer; the massless-cell sweeps one thread per column.
#include "ens_internal.h"

namespace fimhip {

#define BC2(a, col, k) (a)[(size_t)(col) * nvl + (k)]
#define BE3(a, col, e, k) (a)[((size_t)(col) * NPP + (e)) * nvl + (k)]
#define BTR(a, t, col, k) (a)[((size_t)(t) * ncol + (col)) * nvl + (k)]
#define BTE(a, t, col, e, k) (a)[(((size_t)(t) * ncol + (col)) * NPP + (e)) * nvl + (k)]
#define BPROX(i, e) ldc(g.d.prox, (i) * NPP + (e))
#define BPROXS(i, e) ldc(g.d.proxs, (i) * NPP + (e))
#define BNP(i) ldc(g.d.nprox, (i))

namespace {

constexpr int BC = 4;   // columns per block
constexpr float ATHIRD = 1.f / 3.f, EPSIL = 1.e-11f, ONEMU = 1.e-6f;

__device__ __forceinline__ int b_cell() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * BC + threadIdx.y));
}
// a step that moves the class-B tracers (dyn_run.F90:396)
__device__ __forceinline__ bool b_active(const EnsState &s, int dtratio) {
  return load_ctl(s.ctl).its % dtratio == 0;
}

// transp0 (transp3d.F90:22-41); unconditional at start-up
__global__ void __launch_bounds__(64 * BC) k_eb_transp0(EnsGrid g, EnsState s, EnsClassB b, int always) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = b_cell();
  if (ipn >= nip || k >= nvl) return;
  if (!always && !b_active(s, b.dtratio)) return;
  const int col = blockIdx.y * nip + ipn;
  BC2(b.dpinit, col, k) = BC2(s.dp, col, k);
#pragma unroll
  for (int e = 0; e < NPP; e++) BE3(b.cumufx, col, e, k) = 0.f;
}

// transp1 (transp3d.F90:44-78): this step's flux into the ring, then the
// AB-weighted sum of the ring's three slots added to the integral (FP64 sum,
// rounded once)
__global__ void __launch_bounds__(64 * BC) k_eb_transp1(EnsGrid g, EnsState s, EnsClassB b) {
  const int nip = g.d.nip, nvl = g.d.nvl, ncol = s.ncol;
  const int k = threadIdx.x, ipn = b_cell();
  if (ipn >= nip || k >= nvl) return;
  const StepCtl c = load_ctl(s.ctl);
  const int col = blockIdx.y * nip + ipn;
  const int np = BNP(ipn);
  float *rn = b.ring + (size_t)(c.nf - 1) * ncol * NPP * nvl;
  const float *ro = b.ring + (size_t)(c.of - 1) * ncol * NPP * nvl;
  const float *rv = b.ring + (size_t)(c.vof - 1) * ncol * NPP * nvl;
  for (int e = 0; e < np; e++) {
    const float m = BE3(s.massfx, col, e, k);
    const double inc = c.ab1 * (double)m + c.ab2 * (double)BE3(ro, col, e, k) + c.ab3 * (double)BE3(rv, col, e, k);
    BE3(rn, col, e, k) = m;
    BE3(b.cumufx, col, e, k) = BE3(b.cumufx, col, e, k) + (float)inc;
  }
}

// transp2 (transp3d.F90:137-169): the vertical mass flux w, summed upward
// from the bottom layer (FP64 wavefront scan)
__global__ void __launch_bounds__(64 * BC) k_eb_w(EnsGrid g, EnsState s, EnsClassB b) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = b_cell();
  if (ipn >= nip) return;
  if (!b_active(s, b.dtratio)) return;
  const int col = blockIdx.y * nip + ipn;
  const bool live = k < nvl;
  double vdv = 0.;
  if (live) {
    const int np = BNP(ipn);
    double hordiv = 0.;
    for (int e = 0; e < np; e++) hordiv = hordiv + (double)BE3(b.cumufx, col, e, k);
    vdv = (BC2(b.dpinit, col, k) - BC2(s.dp, col, k)) - hordiv * ldc(g.d.rarea, ipn);
  }
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(vdv, d, 64);
    if (k >= d) vdv += y;
  }
  if (live) BC2(b.w, col, k) = vdv;
}

// fct3d.F90:127-140: massless cells filled from the layers above and below,
// one thread per column and class-B tracer, the column in registers, loads
// and stores staged through LDS in chunks of 16 levels.  Needs nvl == NVL.
template <int NVL>
__global__ void __launch_bounds__(64) k_eb_fill(EnsState s, EnsClassB b) {
  constexpr int CH = 16, NCH = NVL / CH;
  static_assert(NVL % CH == 0, "whole chunks");
  __shared__ float sf[CH][65];
  __shared__ double sth[CH][65];
  if (!b_active(s, b.dtratio)) return;
  const int nvl = NVL, ncol = s.ncol;
  const int lane = threadIdx.x, c0 = blockIdx.x * 64, n = blockIdx.y;
  const float *trb = b.trb + (size_t)n * ncol * nvl;
  float *fld = b.fld + (size_t)n * ncol * nvl;
  float x[NVL];
#pragma unroll
  for (int ch = NCH - 1; ch >= 0; ch--) {
    const int k0 = ch * CH;
    __syncthreads();
    for (int i = 0; i < 16; i++) {
      const int cc = i * 4 + (lane >> 4), kk = lane & 15, gc = c0 + cc;
      const bool in = gc < ncol;
      sf[kk][cc] = in ? trb[(size_t)gc * nvl + k0 + kk] : 0.f;
      sth[kk][cc] = in ? b.dpinit[(size_t)gc * nvl + k0 + kk] : 1.;
    }
    __syncthreads();
#pragma unroll
    for (int kk = CH - 1; kk >= 0; kk--) {
      const int k = k0 + kk;
      if (k == NVL - 1) {
        x[k] = sf[kk][lane];
      } else {
        const float t = (float)sth[kk][lane];
        x[k] = (sf[kk][lane] * t + x[k + 1] * ONEMU) / (t + ONEMU);
      }
    }
  }
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const int k0 = ch * CH;
    __syncthreads();
    for (int i = 0; i < 16; i++) {
      const int cc = i * 4 + (lane >> 4), kk = lane & 15, gc = c0 + cc;
      sth[kk][cc] = gc < ncol ? b.dpinit[(size_t)gc * nvl + k0 + kk] : 1.;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < CH; kk++) {
      const int k = k0 + kk;
      if (k > 0) {
        const float t = (float)sth[kk][lane];
        x[k] = (x[k] * t + x[k - 1] * ONEMU) / (t + ONEMU);
      }
      sf[kk][lane] = x[k];
    }
    __syncthreads();
    for (int i = 0; i < 16; i++) {
      const int cc = i * 4 + (lane >> 4), kk = lane & 15, gc = c0 + cc;
      if (gc < ncol) fld[(size_t)gc * nvl + k0 + kk] = sf[kk][cc];
    }
  }
}

// fct3d.F90:141-260: (the massless-cell sweeps in lane 0 unless FILLED),
// the parabolic fit of each layer and the vertical tracer flux over the
// upstream slab
template <bool FILLED>
__global__ void __launch_bounds__(64 * BC) k_eb_vert(EnsGrid g, EnsState s, EnsClassB b) {
  __shared__ float sf[BC][NVL_MAX], sth[BC][NVL_MAX];
  __shared__ float sa[BC][NVL_MAX], sb[BC][NVL_MAX], sc[BC][NVL_MAX];
  const int nip = g.d.nip, nvl = g.d.nvl, ncol = s.ncol;
  const int k = threadIdx.x, c = threadIdx.y, ipn = b_cell(), n = blockIdx.z;
  if (!b_active(s, b.dtratio)) return;         // the same for every wave of the grid
  const int col = blockIdx.y * nip + ipn;
  const bool live = ipn < nip && k < nvl;       // every wave of the block reaches the barriers
  if (live) {
    sf[c][k] = FILLED ? BTR(b.fld, n, col, k) : BTR(b.trb, n, col, k);
    sth[c][k] = (float)BC2(b.dpinit, col, k);
  }
  __syncthreads();
  float *f = sf[c];
  const float *th = sth[c];
  if (!FILLED && k == 0 && ipn < nip) {
    for (int kk = nvl - 2; kk >= 0; kk--) f[kk] = (f[kk] * th[kk] + f[kk + 1] * ONEMU) / (th[kk] + ONEMU);
    for (int kk = 1; kk < nvl; kk++) f[kk] = (f[kk] * th[kk] + f[kk - 1] * ONEMU) / (th[kk] + ONEMU);
  }
  __syncthreads();
  if (live) {
    float a = f[k], bb = 0.f, cc = 0.f;
    if (k > 0 && k < nvl - 1) {
      const float dxlft = fmaxf(EPSIL, th[k - 1]), dxmid = fmaxf(EPSIL, th[k]), dxrgt = fmaxf(EPSIL, th[k + 1]);
      const float fk = f[k];
      const float yl = (dxlft * fk + dxmid * f[k - 1]) / (dxlft + dxmid);
      const float yr = (dxrgt * fk + dxmid * f[k + 1]) / (dxrgt + dxmid);
      const float mid = .5f * (yl + yr) - fk;
      a = 1.5f * fk - .25f * (yl + yr);
      bb = yr - yl;
      cc = 6.f * mid;
      if (fabsf(yr - yl) < 6.f * fabsf(mid)) {
        if (fabsf(yr - yl) > 2.f * fabsf(mid)) {
          if ((yr - yl) * mid > 0.f) {
            a = .25f * (3.f * fk + yl);
            cc = 3.f * (fk - yl);
            bb = cc;
          } else {
            a = .25f * (3.f * fk + yr);
            cc = 3.f * (fk - yr);
            bb = -cc;
          }
        } else {
          a = fk; bb = 0.f; cc = 0.f;
        }
      }
    }
    sa[c][k] = a; sb[c][k] = bb; sc[c][k] = cc;
  }
  __syncthreads();
  if (!live) return;
  const float *A = sa[c], *B = sb[c], *Cc = sc[c];
  float vfx = 0.f;
  if (k < nvl - 1) {
    const float wk = (float)BC2(b.w, col, k);
    float slab = ONEMU, amount = 0.f;
    if (wk < 0.f) {
      amount = slab * f[k + 1];
      int kp = k;
      for (;;) {
        kp = kp + 1;
        if (slab >= -wk) break;
        if (th[kp] > 0.f) {
          const float dslab = fminf(slab + th[kp], -wk) - fminf(slab, -wk);
          const float dx = dslab / th[kp];
          const float fcdx = A[kp] + B[kp] * .5f * (dx - 1.f) + Cc[kp] * (.25f - dx * (.5f - dx * ATHIRD));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
        if (!(kp < nvl - 1)) break;
      }
    } else if (wk > 0.f) {
      amount = slab * f[k];
      int kp = k + 1;
      for (;;) {
        kp = kp - 1;
        if (slab >= wk) break;
        if (th[kp] > 0.f) {
          const float dslab = fminf(slab + th[kp], wk) - fminf(slab, wk);
          const float dx = dslab / th[kp];
          const float fcdx = A[kp] + B[kp] * .5f * (1.f - dx) + Cc[kp] * (.25f - dx * (.5f - dx * ATHIRD));
          amount = amount + fcdx * dslab;
          slab = slab + dslab;
        }
        if (!(kp > 1)) break;
      }
    }
    vfx = wk * amount / slab;
  }
  BTR(b.vfx, n, col, k) = vfx;
  if (!FILLED) BTR(b.fld, n, col, k) = f[k];
}

// fct3d.F90:276-370: low-order and antidiffusive fluxes, local bounds, the
// low-order update (FP64) and the limiter ratios
__global__ void __launch_bounds__(64 * BC) k_eb_lo(EnsGrid g, EnsState s, EnsClassB b) {
  const int nip = g.d.nip, nvl = g.d.nvl, ncol = s.ncol;
  const int k = threadIdx.x, ipn = b_cell(), n = blockIdx.z;
  if (ipn >= nip || k >= nvl) return;
  if (!b_active(s, b.dtratio)) return;
  const int base = blockIdx.y * nip, col = base + ipn;
  const int np = BNP(ipn);
  const float fl = BTR(b.fld, n, col, k);
  float mx = fl, mn = fl;
  double hd = 0., totin = 0., totou = 0.;
  for (int e = 0; e < np; e++) {
    const int xc = base + BPROX(ipn, e), edx = BPROXS(ipn, e);
    const float ui = BE3(b.cumufx, col, e, k), ux = BE3(b.cumufx, xc, edx, k);
    const float fx = BTR(b.fld, n, xc, k);
    const float flx = .5f * ((ui + fabsf(ui)) * fl - (ux + fabsf(ux)) * fx);
    const float an = .5f * (fl + fx) * ui - flx;
    BTE(b.uan, n, col, e, k) = an;
    mx = fmaxf(mx, fx);
    mn = fminf(mn, fx);
    hd = hd + (double)flx;
    totin = totin - (double)fminf(0.f, an);
    totou = totou + (double)fmaxf(0.f, an);
  }
  const float vk = BTR(b.vfx, n, col, k);
  if (k < nvl - 1) {
    const float wk = (float)BC2(b.w, col, k);
    if (wk < 0.f) {
      mx = fmaxf(mx, vk / wk);
      mn = fminf(mn, vk / wk);
    }
  }
  double vdv = vk;
  if (k > 0) {
    const float vm = BTR(b.vfx, n, col, k - 1), wm = (float)BC2(b.w, col, k - 1);
    if (wm > 0.f) {
      mx = fmaxf(mx, vm / wm);
      mn = fminf(mn, vm / wm);
    }
    vdv = (double)vk - (double)vm;
  }
  const double rar = ldc(g.d.rarea, ipn);
  const double thko = BC2(b.dpinit, col, k), thkn = BC2(s.dp, col, k);
  const double q = (double)fl * thko - hd * rar - vdv;
  const double amount = fmax(fmax(0., (double)mn * thkn), fmin(q, (double)mx * thkn));
  const float fl1 = (float)(((double)fl * ONEMU + amount) / (ONEMU + thkn));
  BTR(b.fmx, n, col, k) = mx;
  BTR(b.fmn, n, col, k) = mn;
  BTR(b.fld1, n, col, k) = fl1;
  BTR(b.flp, n, col, k) = (float)(((double)mx - fl1) * thkn / (totin + EPSIL) * rar);
  BTR(b.fln, n, col, k) = (float)(((double)fl1 - mn) * thkn / (totou + EPSIL) * rar);
}

// fct3d.F90:372-405: limited antidiffusive fluxes and the final update
__global__ void __launch_bounds__(64 * BC) k_eb_clip(EnsGrid g, EnsState s, EnsClassB b) {
  const int nip = g.d.nip, nvl = g.d.nvl, ncol = s.ncol;
  const int k = threadIdx.x, ipn = b_cell(), n = blockIdx.z;
  if (ipn >= nip || k >= nvl) return;
  if (!b_active(s, b.dtratio)) return;
  const int base = blockIdx.y * nip, col = base + ipn;
  const int np = BNP(ipn);
  const float lp = BTR(b.flp, n, col, k), ln = BTR(b.fln, n, col, k);
  double hd = 0.;
  for (int e = 0; e < np; e++) {
    const int xc = base + BPROX(ipn, e);
    const float an = BTE(b.uan, n, col, e, k);
    const float clip = an >= 0.f ? fminf(fminf(1.f, ln), BTR(b.flp, n, xc, k))
                                 : fminf(fminf(1.f, lp), BTR(b.fln, n, xc, k));
    hd = hd + (double)(an * clip);
  }
  const double thkn = BC2(s.dp, col, k);
  const double fl1 = BTR(b.fld1, n, col, k);
  const double q = fl1 * thkn - hd * ldc(g.d.rarea, ipn);
  const double amount = fmax(fmax(0., (double)BTR(b.fmn, n, col, k) * thkn),
                             fmin(q, (double)BTR(b.fmx, n, col, k) * thkn));
  BTR(b.trb, n, col, k) = (float)((fl1 * ONEMU + amount) / (ONEMU + thkn));
}

}  // namespace

void ens_launch_class_b_init(const EnsGrid &g, const EnsState &s, const EnsClassB &b, hipStream_t st) {
  const dim3 gr((g.d.nip + BC - 1) / BC, s.members), bl(64, BC);
  k_eb_transp0<<<gr, bl, 0, st>>>(g, s, b, 1);
}

// transp1 every step; transp2 (w, fct3d) and transp0 on the steps that move
// the class-B tracers (decided on the device)
void ens_launch_class_b_step(const EnsGrid &g, const EnsState &s, const EnsClassB &b, hipStream_t st) {
  const dim3 gr((g.d.nip + BC - 1) / BC, s.members), bl(64, BC);
  const dim3 gt((g.d.nip + BC - 1) / BC, s.members, b.ntrb);
  KT_LAUNCH("eb_transp1", st, k_eb_transp1<<<gr, bl, 0, st>>>(g, s, b));
  KT_LAUNCH("eb_w", st, k_eb_w<<<gr, bl, 0, st>>>(g, s, b));
  if (g.d.nvl == NVL_MAX) {
    KT_LAUNCH("eb_fill", st, k_eb_fill<NVL_MAX><<<dim3((s.ncol + 63) / 64, b.ntrb), 64, 0, st>>>(s, b));
    KT_LAUNCH("eb_vert", st, k_eb_vert<true><<<gt, bl, 0, st>>>(g, s, b));
  } else {
    KT_LAUNCH("eb_vert", st, k_eb_vert<false><<<gt, bl, 0, st>>>(g, s, b));
  }
  KT_LAUNCH("eb_lo", st, k_eb_lo<<<gt, bl, 0, st>>>(g, s, b));
  KT_LAUNCH("eb_clip", st, k_eb_clip<<<gt, bl, 0, st>>>(g, s, b));
  KT_LAUNCH("eb_transp0", st, k_eb_transp0<<<gr, bl, 0, st>>>(g, s, b, 0));
}

}  // namespace fimhip
