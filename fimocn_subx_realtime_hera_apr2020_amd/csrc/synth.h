// Seeded synthetic FIM states, computed identically on the host and on the
// GPU: every random number is a splitmix64 hash of (seed, member, field,
// cell, level), and every transcendental function used per cell-level is
// built from IEEE operations only (fim_pow.h), so the bytes do not depend on
// the device, on the batching of members or on the tiling of cells.
// Per-cell profiles that need sin/cos of the grid coordinates (solid-body
// wind, cosine bell) are computed once on the host and passed in.
#pragma once
#include <stdint.h>
#include "fim_pow.h"

#define FIM_SYN_DP 1
#define FIM_SYN_TH 2
#define FIM_SYN_U 3
#define FIM_SYN_V 4
#define FIM_SYN_TR 10
#define FIM_SYN_MEMBER 100

typedef struct {
  int nip, nvl, ntra;
  uint64_t seed;
  int member;
  int wind_noise;       // add N(0,1) m/s to the solid-body wind
  int perturb;          // ensemble member perturbations
  int lognormal;        // tracers 2..T: log-normal x bell instead of bell + U[0,1e-2]
} fim_synth_params;

FIM_HD static inline uint64_t fim_mix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

FIM_HD static inline uint64_t fim_field_key(uint64_t seed, int member, int field) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  uint64_t key = fim_mix64(seed * G + (uint64_t)member);
  return fim_mix64(key ^ ((uint64_t)field * G));
}

// U[0,1) for one (cell, level)
FIM_HD static inline double fim_uniform(uint64_t key, int cell, int lev) {
  const uint64_t G = 0x9E3779B97F4A7C15ULL;
  uint64_t x = fim_mix64(key ^ fim_mix64((uint64_t)cell * 0x100000ULL + (uint64_t)lev + G));
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// cos(2 pi u) for u in [0,1): reduce to an octant, then Taylor polynomials
FIM_HD static inline double fim_cos2pi(double u) {
  const double TWO_PI = 6.283185307179586;
  double q = u * 8.0;                 // octant index in [0, 8)
  int oct = (int)q;
  double f = u - (double)oct * 0.125; // in [0, 1/8)
  // angle within the octant, measured from the octant's start
  double s, c;
  int quad = oct >> 1;
  bool odd = oct & 1;
  double a = odd ? (0.125 - f) * TWO_PI : f * TWO_PI;   // in [0, pi/4]
  double a2 = a * a;
  // sin and cos on [0, pi/4] to ~1e-17
  s = a * (1.0 + a2 * (-1.0 / 6 + a2 * (1.0 / 120 + a2 * (-1.0 / 5040 + a2 * (1.0 / 362880 +
      a2 * (-1.0 / 39916800 + a2 * (1.0 / 6227020800.0)))))));
  c = 1.0 + a2 * (-0.5 + a2 * (1.0 / 24 + a2 * (-1.0 / 720 + a2 * (1.0 / 40320 + a2 * (-1.0 / 3628800 +
      a2 * (1.0 / 479001600.0 + a2 * (-1.0 / 87178291200.0)))))));
  // within the quadrant the angle is f' = a (even octant) or pi/2 - a (odd)
  double cq = odd ? s : c;            // cos of the angle within the quadrant
  double sq = odd ? c : s;
  switch (quad) {
    case 0: return cq;
    case 1: return -sq;
    case 2: return -cq;
    default: return sq;
  }
}

FIM_HD static inline double fim_normal(uint64_t key1, uint64_t key2, int cell, int lev) {
  double u1 = fim_uniform(key1, cell, lev);
  double u2 = fim_uniform(key2, cell, lev);
  double lh, ll;
  fim_log_dd(1.0 - u1, &lh, &ll);     // 1 - u1 in (0, 1]
  double r = -2.0 * lh;
  r = r > 0.0 ? sqrt(r) : 0.0;
  return r * fim_cos2pi(u2);
}

// Fill one cell column.  u0c, v0c: solid-body wind at the cell; bell: the
// cosine bell at the cell; pint: the nvl+1 reference interface pressures.
FIM_HD static inline void fim_synth_column(const fim_synth_params &p, int cell, double u0c, double v0c,
                                           double bell, const double *pint, double *us, double *vs,
                                           double *dp, double *pr, double *ex, double *ph, double *tr,
                                           size_t trstride) {
  const int nvl = p.nvl;
  const double CP = 1004.6855, RD = 287.0586, P1000 = 100000.;
  const uint64_t kdp = fim_field_key(p.seed, 0, FIM_SYN_DP);
  const uint64_t kdpm = fim_field_key(p.seed, p.member, FIM_SYN_MEMBER + FIM_SYN_DP);
  for (int k = 0; k < nvl; k++) {
    double d = (pint[k] - pint[k + 1]) * (1.0 + 0.05 * (2.0 * fim_uniform(kdp, cell, k) - 1.0));
    if (p.perturb) d = d * (1.0 + 0.02 * (2.0 * fim_uniform(kdpm, cell, k) - 1.0));
    dp[k] = d;
  }
  pr[nvl] = pint[nvl];
  for (int k = nvl - 1; k >= 0; k--) pr[k] = pr[k + 1] + dp[k];
  for (int k = 0; k <= nvl; k++) {
    ex[k] = CP * fim_pow(pr[k] / P1000, RD / CP);
    ph[k] = 0.0;
  }
  const uint64_t kth = fim_field_key(p.seed, 0, FIM_SYN_TH);
  const uint64_t kthm1 = fim_field_key(p.seed, p.member, FIM_SYN_MEMBER + FIM_SYN_TH);
  const uint64_t kthm2 = fim_field_key(p.seed, p.member, FIM_SYN_MEMBER + FIM_SYN_TH + 50);
  const uint64_t ku1 = fim_field_key(p.seed, 0, FIM_SYN_U), ku2 = fim_field_key(p.seed, 0, FIM_SYN_U + 50);
  const uint64_t kv1 = fim_field_key(p.seed, 0, FIM_SYN_V), kv2 = fim_field_key(p.seed, 0, FIM_SYN_V + 50);
  const uint64_t kum1 = fim_field_key(p.seed, p.member, FIM_SYN_MEMBER + FIM_SYN_U);
  const uint64_t kum2 = fim_field_key(p.seed, p.member, FIM_SYN_MEMBER + FIM_SYN_U + 50);
  const uint64_t kvm1 = fim_field_key(p.seed, p.member, FIM_SYN_MEMBER + FIM_SYN_V);
  const uint64_t kvm2 = fim_field_key(p.seed, p.member, FIM_SYN_MEMBER + FIM_SYN_V + 50);
  const double dnv = (double)(nvl > 1 ? nvl - 1 : 1);
  for (int k = 0; k < nvl; k++) {
    double th = 250.0 + 350.0 * (double)k / dnv + (fim_uniform(kth, cell, k) - 0.5);
    if (p.perturb) th = th + 0.5 * fim_normal(kthm1, kthm2, cell, k);
    tr[k] = th;
    double u = u0c, v = v0c;
    if (p.wind_noise) {
      u = u + fim_normal(ku1, ku2, cell, k);
      v = v + fim_normal(kv1, kv2, cell, k);
    }
    if (p.perturb) {
      u = u + fim_normal(kum1, kum2, cell, k);
      v = v + fim_normal(kvm1, kvm2, cell, k);
    }
    us[k] = u;
    vs[k] = v;
  }
  for (int t = 1; t < p.ntra; t++) {
    double *trt = tr + (size_t)t * trstride;
    const uint64_t k1 = fim_field_key(p.seed, 0, FIM_SYN_TR + t);
    const uint64_t k2 = fim_field_key(p.seed, 0, FIM_SYN_TR + t + 50);
    for (int k = 0; k < nvl; k++) {
      if (p.lognormal) {
        double z = fim_normal(k1, k2, cell, k);
        trt[k] = fim_exp_dd(-20.72326583694641 + 1.5 * z, 0.0) * bell;   // exp(ln 1e-9 + 1.5 z)
      } else {
        trt[k] = bell + 1.0e-2 * fim_uniform(k1, cell, k);
      }
    }
  }
}
