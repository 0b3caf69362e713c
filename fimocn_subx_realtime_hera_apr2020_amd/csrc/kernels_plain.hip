// FP64 kernels of the FIM dynamics hot path, one kernel family per reference
// routine, following the reference's operation order exactly (build with
// -ffp-contract=off; x**y through fim_pow).  Mapping: one wavefront lane per
// level and one wavefront per cell column (block = 64 x 4 columns), so every
// neighbour-column gather through prox is a coalesced 64-element load.  The
// column-serial parts of hybgen and hystat run one thread per column.
#include "fimhip_internal.h"
#include "fim_pow.h"

namespace fimhip {

#define C2(a, ipn, k) (a)[(size_t)(ipn) * nvl + (k)]
#define L2(a, ipn, k) (a)[(size_t)(ipn) * nvlp1 + (k)]
#define E3(a, ipn, e, k) (a)[((size_t)(ipn) * NPP + (e)) * nvl + (k)]
#define AB3(a, s, ipn, k) (a)[((size_t)(s) * nip + (ipn)) * nvl + (k)]
#define TR(a, t, ipn, k) (a)[((size_t)(t) * nip + (ipn)) * nvl + (k)]
#define TAB(a, t, s, ipn, k) (a)[(((size_t)(t) * 3 + (s)) * nip + (ipn)) * nvl + (k)]
#define MF(a, s, ipn, e, k) (a)[(((size_t)(s) * nip + (ipn)) * NPP + (e)) * nvl + (k)]
#define PROX(ipn, e) ldc(g.prox, (ipn) * NPP + (e))
#define PROXS(ipn, e) ldc(g.proxs, (ipn) * NPP + (e))
#define CS(ipn, e, i) ldc(g.cs, ((ipn) * NPP + (e)) * 4 + (i))
#define SN(ipn, e, i) ldc(g.sn, ((ipn) * NPP + (e)) * 4 + (i))
#define SVC(ipn, e, i) ldc(g.svc, ((ipn) * NPP + (e)) * 2 + (i))
#define SVE(ipn, e, i) ldc(g.sve, ((ipn) * NPP + (e)) * 2 + (i))
#define HF(ipn, i, j) ldc(g.hfcoef, ((ipn) * NPP + (i)) * NPP + (j))
#define GNP(ipn) ldc(g.nprox, (ipn))
#define GV(a, ipn) ldc(g.a, (ipn))

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

constexpr int COLS = 4;   // columns per block

// The column index of a block row is the same for all 64 lanes of a wave
// (blockDim.x == 64 == wavefront size); making it wave-uniform lets every
// grid-table access (prox, cs, sn, sidevec, hfcoef, area) go through
// scalar loads instead of vector gathers.
__device__ __forceinline__ int wave_col() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * COLS + threadIdx.y));
}

static inline dim3 col_grid(int nip) { return dim3((nip + COLS - 1) / COLS); }
static inline dim3 col_block() { return dim3(NVL_MAX, COLS); }

// ---------------------------------------------------------------------------
// edgvar1.F90:160-421 (eqwgt branch) + tracer-1 loop 425-511.  The values of
// the cell and its NP neighbours are loaded once into registers (slot j of
// the stencil at index j + 1, the cell itself at 0) and the edge loop is
// unrolled over them; every edge value is formed by the reference's
// operations in its order.
template <int NP>
__device__ __forceinline__ void edgvar1_col(const DevGrid &g, const DevState &s, int ipn, int k) {
  const int nip = g.nip, nvl = g.nvl, nvlp1 = nvl + 1;
  const double divby6 = 1. / 6., eps = 0.001;
  double cu[NP + 1], cv[NP + 1], cm[NP + 1], ct[NP + 1], cp0[NP + 1], cp1[NP + 1];
#pragma unroll
  for (int j = 0; j <= NP; j++) {
    const int c = j == 0 ? ipn : PROX(ipn, j - 1);
    cu[j] = C2(s.us, c, k);
    cv[j] = C2(s.vs, c, k);
    cm[j] = C2(s.mp, c, k);
    ct[j] = TR(s.tr, 0, c, k);
    cp0[j] = L2(s.pr, c, k);
    cp1[j] = L2(s.pr, c, k + 1);
  }
#pragma unroll
  for (int edg = 0; edg < NP; edg++) {
    const int i2 = edg + 1, i3t = (edg - 1 + NP) % NP + 1, i3l = (edg + 1) % NP + 1;
    double u3trl = CS(ipn, edg, 2) * cu[i3t] + SN(ipn, edg, 2) * cv[i3t];
    double v3trl = -SN(ipn, edg, 2) * cu[i3t] + CS(ipn, edg, 2) * cv[i3t];
    double u1ldg = CS(ipn, edg, 0) * cu[0] + SN(ipn, edg, 0) * cv[0];
    double v1ldg = -SN(ipn, edg, 0) * cu[0] + CS(ipn, edg, 0) * cv[0];
    double u2ldg = CS(ipn, edg, 1) * cu[i2] + SN(ipn, edg, 1) * cv[i2];
    double v2ldg = -SN(ipn, edg, 1) * cu[i2] + CS(ipn, edg, 1) * cv[i2];
    double u3ldg = CS(ipn, edg, 3) * cu[i3l] + SN(ipn, edg, 3) * cv[i3l];
    double v3ldg = -SN(ipn, edg, 3) * cu[i3l] + CS(ipn, edg, 3) * cv[i3l];
    double ue = (2. * (u1ldg + u2ldg) + u3trl + u3ldg) * divby6;
    double ve = (2. * (v1ldg + v2ldg) + v3trl + v3ldg) * divby6;
    double m_1ldg = sqrt(u1ldg * u1ldg + v1ldg * v1ldg);
    double m_2ldg = sqrt(u2ldg * u2ldg + v2ldg * v2ldg);
    double m_3trl = sqrt(u3trl * u3trl + v3trl * v3trl);
    double m_3ldg = sqrt(u3ldg * u3ldg + v3ldg * v3ldg);
    double m_1 = sqrt(ue * ue + ve * ve);
    double m_2 = (2. * (m_1ldg + m_2ldg) + m_3trl + m_3ldg) * divby6;
    ue = ue * m_2 / dmax(eps, m_1);
    ve = ve * m_2 / dmax(eps, m_1);
    __builtin_nontemporal_store(ue, &E3(s.u_edg, ipn, edg, k));
    __builtin_nontemporal_store(ve, &E3(s.v_edg, ipn, edg, k));
    E3(s.lp_edg, ipn, edg, k) =
        (cp0[0] + cp1[0] + cp0[i2] + cp1[i2] + .5 * (cp0[i3t] + cp1[i3t] + cp0[i3l] + cp1[i3l])) * divby6;
    __builtin_nontemporal_store((2. * (cm[0] + cm[i2]) + cm[i3t] + cm[i3l]) * divby6, &E3(s.mont_edg, ipn, edg, k));
    __builtin_nontemporal_store(.5 * (ue * ue + ve * ve), &E3(s.uvsq_edg, ipn, edg, k));
    __builtin_nontemporal_store((2. * (ct[0] + ct[i2]) + ct[i3t] + ct[i3l]) * divby6, &E3(s.trc_edg, ipn, edg, k));
  }
  if (NP < NPP) __builtin_nontemporal_store(0., &E3(s.trc_edg, ipn, NPP - 1, k));
}

__global__ void __launch_bounds__(256) k_edgvar1(DevGrid g, DevState s) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  if (GNP(ipn) == 6) edgvar1_col<6>(g, s, ipn, k);
  else edgvar1_col<5>(g, s, ipn, k);
}

// ---------------------------------------------------------------------------
// edgvar3.F90:65-163: LS reconstruction of each edge (edgvalEval) and the
// two-sided average of loop 2.  k_edgvar3 (the drop-in entry) fuses them:
// each cell-level recomputes its neighbour's reconstruction of the shared
// edge (2-ring gathers).  k_edgvar3_own (the device-resident step) stores
// only each cell's own reconstructions, every field of the cell in one
// thread, and the readers (cnuity12, massadv1, trcadv1) form
// 0.5*(own + neighbour's) themselves, the same operands in the same order,
// so every result keeps its bits.
__global__ void __launch_bounds__(256) k_edgvar3(DevGrid g, DevState s) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  const int f = blockIdx.y;   // 0: dp, f >= 1: tracer f
  if (ipn >= nip || k >= nvl) return;
  const double *fld = f == 0 ? s.dp : s.tr + (size_t)f * nip * nvl;
  double *out = f == 0 ? s.dp_edg : s.trc_edg + (size_t)f * nip * NPP * nvl;
  const int np = GNP(ipn);
  const double b0 = fld[(size_t)ipn * nvl + k];
  double nbv[NPP];
  for (int j = 0; j < NPP; j++) nbv[j] = fld[(size_t)PROX(ipn, j) * nvl + k];
  for (int i = 0; i < np; i++) {
    // this cell's reconstruction of edge i (edgvalEval)
    double own = b0;
    for (int j = 0; j < np; j++) own = own + (nbv[j] - b0) * HF(ipn, i, j);
    // the neighbour's reconstruction of the same edge
    const int nb = PROX(ipn, i), eo = PROXS(ipn, i), nn = GNP(nb);
    const double c0 = nbv[i];
    double oth = c0;
    for (int j = 0; j < nn; j++) oth = oth + (fld[(size_t)PROX(nb, j) * nvl + k] - c0) * HF(nb, eo, j);
    E3(out, ipn, i, k) = 0.5 * (own + oth);
  }
  for (int i = np; i < NPP; i++) E3(out, ipn, i, k) = 0.;
}

template <int NP>
__device__ __forceinline__ void edgvar3_own_col(const DevGrid &g, const DevState &s, int ipn, int k) {
  const int nip = g.nip, nvl = g.nvl;
  int nb[NP];
#pragma unroll
  for (int j = 0; j < NP; j++) nb[j] = PROX(ipn, j);
  for (int f = 0; f < g.ntra; f++) {   // 0: dp, f >= 1: tracer f
    const double *fld = f == 0 ? s.dp : s.tr + (size_t)f * nip * nvl;
    double *out = f == 0 ? s.dp_edg : s.trc_edg + (size_t)f * nip * NPP * nvl;
    const double b0 = C2(fld, ipn, k);
    double nbv[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) nbv[j] = C2(fld, nb[j], k);
#pragma unroll
    for (int i = 0; i < NP; i++) {
      double own = b0;
#pragma unroll
      for (int j = 0; j < NP; j++) own = own + (nbv[j] - b0) * HF(ipn, i, j);
      E3(out, ipn, i, k) = own;
    }
    if (NP < NPP) E3(out, ipn, NPP - 1, k) = 0.;
  }
}

__global__ void __launch_bounds__(256) k_edgvar3_own(DevGrid g, DevState s) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  if (GNP(ipn) == 6) edgvar3_own_col<6>(g, s, ipn, k);
  else edgvar3_own_col<5>(g, s, ipn, k);
}

// an edge value of edgvar3's fields as the readers see it: stored averaged
// (TWO_SIDED), or formed here from both cells' own reconstructions
template <bool TWO_SIDED>
__device__ __forceinline__ double edge_val(const double *a, int ipn, int edg, int ipx, int edx, int k, int nvl) {
  if (TWO_SIDED) return E3(a, ipn, edg, k);
  return 0.5 * (E3(a, ipn, edg, k) + E3(a, ipx, edx, k));
}

// ---------------------------------------------------------------------------
// cnuity1 + cnuity2 (cnuity_compute.F90:17-150)
template <bool TWO_SIDED>
__global__ void __launch_bounds__(256) k_cnuity12(DevGrid g, DevState s, AB ab) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  const int nf = ab.nf - 1, of = ab.of - 1, vof = ab.vof - 1;
  const int np = GNP(ipn);
  const double dpi = C2(s.dp, ipn, k);
  double acc = 0.;
  for (int edg = 0; edg < np; edg++) {
    const int ipx = PROX(ipn, edg), edx = PROXS(ipn, edg);
    double vn = SVE(ipn, edg, 1) * E3(s.u_edg, ipn, edg, k) - SVE(ipn, edg, 0) * E3(s.v_edg, ipn, edg, k);
    double vx = SVE(ipx, edx, 1) * E3(s.u_edg, ipx, edx, k) - SVE(ipx, edx, 0) * E3(s.v_edg, ipx, edx, k);
    double flxhi = vn * edge_val<TWO_SIDED>(s.dp_edg, ipn, edg, ipx, edx, k, nvl);
    double m = 0.5 * ((vn + fabs(vn)) * dpi - (vx + fabs(vx)) * C2(s.dp, ipx, k));
    __builtin_nontemporal_store(m, &MF(s.massfx, nf, ipn, edg, k));
    __builtin_nontemporal_store(flxhi - m, &E3(s.antifx, ipn, edg, k));
    acc = acc + m;
  }
  if (np < NPP) {
    __builtin_nontemporal_store(0., &MF(s.massfx, nf, ipn, NPP - 1, k));
    __builtin_nontemporal_store(0., &E3(s.antifx, ipn, NPP - 1, k));
  }
  double dplo = -acc * GV(rarea, ipn);
  __builtin_nontemporal_store(dplo, &AB3(s.dpl_tdcy, nf, ipn, k));
  C2(s.delp_lo, ipn, k) = dpi + ab.ab1 * dplo + ab.ab2 * AB3(s.dpl_tdcy, of, ipn, k) +
                          ab.ab3 * AB3(s.dpl_tdcy, vof, ipn, k);
}

// cnuity3 (cnuity_compute.F90:152-259)
__global__ void __launch_bounds__(256) k_cnuity3(DevGrid g, DevState s, AB ab) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  const int of = ab.of - 1, vof = ab.vof - 1;
  const double thshld = 1.e-11;
  const int np = GNP(ipn);
  double pp = 0., pm = 0.;
  for (int edg = 0; edg < np; edg++) {
    double a = E3(s.antifx, ipn, edg, k);
    if (a <= 0.) pp = pp - a; else pm = pm + a;
  }
  const double lo = C2(s.delp_lo, ipn, k);
  double dpmax = dmax(lo, C2(s.dp, ipn, k));
  double dpmin = dmin(lo, C2(s.dp, ipn, k));
  for (int e = 0; e < NPP; e++) {
    int x = PROX(ipn, e);
    double a = C2(s.delp_lo, x, k), b = C2(s.dp, x, k);
    dpmax = dmax(dpmax, dmax(a, b));
    dpmin = dmin(dpmin, dmin(a, b));
  }
  dpmax = dmax(0., dpmax);
  dpmin = dmax(0., dpmin);
  const double to = AB3(s.dp_tdcy, of, ipn, k), tv = AB3(s.dp_tdcy, vof, ipn, k);
  double q_plus = (dpmax - lo - (ab.ab2 * dmin(0., to) + ab.ab3 * dmax(0., tv))) / ab.ab1 * GV(area, ipn);
  double q_mnus = (lo - dpmin + (ab.ab2 * dmax(0., to) + ab.ab3 * dmin(0., tv))) / ab.ab1 * GV(area, ipn);
  __builtin_nontemporal_store(dmax(0., dmin(1., q_plus / dmax(pp, thshld))), &C2(s.r_plus, ipn, k));
  __builtin_nontemporal_store(dmax(0., dmin(1., q_mnus / dmax(pm, thshld))), &C2(s.r_mnus, ipn, k));
}

// cnuity5 (cnuity_compute.F90:376-451), which is also update_pres of mass_adv
// (massadv_compute.F90:73-171): the top-down pressure and column-tendency
// sums run serially in lane 0 in the reference summation order over the new
// thickness (dpn) and layer tendency (lt) each lane left in LDS; then the
// Exner function and omega per lane.  Called by every thread of the block.
__device__ __forceinline__ void update_pres_col(const DevGrid &g, const DevState &s, int ipn, bool live,
                                                const double *dpn, const double *lt, double *pr, double *om) {
  const int nip = g.nip, nvl = g.nvl, nvlp1 = nvl + 1;
  const int k = threadIdx.x;
  __syncthreads();
  if (k == 0 && ipn < nip) {
    double p = L2(s.pr, ipn, nvl);
    double coltend = 0.;
    for (int kk = nvl - 1; kk >= 0; kk--) {
      p = p + dpn[kk];
      pr[kk] = p;
      const double t = lt[kk];
      om[kk] = coltend + 0.5 * t;
      coltend = coltend + t;
    }
  }
  __syncthreads();
  if (live) {
    const double p = pr[k];
    __builtin_nontemporal_store(p, &L2(s.pr, ipn, k));
    __builtin_nontemporal_store(CP * fim_pow(p / P1000, RD / CP), &L2(s.ex, ipn, k));
    const int np = GNP(ipn);
    double dpdx = 0., dpdy = 0.;
    for (int edg = 0; edg < np; edg++) {
      dpdx = dpdx + E3(s.lp_edg, ipn, edg, k) * SVC(ipn, edg, 1);
      dpdy = dpdy - E3(s.lp_edg, ipn, edg, k) * SVC(ipn, edg, 0);
    }
    __builtin_nontemporal_store(om[k] + (C2(s.us, ipn, k) * dpdx + C2(s.vs, ipn, k) * dpdy) * GV(rarea, ipn), &C2(s.ws, ipn, k));
  }
}

// cnuity4 + cnuity5 (cnuity_compute.F90:261-451), one block row per column:
// the top-down pressure and column-tendency sums are serial in lane 0 (the
// reference summation order), the Exner function is per lane.
__global__ void __launch_bounds__(256) k_cnuity45(DevGrid g, DevState s, AB ab) {
  __shared__ double sh_dp[COLS][NVL_MAX + 1];
  __shared__ double sh_lt[COLS][NVL_MAX + 1];
  __shared__ double sh_pr[COLS][NVL_MAX + 1];
  __shared__ double sh_om[COLS][NVL_MAX + 1];
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, c = threadIdx.y;
  const int ipn = wave_col();
  const bool live = ipn < nip && k < nvl;
  const int nf = ab.nf - 1, of = ab.of - 1, vof = ab.vof - 1;
  double dtn = 0.;
  if (live) {
    const int np = GNP(ipn);
    double at = 0.;
    for (int edg = 0; edg < np; edg++) {
      const int x = PROX(ipn, edg);
      double a = E3(s.antifx, ipn, edg, k), clip;
      if (a >= 0.) clip = dmin(C2(s.r_mnus, ipn, k), C2(s.r_plus, x, k));
      else clip = dmin(C2(s.r_plus, ipn, k), C2(s.r_mnus, x, k));
      a = a * clip;
      __builtin_nontemporal_store(MF(s.massfx, nf, ipn, edg, k) + a, &MF(s.massfx, nf, ipn, edg, k));
      at = at + a;
    }
    at = -at * GV(rarea, ipn);
    dtn = AB3(s.dpl_tdcy, nf, ipn, k) + at;
    __builtin_nontemporal_store(dtn, &AB3(s.dp_tdcy, nf, ipn, k));
    const double to = AB3(s.dp_tdcy, of, ipn, k), tv = AB3(s.dp_tdcy, vof, ipn, k);
    double dpn = C2(s.dp, ipn, k) + ab.ab1 * dtn + ab.ab2 * to + ab.ab3 * tv;
    __builtin_nontemporal_store(dpn, &C2(s.dp, ipn, k));
    sh_dp[c][k] = dpn;
    sh_lt[c][k] = (ab.ab1 * dtn + ab.ab2 * to + ab.ab3 * tv) / g.dt;
  }
  update_pres_col(g, s, ipn, live, sh_dp[c], sh_lt[c], sh_pr[c], sh_om[c]);
}

// ---------------------------------------------------------------------------
// mass_adv, the flux-limited mass transport of flux_conserv_schm = 2
// (massadv.F90:21-160, dyn_run.F90:377-386).
//
// k_massadv1: edgeNormVel (flux_limiter.F90:78-103) and limiter
// (flux_limiter.F90:132-260, M1 form) for one column: each edge depth is
// clamped to the four cells around the edge (q_star, stored back into
// dp_edg), then the inflow sums give the bounds q_min_out / q_max_out of the
// column's outflow edge values (kept in r_plus / r_mnus).  Normal winds and
// Courant numbers stay in registers.  Without TWO_SIDED, dp_edg holds the
// cells' own reconstructions, which the neighbours still read, so q_star goes
// to antifx (unused by mass_adv) and massadv2 takes it from there.
template <bool TWO_SIDED>
__global__ void __launch_bounds__(256) k_massadv1(DevGrid g, DevState s, AB ab) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  const int of = ab.of - 1, vof = ab.vof - 1;
  const double dt = g.dt, rar = GV(rarea, ipn);
  const double max_out_bound = 100000.0;
  const int np = GNP(ipn);
  const double gv = C2(s.dp, ipn, k);
  const double to = AB3(s.dp_tdcy, of, ipn, k), tv = AB3(s.dp_tdcy, vof, ipn, k);
  double q_min_in_min = 1.7976931348623157e308, q_max_in_max = 0.;   // huge(), 0
  double cq_min_in_sum = 0., cq_max_in_sum = 0., c_num_in_sum = 0., c_num_out_sum = 0.;
  for (int edg = 0; edg < np; edg++) {
    const int ipx = PROX(ipn, edg);
    const int nbm1 = PROX(ipn, edg == 0 ? np - 1 : edg - 1), nbp1 = PROX(ipn, edg + 1 == np ? 0 : edg + 1);
    const double vn = SVE(ipn, edg, 1) * E3(s.u_edg, ipn, edg, k) - SVE(ipn, edg, 0) * E3(s.v_edg, ipn, edg, k);
    const double cn = fabs(vn) * dt * rar;
    const double b = C2(s.dp, ipx, k), cc = C2(s.dp, nbm1, k), d = C2(s.dp, nbp1, k);
    const double q_min = dmin(dmin(gv, b), dmin(cc, d)), q_max = dmax(dmax(gv, b), dmax(cc, d));
    const double qs = dmax(q_min, dmin(q_max, edge_val<TWO_SIDED>(s.dp_edg, ipn, edg, ipx, PROXS(ipn, edg), k, nvl)));
    E3(TWO_SIDED ? s.dp_edg : s.antifx, ipn, edg, k) = qs;
    if (vn > 0.) {
      c_num_out_sum = c_num_out_sum + cn;
    } else {
      const double qn = dmin(qs, b), qx = dmax(qs, b);
      c_num_in_sum = c_num_in_sum + cn;
      q_min_in_min = dmin(q_min_in_min, qn);
      q_max_in_max = dmax(q_max_in_max, qx);
      cq_min_in_sum = cq_min_in_sum + cn * qn;
      cq_max_in_sum = cq_max_in_sum + cn * qx;
    }
  }
  const double q_min_next = dmin(gv, q_min_in_min), q_max_next = dmax(gv, q_max_in_max);
  double qmo, qxo;
  if (c_num_out_sum != 0.) {
    const double at = ab.ab2 * to, av = ab.ab3 * tv;
    qmo = (gv + ab.ab1 / dt * cq_max_in_sum - q_max_next * (1.0 + c_num_in_sum - c_num_out_sum) + at + av) /
          c_num_out_sum / ab.ab1 * dt;
    qxo = (gv + ab.ab1 / dt * cq_min_in_sum - q_min_next * (1.0 + c_num_in_sum - c_num_out_sum) + at + av) /
          c_num_out_sum / ab.ab1 * dt;
  } else {
    qmo = -max_out_bound;
    qxo = max_out_bound;
  }
  __builtin_nontemporal_store(qmo, &C2(s.r_plus, ipn, k));
  __builtin_nontemporal_store(qxo, &C2(s.r_mnus, ipn, k));
}

// k_massadv2: the last limiter clamp of each edge value by the upwind
// column's outflow bounds (flux_limiter.F90:244-258), massflux_newdepth
// (massadv_compute.F90:22-70: mass flux, dp tendency, Adams-Bashforth depth
// update kept non-negative) and update_pres.  The surface pressure the
// reference sums there is a local of mass_adv that nothing reads, so it is
// not formed.
template <bool TWO_SIDED>
__global__ void __launch_bounds__(256) k_massadv2(DevGrid g, DevState s, AB ab) {
  __shared__ double sh_dp[COLS][NVL_MAX + 1];
  __shared__ double sh_lt[COLS][NVL_MAX + 1];
  __shared__ double sh_pr[COLS][NVL_MAX + 1];
  __shared__ double sh_om[COLS][NVL_MAX + 1];
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x, c = threadIdx.y;
  const int ipn = wave_col();
  const bool live = ipn < nip && k < nvl;
  const int nf = ab.nf - 1, of = ab.of - 1, vof = ab.vof - 1;
  if (live) {
    const int np = GNP(ipn);
    double acc = 0.;
    for (int edg = 0; edg < np; edg++) {
      const int ipx = PROX(ipn, edg);
      const double vn = SVE(ipn, edg, 1) * E3(s.u_edg, ipn, edg, k) - SVE(ipn, edg, 0) * E3(s.v_edg, ipn, edg, k);
      const int up = vn > 0. ? ipn : ipx;
      const double qs =
          dmax(C2(s.r_plus, up, k), dmin(C2(s.r_mnus, up, k), E3(TWO_SIDED ? s.dp_edg : s.antifx, ipn, edg, k)));
      __builtin_nontemporal_store(qs, &E3(s.dp_edg, ipn, edg, k));
      const double m = vn * qs;
      __builtin_nontemporal_store(m, &MF(s.massfx, nf, ipn, edg, k));
      acc = acc + m;
    }
    if (np < NPP) __builtin_nontemporal_store(0., &MF(s.massfx, nf, ipn, NPP - 1, k));
    const double dtn = -acc * GV(rarea, ipn);
    __builtin_nontemporal_store(dtn, &AB3(s.dp_tdcy, nf, ipn, k));
    const double to = AB3(s.dp_tdcy, of, ipn, k), tv = AB3(s.dp_tdcy, vof, ipn, k);
    const double d = C2(s.dp, ipn, k) + ab.ab1 * dtn + ab.ab2 * to + ab.ab3 * tv;
    const double dpn = dmax(d, 0.0);
    __builtin_nontemporal_store(dpn, &C2(s.dp, ipn, k));
    sh_dp[c][k] = dpn;
    sh_lt[c][k] = (ab.ab1 * dtn + ab.ab2 * to + ab.ab3 * tv) / g.dt;
  }
  update_pres_col(g, s, ipn, live, sh_dp[c], sh_lt[c], sh_pr[c], sh_om[c]);
}

// ---------------------------------------------------------------------------
// trcadv1-3 (trcadv_compute.F90:19-315); blockIdx.y = tracer.  Tracer 1's
// edge values come from edgvar1 (symmetric as stored); the others from
// edgvar3, stored averaged or per cell (TWO_SIDED as in k_edgvar3).
template <bool TWO_SIDED>
__global__ void __launch_bounds__(256) k_trcadv1(DevGrid g, DevState s, AB ab) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  const int t = blockIdx.y;
  if (ipn >= nip || k >= nvl) return;
  const int nf = ab.nf - 1, of = ab.of - 1, vof = ab.vof - 1;
  const double thshld = 1.e-11;
  const size_t esz = (size_t)nip * NPP * nvl, csz = (size_t)nip * nvl;
  const double *tre = s.trc_edg + t * esz;
  const double *tracr = s.tr + t * csz;
  double *antiflx = s.antiflx + t * esz;
  const int np = GNP(ipn);
  const double ti = C2(tracr, ipn, k);
  double acc = 0.;
  for (int edg = 0; edg < np; edg++) {
    const int ipx = PROX(ipn, edg), edx = PROXS(ipn, edg);
    double m = MF(s.massfx, nf, ipn, edg, k), mx = MF(s.massfx, nf, ipx, edx, k);
    double flx_hi = m * (t == 0 ? E3(tre, ipn, edg, k) : edge_val<TWO_SIDED>(tre, ipn, edg, ipx, edx, k, nvl));
    double fl = 0.5 * ((m + fabs(m)) * ti - (mx + fabs(mx)) * C2(tracr, ipx, k));
    E3(antiflx, ipn, edg, k) = flx_hi - fl;
    acc = acc + fl;
  }
  if (np < NPP) E3(antiflx, ipn, NPP - 1, k) = 0.;
  double tlo = -acc * GV(rarea, ipn);
  __builtin_nontemporal_store(tlo, &TAB(s.trl_tdcy, t, nf, ipn, k));
  double trcdp_lo = C2(s.trdp + t * csz, ipn, k) + ab.ab1 * tlo + ab.ab2 * TAB(s.trl_tdcy, t, of, ipn, k) +
                    ab.ab3 * TAB(s.trl_tdcy, t, vof, ipn, k);
  double mx = ti, mn = ti;
  for (int e = 0; e < NPP; e++) {
    double v = C2(tracr, PROX(ipn, e), k);
    mx = dmax(mx, v);
    mn = dmin(mn, v);
  }
  mx = dmax(0., mx);
  mn = dmax(0., mn);
  __builtin_nontemporal_store(mx, &C2(s.trmax + t * csz, ipn, k));
  __builtin_nontemporal_store(mn, &C2(s.trmin + t * csz, ipn, k));
  __builtin_nontemporal_store(dmax(mn, dmin(mx, trcdp_lo / dmax(thshld, C2(s.dp, ipn, k)))), &C2(s.trcr_lo + t * csz, ipn, k));
}

__global__ void __launch_bounds__(256) k_trcadv2(DevGrid g, DevState s, AB ab) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  const int t = blockIdx.y;
  if (ipn >= nip || k >= nvl) return;
  const int of = ab.of - 1, vof = ab.vof - 1;
  const double thshld = 1.e-11;
  const size_t esz = (size_t)nip * NPP * nvl, csz = (size_t)nip * nvl;
  const double *antiflx = s.antiflx + t * esz;
  const double *lo_ = s.trcr_lo + t * csz;
  const int np = GNP(ipn);
  double sp = 0., sm = 0.;
  for (int edg = 0; edg < np; edg++) {
    double a = E3(antiflx, ipn, edg, k);
    sp = sp - dmin(0., a);
    sm = sm + dmax(0., a);
  }
  const double lo = C2(lo_, ipn, k);
  double mx = dmax(C2(s.trmax + t * csz, ipn, k), lo);
  double mn = dmin(C2(s.trmin + t * csz, ipn, k), lo);
  for (int e = 0; e < NPP; e++) {
    double v = C2(lo_, PROX(ipn, e), k);
    mx = dmax(mx, v);
    mn = dmin(mn, v);
  }
  mx = dmax(0., mx);
  mn = dmax(0., mn);
  __builtin_nontemporal_store(mx, &C2(s.trmax + t * csz, ipn, k));
  __builtin_nontemporal_store(mn, &C2(s.trmin + t * csz, ipn, k));
  const double dp = C2(s.dp, ipn, k);
  const double to = TAB(s.trc_tdcy, t, of, ipn, k), tv = TAB(s.trc_tdcy, t, vof, ipn, k);
  double q_plus = ((mx - lo) * dp - (ab.ab2 * dmin(0., to) + ab.ab3 * dmax(0., tv))) / ab.ab1 * GV(area, ipn);
  double q_mnus = ((lo - mn) * dp + (ab.ab2 * dmax(0., to) + ab.ab3 * dmin(0., tv))) / ab.ab1 * GV(area, ipn);
  __builtin_nontemporal_store(dmax(0., dmin(1., q_plus / dmax(sp, thshld))), &C2(s.tr_rp + t * csz, ipn, k));
  __builtin_nontemporal_store(dmax(0., dmin(1., q_mnus / dmax(sm, thshld))), &C2(s.tr_rm + t * csz, ipn, k));
}

__global__ void __launch_bounds__(256) k_trcadv3(DevGrid g, DevState s, AB ab) {
  const int nip = g.nip, nvl = g.nvl;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  const int t = blockIdx.y;
  if (ipn >= nip || k >= nvl) return;
  const int nf = ab.nf - 1, of = ab.of - 1, vof = ab.vof - 1;
  const double thshld = 1.e-11;
  const size_t esz = (size_t)nip * NPP * nvl, csz = (size_t)nip * nvl;
  const double *antiflx = s.antiflx + t * esz;
  const double *rp = s.tr_rp + t * csz, *rm = s.tr_rm + t * csz;
  const int np = GNP(ipn);
  double at = 0.;
  for (int edg = 0; edg < np; edg++) {
    const int x = PROX(ipn, edg);
    double a = E3(antiflx, ipn, edg, k);
    if (a >= 0.) a = a * dmin(C2(rm, ipn, k), C2(rp, x, k));
    else a = a * dmin(C2(rp, ipn, k), C2(rm, x, k));
    at = at + a;
  }
  at = -at * GV(rarea, ipn);
  double tt = TAB(s.trl_tdcy, t, nf, ipn, k) + at;
  __builtin_nontemporal_store(tt, &TAB(s.trc_tdcy, t, nf, ipn, k));
  double *trcdp = s.trdp + t * csz;
  double tdp = C2(trcdp, ipn, k) + ab.ab1 * tt + ab.ab2 * TAB(s.trc_tdcy, t, of, ipn, k) +
               ab.ab3 * TAB(s.trc_tdcy, t, vof, ipn, k);
  C2(trcdp, ipn, k) = tdp;
  C2(s.tr + t * csz, ipn, k) = dmax(C2(s.trmin + t * csz, ipn, k),
                                    dmin(C2(s.trmax + t * csz, ipn, k), tdp / dmax(thshld, C2(s.dp, ipn, k))));
}

// ---------------------------------------------------------------------------
// momtum (momtum.F90:101-158; dissip momtum_compute.F90:102-229; momtum1
// 231-450).  Reads the winds from us/vs and writes the new winds to u2/v2,
// because dissip reads the neighbours' winds before any of them is updated.
__global__ void __launch_bounds__(256) k_momtum(DevGrid g, DevState s, AB ab) {
  __shared__ double sh_q[COLS][3];
  const int nip = g.nip, nvl = g.nvl, nvlp1 = nvl + 1;
  const int k = threadIdx.x, c = threadIdx.y;
  const int ipn = wave_col();
  const bool live = ipn < nip && k < nvl;
  const int nf = ab.nf - 1, of = ab.of - 1, vof = ab.vof - 1;
  const double thshld = 1.e-11;
  double un = 0., vn = 0.;
  if (live) {
    const int np = GNP(ipn);
    const double wui = C2(s.us, ipn, k) + GV(grndspd, ipn), wvi = C2(s.vs, ipn, k);
    const double dpi = C2(s.dp, ipn, k);
    const double factor = GV(rarea, ipn) / dmax(dpi, thshld);
    double wu = 0., wv = 0.;
    for (int edg = 0; edg < np; edg++) {
      const int ipx = PROX(ipn, edg);
      const double wux = C2(s.us, ipx, k) + GV(grndspd, ipx), wvx = C2(s.vs, ipx, k);
      double uxy1 = CS(ipn, edg, 0) * wui + SN(ipn, edg, 0) * wvi;
      double vxy1 = -SN(ipn, edg, 0) * wui + CS(ipn, edg, 0) * wvi;
      double uxy2 = CS(ipn, edg, 1) * wux + SN(ipn, edg, 1) * wvx;
      double vxy2 = -SN(ipn, edg, 1) * wux + CS(ipn, edg, 1) * wvx;
      double a = dmax(dpi, thshld), b = dmax(C2(s.dp, ipx, k), thshld);
      double hh = a * b / (a + b);
      double ufx = (uxy2 - uxy1) * g.dfflen[k] * GV(sideln, ipn * NPP + edg) * 2. * hh;
      double vfx = (vxy2 - vxy1) * g.dfflen[k] * GV(sideln, ipn * NPP + edg) * 2. * hh;
      double uflux = CS(ipn, edg, 0) * ufx - SN(ipn, edg, 0) * vfx;
      double vflux = SN(ipn, edg, 0) * ufx + CS(ipn, edg, 0) * vfx;
      wu = wu + uflux * factor;
      wv = wv + vflux * factor;
    }
    const double u1 = C2(s.us, ipn, k) + wu, v1 = C2(s.vs, ipn, k) + wv;
    double px = 0., py = 0., rv = 0.;
    const double exm = .5 * (L2(s.ex, ipn, k) + L2(s.ex, ipn, k + 1));
    for (int edg = 0; edg < np; edg++) {
      double b = E3(s.mont_edg, ipn, edg, k) + E3(s.uvsq_edg, ipn, edg, k);
      double te = E3(s.trc_edg, ipn, edg, k);
      px = px - b * SVC(ipn, edg, 1) + exm * te * SVC(ipn, edg, 1);
      py = py + b * SVC(ipn, edg, 0) - exm * te * SVC(ipn, edg, 0);
      rv = rv + (SVE(ipn, edg, 0) * E3(s.u_edg, ipn, edg, k) + SVE(ipn, edg, 1) * E3(s.v_edg, ipn, edg, k));
    }
    px = px * GV(rarea, ipn);
    py = py * GV(rarea, ipn);
    rv = rv * GV(rarea, ipn);
    __builtin_nontemporal_store(rv, &C2(s.relvor, ipn, k));
    double uzeta = (GV(corio, ipn) + rv) * u1;
    double vzeta = (GV(corio, ipn) + rv) * v1;
    double ut = px + vzeta, vt = py - uzeta;
    __builtin_nontemporal_store(ut, &AB3(s.u_tdcy, nf, ipn, k));
    __builtin_nontemporal_store(vt, &AB3(s.v_tdcy, nf, ipn, k));
    un = u1 + ab.ab1 * ut + ab.ab2 * AB3(s.u_tdcy, of, ipn, k) + ab.ab3 * AB3(s.u_tdcy, vof, ipn, k);
    vn = v1 + ab.ab1 * vt + ab.ab2 * AB3(s.v_tdcy, of, ipn, k) + ab.ab3 * AB3(s.v_tdcy, vof, ipn, k);
    if (k >= nvl - 3) sh_q[c][nvl - 1 - k] = un * un + vn * vn;
  }
  __syncthreads();
  if (live) {
    // Rayleigh damping of the top layers (momtum_compute.F90:406-416)
    double qq = dmax(dmax(sh_q[c][0], sh_q[c][1]), sh_q[c][2]) / (RLTHRESH * RLTHRESH);
    const int ndamp = (int)(0.25 * nvl);
    const int k1 = k + 1;
    if (qq > 1. && k1 >= nvl - ndamp) {
      qq = sqrt((qq - 1.) * qq);
      double factor = 1. - (g.dt / 86400.) * RLEIGH * qq * fim_pow(2., RLTAPER * (double)(k1 - nvl));
      un = un * factor;
      vn = vn * factor;
    }
    __builtin_nontemporal_store(un, &C2(s.u2, ipn, k));
    __builtin_nontemporal_store(vn, &C2(s.v2, ipn, k));
  }
}

// ---------------------------------------------------------------------------
// hystat.F90:23-115 (column part), one wavefront per column: the column is
// staged in LDS with coalesced loads, lane 0 runs the serial surface-pressure
// and Montgomery-potential sums in the reference order, and every lane then
// forms its level's geopotential.
__global__ void __launch_bounds__(256) k_hystat(DevGrid g, DevState s, int its) {
  __shared__ double sh_ex[COLS][NVL_MAX + 1];
  __shared__ double sh_th[COLS][NVL_MAX];
  __shared__ double sh_dp[COLS][NVL_MAX];
  __shared__ double sh_mp[COLS][NVL_MAX];
  const int nip = g.nip, nvl = g.nvl, nvlp1 = nvl + 1;
  const int k = threadIdx.x, c = threadIdx.y;
  const int ipn = wave_col();
  const bool live = ipn < nip;
  if (live) {
    for (int kk = k; kk < nvlp1; kk += NVL_MAX) sh_ex[c][kk] = L2(s.ex, ipn, kk);
    if (k < nvl) {
      sh_th[c][k] = TR(s.tr, 0, ipn, k);
      sh_dp[c][k] = C2(s.dp, ipn, k);
    }
  }
  __syncthreads();
  if (live && k == 0) {
    double work = s.psrf[ipn];
    double ps = sh_dp[c][nvl - 1];
    for (int kk = nvl - 1; kk >= 1; kk--) ps = ps + sh_dp[c][kk - 1];
    s.psrf[ipn] = ps;
    if (its > 0) s.ptdcy[(size_t)(its % 2) * nip + ipn] = ps - work;
    double mp = sh_ex[c][0] * sh_th[c][0] + L2(s.ph, ipn, 0);
    sh_mp[c][0] = mp;
    for (int kk = 2; kk <= nvl; kk++) {
      mp = mp + sh_ex[c][kk - 1] * (sh_th[c][kk - 1] - sh_th[c][kk - 2]);
      sh_mp[c][kk - 1] = mp;
    }
  }
  __syncthreads();
  if (live && k < nvl) {
    const double mp = sh_mp[c][k];
    __builtin_nontemporal_store(mp, &C2(s.mp, ipn, k));
    __builtin_nontemporal_store(mp - sh_th[c][k] * sh_ex[c][k + 1], &L2(s.ph, ipn, k + 1));
  }
}

// tracer floors and tracer*thickness (hystat.F90:82-95), per cell-level
__global__ void __launch_bounds__(256) k_hystat_tr(DevGrid g, DevState s) {
  const int nip = g.nip, nvl = g.nvl, ntra = g.ntra;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  if (ntra > 1) TR(s.tr, 1, ipn, k) = dmax(QVMIN, TR(s.tr, 1, ipn, k));
  if (ntra > 2) TR(s.tr, 2, ipn, k) = dmax(QWMIN, TR(s.tr, 2, ipn, k));
  const double dp = C2(s.dp, ipn, k);
  for (int t = 0; t < ntra; t++) TR(s.trdp, t, ipn, k) = TR(s.tr, t, ipn, k) * dp;
}

// pressure from the Exner function after the start-up hybgen (dyn_init.F90:884-893)
__global__ void __launch_bounds__(256) k_pres_from_exner(DevGrid g, DevState s) {
  const int nip = g.nip, nvl = g.nvl, nvlp1 = nvl + 1;
  const int ipn = wave_col();
  if (ipn >= nip) return;
  for (int k = threadIdx.x; k < nvlp1; k += blockDim.x)
    __builtin_nontemporal_store(P1000 * fim_pow(L2(s.ex, ipn, k) / CP, CP / RD), &L2(s.pr, ipn, k));
}

__global__ void __launch_bounds__(256) k_trdp(DevGrid g, DevState s) {
  const int nip = g.nip, nvl = g.nvl, ntra = g.ntra;
  const int k = threadIdx.x;
  const int ipn = wave_col();
  if (ipn >= nip || k >= nvl) return;
  for (int t = 0; t < ntra; t++) TR(s.trdp, t, ipn, k) = TR(s.tr, t, ipn, k) * C2(s.dp, ipn, k);
}

// ---------------------------------------------------------------------------
void launch_edgvar1(const DevGrid &g, const DevState &s, hipStream_t st) {
  KT_LAUNCH("k_edgvar1", st, k_edgvar1<<<col_grid(g.nip), col_block(), 0, st>>>(g, s));
}
void launch_edgvar3(const DevGrid &g, const DevState &s, hipStream_t st, bool two_sided) {
  dim3 grid((g.nip + COLS - 1) / COLS, g.ntra);
  if (two_sided) KT_LAUNCH("k_edgvar3", st, k_edgvar3<<<grid, col_block(), 0, st>>>(g, s));
  else KT_LAUNCH("k_edgvar3", st, k_edgvar3_own<<<col_grid(g.nip), col_block(), 0, st>>>(g, s));
}
void launch_cnuity(const DevGrid &g, const DevState &s, const AB &ab, hipStream_t st, bool two_sided) {
  if (two_sided) KT_LAUNCH("k_cnuity12", st, k_cnuity12<true><<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
  else KT_LAUNCH("k_cnuity12", st, k_cnuity12<false><<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
  KT_LAUNCH("k_cnuity3", st, k_cnuity3<<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
  KT_LAUNCH("k_cnuity45", st, k_cnuity45<<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
}
void launch_mass_adv(const DevGrid &g, const DevState &s, const AB &ab, hipStream_t st, bool two_sided) {
  if (two_sided) {
    KT_LAUNCH("k_massadv1", st, k_massadv1<true><<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
    KT_LAUNCH("k_massadv2", st, k_massadv2<true><<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
  } else {
    KT_LAUNCH("k_massadv1", st, k_massadv1<false><<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
    KT_LAUNCH("k_massadv2", st, k_massadv2<false><<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
  }
}
void launch_trcadv(const DevGrid &g, const DevState &s, const AB &ab, hipStream_t st, bool two_sided) {
  dim3 grid((g.nip + COLS - 1) / COLS, g.ntra);
  if (two_sided) KT_LAUNCH("k_trcadv1", st, k_trcadv1<true><<<grid, col_block(), 0, st>>>(g, s, ab));
  else KT_LAUNCH("k_trcadv1", st, k_trcadv1<false><<<grid, col_block(), 0, st>>>(g, s, ab));
  KT_LAUNCH("k_trcadv2", st, k_trcadv2<<<grid, col_block(), 0, st>>>(g, s, ab));
  KT_LAUNCH("k_trcadv3", st, k_trcadv3<<<grid, col_block(), 0, st>>>(g, s, ab));
}
void launch_momtum(const DevGrid &g, DevState &s, const AB &ab, hipStream_t st) {
  KT_LAUNCH("k_momtum", st, k_momtum<<<col_grid(g.nip), col_block(), 0, st>>>(g, s, ab));
  double *t = s.us; s.us = s.u2; s.u2 = t;
  t = s.vs; s.vs = s.v2; s.v2 = t;
}
void launch_hystat(const DevGrid &g, const DevState &s, int its, hipStream_t st) {
  KT_LAUNCH("k_hystat", st, k_hystat<<<col_grid(g.nip), col_block(), 0, st>>>(g, s, its));
  KT_LAUNCH("k_hystat_tr", st, k_hystat_tr<<<col_grid(g.nip), col_block(), 0, st>>>(g, s));
}
void launch_pres_from_exner(const DevGrid &g, const DevState &s, hipStream_t st) {
  KT_LAUNCH("k_pres_from_exner", st, k_pres_from_exner<<<col_grid(g.nip), col_block(), 0, st>>>(g, s));
}
void launch_trdp(const DevGrid &g, const DevState &s, hipStream_t st) {
  KT_LAUNCH("k_trdp", st, k_trdp<<<col_grid(g.nip), col_block(), 0, st>>>(g, s));
}

}  // namespace fimhip
