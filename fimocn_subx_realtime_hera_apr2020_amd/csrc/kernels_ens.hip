// Kernels of the batched mixed-precision form (precision split and layout in
// ens_internal.h).  Each kernel follows the same reference routine as its
// FP64 counterpart in kernels_plain.hip; the differences are the arithmetic
// type, the member index (blockIdx.y), the one-sided mass and tracer fluxes
// that make the FP32 fluxes cancel exactly, and parallel column scans in
// place of the lane-0 serial sums (the sums stay FP64).
//
// Mapping: one wavefront per column, lane = level; COLS columns per block;
// blockIdx.y = member, blockIdx.z = tracer where a kernel loops over tracers.
#include "ens_internal.h"
#include "fim_pow.h"

namespace fimhip {

const char *const ens_kernel_names[EK_NKERNELS] = {
    "e_edgvar1", "e_edgvar3", "e_cnuity12", "e_cnuity3", "e_cnuity45",
    "e_trcadv1", "e_trcadv2", "e_trcadv3", "e_momtum", "e_hystat", "e_massadv1", "e_massadv2", "e_trc_fct"};

#define EC(a, col, k) (a)[(size_t)(col) * nvl + (k)]
#define EL(a, col, k) (a)[(size_t)(col) * nvlp1 + (k)]
#define EE(a, col, e, k) (a)[((size_t)(col) * NPP + (e)) * nvl + (k)]
#define EAB(a, sl, col, k) (a)[((size_t)(sl) * ncol + (col)) * nvl + (k)]
#define ETR(a, t, col, k) (a)[((size_t)(t) * ncol + (col)) * nvl + (k)]
#define ETAB(a, t, sl, col, k) (a)[(((size_t)(t) * 3 + (sl)) * ncol + (col)) * nvl + (k)]
#define ETE(a, t, col, e, k) (a)[(((size_t)(t) * ncol + (col)) * NPP + (e)) * nvl + (k)]
#define GPROX(i, e) ldc(g.d.prox, (i) * NPP + (e))
#define GPROXS(i, e) ldc(g.d.proxs, (i) * NPP + (e))
#define FCS(i, e, j) ldc(g.cs, ((i) * NPP + (e)) * 4 + (j))
#define FSN(i, e, j) ldc(g.sn, ((i) * NPP + (e)) * 4 + (j))
#define FSVC(i, e, j) ldc(g.svc, ((i) * NPP + (e)) * 2 + (j))
#define FSVE(i, e, j) ldc(g.sve, ((i) * NPP + (e)) * 2 + (j))
#define FHF(i, a, b) ldc(g.hfcoef, ((i) * NPP + (a)) * NPP + (b))
#define GNP(i) ldc(g.d.nprox, (i))
#define GF(a, i) ldc(g.a, (i))
#define GD(a, i) ldc(g.d.a, (i))

// The cell a block row works on, wave-uniform.  With SWZ the block index is
// permuted so that the blocks the hardware sends to XCD x (block ids
// congruent to x mod 8) cover one contiguous eighth of the cells; the host
// pads gridDim.x to a multiple of 8 for it.
template <int COLS, bool SWZ>
__device__ __forceinline__ int tile_cell() {
  int bx = blockIdx.x;
  if (SWZ) {
    const int per = gridDim.x >> 3;
    bx = (bx & 7) * per + (bx >> 3);
  }
  return __builtin_amdgcn_readfirstlane(bx * COLS + (int)threadIdx.y);
}

__device__ __forceinline__ double wave_suffix_sum(double x, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_down(x, d, 64);
    if (lane + d < 64) x += y;
  }
  return x;
}
__device__ __forceinline__ double wave_prefix_sum(double x, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// The edge-wind, edgvar1 and momtum bodies are compiled with a*b+c
// contracted to fused multiply-adds (one rounding instead of two): none of
// their results has to match another cell's bit for bit.  The flux code of
// cnuity and trcadv, where the two cells of an edge must form exactly
// opposite fluxes, and the shared theta edge value keep separate roundings.
//
// Hardware square root and reciprocal (v_sqrt_f32, v_rcp_f32: about 1 ulp)
// in place of the correctly rounded sequences of sqrtf and '/', which cost
// about ten instructions each; every reciprocal below is of a value bounded
// away from zero (a threshold or eps under fmaxf).
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// The edge wind of cell c at its slot e (edgvar1.F90:162-421, eqwgt): the
// winds of the cell (i), the cell across the edge (x) and the two cells at
// the edge's ends (t: slot e-1, l: slot e+1) rotated into the edge frame with
// cell c's coefficients, trapezoid-weighted, the magnitude rescaled to the
// weighted mean of the four magnitudes.  edgvar1 (normal component) and
// momtum (kinetic energy, vorticity) both form edge winds through this one
// function.  The reference rescales by the magnitudes of the four rotated
// winds; the rotation coefficients of a stencil cell satisfy cs^2 + sn^2 = 1
// (to float rounding, within 3.4e-7 on the grid), so those are the
// magnitudes of the cells' own winds (mi, mx, mt, ml), which the callers form
// once per cell rather than once per edge and stencil position.
__device__ __forceinline__ float2 e_edge_wind(const EnsGrid &g, int c, int e, float ui, float vi, float ux, float vx,
                                              float ut, float vt, float ul, float vl, float mi, float mx, float mt,
                                              float ml) {
#pragma clang fp contract(fast)
  const float d6 = 1.f / 6.f, eps = 0.001f;
  const float c0 = FCS(c, e, 0), s0 = FSN(c, e, 0), c1 = FCS(c, e, 1), s1 = FSN(c, e, 1);
  const float c2 = FCS(c, e, 2), s2 = FSN(c, e, 2), c3 = FCS(c, e, 3), s3 = FSN(c, e, 3);
  const float u3trl = c2 * ut + s2 * vt, v3trl = -s2 * ut + c2 * vt;
  const float u1 = c0 * ui + s0 * vi, v1 = -s0 * ui + c0 * vi;
  const float u2 = c1 * ux + s1 * vx, v2 = -s1 * ux + c1 * vx;
  const float u3l = c3 * ul + s3 * vl, v3l = -s3 * ul + c3 * vl;
  float ue = (2.f * (u1 + u2) + u3trl + u3l) * d6;
  float ve = (2.f * (v1 + v2) + v3trl + v3l) * d6;
  const float m1 = fsqrt(ue * ue + ve * ve);
  const float m2 = (2.f * (mi + mx) + mt + ml) * d6;
  const float sc = m2 * frcp(fmaxf(eps, m1));
  return make_float2(ue * sc, ve * sc);
}

// Theta on the edge between cell i and x (edgvar1.F90:425-511, eqwgt): the
// trapezoid (2(i + x) + t + l)/6 with the two outer stencil cells t, l (slots
// e-1, e+1 of i) added in column order, so both cells of the edge form the
// same bits.
__device__ __forceinline__ float e_theta_edge(float thi, float thx, float tht, float thl, int nbt, int nbl) {
  float ta = tht, tb = thl;
  if (nbt > nbl) { const float t = ta; ta = tb; tb = t; }
  return (2.f * (thi + thx) + ta + tb) * (1.f / 6.f);
}

// ---------------------------------------------------------------------------
// edgvar1 (edgvar1.F90:162-421, 425-511): the edge pressure, and of the
// edge wind only its normal component sidevec_e x (u,v), the one cnuity1
// uses (cnuity_compute.F90:17-95); momtum forms the edge winds it needs
// itself (e_edge_wind).  Theta's edge values are formed where they are used
// (trcadv1, momtum: e_theta_edge) from the cells' theta; edgvar1 keeps the
// step's starting theta in th0 for momtum.
// cnuity5's pressure gradient (cnuity_compute.F90:421-428): the line integral
// of the mid-layer edge pressures lp_edg along the cell's edge vectors.  It
// depends only on the pressure at the start of the step, so edgvar1 forms it
// once per cell-level (two values) where it forms the edge pressures, instead
// of storing the six edge pressures for cnuity5; separate roundings, in the
// reference's edge order.
template <int NP>
__device__ __forceinline__ float2 e_pres_grad(const EnsGrid &g, int ipn, const float *lp) {
#pragma clang fp contract(off)
  float dpdx = 0.f, dpdy = 0.f;
#pragma unroll
  for (int e = 0; e < NP; e++) {
    dpdx = dpdx + lp[e] * FSVC(ipn, e, 1);
    dpdy = dpdy - lp[e] * FSVC(ipn, e, 0);
  }
  return make_float2(dpdx, dpdy);
}

template <int NP>
__device__ __forceinline__ void e_edgvar1_col(const EnsGrid &g, const EnsState &s, int ipn, int base, int k) {
#pragma clang fp contract(fast)
  const int nvl = g.d.nvl, nvlp1 = nvl + 1, ncol = s.ncol;
  const int col = base + ipn;
  int nb[NP];
  float nu[NP], nv[NP];
  double np0[NP], np1[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) {
    nb[e] = base + GPROX(ipn, e);
    nu[e] = EC(s.us, nb[e], k);
    nv[e] = EC(s.vs, nb[e], k);
    np0[e] = EL(s.pr, nb[e], k);
    np1[e] = EL(s.pr, nb[e], k + 1);
  }
  const float ui = EC(s.us, col, k), vi = EC(s.vs, col, k), thi = ETR(s.tr, 0, col, k);
  const double pi0 = EL(s.pr, col, k), pi1 = EL(s.pr, col, k + 1);
  const float mi = fsqrt(ui * ui + vi * vi);
  float mn[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) mn[e] = fsqrt(nu[e] * nu[e] + nv[e] * nv[e]);
  float lp[NP];
  const bool wc = col_wet(s.wet, col);
#pragma unroll
  for (int e = 0; e < NP; e++) {
    const int et = (e + NP - 1) % NP, el = (e + 1) % NP;
    const float2 w = e_edge_wind(g, ipn, e, ui, vi, nu[e], nv[e], nu[et], nv[et], nu[el], nv[el], mi, mn[e], mn[et],
                                 mn[el]);
    const float vnrm = FSVE(ipn, e, 1) * w.x - FSVE(ipn, e, 0) * w.y;
    __builtin_nontemporal_store((wc && col_wet(s.wet, nb[e])) ? vnrm : 0.f, &EE(s.vn_edg, col, e, k));   // no flux through a land edge
    lp[e] = (float)((pi0 + pi1 + np0[e] + np1[e] + .5 * (np0[et] + np1[et] + np0[el] + np1[el])) * (1. / 6.));
  }
  const float2 pg = e_pres_grad<NP>(g, ipn, lp);
  __builtin_nontemporal_store(pg.x, &EC(s.pgx, col, k));
  __builtin_nontemporal_store(pg.y, &EC(s.pgy, col, k));
  // theta at the start of the step, for momtum's edge values (the tracer
  // itself is updated by trcadv before momtum runs)
  __builtin_nontemporal_store(thi, &EC(s.th0, col, k));
}

template <int COLS, bool SWZ>
// waves_per_eu: the register allocation otherwise lands just above an
// occupancy step (84 VGPRs, 5 waves per SIMD); 80 VGPRs give 6, no spills
__global__ void __launch_bounds__(64 * COLS) __attribute__((amdgpu_waves_per_eu(6))) k_e_edgvar1(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_edgvar1_col<6>(g, s, ipn, base, k);
  else e_edgvar1_col<5>(g, s, ipn, base, k);
}

// edgvar3 (edgvar3.F90:65-110, edgvalEval): this cell's least-squares
// reconstruction of each edge.  The two-sided average of edgvar3.F90:111-132
// is formed by the readers (cnuity, trcadv) from both cells' values; the
// tracers' sums are fused multiply-adds in stencil order, as the fused tracer
// transport forms them for both cells of an edge.  The
// thickness is reconstructed as a deviation from the cell's own value, the
// differences taken in FP64.
template <int NP>
__device__ __forceinline__ void e_edgvar3_col(const EnsGrid &g, const EnsState &s, int ipn, int base, int k) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int col = base + ipn;
  float hf[NP][NP];
  int nb[NP];
#pragma unroll
  for (int i = 0; i < NP; i++) {
    nb[i] = base + GPROX(ipn, i);
#pragma unroll
    for (int j = 0; j < NP; j++) hf[i][j] = FHF(ipn, i, j);
  }
  float d[NP], out[NP];
  {
    const double b0 = EC(s.dp, col, k);
#pragma unroll
    for (int j = 0; j < NP; j++) d[j] = (float)(EC(s.dp, nb[j], k) - b0);
#pragma unroll
    for (int i = 0; i < NP; i++) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < NP; j++) acc = acc + d[j] * hf[i][j];
      out[i] = (float)(b0 + (double)acc);
    }
#pragma unroll
    for (int i = 0; i < NP; i++) __builtin_nontemporal_store(out[i], &EE(s.dp_edg, col, i, k));
    if (NP < NPP) __builtin_nontemporal_store(0.f, &EE(s.dp_edg, col, NPP - 1, k));
  }
  for (int f = 1; f <= s.edg3_ntr; f++) {
    const float b0 = ETR(s.tr, f, col, k);
#pragma unroll
    for (int j = 0; j < NP; j++) d[j] = ETR(s.tr, f, nb[j], k) - b0;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < NP; j++) acc = __builtin_fmaf(d[j], hf[i][j], acc);
      out[i] = b0 + acc;
    }
#pragma unroll
    for (int i = 0; i < NP; i++) __builtin_nontemporal_store(out[i], &ETE(s.trc_edg, f - 1, col, i, k));
    if (NP < NPP) __builtin_nontemporal_store(0.f, &ETE(s.trc_edg, f - 1, col, NPP - 1, k));
  }
}

template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_edgvar3(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_edgvar3_col<6>(g, s, ipn, base, k);
  else e_edgvar3_col<5>(g, s, ipn, base, k);
}

// ---------------------------------------------------------------------------
// cnuity1 + cnuity2 (cnuity_compute.F90:17-150).  The high-order flux of an
// edge is formed by the cell with the lower index and negated by the other;
// the low-order (donor-cell) flux is antisymmetric as written.  Both cells
// store their (negated) copies of the fluxes: storing them on the owner only
// and reading the owner's slot from the other cell halves the writes but
// measured slower overall (the scattered single-use reads stream worse than
// each cell's own slots).
template <int NP>
__device__ __forceinline__ void e_cnuity12_col(const EnsGrid &g, const EnsState &s, const StepCtl &c, int ipn,
                                               int base, int k) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const int col = base + ipn;
  const double dpi = EC(s.dp, col, k);
  const float dpf = (float)dpi;
  const double lo_of = (double)EAB(s.dpl_tdcy, of, col, k), lo_vof = (double)EAB(s.dpl_tdcy, vof, col, k);
  int x[NP];
  float vn[NP], vx[NP], de[NP], dx[NP], dpx[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) {
    x[e] = GPROX(ipn, e);
    const int xc = base + x[e], ex = GPROXS(ipn, e);
    vn[e] = EE(s.vn_edg, col, e, k);
    vx[e] = EE(s.vn_edg, xc, ex, k);
    de[e] = EE(s.dp_edg, col, e, k);
    dx[e] = EE(s.dp_edg, xc, ex, k);
    dpx[e] = (float)EC(s.dp, xc, k);
  }
  double acc = 0.;
  float pp = 0.f, pm = 0.f;     // cnuity3's sums of the antidiffusive in- and outflows
#pragma unroll
  for (int e = 0; e < NP; e++) {
    const float dpe = .5f * (de[e] + dx[e]);
    const float hi = ipn < x[e] ? vn[e] * dpe : -(vx[e] * dpe);
    const float m = .5f * ((vn[e] + fabsf(vn[e])) * dpf - (vx[e] + fabsf(vx[e])) * dpx[e]);
    const float a = hi - m;
    __builtin_nontemporal_store(m, &EE(s.massfx, col, e, k));
    __builtin_nontemporal_store(a, &EE(s.antifx, col, e, k));
    if (a <= 0.f) pp = pp - a; else pm = pm + a;
    acc = acc + (double)m;
  }
  __builtin_nontemporal_store(pp, &EC(s.r_plus, col, k));
  __builtin_nontemporal_store(pm, &EC(s.r_mnus, col, k));
  if (NP < NPP) {
    __builtin_nontemporal_store(0.f, &EE(s.massfx, col, NPP - 1, k));
    __builtin_nontemporal_store(0.f, &EE(s.antifx, col, NPP - 1, k));
  }
  const double dplo = -acc * GD(rarea, ipn);
  EAB(s.dpl_tdcy, nf, col, k) = (float)dplo;
  __builtin_nontemporal_store((float)(dpi + c.ab1 * dplo + c.ab2 * lo_of + c.ab3 * lo_vof), &EC(s.delp_lo, col, k));
}

template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_cnuity12(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const StepCtl c = load_ctl(s.ctl);
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_cnuity12_col<6>(g, s, c, ipn, base, k);
  else e_cnuity12_col<5>(g, s, c, ipn, base, k);
}

// cnuity3 (cnuity_compute.F90:152-259); the in- and outflow sums were left
// in r_plus/r_mnus by cnuity12
template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_cnuity3(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl, ncol = s.ncol;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const StepCtl c = load_ctl(s.ctl);
  const int of = c.of - 1, vof = c.vof - 1;
  const int base = blockIdx.y * nip, col = base + ipn;
  const float thshld = 1.e-11f;
  const float pp = EC(s.r_plus, col, k), pm = EC(s.r_mnus, col, k);
  const float lo = EC(s.delp_lo, col, k), dpi = (float)EC(s.dp, col, k);
  float mx = fmaxf(lo, dpi), mn = fminf(lo, dpi);
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const int xc = base + GPROX(ipn, e);
    const float a = EC(s.delp_lo, xc, k), b = (float)EC(s.dp, xc, k);
    mx = fmaxf(mx, fmaxf(a, b));
    mn = fminf(mn, fminf(a, b));
  }
  mx = fmaxf(0.f, mx);
  mn = fmaxf(0.f, mn);
  const double to = EAB(s.dp_tdcy, of, col, k), tv = EAB(s.dp_tdcy, vof, col, k);
  const double rab1 = 1. / c.ab1;
  const float qp = (float)(((double)mx - lo - (c.ab2 * fmin(0., to) + c.ab3 * fmax(0., tv))) * rab1) * GF(area, ipn);
  const float qm = (float)(((double)lo - mn + (c.ab2 * fmax(0., to) + c.ab3 * fmin(0., tv))) * rab1) * GF(area, ipn);
  __builtin_nontemporal_store(fmaxf(0.f, fminf(1.f, qp * frcp(fmaxf(pp, thshld)))), &EC(s.r_plus, col, k));
  __builtin_nontemporal_store(fmaxf(0.f, fminf(1.f, qm * frcp(fmaxf(pm, thshld)))), &EC(s.r_mnus, col, k));
}

// cnuity5 (cnuity_compute.F90:376-451), also update_pres of mass_adv
// (massadv_compute.F90:73-171): from the new thickness dpn and the layer
// tendency lt of each lane, the top-down pressure and the column tendency as
// wavefront suffix sums (FP64), then the Exner function and omega.  Every
// lane of the wavefront calls it (lanes at or above nvl with live = false).
template <int NP>
__device__ __forceinline__ void e_update_pres_col(const EnsGrid &g, const EnsState &s, int ipn, int col, int k,
                                                  bool live, double dpn, double lt) {
  const int nvl = g.d.nvl, nvlp1 = nvl + 1;
  const double sp = wave_suffix_sum(dpn, k);
  const double sl = wave_suffix_sum(lt, k);
  if (!live) return;
  const float dpdx = EC(s.pgx, col, k), dpdy = EC(s.pgy, col, k);   // formed by edgvar1
  const float ui = EC(s.us, col, k), vi = EC(s.vs, col, k);
  const double p = EL(s.pr, col, nvl) + sp;
  __builtin_nontemporal_store(p, &EL(s.pr, col, k));
  __builtin_nontemporal_store(CP * fim_pow_fast(p / P1000, RD / CP), &EL(s.ex, col, k));
  const float om = (float)(sl - .5 * lt);
  __builtin_nontemporal_store(om + (ui * dpdx + vi * dpdy) * GF(rarea, ipn), &EC(s.ws, col, k));
}

// cnuity4 + cnuity5 (cnuity_compute.F90:261-451).  The thickness tendency is
// formed in FP64 from the final (antisymmetric) FP32 fluxes, the thickness is
// advanced in FP64, and the top-down pressure sum and the column tendency for
// omega are wavefront suffix sums.
template <int NP>
__device__ __forceinline__ void e_cnuity45_col(const EnsGrid &g, const EnsState &s, const StepCtl &c, int ipn,
                                               int base, int k) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const int col = base + ipn;
  const bool live = k < nvl;
  double dpn = 0., lt = 0.;
  if (live) {
    const float rp = EC(s.r_plus, col, k), rm = EC(s.r_mnus, col, k);
    const double t_of = EAB(s.dp_tdcy, of, col, k), t_vof = EAB(s.dp_tdcy, vof, col, k), dp0 = EC(s.dp, col, k);
    float a[NP], m[NP], r[NP];
#pragma unroll
    for (int e = 0; e < NP; e++) {
      a[e] = EE(s.antifx, col, e, k);
      m[e] = EE(s.massfx, col, e, k);
    }
#pragma unroll
    for (int e = 0; e < NP; e++) {
      const float *src = a[e] >= 0.f ? s.r_plus : s.r_mnus;
      r[e] = src[(size_t)(base + GPROX(ipn, e)) * nvl + k];
    }
    double at = 0.;
#pragma unroll
    for (int e = 0; e < NP; e++) {
      const float mf = m[e] + a[e] * fminf(a[e] >= 0.f ? rm : rp, r[e]);
      __builtin_nontemporal_store(mf, &EE(s.massfx, col, e, k));
      at = at + (double)mf;
    }
    const double dtn = -at * GD(rarea, ipn);
    EAB(s.dp_tdcy, nf, col, k) = dtn;
    const double inc = c.ab1 * dtn + c.ab2 * t_of + c.ab3 * t_vof;
    dpn = dp0 + inc;
    __builtin_nontemporal_store(dpn, &EC(s.dp, col, k));
    lt = inc / g.d.dt;
  }
  e_update_pres_col<NP>(g, s, ipn, col, k, live, dpn, lt);
}

template <int COLS, bool SWZ>
// waves_per_eu: 106 SGPRs would hold occupancy at 7 waves per SIMD; the
// request fits 76 SGPRs and 42 VGPRs (8 waves), no spills
__global__ void __launch_bounds__(64 * COLS) __attribute__((amdgpu_waves_per_eu(8))) k_e_cnuity45(EnsGrid g, EnsState s) {
  const int nip = g.d.nip;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip) return;
  const StepCtl c = load_ctl(s.ctl);
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_cnuity45_col<6>(g, s, c, ipn, base, k);
  else e_cnuity45_col<5>(g, s, c, ipn, base, k);
}

// ---------------------------------------------------------------------------
// mass_adv, the flux-limited mass transport of flux_conserv_schm = 2
// (massadv.F90:21-160; flux_limiter.F90:78-260, M1 form;
// massadv_compute.F90:22-171).  The normal wind of an edge is the lower
// index cell's, negated on the other side, and the edge depth is the average
// of both cells' reconstructions clamped to the four cells around the edge
// (q_star: the same bits from both sides), so the final limited edge depth,
// taken from the upwind cell's bounds, is the same on both sides and the
// FP32 mass flux vnorm * depth cancels exactly across the edge.
//
// massadv1: q_star (kept in antifx) and the column's outflow bounds
// q_min_out / q_max_out (in r_plus / r_mnus).  The inflow sums and the bound
// formula are FP64.
template <int NP>
__device__ __forceinline__ void e_massadv1_col(const EnsGrid &g, const EnsState &s, const StepCtl &c, int ipn,
                                               int base, int k) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int of = c.of - 1, vof = c.vof - 1;
  const int col = base + ipn;
  const double gv = EC(s.dp, col, k);
  const float gvf = (float)gv;
  const double t_of = EAB(s.dp_tdcy, of, col, k), t_vof = EAB(s.dp_tdcy, vof, col, k);
  int x[NP];
  float vn[NP], vx[NP], de[NP], dx[NP], dpx[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) {
    x[e] = GPROX(ipn, e);
    const int xc = base + x[e], ex = GPROXS(ipn, e);
    vn[e] = EE(s.vn_edg, col, e, k);
    vx[e] = EE(s.vn_edg, xc, ex, k);
    de[e] = EE(s.dp_edg, col, e, k);
    dx[e] = EE(s.dp_edg, xc, ex, k);
    dpx[e] = (float)EC(s.dp, xc, k);
  }
  const float cdt = (float)(g.d.dt * GD(rarea, ipn));
  double c_in = 0., c_out = 0., cq_min = 0., cq_max = 0.;
  float q_min_in = 3.402823466e38f, q_max_in = 0.f;    // huge(), 0
#pragma unroll
  for (int e = 0; e < NP; e++) {
    const float vnorm = ipn < x[e] ? vn[e] : -vx[e];
    const float cn = fabsf(vnorm) * cdt;
    const float nm = dpx[(e + NP - 1) % NP], np1 = dpx[(e + 1) % NP];
    const float lo4 = fminf(fminf(gvf, dpx[e]), fminf(nm, np1)), hi4 = fmaxf(fmaxf(gvf, dpx[e]), fmaxf(nm, np1));
    const float qs = fmaxf(lo4, fminf(hi4, .5f * (de[e] + dx[e])));
    __builtin_nontemporal_store(qs, &EE(s.antifx, col, e, k));
    if (vnorm > 0.f) {
      c_out = c_out + (double)cn;
    } else {
      const float qn = fminf(qs, dpx[e]), qx = fmaxf(qs, dpx[e]);
      c_in = c_in + (double)cn;
      q_min_in = fminf(q_min_in, qn);
      q_max_in = fmaxf(q_max_in, qx);
      cq_min = cq_min + (double)cn * (double)qn;
      cq_max = cq_max + (double)cn * (double)qx;
    }
  }
  float qmo = -100000.f, qxo = 100000.f;
  if (c_out != 0.) {
    const double q_min_next = fmin(gv, (double)q_min_in), q_max_next = fmax(gv, (double)q_max_in);
    const double rest = gv + c.ab2 * t_of + c.ab3 * t_vof, keep = 1.0 + c_in - c_out;
    const double w = g.d.dt / (c_out * c.ab1), a = c.ab1 / g.d.dt;
    qmo = (float)((rest + a * cq_max - q_max_next * keep) * w);
    qxo = (float)((rest + a * cq_min - q_min_next * keep) * w);
  }
  __builtin_nontemporal_store(qmo, &EC(s.r_plus, col, k));
  __builtin_nontemporal_store(qxo, &EC(s.r_mnus, col, k));
}

template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_massadv1(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const StepCtl c = load_ctl(s.ctl);
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_massadv1_col<6>(g, s, c, ipn, base, k);
  else e_massadv1_col<5>(g, s, c, ipn, base, k);
}

// massadv2: the final clamp by the upwind column's bounds, the mass flux,
// the FP64 thickness tendency and Adams-Bashforth update kept non-negative
// (massflux_newdepth), and update_pres.
template <int NP>
__device__ __forceinline__ void e_massadv2_col(const EnsGrid &g, const EnsState &s, const StepCtl &c, int ipn,
                                               int base, int k) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const int col = base + ipn;
  const bool live = k < nvl;
  double dpn = 0., lt = 0.;
  if (live) {
    const double t_of = EAB(s.dp_tdcy, of, col, k), t_vof = EAB(s.dp_tdcy, vof, col, k), dp0 = EC(s.dp, col, k);
    float vnorm[NP], qs[NP], lo[NP], hi[NP];
#pragma unroll
    for (int e = 0; e < NP; e++) {
      const int x = GPROX(ipn, e), xc = base + x, ex = GPROXS(ipn, e);
      const float vn = EE(s.vn_edg, col, e, k), vx = EE(s.vn_edg, xc, ex, k);
      vnorm[e] = ipn < x ? vn : -vx;
      qs[e] = EE(s.antifx, col, e, k);
      const size_t up = (size_t)(vnorm[e] > 0.f ? col : xc) * nvl + k;
      lo[e] = s.r_plus[up];
      hi[e] = s.r_mnus[up];
    }
    double at = 0.;
#pragma unroll
    for (int e = 0; e < NP; e++) {
      const float mf = vnorm[e] * fmaxf(lo[e], fminf(hi[e], qs[e]));
      __builtin_nontemporal_store(mf, &EE(s.massfx, col, e, k));
      at = at + (double)mf;
    }
    if (NP < NPP) __builtin_nontemporal_store(0.f, &EE(s.massfx, col, NPP - 1, k));
    const double dtn = -at * GD(rarea, ipn);
    EAB(s.dp_tdcy, nf, col, k) = dtn;
    const double inc = c.ab1 * dtn + c.ab2 * t_of + c.ab3 * t_vof;
    dpn = fmax(dp0 + inc, 0.);
    __builtin_nontemporal_store(dpn, &EC(s.dp, col, k));
    lt = inc / g.d.dt;
  }
  e_update_pres_col<NP>(g, s, ipn, col, k, live, dpn, lt);
}

template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_massadv2(EnsGrid g, EnsState s) {
  const int nip = g.d.nip;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip) return;
  const StepCtl c = load_ctl(s.ctl);
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_massadv2_col<6>(g, s, c, ipn, base, k);
  else e_massadv2_col<5>(g, s, c, ipn, base, k);
}

// ---------------------------------------------------------------------------
// trcadv1-3 (trcadv_compute.F90:19-315), every tracer of a column in one
// thread, so the mass fluxes and the thickness are read once.  The final mass
// flux is antisymmetric, so the neighbour's flux through the shared edge is
// the negated own one.  Tracer edge values are the same bits on both sides
// of an edge: tracer 1 from the edge-symmetric stencil (e_theta_edge, formed
// here from the tracer values this kernel loads anyway), the others as the
// average of both cells' reconstructions (stored by edgvar3; forming the
// neighbour's here from its ring instead measured 7 ms slower at 16 x G7 x
// 64 than the stored values; forming the cell's own one here from the
// neighbour values already loaded saved no time either: 81 VGPRs instead of
// 63, occupancy 5 instead of 7).  trcadv1 leaves trcadv2's sums of the
// antidiffusive out- and inflows in tr_rp/tr_rm, and trcadv2 forms the local
// bounds of trcadv1 again from the tracer values instead of reading them.

template <int NP>
__device__ __forceinline__ void e_trcadv1_col(const EnsGrid &g, const EnsState &s, const StepCtl &c, int ipn,
                                              int base, int k) {
  const int nvl = g.d.nvl, ncol = s.ncol, ntra = g.d.ntra;
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const int col = base + ipn;
  const float thshld = 1.e-11f;
  float m[NP];
  int xs[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) {
    xs[e] = base + GPROX(ipn, e);
    m[e] = EE(s.massfx, col, e, k);
  }
  const float rdp = frcp(fmaxf(thshld, (float)EC(s.dp, col, k))), rar = GF(rarea, ipn);
  const float a1 = (float)c.ab1, a2 = (float)c.ab2, a3 = (float)c.ab3;
  for (int t = 0; t < ntra; t++) {
    // every load of this tracer before its first store
    const float *trf = s.tr + (size_t)t * ncol * nvl;
    const float ti = trf[(size_t)col * nvl + k];
    float v[NP], te[NP];
#pragma unroll
    for (int e = 0; e < NP; e++) v[e] = trf[(size_t)xs[e] * nvl + k];
    const float tdp0 = ETR(s.trdp, t, col, k), tlo_of = ETAB(s.trl_tdcy, t, of, col, k);
    const float tlo_vof = ETAB(s.trl_tdcy, t, vof, col, k);
    if (t == 0) {
#pragma unroll
      for (int e = 0; e < NP; e++) {
        const int et = (e + NP - 1) % NP, el = (e + 1) % NP;
        te[e] = e_theta_edge(ti, v[e], v[et], v[el], xs[et], xs[el]);
      }
    } else {
      float eo[NP], en[NP];
#pragma unroll
      for (int e = 0; e < NP; e++) {
        eo[e] = ETE(s.trc_edg, t - 1, col, e, k);
        en[e] = ETE(s.trc_edg, t - 1, xs[e], GPROXS(ipn, e), k);
      }
#pragma unroll
      for (int e = 0; e < NP; e++) te[e] = .5f * (eo[e] + en[e]);
    }
    float acc = 0.f, sp = 0.f, sm = 0.f, mx = ti, mn = ti;
#pragma unroll
    for (int e = 0; e < NP; e++) {
      mx = fmaxf(mx, v[e]);
      mn = fminf(mn, v[e]);
    }
#pragma unroll
    for (int e = 0; e < NP; e++) {
      const float fl = .5f * ((m[e] + fabsf(m[e])) * ti - (fabsf(m[e]) - m[e]) * v[e]);
      const float a = m[e] * te[e] - fl;
      __builtin_nontemporal_store(a, &ETE(s.antiflx, t, col, e, k));
      sp = sp - fminf(0.f, a);
      sm = sm + fmaxf(0.f, a);
      acc = acc + fl;
    }
    if (NP < NPP) __builtin_nontemporal_store(0.f, &ETE(s.antiflx, t, col, NPP - 1, k));
    const float tlo = -acc * rar;
    __builtin_nontemporal_store(tlo, &ETAB(s.trl_tdcy, t, nf, col, k));
    const float lo = tdp0 + a1 * tlo + a2 * tlo_of + a3 * tlo_vof;
    mx = fmaxf(0.f, mx);
    mn = fmaxf(0.f, mn);
    __builtin_nontemporal_store(fmaxf(mn, fminf(mx, lo * rdp)), &ETR(s.trcr_lo, t, col, k));
    __builtin_nontemporal_store(sp, &ETR(s.tr_rp, t, col, k));
    __builtin_nontemporal_store(sm, &ETR(s.tr_rm, t, col, k));
  }
}

template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_trcadv1(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const StepCtl c = load_ctl(s.ctl);
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_trcadv1_col<6>(g, s, c, ipn, base, k);
  else e_trcadv1_col<5>(g, s, c, ipn, base, k);
}

template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_trcadv2(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl, ncol = s.ncol, ntra = g.d.ntra;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const StepCtl c = load_ctl(s.ctl);
  const int of = c.of - 1, vof = c.vof - 1;
  const int base = blockIdx.y * nip, col = base + ipn;
  const float thshld = 1.e-11f;
  int xs[NPP];
#pragma unroll
  for (int e = 0; e < NPP; e++) xs[e] = base + GPROX(ipn, e);
  const float dp = (float)EC(s.dp, col, k), area = GF(area, ipn);
  const float a2 = (float)c.ab2, a3 = (float)c.ab3, ra1 = (float)(1. / c.ab1) * area;
  for (int t = 0; t < ntra; t++) {
    const float sp = ETR(s.tr_rp, t, col, k), sm = ETR(s.tr_rm, t, col, k);
    const float ti = ETR(s.tr, t, col, k), lo = ETR(s.trcr_lo, t, col, k);
    float mx = ti, mn = ti;
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      const float v = ETR(s.tr, t, xs[e], k);
      mx = fmaxf(mx, v);
      mn = fminf(mn, v);
    }
    mx = fmaxf(fmaxf(0.f, mx), lo);
    mn = fminf(fmaxf(0.f, mn), lo);
#pragma unroll
    for (int e = 0; e < NPP; e++) {
      const float v = ETR(s.trcr_lo, t, xs[e], k);
      mx = fmaxf(mx, v);
      mn = fminf(mn, v);
    }
    mx = fmaxf(0.f, mx);
    mn = fmaxf(0.f, mn);
    __builtin_nontemporal_store(mx, &ETR(s.trmax, t, col, k));
    __builtin_nontemporal_store(mn, &ETR(s.trmin, t, col, k));
    const float to = ETAB(s.trc_tdcy, t, of, col, k), tv = ETAB(s.trc_tdcy, t, vof, col, k);
    const float qp = ((mx - lo) * dp - (a2 * fminf(0.f, to) + a3 * fmaxf(0.f, tv))) * ra1;
    const float qm = ((lo - mn) * dp + (a2 * fmaxf(0.f, to) + a3 * fminf(0.f, tv))) * ra1;
    __builtin_nontemporal_store(fmaxf(0.f, fminf(1.f, qp * frcp(fmaxf(sp, thshld)))), &ETR(s.tr_rp, t, col, k));
    __builtin_nontemporal_store(fmaxf(0.f, fminf(1.f, qm * frcp(fmaxf(sm, thshld)))), &ETR(s.tr_rm, t, col, k));
  }
}

// trcadv3: the limiter's neighbour value is gathered through a per-lane
// address (R+ or R- of the neighbour by the sign of the flux) and every load
// of a tracer is issued before the first is used.
template <int NP>
__device__ __forceinline__ void e_trcadv3_col(const EnsGrid &g, const EnsState &s, const StepCtl &c, int ipn,
                                              int base, int k) {
  const int nvl = g.d.nvl, ncol = s.ncol, ntra = g.d.ntra;
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const int col = base + ipn;
  const float thshld = 1.e-11f;
  int xs[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) xs[e] = base + GPROX(ipn, e);
  const float rdp = frcp(fmaxf(thshld, (float)EC(s.dp, col, k))), rar = GF(rarea, ipn);
  const float a1 = (float)c.ab1, a2 = (float)c.ab2, a3 = (float)c.ab3;
  for (int t = 0; t < ntra; t++) {
    const size_t tb = (size_t)t * ncol;
    const float rp = ETR(s.tr_rp, t, col, k), rm = ETR(s.tr_rm, t, col, k);
    const float trl = ETAB(s.trl_tdcy, t, nf, col, k), tco = ETAB(s.trc_tdcy, t, of, col, k);
    const float tcv = ETAB(s.trc_tdcy, t, vof, col, k), tdp0 = ETR(s.trdp, t, col, k);
    const float tmn = ETR(s.trmin, t, col, k), tmx = ETR(s.trmax, t, col, k);
    float a[NP], r[NP];
#pragma unroll
    for (int e = 0; e < NP; e++) a[e] = ETE(s.antiflx, t, col, e, k);
#pragma unroll
    for (int e = 0; e < NP; e++) {
      const float *src = a[e] >= 0.f ? s.tr_rp : s.tr_rm;
      r[e] = src[(tb + xs[e]) * nvl + k];
    }
    float at = 0.f;
#pragma unroll
    for (int e = 0; e < NP; e++) at = at + a[e] * fminf(a[e] >= 0.f ? rm : rp, r[e]);
    at = -at * rar;
    const float tt = trl + at;
    const float tdp = tdp0 + a1 * tt + a2 * tco + a3 * tcv;
    __builtin_nontemporal_store(tt, &ETAB(s.trc_tdcy, t, nf, col, k));
    __builtin_nontemporal_store(tdp, &ETR(s.trdp, t, col, k));
    if (col_wet(s.wet, col)) __builtin_nontemporal_store(fmaxf(tmn, fminf(tmx, tdp * rdp)), &ETR(s.tr, t, col, k));   // land keeps its tracers
  }
}

template <int COLS, bool SWZ>
// waves_per_eu: 66 VGPRs -> 56 (8 waves per SIMD instead of 7), no spills
__global__ void __launch_bounds__(64 * COLS) __attribute__((amdgpu_waves_per_eu(8))) k_e_trcadv3(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip || k >= nvl) return;
  const StepCtl c = load_ctl(s.ctl);
  const int base = blockIdx.y * nip;
  if (GNP(ipn) == 6) e_trcadv3_col<6>(g, s, c, ipn, base, k);
  else e_trcadv3_col<5>(g, s, c, ipn, base, k);
}

// ---------------------------------------------------------------------------
// momtum (momtum.F90:101-158; dissip momtum_compute.F90:102-229; momtum1
// 231-450).  Diffusion, vorticity, Coriolis and the AB update in FP32; the
// pressure-gradient line integral of Montgomery potential and Exner function
// in FP64.  The edge winds, Montgomery potential and kinetic energy are
// formed here from the cell values, with the operations edgvar1 uses for them
// (edgvar1.F90:162-421), instead of being stored by edgvar1 and read back.  New winds go to u2/v2 (hybgen's remap reads them from there).
template <int NP>
__device__ __forceinline__ void e_momtum_col(const EnsGrid &g, const EnsState &s, const StepCtl &c, int ipn,
                                             int base, int k, float &un, float &vn) {
#pragma clang fp contract(fast)
  const int nvl = g.d.nvl, nvlp1 = nvl + 1, ncol = s.ncol;
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const int col = base + ipn;
  const float thshld = 1.e-11f;
  int nb[NP];
  float nu[NP], nv[NP], ndp[NP], nth[NP];
  double nmp[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) {
    nb[e] = base + GPROX(ipn, e);
    nu[e] = EC(s.us, nb[e], k);
    nv[e] = EC(s.vs, nb[e], k);
    ndp[e] = (float)EC(s.dp, nb[e], k);
    nmp[e] = EC(s.mp, nb[e], k);
    nth[e] = EC(s.th0, nb[e], k);
  }
  const float thi = EC(s.th0, col, k);
  const float ui = EC(s.us, col, k), vi = EC(s.vs, col, k);
  const float uo = EAB(s.u_tdcy, of, col, k), uvo = EAB(s.u_tdcy, vof, col, k);
  const float vo = EAB(s.v_tdcy, of, col, k), vvo = EAB(s.v_tdcy, vof, col, k);
  const float wui = ui + GF(grndspd, ipn), wvi = vi;
  const float a = fmaxf((float)EC(s.dp, col, k), thshld);
  const float factor = GF(rarea, ipn) * frcp(a);
  const float dff = g.dfflen[k];
  float wu = 0.f, wv = 0.f;
#pragma unroll
  for (int e = 0; e < NP; e++) {
    if (!col_wet(s.wet, nb[e])) continue;          // no diffusion through a land edge
    const float wux = nu[e] + GF(grndspd, nb[e] - base), wvx = nv[e];
    const float c0 = FCS(ipn, e, 0), s0 = FSN(ipn, e, 0), c1 = FCS(ipn, e, 1), s1 = FSN(ipn, e, 1);
    const float uxy1 = c0 * wui + s0 * wvi, vxy1 = -s0 * wui + c0 * wvi;
    const float uxy2 = c1 * wux + s1 * wvx, vxy2 = -s1 * wux + c1 * wvx;
    const float b = fmaxf(ndp[e], thshld);
    const float hh = a * b * frcp(a + b);
    const float w = dff * GF(sideln, ipn * NPP + e) * 2.f * hh;
    const float ufx = (uxy2 - uxy1) * w, vfx = (vxy2 - vxy1) * w;
    wu = wu + (c0 * ufx - s0 * vfx) * factor;
    wv = wv + (s0 * ufx + c0 * vfx) * factor;
  }
  const float u1 = ui + wu, v1 = vi + wv;
  double px = 0., py = 0.;
  float rv = 0.f;
  const double exm = .5 * (EL(s.ex, col, k) + EL(s.ex, col, k + 1));
  const double mpi = EC(s.mp, col, k);
  const float mi = fsqrt(ui * ui + vi * vi);
  float mn[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) mn[e] = fsqrt(nu[e] * nu[e] + nv[e] * nv[e]);
#pragma unroll
  for (int e = 0; e < NP; e++) {
    const int et = (e + NP - 1) % NP, el = (e + 1) % NP;
    const float2 w = e_edge_wind(g, ipn, e, ui, vi, nu[e], nv[e], nu[et], nv[et], nu[el], nv[el], mi, mn[e], mn[et],
                                 mn[el]);
    const float ue = w.x, ve = w.y;
    const double mont = (2. * (mpi + nmp[e]) + nmp[et] + nmp[el]) * (1. / 6.);
    const double b = mont + (double)(.5f * (ue * ue + ve * ve));
    const double te = (double)e_theta_edge(thi, nth[e], nth[et], nth[el], nb[et], nb[el]);
    const double q = exm * te - b;          // -(M + KE) + Exner * theta, times sidevec_c
    px = px + q * (double)FSVC(ipn, e, 1);
    py = py - q * (double)FSVC(ipn, e, 0);
    rv = rv + (FSVE(ipn, e, 0) * ue + FSVE(ipn, e, 1) * ve);
  }
  const float pgx = (float)(px * GD(rarea, ipn)), pgy = (float)(py * GD(rarea, ipn));
  rv = rv * GF(rarea, ipn);
  const float f = GF(corio, ipn) + rv;
  const float ut = pgx + f * v1, vt = pgy - f * u1;
  un = u1 + (float)c.ab1 * ut + (float)c.ab2 * uo + (float)c.ab3 * uvo;
  vn = v1 + (float)c.ab1 * vt + (float)c.ab2 * vo + (float)c.ab3 * vvo;
  __builtin_nontemporal_store(rv, &EC(s.relvor, col, k));
  EAB(s.u_tdcy, nf, col, k) = ut;
  EAB(s.v_tdcy, nf, col, k) = vt;
}

template <int COLS, bool SWZ>
// waves_per_eu: 101 VGPRs -> 91 (5 waves per SIMD instead of 4), scalar
// registers spilled to vector lanes only
__global__ void __launch_bounds__(64 * COLS) __attribute__((amdgpu_waves_per_eu(5))) k_e_momtum(EnsGrid g, EnsState s) {
  const int nip = g.d.nip, nvl = g.d.nvl, ncol = s.ncol;
  const int k = threadIdx.x, ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip) return;
  const bool live = k < nvl;
  const StepCtl c = load_ctl(s.ctl);
  const int base = blockIdx.y * nip, col = base + ipn;
  if (!col_wet(s.wet, col)) {         // a land column keeps its winds
    if (live) {
      const int nf = c.nf - 1;
      __builtin_nontemporal_store(EC(s.us, col, k), &EC(s.u2, col, k));
      __builtin_nontemporal_store(EC(s.vs, col, k), &EC(s.v2, col, k));
      EAB(s.u_tdcy, nf, col, k) = 0.f;
      EAB(s.v_tdcy, nf, col, k) = 0.f;
      __builtin_nontemporal_store(0.f, &EC(s.relvor, col, k));
    }
    return;
  }
  float un = 0.f, vn = 0.f;
  if (live) {
    if (GNP(ipn) == 6) e_momtum_col<6>(g, s, c, ipn, base, k, un, vn);
    else e_momtum_col<5>(g, s, c, ipn, base, k, un, vn);
  }
  // Rayleigh damping of the top layers (momtum_compute.F90:406-416): the
  // largest squared speed of the top three layers, from their lanes
  const float q = (live && k >= nvl - 3) ? un * un + vn * vn : 0.f;
  float qmax = q;
  for (int d = 1; d < 64; d <<= 1) qmax = fmaxf(qmax, __shfl_xor(qmax, d, 64));
  if (!live) return;
  float qq = qmax / (float)(RLTHRESH * RLTHRESH);
  const int ndamp = (int)(0.25 * nvl);
  const int k1 = k + 1;
  if (qq > 1.f && k1 >= nvl - ndamp) {
    qq = sqrtf((qq - 1.f) * qq);
    const float fac = (float)(1. - (g.d.dt / 86400.) * RLEIGH * (double)qq * fim_pow(2., RLTAPER * (double)(k1 - nvl)));
    un = un * fac;
    vn = vn * fac;
  }
  __builtin_nontemporal_store(un, &EC(s.u2, col, k));
  __builtin_nontemporal_store(vn, &EC(s.v2, col, k));
}

// ---------------------------------------------------------------------------
// hystat (hystat.F90:23-115) with the tracer floors and tracer mass: the
// surface-pressure sum and the Montgomery-potential recurrence are FP64
// wavefront scans.
__device__ __forceinline__ void e_hystat_col(const EnsGrid &g, const EnsState &s, int col, int k, int its) {
  const int nip = g.d.nip, nvl = g.d.nvl, nvlp1 = nvl + 1, ncol = s.ncol, ntra = g.d.ntra;
  const bool live = k < nvl;
  const double dp = live ? EC(s.dp, col, k) : 0.;
  const double th = live ? (double)ETR(s.tr, 0, col, k) : 0.;
  const double exk = live ? EL(s.ex, col, k) : 0.;
  const double exk1 = live ? EL(s.ex, col, k + 1) : 0.;
  double ps = dp;
  for (int d = 1; d < 64; d <<= 1) ps += __shfl_xor(ps, d, 64);
  const double thm = __shfl_up(th, 1, 64);
  const double inc = k == 0 ? exk * th + EL(s.ph, col, 0) : (live ? exk * (th - thm) : 0.);
  const double mp = wave_prefix_sum(inc, k);
  if (k == 0) {
    const double old = s.psrf[col];
    s.psrf[col] = ps;
    if (its > 0) s.ptdcy[(size_t)(its % 2) * ncol + col] = ps - old;
  }
  if (!live) return;
  __builtin_nontemporal_store(mp, &EC(s.mp, col, k));
  __builtin_nontemporal_store(mp - th * exk1, &EL(s.ph, col, k + 1));
  // tracers 2 and 3 clamped (qv, qw), then trdp = tr * dp; four tracers'
  // loads are issued before their stores
  const float dpf = (float)dp;
  for (int t0 = 0; t0 < ntra; t0 += 4) {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = ETR(s.tr, min(t0 + j, ntra - 1), col, k);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int t = t0 + j;
      if (t >= ntra) break;
      float x = v[j];
      if (t == 1 || t == 2) {
        x = fmaxf(t == 1 ? (float)QVMIN : (float)QWMIN, x);
        __builtin_nontemporal_store(x, &ETR(s.tr, t, col, k));
      }
      __builtin_nontemporal_store(x * dpf, &ETR(s.trdp, t, col, k));
    }
  }
  (void)nip;
}

template <int COLS, bool SWZ>
__global__ void __launch_bounds__(64 * COLS) k_e_hystat(EnsGrid g, EnsState s) {
  const int nip = g.d.nip;
  const int ipn = tile_cell<COLS, SWZ>();
  if (ipn >= nip) return;
  e_hystat_col(g, s, blockIdx.y * nip + ipn, threadIdx.x, load_ctl(s.ctl).its);
}

// start-up hystat(its = 0) (dyn_init.F90:1037)
__global__ void __launch_bounds__(256) k_e_hystat0(EnsGrid g, EnsState s) {
  const int nip = g.d.nip;
  const int ipn = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + threadIdx.y));
  if (ipn >= nip) return;
  e_hystat_col(g, s, blockIdx.y * nip + ipn, threadIdx.x, 0);
}

// abstart (abstart.F90:38-66) on the device
__global__ void k_e_abstart(StepCtl *ctl, double dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  StepCtl c = *ctl;
  c.its = c.its + 1;
  c.nf = c.nf + 1; if (c.nf == 4) c.nf = 1;
  c.of = c.of + 1; if (c.of == 4) c.of = 1;
  c.vof = c.vof + 1; if (c.vof == 4) c.vof = 1;
  if (c.its == 1) { c.ab1 = dt; c.ab2 = 0.; c.ab3 = 0.; }
  if (c.its == 2) { c.ab1 = dt * 1.5; c.ab2 = -dt * .5; c.ab3 = 0.; }
  if (c.its == 3) { c.ab1 = dt * 23. / 12.; c.ab2 = -dt * 16. / 12.; c.ab3 = dt * 5. / 12.; }
  *ctl = c;
}

// ---------------------------------------------------------------------------
// start-up helpers
__global__ void __launch_bounds__(256) k_e_pres_from_exner(EnsGrid g, EnsState s) {
  const size_t n = (size_t)s.ncol * (g.d.nvl + 1);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    s.pr[i] = P1000 * fim_pow(s.ex[i] / CP, CP / RD);
}

__global__ void __launch_bounds__(256) k_e_trdp(EnsGrid g, EnsState s) {
  const size_t cz = (size_t)s.ncol * g.d.nvl;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cz; i += (size_t)gridDim.x * blockDim.x) {
    const float dp = (float)s.dp[i];
    for (int t = 0; t < g.d.ntra; t++) s.trdp[(size_t)t * cz + i] = s.tr[(size_t)t * cz + i] * dp;
  }
}

// one member's FP64 state (plain layout) into member slot m
__global__ void __launch_bounds__(256) k_e_convert(EnsGrid g, EnsState s, int m, DevState src) {
  const int nip = g.d.nip, nvl = g.d.nvl, ntra = g.d.ntra;
  const size_t cz = (size_t)nip * nvl, lz = (size_t)nip * (nvl + 1);
  const size_t ecz = (size_t)s.ncol * nvl;
  const size_t oc = (size_t)m * cz, ol = (size_t)m * lz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < lz; i += (size_t)gridDim.x * blockDim.x) {
    s.pr[ol + i] = src.pr[i];
    s.ex[ol + i] = src.ex[i];
    s.ph[ol + i] = src.ph[i];
    if (i < cz) {
      s.us[oc + i] = (float)src.us[i];
      s.vs[oc + i] = (float)src.vs[i];
      s.dp[oc + i] = src.dp[i];
      for (int t = 0; t < ntra; t++) s.tr[(size_t)t * ecz + oc + i] = (float)src.tr[(size_t)t * cz + i];
    }
  }
}

// ---------------------------------------------------------------------------
// globsum (globsum.F90:70-101) per member: FP64 block partials, and FP32
// Kahan-compensated partials (value + compensation) of the same sums.
struct KPair { float s, c; };
__device__ __forceinline__ KPair kp_add(KPair a, float x) {    // Kahan step
  const float y = x - a.c;
  const float t = a.s + y;
  a.c = (t - a.s) - y;
  a.s = t;
  return a;
}
__device__ __forceinline__ KPair kp_merge(KPair a, KPair b) {  // a + b, compensations carried
  const float s = a.s + b.s;
  const float bb = s - a.s;
  const float err = (a.s - (s - bb)) + (b.s - bb);             // two-sum error of a.s + b.s
  KPair r;
  r.s = s;
  r.c = a.c + b.c - err;
  return r;
}

__global__ void __launch_bounds__(256) k_e_globsum(EnsGrid g, EnsState s, double *p64, float *p32) {
  __shared__ double sh[256];
  __shared__ float shs[256], shc[256];
  const int nip = g.d.nip, nvl = g.d.nvl, ntra = g.d.ntra, ncol = s.ncol;
  const int m = blockIdx.y, nb = gridDim.x;
  for (int q = 0; q <= ntra; q++) {
    double acc = 0.;
    KPair kp{0.f, 0.f};
    for (int ipn = blockIdx.x * 256 + threadIdx.x; ipn < nip; ipn += nb * 256) {
      const int col = m * nip + ipn;
      double colsum = 0.;
      for (int k = 0; k < nvl; k++) {
        const double dp = EC(s.dp, col, k);
        double v = g.d.area[ipn] * dp;
        float vf = GF(area, ipn) * (float)dp;
        if (q > 0) {
          const float tr = ETR(s.tr, q - 1, col, k);
          v = v * (double)tr;
          vf = vf * tr;
        }
        colsum = colsum + v;
        kp = kp_add(kp, vf);
      }
      acc = acc + colsum;
    }
    sh[threadIdx.x] = acc;
    shs[threadIdx.x] = kp.s;
    shc[threadIdx.x] = kp.c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        KPair r = kp_merge(KPair{shs[threadIdx.x], shc[threadIdx.x]}, KPair{shs[threadIdx.x + w], shc[threadIdx.x + w]});
        shs[threadIdx.x] = r.s;
        shc[threadIdx.x] = r.c;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const size_t o = ((size_t)m * (ntra + 1) + q) * nb + blockIdx.x;
      p64[o] = sh[0];
      p32[2 * o] = shs[0];
      p32[2 * o + 1] = shc[0];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host launchers
static dim3 ens_grid(int nip, int members, int z, int cols, bool swz) {
  int gx = (nip + cols - 1) / cols;
  if (swz) gx = (gx + 7) / 8 * 8;
  return dim3(gx, members, z);
}

#define ENS_DISPATCH(KNAME, z)                                                                      \
  do {                                                                                              \
    const dim3 gr = ens_grid(g.d.nip, s.members, (z), tile_cols(tile), tile_swz(tile));             \
    const dim3 bl(64, tile_cols(tile));                                                            \
    switch (tile) {                                                                                 \
      case 0: KNAME<1, false><<<gr, bl, 0, st>>>(g, s); break;                                      \
      case 1: KNAME<2, false><<<gr, bl, 0, st>>>(g, s); break;                                      \
      case 2: KNAME<4, false><<<gr, bl, 0, st>>>(g, s); break;                                      \
      case 3: KNAME<8, false><<<gr, bl, 0, st>>>(g, s); break;                                      \
      case 4: KNAME<1, true><<<gr, bl, 0, st>>>(g, s); break;                                       \
      case 5: KNAME<2, true><<<gr, bl, 0, st>>>(g, s); break;                                       \
      case 6: KNAME<4, true><<<gr, bl, 0, st>>>(g, s); break;                                       \
      default: KNAME<8, true><<<gr, bl, 0, st>>>(g, s); break;                                      \
    }                                                                                               \
  } while (0)

void ens_launch(int kernel, int tile, const EnsGrid &g, const EnsState &s, hipStream_t st) {
  const int ntra = g.d.ntra;
  const char *name = ens_kernel_names[kernel];
  if (::fimhip::g_ktimer) kt_begin(name, st);
  switch (kernel) {
    case EK_EDGVAR1: ENS_DISPATCH(k_e_edgvar1, 1); break;
    case EK_EDGVAR3: ENS_DISPATCH(k_e_edgvar3, 1); break;
    case EK_CNUITY12: ENS_DISPATCH(k_e_cnuity12, 1); break;
    case EK_CNUITY3: ENS_DISPATCH(k_e_cnuity3, 1); break;
    case EK_CNUITY45: ENS_DISPATCH(k_e_cnuity45, 1); break;
    case EK_TRCADV1: ENS_DISPATCH(k_e_trcadv1, 1); break;
    case EK_TRCADV2: ENS_DISPATCH(k_e_trcadv2, 1); break;
    case EK_TRCADV3: ENS_DISPATCH(k_e_trcadv3, 1); break;
    case EK_MOMTUM: ENS_DISPATCH(k_e_momtum, 1); break;
    case EK_MASSADV1: ENS_DISPATCH(k_e_massadv1, 1); break;
    case EK_MASSADV2: ENS_DISPATCH(k_e_massadv2, 1); break;
    case EK_TRCFCT: ens_launch_trc_fct(tile, g, s, st); break;
    default: ENS_DISPATCH(k_e_hystat, 1); break;
  }
  if (::fimhip::g_ktimer) kt_end(st);
}

void ens_launch_abstart(const EnsState &s, double dt, hipStream_t st) {
  k_e_abstart<<<1, 64, 0, st>>>(s.ctl, dt);
}
void ens_launch_pres_from_exner(const EnsGrid &g, const EnsState &s, hipStream_t st) {
  k_e_pres_from_exner<<<2048, 256, 0, st>>>(g, s);
}
void ens_launch_trdp(const EnsGrid &g, const EnsState &s, hipStream_t st) {
  k_e_trdp<<<2048, 256, 0, st>>>(g, s);
}
void ens_launch_hystat_init(const EnsGrid &g, const EnsState &s, hipStream_t st) {
  k_e_hystat0<<<dim3((g.d.nip + 3) / 4, s.members), dim3(64, 4), 0, st>>>(g, s);
}
void ens_launch_convert_member(const EnsGrid &g, const EnsState &s, int member, const DevState &src,
                               hipStream_t st) {
  k_e_convert<<<2048, 256, 0, st>>>(g, s, member, src);
}
void ens_launch_globsum(const EnsGrid &g, const EnsState &s, double *part64, float *part32, int nblocks,
                        hipStream_t st) {
  k_e_globsum<<<dim3(nblocks, s.members), 256, 0, st>>>(g, s, part64, part32);
}

}  // namespace fimhip
