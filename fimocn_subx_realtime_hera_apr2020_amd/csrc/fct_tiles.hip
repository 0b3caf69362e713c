// The tracer flux-corrected transport of a step in one kernel (the batched
// mixed-precision form): trcadv1-3 (trcadv_compute.F90:19-315) together with
// the least-squares edge reconstructions of tracers 2..ntra that edgvar3
// forms for them (edgvar3.F90:65-163), without any intermediate array in HBM.
//
// The cells are partitioned into patches of up to FCT_CORE_MAX cells.  One
// block handles one patch of one member for a group of tracers over a chunk
// of levels: it stages the tracers of every cell within three neighbour
// rings of the patch in LDS (one wavefront per cell, its lanes along the
// tracers and levels), then
//   stage A  rings 0-2: the low-order fluxes, tendency and bounded
//            low-order tracer (trcadv1); rings 0-1 also the two-sided edge
//            values and antidiffusive fluxes and their in/outflow sums;
//   stage B  rings 0-1: the local bounds and the limiter ratios R+ and R-
//            (trcadv2);
//   stage C  ring 0: the limited fluxes, tendency, AB update and tracer
//            (trcadv3),
// with a block barrier between stages.  A ring's values are recomputed by
// each patch that needs them, so a cell's results in rings 1-2 exist in
// several blocks but are written to HBM only by the block whose core holds
// the cell.  Every operation is the one the three-kernel chain performs
// (kernels_ens.hip), in the same order, so the results are bit-identical to
// it; in particular each edge's flux is formed from the same operands on
// both sides and the tracer mass still cancels exactly across every edge.
//
// The block reads the tracer and tracer mass of cells that other blocks
// update in the same launch, so the kernel reads one buffer and writes the
// other of a pair (the instance alternates them every step).
#include <algorithm>
#include <vector>

#include "ens_internal.h"
#include "host_common.h"

namespace fimhip {

// ---------------------------------------------------------------------------
// host: patches and their halo rings
//
// Patches grow breadth-first from the lowest unassigned cell over unassigned
// cells, up to FCT_CORE_MAX cells; the three rings around the patch are the
// cells at neighbour distance 1, 2 and 3 from it, each ring in index order.
// A patch whose rings would exceed the LDS tile gives up its last cells.
// On the IJ-ordered icosahedral grids this gives compact, roughly hexagonal
// patches (the three rings hold about twice the core's cells).
int ens_build_tiles(const HostGrid &h, int nip, TileHost &out) {
  std::vector<int> asg(nip, -1), mark(nip, -1);
  out.off.assign(1, 0);
  out.cnt.clear();
  out.ent.clear();
  std::vector<int> core, ring[4], q;
  int stamp = 0;
  for (int seed = 0; seed < nip; seed++) {
    if (asg[seed] >= 0) continue;
    core.assign(1, seed);
    q.assign(1, seed);
    ++stamp;
    mark[seed] = stamp;
    for (size_t qi = 0; qi < q.size() && (int)core.size() < FCT_CORE_MAX; qi++) {
      const int c = q[qi];
      int nb[NPP];
      for (int e = 0; e < NPP; e++) nb[e] = h.prox[(size_t)c * NPP + e];
      std::sort(nb, nb + NPP);
      for (int e = 0; e < NPP && (int)core.size() < FCT_CORE_MAX; e++) {
        const int x = nb[e];
        if (asg[x] >= 0 || mark[x] == stamp) continue;
        mark[x] = stamp;
        core.push_back(x);
        q.push_back(x);
      }
    }
    for (;;) {
      // rings by distance from the core
      ++stamp;
      for (int c : core) mark[c] = stamp;
      ring[0] = core;
      for (int r = 1; r <= 3; r++) {
        ring[r].clear();
        for (int c : ring[r - 1])
          for (int e = 0; e < NPP; e++) {
            const int x = h.prox[(size_t)c * NPP + e];
            if (mark[x] != stamp) { mark[x] = stamp; ring[r].push_back(x); }
          }
        std::sort(ring[r].begin(), ring[r].end());
      }
      const int n0 = (int)ring[0].size(), n1 = n0 + (int)ring[1].size(), n2 = n1 + (int)ring[2].size(),
                n3 = n2 + (int)ring[3].size();
      if (n1 <= FCT_R_MAX && n2 <= FCT_LO_MAX && n3 <= FCT_TILE_MAX) {
        // local indices of the tile cells, and the neighbour table of rings 0-2
        std::vector<int> cells;
        for (int r = 0; r < 4; r++) cells.insert(cells.end(), ring[r].begin(), ring[r].end());
        for (int i = 0; i < n3; i++) mark[cells[i]] = -2 - i;       // local index i, encoded
        for (int i = 0; i < n3; i++) {
          const int c = cells[i];
          int nbl[NPP] = {0, 0, 0, 0, 0, 0}, sl = 0;
          if (i < n2)
            for (int e = 0; e < NPP; e++) {
              const int m = mark[h.prox[(size_t)c * NPP + e]];
              if (m > -2) return FIMHIP_EINVAL;        // a ring-2 cell's neighbour outside the tile
              nbl[e] = -2 - m;
            }
          for (int e = 0; e < NPP; e++) sl |= h.proxs[(size_t)c * NPP + e] << (4 * e);
          out.ent.push_back({c, nbl[0] | nbl[1] << 8 | nbl[2] << 16 | nbl[3] << 24,
                             nbl[4] | nbl[5] << 8 | h.nprox[c] << 16, sl});
        }
        for (int i = 0; i < n3; i++) mark[cells[i]] = 0;
        for (int c : core) asg[c] = (int)out.cnt.size();
        out.cnt.push_back({n0, n1, n2, n3});
        out.off.push_back((int)out.ent.size());
        break;
      }
      if (core.size() == 1) return FIMHIP_EINVAL;     // one cell's rings exceed the tile
      core.pop_back();
    }
  }
  return FIMHIP_OK;
}

int ens_upload_tiles(Arena &a, const TileHost &t, EnsTiles &d) {
  int r;
  std::vector<int4> cnt(t.cnt.size()), ent(t.ent.size());
  for (size_t i = 0; i < t.cnt.size(); i++) cnt[i] = make_int4(t.cnt[i][0], t.cnt[i][1], t.cnt[i][2], t.cnt[i][3]);
  for (size_t i = 0; i < t.ent.size(); i++) ent[i] = make_int4(t.ent[i][0], t.ent[i][1], t.ent[i][2], t.ent[i][3]);
  int *off = nullptr;
  int4 *c = nullptr, *e = nullptr;
  if ((r = a.upload(&off, t.off)) || (r = a.upload(&c, cnt)) || (r = a.upload(&e, ent))) return r;
  d.npatch = (int)t.cnt.size();
  d.off = off;
  d.cnt = c;
  d.ent = e;
  return FIMHIP_OK;
}

}  // namespace fimhip

// The patch tables of a grid, for inspection: sizes[0] = number of patches,
// sizes[1] = number of tile-cell entries; cnt (4 per patch: n0..n3) and ent
// (4 per entry, see EnsTiles) are written when not NULL and large enough.
extern "C" int fimhip_ens_patches(const fimhip_grid_desc *gd, int *sizes, int *cnt, long long cnt_len, int *ent,
                                  long long ent_len) {
  using namespace fimhip;
  if (!gd || !sizes || gd->nip < 12 || !gd->nprox || !gd->prox || !gd->proxs || !gd->lat || !gd->lon || !gd->area)
    return FIMHIP_EINVAL;
  HostGrid h;
  int r = host_grid_prepare(gd, h);
  if (r) return r;
  TileHost t;
  if ((r = ens_build_tiles(h, gd->nip, t))) return r;
  sizes[0] = (int)t.cnt.size();
  sizes[1] = (int)t.ent.size();
  if (cnt) {
    if (cnt_len < 4 * (long long)t.cnt.size()) return FIMHIP_EINVAL;
    for (size_t i = 0; i < t.cnt.size(); i++)
      for (int j = 0; j < 4; j++) cnt[4 * i + j] = t.cnt[i][j];
  }
  if (ent) {
    if (ent_len < 4 * (long long)t.ent.size()) return FIMHIP_EINVAL;
    for (size_t i = 0; i < t.ent.size(); i++)
      for (int j = 0; j < 4; j++) ent[4 * i + j] = t.ent[i][j];
  }
  return FIMHIP_OK;
}

namespace fimhip {

// ---------------------------------------------------------------------------
// device
//
// A block covers one patch of one member for TG tracers over LC = 64 / TG
// levels: lane l of a wavefront works on tracer (group * TG + l / LC) at
// level (chunk * LC + l % LC) of one tile cell, and wavefront w takes the
// tile cells w, w + NW, w + 2 NW, ...  With one cell per wavefront the
// cell's table entry, metric terms and least-squares coefficients are
// wave-uniform (scalar loads) and shared by the TG tracers, and so are its
// mass fluxes and thickness (one load per level for all TG tracers).  LDS
// holds one value per lane and tile cell.

#define EC(a, col, k) (a)[(size_t)(col) * nvl + (k)]
#define EE(a, col, e, k) (a)[((size_t)(col) * NPP + (e)) * nvl + (k)]
#define FHF(i, a, b) ldc(g.hfcoef, ((size_t)(i) * NPP + (a)) * NPP + (b))
// a lane's tracer field at column col: toff = t * ncol * nvl + k
#define TRV(a, col) (a)[toff + (size_t)(col) * nvl]
// a lane's AB slot sl of a tracer tendency: aoff = t * 3 * ncol * nvl + k
#define TAB(a, sl, col) (a)[aoff + ((size_t)(sl) * ncol + (col)) * nvl]

// a table entry of four ints (one load; a scalar load when the index is
// wave-uniform)
__device__ __forceinline__ int4 ldc4(const int4 *p, size_t i) {
  const int *q = (const int *)p + 4 * i;
  return make_int4(ldc(q, 0), ldc(q, 1), ldc(q, 2), ldc(q, 3));
}

// fields of a tile entry: local neighbour e, the cell's edge count, the
// neighbour's index of edge e (proxs)
__device__ __forceinline__ int lnb(const int4 &en, int e) {
  return e < 4 ? (en.y >> (8 * e)) & 0xff : (en.z >> (8 * (e - 4))) & 0xff;
}
__device__ __forceinline__ int enp(const int4 &en) { return (en.z >> 16) & 0xff; }
__device__ __forceinline__ int eslot(const int4 &en, int e) { return (en.w >> (4 * e)) & 0xf; }
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Theta on the edge between cell i and x (edgvar1.F90:425-511, eqwgt), the
// outer stencil cells t, l added in column order (as e_theta_edge).
__device__ __forceinline__ float fct_theta_edge(float thi, float thx, float tht, float thl, int nbt, int nbl) {
  float ta = tht, tb = thl;
  if (nbt > nbl) { const float t = ta; ta = tb; tb = t; }
  return (2.f * (thi + thx) + ta + tb) * (1.f / 6.f);
}

__device__ __forceinline__ float fct_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// The HBM inputs of stage A for one tile cell (a pentagon's sixth mass flux
// is 0), loaded a pass ahead of their use.
struct FctInA {
  float m[NPP], tdp0, tlo_of, tlo_vof;
  double dp;
};
__device__ __forceinline__ FctInA fct_load_a(const EnsState &s, const StepCtl &c, int nvl, int col, int k,
                                             size_t toff, size_t aoff, bool live, const float *trdp_src) {
  const int ncol = s.ncol, of = c.of - 1, vof = c.vof - 1;
  FctInA r;
  if (live) {
#pragma unroll
    for (int e = 0; e < NPP; e++) r.m[e] = EE(s.massfx, col, e, k);
    r.tdp0 = TRV(trdp_src, col);
    r.tlo_of = TAB(s.trl_tdcy, of, col);
    r.tlo_vof = TAB(s.trl_tdcy, vof, col);
    r.dp = EC(s.dp, col, k);
  } else {
#pragma unroll
    for (int e = 0; e < NPP; e++) r.m[e] = 0.f;
    r.tdp0 = r.tlo_of = r.tlo_vof = 0.f;
    r.dp = 1.;
  }
  return r;
}
// The HBM inputs of stage B for one tile cell.
struct FctInB {
  double dp;
  float to, tv;
};
__device__ __forceinline__ FctInB fct_load_b(const EnsState &s, const StepCtl &c, int nvl, int col, int k,
                                             size_t aoff, bool live) {
  const int ncol = s.ncol, of = c.of - 1, vof = c.vof - 1;
  FctInB r{0., 0.f, 0.f};
  if (live) {
    r.dp = EC(s.dp, col, k);
    r.to = TAB(s.trc_tdcy, of, col);
    r.tv = TAB(s.trc_tdcy, vof, col);
  }
  return r;
}

// Stage A of tile cell i with NP edges (trcadv1): the low-order fluxes and
// tendency, the bounded low-order tracer (to Slo); for rings 0-1 the two-
// sided edge values and antidiffusive fluxes a_out and their out/inflow
// sums.  Also returns the low-order tendency (written to HBM for the core)
// and 1/dp.  Lanes of tracer 1 (theta) form its edge values from the edge
// stencil, the others from the least-squares reconstructions of both cells
// of the edge (edgvar3.F90:65-132).
template <int NP>
__device__ __forceinline__ void fct_stage_a(const EnsGrid &g, const EnsState &s, const StepCtl &c, const int4 &en,
                                            int i, int n0, int n1, int t, size_t aoff, int base, int lane,
                                            bool live, const FctInA &in, const float *Str, float *Slo,
                                            const int4 *Sent, float *a_out, float &sp_out, float &sm_out,
                                            float &tlo_out, float &rdp_out) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int nf = c.nf - 1;
  const int cell = uni(en.x), col = base + cell;
  const float thshld = 1.e-11f;
  int l[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) l[e] = uni(lnb(en, e));
  float m[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) m[e] = in.m[e];
  const float tdp0 = in.tdp0, tlo_of = in.tlo_of, tlo_vof = in.tlo_vof;
  const float rdp = live ? fct_rcp(fmaxf(thshld, (float)in.dp)) : 0.f;
  const float ti = Str[i * 64 + lane];
  float v[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) v[e] = Str[l[e] * 64 + lane];
  const float rar = ldc(g.rarea, cell);
  const float a1 = (float)c.ab1, a2 = (float)c.ab2, a3 = (float)c.ab3;
  float acc = 0.f, mx = ti, mn = ti, fl[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) {
    mx = fmaxf(mx, v[e]);
    mn = fminf(mn, v[e]);
  }
#pragma unroll
  for (int e = 0; e < NP; e++) {
    fl[e] = .5f * ((m[e] + fabsf(m[e])) * ti - (fabsf(m[e]) - m[e]) * v[e]);
    acc = acc + fl[e];
  }
  const float tlo = -acc * rar;
  const float lo = tdp0 + a1 * tlo + a2 * tlo_of + a3 * tlo_vof;
  mx = fmaxf(0.f, mx);
  mn = fmaxf(0.f, mn);
  Slo[i * 64 + lane] = fmaxf(mn, fminf(mx, lo * rdp));
  tlo_out = tlo;
  rdp_out = rdp;
  if (i < n0 && live) TAB(s.trl_tdcy, nf, col) = tlo;
  if (i >= n1) return;
  float te[NP];
  if (t == 0) {
#pragma unroll
    for (int e = 0; e < NP; e++) {
      const int et = (e + NP - 1) % NP, el = (e + 1) % NP;
      te[e] = fct_theta_edge(ti, v[e], v[et], v[el], Sent[l[et]].x, Sent[l[el]].x);
    }
  } else {
    // this cell's reconstructions (edgvar3.F90:65-110), then the neighbour's
    // of the same edge, averaged (edgvar3.F90:111-132)
    float d[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) d[j] = v[j] - ti;
#pragma unroll
    for (int e = 0; e < NP; e++) {
      float ac = 0.f;
#pragma unroll
      for (int j = 0; j < NP; j++) ac = __builtin_fmaf(d[j], FHF(cell, e, j), ac);
      const float eo = ti + ac;
      const int4 ex = Sent[l[e]];
      const int xcell = uni(ex.x), xs = uni(eslot(en, e)), xnp = uni(enp(ex));
      const float tx = v[e];
      float ax = 0.f;
#pragma unroll
      for (int j = 0; j < NPP; j++) {
        if (j < xnp) ax = __builtin_fmaf(Str[uni(lnb(ex, j)) * 64 + lane] - tx, FHF(xcell, xs, j), ax);
      }
      const float en_ = tx + ax;
      te[e] = .5f * (eo + en_);
    }
  }
  float sp = 0.f, sm = 0.f;
#pragma unroll
  for (int e = 0; e < NP; e++) {
    const float a = m[e] * te[e] - fl[e];
    a_out[e] = a;
    sp = sp - fminf(0.f, a);
    sm = sm + fmaxf(0.f, a);
  }
  if (NP < NPP) a_out[NPP - 1] = 0.f;
  sp_out = sp;
  sm_out = sm;
}

// Stage B (trcadv2): local bounds from the tracer and the low-order tracer
// of the cell and its neighbours, and the limiter ratios.
__device__ __forceinline__ void fct_stage_b(const EnsGrid &g, const StepCtl &c, const int4 &en, int i, int lane,
                                            const FctInB &in, const float *Str, const float *Slo, float sp,
                                            float sm, float &rp, float &rm, float &mx_o, float &mn_o) {
  const int cell = uni(en.x);
  const float thshld = 1.e-11f;
  const float dp = (float)in.dp, to = in.to, tv = in.tv;
  const float area = ldc(g.area, cell);
  const float a2 = (float)c.ab2, a3 = (float)c.ab3, ra1 = (float)(1. / c.ab1) * area;
  const float ti = Str[i * 64 + lane], lo = Slo[i * 64 + lane];
  float mx = ti, mn = ti;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const float v = Str[uni(lnb(en, e)) * 64 + lane];
    mx = fmaxf(mx, v);
    mn = fminf(mn, v);
  }
  mx = fmaxf(fmaxf(0.f, mx), lo);
  mn = fminf(fmaxf(0.f, mn), lo);
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const float v = Slo[uni(lnb(en, e)) * 64 + lane];
    mx = fmaxf(mx, v);
    mn = fminf(mn, v);
  }
  mx = fmaxf(0.f, mx);
  mn = fmaxf(0.f, mn);
  const float qp = ((mx - lo) * dp - (a2 * fminf(0.f, to) + a3 * fmaxf(0.f, tv))) * ra1;
  const float qm = ((lo - mn) * dp + (a2 * fmaxf(0.f, to) + a3 * fminf(0.f, tv))) * ra1;
  rp = fmaxf(0.f, fminf(1.f, qp * fct_rcp(fmaxf(sp, thshld))));
  rm = fmaxf(0.f, fminf(1.f, qm * fct_rcp(fmaxf(sm, thshld))));
  mx_o = mx;
  mn_o = mn;
}

// blockIdx.x: (level chunk, tracer group, patch), the chunk fastest, in an
// order that gives each XCD a contiguous range (the host pads gridDim.x to a
// multiple of 8); blockIdx.y: member.  trs/tds: the tracer and tracer-mass
// buffers read, trd/tdd: those written.
// LDS: the tile's tracer (rings 0-3), its bounded low-order tracer (rings
// 0-2), and per ring 0-1 cell a pair that holds the antidiffusive out/inflow
// sums after stage A and the limiter ratios R+/R- after stage B.
template <int TG, int NW, bool SWZ>
__global__ void __launch_bounds__(64 * NW) k_e_trc_fct(EnsGrid g, EnsState s, int nchunk, int ngrp,
                                                       const float *trs, const float *tds, float *trd, float *tdd) {
  constexpr int LC = 64 / TG;                   // levels per block
  constexpr int J0 = (FCT_TILE_MAX + NW - 1) / NW, JA = (FCT_LO_MAX + NW - 1) / NW;
  constexpr int JB = (FCT_R_MAX + NW - 1) / NW, JC = (FCT_CORE_MAX + NW - 1) / NW;
  __shared__ float Str[FCT_TILE_MAX * 64];
  __shared__ float Slo[FCT_LO_MAX * 64];
  __shared__ float Sr[FCT_R_MAX * 2 * 64];
  __shared__ int4 Sent[FCT_TILE_MAX];          // the tile's entries
  const int nvl = g.d.nvl, nip = g.d.nip, ncol = s.ncol, ntra = g.d.ntra;
  const int w = threadIdx.y, lane = threadIdx.x;
  const EnsTiles &ts = g.tiles;
  int bx = blockIdx.x;
  if (SWZ) {
    const int per = gridDim.x >> 3;
    bx = (bx & 7) * per + (bx >> 3);
  }
  const int chunk = bx % nchunk, pg = bx / nchunk, grp = pg % ngrp, p = pg / ngrp;
  if (p >= ts.npatch) return;          // the whole block: before any barrier
  const int t = grp * TG + lane / LC, k = chunk * LC + lane % LC;
  const bool live = k < nvl && t < ntra;
  const int tc = t < ntra ? t : 0, kc = k < nvl ? k : 0;   // in-range for the offsets of dead lanes
  const size_t toff = (size_t)tc * ncol * nvl + kc, aoff = (size_t)tc * 3 * ncol * nvl + kc;
  const int base = blockIdx.y * nip;
  const int4 cn = ldc4(ts.cnt, p);
  const int off = ldc(ts.off, p);
  const int n0 = cn.x, n1 = cn.y, n2 = cn.z, n3 = cn.w;
  const StepCtl c = load_ctl(s.ctl);
  // stage 0: the tracer columns of the tile, every load issued before the
  // first is stored
  {
    int4 ev[J0];
#pragma unroll
    for (int j = 0; j < J0; j++) ev[j] = ldc4(ts.ent, (size_t)off + uni(min(w + j * NW, n3 - 1)));
    float v[J0];
#pragma unroll
    for (int j = 0; j < J0; j++) v[j] = live ? TRV(trs, base + ev[j].x) : 0.f;
#pragma unroll
    for (int j = 0; j < J0; j++) {
      const int i = w + j * NW;
      if (i < n3) {
        Str[i * 64 + lane] = v[j];
        if (lane == 0) Sent[i] = ev[j];
      }
    }
  }
  __syncthreads();
  // stage A, each pass's HBM inputs loaded during the pass before; the core
  // cells (the first passes) keep their fluxes, low-order tendency, tracer
  // mass and 1/dp for stage C
  float a[JC][NPP], tlo[JC], tdp0[JC], rdp[JC];
  {
    FctInA cur, nxt;
    {
      const int i = uni(min(w, n2 - 1));
      cur = fct_load_a(s, c, nvl, base + uni(Sent[i].x), k, toff, aoff, live, tds);
    }
#pragma unroll
    for (int j = 0; j < JA; j++) {
      const int i = uni(w + j * NW);
      if (j + 1 < JA) {
        const int in = uni(min(w + (j + 1) * NW, n2 - 1));
        nxt = fct_load_a(s, c, nvl, base + uni(Sent[in].x), k, toff, aoff, live, tds);
      }
      if (i < n2) {
        const int4 en = Sent[i];
        float aa[NPP], ssp = 0.f, ssm = 0.f, tl = 0.f, rd = 0.f;
        if (uni(enp(en)) == 6)
          fct_stage_a<6>(g, s, c, en, i, n0, n1, t, aoff, base, lane, live, cur, Str, Slo, Sent, aa, ssp, ssm,
                         tl, rd);
        else
          fct_stage_a<5>(g, s, c, en, i, n0, n1, t, aoff, base, lane, live, cur, Str, Slo, Sent, aa, ssp, ssm,
                         tl, rd);
        if (i < n1) {
          Sr[(2 * i) * 64 + lane] = ssp;
          Sr[(2 * i + 1) * 64 + lane] = ssm;
        }
        if (j < JC) {
          const int jj = j < JC ? j : 0;
#pragma unroll
          for (int e = 0; e < NPP; e++) a[jj][e] = aa[e];
          tlo[jj] = tl;
          tdp0[jj] = cur.tdp0;
          rdp[jj] = rd;
        }
      }
      if (j + 1 < JA) cur = nxt;
    }
  }
  __syncthreads();
  // stage B: R+/R- replace each cell's own sums in Sr; inputs a pass ahead
  float tmx[JC], tmn[JC], tco[JC], tcv[JC];
  {
    FctInB cur, nxt;
    {
      const int i = uni(min(w, n1 - 1));
      cur = fct_load_b(s, c, nvl, base + uni(Sent[i].x), k, aoff, live);
    }
#pragma unroll
    for (int j = 0; j < JB; j++) {
      const int i = uni(w + j * NW);
      if (j + 1 < JB) {
        const int in = uni(min(w + (j + 1) * NW, n1 - 1));
        nxt = fct_load_b(s, c, nvl, base + uni(Sent[in].x), k, aoff, live);
      }
      if (i < n1) {
        const int4 en = Sent[i];
        float mx, mn, rp, rm;
        fct_stage_b(g, c, en, i, lane, cur, Str, Slo, Sr[(2 * i) * 64 + lane], Sr[(2 * i + 1) * 64 + lane], rp,
                    rm, mx, mn);
        Sr[(2 * i) * 64 + lane] = rp;
        Sr[(2 * i + 1) * 64 + lane] = rm;
        if (j < JC) {
          const int jj = j < JC ? j : 0;
          tmx[jj] = mx;
          tmn[jj] = mn;
          tco[jj] = cur.to;
          tcv[jj] = cur.tv;
        }
      }
      if (j + 1 < JB) cur = nxt;
    }
  }
  __syncthreads();
  // stage C: the core
  const int nf = c.nf - 1;
  const float a1 = (float)c.ab1, a2 = (float)c.ab2, a3 = (float)c.ab3;
#pragma unroll
  for (int j = 0; j < JC; j++) {
    const int i = uni(w + j * NW);
    if (i < n0 && live) {
      const int4 en = Sent[i];
      const int cell = uni(en.x), col = base + cell;
      const int np = uni(enp(en));
      const float rar = ldc(g.rarea, cell);
      const float rpi = Sr[(2 * i) * 64 + lane], rmi = Sr[(2 * i + 1) * 64 + lane];
      float r[NPP];
#pragma unroll
      for (int e = 0; e < NPP; e++) {
        const int x = uni(lnb(en, e));
        r[e] = Sr[(2 * x + (a[j][e] >= 0.f ? 0 : 1)) * 64 + lane];
      }
      float at = 0.f;
#pragma unroll
      for (int e = 0; e < NPP; e++)
        if (e < np) at = at + a[j][e] * fminf(a[j][e] >= 0.f ? rmi : rpi, r[e]);
      at = -at * rar;
      const float tt = tlo[j] + at;
      const float tdp = tdp0[j] + a1 * tt + a2 * tco[j] + a3 * tcv[j];
      TAB(s.trc_tdcy, nf, col) = tt;
      TRV(tdd, col) = tdp;
      // a land column keeps its tracers
      TRV(trd, col) = col_wet(s.wet, col) ? fmaxf(tmn[j], fminf(tmx[j], tdp * rdp[j])) : Str[i * 64 + lane];
    }
  }
}

// Launch shapes (the autotuner's tiles of this kernel): tile & 1 selects 16
// or 8 wavefronts per block, tile & 2 half the tracers per block (twice the
// levels), tile & 4 the XCD-contiguous block order.  The tracers per block
// are 4 for ntra >= 3, else ntra.  Every shape performs the same operations,
// so the results do not depend on it.
template <int TG, int NW, bool SWZ>
static void launch_fct(const EnsGrid &g, const EnsState &s, hipStream_t st) {
  constexpr int LC = 64 / TG;
  const int nchunk = (g.d.nvl + LC - 1) / LC, ngrp = (g.d.ntra + TG - 1) / TG;
  int gx = g.tiles.npatch * nchunk * ngrp;
  if (SWZ) gx = (gx + 7) / 8 * 8;
  k_e_trc_fct<TG, NW, SWZ><<<dim3(gx, s.members), dim3(64, NW), 0, st>>>(g, s, nchunk, ngrp, s.tr, s.trdp,
                                                                         s.tr_alt, s.trdp_alt);
}

template <int TG>
static void launch_fct_tg(int tile, const EnsGrid &g, const EnsState &s, hipStream_t st) {
  switch (tile & 5) {
    case 0: launch_fct<TG, 16, false>(g, s, st); break;
    case 1: launch_fct<TG, 8, false>(g, s, st); break;
    case 4: launch_fct<TG, 16, true>(g, s, st); break;
    default: launch_fct<TG, 8, true>(g, s, st); break;
  }
}

void ens_launch_trc_fct(int tile, const EnsGrid &g, const EnsState &s, hipStream_t st) {
  int tg = g.d.ntra >= 3 ? 4 : g.d.ntra;
  if ((tile & 2) && tg > 1) tg >>= 1;
  if (tg == 4) launch_fct_tg<4>(tile, g, s, st);
  else if (tg == 2) launch_fct_tg<2>(tile, g, s, st);
  else launch_fct_tg<1>(tile, g, s, st);
}

}  // namespace fimhip
