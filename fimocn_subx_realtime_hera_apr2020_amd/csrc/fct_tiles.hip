// The tracer flux-corrected transport of a step in one kernel (the batched
// mixed-precision form): trcadv1-3 (trcadv_compute.F90:19-315) together with
// the least-squares edge reconstructions of tracers 2..ntra that edgvar3
// forms for them (edgvar3.F90:65-163), without any intermediate array in HBM.
//
// The cells are partitioned into patches of up to FCT_CORE_MAX cells.  One
// block handles one patch of one tracer of one member over a chunk of levels:
// it stages the tracer column chunk of every cell within three neighbour
// rings of the patch in LDS (lanes along the level, the reference's
// (nvl,nip) layout), then
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
// A block covers LC levels of its patch's columns (LC = 64: one level chunk,
// one cell per wavefront; LC = 32: two chunks, two cells per wavefront, half
// the LDS per block so two blocks share a CU, and a 32-layer instance fills
// every lane).  Lane l of wavefront w works on level chunk * LC + l % LC of
// the tile cells (w * CPW + l / LC) + j * NW * CPW, j = 0, 1, ... (NW
// wavefronts per block).

#define EC(a, col, k) (a)[(size_t)(col) * nvl + (k)]
#define EE(a, col, e, k) (a)[((size_t)(col) * NPP + (e)) * nvl + (k)]
#define ETR(a, t, col, k) (a)[((size_t)(t) * ncol + (col)) * nvl + (k)]
#define ETAB(a, t, sl, col, k) (a)[(((size_t)(t) * 3 + (sl)) * ncol + (col)) * nvl + (k)]
#define FHF(i, a, b) ldc(g.hfcoef, ((size_t)(i) * NPP + (a)) * NPP + (b))

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

// a value that is wave-uniform when a wavefront holds one cell
template <int CPW>
__device__ __forceinline__ int fct_uni(int x) { return CPW == 1 ? __builtin_amdgcn_readfirstlane(x) : x; }

// Theta on the edge between cell i and x (edgvar1.F90:425-511, eqwgt), the
// outer stencil cells t, l added in column order (as e_theta_edge).
__device__ __forceinline__ float fct_theta_edge(float thi, float thx, float tht, float thl, int nbt, int nbl) {
  float ta = tht, tb = thl;
  if (nbt > nbl) { const float t = ta; ta = tb; tb = t; }
  return (2.f * (thi + thx) + ta + tb) * (1.f / 6.f);
}

__device__ __forceinline__ float fct_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Stage A of tile cell i with NP edges (trcadv1): the low-order fluxes and
// tendency, the bounded low-order tracer (to Slo); for rings 0-1 the two-
// sided edge values and antidiffusive fluxes a_out and their out/inflow
// sums.  Also returns the low-order tendency (written to HBM for the core),
// the tracer mass at the start and 1/dp.
template <int NP, int LC, int CPW>
__device__ __forceinline__ void fct_stage_a(const EnsGrid &g, const EnsState &s, const StepCtl &c, const int4 &en,
                                            int i, int n0, int n1, int t, int base, int k, int kk, bool live,
                                            const float *__restrict__ trdp_src, const float *Str, float *Slo,
                                            const int4 *Sent, float *a_out, float &sp_out,
                                            float &sm_out, float &tlo_out, float &tdp0_out, float &rdp_out) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const int cell = en.x, col = base + cell;
  const float thshld = 1.e-11f;
  int l[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) l[e] = lnb(en, e);
  float m[NP];
  float tdp0 = 0.f, tlo_of = 0.f, tlo_vof = 0.f, rdp = 0.f;
  if (live) {
#pragma unroll
    for (int e = 0; e < NP; e++) m[e] = EE(s.massfx, col, e, k);
    tdp0 = ETR(trdp_src, t, col, k);
    tlo_of = ETAB(s.trl_tdcy, t, of, col, k);
    tlo_vof = ETAB(s.trl_tdcy, t, vof, col, k);
    rdp = fct_rcp(fmaxf(thshld, (float)EC(s.dp, col, k)));
  } else {
#pragma unroll
    for (int e = 0; e < NP; e++) m[e] = 0.f;
  }
  const float ti = Str[i * LC + kk];
  float v[NP];
#pragma unroll
  for (int e = 0; e < NP; e++) v[e] = Str[l[e] * LC + kk];
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
  Slo[i * LC + kk] = fmaxf(mn, fminf(mx, lo * rdp));
  tlo_out = tlo;
  tdp0_out = tdp0;
  rdp_out = rdp;
  if (i < n0 && live) ETAB(s.trl_tdcy, t, nf, col, k) = tlo;
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
    int4 ex[NP];
#pragma unroll
    for (int e = 0; e < NP; e++) ex[e] = Sent[l[e]];
#pragma unroll
    for (int e = 0; e < NP; e++) {
      float ac = 0.f;
#pragma unroll
      for (int j = 0; j < NP; j++) ac = ac + d[j] * FHF(cell, e, j);
      const float eo = ti + ac;
      const int xcell = ex[e].x, xs = eslot(en, e), xnp = enp(ex[e]);
      const float tx = v[e];
      float ax = 0.f;
#pragma unroll
      for (int j = 0; j < NPP; j++) {
        if (j < xnp) ax = ax + (Str[lnb(ex[e], j) * LC + kk] - tx) * FHF(xcell, xs, j);
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
template <int LC>
__device__ __forceinline__ void fct_stage_b(const EnsGrid &g, const EnsState &s, const StepCtl &c, const int4 &en,
                                            int i, int t, int base, int k, int kk, bool live, const float *Str,
                                            const float *Slo, float sp, float sm, float &rp, float &rm, float &mx_o,
                                            float &mn_o, float &to_o, float &tv_o) {
  const int nvl = g.d.nvl, ncol = s.ncol;
  const int of = c.of - 1, vof = c.vof - 1;
  const int cell = en.x, col = base + cell;
  const float thshld = 1.e-11f;
  float dp = 0.f, to = 0.f, tv = 0.f;
  if (live) {
    dp = (float)EC(s.dp, col, k);
    to = ETAB(s.trc_tdcy, t, of, col, k);
    tv = ETAB(s.trc_tdcy, t, vof, col, k);
  }
  const float area = ldc(g.area, cell);
  const float a2 = (float)c.ab2, a3 = (float)c.ab3, ra1 = (float)(1. / c.ab1) * area;
  const float ti = Str[i * LC + kk], lo = Slo[i * LC + kk];
  float mx = ti, mn = ti;
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const float v = Str[lnb(en, e) * LC + kk];
    mx = fmaxf(mx, v);
    mn = fminf(mn, v);
  }
  mx = fmaxf(fmaxf(0.f, mx), lo);
  mn = fminf(fmaxf(0.f, mn), lo);
#pragma unroll
  for (int e = 0; e < NPP; e++) {
    const float v = Slo[lnb(en, e) * LC + kk];
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
  to_o = to;
  tv_o = tv;
}

// blockIdx.x: (patch, level chunk, tracer) in an order that gives each XCD
// a contiguous range of patches (the host pads gridDim.x to a multiple of
// 8); blockIdx.y: member.  trs/tds: the tracer and tracer-mass buffers read,
// trd/tdd: those written.
// LDS: the tile's tracer (rings 0-3), its bounded low-order tracer (rings
// 0-2), and per ring 0-1 cell a pair that holds the antidiffusive out/inflow
// sums after stage A and the limiter ratios R+/R- after stage B.
// With 8 wavefronts per block the register allocation is capped so that
// two blocks (and their LDS tiles) share a CU.
template <int LC, int NW, bool SWZ>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW == 8 ? 4 : 1))) k_e_trc_fct(EnsGrid g, EnsState s, int nchunk, const float *trs,
                                                       const float *tds, float *trd, float *tdd) {
  constexpr int CPW = 64 / LC;                  // cells per wavefront
  constexpr int STEP = NW * CPW;                // tile cells per pass of the block
  constexpr int J0 = (FCT_TILE_MAX + STEP - 1) / STEP, JA = (FCT_LO_MAX + STEP - 1) / STEP;
  constexpr int JB = (FCT_R_MAX + STEP - 1) / STEP, JC = (FCT_CORE_MAX + STEP - 1) / STEP;
  __shared__ float Str[FCT_TILE_MAX * LC];
  __shared__ float Slo[FCT_LO_MAX * LC];
  __shared__ float Sr[FCT_R_MAX * 2 * LC];
  __shared__ int4 Sent[FCT_TILE_MAX];          // the tile's entries
  const int nvl = g.d.nvl, nip = g.d.nip, ncol = s.ncol, ntra = g.d.ntra;
  const int w = threadIdx.y, kk = threadIdx.x % LC, h = threadIdx.x / LC;
  const EnsTiles &ts = g.tiles;
  int bx = blockIdx.x;
  if (SWZ) {
    const int per = gridDim.x >> 3;
    bx = (bx & 7) * per + (bx >> 3);
  }
  const int t = bx % ntra, pc = bx / ntra, chunk = pc % nchunk, p = pc / nchunk;
  if (p >= ts.npatch) return;          // the whole block: before any barrier
  const int k = chunk * LC + kk;
  const int base = blockIdx.y * nip;
  const int4 cn = ldc4(ts.cnt, p);
  const int off = ldc(ts.off, p);
  const int n0 = cn.x, n1 = cn.y, n2 = cn.z, n3 = cn.w;
  const bool live = k < nvl;
  const StepCtl c = load_ctl(s.ctl);
  const float *trf = trs + (size_t)t * ncol * nvl;
  const int slot = w * CPW + h;
  // stage 0: the tracer columns of the tile, every load issued before the
  // first is stored
  {
    int4 ev[J0];
#pragma unroll
    for (int j = 0; j < J0; j++) ev[j] = ldc4(ts.ent, (size_t)off + fct_uni<CPW>(min(slot + j * STEP, n3 - 1)));
    float v[J0];
#pragma unroll
    for (int j = 0; j < J0; j++) v[j] = live ? trf[(size_t)(base + ev[j].x) * nvl + k] : 0.f;
#pragma unroll
    for (int j = 0; j < J0; j++) {
      const int i = slot + j * STEP;
      if (i < n3) {
        Str[i * LC + kk] = v[j];
        if (kk == 0) Sent[i] = ev[j];
      }
    }
  }
  __syncthreads();
  // stage A; the core cells keep their fluxes for stage C, and when a lane
  // holds at most two core cells also their low-order tendency, tracer mass,
  // 1/dp and old tendencies (otherwise stage C loads them again)
  constexpr bool KEEP = JC <= 2;
  constexpr int JK = KEEP ? JC : 1;
  float a[JC][NPP], tlo[JK], tdp0[JK], rdp[JK];
#pragma unroll
  for (int j = 0; j < JA; j++) {
    const int i = fct_uni<CPW>(slot + j * STEP);
    if (i < n2) {
      const int4 en = Sent[i];
      float aa[NPP], ssp = 0.f, ssm = 0.f, tl = 0.f, td = 0.f, rd = 0.f;
      if (enp(en) == 6)
        fct_stage_a<6, LC, CPW>(g, s, c, en, i, n0, n1, t, base, k, kk, live, tds, Str, Slo, Sent, aa, ssp,
                                ssm, tl, td, rd);
      else
        fct_stage_a<5, LC, CPW>(g, s, c, en, i, n0, n1, t, base, k, kk, live, tds, Str, Slo, Sent, aa, ssp,
                                ssm, tl, td, rd);
      if (i < n1) {
        Sr[(2 * i) * LC + kk] = ssp;
        Sr[(2 * i + 1) * LC + kk] = ssm;
      }
      if (j < JC) {
        const int jj = j < JC ? j : 0;
#pragma unroll
        for (int e = 0; e < NPP; e++) a[jj][e] = aa[e];
        if (KEEP) {
          const int jk = j < JK ? j : 0;
          tlo[jk] = tl;
          tdp0[jk] = td;
          rdp[jk] = rd;
        }
      }
    }
  }
  __syncthreads();
  // stage B: R+/R- replace each cell's own sums in Sr
  float tmx[JC], tmn[JC], tco[JK], tcv[JK];
#pragma unroll
  for (int j = 0; j < JB; j++) {
    const int i = fct_uni<CPW>(slot + j * STEP);
    if (i < n1) {
      const int4 en = Sent[i];
      float mx, mn, to, tv, rp, rm;
      fct_stage_b<LC>(g, s, c, en, i, t, base, k, kk, live, Str, Slo, Sr[(2 * i) * LC + kk],
                      Sr[(2 * i + 1) * LC + kk], rp, rm, mx, mn, to, tv);
      Sr[(2 * i) * LC + kk] = rp;
      Sr[(2 * i + 1) * LC + kk] = rm;
      if (j < JC) {
        const int jj = j < JC ? j : 0;
        tmx[jj] = mx;
        tmn[jj] = mn;
        if (KEEP) {
          const int jk = j < JK ? j : 0;
          tco[jk] = to;
          tcv[jk] = tv;
        }
      }
    }
  }
  __syncthreads();
  // stage C: the core
  const int nf = c.nf - 1, of = c.of - 1, vof = c.vof - 1;
  const float a1 = (float)c.ab1, a2 = (float)c.ab2, a3 = (float)c.ab3;
#pragma unroll
  for (int j = 0; j < JC; j++) {
    const int i = fct_uni<CPW>(slot + j * STEP);
    if (i < n0 && live) {
      const int4 en = Sent[i];
      const int cell = en.x, col = base + cell;
      const int np = enp(en);
      const float rar = ldc(g.rarea, cell);
      const int jk = KEEP && j < JK ? j : 0;
      // the same values stage A and B used (stage A wrote the low-order tendency)
      const float tlo_j = KEEP ? tlo[jk] : ETAB(s.trl_tdcy, t, nf, col, k);
      const float tdp0_j = KEEP ? tdp0[jk] : ETR(tds, t, col, k);
      const float rdp_j = KEEP ? rdp[jk] : fct_rcp(fmaxf(1.e-11f, (float)EC(s.dp, col, k)));
      const float tco_j = KEEP ? tco[jk] : ETAB(s.trc_tdcy, t, of, col, k);
      const float tcv_j = KEEP ? tcv[jk] : ETAB(s.trc_tdcy, t, vof, col, k);
      const float rpi = Sr[(2 * i) * LC + kk], rmi = Sr[(2 * i + 1) * LC + kk];
      float r[NPP];
#pragma unroll
      for (int e = 0; e < NPP; e++) {
        const int x = lnb(en, e);
        r[e] = Sr[(2 * x + (a[j][e] >= 0.f ? 0 : 1)) * LC + kk];
      }
      float at = 0.f;
#pragma unroll
      for (int e = 0; e < NPP; e++)
        if (e < np) at = at + a[j][e] * fminf(a[j][e] >= 0.f ? rmi : rpi, r[e]);
      at = -at * rar;
      const float tt = tlo_j + at;
      const float tdp = tdp0_j + a1 * tt + a2 * tco_j + a3 * tcv_j;
      ETAB(s.trc_tdcy, t, nf, col, k) = tt;
      ETR(tdd, t, col, k) = tdp;
      // a land column keeps its tracers
      ETR(trd, t, col, k) = col_wet(s.wet, col) ? fmaxf(tmn[j], fminf(tmx[j], tdp * rdp_j)) : Str[i * LC + kk];
    }
  }
}

// Launch shapes (the autotuner's tiles of this kernel): tile & 3 selects
// (levels per block, wavefronts per block) = (32, 16), (32, 8), (64, 16),
// (64, 8); tile & 4 the XCD-contiguous block order.  Every shape performs the
// same operations, so the results do not depend on it.
template <int LC, int NW, bool SWZ>
static void launch_fct(const EnsGrid &g, const EnsState &s, hipStream_t st) {
  const int nchunk = (g.d.nvl + LC - 1) / LC;
  int gx = g.tiles.npatch * nchunk * g.d.ntra;
  if (SWZ) gx = (gx + 7) / 8 * 8;
  k_e_trc_fct<LC, NW, SWZ><<<dim3(gx, s.members), dim3(64, NW), 0, st>>>(g, s, nchunk, s.tr, s.trdp, s.tr_alt,
                                                                         s.trdp_alt);
}

void ens_launch_trc_fct(int tile, const EnsGrid &g, const EnsState &s, hipStream_t st) {
  switch (tile & 7) {
    case 0: launch_fct<32, 16, false>(g, s, st); break;
    case 1: launch_fct<32, 8, false>(g, s, st); break;
    case 2: launch_fct<64, 16, false>(g, s, st); break;
    case 3: launch_fct<64, 8, false>(g, s, st); break;
    case 4: launch_fct<32, 16, true>(g, s, st); break;
    case 5: launch_fct<32, 8, true>(g, s, st); break;
    case 6: launch_fct<64, 16, true>(g, s, st); break;
    default: launch_fct<64, 8, true>(g, s, st); break;
  }
}

}  // namespace fimhip
