// C ABI of the batched mixed-precision form (fimhip_ens_*): creation,
// seeded member states, start-up, the graph-captured step, downloads, global
// sums with FP32 Kahan partials, per-kernel launch shapes and their
// autotuner, and per-kernel event timing.
#include <stdio.h>
#include <string.h>

#include <limits>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "ens_internal.h"
#include "fimhip.h"
#include "host_common.h"

using namespace fimhip;

struct fimhip_ens {
  int device = 0;
  hipStream_t stream = nullptr;
  int nip = 0, nvl = 0, ntra = 0, members = 0;
  double dt = 0.;
  EnsGrid g{};
  EnsState s{};
  Arena arena;
  DevState stage{};                  // one member in FP64 (synth / load staging)
  std::vector<double> lat, lon;
  int tile[EK_NKERNELS];
  int flux_conserv_schm = 1;         // 1 = cnuity (FCT), 2 = mass_adv
  EnsClassB cb{};                    // class-B transport (cb.ntrb = 0: off)
  // globsum partials
  int gs_blocks = 0;
  double *gs64 = nullptr;
  float *gs32 = nullptr;
  // captured step
  bool use_graph = true, overlap = true;
  int graph_mode = -1;
  // a second stream for the kernels of a step that do not depend on each
  // other (edgvar1 / edgvar3; momtum / the tracer FCT and the first part of
  // hybgen), forked and joined with events; used in the captured graph
  hipStream_t side = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // one captured step per buffer parity (the fused tracer FCT reads tr/trdp
  // and writes tr_alt/trdp_alt, and the instance swaps them after the step)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int graph_modes[2] = {-1, -1};
  KTimer *kt = nullptr;
  // the tracer FCT as one kernel over patches with recomputed halo rings
  // (fct_tiles.hip), or as the three-kernel chain with edgvar3's tracer
  // reconstructions in HBM; both give the same bits.  The chain is the
  // default (faster at 16 members x G7 x 64 on MI355X); the autotuner keeps
  // whichever form measures faster
  bool fused = false;
};

static void drop_graph(fimhip_ens *e) {
  for (int p = 0; p < 2; p++) {
    if (e->exec[p]) (void)hipGraphExecDestroy(e->exec[p]);
    if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
    e->exec[p] = nullptr;
    e->graph[p] = nullptr;
    e->graph_modes[p] = -1;
  }
  e->graph_mode = -1;
}

// after a step with the fused tracer FCT the new tracers are in the other buffer
static void swap_tracer_buffers(fimhip_ens *e) {
  std::swap(e->s.tr, e->s.tr_alt);
  std::swap(e->s.trdp, e->s.trdp_alt);
}

template <class T>
static int to_float(Arena &a, const T **out, const double *src, size_t n) {
  std::vector<float> h(n);
  for (size_t i = 0; i < n; i++) h[i] = (float)src[i];
  float *p = nullptr;
  int r = a.upload(&p, h);
  *out = p;
  return r;
}

extern "C" int fimhip_ens_create(const fimhip_grid_desc *gd, int members, int device, fimhip_ens **out) {
  if (!gd || !out) return FIMHIP_EINVAL;
  *out = nullptr;
  const int nip = gd->nip, nvl = gd->nvl, ntra = gd->ntra;
  if (members < 1 || members > 4096 || nip < 12 || nvl < 4 || nvl > NVL_MAX || ntra < 1 || ntra > NTRA_MAX ||
      !(gd->dt > 0.))
    return FIMHIP_EINVAL;
  if (!gd->nprox || !gd->prox || !gd->proxs || !gd->cs || !gd->sn || !gd->sidevec_c || !gd->sidevec_e ||
      !gd->sideln || !gd->hfcoef || !gd->area || !gd->lat || !gd->lon || !gd->thetac || !gd->dpsig)
    return FIMHIP_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return FIMHIP_ENODEV;
  HostGrid hg;
  int r = host_grid_prepare(gd, hg);
  if (r) return r;
  fimhip_ens *e = new fimhip_ens();
  e->device = device;
  e->nip = nip; e->nvl = nvl; e->ntra = ntra; e->members = members; e->dt = gd->dt;
  e->lat.assign(gd->lat, gd->lat + nip);
  e->lon.assign(gd->lon, gd->lon + nip);
  for (int k = 0; k < EK_NKERNELS; k++) e->tile[k] = 2;   // 4 columns per block, no swizzle
  e->tile[EK_TRCFCT] = 4;        // 32 levels and 16 wavefronts per block, XCD-contiguous order
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&e->stream) != hipSuccess) {
    delete e;
    return FIMHIP_EHIP;
  }
  e->arena.stream = e->stream;
  if (hipStreamCreate(&e->side) != hipSuccess) { fimhip_ens_destroy(e); return FIMHIP_EHIP; }
  for (auto &v : e->ev)
    if (hipEventCreateWithFlags(&v, hipEventDisableTiming) != hipSuccess) { fimhip_ens_destroy(e); return FIMHIP_EHIP; }
  Arena &a = e->arena;
  EnsGrid &g = e->g;
  if ((r = dev_grid_upload(a, gd, hg, g.d))) { fimhip_ens_destroy(e); return r; }
  if ((r = to_float(a, &g.cs, gd->cs, (size_t)nip * NPP * 4)) || (r = to_float(a, &g.sn, gd->sn, (size_t)nip * NPP * 4)) ||
      (r = to_float(a, &g.svc, gd->sidevec_c, (size_t)nip * NPP * 2)) ||
      (r = to_float(a, &g.sve, gd->sidevec_e, (size_t)nip * NPP * 2)) ||
      (r = to_float(a, &g.sideln, gd->sideln, (size_t)nip * NPP)) ||
      (r = to_float(a, &g.hfcoef, gd->hfcoef, (size_t)nip * NPP * NPP)) ||
      (r = to_float(a, &g.area, gd->area, nip)) || (r = to_float(a, &g.rarea, hg.rarea.data(), nip)) ||
      (r = to_float(a, &g.corio, hg.corio.data(), nip)) || (r = to_float(a, &g.grndspd, hg.grndspd.data(), nip)) ||
      (r = to_float(a, &g.dfflen, hg.dfflen.data(), nvl))) {
    fimhip_ens_destroy(e);
    return r;
  }
  const size_t ncol = (size_t)members * nip;
  const size_t cz = ncol * nvl, lz = ncol * (nvl + 1), ez = ncol * NPP * nvl;
  EnsState &s = e->s;
  s.members = members;
  s.ncol = (int)ncol;
  struct { float **p; size_t n; } f32[] = {
      {&s.us, cz}, {&s.vs, cz}, {&s.u2, cz}, {&s.v2, cz}, {&s.tr, cz * ntra}, {&s.trdp, cz * ntra},
      {&s.tr_alt, cz * ntra}, {&s.trdp_alt, cz * ntra},
      {&s.u_tdcy, 3 * cz}, {&s.v_tdcy, 3 * cz}, {&s.dpl_tdcy, 3 * cz}, {&s.trc_tdcy, 3 * cz * ntra},
      {&s.trl_tdcy, 3 * cz * ntra}, {&s.massfx, ez}, {&s.ws, cz}, {&s.relvor, cz},
      {&s.vn_edg, ez}, {&s.dp_edg, ez}, {&s.pgx, cz}, {&s.pgy, cz}, {&s.trc_edg, ez * (ntra > 1 ? ntra - 1 : 1)}, {&s.th0, cz},
      {&s.antifx, ez}, {&s.delp_lo, cz}, {&s.r_plus, cz}, {&s.r_mnus, cz}, {&s.antiflx, ez * ntra},
      {&s.trcr_lo, cz * ntra}, {&s.trmax, cz * ntra}, {&s.trmin, cz * ntra}, {&s.tr_rp, cz * ntra},
      {&s.tr_rm, cz * ntra}};
  struct { double **p; size_t n; } f64[] = {
      {&s.dp, cz}, {&s.pr, lz}, {&s.ex, lz}, {&s.ph, lz}, {&s.mp, cz}, {&s.sdot, lz}, {&s.exsmo, lz},
      {&s.psrf, ncol}, {&s.ptdcy, 2 * ncol}, {&s.dp_tdcy, 3 * cz}, {&s.thbuf, cz}};
  for (auto &x : f32)
    if ((r = a.alloc(x.p, x.n))) { fimhip_ens_destroy(e); return r; }
  for (auto &x : f64)
    if ((r = a.alloc(x.p, x.n))) { fimhip_ens_destroy(e); return r; }
  if ((r = a.alloc(&s.ctl, 1))) { fimhip_ens_destroy(e); return r; }
  {
    TileHost th;
    if ((r = ens_build_tiles(hg, nip, th)) || (r = ens_upload_tiles(a, th, g.tiles))) {
      fimhip_ens_destroy(e);
      return r;
    }
  }
  // staging member (FP64, plain layout)
  const size_t mcz = (size_t)nip * nvl, mlz = (size_t)nip * (nvl + 1);
  DevState &t = e->stage;
  struct { double **p; size_t n; } st[] = {{&t.us, mcz}, {&t.vs, mcz}, {&t.dp, mcz}, {&t.pr, mlz},
                                           {&t.ex, mlz}, {&t.ph, mlz}, {&t.tr, mcz * ntra}};
  for (auto &x : st)
    if ((r = a.alloc(x.p, x.n))) { fimhip_ens_destroy(e); return r; }
  e->gs_blocks = 256;
  if ((r = a.alloc(&e->gs64, (size_t)members * (ntra + 1) * e->gs_blocks)) ||
      (r = a.alloc(&e->gs32, 2 * (size_t)members * (ntra + 1) * e->gs_blocks))) {
    fimhip_ens_destroy(e);
    return r;
  }
  if (hipStreamSynchronize(e->stream) != hipSuccess) { fimhip_ens_destroy(e); return FIMHIP_EHIP; }
  *out = e;
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_destroy(fimhip_ens *e) {
  if (!e) return FIMHIP_EINVAL;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(e);
  delete e->kt;
  e->arena.release();
  for (auto v : e->ev)
    if (v) (void)hipEventDestroy(v);
  if (e->side) (void)hipStreamDestroy(e->side);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return FIMHIP_OK;
}

static int check_member(const fimhip_ens *e, int m) { return (!e || m < 0 || m >= e->members) ? FIMHIP_EINVAL : 0; }

extern "C" int fimhip_ens_synth(fimhip_ens *e, int member, unsigned long long seed, int synth_member, int flags) {
  if (check_member(e, member)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  int r = synth_into(e->nip, e->nvl, e->ntra, e->lat.data(), e->lon.data(), seed, synth_member, flags, e->stage,
                     e->stream);
  if (r) return r;
  ens_launch_convert_member(e->g, e->s, member, e->stage, e->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_load_member(fimhip_ens *e, int member, const double *us, const double *vs,
                                      const double *dp, const double *pr, const double *ex, const double *ph,
                                      const double *tr) {
  if (check_member(e, member) || !us || !vs || !dp || !pr || !ex || !ph || !tr) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  const size_t cz = (size_t)e->nip * e->nvl, lz = (size_t)e->nip * (e->nvl + 1);
  const DevState &t = e->stage;
  HIPCHK(hipMemcpyAsync(t.us, us, cz * 8, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpyAsync(t.vs, vs, cz * 8, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpyAsync(t.dp, dp, cz * 8, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpyAsync(t.pr, pr, lz * 8, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpyAsync(t.ex, ex, lz * 8, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpyAsync(t.ph, ph, lz * 8, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpyAsync(t.tr, tr, cz * e->ntra * 8, hipMemcpyHostToDevice, e->stream));
  ens_launch_convert_member(e->g, e->s, member, t, e->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  return FIMHIP_OK;
}

// hybgen's view of a state: the step's own copy, whose tracer buffers are
// the new ones after the fused tracer transport
static HybCols<float> hyb_cols(const fimhip_ens *e, const EnsState &s) {
  const DevGrid &d = e->g.d;
  return HybCols<float>{d.nip, s.ncol, d.nvl, d.ntra, s.u2, s.v2, s.us, s.vs, s.tr, s.dp, s.pr, s.ex, s.sdot,
                        s.exsmo, s.thbuf, d.thetac, d.dpsig, d.smoo, d.sideln, d.rarea, d.prox, d.nprox, s.wet};
}

// dyn_init start-up for every member (dyn_init.F90:870-1094): optional
// hybgen(its=0) and pressure refresh, hystat(0), tracer mass, tendencies
// cleared, step counter and AB slots nf=0, of=2, vof=1.
extern "C" int fimhip_ens_init_state(fimhip_ens *e, int init_hybgen) {
  if (!e) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  EnsState &s = e->s;
  const size_t cz = (size_t)s.ncol * e->nvl;
  if (init_hybgen) {
    // the start-up remap reads the winds from u2/v2
    HIPCHK(hipMemcpyAsync(s.u2, s.us, cz * 4, hipMemcpyDeviceToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(s.v2, s.vs, cz * 4, hipMemcpyDeviceToDevice, e->stream));
    launch_hybgen_cols<float>(hyb_cols(e, s), nullptr, 0, e->stream);
    ens_launch_pres_from_exner(e->g, s, e->stream);
  }
  ens_launch_hystat_init(e->g, s, e->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(s.u_tdcy, 0, 3 * cz * 4, e->stream));
  HIPCHK(hipMemsetAsync(s.v_tdcy, 0, 3 * cz * 4, e->stream));
  HIPCHK(hipMemsetAsync(s.dpl_tdcy, 0, 3 * cz * 4, e->stream));
  HIPCHK(hipMemsetAsync(s.dp_tdcy, 0, 3 * cz * 8, e->stream));
  HIPCHK(hipMemsetAsync(s.trc_tdcy, 0, 3 * cz * e->ntra * 4, e->stream));
  HIPCHK(hipMemsetAsync(s.trl_tdcy, 0, 3 * cz * e->ntra * 4, e->stream));
  HIPCHK(hipMemsetAsync(s.massfx, 0, cz * NPP * 4, e->stream));
  StepCtl c{0, 0, 2, 1, 0., 0., 0.};
  HIPCHK(hipMemcpyAsync(s.ctl, &c, sizeof c, hipMemcpyHostToDevice, e->stream));
  if (e->cb.ntrb > 0) {            // dyn_init.F90:1096-1098
    HIPCHK(hipMemsetAsync(e->cb.ring, 0, 3 * cz * NPP * sizeof(float), e->stream));
    ens_launch_class_b_init(e->g, s, e->cb, e->stream);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  return FIMHIP_OK;
}

// one step's launches (DYN_1 then DYN_2, dyn_run.F90:119-425)
// With 'side', edgvar3 runs beside edgvar1, and momtum beside the tracer FCT
// and hybgen up to remap_1d (the first reader of its winds): momtum reads only
// what cnuity and edgvar1 left and writes only what remap_1d and the next
// step read.
static void enqueue_step(fimhip_ens *e, int mode, bool side) {
  const EnsGrid &g = e->g;
  EnsState s = e->s;
  s.edg3_ntr = e->fused ? 0 : e->ntra - 1;
  hipStream_t st = e->stream, sd = e->side;
  ens_launch_abstart(s, e->dt, st);
  if (side) {
    (void)hipEventRecord(e->ev[0], st);
    (void)hipStreamWaitEvent(sd, e->ev[0], 0);
    ens_launch(EK_EDGVAR3, e->tile[EK_EDGVAR3], g, s, sd);
    (void)hipEventRecord(e->ev[1], sd);
    ens_launch(EK_EDGVAR1, e->tile[EK_EDGVAR1], g, s, st);
    (void)hipStreamWaitEvent(st, e->ev[1], 0);
  } else {
    ens_launch(EK_EDGVAR1, e->tile[EK_EDGVAR1], g, s, st);
    ens_launch(EK_EDGVAR3, e->tile[EK_EDGVAR3], g, s, st);
  }
  if (e->flux_conserv_schm == 2) {
    ens_launch(EK_MASSADV1, e->tile[EK_MASSADV1], g, s, st);
    ens_launch(EK_MASSADV2, e->tile[EK_MASSADV2], g, s, st);
  } else {
    for (int k = EK_CNUITY12; k <= EK_CNUITY45; k++) ens_launch(k, e->tile[k], g, s, st);
  }
  if (e->cb.ntrb > 0) ens_launch_class_b_step(g, s, e->cb, st);   // dyn_run.F90:388-405
  if (e->fused) {
    ens_launch(EK_TRCFCT, e->tile[EK_TRCFCT], g, s, st);
    std::swap(s.tr, s.tr_alt);         // the rest of the step works on the new tracers
    std::swap(s.trdp, s.trdp_alt);
  } else {
    for (int k = EK_TRCADV1; k <= EK_TRCADV3; k++) ens_launch(k, e->tile[k], g, s, st);
  }
  // momtum needs nothing of the tracer transport; it runs beside hybgen's
  // column kernels up to the remap (latency-bound, so the two share the
  // chip), rather than beside the bandwidth-bound tracer kernels
  hipEvent_t momtum_done = nullptr;
  if (mode == 0 && side) {
    (void)hipEventRecord(e->ev[2], st);
    (void)hipStreamWaitEvent(sd, e->ev[2], 0);
    ens_launch(EK_MOMTUM, e->tile[EK_MOMTUM], g, s, sd);
    (void)hipEventRecord(e->ev[3], sd);
    momtum_done = e->ev[3];
  }
  if (mode == 0) {
    if (!side) ens_launch(EK_MOMTUM, e->tile[EK_MOMTUM], g, s, st);
    launch_hybgen_cols<float>(hyb_cols(e, s), s.ctl, 0, st, momtum_done);
    ens_launch(EK_HYSTAT, e->tile[EK_HYSTAT], g, s, st);
  }
}

// one step enqueued directly (no graph), the tracer buffers swapped after it
static void step_once(fimhip_ens *e, int mode, bool side) {
  enqueue_step(e, mode, side);
  if (e->fused) swap_tracer_buffers(e);
}

extern "C" int fimhip_ens_step(fimhip_ens *e, int nsteps, int mode) {
  if (!e || nsteps < 0 || (mode != 0 && mode != 1)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  if (!e->use_graph || e->kt) {
    g_ktimer = e->kt;
    for (int i = 0; i < nsteps; i++) step_once(e, mode, false);
    g_ktimer = nullptr;
    HIPCHK(hipGetLastError());
    if (e->kt && e->kt->flush(e->stream)) return FIMHIP_EHIP;
    return FIMHIP_OK;
  }
  if (e->graph_mode != mode) {
    drop_graph(e);
    e->graph_mode = mode;
  }
  for (int i = 0; i < nsteps; i++) {
    // the buffer parity of this step: which of the pair holds the tracers
    const int p = (e->fused && e->s.tr > e->s.tr_alt) ? 1 : 0;
    if (e->graph_modes[p] != mode) {
      HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      enqueue_step(e, mode, e->overlap);
      HIPCHK(hipStreamEndCapture(e->stream, &e->graph[p]));
      HIPCHK(hipGraphInstantiate(&e->exec[p], e->graph[p], nullptr, nullptr, 0));
      e->graph_modes[p] = mode;
    }
    HIPCHK(hipGraphLaunch(e->exec[p], e->stream));
    if (e->fused) swap_tracer_buffers(e);
  }
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_set_fused(fimhip_ens *e, int on) {
  if (!e) return FIMHIP_EINVAL;
  if (e->fused != (on != 0)) drop_graph(e);
  e->fused = on != 0;
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_set_class_b(fimhip_ens *e, int ntrb, int dtratio) {
  if (!e || ntrb < 0 || ntrb > NTRA_MAX || dtratio < 1) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  drop_graph(e);
  if (ntrb == e->cb.ntrb) {
    e->cb.dtratio = dtratio;
    return FIMHIP_OK;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  EnsClassB b{};
  b.ntrb = ntrb;
  b.dtratio = dtratio;
  if (ntrb > 0) {
    // arrays of an earlier setting stay in the arena until the instance is destroyed
    const size_t cz = (size_t)e->s.ncol * e->nvl, ez = cz * NPP, tz = cz * ntrb;
    struct { float **p; size_t n; } f32[] = {{&b.trb, tz}, {&b.ring, 3 * ez}, {&b.cumufx, ez}, {&b.fld, tz},
                                             {&b.vfx, tz}, {&b.fmx, tz}, {&b.fmn, tz}, {&b.fld1, tz},
                                             {&b.flp, tz}, {&b.fln, tz}, {&b.uan, ez * ntrb}};
    struct { double **p; size_t n; } f64[] = {{&b.dpinit, cz}, {&b.w, cz}};
    int r;
    for (auto &x : f32)
      if ((r = e->arena.alloc(x.p, x.n))) return r;
    for (auto &x : f64)
      if ((r = e->arena.alloc(x.p, x.n))) return r;
    HIPCHK(hipMemsetAsync(b.trb, 0, tz * sizeof(float), e->stream));
    HIPCHK(hipMemsetAsync(b.ring, 0, 3 * ez * sizeof(float), e->stream));
    HIPCHK(hipMemsetAsync(b.cumufx, 0, ez * sizeof(float), e->stream));
    HIPCHK(hipMemcpyAsync(b.dpinit, e->s.dp, cz * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  e->cb = b;
  return FIMHIP_OK;
}

// class-B tracers of one member from host FP64 arrays (nvl, nip, ntrb)
extern "C" int fimhip_ens_load_class_b(fimhip_ens *e, int member, const double *trb) {
  if (check_member(e, member) || !trb || e->cb.ntrb < 1) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  const size_t cz = (size_t)e->nip * e->nvl, ecz = (size_t)e->s.ncol * e->nvl;
  std::vector<float> h(cz);
  for (int t = 0; t < e->cb.ntrb; t++) {
    for (size_t i = 0; i < cz; i++) h[i] = (float)trb[(size_t)t * cz + i];
    HIPCHK(hipMemcpyAsync(e->cb.trb + t * ecz + (size_t)member * cz, h.data(), cz * sizeof(float),
                          hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));   // h is reused for the next tracer
  }
  return FIMHIP_OK;
}

// land/sea mask of one member (nip ints, 0 = land); the first call gives the
// instance a mask with every column wet
extern "C" int fimhip_ens_set_wet(fimhip_ens *e, int member, const int *wet) {
  if (check_member(e, member) || !wet) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  int r;
  if (!e->s.wet) {
    std::vector<unsigned char> ones((size_t)e->s.ncol, 1);
    unsigned char *w = nullptr;
    if ((r = e->arena.upload(&w, ones))) return r;
    e->s.wet = w;
    drop_graph(e);
  }
  std::vector<unsigned char> h((size_t)e->nip);
  for (int i = 0; i < e->nip; i++) h[i] = wet[i] != 0;
  HIPCHK(hipMemcpyAsync((unsigned char *)e->s.wet + (size_t)member * e->nip, h.data(), h.size(),
                        hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_set_flux_scheme(fimhip_ens *e, int flux_conserv_schm) {
  if (!e || (flux_conserv_schm != 1 && flux_conserv_schm != 2)) return FIMHIP_EINVAL;
  if (e->flux_conserv_schm != flux_conserv_schm) drop_graph(e);
  e->flux_conserv_schm = flux_conserv_schm;
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_use_graph(fimhip_ens *e, int on) {
  if (!e) return FIMHIP_EINVAL;
  e->use_graph = on != 0;
  e->overlap = on != 2;       // 2: a graph without the second stream
  drop_graph(e);
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_sync(fimhip_ens *e) {
  if (!e) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(hipStreamSynchronize(e->stream));
  return FIMHIP_OK;
}

// element count of one member's field
extern "C" long long fimhip_ens_field_size(const fimhip_ens *e, int field) {
  if (!e) return FIMHIP_EINVAL;
  const long long cz = (long long)e->nip * e->nvl, lz = (long long)e->nip * (e->nvl + 1);
  switch (field) {
    case FIMHIP_US: case FIMHIP_VS: case FIMHIP_DP: case FIMHIP_MP: case FIMHIP_WS: case FIMHIP_RELVOR: return cz;
    case FIMHIP_PR: case FIMHIP_EX: case FIMHIP_PH: case FIMHIP_SDOT: return lz;
    case FIMHIP_TR: case FIMHIP_TRDP: return cz * e->ntra;
    case FIMHIP_TRB: return cz * e->cb.ntrb;
    case FIMHIP_PSRF: return e->nip;
    default: return FIMHIP_EINVAL;
  }
}

template <class T>
static int fetch(fimhip_ens *e, const T *src, size_t n, double *dst) {
  std::vector<T> h(n);
  HIPCHK(hipMemcpyAsync(h.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < n; i++) dst[i] = (double)h[i];
  return 0;
}

// one member's field, widened to FP64, in the plain (reference) layout
extern "C" int fimhip_ens_download(fimhip_ens *e, int member, int field, double *host) {
  if (check_member(e, member) || !host) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  const size_t cz = (size_t)e->nip * e->nvl, lz = (size_t)e->nip * (e->nvl + 1);
  const size_t oc = member * cz, ol = member * lz, ecz = (size_t)e->s.ncol * e->nvl;
  const EnsState &s = e->s;
  switch (field) {
    case FIMHIP_US: return fetch(e, s.us + oc, cz, host);
    case FIMHIP_VS: return fetch(e, s.vs + oc, cz, host);
    case FIMHIP_DP: return fetch(e, s.dp + oc, cz, host);
    case FIMHIP_MP: return fetch(e, s.mp + oc, cz, host);
    case FIMHIP_WS: return fetch(e, s.ws + oc, cz, host);
    case FIMHIP_RELVOR: return fetch(e, s.relvor + oc, cz, host);
    case FIMHIP_PR: return fetch(e, s.pr + ol, lz, host);
    case FIMHIP_EX: return fetch(e, s.ex + ol, lz, host);
    case FIMHIP_PH: return fetch(e, s.ph + ol, lz, host);
    case FIMHIP_SDOT: return fetch(e, s.sdot + ol, lz, host);
    case FIMHIP_PSRF: return fetch(e, s.psrf + (size_t)member * e->nip, e->nip, host);
    case FIMHIP_TR:
    case FIMHIP_TRDP: {
      const float *b = field == FIMHIP_TR ? s.tr : s.trdp;
      for (int t = 0; t < e->ntra; t++) {
        int r = fetch(e, b + t * ecz + oc, cz, host + t * cz);
        if (r) return r;
      }
      return 0;
    }
    case FIMHIP_TRB: {
      if (e->cb.ntrb < 1) return FIMHIP_EINVAL;
      for (int t = 0; t < e->cb.ntrb; t++) {
        int r = fetch(e, e->cb.trb + t * ecz + oc, cz, host + t * cz);
        if (r) return r;
      }
      return 0;
    }
    default: return FIMHIP_EINVAL;
  }
}

// Global mass (sum area*dp/g) and tracer totals (sum area*dp*tr) of every
// member (globsum.F90:70-101): mass[m], trtot[m*ntra + t] from FP64 sums;
// mass32/trtot32 (optional) from FP32 Kahan-compensated sums of the same
// products, the block partials combined in compensated FP32 pairs on the host.
extern "C" int fimhip_ens_globsum(fimhip_ens *e, double *mass, double *trtot, double *mass32, double *trtot32) {
  if (!e || !mass) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  ens_launch_globsum(e->g, e->s, e->gs64, e->gs32, e->gs_blocks, e->stream);
  HIPCHK(hipGetLastError());
  const size_t n = (size_t)e->members * (e->ntra + 1) * e->gs_blocks;
  std::vector<double> h64(n);
  std::vector<float> h32(2 * n);
  HIPCHK(hipMemcpyAsync(h64.data(), e->gs64, n * 8, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpyAsync(h32.data(), e->gs32, 2 * n * 4, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  const int nb = e->gs_blocks, nq = e->ntra + 1;
  for (int m = 0; m < e->members; m++)
    for (int q = 0; q < nq; q++) {
      const size_t o = ((size_t)m * nq + q) * nb;
      double acc = 0.;
      volatile float ss = 0.f, cc = 0.f;   // compensated FP32 pair
      for (int b = 0; b < nb; b++) {
        acc += h64[o + b];
        const float bs = h32[2 * (o + b)], bc = h32[2 * (o + b) + 1];
        const float sum = ss + bs;
        const float bb = sum - ss;
        const float err = (ss - (sum - bb)) + (bs - bb);
        cc = cc + bc - err;
        ss = sum;
      }
      const double k32 = (double)ss - (double)cc;
      if (q == 0) {
        mass[m] = acc / GRVITY;
        if (mass32) mass32[m] = k32 / GRVITY;
      } else {
        if (trtot) trtot[(size_t)m * e->ntra + q - 1] = acc;
        if (trtot32) trtot32[(size_t)m * e->ntra + q - 1] = k32;
      }
    }
  return FIMHIP_OK;
}

// ---- launch shapes -------------------------------------------------------
extern "C" int fimhip_ens_kernel_count(void) { return EK_NKERNELS; }
extern "C" const char *fimhip_ens_kernel_name(int k) {
  return (k >= 0 && k < EK_NKERNELS) ? ens_kernel_names[k] : "";
}
extern "C" int fimhip_ens_set_tile(fimhip_ens *e, int kernel, int tile) {
  if (!e || kernel < 0 || kernel >= EK_NKERNELS || tile < 0 || tile >= ENS_NTILES) return FIMHIP_EINVAL;
  if (e->tile[kernel] != tile) drop_graph(e);
  e->tile[kernel] = tile;
  return FIMHIP_OK;
}
extern "C" int fimhip_ens_get_tile(const fimhip_ens *e, int kernel) {
  if (!e || kernel < 0 || kernel >= EK_NKERNELS) return FIMHIP_EINVAL;
  return e->tile[kernel];
}

// Autotuner: for each launch shape in turn, every kernel runs with that
// shape for 'reps' steps of the given mode under event timing; each kernel
// then keeps the shape with the least mean time.  Both forms of the tracer
// transport are swept (the three-kernel chain with edgvar3's tracer
// reconstructions, then the fused kernel with edgvar3 for the thickness only),
// and the form whose transport kernels sum to less time is kept.  The steps
// advance the state, so callers re-create or reload it afterwards.  times
// (optional) receives EK_NKERNELS x ENS_NTILES mean milliseconds (edgvar3's
// row from the form kept).
extern "C" int fimhip_ens_autotune(fimhip_ens *e, int mode, int reps, double *times) {
  if (!e || reps < 1 || (mode != 0 && mode != 1)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  const int nt = EK_NKERNELS * ENS_NTILES;
  std::vector<double> best[2], tab[2];
  std::vector<int> pick[2];
  KTimer *saved = e->kt;
  for (int f = 0; f < 2; f++) {
    e->fused = f == 1;
    best[f].assign(EK_NKERNELS, 1e30);
    tab[f].assign(nt, std::numeric_limits<double>::quiet_NaN());
    pick[f].assign(e->tile, e->tile + EK_NKERNELS);
    for (int t = 0; t < ENS_NTILES; t++) {
      for (int k = 0; k < EK_NKERNELS; k++) e->tile[k] = t;
      KTimer kt;
      g_ktimer = &kt;
      for (int i = 0; i < reps; i++) step_once(e, mode, false);
      g_ktimer = nullptr;
      HIPCHK(hipGetLastError());
      if (kt.flush(e->stream)) return FIMHIP_EHIP;
      for (int k = 0; k < EK_NKERNELS; k++) {
        auto it = kt.acc.find(ens_kernel_names[k]);
        if (it == kt.acc.end() || it->second.second == 0) continue;
        const double ms = it->second.first / it->second.second;
        tab[f][k * ENS_NTILES + t] = ms;
        if (ms < best[f][k]) { best[f][k] = ms; pick[f][k] = t; }
      }
    }
  }
  e->kt = saved;
  const double chain = best[0][EK_EDGVAR3] + best[0][EK_TRCADV1] + best[0][EK_TRCADV2] + best[0][EK_TRCADV3];
  const double fused = best[1][EK_EDGVAR3] + best[1][EK_TRCFCT];
  const int fsel = fused < chain ? 1 : 0;
  e->fused = fsel == 1;
  for (int k = 0; k < EK_NKERNELS; k++) {
    // edgvar3 does different work in the two forms; the other kernels the same
    const int f = (k == EK_EDGVAR3 || best[fsel][k] <= best[1 - fsel][k]) ? fsel : 1 - fsel;
    e->tile[k] = pick[f][k];
    if (times)
      for (int t = 0; t < ENS_NTILES; t++) times[k * ENS_NTILES + t] = tab[f][k * ENS_NTILES + t];
  }
  drop_graph(e);
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_get_fused(const fimhip_ens *e) {
  if (!e) return FIMHIP_EINVAL;
  return e->fused ? 1 : 0;
}

// ---- event timing ----------------------------------------------------------
extern "C" int fimhip_ens_profile(fimhip_ens *e, int on) {
  if (!e) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(hipStreamSynchronize(e->stream));
  delete e->kt;
  e->kt = on ? new KTimer() : nullptr;
  return FIMHIP_OK;
}

extern "C" int fimhip_ens_profile_report(fimhip_ens *e, char *buf, int len) {
  if (!e || !buf || len < 1) return FIMHIP_EINVAL;
  std::string out;
  if (e->kt)
    for (auto &kv : e->kt->acc) {
      char line[160];
      snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.second, kv.second.first);
      out += line;
    }
  if ((int)out.size() >= len) return FIMHIP_EINVAL;
  memcpy(buf, out.c_str(), out.size() + 1);
  return FIMHIP_OK;
}
