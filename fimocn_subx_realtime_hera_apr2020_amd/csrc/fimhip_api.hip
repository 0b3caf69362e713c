// C ABI of the FIM dynamics library: instance creation, the per-routine
// drop-in entries with reference argument order, and the device-resident
// fast path.  See include/fimhip.h.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fimhip.h"
#include "fimhip_internal.h"
#include "fim_pow.h"
#include "synth.h"
#include "host_common.h"

using namespace fimhip;

// ---------------------------------------------------------------------------
// per-kernel event timing
namespace fimhip {
thread_local KTimer *g_ktimer = nullptr;
void kt_begin(const char *name, hipStream_t st) {
  hipEvent_t e = g_ktimer->get();
  if (e) (void)hipEventRecord(e, st);
  g_ktimer->recs.push_back({name, e, nullptr});
}
void kt_end(hipStream_t st) {
  hipEvent_t e = g_ktimer->get();
  if (e) (void)hipEventRecord(e, st);
  g_ktimer->recs.back().b = e;
}
}  // namespace fimhip

constexpr int GSUM_BLOCKS = 1024;   // blocks of the globsum kernels

struct fimhip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nip = 0, nvl = 0, ntra = 0;
  double dt = 0.;
  DevGrid g{};
  DevState s{};
  AB ab{0, 2, 1, 0., 0., 0.};
  Arena arena;
  double *gsum_partial = nullptr;   // globsum block partials
  int gsum_blocks = 0;
  std::vector<double> lat, lon;      // host copies for the synthetic-state profiles
  KTimer *kt = nullptr;              // non-null while profiling
  int flux_conserv_schm = 1;         // 1 = cnuity (FCT), 2 = mass_adv
  int dtratio = 6;                   // class-B / class-A time step ratio
  DevClassB cb{};                    // class-B tracer transport (cb.ntrb = 0: off)
};

// dyn_init.F90:125 taper for smoo_coeff
static double taper_smoo(int k1, int k2) {
  double d = (double)(k2 - k1);
  return d * d / (9. + d * d);
}

namespace fimhip {
int host_grid_prepare(const fimhip_grid_desc *gd, HostGrid &h) {
  const int nip = gd->nip, nvl = gd->nvl;
  h.prox.resize((size_t)nip * NPP);
  h.proxs.resize((size_t)nip * NPP);
  h.nprox.resize(nip);
  for (int i = 0; i < nip; i++) {
    int n = gd->nprox[i];
    if (n != 5 && n != 6) return FIMHIP_EINVAL;
    h.nprox[i] = n;
    for (int e = 0; e < NPP; e++) {
      int ee = (e < n) ? e : n - 1;
      int p = gd->prox[i * NPP + ee] - 1, q = gd->proxs[i * NPP + ee] - 1;
      if (p < 0 || p >= nip || q < 0 || q >= NPP) return FIMHIP_EINVAL;
      h.prox[i * NPP + e] = p;
      h.proxs[i * NPP + e] = q;
    }
  }
  h.rarea.resize(nip); h.corio.resize(nip); h.grndspd.resize(nip);
  h.dfflen.resize(nvl); h.smoo.resize(nvl);
  for (int i = 0; i < nip; i++) {          // dyn_init.F90:753-757
    h.corio[i] = OMEGX2 * sin(gd->lat[i]);
    h.rarea[i] = 1. / gd->area[i];
    h.grndspd[i] = EARTHRAD * 2. * PI_T * cos(gd->lat[i]) / 86164.;
  }
  int ndamp = (int)(0.1 * nvl);            // dyn_init.F90:309-321
  for (int k = 1; k <= nvl; k++) {
    double wgt = (double)(k - nvl + ndamp) / (double)ndamp;
    wgt = wgt > 0. ? wgt : 0.;
    h.dfflen[k - 1] =
        gd->dt * (VELDFF_BKGND * (1. - wgt) + (VELDFF_BKGND > VELDFF_BOOST ? VELDFF_BKGND : VELDFF_BOOST) * wgt);
    h.smoo[k - 1] = 0.;
  }
  for (int k = KBL + 2; k <= nvl; k++)
    h.smoo[k - 1] = gd->dt * INTFC_SMOOTH * taper_smoo(k, KBL + 1) * taper_smoo(k, nvl + 1);
  return 0;
}

int dev_grid_upload(Arena &a, const fimhip_grid_desc *gd, const HostGrid &h, DevGrid &g) {
  const int nip = gd->nip, nvl = gd->nvl;
  auto vec = [](const double *p, size_t n) { return std::vector<double>(p, p + n); };
  g.nip = nip; g.nvl = nvl; g.ntra = gd->ntra; g.dt = gd->dt;
  int *dprox, *dproxs, *dnprox;
  double *dcs, *dsn, *dsvc, *dsve, *dsideln, *dhf, *darea, *drarea, *dcorio, *dgrnd, *ddff, *dsmoo, *dth, *dds;
  int r = 0;
  if ((r = a.upload(&dprox, h.prox)) || (r = a.upload(&dproxs, h.proxs)) || (r = a.upload(&dnprox, h.nprox)) ||
      (r = a.upload(&dcs, vec(gd->cs, (size_t)nip * NPP * 4))) ||
      (r = a.upload(&dsn, vec(gd->sn, (size_t)nip * NPP * 4))) ||
      (r = a.upload(&dsvc, vec(gd->sidevec_c, (size_t)nip * NPP * 2))) ||
      (r = a.upload(&dsve, vec(gd->sidevec_e, (size_t)nip * NPP * 2))) ||
      (r = a.upload(&dsideln, vec(gd->sideln, (size_t)nip * NPP))) ||
      (r = a.upload(&dhf, vec(gd->hfcoef, (size_t)nip * NPP * NPP))) ||
      (r = a.upload(&darea, vec(gd->area, nip))) || (r = a.upload(&drarea, h.rarea)) ||
      (r = a.upload(&dcorio, h.corio)) || (r = a.upload(&dgrnd, h.grndspd)) ||
      (r = a.upload(&ddff, h.dfflen)) || (r = a.upload(&dsmoo, h.smoo)) ||
      (r = a.upload(&dth, vec(gd->thetac, nvl))) || (r = a.upload(&dds, vec(gd->dpsig, nvl))))
    return r;
  g.prox = dprox; g.proxs = dproxs; g.nprox = dnprox;
  g.cs = dcs; g.sn = dsn; g.svc = dsvc; g.sve = dsve; g.sideln = dsideln; g.hfcoef = dhf;
  g.area = darea; g.rarea = drarea; g.corio = dcorio; g.grndspd = dgrnd;
  g.dfflen = ddff; g.smoo = dsmoo; g.thetac = dth; g.dpsig = dds;
  return 0;
}
}  // namespace fimhip

extern "C" int fimhip_create(const fimhip_grid_desc *gd, int device, fimhip_ctx **out) {
  if (!gd || !out) return FIMHIP_EINVAL;
  *out = nullptr;
  const int nip = gd->nip, nvl = gd->nvl, ntra = gd->ntra;
  if (nip < 12 || nvl < 4 || nvl > NVL_MAX || ntra < 1 || ntra > NTRA_MAX || !(gd->dt > 0.)) return FIMHIP_EINVAL;
  if (!gd->nprox || !gd->prox || !gd->proxs || !gd->cs || !gd->sn || !gd->sidevec_c || !gd->sidevec_e ||
      !gd->sideln || !gd->hfcoef || !gd->area || !gd->lat || !gd->lon || !gd->thetac || !gd->dpsig)
    return FIMHIP_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return FIMHIP_ENODEV;
  HostGrid hg;
  int r = host_grid_prepare(gd, hg);
  if (r) return r;
  fimhip_ctx *c = new fimhip_ctx();
  c->device = device;
  c->nip = nip; c->nvl = nvl; c->ntra = ntra; c->dt = gd->dt;
  c->lat.assign(gd->lat, gd->lat + nip);
  c->lon.assign(gd->lon, gd->lon + nip);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return FIMHIP_EHIP;
  }
  c->arena.stream = c->stream;
  if ((r = dev_grid_upload(c->arena, gd, hg, c->g))) { fimhip_destroy(c); return r; }
  // ---- state and work arrays
  const size_t cz = (size_t)nip * nvl, lz = (size_t)nip * (nvl + 1), ez = (size_t)nip * NPP * nvl;
  DevState &s = c->s;
  struct { double **p; size_t n; } arrs[] = {
      {&s.us, cz}, {&s.vs, cz}, {&s.dp, cz}, {&s.pr, lz}, {&s.ex, lz}, {&s.ph, lz}, {&s.mp, cz},
      {&s.tr, cz * ntra}, {&s.trdp, cz * ntra}, {&s.u_tdcy, 3 * cz}, {&s.v_tdcy, 3 * cz},
      {&s.dp_tdcy, 3 * cz}, {&s.dpl_tdcy, 3 * cz}, {&s.trc_tdcy, 3 * cz * ntra}, {&s.trl_tdcy, 3 * cz * ntra},
      {&s.massfx, 3 * ez}, {&s.ws, cz}, {&s.relvor, cz}, {&s.sdot, lz}, {&s.psrf, (size_t)nip},
      {&s.ptdcy, 2 * (size_t)nip}, {&s.u_edg, ez}, {&s.v_edg, ez}, {&s.dp_edg, ez}, {&s.lp_edg, ez},
      {&s.mont_edg, ez}, {&s.uvsq_edg, ez}, {&s.trc_edg, ez * ntra}, {&s.antifx, ez}, {&s.delp_lo, cz},
      {&s.r_plus, cz}, {&s.r_mnus, cz}, {&s.antiflx, ez * ntra}, {&s.trcr_lo, cz * ntra},
      {&s.trmax, cz * ntra}, {&s.trmin, cz * ntra}, {&s.tr_rp, cz * ntra}, {&s.tr_rm, cz * ntra},
      {&s.u2, cz}, {&s.v2, cz}, {&s.exsmo, lz}, {&s.thbuf, cz}};
  for (auto &a : arrs)
    if ((r = c->arena.alloc(a.p, a.n))) { fimhip_destroy(c); return r; }
  c->gsum_blocks = GSUM_BLOCKS;
  if ((r = c->arena.alloc(&c->gsum_partial, (size_t)c->gsum_blocks * (1 + ntra)))) { fimhip_destroy(c); return r; }
  if (hipStreamSynchronize(c->stream) != hipSuccess) { fimhip_destroy(c); return FIMHIP_EHIP; }
  *out = c;
  return FIMHIP_OK;
}

extern "C" int fimhip_destroy(fimhip_ctx *c) {
  if (!c) return FIMHIP_EINVAL;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  delete c->kt;
  c->arena.release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FIMHIP_OK;
}

// ---------------------------------------------------------------------------
static void ab_rotate(AB &ab, int its, double dt) {   // abstart.F90:38-66
  ab.nf = ab.nf + 1; if (ab.nf == 4) ab.nf = 1;
  ab.of = ab.of + 1; if (ab.of == 4) ab.of = 1;
  ab.vof = ab.vof + 1; if (ab.vof == 4) ab.vof = 1;
  if (its == 1) { ab.ab1 = dt; ab.ab2 = 0.; ab.ab3 = 0.; }
  if (its == 2) { ab.ab1 = dt * 1.5; ab.ab2 = -dt * .5; ab.ab3 = 0.; }
  if (its == 3) { ab.ab1 = dt * 23. / 12.; ab.ab2 = -dt * 16. / 12.; ab.ab3 = dt * 5. / 12.; }
}

extern "C" int fimhip_abstart(fimhip_ctx *c, int its, int itsDFI, int *nf, int *of, int *vof,
                              double *a1, double *a2, double *a3) {
  if (!c || !nf || !of || !vof || !a1 || !a2 || !a3) return FIMHIP_EINVAL;
  (void)itsDFI;   // digifilt is off on this path
  AB ab{*nf, *of, *vof, *a1, *a2, *a3};
  ab_rotate(ab, its, c->dt);
  *nf = ab.nf; *of = ab.of; *vof = ab.vof; *a1 = ab.ab1; *a2 = ab.ab2; *a3 = ab.ab3;
  c->ab = ab;
  return FIMHIP_OK;
}

extern "C" int fimhip_set_ab(fimhip_ctx *c, int nf, int of, int vof, double a1, double a2, double a3) {
  if (!c || nf < 1 || nf > 3 || of < 1 || of > 3 || vof < 1 || vof > 3) return FIMHIP_EINVAL;
  c->ab = AB{nf, of, vof, a1, a2, a3};
  return FIMHIP_OK;
}

extern "C" int fimhip_get_ab(const fimhip_ctx *c, int *nf, int *of, int *vof, double *a1, double *a2,
                             double *a3) {
  if (!c || !nf || !of || !vof || !a1 || !a2 || !a3) return FIMHIP_EINVAL;
  *nf = c->ab.nf; *of = c->ab.of; *vof = c->ab.vof;
  *a1 = c->ab.ab1; *a2 = c->ab.ab2; *a3 = c->ab.ab3;
  return FIMHIP_OK;
}

extern "C" int fimhip_set_class_b(fimhip_ctx *c, int ntrb, int dtratio) {
  if (!c || ntrb < 0 || ntrb > NTRA_MAX || dtratio < 1) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  c->dtratio = dtratio;
  if (ntrb == c->cb.ntrb) return FIMHIP_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  DevClassB b{};
  b.ntrb = ntrb;
  if (ntrb > 0) {
    // arrays of an earlier setting stay in the arena until the instance is destroyed
    const size_t cz = (size_t)c->nip * c->nvl, ez = cz * NPP, tz = cz * ntrb;
    struct { double **p; size_t n; } arrs[] = {
        {&b.trb, tz}, {&b.cumufx, ez}, {&b.dpinit, cz}, {&b.w, cz}, {&b.fld, tz}, {&b.vfx, tz},
        {&b.fmx, tz}, {&b.fmn, tz}, {&b.fld1, tz}, {&b.flp, tz}, {&b.fln, tz}, {&b.uan, ez * ntrb}};
    int r;
    for (auto &a : arrs)
      if ((r = c->arena.alloc(a.p, a.n))) return r;
    double *part = nullptr;
    if ((r = c->arena.alloc(&part, (size_t)c->gsum_blocks * (1 + c->ntra + ntrb)))) return r;
    c->gsum_partial = part;
    HIPCHK(hipMemsetAsync(b.trb, 0, tz * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(b.cumufx, 0, ez * sizeof(double), c->stream));
    HIPCHK(hipMemcpyAsync(b.dpinit, c->s.dp, cz * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->cb = b;
  return FIMHIP_OK;
}

extern "C" int fimhip_set_flux_scheme(fimhip_ctx *c, int flux_conserv_schm) {
  if (!c || (flux_conserv_schm != 1 && flux_conserv_schm != 2)) return FIMHIP_EINVAL;
  c->flux_conserv_schm = flux_conserv_schm;
  return FIMHIP_OK;
}

// ---------------------------------------------------------------------------
static double **field_ptr(fimhip_ctx *c, int f) {
  DevState &s = c->s;
  switch (f) {
    case FIMHIP_US: return &s.us;
    case FIMHIP_VS: return &s.vs;
    case FIMHIP_DP: return &s.dp;
    case FIMHIP_PR: return &s.pr;
    case FIMHIP_EX: return &s.ex;
    case FIMHIP_PH: return &s.ph;
    case FIMHIP_MP: return &s.mp;
    case FIMHIP_TR: return &s.tr;
    case FIMHIP_TRDP: return &s.trdp;
    case FIMHIP_WS: return &s.ws;
    case FIMHIP_RELVOR: return &s.relvor;
    case FIMHIP_SDOT: return &s.sdot;
    case FIMHIP_PSRF: return &s.psrf;
    case FIMHIP_PTDCY: return &s.ptdcy;
    case FIMHIP_U_TDCY: return &s.u_tdcy;
    case FIMHIP_V_TDCY: return &s.v_tdcy;
    case FIMHIP_DP_TDCY: return &s.dp_tdcy;
    case FIMHIP_DPL_TDCY: return &s.dpl_tdcy;
    case FIMHIP_TRC_TDCY: return &s.trc_tdcy;
    case FIMHIP_TRL_TDCY: return &s.trl_tdcy;
    case FIMHIP_MASSFX: return &s.massfx;
    case FIMHIP_U_EDG: return &s.u_edg;
    case FIMHIP_V_EDG: return &s.v_edg;
    case FIMHIP_DP_EDG: return &s.dp_edg;
    case FIMHIP_LP_EDG: return &s.lp_edg;
    case FIMHIP_MONT_EDG: return &s.mont_edg;
    case FIMHIP_UVSQ_EDG: return &s.uvsq_edg;
    case FIMHIP_TRC_EDG: return &s.trc_edg;
    case FIMHIP_TRB: return &c->cb.trb;
    case FIMHIP_CUMUFX: return &c->cb.cumufx;
    case FIMHIP_DPINIT: return &c->cb.dpinit;
  }
  return nullptr;
}

extern "C" long long fimhip_field_size(const fimhip_ctx *c, int f) {
  if (!c) return FIMHIP_EINVAL;
  const long long cz = (long long)c->nip * c->nvl, lz = (long long)c->nip * (c->nvl + 1);
  const long long ez = (long long)c->nip * NPP * c->nvl;
  switch (f) {
    case FIMHIP_US: case FIMHIP_VS: case FIMHIP_DP: case FIMHIP_MP: case FIMHIP_WS: case FIMHIP_RELVOR:
      return cz;
    case FIMHIP_PR: case FIMHIP_EX: case FIMHIP_PH: case FIMHIP_SDOT: return lz;
    case FIMHIP_TR: case FIMHIP_TRDP: return cz * c->ntra;
    case FIMHIP_PSRF: return c->nip;
    case FIMHIP_PTDCY: return 2LL * c->nip;
    case FIMHIP_U_TDCY: case FIMHIP_V_TDCY: case FIMHIP_DP_TDCY: case FIMHIP_DPL_TDCY: return 3 * cz;
    case FIMHIP_TRC_TDCY: case FIMHIP_TRL_TDCY: return 3 * cz * c->ntra;
    case FIMHIP_MASSFX: return 3 * ez;
    case FIMHIP_U_EDG: case FIMHIP_V_EDG: case FIMHIP_DP_EDG: case FIMHIP_LP_EDG: case FIMHIP_MONT_EDG:
    case FIMHIP_UVSQ_EDG: return ez;
    case FIMHIP_TRC_EDG: return ez * c->ntra;
    case FIMHIP_TRB: return cz * c->cb.ntrb;
    case FIMHIP_CUMUFX: return c->cb.ntrb ? ez : 0;
    case FIMHIP_DPINIT: return c->cb.ntrb ? cz : 0;
  }
  return FIMHIP_EINVAL;
}

static int h2d(fimhip_ctx *c, int f, const double *h) {
  if (!h) return FIMHIP_EINVAL;
  double **p = field_ptr(c, f);
  if (!*p) return FIMHIP_EINVAL;     // class-B fields before fimhip_set_class_b
  HIPCHK(hipMemcpyAsync(*p, h, fimhip_field_size(c, f) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return 0;
}
static int d2h(fimhip_ctx *c, int f, double *h) {
  if (!h) return FIMHIP_EINVAL;
  double **p = field_ptr(c, f);
  if (!*p) return FIMHIP_EINVAL;
  HIPCHK(hipMemcpyAsync(h, *p, fimhip_field_size(c, f) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return 0;
}

// One Adams-Bashforth slot (1-based, as nf/of/vof) of a tendency or mass-flux
// field, whose outermost index is the slot ((nvl,nip,3), (nvl,npp,nip,3)) or
// the slot within each tracer ((nvl,nip,3,ntra)): the drop-ins move only the
// slots a routine reads (of, vof) or writes (nf) across PCIe, and the caller's
// other slots stay as they were, as in the reference.
static int ab_slot(fimhip_ctx *c, int f, double *h, int slot, bool to_device) {
  if (!h || slot < 1 || slot > 3) return FIMHIP_EINVAL;
  double **p = field_ptr(c, f);
  if (!*p) return FIMHIP_EINVAL;
  const int ntr = (f == FIMHIP_TRC_TDCY || f == FIMHIP_TRL_TDCY) ? c->ntra : 1;
  const size_t chunk = (size_t)fimhip_field_size(c, f) / (3 * (size_t)ntr);
  for (int t = 0; t < ntr; t++) {
    const size_t off = ((size_t)t * 3 + (slot - 1)) * chunk;
    if (to_device)
      HIPCHK(hipMemcpyAsync(*p + off, h + off, chunk * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else
      HIPCHK(hipMemcpyAsync(h + off, *p + off, chunk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  return 0;
}
// the slots an AB update reads (of, vof) to the device; the one it writes (nf) back
static int ab_in(fimhip_ctx *c, int f, const double *h) {
  int r = ab_slot(c, f, const_cast<double *>(h), c->ab.of, true);
  return r ? r : ab_slot(c, f, const_cast<double *>(h), c->ab.vof, true);
}
static int ab_nf_in(fimhip_ctx *c, int f, const double *h) { return ab_slot(c, f, const_cast<double *>(h), c->ab.nf, true); }
static int ab_out(fimhip_ctx *c, int f, double *h) { return ab_slot(c, f, h, c->ab.nf, false); }

extern "C" int fimhip_upload(fimhip_ctx *c, int f, const double *h) {
  if (!c || f < 0 || f >= FIMHIP_NFIELDS) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  int r = h2d(c, f, h);
  if (r) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fimhip_download(fimhip_ctx *c, int f, double *h) {
  if (!c || f < 0 || f >= FIMHIP_NFIELDS) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  int r = d2h(c, f, h);
  if (r) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int finish(fimhip_ctx *c) {
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

#define TRY(x) do { int r_ = (x); if (r_) return r_; } while (0)

// ---------------------------------------------------------------------------
// per-routine drop-in entries (reference argument order, host arrays)
extern "C" int fimhip_edgvar1_r8(fimhip_ctx *c, int its, const double *u_vel, const double *v_vel,
                                 const double *delp, const double *pres, const double *exner, const double *geop,
                                 const double *montg, const double *tracr, double *u_edg, double *v_edg,
                                 double *dp_edg, double *lp_edg, double *geop_edg, double *mont_edg,
                                 double *uvsq_edg, double *trc_edg) {
  (void)its; (void)delp; (void)exner; (void)geop; (void)dp_edg; (void)geop_edg;
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_US, u_vel)); TRY(h2d(c, FIMHIP_VS, v_vel)); TRY(h2d(c, FIMHIP_PR, pres));
  TRY(h2d(c, FIMHIP_MP, montg)); TRY(h2d(c, FIMHIP_TR, tracr));
  // trc_edg is intent(out) for tracer 1 only; the other tracers pass through
  TRY(h2d(c, FIMHIP_TRC_EDG, trc_edg));
  launch_edgvar1(c->g, c->s, c->stream);
  TRY(d2h(c, FIMHIP_U_EDG, u_edg)); TRY(d2h(c, FIMHIP_V_EDG, v_edg)); TRY(d2h(c, FIMHIP_LP_EDG, lp_edg));
  TRY(d2h(c, FIMHIP_MONT_EDG, mont_edg)); TRY(d2h(c, FIMHIP_UVSQ_EDG, uvsq_edg));
  TRY(d2h(c, FIMHIP_TRC_EDG, trc_edg));
  return finish(c);
}

extern "C" int fimhip_edgvar3_r8(fimhip_ctx *c, int its, const double *u_vel, const double *v_vel,
                                 const double *delp, const double *pres, const double *exner, const double *geop,
                                 const double *montg, const double *tracr, double *u_edg, double *v_edg,
                                 double *dp_edg, double *lp_edg, double *geop_edg, double *mont_edg,
                                 double *uvsq_edg, double *trc_edg) {
  (void)its; (void)u_vel; (void)v_vel; (void)pres; (void)exner; (void)geop; (void)montg;
  (void)u_edg; (void)v_edg; (void)lp_edg; (void)geop_edg; (void)mont_edg; (void)uvsq_edg;
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_DP, delp)); TRY(h2d(c, FIMHIP_TR, tracr));
  TRY(h2d(c, FIMHIP_TRC_EDG, trc_edg));   // tracer 1 edges pass through
  launch_edgvar3(c->g, c->s, c->stream, true);
  TRY(d2h(c, FIMHIP_DP_EDG, dp_edg)); TRY(d2h(c, FIMHIP_TRC_EDG, trc_edg));
  return finish(c);
}

extern "C" int fimhip_cnuity_r8(fimhip_ctx *c, int its, const double *u_velo, const double *v_velo,
                                const double *u_edg, const double *v_edg, const double *dp_edg,
                                const double *lp_edg, double *delp, double *pres, double *exner,
                                double *dp_tndcy, double *dplo_tndcy, double *massfx, double *omega,
                                int timing) {
  (void)its; (void)timing;
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_US, u_velo)); TRY(h2d(c, FIMHIP_VS, v_velo)); TRY(h2d(c, FIMHIP_U_EDG, u_edg));
  TRY(h2d(c, FIMHIP_V_EDG, v_edg)); TRY(h2d(c, FIMHIP_DP_EDG, dp_edg)); TRY(h2d(c, FIMHIP_LP_EDG, lp_edg));
  TRY(h2d(c, FIMHIP_DP, delp)); TRY(h2d(c, FIMHIP_PR, pres)); TRY(h2d(c, FIMHIP_EX, exner));
  TRY(ab_in(c, FIMHIP_DP_TDCY, dp_tndcy)); TRY(ab_in(c, FIMHIP_DPL_TDCY, dplo_tndcy));
  TRY(ab_nf_in(c, FIMHIP_MASSFX, massfx));     // a pentagon's sixth edge passes through
  launch_cnuity(c->g, c->s, c->ab, c->stream, true);
  TRY(d2h(c, FIMHIP_DP, delp)); TRY(d2h(c, FIMHIP_PR, pres)); TRY(d2h(c, FIMHIP_EX, exner));
  TRY(ab_out(c, FIMHIP_DP_TDCY, dp_tndcy)); TRY(ab_out(c, FIMHIP_DPL_TDCY, dplo_tndcy));
  TRY(ab_out(c, FIMHIP_MASSFX, massfx)); TRY(d2h(c, FIMHIP_WS, omega));
  return finish(c);
}

extern "C" int fimhip_mass_adv_r8(fimhip_ctx *c, int its, const double *u_velo, const double *v_velo,
                                  const double *u_edg, const double *v_edg, double *dp_edg, const double *lp_edg,
                                  double *delp, double *pres, double *exner, double *dp_tndcy, double *massfx,
                                  double *omega, int timing) {
  (void)its; (void)timing;
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_US, u_velo)); TRY(h2d(c, FIMHIP_VS, v_velo)); TRY(h2d(c, FIMHIP_U_EDG, u_edg));
  TRY(h2d(c, FIMHIP_V_EDG, v_edg)); TRY(h2d(c, FIMHIP_DP_EDG, dp_edg)); TRY(h2d(c, FIMHIP_LP_EDG, lp_edg));
  TRY(h2d(c, FIMHIP_DP, delp)); TRY(h2d(c, FIMHIP_PR, pres)); TRY(h2d(c, FIMHIP_EX, exner));
  TRY(ab_in(c, FIMHIP_DP_TDCY, dp_tndcy)); TRY(ab_nf_in(c, FIMHIP_MASSFX, massfx));
  launch_mass_adv(c->g, c->s, c->ab, c->stream, true);
  TRY(d2h(c, FIMHIP_DP_EDG, dp_edg)); TRY(d2h(c, FIMHIP_DP, delp)); TRY(d2h(c, FIMHIP_PR, pres));
  TRY(d2h(c, FIMHIP_EX, exner)); TRY(ab_out(c, FIMHIP_DP_TDCY, dp_tndcy)); TRY(ab_out(c, FIMHIP_MASSFX, massfx));
  TRY(d2h(c, FIMHIP_WS, omega));
  return finish(c);
}

// class-B transport drop-ins; the class-B tracers are the fields ntra+1..ntra+ntrb
// of the caller's tracer array (nvl,nip,ntra+ntrb)
extern "C" int fimhip_transp0_r8(fimhip_ctx *c, int its, double *cumufx, const double *dp3d, double *dpinit) {
  (void)its;
  if (!c || c->cb.ntrb < 1) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_DP, dp3d));
  launch_transp0(c->g, c->s, c->cb, c->stream);
  TRY(d2h(c, FIMHIP_CUMUFX, cumufx)); TRY(d2h(c, FIMHIP_DPINIT, dpinit));
  return finish(c);
}

extern "C" int fimhip_transp1_r8(fimhip_ctx *c, int its, int nf, int of, int vof, double adbash1, double adbash2,
                                 double adbash3, double *cumufx, const double *massfx) {
  (void)its;
  if (!c || c->cb.ntrb < 1 || nf < 1 || nf > 3 || of < 1 || of > 3 || vof < 1 || vof > 3) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_CUMUFX, cumufx)); TRY(h2d(c, FIMHIP_MASSFX, massfx));
  launch_transp1(c->g, c->s, c->cb, AB{nf, of, vof, adbash1, adbash2, adbash3}, c->stream);
  TRY(d2h(c, FIMHIP_CUMUFX, cumufx));
  return finish(c);
}

extern "C" int fimhip_transp2_r8(fimhip_ctx *c, int its, double *tracr, const double *cumufx, const double *dpinit,
                                 const double *dpfinl, int timing) {
  (void)its; (void)timing;
  if (!c || c->cb.ntrb < 1 || !tracr) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  const size_t tz = (size_t)c->nip * c->nvl * c->cb.ntrb;
  double *trb_host = tracr + (size_t)c->nip * c->nvl * c->ntra;
  HIPCHK(hipMemcpyAsync(c->cb.trb, trb_host, tz * sizeof(double), hipMemcpyHostToDevice, c->stream));
  TRY(h2d(c, FIMHIP_CUMUFX, cumufx)); TRY(h2d(c, FIMHIP_DPINIT, dpinit)); TRY(h2d(c, FIMHIP_DP, dpfinl));
  launch_transp2(c->g, c->s, c->cb, c->stream);
  HIPCHK(hipMemcpyAsync(trb_host, c->cb.trb, tz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return finish(c);
}

extern "C" int fimhip_trcadv_r8(fimhip_ctx *c, int its, const double *trcr_edg, double *tracr, double *trcdp,
                                double *trc_tdcy, double *trclo_tdcy, const double *massfx, const double *delp,
                                int timing) {
  (void)its; (void)timing;
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_TRC_EDG, trcr_edg)); TRY(h2d(c, FIMHIP_TR, tracr)); TRY(h2d(c, FIMHIP_TRDP, trcdp));
  TRY(ab_in(c, FIMHIP_TRC_TDCY, trc_tdcy)); TRY(ab_in(c, FIMHIP_TRL_TDCY, trclo_tdcy));
  TRY(ab_nf_in(c, FIMHIP_MASSFX, massfx)); TRY(h2d(c, FIMHIP_DP, delp));
  launch_trcadv(c->g, c->s, c->ab, c->stream, true);
  TRY(d2h(c, FIMHIP_TR, tracr)); TRY(d2h(c, FIMHIP_TRDP, trcdp));
  TRY(ab_out(c, FIMHIP_TRC_TDCY, trc_tdcy)); TRY(ab_out(c, FIMHIP_TRL_TDCY, trclo_tdcy));
  return finish(c);
}

extern "C" int fimhip_momtum_r8(fimhip_ctx *c, int its, double *u_velo, double *v_velo, const double *exner,
                                double *relvor, const double *u_edg, const double *v_edg, const double *trcr_edg,
                                const double *geop_edg, const double *mont_edg, const double *uvsq_edg,
                                double *u_tndcy, double *v_tndcy, const double *dp3d) {
  (void)its; (void)geop_edg;
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_US, u_velo)); TRY(h2d(c, FIMHIP_VS, v_velo)); TRY(h2d(c, FIMHIP_EX, exner));
  TRY(h2d(c, FIMHIP_U_EDG, u_edg)); TRY(h2d(c, FIMHIP_V_EDG, v_edg)); TRY(h2d(c, FIMHIP_TRC_EDG, trcr_edg));
  TRY(h2d(c, FIMHIP_MONT_EDG, mont_edg)); TRY(h2d(c, FIMHIP_UVSQ_EDG, uvsq_edg));
  TRY(ab_in(c, FIMHIP_U_TDCY, u_tndcy)); TRY(ab_in(c, FIMHIP_V_TDCY, v_tndcy)); TRY(h2d(c, FIMHIP_DP, dp3d));
  launch_momtum(c->g, c->s, c->ab, c->stream);
  TRY(d2h(c, FIMHIP_US, u_velo)); TRY(d2h(c, FIMHIP_VS, v_velo)); TRY(d2h(c, FIMHIP_RELVOR, relvor));
  TRY(ab_out(c, FIMHIP_U_TDCY, u_tndcy)); TRY(ab_out(c, FIMHIP_V_TDCY, v_tndcy));
  return finish(c);
}

extern "C" int fimhip_hybgen_r8(fimhip_ctx *c, int its, const double *targt, double *us3d, double *vs3d,
                                double *tr3d, double *sdot, double *ex3d, double *dp3d, double *pr3d,
                                int timing) {
  (void)timing;
  if (!c || !targt) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync((void *)c->g.thetac, targt, c->nvl * sizeof(double), hipMemcpyHostToDevice, c->stream));
  TRY(h2d(c, FIMHIP_US, us3d)); TRY(h2d(c, FIMHIP_VS, vs3d)); TRY(h2d(c, FIMHIP_TR, tr3d));
  TRY(h2d(c, FIMHIP_EX, ex3d)); TRY(h2d(c, FIMHIP_DP, dp3d)); TRY(h2d(c, FIMHIP_PR, pr3d));
  launch_hybgen(c->g, c->s, its, c->stream);
  TRY(d2h(c, FIMHIP_US, us3d)); TRY(d2h(c, FIMHIP_VS, vs3d)); TRY(d2h(c, FIMHIP_TR, tr3d));
  TRY(d2h(c, FIMHIP_SDOT, sdot)); TRY(d2h(c, FIMHIP_EX, ex3d)); TRY(d2h(c, FIMHIP_DP, dp3d));
  TRY(d2h(c, FIMHIP_PR, pr3d));
  return finish(c);
}

extern "C" int fimhip_hystat_r8(fimhip_ctx *c, int its, double *ph3d, const double *ex3d, double *mp3d,
                                const double *dp3d, double *tr3d, double *trdp, double *psrf, double *ptdcy) {
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(h2d(c, FIMHIP_PH, ph3d)); TRY(h2d(c, FIMHIP_EX, ex3d)); TRY(h2d(c, FIMHIP_DP, dp3d));
  TRY(h2d(c, FIMHIP_TR, tr3d)); TRY(h2d(c, FIMHIP_PSRF, psrf)); TRY(h2d(c, FIMHIP_PTDCY, ptdcy));
  launch_hystat(c->g, c->s, its, c->stream);
  TRY(d2h(c, FIMHIP_PH, ph3d)); TRY(d2h(c, FIMHIP_MP, mp3d)); TRY(d2h(c, FIMHIP_TR, tr3d));
  TRY(d2h(c, FIMHIP_TRDP, trdp)); TRY(d2h(c, FIMHIP_PSRF, psrf)); TRY(d2h(c, FIMHIP_PTDCY, ptdcy));
  return finish(c);
}

// ---------------------------------------------------------------------------
// real*4 drop-ins: the production FIM build has default real 4, so its
// routines hand over float arrays.  Each _r4 entry widens the caller's arrays
// into FP64 staging copies, runs the _r8 entry (FP64 arithmetic throughout),
// and rounds the arrays the routine may write back to float.  Arrays the _r8
// entry leaves untouched come back bit for bit (float -> double -> float is
// exact).
namespace {
struct R4 {
  float *p;
  size_t n;
  bool out;   // written by the routine: rounded back to the caller's array
};
template <size_t N, class F>
int r4_call(const R4 (&a)[N], F &&f) {
  std::vector<std::vector<double>> st(N);
  double *d[N];
  for (size_t i = 0; i < N; i++) {
    if (!a[i].p) return FIMHIP_EINVAL;
    st[i].assign(a[i].p, a[i].p + a[i].n);
    d[i] = st[i].data();
  }
  const int r = f(d);
  if (r) return r;
  for (size_t i = 0; i < N; i++)
    if (a[i].out)
      for (size_t j = 0; j < a[i].n; j++) a[i].p[j] = (float)st[i][j];
  return FIMHIP_OK;
}
struct Sz {
  size_t c, l, e, t, te, nip, nvl;
  explicit Sz(const fimhip_ctx *x)
      : c((size_t)x->nip * x->nvl), l((size_t)x->nip * (x->nvl + 1)), e(c * NPP), t(c * x->ntra), te(e * x->ntra),
        nip((size_t)x->nip), nvl((size_t)x->nvl) {}
};
}  // namespace

extern "C" int fimhip_abstart_r4(fimhip_ctx *c, int its, int itsDFI, int *nf, int *of, int *vof, float *a1,
                                 float *a2, float *a3) {
  if (!a1 || !a2 || !a3) return FIMHIP_EINVAL;
  double d1 = *a1, d2 = *a2, d3 = *a3;
  const int r = fimhip_abstart(c, its, itsDFI, nf, of, vof, &d1, &d2, &d3);
  if (r) return r;
  *a1 = (float)d1; *a2 = (float)d2; *a3 = (float)d3;
  return FIMHIP_OK;
}

#define FIMHIP_EDGVAR_R4(NAME, R8)                                                                         \
  extern "C" int NAME(fimhip_ctx *c, int its, float *u_vel, float *v_vel, float *delp, float *pres,        \
                      float *exner, float *geop, float *montg, float *tracr, float *u_edg, float *v_edg,  \
                      float *dp_edg, float *lp_edg, float *geop_edg, float *mont_edg, float *uvsq_edg,    \
                      float *trc_edg) {                                                                    \
    if (!c) return FIMHIP_EINVAL;                                                                          \
    const Sz z(c);                                                                                         \
    const R4 a[] = {{u_vel, z.c, false},   {v_vel, z.c, false},    {delp, z.c, false},                     \
                    {pres, z.l, false},    {exner, z.l, false},    {geop, z.l, false},                     \
                    {montg, z.c, false},   {tracr, z.t, false},    {u_edg, z.e, true},                     \
                    {v_edg, z.e, true},    {dp_edg, z.e, true},    {lp_edg, z.e, true},                    \
                    {geop_edg, z.e, true}, {mont_edg, z.e, true},  {uvsq_edg, z.e, true},                  \
                    {trc_edg, z.te, true}};                                                                \
    return r4_call(a, [&](double **d) {                                                                    \
      return R8(c, its, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11], d[12],  \
                d[13], d[14], d[15]);                                                                      \
    });                                                                                                    \
  }
FIMHIP_EDGVAR_R4(fimhip_edgvar1_r4, fimhip_edgvar1_r8)
FIMHIP_EDGVAR_R4(fimhip_edgvar3_r4, fimhip_edgvar3_r8)
#undef FIMHIP_EDGVAR_R4

extern "C" int fimhip_cnuity_r4(fimhip_ctx *c, int its, float *u_velo, float *v_velo, float *u_edg, float *v_edg,
                                float *dp_edg, float *lp_edg, float *delp, float *pres, float *exner,
                                float *dp_tndcy, float *dplo_tndcy, float *massfx, float *omega, int timing) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{u_velo, z.c, false}, {v_velo, z.c, false},  {u_edg, z.e, false},          {v_edg, z.e, false},
                  {dp_edg, z.e, false}, {lp_edg, z.e, false},  {delp, z.c, true},            {pres, z.l, true},
                  {exner, z.l, true},   {dp_tndcy, 3 * z.c, true}, {dplo_tndcy, 3 * z.c, true},
                  {massfx, 3 * z.e, true}, {omega, z.c, true}};
  return r4_call(a, [&](double **d) {
    return fimhip_cnuity_r8(c, its, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11],
                            d[12], timing);
  });
}

extern "C" int fimhip_mass_adv_r4(fimhip_ctx *c, int its, float *u_velo, float *v_velo, float *u_edg,
                                  float *v_edg, float *dp_edg, float *lp_edg, float *delp, float *pres,
                                  float *exner, float *dp_tndcy, float *massfx, float *omega, int timing) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{u_velo, z.c, false}, {v_velo, z.c, false}, {u_edg, z.e, false}, {v_edg, z.e, false},
                  {dp_edg, z.e, true},  {lp_edg, z.e, false}, {delp, z.c, true},   {pres, z.l, true},
                  {exner, z.l, true},   {dp_tndcy, 3 * z.c, true}, {massfx, 3 * z.e, true}, {omega, z.c, true}};
  return r4_call(a, [&](double **d) {
    return fimhip_mass_adv_r8(c, its, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11],
                              timing);
  });
}

extern "C" int fimhip_trcadv_r4(fimhip_ctx *c, int its, float *trcr_edg, float *tracr, float *trcdp,
                                float *trc_tdcy, float *trclo_tdcy, float *massfx, float *delp, int timing) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{trcr_edg, z.te, false}, {tracr, z.t, true}, {trcdp, z.t, true}, {trc_tdcy, 3 * z.t, true},
                  {trclo_tdcy, 3 * z.t, true}, {massfx, 3 * z.e, false}, {delp, z.c, false}};
  return r4_call(a, [&](double **d) {
    return fimhip_trcadv_r8(c, its, d[0], d[1], d[2], d[3], d[4], d[5], d[6], timing);
  });
}

extern "C" int fimhip_momtum_r4(fimhip_ctx *c, int its, float *u_velo, float *v_velo, float *exner, float *relvor,
                                float *u_edg, float *v_edg, float *trcr_edg, float *geop_edg, float *mont_edg,
                                float *uvsq_edg, float *u_tndcy, float *v_tndcy, float *dp3d) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{u_velo, z.c, true},    {v_velo, z.c, true},   {exner, z.l, false},    {relvor, z.c, true},
                  {u_edg, z.e, false},    {v_edg, z.e, false},   {trcr_edg, z.te, false}, {geop_edg, z.e, false},
                  {mont_edg, z.e, false}, {uvsq_edg, z.e, false}, {u_tndcy, 3 * z.c, true},
                  {v_tndcy, 3 * z.c, true}, {dp3d, z.c, false}};
  return r4_call(a, [&](double **d) {
    return fimhip_momtum_r8(c, its, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11],
                            d[12]);
  });
}

extern "C" int fimhip_hybgen_r4(fimhip_ctx *c, int its, float *targt, float *us3d, float *vs3d, float *tr3d,
                                float *sdot, float *ex3d, float *dp3d, float *pr3d, int timing) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{targt, z.nvl, false}, {us3d, z.c, true}, {vs3d, z.c, true}, {tr3d, z.t, true},
                  {sdot, z.l, true},     {ex3d, z.l, true}, {dp3d, z.c, true}, {pr3d, z.l, true}};
  return r4_call(a, [&](double **d) {
    return fimhip_hybgen_r8(c, its, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], timing);
  });
}

extern "C" int fimhip_hystat_r4(fimhip_ctx *c, int its, float *ph3d, float *ex3d, float *mp3d, float *dp3d,
                                float *tr3d, float *trdp, float *psrf, float *ptdcy) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{ph3d, z.l, true}, {ex3d, z.l, false}, {mp3d, z.c, true}, {dp3d, z.c, false},
                  {tr3d, z.t, true}, {trdp, z.t, true},  {psrf, z.nip, true}, {ptdcy, 2 * z.nip, true}};
  return r4_call(a, [&](double **d) {
    return fimhip_hystat_r8(c, its, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]);
  });
}

extern "C" int fimhip_transp0_r4(fimhip_ctx *c, int its, float *cumufx, float *dp3d, float *dpinit) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{cumufx, z.e, true}, {dp3d, z.c, false}, {dpinit, z.c, true}};
  return r4_call(a, [&](double **d) { return fimhip_transp0_r8(c, its, d[0], d[1], d[2]); });
}

extern "C" int fimhip_transp1_r4(fimhip_ctx *c, int its, int nf, int of, int vof, float adbash1, float adbash2,
                                 float adbash3, float *cumufx, float *massfx) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{cumufx, z.e, true}, {massfx, 3 * z.e, false}};
  return r4_call(a, [&](double **d) {
    return fimhip_transp1_r8(c, its, nf, of, vof, adbash1, adbash2, adbash3, d[0], d[1]);
  });
}

extern "C" int fimhip_transp2_r4(fimhip_ctx *c, int its, float *tracr, float *cumufx, float *dpinit, float *dpfinl,
                                 int timing) {
  if (!c) return FIMHIP_EINVAL;
  const Sz z(c);
  const R4 a[] = {{tracr, z.c * (size_t)(c->ntra + c->cb.ntrb), true}, {cumufx, z.e, false}, {dpinit, z.c, false},
                  {dpfinl, z.c, false}};
  return r4_call(a, [&](double **d) { return fimhip_transp2_r8(c, its, d[0], d[1], d[2], d[3], timing); });
}

// ---------------------------------------------------------------------------
// device-resident fast path
extern "C" int fimhip_init_state(fimhip_ctx *c, int init_hybgen) {
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  launch_trdp(c->g, c->s, c->stream);
  if (init_hybgen) {
    launch_hybgen(c->g, c->s, 0, c->stream);
    launch_pres_from_exner(c->g, c->s, c->stream);
  }
  launch_hystat(c->g, c->s, 0, c->stream);
  DevState &s = c->s;
  const size_t cz = (size_t)c->nip * c->nvl, ez = (size_t)c->nip * NPP * c->nvl;
  HIPCHK(hipMemsetAsync(s.u_tdcy, 0, 3 * cz * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(s.v_tdcy, 0, 3 * cz * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(s.dp_tdcy, 0, 3 * cz * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(s.dpl_tdcy, 0, 3 * cz * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(s.trc_tdcy, 0, 3 * cz * c->ntra * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(s.trl_tdcy, 0, 3 * cz * c->ntra * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(s.massfx, 0, 3 * ez * sizeof(double), c->stream));
  c->ab = AB{0, 2, 1, 0., 0., 0.};
  if (c->cb.ntrb > 0) launch_transp0(c->g, c->s, c->cb, c->stream);   // dyn_init.F90:1096-1098
  return finish(c);
}

extern "C" int fimhip_step(fimhip_ctx *c, int its, int mode) {
  if (!c || its < 1 || (mode != 0 && mode != 1)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  struct TimerScope {
    explicit TimerScope(KTimer *t) { g_ktimer = t; }
    ~TimerScope() { g_ktimer = nullptr; }
  } scope(c->kt);
  ab_rotate(c->ab, its, c->dt);
  launch_edgvar1(c->g, c->s, c->stream);
  launch_edgvar3(c->g, c->s, c->stream, false);
  if (c->flux_conserv_schm == 2) launch_mass_adv(c->g, c->s, c->ab, c->stream, false);
  else launch_cnuity(c->g, c->s, c->ab, c->stream, false);
  if (c->cb.ntrb > 0) {             // dyn_run.F90:388-405
    launch_transp1(c->g, c->s, c->cb, c->ab, c->stream);
    if (its % c->dtratio == 0) {
      launch_transp2(c->g, c->s, c->cb, c->stream);
      launch_transp0(c->g, c->s, c->cb, c->stream);
    }
  }
  launch_trcadv(c->g, c->s, c->ab, c->stream, false);
  if (mode == 0) {
    launch_momtum(c->g, c->s, c->ab, c->stream);
    launch_hybgen(c->g, c->s, its, c->stream);
    launch_hystat(c->g, c->s, its, c->stream);
  }
  HIPCHK(hipGetLastError());
  return FIMHIP_OK;
}

extern "C" int fimhip_sync(fimhip_ctx *c) {
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  return finish(c);
}

// globsum: fixed-order block partials (cells strided by the grid), summed in
// block order, so the result is reproducible run to run.  One block's partial
// of sum q: area*dp (q = 0) or area*dp*tr_q over its cells.

__device__ __forceinline__ double gsum_block(const double *area, const double *dp, const double *tq, bool with_tr,
                                             int nip, int nvl, double *sh) {
  double acc = 0.;
  for (int ipn = blockIdx.x * 256 + threadIdx.x; ipn < nip; ipn += gridDim.x * 256) {
    double col = 0.;
    for (int k = 0; k < nvl; k++) {
      double v = area[ipn] * dp[(size_t)ipn * nvl + k];
      if (with_tr) v = v * tq[(size_t)ipn * nvl + k];
      col = col + v;
    }
    acc = acc + col;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256) k_globsum(DevGrid g, DevState s, const double *trb, int ntrb,
                                                 double *part) {
  __shared__ double sh[256];
  const int nip = g.nip, nvl = g.nvl, ntra = g.ntra;
  for (int q = 0; q <= ntra + ntrb; q++) {
    const double *tq = q > ntra ? trb + (size_t)(q - 1 - ntra) * nip * nvl : s.tr + (size_t)(q > 0 ? q - 1 : 0) * nip * nvl;
    const double r = gsum_block(g.area, s.dp, tq, q > 0, nip, nvl, sh);
    if (threadIdx.x == 0) part[(size_t)q * gridDim.x + blockIdx.x] = r;
  }
}

extern "C" int fimhip_globsum(fimhip_ctx *c, double *mass, double *trtot) {
  if (!c || !mass) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  const int nq = c->ntra + c->cb.ntrb;
  k_globsum<<<c->gsum_blocks, 256, 0, c->stream>>>(c->g, c->s, c->cb.trb, c->cb.ntrb, c->gsum_partial);
  std::vector<double> h((size_t)c->gsum_blocks * (1 + nq));
  HIPCHK(hipMemcpyAsync(h.data(), c->gsum_partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TRY(finish(c));
  for (int q = 0; q <= nq; q++) {
    double acc = 0.;
    for (int b = 0; b < c->gsum_blocks; b++) acc += h[(size_t)q * c->gsum_blocks + b];
    if (q == 0) *mass = acc / GRVITY;
    else if (trtot) trtot[q - 1] = acc;
  }
  return FIMHIP_OK;
}

// ---------------------------------------------------------------------------
// globsum over caller-owned device arrays without any allocation: the block
// partials go to a file-scope device array, `scratch` (each GPU holds its own
// copy at its own address), and one thread per sum adds them in block order,
// as fimhip_globsum does on the host, so both give the same bits.  Every call
// on a GPU writes every partial it then reads (nothing needs clearing), and
// calls that overlap on one GPU race on the array.
static __device__ __align__(256) double scratch[GSUM_BLOCKS * FIMHIP_GLOBSUM_NOALLOC_MAX_Q];

constexpr int GSUM_MAX_DEVICES = 64;
static double *g_scratch_dev[GSUM_MAX_DEVICES];   // scratch's device address, per GPU
static std::mutex g_scratch_mu;

__global__ void __launch_bounds__(256) k_globsum_dev(const double *area, const double *dp, const double *tr, int nip,
                                                     int nvl, int nq1, double *part) {
  __shared__ double sh[256];
  for (int q = 0; q < nq1; q++) {
    const double *tq = tr + (size_t)(q > 0 ? q - 1 : 0) * nip * nvl;
    const double r = gsum_block(area, dp, tq, q > 0, nip, nvl, sh);
    if (threadIdx.x == 0) part[(size_t)q * GSUM_BLOCKS + blockIdx.x] = r;
  }
}

__global__ void __launch_bounds__(64) k_globsum_final(const double *part, int nq1, double *out) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q >= nq1) return;
  double acc = 0.;
  for (int b = 0; b < GSUM_BLOCKS; b++) acc += part[(size_t)q * GSUM_BLOCKS + b];
  out[q] = q == 0 ? acc / GRVITY : acc;
}

extern "C" int fimhip_globsum_noalloc(int device, int nip, int nvl, int nq, const double *area, const double *dp,
                                      const double *tr, double *out, void *stream) {
  if (device < 0 || device >= GSUM_MAX_DEVICES || nip < 1 || nvl < 1 || nvl > NVL_MAX || nq < 0 ||
      nq + 1 > FIMHIP_GLOBSUM_NOALLOC_MAX_Q || !area || !dp || !out || (nq > 0 && !tr))
    return FIMHIP_EINVAL;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device >= ndev) return FIMHIP_EINVAL;
  int prev = 0;
  HIPCHK(hipGetDevice(&prev));
  if (prev != device) HIPCHK(hipSetDevice(device));
  hipError_t e = hipSuccess;
  double *part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {       // set-up, once per GPU, with that GPU current
      double *p = nullptr;
      e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
      if (e == hipSuccess) g_scratch_dev[device] = p;
    }
    part = g_scratch_dev[device];
  }
  if (e == hipSuccess) {
    hipStream_t st = (hipStream_t)stream;
    k_globsum_dev<<<GSUM_BLOCKS, 256, 0, st>>>(area, dp, tr, nip, nvl, nq + 1, part);
    k_globsum_final<<<(nq + 1 + 63) / 64, 64, 0, st>>>(part, nq + 1, out);
    e = hipGetLastError();
  }
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    fprintf(stderr, "fimhip: globsum_noalloc: %s\n", hipGetErrorString(e));
    return FIMHIP_EHIP;
  }
  return FIMHIP_OK;
}

// ---------------------------------------------------------------------------
// synthetic states (synth.h)
static void synth_profiles(int nip, int nvl, const double *lat, const double *lon, std::vector<double> &u0,
                           std::vector<double> &v0, std::vector<double> &bell, std::vector<double> &pint) {
  const double uu = 38.6, alpha = 0.25 * M_PI;   // Williamson et al. 1992 test 1
  const double lam_c = 1.5 * M_PI, rb = 1.0 / 3.0;
  u0.resize(nip); v0.resize(nip); bell.resize(nip); pint.resize(nvl + 1);
  for (int i = 0; i < nip; i++) {
    u0[i] = uu * (cos(lat[i]) * cos(alpha) + sin(lat[i]) * cos(lon[i]) * sin(alpha));
    v0[i] = -uu * sin(lon[i]) * sin(alpha);
    double cr = cos(lat[i]) * cos(lon[i] - lam_c);
    cr = cr > 1. ? 1. : (cr < -1. ? -1. : cr);
    double r = acos(cr);
    bell[i] = r < rb ? 0.5 * (1.0 + cos(M_PI * r / rb)) : 0.0;
  }
  for (int k = 0; k <= nvl; k++) pint[k] = 1.0e5 * fim_pow(5.0e-4, (double)k / (double)nvl);
}

static int check_synth(int nip, int nvl, int ntra) {
  return (nip < 1 || nvl < 2 || nvl > NVL_MAX || ntra < 1 || ntra > NTRA_MAX) ? FIMHIP_EINVAL : 0;
}

extern "C" int fimhip_synth_state(int nip, int nvl, int ntra, const double *lat, const double *lon,
                                  unsigned long long seed, int member, int flags, double *us, double *vs,
                                  double *dp, double *pr, double *ex, double *ph, double *tr) {
  if (check_synth(nip, nvl, ntra) || !lat || !lon || !us || !vs || !dp || !pr || !ex || !ph || !tr)
    return FIMHIP_EINVAL;
  std::vector<double> u0, v0, bell, pint;
  synth_profiles(nip, nvl, lat, lon, u0, v0, bell, pint);
  fim_synth_params p{nip, nvl, ntra, (uint64_t)seed, member, flags & 1, (flags >> 1) & 1, (flags >> 2) & 1};
  for (int i = 0; i < nip; i++)
    fim_synth_column(p, i, u0[i], v0[i], bell[i], pint.data(), us + (size_t)i * nvl, vs + (size_t)i * nvl,
                     dp + (size_t)i * nvl, pr + (size_t)i * (nvl + 1), ex + (size_t)i * (nvl + 1),
                     ph + (size_t)i * (nvl + 1), tr + (size_t)i * nvl, (size_t)nip * nvl);
  return FIMHIP_OK;
}

__global__ void __launch_bounds__(64) k_synth(fim_synth_params p, const double *u0, const double *v0,
                                              const double *bell, const double *pint, DevState s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nip) return;
  const int nvl = p.nvl;
  double pl[NVL_MAX + 1];
  for (int k = 0; k <= nvl; k++) pl[k] = pint[k];
  fim_synth_column(p, i, u0[i], v0[i], bell[i], pl, s.us + (size_t)i * nvl, s.vs + (size_t)i * nvl,
                   s.dp + (size_t)i * nvl, s.pr + (size_t)i * (nvl + 1), s.ex + (size_t)i * (nvl + 1),
                   s.ph + (size_t)i * (nvl + 1), s.tr + (size_t)i * nvl, (size_t)p.nip * nvl);
}

namespace fimhip {
// one member's seeded state written into dst (plain FP64 layout) on st
int synth_into(int nip, int nvl, int ntra, const double *lat, const double *lon, unsigned long long seed,
               int member, int flags, const DevState &dst, hipStream_t st) {
  std::vector<double> u0, v0, bell, pint;
  synth_profiles(nip, nvl, lat, lon, u0, v0, bell, pint);
  std::vector<double> packed;
  packed.reserve(3 * (size_t)nip + nvl + 1);
  packed.insert(packed.end(), u0.begin(), u0.end());
  packed.insert(packed.end(), v0.begin(), v0.end());
  packed.insert(packed.end(), bell.begin(), bell.end());
  packed.insert(packed.end(), pint.begin(), pint.end());
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, packed.size() * sizeof(double)));
  hipError_t e = hipMemcpy(d, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    fim_synth_params p{nip, nvl, ntra, (uint64_t)seed, member, flags & 1, (flags >> 1) & 1, (flags >> 2) & 1};
    k_synth<<<(nip + 63) / 64, 64, 0, st>>>(p, d, d + nip, d + 2 * (size_t)nip, d + 3 * (size_t)nip, dst);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  (void)hipFree(d);
  HIPCHK(e);
  return FIMHIP_OK;
}
}  // namespace fimhip

extern "C" int fimhip_synth_device(fimhip_ctx *c, unsigned long long seed, int member, int flags) {
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  return synth_into(c->nip, c->nvl, c->ntra, c->lat.data(), c->lon.data(), seed, member, flags, c->s, c->stream);
}

extern "C" int fimhip_profile(fimhip_ctx *c, int on) {
  if (!c) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  delete c->kt;
  c->kt = on ? new KTimer() : nullptr;
  return FIMHIP_OK;
}

extern "C" int fimhip_profile_report(fimhip_ctx *c, char *buf, int len) {
  if (!c || !buf || len < 1) return FIMHIP_EINVAL;
  buf[0] = 0;
  if (!c->kt) return FIMHIP_OK;
  HIPCHK(hipSetDevice(c->device));
  int r = c->kt->flush(c->stream);
  if (r) return r;
  std::string out;
  char line[256];
  for (auto &kv : c->kt->acc) {
    snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.second, kv.second.first);
    out += line;
  }
  if ((int)out.size() >= len) return FIMHIP_EINVAL;
  memcpy(buf, out.c_str(), out.size() + 1);
  return FIMHIP_OK;
}

extern "C" double fimhip_pow(double x, double y) { return fim_pow(x, y); }
extern "C" const char *fimhip_version(void) { return "fimhip 0.1 gfx950"; }
