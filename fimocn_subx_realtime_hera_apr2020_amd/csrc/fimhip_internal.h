// Internal types of the FIM dynamics library (device state and kernels).
//
// Data layout in HBM follows the reference's Fortran layout with the level
// index fastest (FIMsrc/cntl/module_constants.F90:45-94 and the dummy
// argument shapes of cnuity.F90:42-71):
//   cell field  (nvl,nip)          a[ipn*nvl + k]
//   level field (nvl+1,nip)        a[ipn*(nvl+1) + k]
//   edge field  (nvl,6,nip)        a[(ipn*6 + e)*nvl + k]
//   tracers     (nvl,nip,ntra)     a[(t*nip + ipn)*nvl + k]
//   AB history  (nvl,nip,3)        a[(s*nip + ipn)*nvl + k]
//   tracer AB   (nvl,nip,3,ntra)   a[((t*3 + s)*nip + ipn)*nvl + k]
//   massfx      (nvl,6,nip,3)      a[((s*nip + ipn)*6 + e)*nvl + k]
// Neighbour tables are 0-based on the device, with slot 6 of a pentagon
// equal to slot 5 (dyn_init.F90:546-548).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fimhip {

constexpr int NPP = 6;
constexpr int KBL = 2;
constexpr int NVL_MAX = 64;      // one wavefront lane per level
constexpr int NTRA_MAX = 32;     // tracers per instance

// module_constants.F90:14-24 and the hot-path namelist defaults
// (fimnamelist.F90:100-117)
constexpr double OMEGX2 = 1.4584e-4;
constexpr double EARTHRAD = 6371220.;
constexpr double PI_T = 3.14159265;
constexpr double GRVITY = 9.80665;
constexpr double P1000 = 100000.;
constexpr double CP = 1004.6855;
constexpr double RD = 287.0586;
constexpr double QVMIN = 1.e-10, QWMIN = 1.e-10;
constexpr double INTFC_SMOOTH = 30., PTOP = 30., THKTOP = 30.;
constexpr double RLEIGH = 2., RLTHRESH = 90., RLTAPER = 0.4, SLAK = 0.5;
constexpr double VELDFF_BKGND = 0.3, VELDFF_BOOST = 3.;

// Read-only table access through the constant address space: with a
// wave-uniform index the compiler issues a scalar load even when the kernel
// stores to other arrays in between (a load through a generic pointer after a
// store must otherwise be a vector load, since the store might alias it).
// Only for arrays no kernel writes while it runs (grid tables, step control).
template <typename T>
__device__ __forceinline__ T ldc(const T *p, size_t i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}

struct DevGrid {
  int nip, nvl, ntra;
  double dt;
  const int *prox, *proxs, *nprox;
  const double *cs, *sn, *svc, *sve, *sideln, *hfcoef;
  const double *area, *rarea, *corio, *grndspd;
  const double *dfflen, *smoo, *thetac, *dpsig;
};

struct DevState {
  double *us, *vs, *dp, *pr, *ex, *ph, *mp, *tr, *trdp;
  double *u_tdcy, *v_tdcy, *dp_tdcy, *dpl_tdcy, *trc_tdcy, *trl_tdcy, *massfx;
  double *ws, *relvor, *sdot, *psrf, *ptdcy;
  double *u_edg, *v_edg, *dp_edg, *lp_edg, *mont_edg, *uvsq_edg, *trc_edg;
  // work
  double *antifx, *delp_lo, *r_plus, *r_mnus;
  double *antiflx, *trcr_lo, *trmax, *trmin, *tr_rp, *tr_rm;   // per tracer
  double *u2, *v2, *exsmo, *thbuf;
};

// Class-B tracer transport state (transp3d.F90, fct3d.F90): the tracers
// (nvl,nip,ntrb), the time-integrated mass flux (nvl,6,nip), the thickness at
// the start of the integration, the vertical mass flux (nvl,nip), and per
// class-B tracer fct3d work: the filled field, vertical flux, bounds, the
// low-order update, the limiter ratios (nvl,nip,ntrb each) and the
// antidiffusive fluxes (nvl,6,nip,ntrb).
struct DevClassB {
  int ntrb;
  double *trb, *cumufx, *dpinit, *w;
  double *fld, *vfx, *fmx, *fmn, *fld1, *flp, *fln, *uan;
};

struct AB {
  int nf, of, vof;       // 1-based slots as in module_variables.F90
  double ab1, ab2, ab3;
};

// Step control kept in device memory for the batched form: a one-thread
// kernel advances it at the start of each step (abstart.F90:38-66), and the
// step's kernels read the slots and weights from it, so one captured graph
// serves every step.
struct StepCtl {
  int its, nf, of, vof;   // 1-based slots
  double ab1, ab2, ab3;
};
__device__ __forceinline__ StepCtl load_ctl(const StepCtl *p) {
  const __attribute__((address_space(4))) StepCtl *q = (const __attribute__((address_space(4))) StepCtl *)p;
  StepCtl c;
  c.its = q->its; c.nf = q->nf; c.of = q->of; c.vof = q->vof;
  c.ab1 = q->ab1; c.ab2 = q->ab2; c.ab3 = q->ab3;
  return c;
}

// The columns hybgen works on: every column of every member (ncol = members *
// nip, column = member * nip + cell), winds and tracers of value type V,
// thickness, pressure and Exner function in FP64.
template <typename V>
struct HybCols {
  int nip, ncol, nvl, ntra;
  const V *us_in, *vs_in;   // winds remapped by remap_1d (may be the output arrays)
  V *us, *vs, *tr;
  double *dp, *pr, *ex, *sdot, *exsmo;
  double *thbuf;            // (nvl, ncol) work column for the split form, or null
  const double *thetac, *dpsig, *smoo, *sideln, *rarea;
  const int *prox, *nprox;
  const unsigned char *wet;  // land/sea mask per column (0 = land: left as it is), or null
};
// winds_ready (may be null): an event remap_1d's launch waits for, since it is
// the first of these kernels to read the winds
template <typename V>
void launch_hybgen_cols(const HybCols<V> &h, const StepCtl *ctl, int its, hipStream_t st,
                        hipEvent_t winds_ready = nullptr);

// Per-kernel HIP-event timing: when g_ktimer is set (by the API for an
// instance with profiling on), every launch is bracketed by events on its stream.
struct KTimer;
extern thread_local KTimer *g_ktimer;
void kt_begin(const char *name, hipStream_t st);
void kt_end(hipStream_t st);
#define KT_LAUNCH(name, st, ...)                 \
  do {                                           \
    if (::fimhip::g_ktimer) ::fimhip::kt_begin(name, st); \
    __VA_ARGS__;                                 \
    if (::fimhip::g_ktimer) ::fimhip::kt_end(st); \
  } while (0)

// launch helpers (kernels_plain.hip)
void launch_edgvar1(const DevGrid &g, const DevState &s, hipStream_t st);
// two_sided: edgvar3 stores the two-sided edge averages (the drop-in entries'
// arrays); false: each cell's own reconstructions, averaged by the readers
// (the device-resident step).  Results are the same bits either way.
void launch_edgvar3(const DevGrid &g, const DevState &s, hipStream_t st, bool two_sided);
void launch_cnuity(const DevGrid &g, const DevState &s, const AB &ab, hipStream_t st, bool two_sided);
void launch_mass_adv(const DevGrid &g, const DevState &s, const AB &ab, hipStream_t st, bool two_sided);
void launch_transp0(const DevGrid &g, const DevState &s, const DevClassB &b, hipStream_t st);
void launch_transp1(const DevGrid &g, const DevState &s, const DevClassB &b, const AB &ab, hipStream_t st);
void launch_transp2(const DevGrid &g, const DevState &s, const DevClassB &b, hipStream_t st);
void launch_trcadv(const DevGrid &g, const DevState &s, const AB &ab, hipStream_t st, bool two_sided);
void launch_momtum(const DevGrid &g, DevState &s, const AB &ab, hipStream_t st);
void launch_hybgen(const DevGrid &g, const DevState &s, int its, hipStream_t st);
void launch_hystat(const DevGrid &g, const DevState &s, int its, hipStream_t st);
void launch_pres_from_exner(const DevGrid &g, const DevState &s, hipStream_t st);
void launch_trdp(const DevGrid &g, const DevState &s, hipStream_t st);
// seeded synthetic state of one member into dst (fimhip_api.hip)
int synth_into(int nip, int nvl, int ntra, const double *lat, const double *lon, unsigned long long seed,
               int member, int flags, const DevState &dst, hipStream_t st);

}  // namespace fimhip
