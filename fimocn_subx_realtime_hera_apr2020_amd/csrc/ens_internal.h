// The batched mixed-precision form: several members (ensemble members or
// other instances on the same grid) advanced by the same kernels.
//
// Precision split (per field):
//   FP32  winds, tracers, tracer mass, their Adams-Bashforth histories, the
//         edge values (of the edge wind only its normal component is stored;
//         momtum forms the full edge winds itself), the FCT work arrays and
//         the mass fluxes;
//   FP64  layer thickness and its AB history (mass conservation), interface
//         pressure and Exner function, Montgomery potential (also on the
//         edges, formed where momtum needs it) and geopotential (the
//         pressure-gradient force), and all of hybgen.
// Each edge's mass flux is taken from one side (the lower cell index) and
// negated on the other, so the FP32 fluxes cancel exactly in the global sum.
//
// Layout: the plain form's layout with the cell index replaced by the column
// index col = member * nip + cell (ncol = members * nip):
//   cell (nvl, ncol)   level (nvl+1, ncol)   edge (nvl, 6, ncol)
//   tracer (nvl, ncol, ntra)   AB (nvl, ncol, 3)   tracer AB (nvl, ncol, 3, ntra)
//   tracer edge (nvl, 6, ncol, ntra - 1) for tracers 2..ntra (theta's edge
//   values are formed where they are used); the mass flux keeps only the
//   current slot.
#pragma once
#include "fimhip_internal.h"

namespace fimhip {

// FP32 copies of the grid metrics (the reference stores them in FP32).
struct EnsGrid {
  DevGrid d;                 // FP64 metrics and the neighbour tables
  const float *cs, *sn, *svc, *sve, *sideln, *hfcoef;
  const float *area, *rarea, *corio, *grndspd, *dfflen;
};

struct EnsState {
  int members, ncol;
  float *us, *vs, *u2, *v2;
  double *dp, *pr, *ex, *ph, *mp, *sdot, *exsmo, *psrf, *ptdcy, *thbuf;
  float *tr, *trdp;
  float *u_tdcy, *v_tdcy, *dpl_tdcy;
  double *dp_tdcy;
  float *trc_tdcy, *trl_tdcy;
  float *massfx, *ws, *relvor;
  float *vn_edg, *dp_edg, *trc_edg;   // vn_edg: normal edge wind; trc_edg: tracers 2..ntra
  float *pgx, *pgy;                   // the edge-pressure line integrals of cnuity5 (per cell-level)
  float *th0;                                  // theta at the start of the step
  float *antifx, *delp_lo, *r_plus, *r_mnus;
  float *antiflx, *trcr_lo, *trmax, *trmin, *tr_rp, *tr_rm;
  StepCtl *ctl;
};

// Launch shapes ("tiles"): columns per 256-lane-or-smaller block and whether
// the block index is remapped so that each XCD works on a contiguous range of
// cells (neighbour columns then share that XCD's L2).
constexpr int ENS_NTILES = 8;                       // cols {1,2,4,8} x swizzle {0,1}
inline int tile_cols(int t) { return 1 << (t & 3); }
inline bool tile_swz(int t) { return (t >> 2) & 1; }

enum EnsKernel {
  EK_EDGVAR1 = 0, EK_EDGVAR3, EK_CNUITY12, EK_CNUITY3, EK_CNUITY45, EK_TRCADV1, EK_TRCADV2,
  EK_TRCADV3, EK_MOMTUM, EK_HYSTAT, EK_MASSADV1, EK_MASSADV2, EK_NKERNELS
};
extern const char *const ens_kernel_names[EK_NKERNELS];

// Class-B transport of the batched form: the tracers (nvl, ncol, ntrb) FP32,
// the ring of the last three steps' mass fluxes (nvl, 6, ncol, 3) and their
// AB-weighted time integral (nvl, 6, ncol) FP32, the thickness at the start of
// the integration and the vertical mass flux (nvl, ncol) FP64, and per class-B
// tracer the fct3d work fields (FP32; uan on the edges).
struct EnsClassB {
  int ntrb, dtratio;
  float *trb, *ring, *cumufx;
  double *dpinit, *w;
  float *fld, *vfx, *fmx, *fmn, *fld1, *flp, *fln, *uan;
};
void ens_launch_class_b_init(const EnsGrid &g, const EnsState &s, const EnsClassB &b, hipStream_t st);
void ens_launch_class_b_step(const EnsGrid &g, const EnsState &s, const EnsClassB &b, hipStream_t st);

void ens_launch_abstart(const EnsState &s, double dt, hipStream_t st);
void ens_launch(int kernel, int tile, const EnsGrid &g, const EnsState &s, hipStream_t st);
void ens_launch_pres_from_exner(const EnsGrid &g, const EnsState &s, hipStream_t st);
void ens_launch_trdp(const EnsGrid &g, const EnsState &s, hipStream_t st);
void ens_launch_hystat_init(const EnsGrid &g, const EnsState &s, hipStream_t st);
void ens_launch_convert_member(const EnsGrid &g, const EnsState &s, int member, const DevState &src,
                               hipStream_t st);
void ens_launch_globsum(const EnsGrid &g, const EnsState &s, double *part64, float *part32, int nblocks,
                        hipStream_t st);

}  // namespace fimhip
