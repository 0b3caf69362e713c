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
#include <array>
#include <vector>

#include "fimhip_internal.h"

namespace fimhip {

struct HostGrid;
struct Arena;

// Patches of cells for the fused tracer FCT (fct_tiles.hip): per patch the
// cumulative ends n0..n3 of its core and three neighbour rings in its list
// of tile cells, and per tile cell an entry of four ints: the cell index;
// the local indices in the tile of its neighbours 0-3 (8 bits each); of
// neighbours 4-5 and its edge count (8 bits each); the neighbours' indices of
// the shared edges, proxs (4 bits each).  The local indices are set for
// rings 0-2 only.
struct EnsTiles {
  int npatch;
  const int *off;      // (npatch + 1) first tile cell of each patch
  const int4 *cnt;     // (npatch) n0, n1, n2, n3
  const int4 *ent;
};

// FP32 copies of the grid metrics (the reference stores them in FP32).
struct EnsGrid {
  DevGrid d;                 // FP64 metrics and the neighbour tables
  const float *cs, *sn, *svc, *sve, *sideln, *hfcoef;
  const float *area, *rarea, *corio, *grndspd, *dfflen;
  EnsTiles tiles;            // the patch partition of the fused tracer FCT
};

struct EnsState {
  int members, ncol;
  float *us, *vs, *u2, *v2;
  float *tr_alt, *trdp_alt;          // the other buffer of tr / trdp (the fused FCT writes it)
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
  int edg3_ntr;                       // tracers 2..edg3_ntr+1 reconstructed by edgvar3 (0: thickness only)
  // Land/sea mask per column (0 = land), or null (every column wet).  An edge
  // carries mass and tracer flux only if both of its cells are wet (as the
  // ocean's continuity gates its fluxes, hycom_cnuity.F90:100-125); a land
  // column keeps its thickness, winds and tracers, and the Laplacian
  // momentum diffusion through a land edge is zero.
  const unsigned char *wet;
};
// wet(column) of a mask that may be null
__device__ __forceinline__ bool col_wet(const unsigned char *wet, int col) { return !wet || ldc(wet, col) != 0; }

// Launch shapes ("tiles"): columns per 256-lane-or-smaller block and whether
// the block index is remapped so that each XCD works on a contiguous range of
// cells (neighbour columns then share that XCD's L2).
constexpr int ENS_NTILES = 8;                       // cols {1,2,4,8} x swizzle {0,1}
inline int tile_cols(int t) { return 1 << (t & 3); }
inline bool tile_swz(int t) { return (t >> 2) & 1; }

enum EnsKernel {
  EK_EDGVAR1 = 0, EK_EDGVAR3, EK_CNUITY12, EK_CNUITY3, EK_CNUITY45, EK_TRCADV1, EK_TRCADV2,
  EK_TRCADV3, EK_MOMTUM, EK_HYSTAT, EK_MASSADV1, EK_MASSADV2, EK_TRCFCT, EK_NKERNELS
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

// Sizes of the fused tracer FCT's patches (its LDS tile).
constexpr int FCT_CORE_MAX = 64;               // core cells of a patch
constexpr int FCT_R_MAX = 112;                 // cells in rings 0-1
constexpr int FCT_LO_MAX = 160;                // cells in rings 0-2
constexpr int FCT_TILE_MAX = 208;              // cells in rings 0-3
struct TileHost {
  std::vector<int> off;
  std::vector<std::array<int, 4>> cnt, ent;
};
int ens_build_tiles(const HostGrid &h, int nip, TileHost &out);
int ens_upload_tiles(Arena &a, const TileHost &t, EnsTiles &d);
// the fused tracer FCT: reads s.tr / s.trdp, writes s.tr_alt / s.trdp_alt
void ens_launch_trc_fct(int tile, const EnsGrid &g, const EnsState &s, hipStream_t st);

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
