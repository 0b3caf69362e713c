// Definitions shared by the HYCOM ocean kernels (hycom_cnuity.hip,
// hycom_edgvar.hip, hycom_momtum.hip, hycom_transp.hip, hycom_hybgen.hip):
// the constants of hycom_constants.F90, the grid and state the kernels read,
// the lane-to-column mapping and the wavefront-local column sums, and the
// ocean instance behind the fimhip_ocn_* entries.
#pragma once
#include <vector>

#include "fimhip.h"
#include "host_common.h"

namespace fimhip {
namespace hyc {

constexpr double H_GRVITY = 9.80665;
constexpr double H_ONECM = 0.01 * (1000. * 9.80665);
constexpr double H_ASSLDP = .25;
constexpr double H_LAND = -1.e33;
constexpr double H_EPSIL = 1.e-11;
constexpr double H_ONEM = 1000. * 9.80665;        // rho * grvity (hycom_constants.F90:35)
constexpr double H_TENM = 10. * H_ONEM;
constexpr double H_SLIP = -1.;                    // no-slip walls (hycom_constants.F90:79)
constexpr double H_DIVBY18 = 1. / 18.;            // Simpson's weights (hycom_edgvar.F90:26)

// cinfo bits: 0-2 the number of edges, 3 the cell is ocean, 4+e edge e (< the
// number of edges) leads to an ocean cell
struct HycGrid {
  int nip, kdm;
  const int *prox, *proxs, *cinfo;
  const double *area, *rarea, *sve, *sideln, *rsideln, *rprox_ln;
  const double *cs, *sn;          // (4,6,nip) rotation coefficients of the edge interpolation
};
struct HycState {
  double *um, *vm, *un, *vn, *dpm_edg;
  double *dp, *pres, *massflx, *mssfx_avg;
  // work arrays
  double *antiflx, *dp_lo, *dpold, *dpc, *flxplus, *flxmnus, *rplus, *rmnus, *del2p, *dptmp;
  double *clippd, *rpbot, *aux;
  int *flag;
};

__device__ __forceinline__ double hmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double hmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ int ci_np(int ci) { return ci & 7; }
__device__ __forceinline__ bool ci_wet(int ci) { return (ci >> 3) & 1; }
__device__ __forceinline__ bool ci_open(int ci, int e) { return (ci >> (4 + e)) & 1; }

// the column of the lane (i, its slot cc in the block) and the lane's layer k;
// the host pads the grid to a multiple of 8 blocks, and block b, which the
// hardware sends to XCD b mod 8, takes the (b mod 8)-th eighth of the cells
template <int LPC>
struct Col {
  int i, k, cc;
  bool ok;
};
template <int LPC>
__device__ __forceinline__ Col<LPC> hcol(int ncol) {
  constexpr int CPB = 256 / LPC;
  const int per = gridDim.x >> 3;
  const int bx = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  Col<LPC> c;
  c.cc = threadIdx.x / LPC;
  c.i = bx * CPB + c.cc;
  c.k = threadIdx.x % LPC;
  c.ok = c.i < ncol;
  return c;
}

// orders the wavefront's LDS accesses around an exchange between its lanes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// pres(k+1) = pres(k) + d(k), k = 0..kdm-1, from pres(0) = p0: every lane forms
// the whole chain in layer order and keeps the interface below its layer
__device__ __forceinline__ double col_pres(double *buf, double d, double p0, int k, int kdm) {
  buf[k] = d;
  wave_sync();
  double prev = p0, mine = 0.;
  for (int j = 0; j < kdm; j++) {
    prev = prev + buf[j];
    if (j == k) mine = prev;
  }
  wave_sync();
  return mine;
}
// the six sums over the column's layers, in layer order from 0, of part[e]:
// lane e < NPP of the column returns edge e's
template <int LPC>
__device__ __forceinline__ double col_sums6(double (*buf)[LPC + 1], const double *part, int k, int kdm) {
#pragma unroll
  for (int e = 0; e < NPP; e++) buf[e][k] = part[e];
  wave_sync();
  double acc = 0.;
  if (k < NPP)
    for (int j = 0; j < kdm; j++) acc = acc + buf[k][j];
  wave_sync();
  return acc;
}

// the arrays of the edge interpolation (hycom_edgvar.F90 bclin_edgvar)
struct HycEdg {
  double *uclin, *vclin, *utrop, *vtrop, *geop, *montg, *spcifv;
  double *pr_edg, *geop_edg, *mont_edg, *uvsq_edg, *spcv_edg;
};

// the momentum equation's grid values, forcing and work arrays
// (hycom_momtum.F90, hycom_dissip.F90)
struct HycMom {
  const double *svc;              // sidevec_c (2,6,nip)
  const double *corio, *meshsz;   // omegx2*sin(lat) (dyn_init.F90:753), sqrt(area) (hycom_init.F90:327)
  double *taux, *tauy, *ustarb;   // (nip)
  double *drag, *dpbot;           // (nip): bottom drag; bottom of dissip's pressure column
  double *gradx, *grady, *d4u, *d4v, *du, *dv;   // (kdm,nip)
  double batrop, veldff, smagcf;
};

// the tracer transport's arrays (hycom_transp3d.F90, hycom_fct3d.F90): the
// time-integrated mass fluxes, the thicknesses at both ends of the interval,
// the tracers, and fct3d's work arrays over the nfld = 2 + numtr fields
// ((nfld,nip,kdm); uan (nfld,nip,6,kdm); colsum (nfld,4,nwp) the column
// parts of the conservation sums by ocean rank)
struct HycTr {
  double *cumuflx, *dpinit, *dpfinl, *temp, *saln, *passv;
  double *w, *fld, *flo, *vfx, *vdv, *fmx, *fmn, *flp, *fln, *uan, *colsum, *flxdv, *q;
  const int *rank;                // (nip) an ocean cell's rank among the nwet ocean cells
  int nfld, numtr, nwet, nwp;    // nwp: the padded row length of colsum
  double temdff;
};

// the hybrid grid generator's tables (hycom_constants' theta, salmin, dpmin;
// the latitude in degrees) and the ocean columns it runs on; dens (kdm,nip)
struct HycHybState {
  int *cols;
  int ncol;
  double *targt, *salmin, *dpmin, *latij, *dens;
  double *pold;                   // (kdm+1,nip) work: the interfaces before the moves
};

}  // namespace hyc
}  // namespace fimhip

#define HC(a, i, k) (a)[(size_t)(i) * kdm + (k)]
#define HP(i, k) s.pres[(size_t)(i) * (kdm + 1) + (k)]
#define HE(a, i, e, k) (a)[((size_t)(i) * NPP + (e)) * kdm + (k)]
#define HSVE(i, e, c) g.sve[((size_t)(i) * NPP + (e)) * 2 + (c)]

struct fimhip_ocn {
  int device = 0, nip = 0, kdm = 0;
  double dtleap = 0., thkdff = 0., biharm = 0.;
  double totmass = 0., massglb0 = 0., masscor = 0.;
  hipStream_t stream = nullptr;
  fimhip::Arena arena;
  fimhip::hyc::HycGrid g{};
  fimhip::hyc::HycState s{};
  fimhip::hyc::HycEdg ed{};
  fimhip::hyc::HycMom mo{};
  fimhip::hyc::HycTr tr{};
  fimhip::hyc::HycHybState hy{};
  bool janjic = false;               // janjic_ocn (fimnamelist.F90:246)
  std::vector<int> wet;
  std::vector<double> area;
  fimhip::KTimer *kt = nullptr;              // non-null while profiling
};

// the device array of a FIMHIP_OCN_* field and its element count
double *ocn_field(fimhip_ocn *o, int f, size_t *n);
// blocks of a launch over all cells with LPC lanes per column (a multiple of 8)
inline unsigned ocn_blocks(const fimhip_ocn *o, int lpc) {
  const int cpb = 256 / lpc;
  const int nblk = (o->nip + cpb - 1) / cpb;
  return (unsigned)((nblk + 7) / 8 * 8);
}
