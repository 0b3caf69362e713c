/* fimhip.h - C ABI of the MI355X (gfx950) FIM horizontal-dynamics library.
 *
 * A single-task serial FIM build calls these entry points in place of the
 * reference subroutines of FIMsrc/fim/horizontal (through a Fortran module of
 * bind(C) interfaces).  All arrays are host arrays in the
 * reference's Fortran layout, level index fastest:
 *   cell (nvl,nip)   level (nvl+1,nip)   edge (nvl,6,nip)   tracers (nvl,nip,ntra)
 *   tendencies (nvl,nip,3)   tracer tendencies (nvl,nip,3,ntra)   massfx (nvl,6,nip,3)
 * The _r8 entries take REAL*8 arrays (the -fdefault-real-8 build).  Per-routine
 * entries are synchronous: outputs are valid on return.  Every function
 * returns 0 on success and a negative FIMHIP_E* code on error; the library
 * never aborts the caller.  SMS halo exchanges of the reference are no-ops
 * here (one GPU holds the whole grid).
 */
#ifndef FIMHIP_H
#define FIMHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FIMHIP_OK 0
#define FIMHIP_EINVAL (-1)      /* bad argument or size */
#define FIMHIP_EHIP (-2)        /* a HIP runtime call failed */
#define FIMHIP_ENOMEM (-3)      /* device allocation failed */
#define FIMHIP_EDUP (-4)        /* duplicate edge fluxes differ (hycom_cnuity.F90:1030-1075
                                   stops the program there) */
#define FIMHIP_ENODEV (-88)     /* no usable GPU (GPUinit.cu:27-34 uses -88) */

typedef struct fimhip_ctx fimhip_ctx;

/* Grid and vertical tables, as module_constants holds them after dyn_init
 * (dyn_init.F90:373-548, 789-818).  Neighbour indices are 1-based as in
 * glvl.dat; slot 6 of a pentagon is ignored. */
typedef struct {
  int nip, nvl, ntra;           /* nvl <= 64, ntra <= 32 */
  double dt;                    /* model time step (s) */
  const int *nprox;             /* (nip) */
  const int *prox;              /* (6,nip) */
  const int *proxs;             /* (6,nip) */
  const double *cs, *sn;        /* (4,6,nip) */
  const double *sidevec_c;      /* (2,6,nip) */
  const double *sidevec_e;      /* (2,6,nip) */
  const double *sideln;         /* (6,nip) */
  const double *hfcoef;         /* (6,6,nip) */
  const double *area;           /* (nip) m^2 */
  const double *lat, *lon;      /* (nip) radians */
  const double *thetac;         /* (nvl) hybgen target theta (theta_coor.txt) */
  const double *dpsig;          /* (nvl) minimum layer thickness, Pa (dpsig.txt) */
} fimhip_grid_desc;

/* Create a device instance on GPU 'device' (replaces GPUinit + copytoGPUinit,
 * dyn_init.F90:199, 1198).  Derives corio, rarea, grndspd (dyn_init.F90:753-757),
 * dfflen and smoo_coeff (dyn_init.F90:309-321). */
int fimhip_create(const fimhip_grid_desc *grid, int device, fimhip_ctx **out);
int fimhip_destroy(fimhip_ctx *ctx);

/* abstart.F90:14 - rotate the Adams-Bashforth slots and set the weights;
 * the instance keeps its own copy, and the caller's module copy is updated. */
int fimhip_abstart(fimhip_ctx *ctx, int its, int itsDFI, int *nf, int *of, int *vof,
                   double *adbash1, double *adbash2, double *adbash3);
/* set the AB slots and weights explicitly (restart, or a caller that keeps them) */
int fimhip_set_ab(fimhip_ctx *ctx, int nf, int of, int vof, double adbash1, double adbash2, double adbash3);
/* the instance's current AB slots and weights (module_variables.F90 nf, of,
 * vof, adbash1..3), e.g. for a restart file (restart_dyn.F90:117) */
int fimhip_get_ab(const fimhip_ctx *ctx, int *nf, int *of, int *vof, double *adbash1, double *adbash2,
                  double *adbash3);

/* edgvar1.F90:27 (eqwgt, janjic=0, hiOrderFluxComp) - geop_edg and dp_edg
 * are not written (the reference leaves them untouched on this path). */
int fimhip_edgvar1_r8(fimhip_ctx *ctx, int its, const double *u_vel, const double *v_vel,
                      const double *delp, const double *pres, const double *exner, const double *geop,
                      const double *montg, const double *tracr, double *u_edg, double *v_edg,
                      double *dp_edg, double *lp_edg, double *geop_edg, double *mont_edg,
                      double *uvsq_edg, double *trc_edg);
/* edgvar3.F90:25 - dp_edg and trc_edg(:,:,:,2:ntra) */
int fimhip_edgvar3_r8(fimhip_ctx *ctx, int its, const double *u_vel, const double *v_vel,
                      const double *delp, const double *pres, const double *exner, const double *geop,
                      const double *montg, const double *tracr, double *u_edg, double *v_edg,
                      double *dp_edg, double *lp_edg, double *geop_edg, double *mont_edg,
                      double *uvsq_edg, double *trc_edg);
/* cnuity.F90:23 */
int fimhip_cnuity_r8(fimhip_ctx *ctx, int its, const double *u_velo, const double *v_velo,
                     const double *u_edg, const double *v_edg, const double *dp_edg,
                     const double *lp_edg, double *delp, double *pres, double *exner,
                     double *dp_tndcy, double *dplo_tndcy, double *massfx, double *omega,
                     int TimingBarriers);
/* massadv.F90:21 (mass_adv: the flux-limited mass transport run when
 * flux_conserv_schm = 2, dyn_run.F90:377-386).  dp_edg is limited in place
 * (flux_limiter.F90:132-260); dplo_tndcy is not an argument, as in the
 * reference. */
int fimhip_mass_adv_r8(fimhip_ctx *ctx, int its, const double *u_velo, const double *v_velo,
                       const double *u_edg, const double *v_edg, double *dp_edg, const double *lp_edg,
                       double *delp, double *pres, double *exner, double *dp_tndcy, double *massfx,
                       double *omega, int TimingBarriers);
/* Class-B tracers: ntrb tracers advected every dtratio steps by the 3-D FCT
 * scheme (ntrb and dtratio of module_control.F90:67 / fimnamelist.F90:87;
 * dyn_run.F90:388-405).  Allocates the class-B state on the device; with
 * ntrb > 0, fimhip_step and fimhip_init_state run transp0/1/2 as dyn_run and
 * dyn_init do, and fimhip_globsum returns ntra + ntrb tracer totals. */
int fimhip_set_class_b(fimhip_ctx *ctx, int ntrb, int dtratio);
/* transp3d.F90:22 (transp0) */
int fimhip_transp0_r8(fimhip_ctx *ctx, int its, double *cumufx, const double *dp3d, double *dpinit);
/* transp3d.F90:44 (transp1) */
int fimhip_transp1_r8(fimhip_ctx *ctx, int its, int nf, int of, int vof, double adbash1, double adbash2,
                      double adbash3, double *cumufx, const double *massfx);
/* transp3d.F90:81 (transp2, calling fct3d.F90:4); tracr (nvl,nip,ntra+ntrb),
 * of which the class-B fields ntra+1..ntra+ntrb are transported */
int fimhip_transp2_r8(fimhip_ctx *ctx, int its, double *tracr, const double *cumufx, const double *dpinit,
                      const double *dpfinl, int TimingBarriers);
/* trcadv.F90:19 */
int fimhip_trcadv_r8(fimhip_ctx *ctx, int its, const double *trcr_edg, double *tracr, double *trcdp,
                     double *trc_tdcy, double *trclo_tdcy, const double *massfx, const double *delp,
                     int TimingBarriers);
/* momtum.F90:16 */
int fimhip_momtum_r8(fimhip_ctx *ctx, int its, double *u_velo, double *v_velo, const double *exner,
                     double *relvor, const double *u_edg, const double *v_edg, const double *trcr_edg,
                     const double *geop_edg, const double *mont_edg, const double *uvsq_edg,
                     double *u_tndcy, double *v_tndcy, const double *dp3d);
/* hybgen.F90:20 */
int fimhip_hybgen_r8(fimhip_ctx *ctx, int its, const double *targt, double *us3d, double *vs3d,
                     double *tr3d, double *sdot, double *ex3d, double *dp3d, double *pr3d,
                     int TimingBarriers);
/* hystat.F90:23 */
int fimhip_hystat_r8(fimhip_ctx *ctx, int its, double *ph3d, const double *ex3d, double *mp3d,
                     const double *dp3d, double *tr3d, double *trdp, double *psrf, double *ptdcy);

/* ---- real*4 drop-ins ------------------------------------------------------
 * The same routines for the production build (default real 4): the argument
 * lists of the _r8 entries above with float arrays (and float AB weights).
 * The arrays are widened to FP64, the _r8 entry runs, and the arrays the
 * routine writes are rounded back to float; input arrays are not written. */
int fimhip_abstart_r4(fimhip_ctx *ctx, int its, int itsDFI, int *nf, int *of, int *vof,
                      float *adbash1, float *adbash2, float *adbash3);
int fimhip_edgvar1_r4(fimhip_ctx *ctx, int its, float *u_vel, float *v_vel, float *delp, float *pres,
                      float *exner, float *geop, float *montg, float *tracr, float *u_edg, float *v_edg,
                      float *dp_edg, float *lp_edg, float *geop_edg, float *mont_edg, float *uvsq_edg,
                      float *trc_edg);
int fimhip_edgvar3_r4(fimhip_ctx *ctx, int its, float *u_vel, float *v_vel, float *delp, float *pres,
                      float *exner, float *geop, float *montg, float *tracr, float *u_edg, float *v_edg,
                      float *dp_edg, float *lp_edg, float *geop_edg, float *mont_edg, float *uvsq_edg,
                      float *trc_edg);
int fimhip_cnuity_r4(fimhip_ctx *ctx, int its, float *u_velo, float *v_velo, float *u_edg, float *v_edg,
                     float *dp_edg, float *lp_edg, float *delp, float *pres, float *exner, float *dp_tndcy,
                     float *dplo_tndcy, float *massfx, float *omega, int TimingBarriers);
int fimhip_mass_adv_r4(fimhip_ctx *ctx, int its, float *u_velo, float *v_velo, float *u_edg, float *v_edg,
                       float *dp_edg, float *lp_edg, float *delp, float *pres, float *exner, float *dp_tndcy,
                       float *massfx, float *omega, int TimingBarriers);
int fimhip_trcadv_r4(fimhip_ctx *ctx, int its, float *trcr_edg, float *tracr, float *trcdp, float *trc_tdcy,
                     float *trclo_tdcy, float *massfx, float *delp, int TimingBarriers);
int fimhip_momtum_r4(fimhip_ctx *ctx, int its, float *u_velo, float *v_velo, float *exner, float *relvor,
                     float *u_edg, float *v_edg, float *trcr_edg, float *geop_edg, float *mont_edg,
                     float *uvsq_edg, float *u_tndcy, float *v_tndcy, float *dp3d);
int fimhip_hybgen_r4(fimhip_ctx *ctx, int its, float *targt, float *us3d, float *vs3d, float *tr3d, float *sdot,
                     float *ex3d, float *dp3d, float *pr3d, int TimingBarriers);
int fimhip_hystat_r4(fimhip_ctx *ctx, int its, float *ph3d, float *ex3d, float *mp3d, float *dp3d, float *tr3d,
                     float *trdp, float *psrf, float *ptdcy);
int fimhip_transp0_r4(fimhip_ctx *ctx, int its, float *cumufx, float *dp3d, float *dpinit);
int fimhip_transp1_r4(fimhip_ctx *ctx, int its, int nf, int of, int vof, float adbash1, float adbash2,
                      float adbash3, float *cumufx, float *massfx);
int fimhip_transp2_r4(fimhip_ctx *ctx, int its, float *tracr, float *cumufx, float *dpinit, float *dpfinl,
                      int TimingBarriers);

/* ---- device-resident fast path ------------------------------------------ */
enum {
  FIMHIP_US = 0, FIMHIP_VS, FIMHIP_DP, FIMHIP_PR, FIMHIP_EX, FIMHIP_PH, FIMHIP_MP, FIMHIP_TR,
  FIMHIP_TRDP, FIMHIP_WS, FIMHIP_RELVOR, FIMHIP_SDOT, FIMHIP_PSRF, FIMHIP_PTDCY,
  FIMHIP_U_TDCY, FIMHIP_V_TDCY, FIMHIP_DP_TDCY, FIMHIP_DPL_TDCY, FIMHIP_TRC_TDCY, FIMHIP_TRL_TDCY,
  FIMHIP_MASSFX, FIMHIP_U_EDG, FIMHIP_V_EDG, FIMHIP_DP_EDG, FIMHIP_LP_EDG, FIMHIP_MONT_EDG,
  FIMHIP_UVSQ_EDG, FIMHIP_TRC_EDG,
  FIMHIP_TRB, FIMHIP_CUMUFX, FIMHIP_DPINIT,   /* class-B transport (after fimhip_set_class_b) */
  FIMHIP_NFIELDS
};
/* The edge fields are intermediates.  After the per-routine entries they hold
   what the reference's arrays hold.  After fimhip_step, DP_EDG and TRC_EDG
   (tracers 2..ntra) hold each cell's own reconstructions (the step's readers
   form the two-sided averages), except that DP_EDG holds mass_adv's final
   edge depths under flux_conserv_schm = 2. */
/* number of doubles of a field */
long long fimhip_field_size(const fimhip_ctx *ctx, int field);
int fimhip_upload(fimhip_ctx *ctx, int field, const double *host);
int fimhip_download(fimhip_ctx *ctx, int field, double *host);
/* dyn_init start-up: optional hybgen(its=0) + pressure refresh, hystat,
 * zero tendencies and mass fluxes, AB slots nf=0, of=2, vof=1
 * (dyn_init.F90:870-1094) */
int fimhip_init_state(fimhip_ctx *ctx, int init_hybgen);
/* one dynamics step on the device, asynchronous: mode 0 = DYN_1(its) then
 * DYN_2(its) (abstart, edgvar1, edgvar3, cnuity, trcadv, momtum, hybgen,
 * hystat: dyn_run.F90:119-425); mode 1 = transport only (abstart, edgvar1,
 * edgvar3, cnuity, trcadv) */
int fimhip_step(fimhip_ctx *ctx, int its, int mode);
/* the namelist switch flux_conserv_schm (fimnamelist.F90; dyn_run.F90:377-386)
 * for fimhip_step: 1 = cnuity (FCT, the default), 2 = mass_adv */
int fimhip_set_flux_scheme(fimhip_ctx *ctx, int flux_conserv_schm);
int fimhip_sync(fimhip_ctx *ctx);
/* globsum.F90:70-101: column mass (sum area*dp/g) and the totals of the
 * ntra (+ ntrb class-B) tracers, FP64 */
int fimhip_globsum(fimhip_ctx *ctx, double *mass, double *trtot);

/* No counterpart in the reference: fimhip_globsum over caller-owned device
 * arrays, for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  It allocates
 * nothing: its block partials live in a file-scope device array of the
 * library (1024 x FIMHIP_GLOBSUM_NOALLOC_MAX_Q doubles = 512 KiB, held on each
 * GPU it runs on until the process ends).  area (nip), dp (nvl,nip) and tr
 * (nvl,nip,nq) are FP64 device arrays on GPU 'device'; out is a device array
 * of 1 + nq doubles receiving the mass (sum area*dp/g) and the nq tracer
 * totals, the same bits fimhip_globsum returns for the same arrays.  Runs
 * asynchronously on 'stream' (a hipStream_t; NULL = the default stream).
 * nq + 1 > FIMHIP_GLOBSUM_NOALLOC_MAX_Q, nvl > 64 or a null array is
 * FIMHIP_EINVAL.  Calls that overlap on one GPU race on the partials. */
#define FIMHIP_GLOBSUM_NOALLOC_MAX_Q 64
int fimhip_globsum_noalloc(int device, int nip, int nvl, int nq, const double *area, const double *dp,
                           const double *tr, double *out, void *stream);

/* per-kernel timing with HIP events on the instance's stream, for steps run
 * while profiling is on; the report lists "kernel launches total_ms" lines */
int fimhip_profile(fimhip_ctx *ctx, int on);
int fimhip_profile_report(fimhip_ctx *ctx, char *buf, int len);

/* ---- grid generation ----------------------------------------------------
 * The reference prep chain grid -> ginfo -> getlvl (FIMsrc/prep/grid/icos.F90:27,
 * GridGen.F90:1, getlvl.F90:2) for glvl = 0..10, all bisection, modified
 * recursive construction (gtype=2), IJ order (curve=0).  Outputs are the
 * records of glvl.dat in its layout (getlvl.F90:296-346); nip = 10*4^glvl+2:
 *   lat, lon, area (nip)  nprox, inv_perm (nip)  prox, proxs (6,nip), 1-based
 *   cs, sn (4,6,nip)  sidevec_c, sidevec_e (2,6,nip)  sideln, rprox_ln (6,nip)
 *   corner_xy (6,2,nip)  hfcoef (6,6,nip).  Any output pointer may be NULL. */
int fimhip_icos_grid(int glvl, float *lat, float *lon, int *nprox, int *prox, int *proxs, float *area,
                     float *cs, float *sn, float *sidevec_c, float *sidevec_e, float *sideln,
                     float *rprox_ln, float *corner_xy, float *hfcoef, int *inv_perm);
/* the ginfo intermediate file's contents (icos_grid_info_level.dat,
 * GridGen.F90:437-497, single precision): lat, lon (nip, radians),
 * prox (6,nip), nprox (nip), icos_edge (6,2,2,nip): the Voronoi corner at
 * each end of each edge as (lat, lon), inv_perm (nip) */
int fimhip_icos_grid_info(int glvl, float *lat, float *lon, int *prox, int *nprox, float *icos_edge,
                          int *inv_perm);
/* getlvl's least-squares pseudo-inverse (LSfitting.F90:73-130,
 * pseudoInverse, with num_basisfuns = 5): xy holds nobs = 5 or 6 (x, y)
 * pairs; inv (5 x nobs, basis function fastest) receives V S^-1 U^T of the
 * observation matrix [x, y, x*x, x*y, y*y], by the single-precision
 * sgesvd('A','A') the reference vendors (prep/lsfit/svd.f), bit for bit.
 * FIMHIP_EINVAL for other sizes, null pointers, a matrix sgesvd would have to
 * rescale (largest entry outside [2^-40, 2^40]) or an iteration that does
 * not converge.  Host only. */
int fimhip_ls_pseudo_inverse(const float *xy, int nobs, float *inv);

/* ---- synthetic seeded states --------------------------------------------
 * One member's state, bit-identical on the host and on the device: splitmix64
 * hashes of (seed, member, field, cell, level).  flags: bit 0 adds N(0,1) m/s
 * wind noise, bit 1 applies the ensemble-member perturbations (theta N(0,0.5 K),
 * u,v N(0,1), dp x (1+0.02 U[-1,1])), bit 2 makes tracers 2..T log-normal
 * (median 1e-9, sigma 1.5) times the cosine bell.  Arrays as fimhip_upload. */
int fimhip_synth_state(int nip, int nvl, int ntra, const double *lat, const double *lon,
                       unsigned long long seed, int member, int flags, double *us, double *vs, double *dp,
                       double *pr, double *ex, double *ph, double *tr);
/* the same, written straight into the instance's device state */
int fimhip_synth_device(fimhip_ctx *ctx, unsigned long long seed, int member, int flags);

/* ---- batched mixed-precision form ---------------------------------------
 * Several members on the same grid (an ensemble, or instances that differ
 * only in their state) advanced together by one set of kernel launches per
 * step.  Winds, tracers, tracer mass, their AB histories, the edge values and
 * the FCT work arrays are FP32; layer thickness and its AB history, pressure,
 * Exner function, Montgomery potential, geopotential and all of hybgen are
 * FP64.  Each edge's FP32 mass flux is formed by one of its two cells and
 * negated by the other, so the mass fluxes cancel exactly in the global sum.
 * No reference routine has this form; it runs the same step as fimhip_step
 * (dyn_run.F90:119-425) for every member.  Fields are exchanged per member in
 * the plain layout, widened to FP64. */
typedef struct fimhip_ens fimhip_ens;
int fimhip_ens_create(const fimhip_grid_desc *grid, int members, int device, fimhip_ens **out);
int fimhip_ens_destroy(fimhip_ens *ens);
/* seeded state (as fimhip_synth_state with synth_member and flags) into member 'member' */
int fimhip_ens_synth(fimhip_ens *ens, int member, unsigned long long seed, int synth_member, int flags);
/* host FP64 arrays of one member (layouts as fimhip_upload) */
int fimhip_ens_load_member(fimhip_ens *ens, int member, const double *us, const double *vs, const double *dp,
                           const double *pr, const double *ex, const double *ph, const double *tr);
/* dyn_init start-up for every member (dyn_init.F90:870-1094); resets the step counter */
int fimhip_ens_init_state(fimhip_ens *ens, int init_hybgen);
/* nsteps dynamics steps of every member, asynchronous; the step counter and
 * the AB slots advance on the device (abstart.F90:38-66), so the step is
 * captured once as a HIP graph and replayed (mode 0 full, 1 transport only) */
int fimhip_ens_step(fimhip_ens *ens, int nsteps, int mode);
int fimhip_ens_use_graph(fimhip_ens *ens, int on);
/* the tracer transport of the step (trcadv with edgvar3's tracer
 * reconstructions, trcadv_compute.F90:19-315, edgvar3.F90:65-163): on = one
 * kernel over patches of cells with their halo rings staged and recomputed
 * on chip; off = one kernel per stage with the intermediates in device memory
 * (the default).  Both give the same bits; fimhip_ens_autotune keeps the
 * faster, fimhip_ens_get_fused reports which is in use. */
int fimhip_ens_set_fused(fimhip_ens *ens, int on);
int fimhip_ens_get_fused(const fimhip_ens *ens);
/* land/sea mask of member 'member' (wet[nip], 0 = land), as the ocean's
 * continuity gates its fluxes (hycom_cnuity.F90:100-125): an edge carries mass
 * and tracer flux only if both of its cells are wet, the Laplacian momentum
 * diffusion through a land edge is zero, and land columns keep their
 * thickness, winds and tracers (no momentum update, no tracer update, no
 * hybgen).  Without a call every column is wet. */
int fimhip_ens_set_wet(fimhip_ens *ens, int member, const int *wet);
/* the patch partition behind the fused tracer transport, for inspection:
 * sizes[0] = patches, sizes[1] = tile entries; cnt (4 ints per patch: the
 * cumulative ends of the core and of neighbour rings 1, 2, 3 in the patch's
 * entries) and ent (4 ints per entry: the 0-based cell; the local indices in
 * the patch's entries of neighbours 0-3, 8 bits each; of neighbours 4-5 and
 * the cell's edge count, 8 bits each; the neighbours' indices of the shared
 * edges, 4 bits each; the local indices are set for rings 0-2) are written
 * when given with room (len in ints).  Host only. */
int fimhip_ens_patches(const fimhip_grid_desc *grid, int *sizes, int *cnt, long long cnt_len, int *ent,
                       long long ent_len);
/* flux_conserv_schm for the steps that follow (dyn_run.F90:377-386): 1 = FCT
 * cnuity (default), 2 = mass_adv (massadv.F90:21-160), with the edge's normal
 * wind taken from its lower-index cell so the FP32 mass fluxes still cancel */
int fimhip_ens_set_flux_scheme(fimhip_ens *ens, int flux_conserv_schm);
/* class-B tracers in the batched form (as fimhip_set_class_b): FP32 tracers,
 * FP64 thickness and vertical mass flux, moved every dtratio steps on the
 * device's step counter; with ntrb > 0, fimhip_ens_init_state restarts the
 * integration.  load_class_b takes one member's (nvl,nip,ntrb) FP64 array;
 * fimhip_ens_download returns them as FIMHIP_TRB. */
int fimhip_ens_set_class_b(fimhip_ens *ens, int ntrb, int dtratio);
int fimhip_ens_load_class_b(fimhip_ens *ens, int member, const double *trb);
int fimhip_ens_sync(fimhip_ens *ens);
/* per-member element count of FIMHIP_US, VS, DP, PR, EX, PH, MP, TR, TRDP, WS, RELVOR, SDOT, PSRF, TRB */
long long fimhip_ens_field_size(const fimhip_ens *ens, int field);
int fimhip_ens_download(fimhip_ens *ens, int member, int field, double *host);
/* globsum.F90:70-101 per member: mass[m] and trtot[m*ntra+t] from FP64 sums;
 * mass32/trtot32 (may be NULL) from FP32 Kahan-compensated sums */
int fimhip_ens_globsum(fimhip_ens *ens, double *mass, double *trtot, double *mass32, double *trtot32);
/* launch shapes: kernel k in 0..fimhip_ens_kernel_count()-1, tile 0..7 =
 * (1,2,4,8 columns per block) x (XCD-contiguous block order off/on) */
int fimhip_ens_kernel_count(void);
const char *fimhip_ens_kernel_name(int kernel);
int fimhip_ens_set_tile(fimhip_ens *ens, int kernel, int tile);
int fimhip_ens_get_tile(const fimhip_ens *ens, int kernel);
/* time every tile of every kernel over reps steps and keep each kernel's
 * fastest, for both forms of the tracer transport, and keep the faster form;
 * advances the state (reload it afterwards); times may be NULL or receive
 * kernel_count x 8 mean milliseconds */
int fimhip_ens_autotune(fimhip_ens *ens, int mode, int reps, double *times);
int fimhip_ens_profile(fimhip_ens *ens, int on);
int fimhip_ens_profile_report(fimhip_ens *ens, char *buf, int len);

/* ---------------------------------------------------------------------------
 * The HYCOM ocean's continuity equation (hycom_cnuity.F90:17, subroutine
 * cnuity): FCT layer thickness transport on ocean cells (wet > 0), the clipped
 * antidiffusive flux returned as a barotropic correction, interface smoothing
 * by bolus fluxes (thkdff != 0; Laplacian for biharm = 0, else blended with
 * the biharmonic form) and the leapfrog time smoothing, in FP64 with the
 * reference's operations in its order.  The global mass drift correction
 * (masscor, 977-998) is kept by the instance (totmass, massglb0, masscor of
 * hycom_variables) and summed over the ocean cells in cell order.
 * Arrays (Fortran layouts of hycom_cnuity.F90:21-37):
 *   dp (kdm,nip,2)   pres (kdm+1,nip)   edge (kdm,6,nip): um/vm/un/vn/dpm_edg,
 *   massflx, mssfx_avg.  A pentagon's sixth edge slot is ignored on input and
 *   massflx holds -1e33 there and on every edge of a land cell (LAND_SPVAL).
 * Returns FIMHIP_EDUP where the reference's chkdup stops the run; the arrays
 * then hold the step's results. */
typedef struct fimhip_ocn fimhip_ocn;
#define FIMHIP_OCN_DP 0
#define FIMHIP_OCN_PRES 1
#define FIMHIP_OCN_UM_EDG 2
#define FIMHIP_OCN_VM_EDG 3
#define FIMHIP_OCN_UN_EDG 4
#define FIMHIP_OCN_VN_EDG 5
#define FIMHIP_OCN_DPM_EDG 6
#define FIMHIP_OCN_MASSFLX 7
#define FIMHIP_OCN_MSSFX_AVG 8
/* the edge interpolation's inputs (hycom_edgvar.F90:128-160): uclin, vclin
 * (kdm,nip,2), utrop, vtrop (3,nip), geop (kdm+1,nip), montg, spcifv
 * (kdm,nip); and outputs beside um/vm/un/vn_edg and dp_edg (= DPM_EDG, the
 * array cnuity reads): pr_edg (kdm+1,6,nip), geop_edg, mont_edg, uvsq_edg,
 * spcv_edg (kdm,6,nip) */
#define FIMHIP_OCN_UCLIN 9
#define FIMHIP_OCN_VCLIN 10
#define FIMHIP_OCN_UTROP 11
#define FIMHIP_OCN_VTROP 12
#define FIMHIP_OCN_GEOP 13
#define FIMHIP_OCN_MONTG 14
#define FIMHIP_OCN_SPCIFV 15
#define FIMHIP_OCN_PR_EDG 16
#define FIMHIP_OCN_GEOP_EDG 17
#define FIMHIP_OCN_MONT_EDG 18
#define FIMHIP_OCN_UVSQ_EDG 19
#define FIMHIP_OCN_SPCV_EDG 20
/* the momentum equation's wind stress and bottom friction velocity (nip) */
#define FIMHIP_OCN_TAUX 21
#define FIMHIP_OCN_TAUY 22
#define FIMHIP_OCN_USTARB 23
/* the tracer transport's arrays: cumuflx (kdm,6,nip), dpinit, dpfinl
 * (kdm,nip) and passv_tr (kdm,nip,numtr), which fimhip_ocn_set_transport
 * allocates, and temp, saln (kdm,nip) */
#define FIMHIP_OCN_CUMUFLX 24
#define FIMHIP_OCN_DPINIT 25
#define FIMHIP_OCN_DPFINL 26
#define FIMHIP_OCN_TEMP 27
#define FIMHIP_OCN_SALN 28
#define FIMHIP_OCN_PASSV 29
/* the potential density (sigma_1) the grid generator reads and writes
 * (kdm,nip) */
#define FIMHIP_OCN_DENS 30
/* hycom_init's constants (hycom_constants.F90 wet, dtleap; fimnamelist kdm,
 * thkdff, biharm) on the grid of the descriptor (nprox, prox, proxs, cs, sn,
 * sidevec_e, sidevec_c, sideln, area, lat; nvl, ntra, dt, hfcoef, lon,
 * thetac and dpsig are not read) and rprox_ln (6,nip) of glvl.dat;
 * kdm <= 64, wet >= 0 */
int fimhip_ocn_create(const fimhip_grid_desc *grid, const double *rprox_ln, int kdm, const int *wet,
                      double dtleap, double thkdff, double biharm, int device, fimhip_ocn **out);
int fimhip_ocn_destroy(fimhip_ocn *ocn);
/* hycom_cnuity.F90:21 with host arrays: dp, pres, massflx and mssfx_avg are
 * updated in place */
int fimhip_ocn_cnuity_r8(fimhip_ocn *ocn, int nstep, int leapm, int leapn, const double *um_edg,
                         const double *vm_edg, const double *un_edg, const double *vn_edg, const double *dpm_edg,
                         double *dp, double *pres, double *massflx, double *mssfx_avg);
/* the same step on the device-resident arrays (fimhip_ocn_upload first) */
int fimhip_ocn_cnuity(fimhip_ocn *ocn, int nstep, int leapm, int leapn);
long long fimhip_ocn_field_size(const fimhip_ocn *ocn, int field);
/* waits for the instance's asynchronous calls (edgvar, momtum, transp0/1/2,
 * hybgn1 on the device-resident arrays); no counterpart in the reference,
 * whose routines return when done */
int fimhip_ocn_sync(fimhip_ocn *ocn);
int fimhip_ocn_upload(fimhip_ocn *ocn, int field, const double *host);
int fimhip_ocn_download(fimhip_ocn *ocn, int field, double *host);
/* hycom_edgvar.F90:128 bclin_edgvar(nstep, leapm, leapn, uclin, vclin, utrop,
 * vtrop, dp, pres, geop, montg, spcifv, um_edg, vm_edg, un_edg, vn_edg,
 * dp_edg, pr_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg) with host arrays:
 * pres and the edge arrays are updated (geop_edg only with janjic_ocn, on
 * the ocean's edges) */
int fimhip_ocn_edgvar_r8(fimhip_ocn *ocn, int nstep, int leapm, int leapn, const double *uclin,
                         const double *vclin, const double *utrop, const double *vtrop, const double *dp,
                         double *pres, const double *geop, const double *montg, const double *spcifv,
                         double *um_edg, double *vm_edg, double *un_edg, double *vn_edg, double *dp_edg,
                         double *pr_edg, double *geop_edg, double *mont_edg, double *uvsq_edg, double *spcv_edg);
/* the same on the device-resident arrays, asynchronous (fimhip_ocn_cnuity
 * and the downloads wait for it); its edge values are those
 * fimhip_ocn_cnuity reads (dp_edg is FIMHIP_OCN_DPM_EDG) */
int fimhip_ocn_edgvar(fimhip_ocn *ocn, int leapm, int leapn);
/* hycom_momtum.F90:16 momtum(nstep, leapm, leapn, utrop, vtrop, uclin, vclin,
 * um_edg, vm_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg, taux, tauy, ubforc,
 * vbforc, dp, pres, ustarb) with hycom_dissip's del4prep and dissip, on host
 * arrays: uclin, vclin (both levels), pres and ustarb are updated; ubforc and
 * vbforc are not written (modesplit is off, hycom_control.F90:5).  Needs
 * fimhip_ocn_set_momtum first. */
int fimhip_ocn_momtum_r8(fimhip_ocn *ocn, int nstep, int leapm, int leapn, const double *utrop,
                         const double *vtrop, double *uclin, double *vclin, const double *um_edg,
                         const double *vm_edg, const double *geop_edg, const double *mont_edg,
                         const double *uvsq_edg, const double *spcv_edg, const double *taux, const double *tauy,
                         double *ubforc, double *vbforc, const double *dp, double *pres, double *ustarb);
/* the same on the device-resident arrays, asynchronous */
int fimhip_ocn_momtum(fimhip_ocn *ocn, int leapm, int leapn);
/* the barotropic time step batrop (hycom_init.F90:46, dt/subcycle_ocn) and
 * the namelist's veldff and smagcf (fimnamelist.F90:253-254; 0.1 and 0.3
 * until set) */
int fimhip_ocn_set_momtum(fimhip_ocn *ocn, double batrop, double veldff, double smagcf);
/* janjic_ocn (fimnamelist.F90:246, default off): the geopotential on the
 * edges from hydrostatic integration in the neighbouring columns */
int fimhip_ocn_set_janjic(fimhip_ocn *ocn, int on);
/* the tracer transport's numtr passive tracers (0..8; hycom_control.F90:3,
 * 1 there) and temdff (fimnamelist.F90:255, the tracers' mixing velocity in
 * m/s, 0.02 there; dffusn_lyr runs where > 0): allocates the transport's
 * arrays, once per instance (a later call may change temdff only).  transp1
 * and transp2 use fimhip_ocn_set_momtum's batrop. */
int fimhip_ocn_set_transport(fimhip_ocn *ocn, int numtr, double temdff);
/* hycom_transp3d.F90:18 transp0(nstep, leapn, cumuflx, dp, dpinit) on host
 * arrays: cumuflx restarts (0 on the ocean, -1e33 on land), dpinit (ocean
 * columns) = dp(:,:,leapn) */
int fimhip_ocn_transp0_r8(fimhip_ocn *ocn, int nstep, int leapn, double *cumuflx, const double *dp,
                          double *dpinit);
/* hycom_transp3d.F90:54 transp1(nstep, cumuflx, massflx): cumuflx +=
 * massflx*2*batrop on the ocean's edges */
int fimhip_ocn_transp1_r8(fimhip_ocn *ocn, int nstep, double *cumuflx, const double *massflx);
/* hycom_transp3d.F90:88 transp2(nstep, temp, saln, passv_tr, cumuflx, dpinit,
 * dpfinl) with hycom_fct3d.F90 fct3d and hycom_dffusn.F90 dffusn_lyr:
 * temp, saln and passv_tr (NULL when numtr = 0) are advected in place */
int fimhip_ocn_transp2_r8(fimhip_ocn *ocn, int nstep, double *temp, double *saln, double *passv_tr,
                          const double *cumuflx, const double *dpinit, const double *dpfinl);
/* the same on the device-resident arrays, asynchronous: transp0 from
 * FIMHIP_OCN_DP's level leapn, transp1 with FIMHIP_OCN_MASSFLX, transp2 with
 * dpfinl = FIMHIP_OCN_DP's level leapn (hycom_run.F90:398-400) */
int fimhip_ocn_transp0(fimhip_ocn *ocn, int leapn);
int fimhip_ocn_transp1(fimhip_ocn *ocn);
int fimhip_ocn_transp2(fimhip_ocn *ocn, int leapn);
/* the hybrid grid generator's tables: hycom_constants' theta (the target
 * densities, which hycom_run passes as targt), salmin and dpmin (m), kdm
 * each (hycom_init.F90:558-566, ocn_target.txt, ocn_dpmin.txt), and the
 * latitude in degrees (nip; dyn_init.F90:754 deg_lat).  Needs
 * fimhip_ocn_set_momtum's batrop. */
int fimhip_ocn_set_hybgen(fimhip_ocn *ocn, const double *targt, const double *salmin, const double *dpmin,
                          const double *latij);
/* hycom_hybgn1.F90:9 hybgn1(nstep, leap, targt, dens, dp, pres, temp, saln,
 * uvel, vvel, latij, passv_tr) with hybgn1_1d, ppmad3, pcwise, equilb and
 * hycom_sigetc's equation of state, on host arrays: dens, temp, saln,
 * pres, and level leap of dp, uvel, vvel (kdm,nip,2) are updated; passv_tr
 * (kdm,nip,numtr of fimhip_ocn_set_transport) is remapped too when not NULL
 * (hycom_run.F90:516 passes none).  salmin and dpmin from
 * fimhip_ocn_set_hybgen. */
int fimhip_ocn_hybgn1_r8(fimhip_ocn *ocn, int nstep, int leap, const double *targt, double *dens, double *dp,
                         double *pres, double *temp, double *saln, double *uvel, double *vvel, const double *latij,
                         double *passv_tr);
/* the same on the device-resident arrays (FIMHIP_OCN_DENS, _DP, _PRES,
 * _TEMP, _SALN, _UCLIN, _VCLIN, with_tracers: _PASSV), asynchronous */
int fimhip_ocn_hybgn1(fimhip_ocn *ocn, int leap, int with_tracers);
/* hycom_constants' dtleap for the calls that follow (hycom_run.F90:136 sets
 * it every step: batrop at the first, 2*batrop after) */
int fimhip_ocn_set_dtleap(fimhip_ocn *ocn, double dtleap);
/* hycom_variables totmass, massglb0, masscor (a restart sets them) */
int fimhip_ocn_mass(const fimhip_ocn *ocn, double *totmass, double *massglb0, double *masscor);
int fimhip_ocn_set_mass(fimhip_ocn *ocn, double totmass, double massglb0, double masscor);
int fimhip_ocn_profile(fimhip_ocn *ocn, int on);
int fimhip_ocn_profile_report(fimhip_ocn *ocn, char *buf, int len);

/* x**y as computed on the device (the host build of the same routine) */
double fimhip_pow(double x, double y);
const char *fimhip_version(void);

#ifdef __cplusplus
}
#endif
#endif
