"""Benchmark of the FIM horizontal dynamics step on one MI355X.

Headline workload: a 16-member ensemble on the G7 grid (163842 cells) x 64
layers with 4 tracers, dt = 90 s (the reference's step for this grid), full
dynamics steps (abstart, edgvar1,
edgvar3, cnuity, trcadv, momtum, hybgen, hystat), in the batched
mixed-precision form (FP32 winds, tracers and edge values; FP64 thickness,
pressure, Exner function, Montgomery potential and hybgen), the step replayed
from a captured HIP graph with autotuned launch shapes.  Members are seeded
synthetic states (wind noise and member perturbations) generated on the
device.  value = members * nip * nvl * steps / wall seconds
(cell-level-steps/s), state resident in HBM before the timed region.

Beside it, in the same run: the FP64 plain path (one member of the same
grid, bit-identical to the checker), the accuracy of the FP32 form against
FP64 after 96 steps and the global mass drift, and the C restatement timed on
the host's cores.

Prints ONE JSON line.  Besides the contract fields it carries
  roofline:     the headline form's whole step against the compulsory-byte
                model (18 + 8T values per cell-level-step: 220 B in the
                mixed form, 400 B in FP64) at 8 TB/s; traffic = the step's
                PMC-measured HBM bytes (committed profile) where available
  dominant_kernel: the kernel with the most device time (HIP events on the
                library's stream): its own input/output bytes per launch over
                its mean launch duration
  plain_fp64:   the FP64 path's throughput, roofline and dominant kernel; under
                mass_adv the same step with the flux-limited mass transport
                (flux_conserv_schm = 2) and its two kernels; under class_b the
                step with 8 class-B tracers moved every 6 steps by fct3d
                (ensemble.mass_adv, ensemble.class_b: the same for the
                batched form, 4 class-B tracers per member)
  uneven_batch: G6 and G7 concatenated into one instance (two resolutions in
                one launch), mixed form and FP64
  hycom:        the HYCOM ocean (FP64, bitwise the checker's) on G8 x 32
                layers: its continuity equation alone and the edge
                interpolation followed by the continuity equation, each with
                throughput, per-kernel times, its dominant kernel against its
                algorithmic bytes, and the checker timed on one host core
  accuracy_96:  relative L2 of member 0 against FP64 after 96 steps, mass drift
  cpu_baseline: the C restatement (oracle/fim_oracle.c, FP64, OpenMP) on a
                bounded sample of the same grid
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
SEED = 20240601


def plain_kernel_bytes(T):
    """Compulsory HBM bytes per cell-level of each FP64 kernel launch: the
    distinct array elements it must read plus those it writes (8 B each)."""
    v = {
        "k_edgvar1": 5 + 36,                 # us vs pr mp tr1 -> u v lp mont uvsq trc1 on 6 edges
        "k_edgvar3": 7 * T,                  # dp, tracers 2..T -> 6 edge values each
        "k_cnuity12": 18 + 3 + 6 + 6 + 2,    # u/v/dp edges, dp, dplo(of,vof) -> massfx, antifx, dplo, delp_lo
        "k_cnuity3": 6 + 4 + 2,              # antifx, delp_lo, dp, dp_tdcy(of,vof) -> r+, r-
        "k_cnuity45": 27 + 11,               # antifx r+- massfx dplo dptdcy dp pr lp_edg u v -> massfx dptdcy dp pr ex omega
        "k_massadv1": 12 + 6 + 1 + 2 + 6 + 2,   # u/v/dp edges, dp, dp_tdcy(of,vof) -> limited dp edges, q_min/max_out
        "k_massadv2": 12 + 2 + 6 + 2 + 1 + 6 + 2 + 6 + 6 + 1 + 1 + 1 + 1 + 1,
        # u/v/dp edges, q_min/max_out, dp_tdcy(of,vof), dp, lp_edg, u, v -> dp edges, massfx, dp_tdcy, dp, pr, ex, omega
        "k_trcadv1": 27 * T,
        "k_trcadv2": 16 * T,
        "k_trcadv3": 18 * T,
        "k_momtum": 3 + 1 + 30 + 4 + 5,
        "k_regrid": 3,                       # theta, Exner -> Exner (sdot)
        "k_regrid_tail": 2,                  # Exner in sdot, in place
        "k_dffusn_lev": 2,
        "k_remap_inflate": 2,                # smoothed Exner, in place
        "k_remap": 12 + 2 * T,               # ex pr exsmo u v tracers -> dp u v tracers 2..T theta-work depths pr sdot ex
        "k_remap_theta": 3,                  # theta work, layer depths -> theta
        "k_hystat": 5,
        "k_hystat_tr": 2 * T + 3,
    }
    return {k: 8 * n for k, n in v.items()}


def class_b_kernel_bytes(B):
    """Compulsory bytes per cell-level of the class-B transport launches (FP64)
    for B class-B tracers; arrays shared by the tracers of a launch counted once."""
    v = {
        "k_transp1": 6 + 18 + 6,             # cumufx, massfx (3 slots) -> cumufx
        "k_transp2_w": 6 + 2 + 1,            # cumufx, dpinit, dp -> w
        "k_transp0": 1 + 1 + 6,              # dp -> dpinit, cumufx
        "k_fct3d_vert": 2 + 3 * B,           # dpinit, w; per tracer trb -> filled field, vertical flux
        "k_fct3d_lo": 9 + 13 * B,            # cumufx w dpinit dp; per tracer field, vfx -> uan(6) bounds update ratios
        "k_fct3d_clip": 1 + 12 * B,          # dp; per tracer ratios, uan(6), update, bounds -> trb
    }
    return {k: 8 * n for k, n in v.items()}


def ens_kernel_bytes(T, fused=False):
    """The same for the mixed-precision kernels (4 B per FP32 value, 8 B per
    FP64 value; inputs shared by the tracers of one launch counted once).
    With the fused tracer transport edgvar3 reconstructs the thickness only."""
    return {
        "e_edgvar1": 20 + 36,                # us vs tr1 (f32) pr (f64) -> normal wind edges, pressure gradient, step-start theta (f32)
        # dp (f64) -> dp_edg; tracers 2..T -> trc_edg (the fused transport forms those itself)
        "e_edgvar3": 32 if fused else 32 + 28 * (T - 1),
        "e_cnuity12": 104,                   # normal wind and dp edges, dp, dplo(of,vof) -> massfx, antifx, dplo, delp_lo, sums
        "e_cnuity3": 44,                     # in/outflow sums, delp_lo, dp, dp_tdcy(of,vof) -> r+, r-
        "e_cnuity45": 164,                   # antifx r+- massfx dp_tdcy dp pr pressure gradient u v -> massfx dp_tdcy dp pr ex omega
        "e_massadv1": 104,                   # normal wind and dp edges, dp, dp_tdcy(of,vof) -> q_star edges, q_min/max_out
        "e_massadv2": 156,                   # normal wind, q_star, bounds, dp_tdcy, dp, pressure gradient, u v -> massfx dp_tdcy dp pr ex omega
        "e_trcadv1": 8 + 80 * T,             # massfx dp; per tracer trc_edg (not theta's) tr trdp trl -> antiflx trl lo sums
        "e_trcadv2": 8 + 40 * T,
        "e_trcadv3": 8 + 68 * T,
        "e_trc_fct": 32 + 40 * T,            # massfx dp; per tracer tr trdp trl(of,vof) trc(of,vof) -> tr trdp trl trc
        "e_momtum": 80,                      # u v dp ex mp, step-start theta, tendencies -> relvor tendency u2 v2
        "k_regrid": 20,
        "k_regrid_tail": 16,
        "k_dffusn_lev": 16,
        "k_remap_inflate": 16,
        "k_remap": 84 + 8 * T,
        "k_remap_theta": 20,
        "e_hystat": 40 + 8 * T,
    }


def ens_step_bytes(T, mode):
    """Compulsory bytes per cell-level-step of the mixed form: the
    18 + 8T values (10 + 8T transport-only) at 4 B, except the arrays kept in
    FP64: the thickness (read and written) and its AB history (read twice,
    written once), 5 values, +4 B each."""
    v = (18 + 8 * T) if mode == 0 else (10 + 8 * T)
    return 4 * v + 4 * 5


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--glvl", type=int, default=7)
    ap.add_argument("--nvl", type=int, default=64)
    ap.add_argument("--ntra", type=int, default=4)
    ap.add_argument("--members", type=int, default=16)
    ap.add_argument("--dt", type=float, default=None,
                    help="time step (s); default the reference's 1.2*4800/(2**glvl/2) (90 s at G7)")
    ap.add_argument("--mode", type=int, default=0, help="0 full step, 1 transport only")
    ap.add_argument("--form", choices=["both", "ens", "plain", "hycom"], default="both",
                    help="hycom: only the ocean continuity leg")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=96, help="0 skips the accuracy leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the mass_adv and class-B legs")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (for PMC byte calibration)")
    return ap.parse_args()


PMC_TABLE = os.path.join(ROOT, "profiles", "pmc_traffic_G7x64.md")


def pmc_traffic(kernel, form, a):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (copy-calibrated) for this kernel on the G7 x 64 x 4-tracer workload, from
    the committed summary; None for any other workload."""
    if (a.glvl, a.nvl, a.ntra) != (7, 64, 4) or (form == "ens16" and a.members != 16):
        return None
    if not os.path.exists(PMC_TABLE):
        return None
    for line in open(PMC_TABLE):
        f = [x.strip() for x in line.strip().strip("|").split("|")]
        if len(f) == 7 and f[0] == "`%s`" % kernel and f[1] == form:
            return float(f[4])
    return None


def dominant_kernel(prof, kbytes, cells, traffic_of=None):
    """The kernel with the most device time: its own input/output bytes per
    launch (bench's per-kernel model, intermediates included) over its mean
    launch time; traffic is its measured HBM bytes per launch."""
    dname, (dn, dms) = max(prof.items(), key=lambda kv: kv[1][1])
    dmean = dms / dn * 1e-3
    dbytes = kbytes.get(dname, 0) * cells
    return {"kernel": dname, "io_GBs": dbytes / dmean / 1e9, "io_frac": dbytes / dmean / HBM_PEAK,
            "traffic": traffic_of(dname) if traffic_of else None,
            "bytes_per_launch": dbytes, "mean_launch_ms": dmean * 1e3}


def step_roofline(value, b_cell, cells, ms_per_step, prof, steps_profiled, traffic_of=None):
    """The whole step against the compulsory-byte model: achieved =
    cell-level-steps/s x compulsory bytes per cell-level-step (18 + 8T values:
    the state read and written once, the Adams-Bashforth histories read twice
    and written once), frac = achieved / 8 TB/s.  traffic = the measured HBM
    bytes of one step, the sum over the step's kernels of their PMC-measured
    bytes per launch times launches per step (None unless every kernel of the
    step has a committed measurement for this workload)."""
    traffic = None
    if traffic_of is not None and prof:
        tot = 0.
        for k, (n, _) in prof.items():
            t = traffic_of(k)
            if t is None:
                tot = None
                break
            tot += t * n / steps_profiled
        traffic = tot
    achieved = value * b_cell
    return {"scope": "step", "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "compulsory_bytes_per_cell_level": b_cell,
            "compulsory_bytes_per_step": b_cell * cells, "ms_per_step": ms_per_step}


def kernel_table(prof, kbytes, cells):
    tot = sum(v[1] for v in prof.values())
    out = {}
    for k, (n, ms) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        mean = ms / n
        out[k] = {"launches": n, "mean_ms": mean, "share": ms / tot}
        if k in kbytes:
            out[k]["GBs"] = kbytes[k] * cells / (mean * 1e-3) / 1e9
    return out


def run_plain(a, grid, th, ds, fimhip):
    import torch
    m = fimhip.Model(grid, a.nvl, a.ntra, a.dt, th, ds, device=0)
    # bring the GPU to its steady clocks (at least 2 s of steps, the member
    # re-seeded every 80 steps) before timing, then time at least 80 steps
    # (about a second) from a fresh state, so the reference leg is not
    # measured on a cold GPU; every step stays within the 96 steps the
    # synthetic state is checked for
    tw = time.perf_counter()
    while time.perf_counter() - tw < 2.0:
        m.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
        m.init_state(True)
        for its in range(1, 81):
            m.step(its, a.mode)
        m.sync()
    m.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    m.init_state(True)
    its = 0
    for _ in range(a.warmup):
        its += 1
        m.step(its, a.mode)
    m.sync()
    torch.cuda.synchronize()
    n_timed = max(a.steps, 80)
    t0 = time.perf_counter()
    for _ in range(n_timed):
        its += 1
        m.step(its, a.mode)
    m.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-kernel event timing over as many further steps, from a fresh state
    m.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    m.init_state(True)
    its = 0
    m.profile(True)
    for _ in range(a.steps):
        its += 1
        m.step(its, a.mode)
    m.sync()
    prof = m.profile_report()
    m.profile(False)
    cells = grid.nip * a.nvl
    kb = plain_kernel_bytes(a.ntra)
    v_step = (18 + 8 * a.ntra) if a.mode == 0 else (10 + 8 * a.ntra)
    value = cells * n_timed / wall
    tr_of = (lambda k: pmc_traffic(k, "plain", a)) if a.mode == 0 else None
    res = {"value": value, "ms_per_step": wall / n_timed * 1e3, "members": 1, "timed_steps": n_timed,
           "roofline": step_roofline(value, 8 * v_step, cells, wall / n_timed * 1e3, prof, a.steps, tr_of),
           "dominant_kernel": dominant_kernel(prof, kb, cells, tr_of),
           "kernels": kernel_table(prof, kb, cells)}
    if not a.headline_only:
        res["mass_adv"] = run_plain_mass_adv(a, m, cells, kb, fimhip)
        res["class_b"] = run_plain_class_b(a, m, grid, cells, fimhip)
    return m, res


def run_plain_class_b(a, m, grid, cells, fimhip, ntrb=8, dtratio=6):
    """The same FP64 step with ntrb class-B tracers (log-normal fields, the
    chemistry-load case) moved every dtratio steps by transp2/fct3d: step time
    over 2 * dtratio steps and the class-B kernels."""
    import torch
    m.set_class_b(ntrb, dtratio)
    m.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    lg = fimhip.synth_state(grid.lat, grid.lon, a.nvl, 1 + ntrb, seed=7, member=3, flags=fimhip.SYNTH_LOGNORMAL)
    m.upload("trb", lg["tr"][1:])
    del lg
    m.init_state(True)
    its = 0
    for _ in range(dtratio):
        its += 1
        m.step(its, a.mode)
    m.sync()
    torch.cuda.synchronize()
    n_timed = 2 * dtratio
    t0 = time.perf_counter()
    for _ in range(n_timed):
        its += 1
        m.step(its, a.mode)
    m.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    m.profile(True)
    for _ in range(n_timed):
        its += 1
        m.step(its, a.mode)
    m.sync()
    prof = m.profile_report()
    m.profile(False)
    m.set_class_b(0, dtratio)
    kt = kernel_table(prof, class_b_kernel_bytes(ntrb), cells)
    kt = {k: v for k, v in kt.items() if k.startswith(("k_transp", "k_fct3d"))}
    per_call = sum(v["mean_ms"] for k, v in kt.items() if k in ("k_transp2_w", "k_fct3d_fill", "k_fct3d_vert",
                                                                  "k_fct3d_lo", "k_fct3d_clip"))
    return {"ntrb": ntrb, "dtratio": dtratio, "ms_per_step": wall / n_timed * 1e3, "timed_steps": n_timed,
            "transp2_ms_per_call": per_call, "kernels": kt}


def run_plain_mass_adv(a, m, cells, kb, fimhip):
    """The same FP64 step with the flux-limited mass transport in place of
    FCT cnuity (flux_conserv_schm = 2): step time and its two kernels."""
    import torch
    m.set_flux_scheme(2)
    m.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    m.init_state(True)
    its = 0
    for _ in range(a.warmup):
        its += 1
        m.step(its, a.mode)
    m.sync()
    torch.cuda.synchronize()
    n_timed = 40
    t0 = time.perf_counter()
    for _ in range(n_timed):
        its += 1
        m.step(its, a.mode)
    m.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    m.profile(True)
    for _ in range(a.steps):
        its += 1
        m.step(its, a.mode)
    m.sync()
    prof = m.profile_report()
    m.profile(False)
    m.set_flux_scheme(1)
    kt = kernel_table(prof, kb, cells)
    return {"value": cells * n_timed / wall, "ms_per_step": wall / n_timed * 1e3, "timed_steps": n_timed,
            "kernels": {k: kt[k] for k in ("k_massadv1", "k_massadv2") if k in kt}}


def synth_ens(e, members, fimhip):
    for mm in range(members):
        e.synth(mm, seed=SEED, synth_member=mm, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    e.init_state(True)


def run_ens(a, grid, th, ds, fimhip):
    import torch
    M = a.members
    e = fimhip.Ensemble(grid, M, a.nvl, a.ntra, a.dt, th, ds, device=0)
    tune = None
    if not a.no_autotune:
        synth_ens(e, M, fimhip)
        tt = e.autotune(a.mode, 2)
        names = e.kernel_names()
        tiles = e.tiles()
        tune = {n: {"tile": int(tiles[n]), "ms": [round(float(x), 4) for x in tt[k]]}
                for k, n in enumerate(names) if all(x == x for x in tt[k])}   # kernels the step ran
        tune["tracer_transport"] = "fused" if e.fused() else "chain"
    synth_ens(e, M, fimhip)
    e.step(a.warmup, a.mode)
    e.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.step(a.steps, a.mode)
    e.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e.profile(True)
    e.step(a.steps, a.mode)
    e.sync()
    prof = e.profile_report()
    e.profile(False)
    cells = M * grid.nip * a.nvl
    kb = ens_kernel_bytes(a.ntra, e.fused())
    value = cells * a.steps / wall
    b_step = ens_step_bytes(a.ntra, a.mode)
    tr_of = (lambda k: pmc_traffic(k, "ens16", a)) if a.mode == 0 else None
    res = {"value": value, "ms_per_step": wall / a.steps * 1e3, "members": M,
           "roofline": step_roofline(value, b_step, cells, wall / a.steps * 1e3, prof, a.steps, tr_of),
           "dominant_kernel": dominant_kernel(prof, kb, cells, tr_of),
           "kernels": kernel_table(prof, kb, cells), "tiles": tune}
    if not a.headline_only:
        res["mass_adv"] = run_ens_mass_adv(a, e, cells, kb, fimhip)
        res["class_b"] = run_ens_class_b(a, e, grid, cells, fimhip)
    return e, res


def run_ens_class_b(a, e, grid, cells, fimhip, ntrb=4, dtratio=6):
    """The batched step with ntrb class-B tracers per member (log-normal
    fields) moved every dtratio steps: step time over 2 * dtratio steps and
    the class-B kernels (the transport kernels of the other steps return at
    once, so their mean includes those launches)."""
    import torch
    e.set_class_b(ntrb, dtratio)
    for mm in range(a.members):
        e.synth(mm, seed=SEED, synth_member=mm, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    lg = fimhip.synth_state(grid.lat, grid.lon, a.nvl, 1 + ntrb, seed=7, member=3, flags=fimhip.SYNTH_LOGNORMAL)
    for mm in range(a.members):
        e.load_class_b(mm, lg["tr"][1:])
    del lg
    e.init_state(True)
    e.step(dtratio, a.mode)
    e.sync()
    torch.cuda.synchronize()
    n_timed = 2 * dtratio
    t0 = time.perf_counter()
    e.step(n_timed, a.mode)
    e.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e.profile(True)
    e.step(n_timed, a.mode)
    e.sync()
    prof = e.profile_report()
    e.profile(False)
    e.set_class_b(0, dtratio)
    kt = kernel_table(prof, {}, cells)
    return {"ntrb": ntrb, "dtratio": dtratio, "ms_per_step": wall / n_timed * 1e3, "timed_steps": n_timed,
            "kernels": {k: v for k, v in kt.items() if k.startswith("eb_")}}


def run_ens_mass_adv(a, e, cells, kb, fimhip):
    """The same batched step with the flux-limited mass transport in place of
    FCT cnuity (flux_conserv_schm = 2): step time and its two kernels, after
    tuning their launch shapes."""
    import torch
    e.set_flux_scheme(2)
    if not a.no_autotune:
        synth_ens(e, a.members, fimhip)
        e.autotune(a.mode, 2)
    synth_ens(e, a.members, fimhip)
    e.step(a.warmup, a.mode)
    e.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.step(a.steps, a.mode)
    e.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e.profile(True)
    e.step(a.steps, a.mode)
    e.sync()
    prof = e.profile_report()
    e.profile(False)
    e.set_flux_scheme(1)
    kt = kernel_table(prof, kb, cells)
    return {"value": cells * a.steps / wall, "ms_per_step": wall / a.steps * 1e3,
            "kernels": {k: kt[k] for k in ("e_massadv1", "e_massadv2") if k in kt}}


def run_uneven(a, th, ds, fimhip):
    """The uneven batch of one launch: the G6 and G7 atmosphere grids
    concatenated into one instance (204804 cells x nvl layers, 4 tracers, at
    the finer grid's time step) and the ocean instance on G8 (655362 cells x
    32 layers, temperature and salinity, ~30 % land, dt = 45 s), both in the
    mixed-precision form, stepped side by side (each step both instances'
    captured steps enqueued on their own streams before any wait);
    beside it each instance alone, and the atmosphere part in FP64; and the
    atmosphere beside HYCOM's own ocean dynamics (bclin_edgvar, cnuity,
    momtum in FP64) on the G8 ocean."""
    import torch
    from fimocn_subx_realtime_hera_apr2020_amd import glvl, ocean
    gu = glvl.concat([fimhip.icos_grid(6), fimhip.icos_grid(7)])
    dt = 1.2 * 4800.0 / float(2 ** 7 // 2)
    cells = gu.nip * a.nvl
    out = {"grids": "G6 + G7 atmosphere in one instance; G8 x 32 ocean instance", "nip": gu.nip, "dt": dt}
    # the atmosphere alone, mixed form and FP64
    for form in ("mixed", "fp64"):
        inst = (fimhip.Ensemble(gu, 1, a.nvl, a.ntra, dt, th, ds, device=0) if form == "mixed"
                else fimhip.Model(gu, a.nvl, a.ntra, dt, th, ds, device=0))
        try:
            if form == "mixed":
                inst.synth(0, seed=SEED, synth_member=0, flags=fimhip.SYNTH_WIND_NOISE)
                inst.init_state(True)
                inst.step(a.warmup, a.mode)
                inst.sync()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                inst.step(a.steps, a.mode)
                inst.sync()
            else:
                inst.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE)
                inst.init_state(True)
                for its in range(1, a.warmup + 1):
                    inst.step(its, a.mode)
                inst.sync()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for its in range(a.warmup + 1, a.warmup + a.steps + 1):
                    inst.step(its, a.mode)
                inst.sync()
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        finally:
            inst.close()
        out[form] = {"value": cells * a.steps / wall, "ms_per_step": wall / a.steps * 1e3}
    out["ratio"] = out["mixed"]["value"] / out["fp64"]["value"]
    # the ocean instance, and both together
    go = fimhip.icos_grid(8)
    tho, dso = ocean.vertical_tables()
    dto = 1.2 * 4800.0 / float(2 ** 8 // 2)
    sto = ocean.make_ocean_state(go.lat, go.lon)
    ocells = go.nip * ocean.NVL
    atm = fimhip.Ensemble(gu, 1, a.nvl, a.ntra, dt, th, ds, device=0)
    oce = fimhip.Ensemble(go, 1, ocean.NVL, ocean.NTRA, dto, tho, dso, device=0)
    try:
        atm.synth(0, seed=SEED, synth_member=0, flags=fimhip.SYNTH_WIND_NOISE)
        atm.init_state(True)
        oce.load_member(0, sto)
        oce.set_wet(0, sto["wet"])
        oce.init_state(True)
        del sto
        oce.step(a.warmup, a.mode)
        oce.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        oce.step(a.steps, a.mode)
        oce.sync()
        torch.cuda.synchronize()
        wo = time.perf_counter() - t0
        fimhip.step_concurrently([atm, oce], a.warmup, a.mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fimhip.step_concurrently([atm, oce], a.steps, a.mode)
        torch.cuda.synchronize()
        wt = time.perf_counter() - t0
        m, _, _, _ = oce.globsum()
    finally:
        atm.close()
        oce.close()
    out["ocean"] = {"grid": "G8 x 32 layers, T and S, 30 % land", "nip": go.nip, "dt": dto,
                    "value": ocells * a.steps / wo, "ms_per_step": wo / a.steps * 1e3}
    out["together"] = {"value": (cells + ocells) * a.steps / wt, "ms_per_step": wt / a.steps * 1e3,
                       "cell_levels_per_step": cells + ocells,
                       "note": "atmosphere and ocean instances, each step every instance's captured step "
                               "enqueued on its own stream before any wait"}
    # the same atmosphere beside HYCOM's own dynamics on the G8 ocean
    # (hycom_edgvar/cnuity/momtum, FP64, 32 layers, ~30 % land): each step
    # the atmosphere's captured step is enqueued, then the ocean's three
    # routines on the ocean instance's stream (cnuity waits for its own
    # kernels only: it reads the duplicate-flux check and the mass sums)
    hcase = ocean.hycom_momtum_case(go, kdm=32)
    nwet = int((hcase["wet"] > 0).sum())
    hcells = nwet * 32
    atm = fimhip.Ensemble(gu, 1, a.nvl, a.ntra, dt, th, ds, device=0)
    hyo = fimhip.Ocean(go, 32, hcase["wet"], hcase["dtleap"])
    try:
        atm.synth(0, seed=SEED, synth_member=0, flags=fimhip.SYNTH_WIND_NOISE)
        atm.init_state(True)
        hyo.set_momtum(hcase["batrop"], hcase["veldff"], hcase["smagcf"])
        for n in fimhip.OCN_MOMTUM_FIELDS:
            hyo.upload(n, hcase[n])

        def hycom_step(st):
            leapm = 1 + st % 2
            hyo.step_edgvar(leapm, 3 - leapm)
            hyo.step(st, leapm, 3 - leapm)
            hyo.step_momtum(leapm, 3 - leapm)
        st = 0
        for _ in range(a.warmup):
            st += 1
            hycom_step(st)
        hyo.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            st += 1
            hycom_step(st)
        hyo.sync()
        torch.cuda.synchronize()
        wh = time.perf_counter() - t0
        for _ in range(a.warmup):
            st += 1
            atm.step(1, a.mode)
            hycom_step(st)
        atm.sync()
        hyo.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            st += 1
            atm.step(1, a.mode)
            hycom_step(st)
        atm.sync()
        hyo.sync()
        torch.cuda.synchronize()
        wth = time.perf_counter() - t0
    finally:
        atm.close()
        hyo.close()
    out["hycom_ocean"] = {"grid": "G8 x 32 layers, HYCOM dynamics (bclin_edgvar, cnuity, momtum), FP64, %d ocean "
                                  "cells" % nwet, "value": hcells * a.steps / wh, "ms_per_step": wh / a.steps * 1e3}
    out["together_hycom"] = {"value": (cells + hcells) * a.steps / wth, "ms_per_step": wth / a.steps * 1e3,
                             "cell_levels_per_step": cells + hcells,
                             "note": "the atmosphere instance's captured step and HYCOM's ocean routines on their "
                                     "own streams, side by side"}
    return out


# HBM bytes each kernel of the ocean continuity equation must move per ocean
# cell-level (FP64: edge arrays 6 x 8 B, cell arrays 8 B; a neighbour's value
# counted once; per-cell grid tables and the clipped-flux sums left out)
HYC_KERNEL_BYTES = {"hyc_flux": 352, "hyc_lowsum": 144, "hyc_limits": 48, "hyc_clip": 184, "hyc_recover": 128,
                    "hyc_del2p": 16, "hyc_bolus": 80, "hyc_bclip": 88, "hyc_final": 248, "hyc_check": 48}
# what a step must read and write at least per ocean cell-level: um, vm, un,
# vn, dpm_edg and mssfx_avg in, massflx and mssfx_avg out (8 edge arrays), dp
# at both leapfrog levels and pres in and out
HYC_STEP_BYTES = 8 * 48 + 2 * 24
# the same for the edge interpolation's kernels (neighbour values counted once)
HYC_EDG_KERNEL_BYTES = {
    # one launch per level: uclin vclin -> u, v edges; the first also dp ->
    # pres and u, v edges -> uvsq (both counted half, the mean per launch)
    "hyc_uvedge": 8 * 2 + 2 * 48 + 8 + 24,
    "hyc_edgint": 8 + 8 + 16 + 50 + 3 * 48,    # dp pres montg spcifv -> pr_edg dp_edg mont spcv
    "hyc_janjic": 8 + 8 + 8 + 48,              # pres geop spcifv -> geop_edg
}
HYC_MOM_KERNEL_BYTES = {
    "hyc_mom_pgf": 16 + 8 + 3 * 48 + 8 + 8 + 16,     # dp (2 levels) u v (new) mont uvsq spcv edges pres -> pres gradx grady
    "hyc_del4": 16 + 16,                            # u v (new) -> their del4prep
    "hyc_dissip": 16 + 8 + 16,                      # del4prep u v, dp -> dissipation tendencies
    "hyc_mom_upd": 96 + 16 + 32 + 16 + 8 + 8 + 32,  # um vm edges, tendencies, u v (2 levels), grad, pres, dp -> u v
}
# bclin_edgvar, cnuity and momtum: in u, v and dp (2 levels), pres, montg,
# spcifv, mssfx_avg; out u, v (2 levels), dp, pres, massflx, mssfx_avg (the
# edge values stay inside)
HYC_OCEAN_STEP_BYTES = (32 + 16 + 8 + 16 + 48) + (32 + 16 + 8 + 48 + 48)
# bclin_edgvar then cnuity: in uclin, vclin and dp (2 levels), pres, montg,
# spcifv, mssfx_avg; out the edge values the ocean step goes on to read (um,
# vm, un, vn, dp_edg, mont_edg, uvsq_edg, spcv_edg on 6 edges, pr_edg on 6
# edges x kdm+1 interfaces), massflx, mssfx_avg, dp (2 levels) and pres
HYC_CHAIN_BYTES = (32 + 16 + 8 + 16 + 48) + (8 * 48 + 50) + (2 * 48 + 16 + 8)

# the tracer transport's kernels with nfld = 2 + numtr tracers: the bytes one
# launch must move per ocean cell-level (the fct3d kernels run all tracers in
# one launch; the arrays every tracer shares, cumuflx, w, dpinit, dpfinl,
# counted once)
def hyc_tr_kernel_bytes(nfld):
    return {
        "hyc_tr_init": 48 + 8 + 8,                     # -> cumuflx; dp -> dpinit
        "hyc_tr_accum": 96 + 48,                       # cumuflx += massflx
        "hyc_tr_w": 48 + 16 + 8 + 16 * nfld,           # cumuflx dpinit dpfinl -> w; tracers -> working copies
        "hyc_fct_vert": 16 + 24 * nfld,                # dpinit w; fld -> vertical flux, its divergence
        "hyc_fct_lo": 48 + 24 + 96 * nfld,             # cumuflx w dpinit dpfinl; fld vfx vdv -> uan flo fmx fmn
        "hyc_fct_ratio": 8 + 104 * nfld,               # fmx fmn flo uan -> flp fln fmx fmn
        "hyc_fct_clip": 8 + 96 * nfld,                 # flp fln uan flo fmx fmn -> fld
        "hyc_fct_fix": 16 * nfld,                      # fld -> the tracers
        "hyc_dff_flux": 8 + 16 * 2,                    # dpfinl; temp saln -> their flux divergence
        "hyc_dff_apply": 8 + 24 * 2,
    }


# one transport interval (hycom_run.F90:386-400, 553): transp1 on every
# second of its 30 steps, transp2, transp0; what it must move at least per
# ocean cell-level: massflx in 15 times, cumuflx in and out, dpinit and dpfinl
# in, each tracer in and out
def hyc_tr_interval_bytes(nfld):
    return 15 * 48 + 96 + 16 + 16 * nfld


# the grid generator (hycom_hybgn1.F90), bytes per ocean cell-level: the
# interface moves read dens, temp, saln, dp and write pres, the old
# interfaces, dens, temp, saln; the remap reads pres, the old interfaces,
# dens, temp, saln, u, v and writes temp, saln, dp, pres, u, v.  The whole
# call must move dens, temp, saln, dp, u, v in and out and pres out.
HYC_HYB_KERNEL_BYTES = {"hyc_hyb_move": 4 * 8 + 5 * 8, "hyc_hyb_remap": 7 * 8 + 6 * 8}
HYC_HYB_STEP_BYTES = 2 * 6 * 8 + 8 + 8


def _ocean_leg(a, o, wcl, step, byte_model, reset=None):
    """Time 'step(nstep)' over a.steps steps after a.warmup, then the same
    number under per-kernel HIP-event timing; throughput, kernel table and
    the dominant kernel against its algorithmic bytes.  With reset, every
    step starts from the state reset() restores, outside the timed region."""
    import torch
    nstep = 1
    for _ in range(a.warmup):
        if reset:
            reset()
        step(nstep)
        nstep += 1
    torch.cuda.synchronize()
    wall = 0.
    t0 = time.perf_counter()
    for _ in range(a.steps):
        if reset:
            reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        step(nstep)
        nstep += 1
        if reset:
            torch.cuda.synchronize()
            wall += time.perf_counter() - t0
    torch.cuda.synchronize()
    if not reset:
        wall = time.perf_counter() - t0
    o.profile(True)
    for _ in range(a.steps):
        if reset:
            reset()
        step(nstep)
        nstep += 1
    prof = o.profile_report()
    o.profile(False)
    kern = {}
    for n, (cnt, tot) in prof.items():
        ms = tot / cnt
        b = byte_model.get(n, 0) * wcl
        kern[n] = {"ms": ms, "launches_per_step": cnt / a.steps, "bytes": b, "GB_s": b / (ms * 1e-3) / 1e9}
    dom = max(kern, key=lambda n: kern[n]["ms"] * kern[n]["launches_per_step"])
    dk = kern[dom]
    return {"value": wcl * a.steps / wall, "unit": "ocean cell-level-steps/s", "ms_per_step": wall / a.steps * 1e3,
            "kernel_ms_per_step": sum(v["ms"] * v["launches_per_step"] for v in kern.values()),
            "roofline": {"bound": "hbm", "achieved": dk["GB_s"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": dk["GB_s"] * 1e9 / HBM_PEAK, "kernel": dom,
                         "bytes_per_cell_level": byte_model.get(dom, 0)},
            "kernels": kern}


def run_hycom(a, fimhip, cpu=True):
    """The HYCOM ocean on G8 (655362 cells) x 32 layers, ~30 % land,
    biharmonic interface smoothing, dtleap 1800 s, FP64 on device-resident
    arrays: its continuity equation alone (hycom_cnuity.F90 cnuity); the
    edge interpolation followed by the continuity equation, as hycom_run
    calls them (hycom_edgvar.F90 bclin_edgvar, then cnuity on its edge
    values); and those two followed by the momentum equation
    (hycom_momtum.F90 momtum with hycom_dissip), the velocities it writes
    feeding the next step; and the tracer transport over one 30-step
    interval (hycom_transp3d.F90 transp1 every second step, transp2 with
    hycom_fct3d and hycom_dffusn, transp0); and the hybrid grid generator
    (hycom_hybgn1.F90 hybgn1); and all of them together over a whole 30-step
    interval in hycom_run's order.  The checker's C restatement of each
    routine is timed on one host core over 2 steps (one interval, one call)
    of the same input (not for the whole interval: 30 steps of it take
    minutes on one core)."""
    import numpy as np
    from fimocn_subx_realtime_hera_apr2020_amd import ocean
    g = fimhip.icos_grid(8)
    kdm, thkdff, biharm = 32, 0.05, 0.75
    case = ocean.hycom_momtum_case(g, kdm=kdm)
    # the continuity equation alone reads hycom_cnuity_case's edge values (the
    # same grid, mask and thickness; each edge's two sides carry one velocity)
    cnu = ocean.hycom_cnuity_case(g, kdm=kdm, seed=ocean.EDGVAR_CASE_SEED)
    nwet = int((case["wet"] > 0).sum())
    wcl = nwet * kdm
    workload = ("G8 (%d cells, %d ocean) x %d layers, thkdff %g, biharm %g, dtleap %g s, FP64, resident"
                % (g.nip, nwet, kdm, thkdff, biharm, case["dtleap"]))
    bytes_all = dict(HYC_KERNEL_BYTES, **HYC_EDG_KERNEL_BYTES, **HYC_MOM_KERNEL_BYTES)
    out = {"workload": workload}
    o = fimhip.Ocean(g, kdm, case["wet"], case["dtleap"], thkdff=thkdff, biharm=biharm)
    try:
        for n in fimhip.OCN_CNUITY_FIELDS:
            o.upload(n, cnu[n])
        leg = _ocean_leg(a, o, wcl, lambda st: o.step(st, 1 + st % 2, 2 - st % 2), bytes_all)
        ms = leg["ms_per_step"]
        leg["step_roofline"] = {"compulsory_bytes_per_cell_level": HYC_STEP_BYTES,
                                "achieved": HYC_STEP_BYTES * wcl / (ms * 1e-3) / 1e9, "unit": "GB/s",
                                "frac": HYC_STEP_BYTES * wcl / (ms * 1e-3) / HBM_PEAK}
        out["cnuity"] = leg

        def chain(st):
            o.step_edgvar(1 + st % 2, 2 - st % 2)
            o.step(st, 1 + st % 2, 2 - st % 2)
        for n in fimhip.OCN_EDGVAR_FIELDS:
            o.upload(n, case[n])
        leg = _ocean_leg(a, o, wcl, chain, bytes_all)
        ms = leg["ms_per_step"]
        leg["step_roofline"] = {"compulsory_bytes_per_cell_level": HYC_CHAIN_BYTES,
                                "achieved": HYC_CHAIN_BYTES * wcl / (ms * 1e-3) / 1e9, "unit": "GB/s",
                                "frac": HYC_CHAIN_BYTES * wcl / (ms * 1e-3) / HBM_PEAK}
        out["edgvar_cnuity"] = leg

        def ocean_step(st):
            o.step_edgvar(1 + st % 2, 2 - st % 2)
            o.step(st, 1 + st % 2, 2 - st % 2)
            o.step_momtum(1 + st % 2, 2 - st % 2)
        o.set_momtum(case["batrop"], case["veldff"], case["smagcf"])
        for n in fimhip.OCN_MOMTUM_FIELDS:
            o.upload(n, case[n])
        leg = _ocean_leg(a, o, wcl, ocean_step, bytes_all)
        ms = leg["ms_per_step"]
        leg["step_roofline"] = {"compulsory_bytes_per_cell_level": HYC_OCEAN_STEP_BYTES,
                                "achieved": HYC_OCEAN_STEP_BYTES * wcl / (ms * 1e-3) / 1e9, "unit": "GB/s",
                                "frac": HYC_OCEAN_STEP_BYTES * wcl / (ms * 1e-3) / HBM_PEAK}
        out["ocean_step"] = leg
    finally:
        o.close()
    # the tracer transport, one interval of 30 steps per timed step (transp1
    # 15 times, transp2, transp0), temperature, salinity and one passive
    # tracer; transp1's mass flux scaled so that the interval integrates it
    # to the case's cumuflx size
    tcase = ocean.hycom_transport_case(g, kdm=kdm, numtr=1)
    tcase["massflx_tr"] = np.where(tcase["massflx_tr"] == -1.0e33, -1.0e33,
                                   tcase["massflx_tr"] * (0.2 / 2.0e-4) / (2.0 * tcase["batrop"] * 15))
    tcase["dp"] = np.stack([tcase["dpinit"], tcase["dpfinl"]])
    nfld = 2 + tcase["numtr"]
    o = fimhip.Ocean(g, kdm, tcase["wet"], tcase["dtleap"], thkdff=thkdff, biharm=biharm)
    try:
        o.set_momtum(tcase["batrop"])
        o.set_transport(tcase["numtr"], tcase["temdff"])
        for n in ("temp", "saln", "passv", "dpinit", "dp"):
            o.upload(n, tcase[n])
        o.upload("massflx", tcase["massflx_tr"])

        def interval(st):
            o.step_transp0(1)
            for _ in range(15):
                o.step_transp1()
            o.step_transp2(2)
        leg = _ocean_leg(a, o, wcl, interval, hyc_tr_kernel_bytes(nfld))
        ms = leg["ms_per_step"]
        ib = hyc_tr_interval_bytes(nfld)
        leg["unit"] = "ocean cell-level-intervals/s"
        leg["tracers"] = nfld
        leg["step_roofline"] = {"compulsory_bytes_per_cell_level": ib,
                                "achieved": ib * wcl / (ms * 1e-3) / 1e9, "unit": "GB/s",
                                "frac": ib * wcl / (ms * 1e-3) / HBM_PEAK}
        out["transport"] = leg
    finally:
        o.close()
    # the hybrid grid generator (hybgn1, as hycom_run calls it: no tracers),
    # one call per timed step on alternating levels
    hcase = ocean.hycom_hybgen_case(g, kdm=kdm)
    o = fimhip.Ocean(g, kdm, hcase["wet"], case["dtleap"], thkdff=thkdff, biharm=biharm)
    try:
        o.set_momtum(hcase["batrop"])
        o.set_hybgen(hcase["targt"], hcase["salmin"], hcase["dpmin"], hcase["latij"])
        for n, m in (("dens", "dens"), ("dp", "dp"), ("pres", "pres"), ("temp", "temp"), ("saln", "saln"),
                     ("uvel", "uclin"), ("vvel", "vclin")):
            o.upload(m, hcase[n])
        leg = _ocean_leg(a, o, wcl, lambda st: o.step_hybgn1(1 + st % 2), HYC_HYB_KERNEL_BYTES)
        ms = leg["ms_per_step"]
        hb = HYC_HYB_STEP_BYTES
        leg["step_roofline"] = {"compulsory_bytes_per_cell_level": hb,
                                "achieved": hb * wcl / (ms * 1e-3) / 1e9, "unit": "GB/s",
                                "frac": hb * wcl / (ms * 1e-3) / HBM_PEAK}
        out["hybgen"] = leg
    finally:
        o.close()
    # a whole 30-step interval of the ocean's dynamics in hycom_run's order:
    # bclin_edgvar, cnuity and momtum every step, transp1 every second step,
    # transp2, hybgn1 and transp0 at the end (temperature, salinity and one
    # passive tracer); a timed step is one interval
    import types
    ocase = ocean.hycom_ocean_case(g, kdm=kdm, numtr=1)
    o = fimhip.Ocean(g, kdm, ocase["wet"], ocase["dtleap"], thkdff=thkdff, biharm=biharm)
    try:
        o.set_momtum(ocase["batrop"], ocase["veldff"], ocase["smagcf"])
        o.set_transport(ocase["numtr"], ocase["temdff"])
        o.set_hybgen(ocase["targt"], ocase["salmin"], ocase["dpmin"], ocase["latij"])
        state = {"st": 0}

        def reset():
            # the synthetic state drifts without the routines outside this
            # path (hystat, barotp, the mixing): each interval starts afresh
            for n in fimhip.OCN_MOMTUM_FIELDS + ["temp", "saln", "dens", "passv", "cumuflx", "dpinit"]:
                o.upload(n, ocase[n])
            o.mass = (0.0, 0.0, 0.0)
            o.step_transp0(1)
            state["st"] = 0

        def interval(_):
            for _j in range(30):
                state["st"] += 1
                st = state["st"]
                leapm = 1 + st % 2
                leapn = 3 - leapm
                o.step_edgvar(leapm, leapn)
                o.step(st, leapm, leapn)
                o.step_momtum(leapm, leapn)
                if st % 2 == 0:
                    o.step_transp1()
            o.step_transp2(leapn)
            o.step_hybgn1(leapn)
            o.step_transp0(leapn)
        ia = types.SimpleNamespace(steps=max(1, a.steps // 10), warmup=1 if a.steps > 2 else 0)
        bytes_all = dict(HYC_KERNEL_BYTES, **HYC_EDG_KERNEL_BYTES, **HYC_MOM_KERNEL_BYTES,
                         **hyc_tr_kernel_bytes(2 + ocase["numtr"]), **HYC_HYB_KERNEL_BYTES)
        leg = _ocean_leg(ia, o, wcl, interval, bytes_all, reset=reset)
        leg["unit"] = "ocean cell-level-intervals/s"
        leg["steps_per_interval"] = 30
        leg["ocean_cell_level_steps_per_s"] = leg["value"] * 30
        leg["intervals_timed"] = ia.steps
        out["ocean_interval"] = leg
    finally:
        o.close()
    if cpu:
        from oracle import oracle as O
        for name, with_edgvar, with_momtum in (("cnuity", False, False), ("edgvar_cnuity", True, False),
                                               ("ocean_step", True, True)):
            h = O.HycomCnuity(g, case if with_edgvar else cnu, thkdff=thkdff, biharm=biharm)
            t0 = time.perf_counter()
            for st in (1, 2):
                if with_edgvar:
                    h.edgvar(g, 1 + st % 2, 2 - st % 2)
                h.step(st, 1 + st % 2, 2 - st % 2)
                if with_momtum:
                    h.momtum(g, 1 + st % 2, 2 - st % 2)
            tc = time.perf_counter() - t0
            leg = out[name]
            leg["cpu_checker"] = {"value": wcl * 2 / tc, "unit": "ocean cell-level-steps/s", "cores": 1,
                                  "kind": "port", "sample": "2 steps of the same G8 x 32 input",
                                  "ms_per_step": tc / 2 * 1e3}
            leg["speedup_vs_cpu_checker"] = leg["value"] / leg["cpu_checker"]["value"]
        h = O.HycomCnuity(g, tcase)
        t0 = time.perf_counter()
        h.transport(leapn=1, transp0=True)
        h.transport(transp1_calls=15)
        tc = time.perf_counter() - t0
        leg = out["transport"]
        leg["cpu_checker"] = {"value": wcl / tc, "unit": "ocean cell-level-intervals/s", "cores": 1, "kind": "port",
                              "sample": "1 interval (transp0, 15 x transp1, transp2) of the same G8 x 32 input",
                              "ms_per_step": tc * 1e3}
        leg["speedup_vs_cpu_checker"] = leg["value"] / leg["cpu_checker"]["value"]
        t0 = time.perf_counter()
        O.hycom_hybgn1(hcase, 1)
        tc = time.perf_counter() - t0
        leg = out["hybgen"]
        leg["cpu_checker"] = {"value": wcl / tc, "unit": "ocean cell-level-steps/s", "cores": 1, "kind": "port",
                              "sample": "1 call of the same G8 x 32 input", "ms_per_step": tc * 1e3}
        leg["speedup_vs_cpu_checker"] = leg["value"] / leg["cpu_checker"]["value"]
    return out


def accuracy(a, e, m, grid, fimhip):
    """Member 0 of the ensemble against the FP64 path after n steps, and the
    global mass drift of every member."""
    import numpy as np
    n = a.accuracy_steps
    M = a.members
    synth_ens(e, M, fimhip)
    m.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    m.init_state(True)
    m0, t0, _, _ = e.globsum()
    p0, _ = m.globsum()
    e.step(n, a.mode)
    for its in range(1, n + 1):
        m.step(its, a.mode)
    m.sync()
    e.sync()
    m1, t1, k1, kt1 = e.globsum()
    p1, _ = m.globsum()

    def rl2(x, y):
        return float(np.sqrt(np.sum((x - y) ** 2) / np.sum(y ** 2)))

    err = {f: rl2(e.download(0, f), m.download(f)) for f in ("us", "vs", "dp", "pr", "ex", "ph", "mp")}
    tre, trp = e.download(0, "tr"), m.download("tr")
    for t in range(a.ntra):
        err["tr%d" % (t + 1)] = rl2(tre[t], trp[t])
    return {"steps": n, "rel_l2_vs_fp64": err,
            "mass_drift_max": float(np.max(np.abs(m1 - m0) / m0)),
            "mass_drift_fp64_path": float(abs(p1 - p0) / p0),
            "tracer_drift_max": [float(x) for x in np.max(np.abs(t1 - t0) / np.abs(t0), axis=0)],
            "kahan_fp32_vs_fp64_mass": float(np.max(np.abs(k1 - m1) / m1)),
            "kahan_fp32_vs_fp64_tracers": float(np.max(np.abs(kt1 - t1) / np.abs(t1)))}


def cpu_baseline(m, grid, a, th, ds, fimhip):
    """Time the C restatement on the host: one member of the same grid,
    started from the device's post-start-up state, a few full steps."""
    from oracle import oracle as O
    # every host thread this process may use: the job's share where the
    # environment sets one (OMP_NUM_THREADS), else every logical CPU
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = O.OracleRun(O.grid_arrays(grid, a.nvl, a.dt, th, ds), a.nvl, a.ntra, a.dt, pow_kind=0)
    m.synth(seed=SEED, member=0, flags=fimhip.SYNTH_WIND_NOISE | fimhip.SYNTH_PERTURB)
    m.init_state(True)
    for n in ("us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "trdp"):
        o.a[n][...] = m.download(n)
    o.s.nf, o.s.of, o.s.vof = 0, 2, 1
    t0 = time.perf_counter()
    for its in range(1, a.cpu_steps + 1):
        o.step(its, a.mode)
    wall = time.perf_counter() - t0
    cells = grid.nip * a.nvl
    value = cells * a.cpu_steps / wall
    ratio = fortran_ratio()
    out = {"value": value, "unit": "cell-level-steps/s", "cores": threads,
           "kind": "port",
           "sample": "1 member of G%d x %dL, %d tracers, %d %s FP64 steps (C restatement, OpenMP %d threads), "
                     "%.1f s" % (a.glvl, a.nvl, a.ntra, a.cpu_steps, "full" if a.mode == 0 else "transport",
                                 threads, wall),
           "host": {"cpu_model": cpu_model(), "logical_cpus": os.cpu_count(), "threads_used": threads},
           "fortran_ratio": ratio,
           "fortran_equivalent": (value / ratio) if ratio else None,
           "fortran_note": "the reference Fortran (-O2 -fopenmp -fdefault-real-8) runs the same steps at "
                           "1/fortran_ratio times the restatement's rate (profiles/cpu_calibration_G5x64.md, "
                           "8 threads, G5 x 64)"}
    if not a.headline_only:
        out["schemes"] = cpu_schemes(o, m, grid, a, th, ds, fimhip, O, threads)
    return out


def cpu_model():
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:      # lscpu absent: the model is then unknown
        pass
    return "unknown"


def fortran_ratio():
    """restatement / reference-Fortran throughput at equal threads, from the
    committed calibration (profiles/cpu_calibration_G5x64.md)."""
    p = os.path.join(ROOT, "profiles", "cpu_calibration_G5x64.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("restatement / reference (FP64,") and "thread)" not in line:
            return float(line.split("**")[1])
    return None


def cpu_schemes(o, m, grid, a, th, ds, fimhip, O, threads, ntrb=8):
    """The C restatement on the alternative schemes of the bench legs: the
    same steps with mass_adv, and one class-B transport call (transp2 with
    fct3d of ntrb tracers) after one step's integrated mass flux."""
    cells = grid.nip * a.nvl
    o.g.flux_scheme = 2
    for n in ("us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "trdp"):
        o.a[n][...] = m.download(n)
    for n in ("u_tdcy", "v_tdcy", "dp_tdcy", "dpl_tdcy", "trc_tdcy", "trl_tdcy", "massfx"):
        o.a[n][...] = 0.
    o.s.nf, o.s.of, o.s.vof = 0, 2, 1
    t0 = time.perf_counter()
    for its in range(1, a.cpu_steps + 1):
        o.step(its, a.mode)
    mwall = time.perf_counter() - t0
    o.g.flux_scheme = 1
    ob = O.OracleRun(o.garr, a.nvl, a.ntra, a.dt, pow_kind=0, ntrb=ntrb, dtratio=1)
    for n in ("dp", "massfx"):
        ob.a[n][...] = o.a[n]
    lg = fimhip.synth_state(grid.lat, grid.lon, a.nvl, 1 + ntrb, seed=7, member=3, flags=fimhip.SYNTH_LOGNORMAL)
    ob.a["trb"][...] = lg["tr"][1:]
    del lg
    ob.s.nf, ob.s.of, ob.s.vof = o.s.nf, o.s.of, o.s.vof
    ob.s.ab1, ob.s.ab2, ob.s.ab3 = o.s.ab1, o.s.ab2, o.s.ab3
    ob.a["dpinit"][...] = o.a["dp"] * 1.0001      # a thickness change, so the vertical fluxes are not zero
    ob.call("transp1")
    t0 = time.perf_counter()
    ob.call("transp2")
    bwall = time.perf_counter() - t0
    return {"mass_adv": {"value": cells * a.cpu_steps / mwall, "unit": "cell-level-steps/s",
                         "sample": "%d %s FP64 steps with flux_conserv_schm = 2" % (a.cpu_steps,
                                                                                   "full" if a.mode == 0 else "transport")},
            "class_b_transp2": {"ms_per_call": bwall * 1e3, "ntrb": ntrb,
                                "sample": "one transp2 (fct3d of %d class-B tracers) on G%d x %dL" % (ntrb, a.glvl, a.nvl)},
            "cores": threads, "kind": "port"}


def main():
    a = parse()
    if a.dt is None:        # module_control.F90:214-220 (SubdivNum = 2 at every level)
        a.dt = 1.2 * 4800.0 / float(2 ** a.glvl // 2)
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only; ensembles would shard by member and one large "
              "member by IJ blocks with a 1-ring halo, which is not run here", file=sys.stderr)
        sys.exit(2)
    import torch
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, state

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    grid = fimhip.icos_grid(a.glvl)
    th, ds = state.vertical_tables(a.nvl)
    if a.copy_calibration:
        x = torch.empty(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    step_name = "full" if a.mode == 0 else "transport"
    if a.form == "hycom":
        print(json.dumps({"hycom": run_hycom(a, fimhip, cpu=not a.no_cpu_baseline)}))
        return
    plain = ens = None
    m = e = None
    if a.form in ("both", "plain"):
        m, plain = run_plain(a, grid, th, ds, fimhip)
        print("plain FP64: %.3f ms/step" % plain["ms_per_step"], file=sys.stderr, flush=True)
    if a.form in ("both", "ens"):
        e, ens = run_ens(a, grid, th, ds, fimhip)
        print("ensemble: %.3f ms/step for %d members" % (ens["ms_per_step"], a.members), file=sys.stderr,
              flush=True)
    uneven = None
    if a.form == "both" and not a.headline_only:
        uneven = run_uneven(a, th, ds, fimhip)
        print("uneven batch done", file=sys.stderr, flush=True)
    hyc = None
    if a.form == "both" and not a.headline_only:
        hyc = run_hycom(a, fimhip, cpu=not a.no_cpu_baseline)
        print("HYCOM leg done", file=sys.stderr, flush=True)
    acc = None
    if e is not None and m is not None and a.accuracy_steps > 0:
        acc = accuracy(a, e, m, grid, fimhip)
        print("accuracy leg done", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu_baseline:
        if m is None:
            m = fimhip.Model(grid, a.nvl, a.ntra, a.dt, th, ds, device=0)
        cpu = cpu_baseline(m, grid, a, th, ds, fimhip)
    if e is not None:
        e.close()
    if m is not None:
        m.close()

    head = ens if ens is not None else plain
    if ens is not None:
        workload = ("%d members x G%d (%d cells) x %d layers, %d tracers, dt=%gs, %s steps, FP32 form with FP64 "
                    "thickness/pressure/PGF/hybgen, graph-replayed step, autotuned tiles"
                    % (a.members, a.glvl, grid.nip, a.nvl, a.ntra, a.dt, step_name))
        form, dtype = "fp32-batched", "f32+f64"
    else:
        workload = ("1 member x G%d (%d cells) x %d layers, %d tracers, dt=%gs, %s steps, FP64 plain path"
                    % (a.glvl, grid.nip, a.nvl, a.ntra, a.dt, step_name))
        form, dtype = "plain-fp64", "f64"
    out = {
        "metric": "cell-level-steps/s at G7x64L, FP32+batched vs FP64 plain; % HBM peak",
        "value": head["value"], "unit": "cell-level-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded member states generated on the device)",
        "config": {"workload": workload, "nip": grid.nip, "nvl": a.nvl, "ntra": a.ntra,
                   "members": head["members"], "form": form, "dt": a.dt,
                   "dt_note": "the reference's own time step for this grid (module_control.F90:214-220); at G7 "
                              "a 240 s step overflows the FP64 state by step 6, so it is not timed"},
        "roofline": head["roofline"], "dominant_kernel": head["dominant_kernel"],
        "speedup_vs_plain_fp64": (ens["value"] / plain["value"]) if (ens and plain) else None,
        "plain_fp64": plain, "ensemble": ens, "uneven_batch": uneven, "hycom": hyc,
        "accuracy_96": acc, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
