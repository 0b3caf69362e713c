"""Benchmark of the FIM horizontal dynamics step on one MI355X.

Workload (default): one member of the G7 (163842 cells) x 64-layer state with
4 tracers, dt = 240 s, full dynamics steps (abstart, edgvar1, edgvar3,
cnuity, trcadv, momtum, hybgen, hystat) on the FP64 path, seeded synthetic
state generated on the device.  One step = one pass of the hot path over the
whole state; value = members * nip * nvl * steps / wall seconds
(cell-level-steps/s), with the state resident in HBM before the timed region.

Prints ONE JSON line.  Besides the contract fields it carries
  roofline:     the dominant kernel (most device time in the timed region,
                timed with HIP events on the library's stream): algorithmic
                bytes per launch / mean launch duration against 8 TB/s
  step_hbm:     the whole step against the compulsory-byte model
                (18 + 8T values per cell-level-step, 8 B each in FP64)
  cpu_baseline: the C restatement (oracle/fim_oracle.c, FP64, OpenMP) timed
                on this host's cores on a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12

# Algorithmic (compulsory) HBM values touched per cell-level by each FP64
# kernel of one launch: distinct array elements it must read plus those it
# writes, with T tracers.  Each value is 8 bytes.
def kernel_values(T):
    return {
        "k_edgvar1": 5 + 36,                 # us vs pr mp tr1 -> u v lp mont uvsq trc1 on 6 edges
        "k_edgvar3": 7 * T,                  # dp, tracers 2..T -> 6 edge values each
        "k_cnuity12": 18 + 3 + 6 + 6 + 2,    # u/v/dp edges, dp, dplo(of,vof) -> massfx, antifx, dplo, delp_lo
        "k_cnuity3": 6 + 4 + 2,              # antifx, delp_lo, dp, dp_tdcy(of,vof) -> r+, r-
        "k_cnuity45": 27 + 11,               # antifx r+- massfx dplo dptdcy dp pr lp_edg u v -> massfx dptdcy dp pr ex omega
        "k_trcadv1": 27 * T,
        "k_trcadv2": 16 * T,
        "k_trcadv3": 18 * T,
        "k_momtum": 3 + 1 + 30 + 4 + 5,
        "k_regrid": 4,
        "k_dffusn_lev": 2,
        "k_remap": 11 + 2 * T,
        "k_hystat": 5,
        "k_hystat_tr": 2 * T + 3,
    }


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--glvl", type=int, default=7)
    ap.add_argument("--nvl", type=int, default=64)
    ap.add_argument("--ntra", type=int, default=4)
    ap.add_argument("--dt", type=float, default=240.0)
    ap.add_argument("--mode", type=int, default=0, help="0 full step, 1 transport only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (for PMC byte calibration)")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (one MI355X per box); "
              "ensembles would shard by member, one large member by IJ blocks with a 1-ring halo; not run here", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, state

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    grid = fimhip.icos_grid(a.glvl)
    th, ds = state.vertical_tables(a.nvl)
    m = fimhip.Model(grid, a.nvl, a.ntra, a.dt, th, ds, device=0)
    flags = fimhip.SYNTH_WIND_NOISE
    m.synth(seed=20240601, member=0, flags=flags)
    m.init_state(True)
    if a.copy_calibration:
        x = torch.empty(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
    its = 0
    for _ in range(a.warmup):
        its += 1
        m.step(its, a.mode)
    m.sync()
    m.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        its += 1
        m.step(its, a.mode)
    m.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = m.profile_report()
    m.profile(False)

    nip, nvl, T = grid.nip, a.nvl, a.ntra
    cells = nip * nvl
    value = cells * a.steps / wall
    vals = kernel_values(T)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dname, (dn, dms) = dom
    dmean = dms / dn * 1e-3
    dbytes = vals.get(dname, 0) * 8 * cells
    roof = {"kernel": dname, "bound": "hbm", "achieved": dbytes / dmean / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": dbytes / dmean / HBM_PEAK, "traffic": None,
            "bytes_per_launch": dbytes, "mean_launch_ms": dmean * 1e3}
    v_step = (18 + 8 * T) if a.mode == 0 else (10 + 8 * T)
    step_hbm = {"values_per_cell_level": v_step, "bytes_per_value": 8,
                "achieved_GBs": value * v_step * 8 / 1e9, "frac": value * v_step * 8 / HBM_PEAK}
    kernels = {k: {"launches": n, "mean_ms": ms / n, "share": ms / sum(v[1] for v in prof.values())}
               for k, (n, ms) in sorted(prof.items(), key=lambda kv: -kv[1][1])}

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(m, grid, a, th, ds)
    m.close()

    out = {
        "metric": "cell-level-steps/s at G7x64L, FP32+batched vs FP64 plain; % HBM peak",
        "value": value, "unit": "cell-level-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded, generated on the device)",
        "config": {"workload": "G%d x %dL, %d tracers, dt=%gs, %s step, FP64 plain path, 1 member"
                   % (a.glvl, nvl, T, a.dt, "full" if a.mode == 0 else "transport"),
                   "nip": nip, "nvl": nvl, "ntra": T, "members": 1, "form": "plain-fp64"},
        "roofline": roof, "step_hbm": step_hbm, "kernels": kernels, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(m, grid, a, th, ds):
    """Time the C restatement on the host: one member of the same workload,
    started from the device's post-start-up state, a few full steps."""
    import numpy as np
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = O.OracleRun(O.grid_arrays(grid, a.nvl, a.dt, th, ds), a.nvl, a.ntra, a.dt, pow_kind=0)
    # fresh member state at the step the device started from
    m.synth(seed=20240601, member=0, flags=1)
    m.init_state(True)
    for n in ("us", "vs", "dp", "pr", "ex", "ph", "mp", "tr", "trdp"):
        o.a[n][...] = m.download(n)
    o.s.nf, o.s.of, o.s.vof = 0, 2, 1
    t0 = time.perf_counter()
    for its in range(1, a.cpu_steps + 1):
        o.step(its, a.mode)
    wall = time.perf_counter() - t0
    cells = grid.nip * a.nvl
    return {"value": cells * a.cpu_steps / wall, "unit": "cell-level-steps/s", "cores": threads,
            "kind": "port",
            "sample": "1 member of G%d x %dL, %d tracers, %d full FP64 steps (C restatement, OpenMP %d threads)"
                      % (a.glvl, a.nvl, a.ntra, a.cpu_steps, threads)}


if __name__ == "__main__":
    main()
