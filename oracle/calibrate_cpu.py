"""Time the reference FIM Fortran against the C restatement on the host.

Test infrastructure (run where /root/reference was present when
oracle/build_ref.sh built oracle/_ref):

    python oracle/calibrate_cpu.py [--glvl 5] [--steps 20] [--threads 8]

Both run full dynamics steps (start-up hybgen, then DYN_1 + DYN_2 per step)
from the same seeded state on the reference's own grid.  The reference
programs are timed as wall clock of nsteps minus that of 0 steps (the same
program stopped after the first --skip steps), so start-up, file I/O and the
first steps' page faults cancel; the restatement is timed around the same
steps.  Prints a
markdown table (the committed profiles/cpu_calibration_G5x64.md) with
cell-level-steps/s of:
  dynref_r8omp   the reference, -O2 -fopenmp -fdefault-real-8, OMP threads
  dynref_r8      the reference, -O2 -fdefault-real-8, serial
  dynref_r4      the reference, -O2, default real 4 (the production precision), serial
  restatement    oracle/fim_oracle.c (FP64, OpenMP), the same thread counts
and the ratios restatement / reference at equal thread counts, the factor
that turns the GPU box's restatement baseline into reference-Fortran terms.
"""
import argparse
import os
import platform
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--glvl", type=int, default=5)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--skip", type=int, default=10, help="steps before the timed ones")
    ap.add_argument("--reps", type=int, default=2, help="reference runs per step count (the fastest is kept)")
    a = ap.parse_args()
    from oracle import refrun, oracle as O
    from fimocn_subx_realtime_hera_apr2020_amd import fimhip, glvl, state
    if not refrun.available() or not os.path.exists(os.path.join(refrun.BIN, "dynref_r8omp")):
        sys.exit("oracle/_ref not built")
    path = refrun.ref_glvl(a.glvl)
    g = glvl.read_glvl(path)
    nvl, ntra = 64, 4
    dt = 1.2 * 4800.0 / float(2 ** a.glvl // 2)
    th, ds = state.vertical_tables(nvl)
    st = fimhip.synth_state(g.lat, g.lon, nvl, ntra, seed=20240601, member=0, flags=fimhip.SYNTH_WIND_NOISE)
    cells = g.nip * nvl
    rows = []

    n1, n2 = a.skip, a.skip + a.steps

    def time_ref(prec, threads):
        env_old = os.environ.get("OMP_NUM_THREADS")
        os.environ["OMP_NUM_THREADS"] = str(threads)
        try:
            walls = []
            for n in (n1, n2):
                best = None
                for _ in range(a.reps):
                    t0 = time.perf_counter()
                    refrun.run_dynref(path, st, th, ds, nsteps=n, dt=dt, mode=0, prec=prec)
                    w = time.perf_counter() - t0
                    best = w if best is None else min(best, w)
                walls.append(best)
        finally:
            if env_old is None:
                os.environ.pop("OMP_NUM_THREADS", None)
            else:
                os.environ["OMP_NUM_THREADS"] = env_old
        return cells * a.steps / (walls[1] - walls[0])

    def time_restatement(threads):
        # the OpenMP runtime of liboracle.so reads OMP_NUM_THREADS once, so
        # each thread count runs in its own process
        code = ("import sys, time; sys.path.insert(0, %r)\n"
                "from oracle import oracle as O\n"
                "from fimocn_subx_realtime_hera_apr2020_amd import fimhip, glvl, state\n"
                "g = glvl.read_glvl(%r); th, ds = state.vertical_tables(%d)\n"
                "st = fimhip.synth_state(g.lat, g.lon, %d, %d, seed=20240601, member=0, flags=fimhip.SYNTH_WIND_NOISE)\n"
                "o = O.OracleRun(O.grid_arrays(g, %d, %r, th, ds), %d, %d, %r, pow_kind=0)\n"
                "o.load(st); o.init(True)\n"
                "for its in range(1, %d): o.step(its, 0)\n"
                "t0 = time.perf_counter()\n"
                "for its in range(%d, %d): o.step(its, 0)\n"
                "print(time.perf_counter() - t0)\n"
                % (ROOT, path, nvl, nvl, ntra, nvl, dt, nvl, ntra, dt, n1 + 1, n1 + 1, n2 + 1))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, check=True)
        return cells * a.steps / float(r.stdout.strip().splitlines()[-1])

    res = {
        ("dynref_r8omp", a.threads): time_ref("r8omp", a.threads),
        ("dynref_r8omp", 1): time_ref("r8omp", 1),
        ("dynref_r8", 1): time_ref("r8", 1),
        ("dynref_r4", 1): time_ref("r4", 1),
        ("restatement", a.threads): time_restatement(a.threads),
        ("restatement", 1): time_restatement(1),
    }
    print("# CPU calibration: reference Fortran vs the C restatement, G%d x %d layers\n" % (a.glvl, nvl))
    print("Command (in the build container, reference built by `oracle/build_ref.sh`):\n")
    print("    python oracle/calibrate_cpu.py --glvl %d --steps %d --threads %d --skip %d --reps %d\n"
          % (a.glvl, a.steps, a.threads, a.skip, a.reps))
    print("Host: %s, %d logical CPUs.  Steps %d-%d of a run of full steps at dt = %g s (start-up hybgen "
          "first), 4 tracers, seeded state (member 0, wind noise).  Reference timed as wall(%d steps) - "
          "wall(%d steps), the faster of %d runs each.\n"
          % (cpu_model(), os.cpu_count() or 0, n1 + 1, n2, dt, n2, n1, a.reps))
    print("| program | build | threads | cell-level-steps/s |")
    print("|---|---|---:|---:|")
    builds = {"dynref_r8omp": "-O2 -fopenmp -fdefault-real-8", "dynref_r8": "-O2 -fdefault-real-8",
              "dynref_r4": "-O2 (default real 4)", "restatement": "gcc -O2 -fopenmp, FP64"}
    for (name, thr), v in res.items():
        print("| %s | %s | %d | %.4g |" % (name, builds[name], thr, v))
    r_par = res[("restatement", a.threads)] / res[("dynref_r8omp", a.threads)]
    r_ser = res[("restatement", 1)] / res[("dynref_r8", 1)]
    print("\nrestatement / reference (FP64, %d threads): **%.3f**" % (a.threads, r_par))
    print("\nrestatement / reference (FP64, 1 thread): **%.3f**" % r_ser)
    print("\nreference real 4 / real 8 (1 thread): %.3f" % (res[("dynref_r4", 1)] / res[("dynref_r8", 1)]))
    print("\nThe GPU box's baseline (the restatement on the box's threads) divided by the %d-thread "
          "ratio is its estimate in reference-Fortran terms." % a.threads)


if __name__ == "__main__":
    main()
