"""Drive the reference FIM programs built by oracle/build_ref.sh.

Test infrastructure only (the checker): it runs the reference grid prep chain
(grid -> ginfo -> getlvl) and the reference dynamics driver (dynref_r8 /
dynref_r4) in scratch directories under oracle/_ref/, and reads their output.
Available only where /root/reference was present at build time.
"""
from __future__ import annotations

import os
import resource
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "_ref")
BIN = os.path.join(REF, "bin")
TOP_GRID = os.environ.get("FIM_REFERENCE", "/root/reference") + "/FIMrun/top_grid"

_NAMELISTS = ("chemwrf DIAGnamelist gfsphys ISOBARICnamelist LANDnamelist MODELnamelist "
              "OCEANnamelist OUTPUTnamelist PHYSICSnamelist POSTnamelist QUEUEnamelist "
              "SYSTEMnamelist TIMEnamelist TOPOnamelist wrfphysics WRITETASKnamelist").split()


def available() -> bool:
    return all(os.path.exists(os.path.join(BIN, b)) for b in ("grid", "ginfo", "getlvl", "dynref_r8"))


def ref_glvl(glvl: int) -> str:
    """Path of the reference glvl.dat for an all-bisection grid of level glvl
    (gtype=2 modified recursive, curve=0 IJ order), generated on first use."""
    d = os.path.join(REF, "grids", "G%d" % glvl)
    out = os.path.join(d, "glvl.dat")
    if os.path.exists(out):
        return out
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "FIMnamelist"), "w") as f:
        for g in _NAMELISTS:
            f.write(" &%s\n /\n" % g)
        f.write(" &CNTLnamelist\n   glvl = %d\n   SubdivNum = 20*2\n   nvl = 64\n /\n" % glvl)
        f.write(" &PREPnamelist\n   curve = 0\n   gtype = 2\n /\n")
    with open(TOP_GRID) as src, open(os.path.join(d, "top_grid"), "w") as dst:
        dst.write(src.read())
    for prog in ("grid", "ginfo", "getlvl"):
        r = subprocess.run([os.path.join(BIN, prog)], cwd=d, capture_output=True, text=True)
        if prog == "getlvl" and not os.path.exists(out):
            raise RuntimeError("reference getlvl failed: %s" % r.stdout[-2000:])
    return out


def ref_file(glvl: int, name: str) -> str:
    ref_glvl(glvl)
    return os.path.join(REF, "grids", "G%d" % glvl, name)


MAGIC = 1179208497


def write_state(path, st, thetac, dpsig, *, nsteps, dt, mode=0, dumpstep=0, init_hybgen=True, flux_scheme=1,
                dtratio=6):
    """The state file dynref reads; class-B tracers (st["trb"], if present)
    follow the class-A tracers as in tr3d(:,:,ntra+1:ntra+ntrb)."""
    nip, nvl = st["dp"].shape
    ntra = st["tr"].shape[0]
    trb = st.get("trb")
    ntrb = 0 if trb is None else trb.shape[0]
    assert ntrb < 256 and 0 < dtratio < 256
    flags = (1 if init_hybgen else 0) | (2 if flux_scheme == 2 else 0) | (ntrb << 8) | (dtratio << 16)
    with open(path, "wb") as f:
        f.write(struct.pack("<8i", MAGIC, nvl, nip, ntra, nsteps, mode, dumpstep, flags))
        f.write(struct.pack("<d", dt))
        for a in (thetac, dpsig, st["us"], st["vs"], st["dp"], st["pr"], st["ex"], st["ph"], st["tr"]):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        if ntrb:
            f.write(np.ascontiguousarray(trb, dtype="<f8").tobytes())


def read_records(path) -> dict:
    """Records of a dynref output file; repeated tags (tracers) become lists."""
    out: dict = {}
    with open(path, "rb") as f:
        buf = f.read()
    off = 0
    while off < len(buf):
        tag = buf[off:off + 16].decode().strip()
        (n,) = struct.unpack_from("<q", buf, off + 16)
        v = np.frombuffer(buf, dtype="<f8", count=n, offset=off + 24).copy()
        off += 24 + 8 * n
        if tag in out:
            if not isinstance(out[tag], list):
                out[tag] = [out[tag]]
            out[tag].append(v)
        else:
            out[tag] = v
    return out


def _big_stack():
    # the reference keeps (nvl,npp,nip) work arrays as automatic (stack) arrays
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_dynref(glvl_path, st, thetac, dpsig, *, nsteps, dt, mode=0, dumpstep=0,
               init_hybgen=True, prec="r8", workdir=None, flux_scheme=1, dtratio=6) -> dict:
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    tag = "%d_%d_%d_%s_%d_%d" % (os.getpid(), nsteps, mode, prec, flux_scheme,
                                 0 if st.get("trb") is None else st["trb"].shape[0])
    fin = os.path.join(workdir, "state_%s.bin" % tag)
    fout = os.path.join(workdir, "out_%s.bin" % tag)
    write_state(fin, st, thetac, dpsig, nsteps=nsteps, dt=dt, mode=mode, dumpstep=dumpstep,
                init_hybgen=init_hybgen, flux_scheme=flux_scheme, dtratio=dtratio)
    r = subprocess.run([os.path.join(BIN, "dynref_" + prec), glvl_path, fin, fout],
                       cwd=workdir, capture_output=True, text=True, preexec_fn=_big_stack)
    if r.returncode != 0:
        raise RuntimeError("dynref failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-3000:]))
    rec = read_records(fout)
    rec["_stdout"] = r.stdout
    os.remove(fin)
    os.remove(fout)
    return rec


def run_restref_write(values: bytes, workdir=None) -> bytes:
    """The restart file the reference write_restart_dyn writes for a value
    stream (oracle/ref/restref.F90)."""
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    fin = os.path.join(workdir, "rvals_%d.bin" % os.getpid())
    fout = os.path.join(workdir, "rfile_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(values)
    r = subprocess.run([os.path.join(BIN, "restref_r4"), "write", fin, fout], cwd=workdir, capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("restref write failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-2000:]))
    with open(fout, "rb") as f:
        data = f.read()
    os.remove(fin)
    os.remove(fout)
    return data


def run_restref_read(restart_bytes: bytes, dims, workdir=None) -> bytes:
    """The value stream the reference read_restart_dyn reads from a restart
    file with sizes dims = (nvl, nip, ntra, ntrb)."""
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    fin = os.path.join(workdir, "rfile_in_%d.bin" % os.getpid())
    fout = os.path.join(workdir, "rvals_out_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(restart_bytes)
    with open(fin + ".dims", "wb") as f:
        f.write(struct.pack("<4i", *dims))
    r = subprocess.run([os.path.join(BIN, "restref_r4"), "read", fin, fout], cwd=workdir, capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("restref read failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-2000:]))
    with open(fout, "rb") as f:
        data = f.read()
    for p in (fin, fin + ".dims", fout):
        os.remove(p)
    return data


HYC_MAGIC = 1213809491


HYCE_MAGIC = 1213809477
HYCE_RECORDS = ("pres", "um_edg", "vm_edg", "un_edg", "vn_edg", "dp_edg", "pr_edg", "geop_edg", "mont_edg",
                "uvsq_edg", "spcv_edg")


def run_hycedgref(glvl_path, case: dict, *, nstep: int = 1, leapm: int = 2, janjic: bool = False,
                  chain: bool = False, thkdff: float = 0.05, biharm: float = 0.75, mass=(0.0, 0.0, 0.0),
                  workdir=None) -> dict:
    """Run the reference bclin_edgvar (oracle/ref/hycedgref.F90) once on the
    inputs of ocean.hycom_edgvar_case, and with chain the reference cnuity
    after it; returns the records by name (each one array)."""
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    kdm = case["kdm"]
    nip = case["wet"].shape[0]
    fin = os.path.join(workdir, "hyce_state_%d.bin" % os.getpid())
    fout = os.path.join(workdir, "hyce_out_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(struct.pack("<8i", HYCE_MAGIC, kdm, nip, nstep, leapm, int(janjic), int(chain), 0))
        f.write(struct.pack("<6d", case["dtleap"], thkdff, biharm, *mass))
        f.write(np.ascontiguousarray(case["wet"], dtype="<i4").tobytes())
        for n in ("utrop", "vtrop", "uclin", "vclin", "dp", "pres", "geop", "montg", "spcifv", "massflx",
                  "mssfx_avg"):
            f.write(np.ascontiguousarray(case[n], dtype="<f8").tobytes())
    r = subprocess.run([os.path.join(BIN, "hycedgref"), glvl_path, fin, fout], cwd=workdir, capture_output=True,
                       text=True, preexec_fn=_big_stack)
    if r.returncode != 0:
        raise RuntimeError("hycedgref failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-3000:]))
    rec = read_records(fout)
    rec["_stdout"] = r.stdout
    os.remove(fin)
    os.remove(fout)
    return rec


HYCT_MAGIC = 1213809490


def run_hyctrref(glvl_path, case: dict, *, leapn: int = 1, workdir=None) -> dict:
    """Run the reference tracer transport (oracle/ref/hyctrref.F90) on the
    inputs of ocean.hycom_transport_case: transp0 on copies (records
    cumuflx0, dpinit0), then transp1 and transp2 (cumuflx1, temp, saln,
    passv)."""
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    kdm = case["kdm"]
    nip = case["wet"].shape[0]
    fin = os.path.join(workdir, "hyct_state_%d.bin" % os.getpid())
    fout = os.path.join(workdir, "hyct_out_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(struct.pack("<8i", HYCT_MAGIC, kdm, nip, leapn, case["numtr"], 0, 0, 0))
        f.write(struct.pack("<2d", case["batrop"], case["temdff"]))
        f.write(np.ascontiguousarray(case["wet"], dtype="<i4").tobytes())
        for n in ("dp", "massflx_tr", "cumuflx", "dpinit", "dpfinl", "temp", "saln", "passv"):
            f.write(np.ascontiguousarray(case[n], dtype="<f8").tobytes())
    r = subprocess.run([os.path.join(BIN, "hyctrref"), glvl_path, fin, fout], cwd=workdir, capture_output=True,
                       text=True, preexec_fn=_big_stack)
    if r.returncode != 0:
        raise RuntimeError("hyctrref failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-3000:]))
    rec = read_records(fout)
    rec["_stdout"] = r.stdout
    os.remove(fin)
    os.remove(fout)
    return rec


HYCM_MAGIC = 1213809485


def run_hycmomref(glvl_path, case: dict, edges: dict, *, nstep: int = 1, leapm: int = 2, janjic: bool = False,
                  workdir=None) -> dict:
    """Run the reference momtum (oracle/ref/hycmomref.F90) once on the inputs
    of ocean.hycom_momtum_case, with the edge values in 'edges' (um_edg,
    vm_edg, geop_edg, mont_edg, uvsq_edg, spcv_edg); returns the records."""
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    kdm = case["kdm"]
    nip = case["wet"].shape[0]
    fin = os.path.join(workdir, "hycm_state_%d.bin" % os.getpid())
    fout = os.path.join(workdir, "hycm_out_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(struct.pack("<8i", HYCM_MAGIC, kdm, nip, nstep, leapm, int(janjic), 0, 0))
        f.write(struct.pack("<4d", case["dtleap"], case["batrop"], case["veldff"], case["smagcf"]))
        f.write(np.ascontiguousarray(case["wet"], dtype="<i4").tobytes())
        for n in ("utrop", "vtrop", "uclin", "vclin"):
            f.write(np.ascontiguousarray(case[n], dtype="<f8").tobytes())
        for n in ("um_edg", "vm_edg", "geop_edg", "mont_edg", "uvsq_edg", "spcv_edg"):
            f.write(np.ascontiguousarray(edges[n], dtype="<f8").tobytes())
        for n in ("taux", "tauy", "dp", "pres", "ustarb"):
            f.write(np.ascontiguousarray(case[n], dtype="<f8").tobytes())
    r = subprocess.run([os.path.join(BIN, "hycmomref"), glvl_path, fin, fout], cwd=workdir, capture_output=True,
                       text=True, preexec_fn=_big_stack)
    if r.returncode != 0:
        raise RuntimeError("hycmomref failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-3000:]))
    rec = read_records(fout)
    rec["_stdout"] = r.stdout
    os.remove(fin)
    os.remove(fout)
    return rec


def run_hycomref(glvl_path, case: dict, *, nsteps: int, thkdff: float = 0.05, biharm: float = 0.75,
                 nstep0: int = 1, mass=(0.0, 0.0, 0.0), workdir=None) -> dict:
    """Run the reference HYCOM continuity equation (oracle/ref/hycomref.F90)
    for nsteps leapfrog steps from step nstep0 on the inputs of
    ocean.hycom_cnuity_case, the mass state (totmass, massglb0, masscor)
    starting from 'mass'; returns the records, each a list with one array per
    step."""
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    kdm = case["kdm"]
    nip = case["wet"].shape[0]
    fin = os.path.join(workdir, "hyc_state_%d.bin" % os.getpid())
    fout = os.path.join(workdir, "hyc_out_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(struct.pack("<8i", HYC_MAGIC, kdm, nip, nsteps, nstep0, 0, 0, 0))
        f.write(struct.pack("<6d", case["dtleap"], thkdff, biharm, *mass))
        f.write(np.ascontiguousarray(case["wet"], dtype="<i4").tobytes())
        for n in ("dp", "pres", "um_edg", "vm_edg", "un_edg", "vn_edg", "dpm_edg", "massflx", "mssfx_avg"):
            f.write(np.ascontiguousarray(case[n], dtype="<f8").tobytes())
    r = subprocess.run([os.path.join(BIN, "hycomref"), glvl_path, fin, fout], cwd=workdir, capture_output=True,
                       text=True, preexec_fn=_big_stack)
    if r.returncode != 0:
        raise RuntimeError("hycomref failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-3000:]))
    rec = read_records(fout)
    rec = {k: (v if isinstance(v, list) else [v]) for k, v in rec.items()}
    rec["_stdout"] = r.stdout
    os.remove(fin)
    os.remove(fout)
    return rec


HYCH_MAGIC = 1213809480


def run_hychybref(case: dict, *, leap: int = 1, with_tracers: bool = False, workdir=None) -> dict:
    """Run the reference hybrid grid generator (oracle/ref/hychybref.F90,
    hycom_hybgn1.F90 hybgn1) once on the inputs of ocean.hycom_hybgen_case;
    returns the records (dens, dp, pres, temp, saln, uvel, vvel, passv)."""
    workdir = workdir or os.path.join(REF, "runs")
    os.makedirs(workdir, exist_ok=True)
    kdm = case["kdm"]
    nip = case["wet"].shape[0]
    numtr = case["passv"].shape[0]
    fin = os.path.join(workdir, "hych_state_%d.bin" % os.getpid())
    fout = os.path.join(workdir, "hych_out_%d.bin" % os.getpid())
    with open(fin, "wb") as f:
        f.write(struct.pack("<8i", HYCH_MAGIC, kdm, nip, leap, numtr, int(with_tracers), 0, 0))
        f.write(struct.pack("<d", case["batrop"]))
        f.write(np.ascontiguousarray(case["wet"], dtype="<i4").tobytes())
        for n in ("targt", "salmin", "dpmin", "latij", "dens", "dp", "pres", "temp", "saln", "uvel", "vvel", "passv"):
            f.write(np.ascontiguousarray(case[n], dtype="<f8").tobytes())
    r = subprocess.run([os.path.join(BIN, "hychybref"), fin, fout], cwd=workdir, capture_output=True, text=True,
                       preexec_fn=_big_stack)
    if r.returncode != 0:
        raise RuntimeError("hychybref failed (%d): %s" % (r.returncode, (r.stdout + r.stderr)[-3000:]))
    rec = read_records(fout)
    rec["_stdout"] = r.stdout
    os.remove(fin)
    os.remove(fout)
    return rec
