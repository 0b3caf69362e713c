"""Seeded contents of the dynamics section of a FIM restart file, and the raw
value stream oracle/ref/restref.F90 reads and writes (test infrastructure)."""
import struct

import numpy as np

from fimocn_subx_realtime_hera_apr2020_amd import restart


def make(nvl: int, nip: int, ntra: int, ntrb: int, seed: int = 11) -> restart.RestartDyn:
    """A RestartDyn with every item filled from a seeded generator: reals as
    float32 of several magnitudes, integers and logicals in their ranges."""
    rng = np.random.default_rng(seed)
    rd = restart.RestartDyn.empty(nvl, nip, ntra + ntrb)
    for name, a in rd.header.items():
        if a.dtype.kind == "f":
            a[...] = (rng.standard_normal(a.size) * 10.0 ** rng.integers(-3, 5, a.size)).astype(np.float32)
    h = rd.header
    h["curr_write_time"][0] = 7
    h["nf"][0], h["of"][0], h["vof"][0] = 2, 1, 3
    h["nprox"][...] = rng.choice([5, 6], nip)
    h["proxs"][...] = rng.integers(1, 7, h["proxs"].size)
    h["qdtr_set"][0] = 1
    h["next_date"][...] = [2020, 4, 15, 12]
    h["have_next_sst"][0] = 0
    for name, a in rd.fields.items():
        a[...] = (rng.standard_normal(a.shape) * 10.0 ** rng.integers(-6, 6)).astype(np.float32)
    return rd


def value_stream(rd: restart.RestartDyn, ntrb: int) -> bytes:
    """dims, then the header items in record order, then the whole arrays in
    the order restref.F90 reads them (little endian)."""
    out = bytearray(struct.pack("<4i", rd.nvl, rd.nip, rd.ntr - ntrb, ntrb))
    for rec in restart._header_spec(rd.nvl, rd.nip):
        for name, dt, n in rec:
            out.extend(np.asarray(rd.header[name], dtype="<" + dt.replace("l", "i")).tobytes())
    for f in ("u_tdcy", "v_tdcy", "dp_tdcy", "dpl_tdcy", "tr", "trdp", "trc_tdcy", "trl_tdcy", "us", "vs", "ws",
              "mp", "tk", "dp", "rh", "relvor", "dpinit", "pr", "ex", "ph", "sdot", "massfx", "cumufx"):
        out.extend(np.ascontiguousarray(rd.fields[f], dtype="<f4").tobytes())
    return bytes(out)
