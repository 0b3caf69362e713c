"""The dynamics section of a FIM restart file (restart_dyn.F90:17-169).

A FIM restart file is Fortran unformatted sequential: every record is framed
by 4-byte length markers.  The dynamics section written by write_restart_dyn
comes first: six header records of mixed items, then one record per 2-D
slice of each array (readarr32_r / writearr32_r, real*4), in this order
(restart_dyn.F90:117-163):

    per AB slot n:        u_tdcy, v_tdcy, dp_tdcy, dpl_tdcy        (nvl, nip)
    per tracer t (class A, then class B):
                          tr3d, trdp, then per slot trc_tdcy, trl_tdcy
    us3d vs3d ws3d mp3d tk3d dp3d rh3d relvor dpinit               (nvl, nip)
    pr3d ex3d ph3d sdot                                            (nvl+1, nip)
    per slot massfx, then cumufx                                   (nvl*6, nip)

The coupler, physics and ocean sections follow; they are kept as opaque
bytes so a file can be read, updated and written back whole.

Arrays are returned in this library's layout (level fastest, the reference's
Fortran order reversed): cell fields (nip, nvl), AB histories (3, nip, nvl),
tracers (ntr, nip, nvl), tracer histories (ntr, 3, nip, nvl), mass fluxes
(3, nip, 6, nvl), all float32 as stored.  Records larger than 2 GiB
(sub-records) are not supported.
"""
from __future__ import annotations

import struct

import numpy as np

NPP, NABL = 6, 3


def _header_spec(nvl: int, nip: int):
    """(record, [(name, dtype, count)]) of the six header records
    (restart_dyn.F90:117-122); i4 integer, f4 real, l4 logical."""
    c = nip
    return [
        [("curr_write_time", "i4", 1), ("nf", "i4", 1), ("of", "i4", 1), ("vof", "i4", 1),
         ("adbash1", "f4", 1), ("adbash2", "f4", 1), ("adbash3", "f4", 1), ("thetac", "f4", nvl),
         ("dpsig", "f4", nvl), ("sigak", "f4", nvl + 1), ("sigbk", "f4", nvl + 1)],
        [("lat", "f4", c), ("lon", "f4", c), ("nprox", "i4", c), ("proxs", "i4", NPP * c), ("area", "f4", c),
         ("cs", "f4", 4 * NPP * c), ("sn", "f4", 4 * NPP * c), ("psrf", "f4", c), ("ptdcy", "f4", 2 * c)],
        [("qmstr", "f4", 1), ("qmstrc", "f4", 1), ("qmstrn", "f4", 1), ("qdtr_set", "l4", 1)],
        [("sidevec_c", "f4", 2 * NPP * c), ("sidevec_e", "f4", 2 * NPP * c), ("sideln", "f4", NPP * c),
         ("rprox_ln", "f4", NPP * c)],
        [("sst_prev", "f4", c), ("sst_next", "f4", c), ("fice2d_prev", "f4", c), ("fice2d_next", "f4", c),
         ("fice2d", "f4", c), ("next_date", "i4", 4), ("have_next_sst", "l4", 1)],
        [("rarea", "f4", c), ("rsideln", "f4", NPP * c), ("corio", "f4", c), ("deg_lat", "f4", c),
         ("deg_lon", "f4", c), ("rn2d", "f4", c), ("pw2d", "f4", c), ("rc2d", "f4", c), ("u10m", "f4", c),
         ("v10m", "f4", c), ("runoff", "f4", c)],
    ]


def _array_spec(nvl: int, nip: int, ntr: int):
    """[(name, index, shape of one record in this library's layout)] of the
    array records in file order (restart_dyn.F90:125-163)."""
    c, l, e = (nip, nvl), (nip, nvl + 1), (nip, NPP, nvl)
    out = []
    for n in range(NABL):
        for f in ("u_tdcy", "v_tdcy", "dp_tdcy", "dpl_tdcy"):
            out.append((f, (n,), c))
    for t in range(ntr):
        out.append(("tr", (t,), c))
        out.append(("trdp", (t,), c))
        for n in range(NABL):
            out.append(("trc_tdcy", (t, n), c))
            out.append(("trl_tdcy", (t, n), c))
    for f in ("us", "vs", "ws", "mp", "tk", "dp", "rh", "relvor", "dpinit"):
        out.append((f, (), c))
    for f in ("pr", "ex", "ph", "sdot"):
        out.append((f, (), l))
    for n in range(NABL):
        out.append(("massfx", (n,), e))
    out.append(("cumufx", (), e))
    return out


def _field_shapes(nvl: int, nip: int, ntr: int) -> dict:
    c, l, e = (nip, nvl), (nip, nvl + 1), (nip, NPP, nvl)
    shp = {f: (NABL,) + c for f in ("u_tdcy", "v_tdcy", "dp_tdcy", "dpl_tdcy")}
    shp.update(tr=(ntr,) + c, trdp=(ntr,) + c, trc_tdcy=(ntr, NABL) + c, trl_tdcy=(ntr, NABL) + c,
               massfx=(NABL,) + e, cumufx=e)
    shp.update({f: c for f in ("us", "vs", "ws", "mp", "tk", "dp", "rh", "relvor", "dpinit")})
    shp.update({f: l for f in ("pr", "ex", "ph", "sdot")})
    return shp


class RestartDyn:
    """The dynamics section of a restart file: header items (numpy arrays of
    their stored type, scalars as 1-element arrays), the arrays (float32) and
    the bytes of the sections that follow."""

    def __init__(self, nvl: int, nip: int, ntr: int, header: dict, fields: dict, trailing: bytes = b"",
                 endian: str = "<"):
        self.nvl, self.nip, self.ntr = nvl, nip, ntr
        self.header, self.fields, self.trailing, self.endian = header, fields, trailing, endian

    @classmethod
    def empty(cls, nvl: int, nip: int, ntr: int) -> "RestartDyn":
        header = {}
        for rec in _header_spec(nvl, nip):
            for name, dt, n in rec:
                header[name] = np.zeros(n, dtype=np.dtype(dt.replace("l", "i")))
        fields = {f: np.zeros(s, np.float32) for f, s in _field_shapes(nvl, nip, ntr).items()}
        return cls(nvl, nip, ntr, header, fields)

    @property
    def ab(self):
        """(nf, of, vof, adbash1, adbash2, adbash3), slots 1-based"""
        h = self.header
        return (int(h["nf"][0]), int(h["of"][0]), int(h["vof"][0]), float(h["adbash1"][0]),
                float(h["adbash2"][0]), float(h["adbash3"][0]))


def _records(buf: bytes, endian: str):
    off = 0
    while off < len(buf):
        (n,) = struct.unpack_from(endian + "i", buf, off)
        if n < 0 or off + 8 + n > len(buf):
            raise ValueError("restart: bad or unsupported record marker %d at byte %d" % (n, off))
        (m,) = struct.unpack_from(endian + "i", buf, off + 4 + n)
        if m != n:
            raise ValueError("restart: record markers %d / %d disagree at byte %d" % (n, m, off))
        yield off, buf[off + 4:off + 4 + n]
        off += 8 + n


def read_restart_dyn(path: str, ntr: int) -> RestartDyn:
    """Read the dynamics section of a restart file; ntr = ntra + ntrb tracers
    (the count is not stored in the file).  nvl and nip follow from the sizes
    of the first two records; the byte order from the first marker."""
    with open(path, "rb") as f:
        buf = f.read()
    endian = None
    for e in ("<", ">"):
        (n,) = struct.unpack_from(e + "i", buf, 0)
        if 0 < n and (n // 4 - 9) % 4 == 0 and 8 + n <= len(buf):
            (m,) = struct.unpack_from(e + "i", buf, 4 + n)
            if m == n:
                endian = e
                break
    if endian is None:
        raise ValueError("restart: %s does not start with a dynamics header record" % path)
    recs = _records(buf, endian)
    _, r1 = next(recs)
    nvl = (len(r1) // 4 - 9) // 4
    _, r2 = next(recs)
    if len(r2) % (4 * 61):
        raise ValueError("restart: second record of %d bytes is not 244 * nip" % len(r2))
    nip = len(r2) // (4 * 61)
    spec = _header_spec(nvl, nip)
    header = {}
    raw = [r1, r2] + [next(recs)[1] for _ in range(4)]
    for rec, data in zip(spec, raw):
        want = sum(4 * n for _, _, n in rec)
        if len(data) != want:
            raise ValueError("restart: header record of %d bytes, expected %d" % (len(data), want))
        off = 0
        for name, dt, n in rec:
            header[name] = np.frombuffer(data, dtype=np.dtype(endian + dt.replace("l", "i")), count=n,
                                         offset=off).astype(dt.replace("l", "i"))
            off += 4 * n
    fields = {f: np.zeros(s, np.float32) for f, s in _field_shapes(nvl, nip, ntr).items()}
    end = 0
    for name, idx, shp in _array_spec(nvl, nip, ntr):
        try:
            off, data = next(recs)
        except StopIteration:
            raise ValueError("restart: file ends inside the dynamics section (ntr=%d?)" % ntr) from None
        if len(data) != 4 * int(np.prod(shp)):
            raise ValueError("restart: record of %s has %d bytes, expected %d" % (name, len(data), 4 * np.prod(shp)))
        fields[name][idx] = np.frombuffer(data, dtype=endian + "f4").reshape(shp)
        end = off + 8 + len(data)
    return RestartDyn(nvl, nip, ntr, header, fields, buf[end:], endian)


def write_restart_dyn(path: str, rd: RestartDyn) -> None:
    """Write the dynamics section (as write_restart_dyn does) followed by the
    bytes of the sections that came after it when the file was read."""
    e = rd.endian
    out = bytearray()

    def rec(b: bytes):
        if len(b) > 0x7FFFFFF7:
            raise ValueError("restart: records over 2 GiB (sub-records) are not supported")
        out.extend(struct.pack(e + "i", len(b)))
        out.extend(b)
        out.extend(struct.pack(e + "i", len(b)))

    for spec in _header_spec(rd.nvl, rd.nip):
        b = bytearray()
        for name, dt, n in spec:
            a = np.asarray(rd.header[name]).reshape(-1)
            if a.size != n:
                raise ValueError("restart: header item %s has %d values, expected %d" % (name, a.size, n))
            b.extend(a.astype(e + dt.replace("l", "i")).tobytes())
        rec(bytes(b))
    for name, idx, shp in _array_spec(rd.nvl, rd.nip, rd.ntr):
        rec(np.ascontiguousarray(rd.fields[name][idx], dtype=e + "f4").tobytes())
    out.extend(rd.trailing)
    with open(path, "wb") as f:
        f.write(bytes(out))


# ---- exchange with a device-resident instance --------------------------------
_DEVICE_FIELDS = ("us", "vs", "dp", "pr", "ex", "ph", "mp", "ws", "relvor", "sdot", "u_tdcy", "v_tdcy",
                  "dp_tdcy", "dpl_tdcy", "massfx")


def load_into(model, rd: RestartDyn) -> None:
    """Upload the restart's dynamics state into a fimhip.Model (widened to
    FP64): prognostic fields, tendencies, mass fluxes, the AB slots and
    weights, and, with class-B tracers set up, their fields, the integrated
    mass flux and dpinit."""
    ntra, ntrb = model.ntra, model.ntrb
    if rd.ntr != ntra + ntrb or rd.nip != model.nip or rd.nvl != model.nvl:
        raise ValueError("restart sizes (nvl %d, nip %d, ntr %d) do not match the instance (%d, %d, %d + %d)"
                         % (rd.nvl, rd.nip, rd.ntr, model.nvl, model.nip, ntra, ntrb))
    f = rd.fields
    for n in _DEVICE_FIELDS:
        model.upload(n, f[n].astype(np.float64))
    model.upload("tr", f["tr"][:ntra].astype(np.float64))
    model.upload("trdp", f["trdp"][:ntra].astype(np.float64))
    model.upload("trc_tdcy", f["trc_tdcy"][:ntra].astype(np.float64))
    model.upload("trl_tdcy", f["trl_tdcy"][:ntra].astype(np.float64))
    model.upload("psrf", rd.header["psrf"].astype(np.float64))
    model.upload("ptdcy", rd.header["ptdcy"].astype(np.float64).reshape(2, rd.nip))
    if ntrb:
        model.upload("trb", f["tr"][ntra:].astype(np.float64))
        model.upload("cumufx", f["cumufx"].astype(np.float64))
        model.upload("dpinit", f["dpinit"].astype(np.float64))
    model.set_ab(*rd.ab)


def update_from(model, rd: RestartDyn) -> RestartDyn:
    """Copy the device state of a fimhip.Model into rd (rounded to real*4, as
    the reference stores it), with its AB slots and weights; items the
    library does not hold (tk3d, rh3d, the physics and surface fields) keep
    rd's values."""
    ntra, ntrb = model.ntra, model.ntrb
    f = rd.fields
    for n in _DEVICE_FIELDS:
        f[n][...] = model.download(n)
    f["tr"][:ntra] = model.download("tr")
    f["trdp"][:ntra] = model.download("trdp")
    f["trc_tdcy"][:ntra] = model.download("trc_tdcy")
    f["trl_tdcy"][:ntra] = model.download("trl_tdcy")
    rd.header["psrf"][...] = model.download("psrf").astype(np.float32)
    rd.header["ptdcy"][...] = model.download("ptdcy").reshape(-1).astype(np.float32)
    if ntrb:
        f["tr"][ntra:] = model.download("trb")
        f["cumufx"][...] = model.download("cumufx")
        f["dpinit"][...] = model.download("dpinit")
    nf, of, vof, a1, a2, a3 = model.get_ab()
    h = rd.header
    h["nf"][0], h["of"][0], h["vof"][0] = nf, of, vof
    h["adbash1"][0], h["adbash2"][0], h["adbash3"][0] = a1, a2, a3
    return rd
