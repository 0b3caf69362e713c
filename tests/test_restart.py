"""The dynamics section of the FIM restart file (restart_dyn.F90) against the
reference writer and reader themselves: a committed file written by the
reference write_restart_dyn (tests/golden/restart_dyn_small.bin, made by
tests/golden/make_restart_golden.py), and, where the reference build is
present, live runs at other sizes in both directions."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import restart_values  # noqa: E402
from fimocn_subx_realtime_hera_apr2020_amd import restart  # noqa: E402

GOLDEN = os.path.join(HERE, "golden", "restart_dyn_small.bin")


def _same(a: restart.RestartDyn, b: restart.RestartDyn):
    assert (a.nvl, a.nip, a.ntr) == (b.nvl, b.nip, b.ntr)
    for k in a.header:
        assert np.array_equal(a.header[k], b.header[k]), k
    for k in a.fields:
        assert np.array_equal(a.fields[k], b.fields[k]), k


def test_reader_and_writer_match_reference_file(tmp_path):
    """The reference-written file reads back to exactly the seeded contents,
    and writing those contents reproduces the file byte for byte."""
    want = restart_values.make(4, 12, 2, 1)
    got = restart.read_restart_dyn(GOLDEN, 3)
    _same(want, got)
    assert got.trailing == b""
    out = tmp_path / "r.bin"
    restart.write_restart_dyn(str(out), want)
    assert out.read_bytes() == open(GOLDEN, "rb").read()


@pytest.mark.parametrize("nvl, nip, ntra, ntrb", [(8, 42, 4, 2), (64, 162, 4, 0)])
def test_against_live_reference_writer_and_reader(tmp_path, nvl, nip, ntra, ntrb):
    """Both directions against the reference programs: the file this module
    writes is the reference's byte for byte, and the reference reader reads it
    back to the same values."""
    from oracle import refrun
    if not os.path.exists(os.path.join(refrun.BIN, "restref_r4")):
        pytest.skip("reference build (oracle/_ref) not present")
    rd = restart_values.make(nvl, nip, ntra, ntrb, seed=nvl + nip)
    stream = restart_values.value_stream(rd, ntrb)
    ref_bytes = refrun.run_restref_write(stream)
    out = tmp_path / "r.bin"
    restart.write_restart_dyn(str(out), rd)
    assert out.read_bytes() == ref_bytes
    _same(rd, restart.read_restart_dyn(str(out), ntra + ntrb))
    assert refrun.run_restref_read(out.read_bytes(), (nvl, nip, ntra, ntrb)) == stream


def test_following_sections_kept_and_byte_order(tmp_path):
    """The bytes after the dynamics section (coupler, physics, ocean records)
    come back unchanged; a big-endian file reads as the same contents."""
    rd = restart.read_restart_dyn(GOLDEN, 3)
    tail = b"\x10\x00\x00\x00" + bytes(range(16)) + b"\x10\x00\x00\x00"
    src = tmp_path / "full.bin"
    src.write_bytes(open(GOLDEN, "rb").read() + tail)
    full = restart.read_restart_dyn(str(src), 3)
    assert full.trailing == tail
    out = tmp_path / "again.bin"
    restart.write_restart_dyn(str(out), full)
    assert out.read_bytes() == src.read_bytes()
    rd.endian = ">"
    big = tmp_path / "big.bin"
    restart.write_restart_dyn(str(big), rd)
    back = restart.read_restart_dyn(str(big), 3)
    assert back.endian == ">"
    _same(restart.read_restart_dyn(GOLDEN, 3), back)


def test_malformed_files_are_rejected(tmp_path):
    data = open(GOLDEN, "rb").read()
    cut = tmp_path / "cut.bin"
    cut.write_bytes(data[:len(data) // 2])
    with pytest.raises(ValueError):
        restart.read_restart_dyn(str(cut), 3)
    with pytest.raises(ValueError):          # more tracers than the file holds
        restart.read_restart_dyn(GOLDEN, 4)
    junk = tmp_path / "junk.bin"
    junk.write_bytes(b"\x00" * 64)
    with pytest.raises(ValueError):
        restart.read_restart_dyn(str(junk), 3)
