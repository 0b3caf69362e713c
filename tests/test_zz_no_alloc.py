"""fimhip_globsum_noalloc: the global mass and tracer totals of caller-owned
device arrays, computed with no device allocation (block partials in the
library's file-scope device array), against fimhip_globsum on the same state,
bit for bit, on two calls in a row, the second with a smaller input."""
import ctypes as ct

import numpy as np
import pytest

from conftest import gpu_available
from fimocn_subx_realtime_hera_apr2020_amd import fimhip
from helpers import seeded_case

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _case(level, ntra, steps):
    """A seeded state stepped on the device, its fimhip_globsum totals, and
    the same state's area, thickness and tracers as device tensors."""
    import torch
    g, st, th, ds = seeded_case(level, ntra=ntra)
    m = fimhip.Model(g, 64, ntra, 240.0, th, ds, device=0)
    try:
        m.load(st)
        m.init_state(True)
        for its in range(1, steps + 1):
            m.step(its, 0)
        m.sync()
        mass, trt = m.globsum()
        dp, tr = m.download("dp"), m.download("tr")
    finally:
        m.close()
    dev = torch.device("cuda", 0)
    area = torch.tensor(np.asarray(g.area, dtype=np.float64), device=dev)
    return g.nip, mass, np.asarray(trt), area, torch.tensor(dp, device=dev), torch.tensor(tr, device=dev)


def test_globsum_noalloc_equals_globsum_twice_second_smaller():
    import torch
    lib = fimhip.load_library()
    stream = torch.cuda.current_stream(0)
    sizes = []
    for level, ntra, steps in ((3, 4, 3), (2, 2, 1)):
        nip, mass, trt, area, dp, tr = _case(level, ntra, steps)
        out = torch.full((1 + ntra,), float("nan"), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        print("globsum_noalloc: G%d nip %d ntra %d" % (level, nip, ntra), flush=True)
        rc = lib.fimhip_globsum_noalloc(0, nip, 64, ntra, area.data_ptr(), dp.data_ptr(), tr.data_ptr(),
                                        out.data_ptr(), ct.c_void_p(stream.cuda_stream))
        assert rc == 0
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert got[0] == mass, (got[0], mass)
        assert np.array_equal(got[1:], trt), (got[1:], trt)
        sizes.append(nip * ntra)
    assert sizes[1] < sizes[0]


def test_globsum_noalloc_rejects_sizes_beyond_its_scratch():
    import torch
    lib = fimhip.load_library()
    t = torch.zeros(64 * 42, dtype=torch.float64, device="cuda:0")
    o = torch.zeros(64, dtype=torch.float64, device="cuda:0")
    p = t.data_ptr()
    assert lib.fimhip_globsum_noalloc(0, 42, 64, fimhip.GLOBSUM_NOALLOC_MAX_Q, p, p, p, o.data_ptr(), None) == -1
    assert lib.fimhip_globsum_noalloc(0, 42, 65, 1, p, p, p, o.data_ptr(), None) == -1
    assert lib.fimhip_globsum_noalloc(0, 42, 64, 1, p, p, None, o.data_ptr(), None) == -1
    assert lib.fimhip_globsum_noalloc(0, 42, 64, 0, p, p, None, o.data_ptr(), None) == 0
    torch.cuda.synchronize()
