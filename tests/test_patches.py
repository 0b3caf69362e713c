"""The patch partition behind the fused tracer transport of the batched
form (host tables, no GPU): every cell is in exactly one patch's core; a
patch's rings 1-3 are exactly the cells at neighbour distance 1, 2 and 3 from
its core; the local neighbour table of rings 0-2 points at the right cells
inside the tile (with each cell's edge count and its neighbours' edge
slots); and the sizes stay within the kernel's LDS tile (local indices fit
in 8 bits)."""
import numpy as np
import pytest

from fimocn_subx_realtime_hera_apr2020_amd import fimhip, glvl
from helpers import golden_grid

# the kernel's LDS tile (ens_internal.h): core, rings 0-1, 0-2, 0-3
CORE_MAX, R_MAX, LO_MAX, TILE_MAX = 64, 112, 160, 208
assert TILE_MAX <= 256


def _check(grid):
    cnt, ent = fimhip.ens_patches(grid)
    prox = np.asarray(grid.prox, np.int64) - 1
    proxs = np.asarray(grid.proxs, np.int64) - 1
    nprox = np.asarray(grid.nprox)
    for c in np.where(nprox == 5)[0]:
        prox[c, 5] = prox[c, 4]
        proxs[c, 5] = proxs[c, 4]
    nip = grid.nip
    seen = np.zeros(nip, np.int64)
    off = 0
    for n0, n1, n2, n3 in cnt:
        assert 1 <= n0 <= CORE_MAX and n1 <= R_MAX and n2 <= LO_MAX and n3 <= TILE_MAX
        assert n0 <= n1 <= n2 <= n3
        cells = ent[off:off + n3, 0].astype(np.int64)
        assert cells.min() >= 0 and cells.max() < nip and len(set(cells.tolist())) == n3
        seen[cells[:n0]] += 1
        # rings by breadth-first distance from the core
        dist = {int(c): 0 for c in cells[:n0]}
        front = list(dist)
        for r in (1, 2, 3):
            nxt = sorted({int(x) for c in front for x in prox[c] if int(x) not in dist})
            for x in nxt:
                dist[x] = r
            front = nxt
        ends = (0, n0, n1, n2, n3)
        for r in (1, 2, 3):
            ring = cells[ends[r]:ends[r + 1]].tolist()
            assert ring == sorted(x for x, d in dist.items() if d == r), (r, ring[:5])
        # local neighbour table of rings 0-2
        for i in range(n2):
            e4 = ent[off + i].astype(np.int64) & 0xFFFFFFFF
            loc = [(e4[1] >> 8 * e) & 0xFF for e in range(4)] + [(e4[2] >> 8 * e) & 0xFF for e in range(2)]
            assert (e4[2] >> 16) & 0xFF == nprox[cells[i]]
            assert [(e4[3] >> 4 * e) & 0xF for e in range(6)] == [int(x) for x in proxs[cells[i]]]
            assert max(loc) < n3
            assert [int(cells[l]) for l in loc] == [int(x) for x in prox[cells[i]]]
            if i < n1:
                assert max(loc) < n2          # ring-1 neighbours lie in rings 0-2
            if i < n0:
                assert max(loc) < n1          # core neighbours lie in rings 0-1
        off += n3
    assert off == len(ent)
    assert np.all(seen == 1)
    return cnt


@pytest.mark.parametrize("level", [1, 2, 3])
def test_patches_of_the_golden_grids(level):
    _check(golden_grid(level))


def test_patches_of_g6_and_a_concatenated_grid():
    g6 = fimhip.icos_grid(6)
    cnt = _check(g6)
    # compact patches: the three rings hold about twice the core's cells
    assert cnt[:, 3].sum() / cnt[:, 0].sum() < 3.2
    assert cnt[:, 0].mean() > 60
    _check(glvl.concat([golden_grid(2), golden_grid(3)]))
