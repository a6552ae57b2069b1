"""CPU checks of the batched-ensemble work list: the tile walk of the persistent kernel (restated here from
lbm_ensemble.hip) visits every tile of every member exactly once for any grid, every cell it touches lies
inside its member, and the 64-member workload is reproducible."""
import numpy as np
import pytest

from lbm import configs

TILE_X, TILE_Y = 128, 2


def tile_list(sizes):
    first, t = [], 0
    for W, H in sizes:
        first.append(t)
        t += -(-W // TILE_X) * -(-H // TILE_Y)
    return first, t


def walk(sizes, blocks, m0=0, m1=None):
    """(member, tile-x, tile-y) visited by each workgroup of k_ensemble over members [m0, m1), in order."""
    blocks = max(8, -(-blocks // 8) * 8)
    first, T = tile_list(sizes)
    m1 = len(sizes) if m1 is None else m1
    t0, t1 = first[m0], first[m1] if m1 < len(sizes) else T
    gper = blocks // 8
    per = -(-(t1 - t0) // 8)
    seen = []
    for b in range(blocks):
        xcd, l = b % 8, b // 8
        start = t0 + xcd * per
        end = min(start + per, t1)
        for t in range(start + l, end, gper):
            lo, hi = m0, m1 - 1  # the kernel's binary search
            while lo < hi:
                mid = (lo + hi + 1) >> 1
                if first[mid] <= t:
                    lo = mid
                else:
                    hi = mid - 1
            assert lo == int(np.searchsorted(first, t, side="right")) - 1
            lt = t - first[lo]
            tx_n = -(-sizes[lo][0] // TILE_X)
            seen.append((lo, lt % tx_n, lt // tx_n))
    return seen, t1 - t0


def groups(sizes, group_bytes, elem=4):
    """Member ranges of lbm_ensemble_step_n's cache grouping."""
    out, m0, n = [], 0, len(sizes)
    while m0 < n:
        m1, b = m0 + 1, 9 * elem * sizes[m0][0] * sizes[m0][1]
        if group_bytes > 0:
            while m1 < n and b + 9 * elem * sizes[m1][0] * sizes[m1][1] <= group_bytes:
                b += 9 * elem * sizes[m1][0] * sizes[m1][1]
                m1 += 1
        else:
            m1 = n
        out.append((m0, m1))
        m0 = m1
    return out


@pytest.mark.parametrize("blocks", [1, 8, 24, 2048, 10000])
@pytest.mark.parametrize("sizes", [[(64, 32), (70, 33), (129, 5), (256, 130), (3, 3)], [(2048, 2048)],
                                   [(256, 256)] * 7, [(5, 3)]])
def test_every_tile_once(sizes, blocks):
    seen, T = walk(sizes, blocks)
    assert len(seen) == T and len(set(seen)) == T
    for i, tx, ty in seen:
        W, H = sizes[i]
        assert 0 <= tx * TILE_X < W and 0 <= ty * TILE_Y < H


@pytest.mark.parametrize("group_kb", [0, 1, 300, 2000])
def test_grouped_walk_covers_each_member_once(group_kb):
    sizes = [(64, 32), (70, 33), (129, 5), (256, 130), (384, 256), (128, 512), (3, 3), (200, 64), (512, 96)]
    gs = groups(sizes, group_kb << 10)
    assert gs[0][0] == 0 and gs[-1][1] == len(sizes)
    assert all(a[1] == b[0] for a, b in zip(gs, gs[1:]))
    if group_kb == 1:
        assert len(gs) == len(sizes)
    seen = []
    for m0, m1 in gs:
        part, n = walk(sizes, 24, m0, m1)
        assert len(part) == n and all(m0 <= i < m1 for i, _, _ in part)
        seen += part
    _, T = tile_list(sizes)
    assert len(seen) == T and len(set(seen)) == T


def test_xcd_bands_are_contiguous():
    """Each XCD's workgroups take one contiguous eighth of the list."""
    sizes = [(512, 512), (256, 768), (1024, 256)]
    first, T = tile_list(sizes)
    per = -(-T // 8)
    blocks = 64
    for xcd in range(8):
        ts = []
        for b in range(xcd, blocks, 8):
            ts.extend(range(xcd * per + b // 8, min(xcd * per + per, T), blocks // 8))
        assert sorted(ts) == list(range(xcd * per, min(xcd * per + per, T)))


def test_ensemble_workload_members():
    m1 = configs.ensemble()
    m2 = configs.ensemble()
    assert len(m1) == 64
    assert [(m.W, m.H, m.nu) for m in m1] == [(m.W, m.H, m.nu) for m in m2]
    assert all(m.W in configs.ENSEMBLE_SIZES and m.H in configs.ENSEMBLE_SIZES for m in m1)
    assert all(1e-4 <= m.nu <= 1e-2 for m in m1)
    cells = sum(m.W * m.H for m in m1)
    assert 20e6 < cells < 140e6  # expected 64 * 933^2 = 55.7 M
