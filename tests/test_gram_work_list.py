"""The FP64 Schur Gram work list, built on the host (no GPU needed).

The kernel remaps block b to item lin = x*q + min(x, r) + b // 8 on XCD x = b % 8
(q, r = divmod(items, 8)), and reads (split, tile) of item lin from this list.
Each XCD's run must hold exactly the items of the plain order (split-major,
8 x 8 super-blocks) for the same positions, each once, with the full 128-row
tiles before the diagonal tiles, the strips of a short last tile row and the
items that exit.
"""

import ctypes

import numpy as np
import pytest

from sdpx_mi355x import _lib


def work_list(m, group, split):
    nt = -(-m // 128)
    n = nt * (nt + 1) // 2 * split
    out = np.zeros(2 * n, dtype=np.int32)
    assert _lib.load().sdpx_gram_work_list(m, group, split, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
    return out.reshape(n, 2)


def plain_order(nt, group):
    out = []
    for si in range(0, nt, group):
        for sj in range(0, si + 1, group):
            for ti in range(si, min(nt, si + group)):
                for tj in range(sj, min(sj + group, ti + 1)):
                    out.append((ti, tj))
    return out


def length_class(m, ti, tj):
    rows = m - 128 * ti
    if rows <= 32:
        return 3 if tj % 2 else 2
    if rows <= 64:
        return 2
    return 1 if ti == tj and rows >= 128 else 0


@pytest.mark.parametrize("m", [20, 64, 160, 276, 300, 1000, 4000, 6000])
@pytest.mark.parametrize("split", [1, 3, 16, 31, 32])
def test_work_list_is_a_per_xcd_permutation_with_full_tiles_first(m, split):
    nt = -(-m // 128)
    per = nt * (nt + 1) // 2
    order = plain_order(nt, 8)
    items = work_list(m, 8, split)
    total = per * split
    q, r = divmod(total, 8)
    seen = set()
    for x in range(8):
        lo, n = x * q + min(x, r), q + (x < r)
        got = [(int(s), int(y) >> 16, int(y) & 0xFFFF) for s, y in items[lo:lo + n]]
        want = [(lin // per, *order[lin % per]) for lin in range(lo, lo + n)]
        assert sorted(got) == sorted(want)
        classes = [length_class(m, ti, tj) for _, ti, tj in got]
        assert classes == sorted(classes)
        seen.update(got)
    assert len(seen) == total


def test_work_list_rejects_bad_arguments():
    out = np.zeros(2, dtype=np.int32)
    p = out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    assert _lib.load().sdpx_gram_work_list(0, 8, 1, p) == -1
    assert _lib.load().sdpx_gram_work_list(100, 8, 0, p) == -1
