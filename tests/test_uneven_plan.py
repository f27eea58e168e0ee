"""Host-side planning of uneven batches (C ABI, no GPU): sizes, work lists and their order."""
import ctypes
import struct

import pytest

from symplectic_stiefel_nearness_amd import _lib

HDR = "4i9i4x7q"          # UPlanHeader: count, gram_items, rec_items, max_mp, gram_begin[9], offsets / sizes


def plan_for(shapes, tiles=None):
    L = _lib.load()
    arr = (_lib.SsnTarget * len(shapes))()
    for i, (n, k) in enumerate(shapes):
        cr, rr = (tiles or {}).get(i, (0, 0))
        arr[i] = _lib.SsnTarget(0x10000, 2 * k, 0x20000, 2 * k, 0x30000, 2 * k, 0x40000, 0x50000, 0, n, 2 * k, cr, rr)
    wsb = L.ssn_uneven_workspace_bytes(arr, len(shapes))
    pb = L.ssn_uneven_plan_bytes(arr, len(shapes))
    plan = (ctypes.c_uint8 * pb)()
    ws = 0x7F0000000000
    rc = L.ssn_uneven_plan(arr, len(shapes), ws, wsb, ctypes.addressof(plan), pb)
    assert rc == 0
    raw = bytes(plan)
    h = struct.unpack_from(HDR, raw, 0)
    count, gi, ri, max_mp = h[:4]
    begin = h[4:13]
    targets_off, gram_off, rec_off, eigh_lds, dual_lds, plan_bytes, ws_bytes = h[13:]
    gram = [struct.unpack_from("2i", raw, gram_off + 8 * j) for j in range(gi)]
    rec = [struct.unpack_from("2i", raw, rec_off + 8 * j) for j in range(ri)]
    return dict(count=count, gi=gi, ri=ri, max_mp=max_mp, begin=begin, gram=gram, rec=rec,
                plan_bytes=plan_bytes, ws_bytes=ws_bytes, wsb=wsb, pb=pb)


@pytest.fixture(scope="module", autouse=True)
def _need_lib():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libssn.so not built")


def test_counts_and_groups():
    shapes = [(1024, 8), (8192, 64), (2048, 48), (300, 3)]
    p = plan_for(shapes)
    assert p["count"] == 4 and p["max_mp"] == 128
    assert p["gi"] == 2 + 16 + 4 + 1                  # 512-row Gram items
    assert p["ri"] == 16 + 128 + 32 + 5               # 64-row recovery items
    assert p["begin"] == (0, 3, 3, 3, 3, 3, 7, 7, 23)  # grouped by order mp = 16 k
    assert p["plan_bytes"] == p["pb"] and p["ws_bytes"] == p["wsb"]


def test_largest_work_first_within_groups():
    shapes = [(1000, 8), (3000, 8), (700, 48), (5000, 48)]
    p = plan_for(shapes)
    cost = lambda t, idx, rows: min(rows, shapes[t][0] - idx) * ((2 * shapes[t][1] + 15) // 16 * 16) ** 2
    b = p["begin"]
    for o in range(1, 9):
        grp = p["gram"][b[o - 1]:b[o]]
        costs = [cost(t, 512 * c, 512) for t, c in grp]
        assert costs == sorted(costs, reverse=True)
    rc = [cost(t, r0, 64) for t, r0 in p["rec"]]
    assert rc == sorted(rc, reverse=True)
    # every (target, chunk) and (target, row block) exactly once
    assert sorted(p["gram"]) == sorted((t, c) for t, (n, k) in enumerate(shapes) for c in range((n + 511) // 512))
    assert sorted(p["rec"]) == sorted((t, r) for t, (n, k) in enumerate(shapes) for r in range(0, n, 64))


def test_tiles_change_the_work_lists():
    shapes = [(4096, 16), (4096, 16)]
    p = plan_for(shapes, tiles={0: (1024, 256), 1: (256, 128)})
    assert p["gi"] == 4 + 16 and p["ri"] == 16 + 32
    assert plan_for(shapes)["wsb"] < p["wsb"]          # more Gram chunks need more partials


def test_rejects_bad_targets():
    L = _lib.load()
    arr = (_lib.SsnTarget * 1)()
    arr[0] = _lib.SsnTarget(0x10000, 6, 0x20000, 6, 0x30000, 6, 0x40000, 0x50000, 0, 10, 3, 0, 0)   # odd m
    assert L.ssn_uneven_workspace_bytes(arr, 1) == 0 and L.ssn_uneven_plan_bytes(arr, 1) == 0
    arr[0] = _lib.SsnTarget(0x10000, 130, 0x20000, 130, 0x30000, 130, 0x40000, 0x50000, 0, 10, 130, 0, 0)
    assert L.ssn_uneven_workspace_bytes(arr, 1) == 0
    assert L.ssn_uneven_workspace_bytes(arr, 0) == 0
