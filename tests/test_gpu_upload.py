"""Staged uploads: a plan's upload gathers its device copies in pinned host
memory and the next run issues them on its own stream.  Every way of reaching
the device after an upload must see that batch: a second upload before any
run, a form switch (synchronous copies after the staged ones), the autotuner,
a run on a user stream, and a batch that grows the pinned buffer."""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

from genie_amd import Plan, Table, match_topk, workloads as W

pytestmark = pytest.mark.gpu


def cfg(n, d, q, seed, radius=2):
    w = W.cfg1(seed=seed, n=n, d=d, q=q, k=50, radius=radius)
    return w["rows"], w["ranges"]


def expect(rows, ranges, nq, k):
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, ranges[q_of == q]), nq, k)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_second_upload_before_run_wins(mode):
    rows, ra = cfg(60_000, 24, 40, 901)
    _, rb = cfg(60_000, 24, 40, 902)
    t = Table.from_rows(rows)
    ia, ib = t.compile(ra, 40), t.compile(rb, 40)
    p = Plan(t.device(), 40, max(len(ia), len(ib)), 50, 0, mode=mode)
    p.upload(ia, 40, np.full(40, 50, np.uint32))
    p.upload(ib, 40, np.full(40, 50, np.uint32))
    p.run()
    got = p.download()
    ri, rc = expect(rows, rb, 40, 50)
    assert_topk_equal(got[0], got[1], ri, rc, 50, 40, "second upload")


def test_form_switch_and_autotune_after_upload():
    rows, r = cfg(80_000, 32, 33, 911)
    t = Table.from_rows(rows)
    items = t.compile(r, 33)
    ri, rc = expect(rows, r, 33, 50)
    p = Plan(t.device(), 33, len(items), 50, 0, mode=1)
    p.upload(items, 33, np.full(33, 50, np.uint32))
    p.set_dense(False)  # settles the lane form with synchronous copies after the staged ones
    assert not p.dense_active()
    p.run()
    got = p.download()
    assert_topk_equal(got[0], got[1], ri, rc, 50, 33, "lane form after switch")
    p.set_dense(True)
    assert p.dense_active()
    p.run()
    got = p.download()
    assert_topk_equal(got[0], got[1], ri, rc, 50, 33, "dense form after switch")
    # a fresh upload straight into the autotuner
    p.upload(items, 33, np.full(33, 50, np.uint32))
    p.autotune(reps=1)
    p.run()
    got = p.download()
    assert_topk_equal(got[0], got[1], ri, rc, 50, 33, "after autotune")


def test_user_stream_growing_batches_and_graph():
    """Batches of growing size on a user stream (the pinned buffer grows),
    each checked; then graph replays of the last with eager work between."""
    rows, _ = cfg(100_000, 48, 8, 921)
    t = Table.from_rows(rows)
    # a caller's stream, from the HIP runtime the library itself uses
    hip = C.CDLL("libamdhip64.so.7")
    sh = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(sh), 1) == 0  # non-blocking, default priority
    stream = sh.value
    p = Plan(t.device(), 256, 256 * 48 * 3, 100, 0, mode=2)
    for nq, radius, seed in [(8, 2, 922), (64, 5, 923), (256, 5, 924), (16, 2, 925)]:
        _, r = cfg(100_000, 48, nq, seed, radius)
        items = t.compile(r, nq)
        kq = np.full(nq, 100, np.uint32)
        p.upload(items, nq, kq)
        p.run(stream)
        got = p.download(stream)
        ri, rc = expect(rows, r, nq, 100)
        assert_topk_equal(got[0], got[1], ri, rc, 100, nq, f"{nq} queries, range {radius}")
    p.set_graph(True)
    for _ in range(3):
        p.run(stream)
        again = p.download(stream)
        assert np.array_equal(again[0], got[0]) and np.array_equal(again[1], got[1])
        match_topk(t.device(), items[:200], nq, 5)
    p.close()
    assert hip.hipStreamDestroy(C.c_void_p(stream)) == 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_failed_upload_leaves_no_batch(mode):
    """An upload that fails after a good batch has run leaves the plan without
    a batch: run and download refuse until the next good upload."""
    from genie_amd import GenieError
    rows, r = cfg(40_000, 16, 12, 931)
    t = Table.from_rows(rows)
    items = t.compile(r, 12)
    p = Plan(t.device(), 12, len(items), 50, 0, mode=mode)
    p.upload(items, 12, np.full(12, 50, np.uint32))
    p.run()
    first = p.download()
    bad = np.full(12, 50, np.uint32)
    bad[3] = 51  # above the plan's max_k
    with pytest.raises(GenieError):
        p.upload(items, 12, bad)
    with pytest.raises(GenieError):
        p.run()
    with pytest.raises(GenieError):
        p.download()
    p.upload(items, 12, np.full(12, 50, np.uint32))
    p.run()
    again = p.download()
    assert np.array_equal(first[0], again[0]) and np.array_equal(first[1], again[1])
    ri, rc = expect(rows, r, 12, 50)
    assert_topk_equal(again[0], again[1], ri, rc, 50, 12, "after a failed upload")


def test_match_topk_on_a_user_stream_into_device_buffers():
    """genie_hip_match_topk_stream: ordered on the caller's stream, results
    into host buffers and into device buffers (copied back here), equal to
    the oracle's."""
    from genie_amd._native import _hip as H, _chk_hip, DIM_DTYPE
    rows, r = cfg(80_000, 24, 20, 941)
    t = Table.from_rows(rows)
    items = np.ascontiguousarray(t.compile(r, 20), DIM_DTYPE)
    ri, rc = expect(rows, r, 20, 50)
    hip = C.CDLL("libamdhip64.so.7")
    sh = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(sh), 1) == 0
    ids, cnt = match_topk(t.device(), items, 20, 50, stream=sh.value)
    assert_topk_equal(ids, cnt, ri, rc, 50, 20, "user stream, host buffers")
    d_ids, d_cnt = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_ids), C.c_size_t(20 * 50 * 4)) == 0
    assert hip.hipMalloc(C.byref(d_cnt), C.c_size_t(20 * 50 * 4)) == 0
    kq = np.full(20, 50, np.uint32)
    _chk_hip(H.genie_hip_match_topk_stream(t.device().handle, items.ctypes.data, len(items), 20, kq.ctypes.data, 0,
                                           d_ids.value, d_cnt.value, sh.value))
    gi, gc = np.zeros(1000, np.int32), np.zeros(1000, np.int32)
    assert hip.hipMemcpy(C.c_void_p(gi.ctypes.data), d_ids, C.c_size_t(4000), 2) == 0  # device to host
    assert hip.hipMemcpy(C.c_void_p(gc.ctypes.data), d_cnt, C.c_size_t(4000), 2) == 0
    assert_topk_equal(gi, gc, ri, rc, 50, 20, "user stream, device buffers")
    assert hip.hipFree(d_ids) == 0 and hip.hipFree(d_cnt) == 0
    assert hip.hipStreamDestroy(sh) == 0
