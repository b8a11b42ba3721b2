"""Serving form: continuous batching over a resident index.  Every request's
result is the same alone and in the mix (bit-exact), the memory budget
admission holds, and bad requests are refused without failing others."""
import numpy as np
import pytest

import genie_amd
from genie_amd import GenieError, Server, Table, match_topk, workloads as W

pytestmark = pytest.mark.gpu


def requests_of(table, w, merge_runs=False):
    items = table.compile(w["ranges"], w["n_queries"], merge_runs=merge_runs)
    q = items["query"]
    ptr = np.searchsorted(q, np.arange(w["n_queries"] + 1)).astype(np.uint64)
    return items, ptr


@pytest.mark.parametrize("mode,codec,merged", [(0, 0, False), (1, 0, False), (2, 0, False), (0, 11, False),
                                               (2, 11, True), (1, 11, True)])
def test_alone_equals_in_the_mix(mode, codec, merged):
    """merged: a compressed index's requests carry one item per range (its
    segments merged), as the serving benchmark sends them."""
    w = W.cfg4(n=200_000, d=32, q=60)
    t = Table.from_rows(w["rows"], 1024, codec != 0, codec=codec)
    items, ptr = requests_of(t, w, merged)
    k = w["k"]
    srv = Server(t.device(), executors=3, max_batch=16, max_k=1000, mode=mode)
    alone = []
    for r in range(w["n_queries"]):
        tk = srv.submit(items[ptr[r]:ptr[r + 1]], int(k[r]))
        ids, cnt, lat = srv.wait(tk, int(k[r]))
        alone.append((ids, cnt))
        assert lat >= 0
    # all at once (t = 0) and as a Poisson trace: batches form and overlap
    for arrivals in (np.zeros(w["n_queries"]), np.cumsum(genie_amd.synth.exponential(9, w["n_queries"], 4000.0))):
        lat, ids, cnt, wall = srv.replay(items, ptr, k, arrivals)
        off = 0
        for r in range(w["n_queries"]):
            kr = int(k[r])
            assert np.array_equal(ids[off:off + kr], alone[r][0]) and np.array_equal(cnt[off:off + kr], alone[r][1]), r
            off += kr
        assert np.all(lat > 0) and wall > 0
    st = srv.stats()
    assert st["requests"] == 3 * w["n_queries"] and st["batches"] < st["requests"]
    assert st["max_inflight"] > 1
    # the slowest batch per phase is within the phase's total, and a reset clears it
    sl = srv.slowest(reset=True)
    tot = st["executor_seconds"]
    assert 0 < sl["wait"] <= tot["wait"] and 0 < sl["prepare_upload"] <= tot["prepare_upload"]
    assert sl["queue"] > 0
    assert all(v == 0 for v in srv.slowest(reset=False).values())
    # and the same as one plain batch of everything
    pi, pc = match_topk(t.device(), items, w["n_queries"], k)
    assert np.array_equal(pi, ids) and np.array_equal(pc, cnt)
    srv.close()


@pytest.mark.parametrize("codec", [0, 13])
def test_selective_requests_on_the_hash_form(codec):
    """LSH-style exact-match requests (8,192 buckets per dim): the automatic
    plan runs them on the hash form (compressed: segments decoded once per
    batch); in the mix each equals its lone answer and the one-shot search
    of the whole batch (the hash form's oracle parity is in test_gpu_hash)."""
    rows = genie_amd.synth.rows(301, 300_000, 32, genie_amd.synth.UNIFORM, 0.0, 0.0, 0, 8191, 8192)
    qv, _ = genie_amd.synth.queries(302, rows, 40, redraw_p=0.3, n_values=8192)
    ranges = W.single_ranges(qv, 0)
    t = Table.from_rows(rows, 1024, codec != 0, codec=codec)
    w = dict(ranges=ranges, n_queries=40)
    items, ptr = requests_of(t, w)
    k = np.array([10, 100, 1000, 50] * 10, np.uint32)
    srv = Server(t.device(), executors=3, max_batch=16, max_k=1000, mode=2)
    lat, ids, cnt, wall = srv.replay(items, ptr, k, np.cumsum(genie_amd.synth.exponential(11, 40, 2000.0)))
    pi, pc = match_topk(t.device(), items, 40, k)
    assert np.array_equal(pi, ids) and np.array_equal(pc, cnt)
    off = 0
    for r in (0, 7, 22):
        tk = srv.submit(items[ptr[r]:ptr[r + 1]], int(k[r]))
        a, c, _ = srv.wait(tk, int(k[r]))
        o = int(k[:r].sum())
        assert np.array_equal(a, ids[o:o + int(k[r])]) and np.array_equal(c, cnt[o:o + int(k[r])]), r
    srv.close()


def test_memory_budget_admits_one_at_a_time():
    w = W.cfg1(n=50_000, d=16, q=12, k=100)
    t = Table.from_rows(w["rows"])
    items, ptr = requests_of(t, w)
    srv = Server(t.device(), executors=4, max_batch=8, max_k=100, memory_budget=1)
    lat, ids, cnt, wall = srv.replay(items, ptr, 100, np.zeros(12))
    pi, pc = match_topk(t.device(), items, 12, 100)
    assert np.array_equal(pi, ids) and np.array_equal(pc, cnt)
    assert srv.stats()["max_inflight"] == 1
    srv.close()


def test_bad_request_is_refused():
    w = W.cfg0(n=500, d=8, q=2)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 2)
    srv = Server(t.device(), executors=2, max_batch=4, max_k=50)
    bad = items[:3].copy()
    bad["weight"][0] = 0.5
    with pytest.raises(GenieError):
        srv.submit(bad, 10)
    with pytest.raises(GenieError):
        srv.submit(items[:3], 51)
    tk = srv.submit(items[items["query"] == 0], 10)
    ids, cnt, _ = srv.wait(tk, 10)
    ref = match_topk(t.device(), items[items["query"] == 0], 1, 10)
    assert np.array_equal(ids, ref[0]) and np.array_equal(cnt, ref[1])
    srv.close()
