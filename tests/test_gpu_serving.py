"""Serving form: continuous batching over a resident index.  Every request's
result equals the exhaustive oracle's and is the same alone and in the mix
(bit-exact), the memory budget admission holds, a hash-form table that fills
up is settled before its results are handed back, and bad requests are
refused without failing others."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

import genie_amd
from genie_amd import GenieError, Server, Table, workloads as W
from genie_amd._native import RANGE_DTYPE

pytestmark = pytest.mark.gpu


def requests_of(table, w, merge_runs=False):
    items = table.compile(w["ranges"], w["n_queries"], merge_runs=merge_runs)
    q = items["query"]
    ptr = np.searchsorted(q, np.arange(w["n_queries"] + 1)).astype(np.uint64)
    return items, ptr


def oracle_of(rows, ranges, nq, k):
    """Exhaustive closed-form counts from the raw rows, top k per request."""
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, ranges[q_of == q]), nq, k)


@pytest.mark.parametrize("mode,codec,merged", [(0, 0, False), (1, 0, False), (2, 0, False), (0, 11, False),
                                               (2, 11, True), (1, 11, True)])
def test_alone_equals_in_the_mix(mode, codec, merged):
    """merged: a compressed index's requests carry one item per range (its
    segments merged), as the serving benchmark sends them."""
    w = W.cfg4(n=200_000, d=32, q=60)
    t = Table.from_rows(w["rows"], 1024, codec != 0, codec=codec)
    items, ptr = requests_of(t, w, merged)
    k = w["k"]
    ri, rc = oracle_of(w["rows"], w["ranges"], w["n_queries"], k)
    srv = Server(t.device(), executors=3, max_batch=16, max_k=1000, mode=mode)
    alone = []
    for r in range(w["n_queries"]):
        tk = srv.submit(items[ptr[r]:ptr[r + 1]], int(k[r]))
        ids, cnt, lat = srv.wait(tk, int(k[r]))
        alone.append((ids, cnt))
        assert lat >= 0
    assert_topk_equal(np.concatenate([a[0] for a in alone]), np.concatenate([a[1] for a in alone]), ri, rc, k,
                      w["n_queries"], f"served alone, mode {mode} codec {codec}")
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
    # and the oracle's answer for the whole mix
    assert_topk_equal(ids, cnt, ri, rc, k, w["n_queries"], f"served in the mix, mode {mode} codec {codec}")
    srv.close()


@pytest.mark.parametrize("codec", [0, 13])
def test_selective_requests_on_the_hash_form(codec):
    """LSH-style exact-match requests (8,192 buckets per dim): the automatic
    plan runs them on the hash form (compressed: segments decoded once per
    batch); in the mix each equals the oracle's answer and its lone one."""
    rows = genie_amd.synth.rows(301, 300_000, 32, genie_amd.synth.UNIFORM, 0.0, 0.0, 0, 8191, 8192)
    qv, _ = genie_amd.synth.queries(302, rows, 40, redraw_p=0.3, n_values=8192)
    ranges = W.single_ranges(qv, 0)
    t = Table.from_rows(rows, 1024, codec != 0, codec=codec)
    w = dict(ranges=ranges, n_queries=40)
    items, ptr = requests_of(t, w)
    k = np.array([10, 100, 1000, 50] * 10, np.uint32)
    srv = Server(t.device(), executors=3, max_batch=16, max_k=1000, mode=2)
    lat, ids, cnt, wall = srv.replay(items, ptr, k, np.cumsum(genie_amd.synth.exponential(11, 40, 2000.0)))
    ri, rc = oracle_of(rows, ranges, 40, k)
    assert_topk_equal(ids, cnt, ri, rc, k, 40, f"selective served, codec {codec}")
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
    ri, rc = oracle_of(w["rows"], w["ranges"], 12, 100)
    assert_topk_equal(ids, cnt, ri, rc, 100, 12, "one request at a time")
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
    r0 = w["ranges"][w["ranges"]["query"] == 0]
    ri, rc = O.topk(O.count_exhaustive_single(w["rows"], r0), 10)
    assert np.array_equal(ids, ri) and np.array_equal(cnt, rc)
    srv.close()


@pytest.mark.parametrize("executors", [1, 3])
def test_hash_table_full_is_settled_before_hand_back(executors):
    """Request 0 names one key of 20,000 objects, all in its first tile: its
    hash-form table fills up, and the run leaves its rows incomplete until
    the batch is settled on the direct-mapped kernel.  Served alone and in a
    mix of selective requests, every result equals the oracle's."""
    n = 400_000
    rows = genie_amd.synth.rows(221, n, 4, genie_amd.synth.UNIFORM, 0.0, 0.0, 2, 4097, 4096)
    rows[:20_000, 0] = 1
    nq = 24
    r = np.zeros(nq * 4, RANGE_DTYPE)
    r["query"] = np.repeat(np.arange(nq), 4)
    r["dim"] = np.tile(np.arange(4), nq)
    r["low"] = r["up"] = rows[50_000:50_000 + nq].reshape(-1)
    r["weight"] = 1
    r["order"] = r["dim"]
    full = [0, 9, 17]  # requests that name only the crowded key
    keep = ~np.isin(r["query"], full) | (r["dim"] == 0)
    r = r[keep].copy()
    r["low"][np.isin(r["query"], full)] = 1
    r["up"][np.isin(r["query"], full)] = 1
    t = Table.from_rows(rows)
    items = t.compile(r, nq)
    ptr = np.searchsorted(items["query"], np.arange(nq + 1)).astype(np.uint64)
    k = np.full(nq, 100, np.uint32)
    ri, rc = oracle_of(rows, r, nq, k)
    srv = Server(t.device(), executors=executors, max_batch=8, max_k=100, mode=2)
    tk = srv.submit(items[ptr[0]:ptr[1]], 100)
    ids, cnt, _ = srv.wait(tk, 100)
    assert list(ids) == list(range(100)) and np.all(cnt == 1)
    for arrivals in (np.zeros(nq), np.cumsum(genie_amd.synth.exponential(13, nq, 3000.0))):
        lat, ids, cnt, wall = srv.replay(items, ptr, k, arrivals)
        assert_topk_equal(ids, cnt, ri, rc, k, nq, f"table full served, {executors} executors")
    assert srv.stats()["reruns"] >= 1
    srv.close()
