"""Hash form of the plain scan (per (query, tile) open-addressing count tables
in LDS, exact radix select of the tile's top k) against the exhaustive
oracle and the direct-mapped form: selective and forced batches, weights,
ties, large k, compressed and set-valued tables, a table that fills up (its
query is run again on the direct-mapped form), graph replays."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

import genie_amd
from genie_amd import Plan, Table, match_topk, workloads as W
from genie_amd._native import RANGE_DTYPE

pytestmark = pytest.mark.gpu


def plain(table, items, nq, k, hash_mode=-1, graph=False, runs=1):
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
    p = Plan(table.device(), nq, max(1, len(items)), int(kq.max()), 0, mode=0)
    p.set_hash(hash_mode)
    p.upload(items, nq, kq)
    if graph:
        p.set_graph(True)
    out = None
    for _ in range(runs):
        p.run()
        got = p.download()
        if out is not None:
            assert np.array_equal(got[0], out[0]) and np.array_equal(got[1], out[1])
        out = got
    return out, p


def oracle_single(rows, ranges, nq, k):
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, ranges[q_of == q]), nq, k)


def lsh(n, d, q, buckets, seed):
    rows = genie_amd.synth.rows(seed, n, d, genie_amd.synth.UNIFORM, 0.0, 0.0, 0, buckets - 1, buckets)
    qv, _ = genie_amd.synth.queries(seed + 1, rows, q, redraw_p=0.3, n_values=buckets)
    return rows, W.single_ranges(qv, 0)


@pytest.mark.parametrize("n,d,q,buckets,k", [
    (400_000, 32, 40, 4096, 100),    # several tiles per query, ties at count 1 cut by id
    (200_003, 16, 24, 8192, 1000),   # k above most tiles' matches
    (60_000, 64, 9, 2048, 7),        # one tile per query
])
def test_hash_equals_oracle(n, d, q, buckets, k):
    rows, r = lsh(n, d, q, buckets, 201 + d)
    t = Table.from_rows(rows)
    items = t.compile(r, q)
    (b, p) = plain(t, items, q, k)
    assert p.hash_queries() == q
    ri, rc = oracle_single(rows, r, q, k)
    assert_topk_equal(b[0], b[1], ri, rc, k, q, f"hash n={n} d={d}")
    (c, p) = plain(t, items, q, k, hash_mode=0)
    assert p.hash_queries() == 0
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


@pytest.mark.parametrize("n,d,q,buckets,k", [
    (1_000_000, 64, 16, 1024, 1000),  # 16 tiles a query; the top 1,000 ends in ties at count 2
    (1_000_000, 48, 16, 2048, 100),   # 4 tiles a query; ties at count 2 cut by id
])
def test_hash_rank_threshold_across_tiles(n, d, q, buckets, k):
    """A query's tiles run tile-major, and a tile that emitted k candidates
    publishes its k-th (count, id) key: later tiles, of higher ids, leave out
    counts equal to that key's.  Every tile's own k-th is a count-1 tie, so
    the later tiles select from their count-2 slots only; ids and counts
    equal the oracle's and the direct-mapped form's exactly, ties included."""
    rows, r = lsh(n, d, q, buckets, 331 + d)
    t = Table.from_rows(rows)
    items = t.compile(r, q)
    (b, p) = plain(t, items, q, k, hash_mode=1, graph=True, runs=3)
    assert p.hash_queries() == q
    ri, rc = oracle_single(rows, r, q, k)
    assert np.array_equal(b[1], rc) and np.array_equal(b[0], ri)
    (c, p) = plain(t, items, q, k, hash_mode=0)
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


def test_hash_forced_on_dense_queries():
    """Every query on the hash form although each holds a posting per object
    and dim: tiles shrink to 8,192 objects, never more ids than slots."""
    w = W.cfg1(n=100_000, d=16, q=12, k=50)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 12)
    (b, p) = plain(t, items, 12, 50, hash_mode=1)
    assert p.hash_queries() == 12
    ri, rc = oracle_single(w["rows"], w["ranges"], 12, 50)
    assert_topk_equal(b[0], b[1], ri, rc, 50, 12, "forced hash")


def test_hash_weights_and_mixed_k():
    rows, r = lsh(150_000, 24, 18, 2048, 211)
    r = r.copy()
    r["weight"] = np.random.default_rng(9).integers(1, 3000, len(r))  # counts need 17 bits
    t = Table.from_rows(rows)
    items = t.compile(r, 18)
    kq = np.array([1, 10, 100, 1000, 4096, 37] * 3, np.uint32)
    (b, p) = plain(t, items, 18, kq)
    assert p.hash_queries() == 18
    ri, rc = oracle_single(rows, r, 18, kq)
    assert_topk_equal(b[0], b[1], ri, rc, kq, 18, "hash weights")


def test_hash_table_full_falls_back():
    """Query 0 names one key of 20,000 objects, all in its first tile: more
    distinct ids than its table holds.  Its run is redone on the
    direct-mapped form; the other queries stay on the hash form."""
    n = 400_000
    rows = genie_amd.synth.rows(221, n, 4, genie_amd.synth.UNIFORM, 0.0, 0.0, 2, 4097, 4096)
    rows[:20_000, 0] = 1
    nq = 6
    r = np.zeros(nq * 4, RANGE_DTYPE)
    r["query"] = np.repeat(np.arange(nq), 4)
    r["dim"] = np.tile(np.arange(4), nq)
    r["low"] = r["up"] = rows[50_000:50_000 + nq].reshape(-1)
    r["weight"] = 1
    r["order"] = r["dim"]
    r = r[~((r["query"] == 0) & (r["dim"] > 0))].copy()
    r["low"][0] = r["up"][0] = 1
    t = Table.from_rows(rows)
    items = t.compile(r, nq)
    (b, p) = plain(t, items, nq, 100)
    assert p.hash_queries() == nq - 1
    ri, rc = oracle_single(rows, r, nq, 100)
    assert_topk_equal(b[0], b[1], ri, rc, 100, nq, "table full")
    ids, cnt = b[0].reshape(nq, 100), b[1].reshape(nq, 100)
    assert list(ids[0]) == list(range(100)) and np.all(cnt[0] == 1)
    # a new batch on the same plan starts on the hash form again
    sub = items[items["query"] >= 1].copy()
    sub["query"] -= 1
    p.upload(sub, nq - 1, np.full(nq - 1, 100, np.uint32))
    assert p.hash_queries() == nq - 1
    p.run()
    got = p.download()
    assert np.array_equal(got[0], b[0][100:]) and np.array_equal(got[1], b[1][100:])


@pytest.mark.parametrize("codec", [3, 6, 11, 13])
def test_hash_compressed(codec):
    rows, r = lsh(200_000, 16, 14, 4096, 231)
    t = Table.from_rows(rows, 1024, True, codec=codec)
    items = t.compile(r, 14)
    (b, p) = plain(t, items, 14, 60)
    assert p.hash_queries() == 14
    ri, rc = oracle_single(rows, r, 14, 60)
    assert_topk_equal(b[0], b[1], ri, rc, 60, 14, f"hash codec {codec}")


def test_hash_sets_with_duplicates():
    rp, vals = genie_amd.synth.sets(241, 150_000, 5, 40, 60_000, 1.05)
    qp, qv = genie_amd.synth.sets(242, 20, 1, 30, 60_000, 1.05)
    wts = genie_amd.synth.uniform_int(243, 0, len(qv), 1, 4)
    ranges = W.set_ranges(qp, qv, wts)
    t = Table.from_sets(rp, vals)
    items = t.compile(ranges, 20)
    kq = np.array([10, 100, 1000, 3] * 5, np.uint32)
    (b, p) = plain(t, items, 20, kq, hash_mode=1)
    assert p.hash_queries() == 20
    q_of = ranges["query"]
    ri, rc = oracle_topk_batch(lambda q: O.count_exhaustive_biject(rp, vals, ranges[q_of == q]), 20, kq)
    assert_topk_equal(b[0], b[1], ri, rc, kq, 20, "hash sets")


def test_hash_and_direct_in_one_batch_and_graph():
    """Queries 0-2 cover whole dims (direct-mapped form), the rest are
    selective (hash form); graph replays with eager work between them."""
    rows, r = lsh(250_000, 16, 12, 4096, 251)
    r = r.copy()
    m = r["query"] < 3
    r["low"][m] = 0
    r["up"][m] = 4095
    t = Table.from_rows(rows)
    items = t.compile(r, 12)
    (b, p) = plain(t, items, 12, 40)
    assert p.hash_queries() == 9
    ri, rc = oracle_single(rows, r, 12, 40)
    assert_topk_equal(b[0], b[1], ri, rc, 40, 12, "mixed")
    kq = np.full(12, 40, np.uint32)
    g = Plan(t.device(), 12, len(items), 40, 0, mode=0)
    g.upload(items, 12, kq)
    g.set_graph(True)
    for _ in range(3):
        g.run()
        got = g.download()
        assert np.array_equal(got[0], b[0]) and np.array_equal(got[1], b[1])
        match_topk(t.device(), items[:30], 12, 5)


def test_automatic_plan_uses_the_hash_form():
    rows, r = lsh(300_000, 64, 16, 8192, 261)
    t = Table.from_rows(rows)
    items = t.compile(r, 16)
    a = match_topk(t.device(), items, 16, 100)
    ri, rc = oracle_single(rows, r, 16, 100)
    assert_topk_equal(a[0], a[1], ri, rc, 100, 16, "automatic")
    p = Plan(t.device(), 16, len(items), 100, 0, mode=2)
    p.upload(items, 16, np.full(16, 100, np.uint32))
    assert not p.dense_active() and p.hash_queries() == 16


def test_set_hash_rejects_bad_mode():
    w = W.cfg0(n=500, d=8, q=2)
    t = Table.from_rows(w["rows"])
    p = Plan(t.device(), 2, 100, 5, 0, mode=0)
    with pytest.raises(genie_amd.GenieError):
        p.set_hash(3)
