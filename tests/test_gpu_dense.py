"""Dense form of the batched search (bit-sliced counting over per-object key
codes built from the inverted lists) against the exhaustive oracle and the
plain form: identical results; and the batches and indexes that must take the
lane form instead."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

import genie_amd
from genie_amd import Plan, Table, match_topk, workloads as W
from genie_amd._native import RANGE_DTYPE

pytestmark = pytest.mark.gpu


def run(table, items, nq, k, dense=True, graph=False, runs=1):
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
    p = Plan(table.device(), nq, max(1, len(items)), int(kq.max()), 0, mode=1)
    p.set_dense(dense)
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


@pytest.mark.parametrize("n,d,q,k", [
    (150_000, 32, 37, 50),    # a partly filled last query group
    (70_001, 20, 64, 10),     # dims not a multiple of 16, objects not a multiple of the tile
    (5_000, 128, 33, 100),    # fewer tiles than the sample: every tile is sampled
    (300_000, 128, 96, 100),  # full width, three query groups
    (40_000, 7, 5, 1000),     # few dims, large k
])
def test_dense_equals_oracle(n, d, q, k):
    w = W.cfg1(n=n, d=d, q=q, k=k)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], q)
    (b, p) = run(t, items, q, k)
    assert p.dense_active()
    ri, rc = oracle_single(w["rows"], w["ranges"], q, k)
    assert_topk_equal(b[0], b[1], ri, rc, k, q, f"dense n={n} d={d}")


def test_dense_equals_plain_mixed_k_and_normal_data():
    w = W.cfg0(n=20_000, q=70)
    t = Table.from_rows(w["rows"])
    r = w["ranges"].copy()
    r["low"] -= 3  # ranges of 7 values
    r["up"] += 3
    items = t.compile(r, 70)
    kq = np.array([10, 100, 1000] * 23 + [3], np.uint32)
    a = match_topk(t.device(), items, 70, kq)
    (b, p) = run(t, items, 70, kq)
    assert p.dense_active()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_dense_empty_queries_pad():
    w = W.cfg1(n=30_000, d=16, q=10, k=20)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 12)  # queries 10 and 11 have no ranges
    (b, p) = run(t, items, 12, 20)
    assert p.dense_active()
    ids, cnt = b[0].reshape(12, 20), b[1].reshape(12, 20)
    assert np.all(ids[10:] == 0) and np.all(cnt[10:] == 0)
    ri, rc = oracle_single(w["rows"], w["ranges"], 12, 20)
    assert_topk_equal(b[0], b[1], ri, rc, 20, 12, "dense empty")


def test_dense_candidate_regrowth():
    """One wide range on one dim per query: most objects count 1, the sample
    bound stays at 1, and the candidate space has to grow several times."""
    rows = genie_amd.synth.rows(91, 200_000, 4, genie_amd.synth.UNIFORM, 0.0, 0.0, 0, 15, 16)
    nq = 40
    r = np.zeros(nq, RANGE_DTYPE)
    r["query"] = np.arange(nq)
    r["dim"] = np.arange(nq) % 4
    r["low"] = np.arange(nq) % 3
    r["up"] = 12
    r["weight"] = 1
    r["order"] = 0
    t = Table.from_rows(rows)
    items = t.compile(r, nq)
    (b, p) = run(t, items, nq, 10)
    assert p.dense_active()
    ri, rc = oracle_single(rows, r, nq, 10)
    assert_topk_equal(b[0], b[1], ri, rc, 10, nq, "dense regrowth")


def test_dense_asks_grow_when_the_sample_misses():
    """Objects of the dense sample tiles never match, so every query's sample
    histogram is empty and its device-side candidate ask (4k + 1,024) is far
    below the ~167 K objects that count 1.  With k >= 1,000 the index holds
    at most 200 k, so the host total soon gives each query the whole index:
    the asks themselves have to grow, run after run."""
    n, tile = 200_000, 4096
    n_tiles = (n + tile - 1) // tile
    ns = min(n_tiles, max(8, (n_tiles + 31) // 32))
    sample = {i * n_tiles // ns for i in range(ns)}
    obj_tile = np.arange(n) // tile
    rows = np.zeros((n, 2), np.int32)
    rows[:, 0] = [0 if t in sample else 1 for t in obj_tile]
    rows[:, 1] = np.arange(n) % 7
    nq = 6
    r = np.zeros(nq, RANGE_DTYPE)
    r["query"] = np.arange(nq)
    r["dim"] = 0
    r["low"] = 1
    r["up"] = 1
    r["weight"] = 1
    r["order"] = 0
    kq = np.array([2000, 2000, 1000, 2000, 1500, 2000], np.uint32)
    t = Table.from_rows(rows)
    items = t.compile(r, nq)
    p = Plan(t.device(), nq, len(items), 2000, 0, mode=1)
    p.upload(items, nq, kq)
    assert p.dense_active()
    p.run()
    got = p.download()
    ri, rc = oracle_single(rows, r, nq, kq)
    assert_topk_equal(got[0], got[1], ri, rc, kq, nq, "dense asks regrowth")
    # the smallest ids outside the sample tiles, all with count 1
    assert got[0][0] == tile and np.all(got[1][:2000] == 1)


def test_dense_graph_replays_match_eager():
    w = W.cfg1(n=100_000, d=16, q=40, k=32)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 40)
    eager = match_topk(t.device(), items, 40, 32)
    kq = np.full(40, 32, np.uint32)
    p = Plan(t.device(), 40, len(items), 32, 0, mode=1)
    p.upload(items, 40, kq)
    assert p.dense_active()
    p.set_graph(True)
    for _ in range(4):  # replays with eager work and synchronisations between them
        p.run()
        got = p.download()
        assert np.array_equal(eager[0], got[0]) and np.array_equal(eager[1], got[1])
        match_topk(t.device(), items[:50], 40, 5)


def test_dense_autotune_includes_dense():
    w = W.cfg1(n=200_000, d=32, q=32, k=100)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 32)
    a = match_topk(t.device(), items, 32, 100)
    p = Plan(t.device(), 32, len(items), 100, 0, mode=1)
    p.upload(items, 32, np.full(32, 100, np.uint32))
    ms = p.autotune(reps=1)
    assert len(ms) == 14 and ms[12] > 0 and ms[13] > 0  # 12 lane configurations, dense, plain
    assert p.form() == (["lane"] * 12 + ["dense", "plain"])[int(np.argmin(ms))]
    p.run()
    b = p.download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n,d,q,k,wmax", [
    (150_000, 32, 37, 50, 4),     # weights 1-4: 3 planes, a partly filled last group of 10
    (200_000, 128, 45, 30, 7),    # 3 planes at full width (counts up to 896)
    (300_000, 128, 40, 100, 15),  # 4 planes at full width (counts up to 1,920)
    (100_000, 128, 33, 20, 3),    # 2 planes at full width, three groups of 16
    (70_001, 20, 19, 10, 2),      # 2 planes, dims not a multiple of 16, two groups
])
def test_dense_weighted_equals_oracle(n, d, q, k, wmax):
    """Integer weights up to 15: as many weight planes a query in each mask
    word as the largest weight has bits (2-4; 32 / planes queries a group),
    counts assembled over 11 planes."""
    w = W.cfg1(n=n, d=d, q=q, k=k)
    t = Table.from_rows(w["rows"])
    r = w["ranges"].copy()
    r["weight"] = genie_amd.synth.uniform_int(900 + wmax, 0, len(r), 1, wmax)
    items = t.compile(r, q)
    (b, p) = run(t, items, q, k, graph=True, runs=3)
    assert p.dense_active() and p.form() == "dense"
    ri, rc = oracle_single(w["rows"], r, q, k)
    assert_topk_equal(b[0], b[1], ri, rc, k, q, f"weighted dense n={n} d={d} w<={wmax}")


@pytest.mark.parametrize("q,wmax", [
    (700, 1),    # 22 groups of 32: 7 tile chunks a group on 32 workgroups per XCD
    (1056, 1),   # 33 groups: 8 chunks
    (176, 15),   # 22 groups of 8
    (1030, 4),   # 103 groups of 10: 4 chunks
])
def test_dense_group_chunks_equal_oracle(q, wmax):
    """Group counts that the workgroups of an XCD do not divide: the main
    pass splits every group's tiles into chunks and deals out (group, chunk)
    units; each (group, tile) is counted once."""
    w = W.cfg1(n=60_000, d=16, q=q, k=20, seed=q)
    t = Table.from_rows(w["rows"])
    r = w["ranges"].copy()
    r["weight"] = genie_amd.synth.uniform_int(q, 0, len(r), 1, wmax)
    (b, p) = run(t, t.compile(r, q), q, 20)
    assert p.form() == "dense"
    ri, rc = oracle_single(w["rows"], r, q, 20)
    assert_topk_equal(b[0], b[1], ri, rc, 20, q, f"group chunks q={q} w<={wmax}")


@pytest.mark.parametrize("form", ["lane", "plain"])
def test_dense_weighted_batch_on_the_tile_forms(form):
    """A batch the dense form could take, with counts past 255 (128 dims,
    weights up to 15), run with the dense form switched off: the lane and
    plain forms size their counters from the batch's count bound, not from
    the unit-weight dense bound of 128."""
    w = W.cfg1(n=100_000, d=128, q=20, k=50, seed=77)
    t = Table.from_rows(w["rows"])
    r = w["ranges"].copy()
    r["weight"] = genie_amd.synth.uniform_int(78, 0, len(r), 8, 15)
    items = t.compile(r, 20)
    p = Plan(t.device(), 20, len(items), 50, 0, mode=1)
    p.set_dense(False)
    p.upload(items, 20, np.full(20, 50, np.uint32))
    p.set_form(form)
    p.run()
    b = p.download()
    ri, rc = oracle_single(w["rows"], r, 20, 50)
    assert int(rc.max()) > 255
    assert_topk_equal(b[0], b[1], ri, rc, 50, 20, f"weighted batch, {form} form")


def test_dense_weighted_mixed_k_and_compressed():
    w = W.cfg1(n=120_000, d=64, q=30, k=100)
    r = w["ranges"].copy()
    r["weight"] = genie_amd.synth.uniform_int(77, 0, len(r), 1, 4)
    kq = np.array([10, 100, 1000] * 10, np.uint32)
    ri, rc = oracle_single(w["rows"], r, 30, kq)
    for t in (Table.from_rows(w["rows"]), Table.from_rows(w["rows"], 1024, True, codec=11)):
        items = t.compile(r, 30)
        (b, p) = run(t, items, 30, kq)
        assert p.dense_active()
        assert_topk_equal(b[0], b[1], ri, rc, kq, 30, "weighted dense, mixed k")


def test_dense_not_used_where_it_cannot_be_exact():
    w = W.cfg1(n=60_000, d=24, q=10, k=20)
    t = Table.from_rows(w["rows"])
    # weights above 15 (the weighted masks hold four bits a query)
    for wv in (16, 300):
        r = w["ranges"].copy()
        r["weight"][::3] = wv
        (b, p) = run(t, t.compile(r, 10), 10, 20)
        assert not p.dense_active()
        ri, rc = oracle_single(w["rows"], r, 10, 20)
        assert_topk_equal(b[0], b[1], ri, rc, 20, 10, f"weights {wv}")
    # a query covering keys twice (ranges repeated on the same dims)
    r2 = np.concatenate([w["ranges"], w["ranges"][:3]])
    r2 = r2[np.argsort(r2["query"], kind="stable")]
    (b, p) = run(t, t.compile(r2, 10), 10, 20)
    assert not p.dense_active()
    ri, rc = oracle_single(w["rows"], r2, 10, 20)
    assert_topk_equal(b[0], b[1], ri, rc, 20, 10, "overlap")
    # dense switched off
    (b, p) = run(t, t.compile(w["ranges"], 10), 10, 20, dense=False)
    assert not p.dense_active()


def test_dense_not_built_for_other_indexes():
    # more than 128 dims
    w = W.cfg1(n=20_000, d=130, q=4, k=10)
    t = Table.from_rows(w["rows"])
    (b, p) = run(t, t.compile(w["ranges"], 4), 4, 10)
    assert not p.dense_active()
    ri, rc = oracle_single(w["rows"], w["ranges"], 4, 10)
    assert_topk_equal(b[0], b[1], ri, rc, 10, 4, "130 dims")
    # a set-valued (bijectMap) table takes the set form instead (its own tests)
    rp, vals = genie_amd.synth.sets(61, 20_000, 5, 40, 3000, 1.05)
    qp, qv = genie_amd.synth.sets(62, 6, 1, 30, 3000, 1.05)
    sr = W.set_ranges(qp, qv, np.ones(len(qv), np.float32))
    ts = Table.from_sets(rp, vals)
    (b, p) = run(ts, ts.compile(sr, 6), 6, 10)
    assert p.dense_active()
    q_of = sr["query"]
    ri, rc = oracle_topk_batch(lambda q: O.count_exhaustive_biject(rp, vals, sr[q_of == q]), 6, 10)
    assert_topk_equal(b[0], b[1], ri, rc, 10, 6, "sets")


def plain(table, items, nq, k):
    kq = np.full(nq, k, np.uint32)
    p = Plan(table.device(), nq, max(1, len(items)), k, 0, mode=0)
    p.upload(items, nq, kq)
    p.run()
    return p.download()


def test_dense_load_balanced_pieces():
    """Load-balanced items split a range into pieces of max_load postings;
    the pieces of one range together cover whole keys."""
    w = W.cfg1(n=60_000, d=24, q=10, k=20)
    t = Table.from_rows(w["rows"])
    (b, p) = run(t, t.compile(w["ranges"], 10, load_balance=True, max_load=64), 10, 20)
    assert p.dense_active()
    ri, rc = oracle_single(w["rows"], w["ranges"], 10, 20)
    assert_topk_equal(b[0], b[1], ri, rc, 20, 10, "load balance")


@pytest.mark.parametrize("codec", [1, 3, 6, 11, 13])
def test_dense_compressed_equals_oracle(codec):
    """A compressed single-value table is decoded once at index creation for
    its codes; a batch's items (one per segment) run the dense form."""
    w = W.cfg1(n=120_000, d=24, q=40, k=50)
    t = Table.from_rows(w["rows"], 1024, True, codec=codec)
    items = t.compile(w["ranges"], 40)
    (b, p) = run(t, items, 40, 50)
    assert p.dense_active()
    ri, rc = oracle_single(w["rows"], w["ranges"], 40, 50)
    assert_topk_equal(b[0], b[1], ri, rc, 50, 40, f"compressed codec {codec}")


def test_dense_not_used_for_part_of_a_key():
    w = W.cfg1(n=60_000, d=16, q=8, k=20)
    # uncompressed: one item shortened by a posting
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 8).copy()
    i = int(np.argmax(items["end_pos"] - items["start_pos"]))
    items["end_pos"][i] -= 1
    (b, p) = run(t, items, 8, 20)
    assert not p.dense_active()
    a = plain(t, items, 8, 20)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # compressed: one segment of a key left out (4 values per dim: every key
    # has about 60 segments of 256 ids, so no segment is a whole key)
    rows = genie_amd.synth.rows(93, 60_000, 16, genie_amd.synth.UNIFORM, 0.0, 0.0, 0, 3, 4)
    r = np.zeros(8 * 16, RANGE_DTYPE)
    r["query"] = np.repeat(np.arange(8), 16)
    r["dim"] = np.tile(np.arange(16), 8)
    r["low"] = r["up"] = rows[:8].reshape(-1)
    r["weight"] = 1
    r["order"] = r["dim"]
    tc = Table.from_rows(rows, 256, True, codec=11)
    items = tc.compile(r, 8)
    assert len(items) > 50 * 8 * 16
    items = np.delete(items, len(items) // 2)
    (b, p) = run(tc, items, 8, 20)
    assert not p.dense_active()
    a = plain(tc, items, 8, 20)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_dense_speculative_threshold_redo():
    """The sample sees 80 objects at the top count, the whole index only 90
    (k = 100): the speculative threshold leaves the query short, so its group
    is counted again from the safe threshold.  A second query of the same
    group keeps the candidates of the first count."""
    n, d = 200_000, 4
    rows = np.zeros((n, d), np.int32)
    rows[:, 0] = 5
    rows[:, 1] = np.where(np.arange(n) % 7 == 0, 3, 0)
    rows[np.r_[0:80, 4096:4106], :] = 5  # tile 0 is a sample tile, tile 1 is not
    r = np.zeros(d + 1, RANGE_DTYPE)
    r["query"] = [0, 0, 0, 0, 1]
    r["dim"] = [0, 1, 2, 3, 1]
    r["low"] = [5, 5, 5, 5, 3]
    r["up"] = [5, 5, 5, 5, 3]
    r["weight"] = 1
    r["order"] = [0, 1, 2, 3, 0]
    t = Table.from_rows(rows)
    items = t.compile(r, 2)
    (b, p) = run(t, items, 2, 100)
    assert p.dense_active()
    ri, rc = oracle_single(rows, r, 2, 100)
    assert_topk_equal(b[0], b[1], ri, rc, 100, 2, "dense redo")
    ids, cnt = b[0].reshape(2, 100), b[1].reshape(2, 100)
    assert np.all(cnt[0, :90] == 4) and np.all(cnt[0, 90:] == 1)
    assert list(ids[0, 80:90]) == list(range(4096, 4106)) and list(ids[0, 90:]) == list(range(80, 90))
