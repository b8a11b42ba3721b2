"""Set form of the batched search (set-valued tables: each object's forward
list of keys counted against an LDS hash table of its query group's keys,
with the queries' weights as bit planes) against the exhaustive oracle and
the other forms; and the batches and tables that must take another form."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

import genie_amd
from genie_amd import Plan, Table, match_topk, workloads as W

pytestmark = pytest.mark.gpu


def run(table, items, nq, k, mode=1, dense=True, graph=False, runs=1):
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
    p = Plan(table.device(), nq, max(1, len(items)), int(kq.max()), 0, mode=mode)
    if mode:
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


def oracle_sets(rp, vals, ranges, nq, k):
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_biject(rp, vals, ranges[q_of == q]), nq, k)


def sets_case(seed, n, q, vocab, qlen=30, wmax=4):
    rp, vals = genie_amd.synth.sets(seed, n, 5, 40, vocab, 1.05)
    qp, qv = genie_amd.synth.sets(seed + 1, q, 1, qlen, vocab, 1.05)
    wts = genie_amd.synth.uniform_int(seed + 2, 0, len(qv), 1, wmax)
    return rp, vals, W.set_ranges(qp, qv, wts)


@pytest.mark.parametrize("n,q,vocab,k", [
    (100_000, 40, 20_000, [10, 100, 1000, 7]),   # duplicates in documents, weights 1-4, mixed k
    (300_000, 70, 200_000, [50]),                # three query groups, a partly filled last one
    (20_000, 5, 3_000, [1000]),                  # k above most matches
])
def test_set_form_equals_oracle(n, q, vocab, k):
    rp, vals, ranges = sets_case(400 + q, n, q, vocab)
    t = Table.from_sets(rp, vals)
    items = t.compile(ranges, q)
    kq = np.resize(np.asarray(k, np.uint32), q)
    (b, p) = run(t, items, q, kq)
    assert p.dense_active()
    ri, rc = oracle_sets(rp, vals, ranges, q, kq)
    assert_topk_equal(b[0], b[1], ri, rc, kq, q, f"set form n={n}")
    (c, p2) = run(t, items, q, kq, dense=False)
    assert not p2.dense_active()
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


def test_repeated_tokens_add_their_weights():
    """A query naming a token twice counts it with both weights (weights up to
    8 per item; the sums reach 47, in the upper weight planes)."""
    rp, vals, ranges = sets_case(431, 60_000, 12, 10_000, wmax=8)
    r2 = np.concatenate([ranges, ranges[::3]])
    r2 = r2[np.argsort(r2["query"], kind="stable")]
    t = Table.from_sets(rp, vals)
    (b, p) = run(t, t.compile(r2, 12), 12, 40)
    assert p.dense_active()
    ri, rc = oracle_sets(rp, vals, r2, 12, 40)
    assert_topk_equal(b[0], b[1], ri, rc, 40, 12, "repeated tokens")
    # the same batch with every item cut into chunks of at most 300 postings
    items = t.compile(r2, 12, load_balance=True, max_load=300)
    assert len(items) > len(r2)
    (c, p) = run(t, items, 12, 40)
    assert p.dense_active()
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


def test_compressed_set_table_and_automatic_plan_and_graph():
    rp, vals, ranges = sets_case(441, 150_000, 36, 30_000)
    t = Table.from_sets(rp, vals, 1024, True, codec=6)
    items = t.compile(ranges, 36)
    (b, p) = run(t, items, 36, 100, mode=1, graph=True, runs=3)
    assert p.dense_active()
    ri, rc = oracle_sets(rp, vals, ranges, 36, 100)
    assert_topk_equal(b[0], b[1], ri, rc, 100, 36, "compressed sets")
    # the automatic plan times the forms and keeps the faster: same results
    a = match_topk(t.device(), items, 36, 100)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    (c, p2) = run(t, items, 36, 100, mode=2)
    assert np.array_equal(c[0], b[0]) and np.array_equal(c[1], b[1])


def test_counts_near_the_ceiling():
    """Objects holding one token 40 times, queried with weight 51: counts up
    to 2040, the largest the eleven count planes admit (bound 40 x 51)."""
    rp, vals = genie_amd.synth.sets(461, 30_000, 5, 40, 2_000, 1.05)
    heavy = np.arange(100, 30_000, 997)
    lens = np.diff(rp)
    lists = np.split(vals, rp[1:-1])
    for o in heavy:
        lists[o] = np.full(int(lens[o]), 7, np.int32)   # token 7, as often as the list is long
    vals2 = np.concatenate(lists).astype(np.int32)
    r = W.set_ranges(np.array([0, 1, 3, 6], np.uint32), np.array([7, 7, 9, 11, 3, 7], np.int32),
                     np.array([51, 1, 2, 3, 20, 1], np.float32))
    t = Table.from_sets(rp, vals2)
    (b, p) = run(t, t.compile(r, 3), 3, 50)
    assert p.dense_active()
    ri, rc = oracle_sets(rp, vals2, r, 3, 50)
    assert rc.max() > 1500
    assert_topk_equal(b[0], b[1], ri, rc, 50, 3, "ceiling")


def test_set_form_not_used_where_it_cannot_be_exact():
    rp, vals, ranges = sets_case(451, 50_000, 10, 5_000)
    t = Table.from_sets(rp, vals)

    def declined(r, nq, k, where):
        (b, p) = run(t, t.compile(r, nq), nq, k)
        assert not p.dense_active(), where
        ri, rc = oracle_sets(rp, vals, r, nq, k)
        assert_topk_equal(b[0], b[1], ri, rc, k, nq, where)

    # a weight above 63
    r = ranges.copy()
    r["weight"][0] = 64
    declined(r, 10, 20, "weight 64")
    # two items of one token whose weights add up past 63
    r = ranges.copy()
    r = np.concatenate([r, r[:1]])
    r["weight"][0] = 40
    r["weight"][-1] = 30
    r = r[np.argsort(r["query"], kind="stable")]
    declined(r, 10, 20, "summed weight 70")
    # the longest list (40) times a weight of 52 reaches 2080
    r = ranges.copy()
    q3 = np.flatnonzero(r["query"] == 3)
    r["weight"][q3] = 0
    r["weight"][q3[0]] = 52
    declined(r, 10, 20, "bound")
    # a group naming more keys than its table holds
    wide = ranges[:3].copy()
    wide["query"] = 0
    wide["low"] = 0
    wide["up"] = 4_999
    declined(wide, 1, 20, "wide ranges")
    # objects with more than 255 occurrences: no forward lists
    rp2 = np.array([0, 300, 310], np.int64)
    vals2 = np.concatenate([np.arange(300) % 50, np.arange(10)]).astype(np.int32)
    t2 = Table.from_sets(rp2, vals2)
    r3 = W.set_ranges(np.array([0, 2], np.uint32), np.array([3, 7], np.int32), np.ones(2, np.float32))
    (b, p) = run(t2, t2.compile(r3, 1), 1, 2)
    assert not p.dense_active()
    ri, rc = oracle_sets(rp2, vals2, r3, 1, 2)
    assert_topk_equal(b[0], b[1], ri, rc, 2, 1, "long lists")


def test_half_size_workgroups_give_the_same_results():
    """The dense and set kernels at 512 threads per workgroup (taken where a
    runtime places no 1,024-thread workgroup of their size) against the
    oracle and the automatic choice, graph replays included."""
    from helpers import oracle_topk_batch as otb
    w = W.cfg1(seed=471, n=200_000, d=128, q=40, k=100)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 40)
    q_of = w["ranges"]["query"]
    ri, rc = otb(lambda q: O.count_exhaustive_single(w["rows"], w["ranges"][q_of == q]), 40, 100)
    rp, vals, ranges = sets_case(472, 80_000, 36, 12_000)
    ts = Table.from_sets(rp, vals)
    sitems = ts.compile(ranges, 36)
    si, sc = oracle_sets(rp, vals, ranges, 36, 50)
    for table, its, nq, k, ref in [(t, items, 40, 100, (ri, rc)), (ts, sitems, 36, 50, (si, sc))]:
        got = {}
        for nt in (0, 512):
            p = Plan(table.device(), nq, len(its), k, 0, mode=1)
            p.set_dense_threads(nt)
            p.upload(its, nq, np.full(nq, k, np.uint32))
            assert p.dense_active()
            p.set_graph(True)
            for _ in range(3):
                p.run()
                got[nt] = p.download()
                match_topk(table.device(), its[:50], nq, 5)
            assert_topk_equal(got[nt][0], got[nt][1], ref[0], ref[1], k, nq, f"{nt} threads")
        assert np.array_equal(got[0][0], got[512][0]) and np.array_equal(got[0][1], got[512][1])


def test_empty_documents_queries_and_groups():
    """Documents without tokens, queries without items or with zero weights
    only, a whole group of 32 queries naming no key, and k above the matches."""
    rp, vals = genie_amd.synth.sets(481, 40_000, 0, 12, 3_000, 1.05)
    assert (np.diff(rp) == 0).any()
    qp, qv = genie_amd.synth.sets(482, 70, 0, 20, 3_000, 1.05)
    wts = genie_amd.synth.uniform_int(483, 0, len(qv), 1, 5).astype(np.float32)
    r = W.set_ranges(qp, qv, wts)
    r["weight"][r["query"] == 5] = 0          # a query whose items all weigh 0
    r = r[(r["query"] < 32) | (r["query"] >= 64)]  # queries 32-63: no items at all
    t = Table.from_sets(rp, vals)
    kq = np.array([5, 3000, 40] * 23 + [7], np.uint32)
    (b, p) = run(t, t.compile(r, 70), 70, kq)
    assert p.dense_active()
    ri, rc = oracle_sets(rp, vals, r, 70, kq)
    assert_topk_equal(b[0], b[1], ri, rc, kq, 70, "empty cases")
