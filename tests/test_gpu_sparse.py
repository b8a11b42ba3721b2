"""Sparse select of the plain form (counter words touched by a (query, tile)
marked during counting; the per-tile top-k visits only those) against the
exhaustive oracle and the dense select: identical results on selective
batches, on ties trimmed at the tile's k-th count, on a busy query inside a
sparse batch (its tiles fall back to the dense select), on compressed and
set-valued tables."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

import genie_amd
from genie_amd import Plan, Table, workloads as W
from genie_amd._native import RANGE_DTYPE

pytestmark = pytest.mark.gpu


def plain(table, items, nq, k, sparse):
    """The direct-mapped kernels (hash form off) with the given select."""
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
    p = Plan(table.device(), nq, max(1, len(items)), int(kq.max()), 0, mode=0)
    p.set_hash(0)
    p.set_sparse(sparse)
    p.upload(items, nq, kq)
    p.run()
    return p.download(), p


def oracle_single(rows, ranges, nq, k):
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, ranges[q_of == q]), nq, k)


def lsh(n, d, q, buckets, seed, k):
    """Exact-match queries over many buckets per dim: a few hundred postings
    per query and tile, most objects of a tile at count 0 or 1."""
    rows = genie_amd.synth.rows(seed, n, d, genie_amd.synth.UNIFORM, 0.0, 0.0, 0, buckets - 1, buckets)
    qv, _ = genie_amd.synth.queries(seed + 1, rows, q, redraw_p=0.3, n_values=buckets)
    return rows, W.single_ranges(qv, 0)


@pytest.mark.parametrize("n,d,q,buckets,k", [
    (300_000, 32, 48, 2048, 100),    # ties at count 1 trimmed in most tiles
    (200_003, 16, 20, 8192, 1000),   # k above most tiles' matches: every match emitted
    (70_000, 64, 9, 2048, 7),
])
def test_sparse_equals_oracle(n, d, q, buckets, k):
    rows, r = lsh(n, d, q, buckets, 101 + d, k)
    t = Table.from_rows(rows)
    items = t.compile(r, q)
    (b, p) = plain(t, items, q, k, -1)
    assert p.sparse_active()
    ri, rc = oracle_single(rows, r, q, k)
    assert_topk_equal(b[0], b[1], ri, rc, k, q, f"sparse n={n} d={d}")
    (c, p) = plain(t, items, q, k, 0)
    assert not p.sparse_active()
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


def test_busy_query_in_a_sparse_batch():
    """Query 0 covers whole dims (every object of a tile counted): its tiles
    touch every counter word and take the dense select; the others stay
    sparse."""
    rows, r = lsh(150_000, 16, 12, 4096, 131, 50)
    r = r.copy()
    m0 = r["query"] == 0
    r["low"][m0] = 0
    r["up"][m0] = 4095
    t = Table.from_rows(rows)
    items = t.compile(r, 12)
    (b, p) = plain(t, items, 12, 50, 1)
    assert p.sparse_active()
    ri, rc = oracle_single(rows, r, 12, 50)
    assert_topk_equal(b[0], b[1], ri, rc, 50, 12, "busy query")


def test_sparse_weights_and_ties():
    """Integer weights (16-bit count lanes) and many equal counts."""
    rows, r = lsh(120_000, 24, 16, 1024, 141, 64)
    r = r.copy()
    r["weight"] = np.random.default_rng(5).integers(1, 400, len(r))
    t = Table.from_rows(rows)
    items = t.compile(r, 16)
    (b, p) = plain(t, items, 16, 64, 1)
    ri, rc = oracle_single(rows, r, 16, 64)
    assert_topk_equal(b[0], b[1], ri, rc, 64, 16, "sparse weights")


@pytest.mark.parametrize("codec", [3, 11, 13])
def test_sparse_compressed(codec):
    rows, r = lsh(160_000, 16, 12, 2048, 151, 40)
    t = Table.from_rows(rows, 1024, True, codec=codec)
    items = t.compile(r, 12)
    (b, p) = plain(t, items, 12, 40, -1)
    assert p.sparse_active()
    ri, rc = oracle_single(rows, r, 12, 40)
    assert_topk_equal(b[0], b[1], ri, rc, 40, 12, f"sparse codec {codec}")


def test_sparse_sets():
    rp, vals = genie_amd.synth.sets(161, 120_000, 5, 40, 50_000, 1.05)
    qp, qv = genie_amd.synth.sets(162, 24, 1, 30, 50_000, 1.05)
    wts = genie_amd.synth.uniform_int(163, 0, len(qv), 1, 4)
    ranges = W.set_ranges(qp, qv, wts)
    t = Table.from_sets(rp, vals)
    items = t.compile(ranges, 24)
    kq = np.array([10, 100, 1000] * 8, np.uint32)
    (b, p) = plain(t, items, 24, kq, 1)
    q_of = ranges["query"]
    ri, rc = oracle_topk_batch(lambda q: O.count_exhaustive_biject(rp, vals, ranges[q_of == q]), 24, kq)
    assert_topk_equal(b[0], b[1], ri, rc, kq, 24, "sparse sets")
    (c, _) = plain(t, items, 24, kq, 0)
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


def test_set_sparse_rejects_bad_mode():
    w = W.cfg0(n=500, d=8, q=2)
    t = Table.from_rows(w["rows"])
    p = Plan(t.device(), 2, 100, 5, 0, mode=0)
    with pytest.raises(genie_amd.GenieError):
        p.set_sparse(2)
