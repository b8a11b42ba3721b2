"""The workloads at their full sizes against the oracle:

  1M x 128 Zipf(1.1), 1,024 queries, +-2, k = 100 (the headline batch): the
      whole batch, dense and plain forms, uncompressed and compressed; and
      with integer weights 1-4 and 1-15 (the dense form's weighted masks);
  4M x 64 LSH (8,192 buckets), 4,096 exact-match queries, k = 100:
      uncompressed and SERIAL_DELTA_COMP_BP32_VARINT, a fixed sample of 256
      queries against the oracle and compressed = uncompressed over the whole
      batch;
  2M set-valued documents, 16,384 weighted queries with mixed k: the set and
      plain forms, a fixed sample of 256 queries against the oracle's
      index-based search and 16 against its exhaustive count;
  10M x 128 (built like the headline rows) compressed SERIAL_DELTA_BP32 and
      resident, served: a Poisson trace of mixed requests (ranges 0/+-2/+-5,
      k 10/100/1000) in the mix = alone = the oracle on a fixed sample of 256
      requests.

Sizes make these the slowest GPU tests (about a minute each, mostly host
index builds and the oracle)."""
import time

import numpy as np
import pytest

from helpers import (assert_topk_equal, oracle_search_sample, oracle_topk_batch, posting_items, rows_of)
from oracle import oracle as O

import genie_amd
from genie_amd import Plan, Server, Table, match_topk, workloads as W

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _run(index, items, nq, kq, mode, **sw):
    p = Plan(index, nq, len(items), int(np.max(kq)), 0, mode=mode)
    for k, v in sw.items():
        getattr(p, k)(v)
    p.upload(items, nq, kq)
    p.run()
    out = p.download()
    form = p.form()
    p.close()
    return out, form


def test_headline_batch_whole_against_the_oracle():
    t0 = time.time()
    w = W.cfg1()
    nq = w["n_queries"]
    kq = np.full(nq, w["k"], np.uint32)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], nq)
    srt = items[np.argsort(items["query"], kind="stable")]
    ri, rc, _ = O.search_mt(t.export()["inv"].view(np.uint32), srt, w["rows"].shape[0], nq, w["k"])
    _log(f"1M x 128: built and oracle in {time.time() - t0:.1f}s")
    ix = t.device()
    for mode in (1, 0):
        (ids, cnt), form = _run(ix, items, nq, kq, mode)
        assert form == ("dense" if mode == 1 else "plain")
        assert np.array_equal(ids, ri) and np.array_equal(cnt, rc), form
    c = Table.from_rows(w["rows"], 1024, True, codec=11)
    citems = c.compile(w["ranges"], nq, merge_runs=True)
    cix = c.device()
    for mode in (0, 1):
        (ids, cnt), form = _run(cix, citems, nq, kq, mode)
        assert np.array_equal(ids, ri) and np.array_equal(cnt, rc), f"compressed {form}"


@pytest.mark.parametrize("wmax", [4, 15])
def test_headline_batch_weighted_whole_against_the_oracle(wmax):
    t0 = time.time()
    w = W.cfg1()
    nq = w["n_queries"]
    kq = np.full(nq, w["k"], np.uint32)
    r = w["ranges"].copy()
    r["weight"] = genie_amd.synth.uniform_int(1011 + wmax, 0, len(r), 1, wmax)
    t = Table.from_rows(w["rows"])
    items = t.compile(r, nq)
    srt = items[np.argsort(items["query"], kind="stable")]
    ri, rc, _ = O.search_mt(t.export()["inv"].view(np.uint32), srt, w["rows"].shape[0], nq, w["k"])
    _log(f"1M x 128 weights 1-{wmax}: built and oracle in {time.time() - t0:.1f}s")
    ix = t.device()
    for mode in (1, 0):
        (ids, cnt), form = _run(ix, items, nq, kq, mode)
        assert form == ("dense" if mode == 1 else "plain")
        assert np.array_equal(ids, ri) and np.array_equal(cnt, rc), form


def test_lsh_4m_full_batch():
    t0 = time.time()
    w = W.cfg2()
    nq = w["n_queries"]
    kq = np.full(nq, w["k"], np.uint32)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], nq)
    qs = np.unique(np.linspace(0, nq - 1, 256).astype(np.int64))
    ri, rc = oracle_search_sample(t.export()["inv"].view(np.uint32), items, w["rows"].shape[0], w["k"], qs)
    _log(f"4M x 64: built and oracle sample in {time.time() - t0:.1f}s")
    ix = t.device()
    (ua, ub), form = _run(ix, items, nq, kq, 0)
    gi, gc = rows_of(ua, ub, w["k"], nq, qs)
    assert_topk_equal(gi, gc, ri, rc, w["k"], len(qs), "LSH uncompressed plain")
    (ba, bb), form = _run(ix, items, nq, kq, 1, set_form="plain")
    assert np.array_equal(ba, ua) and np.array_equal(bb, ub)
    del ix
    c = Table.from_rows(w["rows"], 1024, True, codec=13)
    cix = c.device()
    for merged in (False, True):
        citems = c.compile(w["ranges"], nq, merge_runs=merged)
        for dm in (-1, 0):
            (ca, cb), _ = _run(cix, citems, nq, kq, 0, set_decode=dm)
            # compressed = uncompressed for every query and slot
            assert np.array_equal(ca, ua) and np.array_equal(cb, ub), f"compressed, merged {merged}, decode {dm}"


def test_sets_2m_full_batch():
    t0 = time.time()
    w = W.cfg3()
    nq, kq = w["n_queries"], w["k"]
    rp, vals = w["row_ptr"], w["vals"]
    t = Table.from_sets(rp, vals)
    items = t.compile(w["ranges"], nq)
    qs = np.unique(np.linspace(0, nq - 1, 256).astype(np.int64))
    ri, rc = oracle_search_sample(t.export()["inv"].view(np.uint32), items, len(rp) - 1, kq, qs)
    # the index-based oracle against the closed form on part of the sample
    q_of = w["ranges"]["query"]
    ei, ec = oracle_topk_batch(lambda j: O.count_exhaustive_biject(rp, vals, w["ranges"][q_of == qs[j]]), 16,
                               kq[qs[:16]])
    li, lc = rows_of(ri, rc, kq[qs], len(qs), np.arange(16))
    assert np.array_equal(ei, li) and np.array_equal(ec, lc)
    _log(f"2M sets: built and oracle sample in {time.time() - t0:.1f}s")
    ix = t.device()
    for mode in (1, 0):
        (ids, cnt), form = _run(ix, items, nq, kq, mode)
        assert form == ("set" if mode == 1 else "plain")
        gi, gc = rows_of(ids, cnt, kq, nq, qs)
        assert_topk_equal(gi, gc, ri, rc, kq[qs], len(qs), f"sets {form}")


def test_serving_10m_compressed_index():
    """The serving workload at its stated size: every sampled request's
    result in a Poisson mix equals its result alone and the oracle's."""
    t0 = time.time()
    n = 10_000_000
    rows = genie_amd.synth.rows(1005, n, 128, genie_amd.synth.ZIPF_PERM, 1.1, 0.0, 0, 255, 256)
    t = Table.from_rows(rows, 1024, True, codec=11)
    nreq = 1536
    sq, _ = genie_amd.synth.queries(7101, rows, nreq, noise=3, lo=0, hi=255)
    radius = np.array([0, 2, 5], np.int32)[genie_amd.synth.uniform_int(7102, 0, nreq, 0, 2)]
    k = np.array([10, 100, 1000], np.uint32)[genie_amd.synth.uniform_int(7103, 1, nreq, 0, 2)]
    ranges = W.single_ranges(sq, radius)
    items = t.compile(ranges, nreq, merge_runs=True)
    ptr = np.searchsorted(items["query"], np.arange(nreq + 1)).astype(np.uint64)
    del rows
    ix = t.device()
    _log(f"10M x 128 compressed: built and resident ({ix.bytes() / 2**30:.2f} GiB) in {time.time() - t0:.1f}s")
    srv = Server(ix, executors=4, max_batch=256, max_k=1000, mode=2)
    arr = np.cumsum(genie_amd.synth.exponential(7104, nreq, 10000.0))
    lat, ids, cnt, wall = srv.replay(items, ptr, k, arr)
    st = srv.stats()
    assert st["max_inflight"] > 1 and st["batches"] < nreq
    qs = np.unique(np.linspace(0, nreq - 1, 256).astype(np.int64))
    gi, gc = rows_of(ids, cnt, k, nreq, qs)
    # alone: one request at a time
    ai, ac = [], []
    for r in qs:
        tk = srv.submit(items[ptr[r]:ptr[r + 1]], int(k[r]))
        a, c, _ = srv.wait(tk, int(k[r]))
        ai.append(a)
        ac.append(c)
    srv.close()
    assert np.array_equal(np.concatenate(ai), gi) and np.array_equal(np.concatenate(ac), gc)
    _log(f"served {nreq} at 10K QPS offered ({nreq / wall:.0f} achieved), p99 {np.percentile(lat, 99):.2f} ms")
    t1 = time.time()
    ex = t.export()
    ri, rc = oracle_search_sample(ex["inv"].view(np.uint32), posting_items(ex, items), n, k, qs)
    _log(f"oracle on {len(qs)} requests in {time.time() - t1:.1f}s")
    assert_topk_equal(gi, gc, ri, rc, k[qs], len(qs), "served 10M")
