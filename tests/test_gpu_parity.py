"""Parity of the HIP engine (through the C ABI) with the oracle, on the GPU.
Counts are bit-exact; ids identical (both sides order count desc, id asc)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_csv
from helpers import assert_topk_equal, split_rows
from oracle import oracle as O

import genie_amd
from genie_amd import DIM_DTYPE, RANGE_DTYPE, GenieError, Plan, Table, match_topk, workloads as W

pytestmark = pytest.mark.gpu


def oracle_single(rows, ranges, n_queries, k):
    res_i, res_c = [], []
    for q in range(n_queries):
        kq = int(k[q]) if np.ndim(k) else int(k)
        i, c = O.topk(O.count_exhaustive_single(rows, ranges[ranges["query"] == q]), kq)
        res_i.append(i)
        res_c.append(c)
    return np.concatenate(res_i), np.concatenate(res_c)


def oracle_sets(row_ptr, vals, ranges, n_queries, k):
    res_i, res_c = [], []
    for q in range(n_queries):
        kq = int(k[q]) if np.ndim(k) else int(k)
        i, c = O.topk(O.count_exhaustive_biject(row_ptr, vals, ranges[ranges["query"] == q]), kq)
        res_i.append(i)
        res_c.append(c)
    return np.concatenate(res_i), np.concatenate(res_c)


PLAIN = {"plain-sparse": (0, 1), "plain-dense": (0, 0), "plain-hash": (1, -1)}  # (hash form, sparse select)


def engine(table, items, n_queries, k, mode):
    """mode "auto": genie_hip_match_topk (dense form when the batch allows,
    else plain); "plain-sparse" / "plain-dense": an explicit plain plan on
    the direct-mapped kernels with the sparse select forced on / off;
    "plain-hash": every query on the hash form; so each plain kernel and
    select is checked whatever the batch; other values: a plan of that
    mode."""
    if mode == "auto":
        return match_topk(table.device(), items, n_queries, k)
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (n_queries,)))
    pm = 0 if mode in PLAIN else mode
    p = Plan(table.device(), n_queries, max(1, len(items)), int(kq.max()), 0, mode=pm)
    if mode in PLAIN:
        p.set_hash(PLAIN[mode][0])
        p.set_sparse(PLAIN[mode][1])
    p.upload(items, n_queries, kq)
    p.run()
    return p.download()


def run_engine(table, ranges, n_queries, k, lb=False, max_load=0, shuffle=False, mode="auto"):
    items = table.compile(ranges, n_queries, lb, max_load)
    if shuffle:
        items = items[np.random.default_rng(0).permutation(len(items))]
    return engine(table, items, n_queries, k, mode)


MODES = ["plain-sparse", "plain-dense", "plain-hash", "auto"]


@pytest.mark.parametrize("mode", MODES)
def test_cfg0_plain(mode):
    w = W.cfg0()
    t = Table.from_rows(w["rows"])
    ids, cnt = run_engine(t, w["ranges"], w["n_queries"], w["k"], mode=mode)
    ri, rc = oracle_single(w["rows"], w["ranges"], w["n_queries"], w["k"])
    assert_topk_equal(ids, cnt, ri, rc, w["k"], w["n_queries"], "cfg0")
    # every query is a copy of a row with U{-3..3} noise per dim, so the row
    # matches about 128/7 dims exactly and is ranked first almost always
    assert np.mean(ids.reshape(w["n_queries"], -1)[:, 0] == w["src"]) > 0.9


def test_sift20_goldens_through_legacy_knn():
    rows = np.array([r[:5] for r in load_csv("sift_20.csv")], np.int32)
    t = Table.from_rows(rows)
    ranges = W.single_ranges(rows[:3], 0)
    ids, cnt = t.knn(ranges, 3, 5)
    assert (ids[0], cnt[0]) == (0, 5) and (ids[1], cnt[1]) == (4, 2)
    assert (ids[5], cnt[5]) == (1, 5) and (ids[10], cnt[10]) == (2, 5)
    ri, rc = oracle_single(rows, ranges, 3, 5)
    assert_topk_equal(ids, cnt, ri, rc, 5, 3, "sift20")


def test_tweets20_exact_counts():
    rows = load_csv("tweets_20.csv")
    rp, vals = split_rows(rows)
    t = Table.from_sets(rp, vals)
    qp, qv = split_rows(rows)
    ranges = W.set_ranges(qp, qv, np.ones(len(qv), np.float32))
    ids, cnt = t.knn(ranges, 20, 5)
    assert sorted(ids[:2]) == [0, 11] and list(cnt[:2]) == [17, 17]
    assert (ids[5], cnt[5]) == (1, 14) and (ids[10], cnt[10]) == (2, 18)
    ri, rc = oracle_sets(rp, vals, ranges, 20, 5)
    assert_topk_equal(ids, cnt, ri, rc, 5, 20, "tweets20")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [70_000, 200_003])
def test_many_tiles_zipf_ranges(n, mode):
    w = W.cfg1(n=n, d=32, q=24, k=100)
    t = Table.from_rows(w["rows"])
    ids, cnt = run_engine(t, w["ranges"], w["n_queries"], w["k"], mode=mode)
    ri, rc = oracle_single(w["rows"], w["ranges"], w["n_queries"], w["k"])
    assert_topk_equal(ids, cnt, ri, rc, w["k"], w["n_queries"], f"zipf n={n}")


def test_weights_widen_count_lanes():
    # bound 32*200 = 6400 > 255 -> 16-bit lanes; 32 * 3000 > 65535 -> 32-bit lanes
    w = W.cfg1(n=90_000, d=32, q=8, k=50)
    t = Table.from_rows(w["rows"])
    for wmax in (200, 3000):
        r = w["ranges"].copy()
        r["weight"] = np.random.default_rng(wmax).integers(1, wmax + 1, len(r))
        ids, cnt = run_engine(t, r, w["n_queries"], w["k"])
        ri, rc = oracle_single(w["rows"], r, w["n_queries"], w["k"])
        assert_topk_equal(ids, cnt, ri, rc, w["k"], w["n_queries"], f"weights<={wmax}")


@pytest.mark.parametrize("mode", MODES)
def test_edge_cases_padding_empty_and_k(mode):
    w = W.cfg0(n=3000, d=16, q=6)
    t = Table.from_rows(w["rows"])
    r = w["ranges"]
    keep = (r["query"] != 2)                      # query 2 has no ranges at all
    r = r[keep].copy()
    r = r[~((r["query"] == 3) & (r["dim"] > 0))]  # query 3: one range -> fewer matches than k
    m4 = r["query"] == 4
    r["low"][m4] = 1000                           # query 4: disjoint ranges only
    r["up"][m4] = 2000
    for k in (1, 7, 300, 4096):
        ids, cnt = run_engine(t, r, 6, k, mode=mode)
        ri, rc = oracle_single(w["rows"], r, 6, k)
        assert_topk_equal(ids, cnt, ri, rc, k, 6, f"edge k={k}")
        c = cnt.reshape(6, k)
        assert np.all(c[2] == 0) and np.all(c[4] == 0)
    # per-query k
    kq = np.array([1, 10, 100, 1000, 5, 4096], np.uint32)
    ids, cnt = run_engine(t, r, 6, kq, mode=mode)
    ri, rc = oracle_single(w["rows"], r, 6, kq)
    assert_topk_equal(ids, cnt, ri, rc, kq, 6, "per-query k")


def test_load_balanced_and_unsorted_items():
    w = W.cfg1(n=120_000, d=16, q=10, k=64)
    t = Table.from_rows(w["rows"])
    ri, rc = oracle_single(w["rows"], w["ranges"], w["n_queries"], w["k"])
    ids, cnt = run_engine(t, w["ranges"], w["n_queries"], w["k"], lb=True, max_load=777, shuffle=True)
    assert_topk_equal(ids, cnt, ri, rc, w["k"], w["n_queries"], "load balanced")


def test_bijectmap_duplicates_and_weights():
    rp, vals = genie_amd.synth.sets(31, 50_000, 5, 40, 3000, 1.05)
    qp, qv = genie_amd.synth.sets(32, 40, 1, 64, 3000, 1.05)
    wts = genie_amd.synth.uniform_int(33, 0, len(qv), 1, 4)
    ranges = W.set_ranges(qp, qv, wts)
    t = Table.from_sets(rp, vals)
    kq = np.array([10, 100, 1000, 7] * 10, np.uint32)
    ids, cnt = run_engine(t, ranges, 40, kq)
    ri, rc = oracle_sets(rp, vals, ranges, 40, kq)
    assert_topk_equal(ids, cnt, ri, rc, kq, 40, "biject")


def test_plan_reuse_matches_one_shot():
    w = W.cfg1(n=80_000, d=24, q=16, k=32)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], w["n_queries"])
    one = match_topk(t.device(), items, w["n_queries"], w["k"])
    p = Plan(t.device(), 64, len(items) + 100, 128)
    p.upload(items, w["n_queries"], w["k"])
    for _ in range(3):
        p.run()
        got = p.download()
        assert np.array_equal(got[0], one[0]) and np.array_equal(got[1], one[1])
    # a smaller batch on the same plan
    sub = items[items["query"] < 5]
    p.upload(sub, 5, 32)
    p.run()
    got = p.download()
    assert np.array_equal(got[0], one[0][:5 * 32]) and np.array_equal(got[1], one[1][:5 * 32])


def test_invalid_inputs_are_rejected():
    w = W.cfg0(n=500, d=8, q=2)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 2)
    bad = items.copy()
    bad["weight"][0] = 1.5
    with pytest.raises(GenieError):
        match_topk(t.device(), bad, 2, 5)
    bad = items.copy()
    bad["end_pos"][0] = 10**9
    with pytest.raises(GenieError):
        match_topk(t.device(), bad, 2, 5)
    with pytest.raises(GenieError):
        match_topk(t.device(), items, 2, 0)
    with pytest.raises(GenieError):
        match_topk(t.device(), items, 1, 5)  # item refers to query 1 of 1


def test_cpp_api_reference_tests():
    exe = os.path.join(ROOT, "genie_amd", "lib", "test_api")
    r = subprocess.run([exe, GOLDEN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


def test_loaded_tables_answer_the_same(tmp_path):
    """SaveTable / LoadTable: the reloaded table's device index gives the same
    results as the original, for rows, sets and a compressed table."""
    w = W.cfg1(n=60_000, d=16, q=12, k=30)
    rp, vals = genie_amd.synth.sets(71, 20_000, 5, 40, 2000, 1.05)
    qp, qv = genie_amd.synth.sets(72, 6, 1, 20, 2000, 1.05)
    sr = W.set_ranges(qp, qv, np.ones(len(qv), np.float32))
    cases = [(Table.from_rows(w["rows"]), w["ranges"], 12),
             (Table.from_rows(w["rows"], 1024, True, codec=13), w["ranges"], 12),
             (Table.from_sets(rp, vals), sr, 6)]
    for i, (t, ranges, nq) in enumerate(cases):
        path = str(tmp_path / f"t{i}.gix")
        t.save(path)
        u = Table.load(path)
        a = match_topk(t.device(), t.compile(ranges, nq), nq, 30)
        b = match_topk(u.device(), u.compile(ranges, nq), nq, 30)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), i


@pytest.mark.parametrize("n,sparse", [(50_000, 0), (50_000, 1), (300_000, 0), (300_000, 1)])
def test_adaptive_threshold_closed_form(n, sparse):
    """updateThreshold's closed form on device output (match_device_utils.h:
    449-473): the final threshold of a query is the largest k-th count of any
    of its tiles that holds k objects counting at least 1.  With the whole
    index in one tile (50,000 objects, tiles of 65,536) that is the query's
    k-th largest count c_(k) -- the reference's thr - 1 = c_(k) -- and 0 for
    a query with fewer than k matching objects."""
    from genie_amd import Plan
    w = W.cfg1(seed=977, n=n, d=20, q=24, k=60)
    r = w["ranges"].copy()
    r = r[~((r["query"] == 5) & (r["dim"] > 0))].copy()  # query 5: one exact value -> few matches
    r["low"][r["query"] == 5] = r["up"][r["query"] == 5]
    t = Table.from_rows(w["rows"])
    items = t.compile(r, 24)
    p = Plan(t.device(), 24, len(items), 60, 0, mode=0)
    p.set_hash(0)
    p.set_sparse(sparse)
    p.upload(items, 24, np.full(24, 60, np.uint32))
    p.run()
    ids, cnt = p.download()
    thr = p.thresholds(24)
    tile = 65536  # direct-mapped tiles at 8-bit counters
    for q in range(24):
        c = O.count_exhaustive_single(w["rows"], r[r["query"] == q]).astype(np.int64)
        want = 0
        for t0 in range(0, n, tile):
            ct = np.sort(c[t0:t0 + tile])[::-1]
            if len(ct) >= 60 and ct[59] >= 1:
                want = max(want, int(ct[59]))
        assert thr[q] == want, (q, thr[q], want)
        if n <= tile:
            ck = int(np.sort(c)[::-1][59])
            assert thr[q] == (ck if ck >= 1 else 0)
            assert thr[q] == (cnt[q * 60 + 59] if np.count_nonzero(c) >= 60 else 0)
    assert thr[5] == 0 or np.count_nonzero(O.count_exhaustive_single(w["rows"], r[r["query"] == 5])) >= 60
