"""Pins the CPU oracle (oracle/) before anything is checked against it:
golden vectors of the reference's own tests, the reference's own inv_list
compiled from its sources (oracle/_ref, when the reference tree is present),
codec known-answer vectors from the format definitions, round trips, and
closed forms / brute force for counting and top-k."""
import os

import numpy as np
import pytest

from conftest import ROOT, load_csv
from helpers import split_rows
from oracle import oracle as O

RANGE = O.RANGE_DTYPE


def sift20():
    return np.array([r[:5] for r in load_csv("sift_20.csv")], np.int32)


def ranges_exact(qrows, dims=None, radius=0, biject=False):
    out = []
    for qi, row in enumerate(qrows):
        for j, v in enumerate(row):
            if dims is not None and j >= dims:
                break
            out.append((qi, 0 if biject else j, v - radius, v + radius, 1.0, j))
    return np.array(out, RANGE)


# ---- index layout --------------------------------------------------------

def test_sift20_inv_golden():
    # test/test_csv_input_sift_data.cc:66-71
    ix = O.OracleIndex.single(sift20())
    assert list(ix.inv[:6]) == [8, 9, 7, 0, 2, 4]


def test_tweets20_inv_golden():
    # test/test_csv_input_tweets.cc:66-71
    rp, vals = split_rows(load_csv("tweets_20.csv"))
    ix = O.OracleIndex.biject(rp, vals)
    assert list(ix.inv[:6]) == [0, 11, 0, 11, 0, 11]


def test_tweets20_multiload_table0_golden():
    # test/test_csv_multiload_tweets.cc:68-73: first table = rows 0..9
    rp, vals = split_rows(load_csv("tweets_20.csv")[:10])
    ix = O.OracleIndex.biject(rp, vals)
    assert list(ix.inv[:6]) == [0, 0, 0, 0, 0, 7]


@pytest.mark.parametrize("lb,maxlen", [(False, 0), (True, 3), (True, 7)])
def test_layout_closed_form(lb, maxlen):
    rng = np.random.default_rng(5)
    rows = rng.integers(-3, 9, size=(200, 6)).astype(np.int32)
    rows[:, 2] = 4  # a single-valued column
    ix = O.OracleIndex.single(rows, maxlen, lb)
    assert len(ix.inv) == rows.size                      # sum of list sizes = N * D
    assert np.all(np.diff(ix.inv_pos) > 0)               # segments non-empty, increasing
    for d in range(rows.shape[1]):
        lo, hi = ix.lo[d], ix.hi[d]
        assert lo == rows[:, d].min() and hi == rows[:, d].max()
        ks = ix.keyseg[ix.keyseg_off[d]:ix.keyseg_off[d + 1]]
        assert len(ks) == hi - lo + 2
        for v in range(lo, hi + 1):
            a, b = ix.inv_pos[ks[v - lo]], ix.inv_pos[ks[v - lo + 1]]
            want = np.nonzero(rows[:, d] == v)[0]
            assert np.array_equal(ix.inv[a:b], want)    # ascending ids of value v
            if lb:
                segs = ix.inv_pos[ks[v - lo]:ks[v - lo + 1] + 1]
                assert np.all(np.diff(segs) <= maxlen)


def _ref_lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libref_invlist.so")
    if not os.path.exists(path):
        pytest.skip("reference inv_list not built here (oracle/_ref needs the reference tree)")
    import ctypes as C
    lib = C.CDLL(path)
    return lib


def test_inversion_matches_reference_inv_list():
    import ctypes as C
    lib = _ref_lib()
    rng = np.random.default_rng(11)
    rows = np.concatenate([sift20(), rng.integers(0, 40, size=(20, 5)).astype(np.int32)])
    ix = O.OracleIndex.single(rows)
    for d in range(rows.shape[1]):
        col = np.ascontiguousarray(rows[:, d])
        lo, hi = C.c_int32(), C.c_int32()
        lens = np.zeros(256, np.int32)
        ids = np.zeros(len(col), np.int32)
        rc = lib.ref_invert(col.ctypes.data_as(C.c_void_p), len(col), C.byref(lo), C.byref(hi),
                            lens.ctypes.data_as(C.c_void_p), C.c_int64(256), ids.ctypes.data_as(C.c_void_p),
                            C.c_int64(len(ids)))
        assert rc == 0
        assert (lo.value, hi.value) == (ix.lo[d], ix.hi[d])
        ks = ix.keyseg[ix.keyseg_off[d]:ix.keyseg_off[d + 1]]
        a, b = ix.inv_pos[ks[0]], ix.inv_pos[ks[-1]]
        assert np.array_equal(ix.inv[a:b], ids)


def test_biject_inversion_matches_reference_inv_list():
    import ctypes as C
    lib = _ref_lib()
    rp, vals = split_rows(load_csv("tweets_20.csv"))
    ix = O.OracleIndex.biject(rp, vals)
    lo, hi = C.c_int32(), C.c_int32()
    lens = np.zeros(4096, np.int32)
    ids = np.zeros(len(vals), np.int32)
    rc = lib.ref_invert_biject(rp.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p), len(rp) - 1,
                               C.byref(lo), C.byref(hi), lens.ctypes.data_as(C.c_void_p), C.c_int64(4096),
                               ids.ctypes.data_as(C.c_void_p), C.c_int64(len(ids)))
    assert rc == 0
    assert (lo.value, hi.value) == (ix.lo[0], ix.hi[0])
    assert np.array_equal(ix.inv, ids)


# ---- codecs ---------------------------------------------------------------

def test_varint_known_answers():
    # 7-bit groups, high bit on the LAST byte, zero padding to a word
    assert list(O.encode(4, [127]).view(np.uint8)) == [0xFF, 0, 0, 0]
    assert list(O.encode(4, [128]).view(np.uint8)) == [0x00, 0x81, 0, 0]
    assert list(O.encode(4, [300, 1]).view(np.uint8)) == [0x2C, 0x82, 0x81, 0]
    assert list(O.encode(4, [0xFFFFFFFF]).view(np.uint8)) == [0x7F, 0x7F, 0x7F, 0x7F, 0x8F, 0, 0, 0]


def test_bp32_known_answers():
    assert list(O.encode(3, [1] * 32)) == [32, 0x01000000, 0xFFFFFFFF]
    # widths 3 and 0 in one miniblock each; value j at bits [3j, 3j+3)
    vals = [5] * 32 + [0] * 32
    w = O.encode(3, vals)
    assert w[0] == 64 and w[1] == (3 << 24) | (0 << 16)
    bits = 0
    for j in range(32):
        bits |= 5 << (3 * j)
    assert list(w[2:5]) == [bits & 0xFFFFFFFF, (bits >> 32) & 0xFFFFFFFF, (bits >> 64) & 0xFFFFFFFF]
    # a full block of 128 then a 1-value tail
    w = O.encode(3, list(range(129)))
    assert w[0] == 129 and w[1] == (5 << 24) | (6 << 16) | (7 << 8) | 7


def test_delta_composite_serial_known_answers():
    assert list(O.encode(2, [5, 7, 7, 10])) == [5, 2, 0, 3]
    # composite: first floor(n/8)*8 values in BP32 after a length word, tail in copy
    vals = list(range(1, 11))
    w = O.encode(5, vals)
    bp = O.encode(3, vals[:8])
    assert w[0] == len(bp) and list(w[1:1 + len(bp)]) == list(bp) and list(w[1 + len(bp):]) == [9, 10]
    # serial: delta first
    assert list(O.encode(11, [3, 4, 6])) == list(O.encode(3, [3, 1, 2]))


def mixed_values(n, seed=666):
    """The value mix of test_compression_codecs.cu:180-204 (1-7, 8-14, 15-21,
    22-28 and 29-32 bit magnitudes), from numpy's generator."""
    rng = np.random.default_rng(seed)
    kind = rng.integers(0, 5, n)
    r = rng.integers(0, 2**31 - 1, n, dtype=np.int64)
    out = np.where(kind == 0, r % (1 << 7),
          np.where(kind == 1, (r + (1 << 7)) % (1 << 15),
          np.where(kind == 2, (r + (1 << 14)) % (1 << 22),
          np.where(kind == 3, (r + (1 << 21)) % (1 << 28), r + (1 << 28)))))
    return (out & 0xFFFFFFFF).astype(np.uint32)


@pytest.mark.parametrize("codec", range(1, 14))
def test_codec_round_trip(codec):
    vals = mixed_values(1024)
    for n in list(range(1, 140)) + list(range(140, 1025, 37)) + [1024]:
        w = O.encode(codec, vals[:n])
        assert np.array_equal(O.decode(codec, w), vals[:n]), (codec, n)
        srt = np.sort(vals[:n])
        assert np.array_equal(O.decode(codec, O.encode(codec, srt)), srt)


# ---- query compilation ------------------------------------------------------

def test_compile_closed_form():
    rng = np.random.default_rng(3)
    rows = rng.integers(0, 30, size=(300, 8)).astype(np.int32)
    ix = O.OracleIndex.single(rows)
    ranges = np.array([(0, 5, 3, 9, 2.0, 0), (0, 1, -4, 2, 1.0, 1), (0, 1, 28, 40, 1.0, 2),
                       (0, 2, 50, 60, 1.0, 3), (0, 9, 0, 5, 1.0, 4), (0, 3, 7, 6, 1.0, 5)], RANGE)
    items = ix.compile(ranges)
    # disjoint (2, 50..60), unknown dim 9 and low > up are dropped; ordered by dim
    assert len(items) == 3
    assert list(items["order"]) == [1, 2, 0]
    for it, (d, lo, up) in zip(items, [(1, -4, 2), (1, 28, 40), (5, 3, 9)]):
        got = np.sort(ix.inv[it["start_pos"]:it["end_pos"]])
        want = np.nonzero((rows[:, d] >= lo) & (rows[:, d] <= up))[0]
        assert np.array_equal(got, want)


def test_compile_load_balance_and_compressed():
    rng = np.random.default_rng(4)
    rows = rng.integers(0, 4, size=(1000, 3)).astype(np.int32)
    ranges = np.array([(0, 0, 1, 2, 1.0, 0), (0, 2, 0, 0, 1.0, 1)], RANGE)
    ix = O.OracleIndex.single(rows)
    plain = ix.compile(ranges)
    split = ix.compile(ranges, max_load=97)
    assert np.all(split["end_pos"] - split["start_pos"] <= 97)
    # splits tile each original span exactly (query.cc:141-191)
    for it in plain:
        parts = split[(split["start_pos"] >= it["start_pos"]) & (split["end_pos"] <= it["end_pos"])]
        assert parts["start_pos"][0] == it["start_pos"] and parts["end_pos"][-1] == it["end_pos"]
        assert np.all(parts["start_pos"][1:] == parts["end_pos"][:-1])
    ixl = O.OracleIndex.single(rows, 64, True)
    words, cpos = ixl.compress(13)
    citems = ixl.compile(ranges, compr_pos=cpos)
    # one item per segment; decoded ids are exactly the postings of the range
    for r in ranges:
        ids = []
        for it in citems[citems["order"] == r["order"]]:
            ids.append(O.decode(13, words[it["start_pos"]:it["end_pos"]]))
        got = np.sort(np.concatenate(ids))
        want = np.nonzero((rows[:, r["dim"]] >= r["low"]) & (rows[:, r["dim"]] <= r["up"]))[0]
        assert np.array_equal(got, want)


# ---- counting and top-k -----------------------------------------------------

def test_count_index_equals_exhaustive_single():
    rng = np.random.default_rng(7)
    rows = rng.integers(0, 12, size=(500, 10)).astype(np.int32)
    ix = O.OracleIndex.single(rows, 17, True)
    ranges = np.array([(0, d, int(v) - 1, int(v) + 1, float(w), d) for d, (v, w) in
                       enumerate(zip(rng.integers(0, 12, 10), rng.integers(1, 5, 10)))], RANGE)
    ranges = np.concatenate([ranges, np.array([(0, 3, 0, 20, 3.0, 10)], RANGE)])  # overlapping range
    want = O.count_exhaustive_single(rows, ranges)
    # brute force in numpy
    bf = np.zeros(len(rows), np.uint32)
    for r in ranges:
        bf += ((rows[:, r["dim"]] >= r["low"]) & (rows[:, r["dim"]] <= r["up"])) * np.uint32(r["weight"])
    assert np.array_equal(want, bf)
    for ml in (0, 13):
        items = ix.compile(ranges, max_load=ml)
        assert np.array_equal(O.count_index(ix.inv.view(np.uint32), items, len(rows)), want)
    words, cpos = ix.compress(11, 17)
    items = ix.compile(ranges, compr_pos=cpos)
    assert np.array_equal(O.count_index(words, items, len(rows), codec=11), want)


def test_count_index_equals_exhaustive_biject():
    rows = load_csv("tweets_20.csv")
    rp, vals = split_rows(rows)
    ix = O.OracleIndex.biject(rp, vals)
    for q in range(len(rows)):
        ranges = ranges_exact([rows[q]], biject=True)
        want = O.count_exhaustive_biject(rp, vals, ranges)
        got = O.count_index(ix.inv.view(np.uint32), ix.compile(ranges), len(rows))
        assert np.array_equal(got, want)
        bf = np.array([sum(r.count(t) for t in rows[q]) for r in rows], np.uint32)
        assert np.array_equal(want, bf)


def test_topk_tie_rule_and_padding():
    counts = np.array([0, 3, 5, 3, 0, 5, 1, 3], np.uint32)
    ids, cnt = O.topk(counts, 4)
    assert list(cnt) == [5, 5, 3, 3] and list(ids) == [2, 5, 1, 3]
    ids, cnt = O.topk(counts, 8)
    assert list(cnt) == [5, 5, 3, 3, 3, 1, 0, 0] and list(ids) == [2, 5, 1, 3, 7, 6, 0, 0]
    rng = np.random.default_rng(2)
    c = rng.integers(0, 6, 5000).astype(np.uint32)
    ids, cnt = O.topk(c, 300)
    order = np.lexsort((np.arange(len(c)), -c.astype(np.int64)))
    order = order[c[order] > 0][:300]
    assert np.array_equal(ids[:len(order)], order) and np.array_equal(cnt[:len(order)], c[order])


def test_sift20_topk_goldens():
    # test/test_csv_input_sift_data.cc:77-86 (k=5, exact match, first 3 rows)
    rows = sift20()
    for q, (i0, c0) in enumerate([(0, 5), (1, 5), (2, 5)]):
        ids, cnt = O.topk(O.count_exhaustive_single(rows, ranges_exact([rows[q]])), 5)
        assert (ids[0], cnt[0]) == (i0, c0)
        if q == 0:
            assert (ids[1], cnt[1]) == (4, 2)


def test_tweets20_exact_counts():
    # test/test_csv_input_tweets.cc:77-85 asserts 16/16/14/16 from 4-bit
    # counters that wrap; the exact counts are 17/17/14/18
    rows = load_csv("tweets_20.csv")
    rp, vals = split_rows(rows)
    res = [O.topk(O.count_exhaustive_biject(rp, vals, ranges_exact([rows[q]], biject=True)), 5) for q in range(3)]
    assert sorted(res[0][0][:2]) == [0, 11] and list(res[0][1][:2]) == [17, 17]
    assert (res[1][0][0], res[1][1][0]) == (1, 14)
    assert (res[2][0][0], res[2][1][0]) == (2, 18)


def test_search_mt_matches_single_thread():
    rng = np.random.default_rng(9)
    rows = rng.integers(0, 20, size=(3000, 16)).astype(np.int32)
    ix = O.OracleIndex.single(rows)
    qrows = rows[rng.integers(0, 3000, 12)]
    ranges = ranges_exact(qrows, radius=1)
    items = np.concatenate([ix.compile(ranges[ranges["query"] == q]) for q in range(12)])
    ids, cnt, th = O.search_mt(ix.inv.view(np.uint32), items, len(rows), 12, 25, n_threads=4)
    assert th == 4
    for q in range(12):
        i, c = O.topk(O.count_exhaustive_single(rows, ranges[ranges["query"] == q]), 25)
        assert np.array_equal(ids[q * 25:(q + 1) * 25], i) and np.array_equal(cnt[q * 25:(q + 1) * 25], c)
