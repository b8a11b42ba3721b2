"""Compressed posting lists on the GPU: the wave decoders against the oracle's
codecs, and compressed search against uncompressed search and the oracle
(bit-exact counts; compressed and uncompressed layouts compared as in
test/test_compression_kernels.cu:236-258, for every query and slot)."""
import numpy as np
import pytest

from helpers import assert_topk_equal
from oracle import oracle as O
from test_oracle import mixed_values

import genie_amd
from genie_amd import Table, match_topk, workloads as W
from genie_amd._native import decode as gpu_decode

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("codec", range(1, 14))
def test_device_decoders_match_oracle(codec):
    vals = mixed_values(1024, seed=666 + codec)
    lengths = list(range(1, 130)) + list(range(130, 1025, 29)) + [1023, 1024]
    words, cpos = [], [0]
    inputs = []
    for i, n in enumerate(lengths):
        x = vals[:n] if i % 2 else np.sort(vals[:n])
        w = O.encode(codec, x)
        if len(w) > 1024:  # the engine decodes segments of at most 1,024 words
            continue
        inputs.append(x)
        words.append(w)
        cpos.append(cpos[-1] + len(w))
    out = gpu_decode(codec, np.concatenate(words), np.array(cpos, np.uint32))
    for x, y in zip(inputs, out):
        assert y is not None and np.array_equal(x, y), (codec, len(x))


def _oracle(rows, ranges, nq, k):
    from helpers import oracle_topk_batch
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, ranges[q_of == q]), nq, k)


def _compare(rows, ranges, nq, k, codec, max_length=1024, label="", oracle=True):
    """Compressed against uncompressed (bit-exact, every query and slot), and
    both against the exhaustive oracle; the compressed index in the plain
    form with its segments decoded once per batch (the default) and decoded
    by the kernels themselves."""
    from genie_amd import Plan
    plain = Table.from_rows(rows)
    comp = Table.from_rows(rows, max_length, True, codec=codec)
    ids_u, cnt_u = match_topk(plain.device(), plain.compile(ranges, nq), nq, k)
    items = comp.compile(ranges, nq)
    ids_c, cnt_c = match_topk(comp.device(), items, nq, k)
    assert np.array_equal(cnt_u, cnt_c) and np.array_equal(ids_u, ids_c), f"{label} codec {codec}"
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
    for dm in (-1, 0):
        p = Plan(comp.device(), nq, len(items), int(kq.max()), 0, mode=0)
        p.set_decode(dm)
        p.upload(items, nq, kq)
        assert p.decode_active() == (dm == -1)
        p.run()
        got = p.download()
        assert np.array_equal(got[0], ids_u) and np.array_equal(got[1], cnt_u), f"{label} codec {codec} decode {dm}"
    if oracle:
        ri, rc = _oracle(rows, ranges, nq, k)
        assert_topk_equal(ids_c, cnt_c, ri, rc, k, nq, f"{label} codec {codec} vs oracle")
    return ids_c, cnt_c


@pytest.mark.parametrize("codec", range(1, 14))
def test_compressed_search_every_codec(codec):
    w = W.cfg1(n=60_000, d=16, q=6, k=40)
    _compare(w["rows"], w["ranges"], w["n_queries"], w["k"], codec, 512, "zipf")


def test_lsh_heavyweight_many_tiles():
    # LSH-like: uniform buckets, exact match, SERIAL_DELTA_COMP_BP32_VARINT
    w = W.cfg2(n=300_000, d=32, q=24, k=100, buckets=2048)
    _compare(w["rows"], w["ranges"], w["n_queries"], w["k"], 13, 1024, "lsh")


def test_zipf_middleweight_ranges_and_mixed_k():
    w = W.cfg4(n=150_000, d=24, q=20)
    _compare(w["rows"], w["ranges"], w["n_queries"], w["k"], 11, 1024, "cfg4-like")


def test_compressed_bijectmap_weights():
    rp, vals = genie_amd.synth.sets(41, 40_000, 5, 40, 2000, 1.05)
    qp, qv = genie_amd.synth.sets(42, 16, 1, 64, 2000, 1.05)
    ranges = W.set_ranges(qp, qv, genie_amd.synth.uniform_int(43, 0, len(qv), 1, 4))
    plain = Table.from_sets(rp, vals)
    comp = Table.from_sets(rp, vals, 1024, True, codec=6)
    a = match_topk(plain.device(), plain.compile(ranges, 16), 16, 100)
    b = match_topk(comp.device(), comp.compile(ranges, 16), 16, 100)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_corrupt_segment_is_rejected():
    w = W.cfg0(n=600, d=4, q=1)
    t = Table.from_rows(w["rows"], 64, True, codec=3)
    e = t.export()
    words = e["words"].copy()
    words[1] ^= 0x1F000000  # first BP32 header: an impossible bit width
    with pytest.raises(genie_amd.GenieError):
        genie_amd.Index.create(600, 4, e["inv"], _key_start(e), _key_dim(e), e["inv_pos"], True, 3, words, e["cpos"])
    # the untouched words build fine
    genie_amd.Index.create(600, 4, e["inv"], _key_start(e), _key_dim(e), e["inv_pos"], True, 3, e["words"], e["cpos"])


def _key_start(e):
    ks = [0]
    for d in range(len(e["lo"])):
        seg = e["keyseg"][_ks_off(e, d):_ks_off(e, d + 1)]
        for v in range(e["hi"][d] - e["lo"][d] + 1):
            ks.append(e["inv_pos"][seg[v + 1]])
    return np.array(ks, np.uint64)


def _key_dim(e):
    return np.concatenate([np.full(e["hi"][d] - e["lo"][d] + 1, d, np.uint32) for d in range(len(e["lo"]))])


def _ks_off(e, d):
    return int(sum(e["hi"][j] - e["lo"][j] + 2 for j in range(d)))


@pytest.mark.parametrize("table_kind", ["rows", "sets"])
def test_merged_segment_items_give_the_same_results(table_kind):
    """Items spanning several whole segments (a range's segments merged into
    one item) against one item per segment, in every form and against the
    oracle: plain (automatic, direct-mapped, hash), batched (dense or set
    form, lane form) and automatic."""
    from genie_amd import Plan
    from helpers import oracle_topk_batch
    if table_kind == "rows":
        w = W.cfg1(seed=931, n=150_000, d=24, q=40, k=50, radius=3)
        r = w["ranges"].copy()
        r["weight"][::5] = 2
        t = Table.from_rows(w["rows"], 256, True, codec=11)
        q_of = r["query"]
        ri, rc = oracle_topk_batch(lambda q: O.count_exhaustive_single(w["rows"], r[q_of == q]), 40, 50)
    else:
        rp, vals = genie_amd.synth.sets(932, 80_000, 5, 40, 4_000, 1.05)
        qp, qv = genie_amd.synth.sets(933, 40, 1, 30, 4_000, 1.05)
        r = W.set_ranges(qp, qv, genie_amd.synth.uniform_int(934, 0, len(qv), 1, 4))
        t = Table.from_sets(rp, vals, 256, True, codec=6)
        q_of = r["query"]
        ri, rc = oracle_topk_batch(lambda q: O.count_exhaustive_biject(rp, vals, r[q_of == q]), 40, 50)
    per_seg = t.compile(r, 40)
    merged = t.compile(r, 40, merge_runs=True)
    assert len(merged) < len(per_seg)
    kq = np.full(40, 50, np.uint32)
    for mode, hm, dense in [(0, -1, True), (0, 0, True), (0, 1, True), (1, -1, True), (1, -1, False), (2, -1, True)]:
        res = []
        for items in (per_seg, merged):
            p = Plan(t.device(), 40, len(per_seg), 50, 0, mode=mode)
            p.set_hash(hm)
            if mode:
                p.set_dense(dense)
            p.upload(items, 40, kq)
            p.run()
            res.append(p.download())
        where = f"{table_kind} mode {mode} hash {hm} dense {dense}"
        assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1]), where
        assert_topk_equal(res[1][0], res[1][1], ri, rc, 50, 40, where)


def test_decode_once_edge_batches():
    """The plain form's decode-once path on batches with zero-weight items,
    empty items, queries without items, load-balanced per-segment items and
    mixed k, one plan across several batches (the scratch keeps segments of
    earlier batches that later ones do not name): every batch equals the
    oracle's results and the in-kernel decode's."""
    from genie_amd import Plan
    from helpers import oracle_topk_batch
    w = W.cfg4(seed=961, n=90_000, d=20, q=30)
    rows = w["rows"]
    t = Table.from_rows(rows, 256, True, codec=11)
    p = Plan(t.device(), 40, 40 * 20 * 200, 1000, 0, mode=0)
    q_of0 = w["ranges"]["query"]
    for b, (nq, lb, zero_every) in enumerate([(30, False, 0), (12, True, 4), (40, False, 3), (5, False, 0)]):
        r = w["ranges"][q_of0 < min(nq, 30)].copy()
        if zero_every:
            r["weight"][::zero_every] = 0
        if nq > 30:  # queries 30.. have no ranges
            pass
        items = t.compile(r, nq, load_balance=lb, max_load=300 if lb else 0)
        empty = items[:3].copy()
        empty["end_pos"] = empty["start_pos"]  # empty spans
        items = np.concatenate([items, empty])
        kq = np.array([(10, 100, 1000, 1)[(i + b) % 4] for i in range(nq)], np.uint32)
        q_of = r["query"]
        ri, rc = oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, r[q_of == q]), nq, kq)
        got = []
        for dm in (-1, 0):
            p.set_decode(dm)
            p.upload(items, nq, kq)
            assert p.decode_active() == (dm == -1)
            p.run()
            got.append(p.download())
        assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1]), b
        assert_topk_equal(got[0][0], got[0][1], ri, rc, kq, nq, f"decode-once batch {b}")
