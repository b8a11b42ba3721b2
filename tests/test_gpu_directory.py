"""Graded tile and segment directories on the GPU: lists whose lengths put
their keys at every row grade (full-resolution rows, rows of one entry per
2, 4 and 8 directory tiles, and no row), an object count that is not a
multiple of the 8,192-object directory tile, and every form whose tiles the
rows clip (direct-mapped plain kernel with either select, hash form with its
tiles of 16-directory-tile multiples, lane form with 8,192- and 16,384-object
tiles), uncompressed and compressed, against the exhaustive oracle."""
import numpy as np
import pytest

from helpers import assert_topk_equal
from oracle import oracle as O

from genie_amd import Plan, Table, workloads as W

pytestmark = pytest.mark.gpu

N = 2_000_003  # 245 directory tiles, the last one partial
D = 4


def graded_rows(seed=2024):
    """Per dim: 200 keys of 300 ids (tile-directory grade 3: 32-entry rows),
    200 of 700 (grade 2), 200 of 1,500 (grade 1), 100 of 3,000 (grade 0, full
    rows), 300 of 100 (no row), 10 of 6,000 and 10 of 12,000 (compressed:
    segment-directory rows of grades 2 and 1) and the rest of the objects in
    24 long keys; objects are dealt to keys by a random permutation, so every
    list spans the whole id range."""
    rng = np.random.default_rng(seed)
    sizes = [300] * 200 + [700] * 200 + [1500] * 200 + [3000] * 100 + [100] * 300 + [6000] * 10 + [12000] * 10
    rest = N - sum(sizes)
    big = [rest // 24] * 24
    big[-1] += rest - sum(big)
    sizes = np.array(sizes + big, np.int64)
    rows = np.empty((N, D), np.int32)
    for d in range(D):
        vals = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
        rows[:, d] = vals[rng.permutation(N)]
    return rows, len(sizes)


def graded_ranges(n_keys, q=12, seed=7):
    rng = np.random.default_rng(seed)
    # centres over the short and mid-length keys (0..1019) and a few long ones
    qv = rng.integers(0, 1020, (q, D)).astype(np.int32)
    qv[::4, 0] = n_keys - 3
    return W.single_ranges(qv, 2)


def oracle(rows, ranges, nq, k):
    ids, cnt = [], []
    for q in range(nq):
        i, c = O.topk(O.count_exhaustive_single(rows, ranges[ranges["query"] == q]), k)
        ids.append(i)
        cnt.append(c)
    return np.concatenate(ids), np.concatenate(cnt)


@pytest.fixture(scope="module")
def graded():
    rows, nk = graded_rows()
    ranges = graded_ranges(nk)
    nq, k = 12, 60
    return rows, ranges, nq, k, oracle(rows, ranges, nq, k)


def run_plan(index, items, nq, k, mode, hash_mode=-1, sparse=-1, decode=-1, dense=True, tile=0, block=1024):
    kq = np.full(nq, k, np.uint32)
    p = Plan(index, nq, len(items), k, 0, mode=mode)
    p.set_hash(hash_mode)
    p.set_sparse(sparse)
    p.set_decode(decode)
    if mode == 1:
        p.set_dense(dense)
        p.set_form("lane")
        if tile:
            p.set_config(block, tile, 65536)
    p.upload(items, nq, kq)
    p.run()
    return p.download()


@pytest.mark.parametrize("form", ["direct-dense", "direct-sparse", "hash", "lanes-8192", "lanes-16384"])
def test_graded_rows_uncompressed(graded, form):
    rows, ranges, nq, k, (ri, rc) = graded
    t = Table.from_rows(rows)
    items = t.compile(ranges, nq)
    if form.startswith("lanes"):
        got = run_plan(t.device(), items, nq, k, 1, dense=False, tile=int(form.split("-")[1]))
    else:
        got = run_plan(t.device(), items, nq, k, 0, hash_mode=1 if form == "hash" else 0,
                       sparse=1 if form == "direct-sparse" else (0 if form == "direct-dense" else -1))
    assert_topk_equal(got[0], got[1], ri, rc, k, nq, f"graded rows, {form}")


@pytest.mark.parametrize("codec", [11, 13])
@pytest.mark.parametrize("decode", [-1, 0])
def test_graded_rows_compressed(graded, codec, decode):
    rows, ranges, nq, k, (ri, rc) = graded
    t = Table.from_rows(rows, 1024, True, codec=codec)
    items = t.compile(ranges, nq, merge_runs=True)
    for hash_mode in (0, 1):
        got = run_plan(t.device(), items, nq, k, 0, hash_mode=hash_mode, decode=decode)
        assert_topk_equal(got[0], got[1], ri, rc, k, nq, f"graded rows, codec {codec}, decode {decode}, "
                                                         f"hash {hash_mode}")
    # the compressed lane kernel keeps a decode scratch per wave: 256 threads
    got = run_plan(t.device(), items, nq, k, 1, dense=False, tile=8192, block=256)
    assert_topk_equal(got[0], got[1], ri, rc, k, nq, f"graded rows, codec {codec}, lane form")
