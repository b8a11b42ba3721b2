"""genie_hip_match_topk_static, the entry point that allocates no device
memory: its results equal genie_hip_match_topk's (and the oracle's) on two
calls in a row, the second with a smaller batch (the static scratch keeps
the first call's contents), on an uncompressed and a compressed index; a
batch that needs more than the static scratch is refused."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

from genie_amd import GenieError, Table, match_topk, workloads as W
from genie_amd._native import match_topk_static, static_scratch_bytes

pytestmark = pytest.mark.gpu


def _oracle(rows, ranges, nq, k):
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, ranges[q_of == q]), nq, k)


@pytest.mark.parametrize("codec", [0, 11])
def test_static_equals_allocating_entry(codec):
    assert static_scratch_bytes() == 32 << 20
    w = W.cfg4(seed=991, n=200_000, d=32, q=48)
    t = Table.from_rows(w["rows"], 1024, codec != 0, codec=codec)
    big = t.compile(w["ranges"], 48, merge_runs=codec != 0)
    k = w["k"]
    ri, rc = _oracle(w["rows"], w["ranges"], 48, k)
    a = match_topk(t.device(), big, 48, k)
    s1 = match_topk_static(t.device(), big, 48, k)
    assert np.array_equal(a[0], s1[0]) and np.array_equal(a[1], s1[1])
    assert_topk_equal(s1[0], s1[1], ri, rc, k, 48, f"static, codec {codec}")
    # a smaller batch next: 7 queries, other ranges and k
    sub = w["ranges"][w["ranges"]["query"] < 7].copy()
    sub["low"] -= 1
    sub["up"] += 1
    small = t.compile(sub, 7, merge_runs=codec != 0)
    k2 = np.array([5, 300, 1, 64, 1000, 17, 100], np.uint32)
    b = match_topk(t.device(), small, 7, k2)
    s2 = match_topk_static(t.device(), small, 7, k2)
    assert np.array_equal(b[0], s2[0]) and np.array_equal(b[1], s2[1])
    ri2, rc2 = _oracle(w["rows"], sub, 7, k2)
    assert_topk_equal(s2[0], s2[1], ri2, rc2, k2, 7, f"static second call, codec {codec}")


def test_static_refuses_a_batch_beyond_its_scratch():
    """128 queries of k = 4,096 over 1M objects (16 tiles of 65,536) need
    128 x 16 x 4,096 x 8 B = 64 MiB of candidate space: refused, and the
    entry answers a small batch right after."""
    w = W.cfg1(seed=992, n=1_000_000, d=8, q=128, k=4096)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 128)
    with pytest.raises(GenieError) as e:
        match_topk_static(t.device(), items, 128, np.full(128, 4096, np.uint32))
    assert "static scratch" in str(e.value)
    one = items[items["query"] == 0]
    a = match_topk(t.device(), one, 1, 50)
    s = match_topk_static(t.device(), one, 1, 50)
    assert np.array_equal(a[0], s[0]) and np.array_equal(a[1], s[1])
