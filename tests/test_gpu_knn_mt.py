"""knn_MT / knn_search_MT (knn.cu:185-283, interface.cc:859-909): several
tables with their own queries, admitted in rounds by estimated device memory
and run concurrently (one plan and stream each).  Every table's results
equal the exhaustive oracle's, whether all tables share one round or each
runs alone, and a budget no table fits in is refused."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

import genie_amd
from genie_amd import GenieError, Table, tables_knn_mt, workloads as W

pytestmark = pytest.mark.gpu

K = 30


def _cases():
    a = W.cfg1(seed=801, n=80_000, d=24, q=20, k=K)
    b = W.cfg4(seed=802, n=60_000, d=16, q=14)
    rp, vals = genie_amd.synth.sets(803, 30_000, 5, 40, 3000, 1.05)
    qp, qv = genie_amd.synth.sets(804, 11, 1, 40, 3000, 1.05)
    sr = W.set_ranges(qp, qv, genie_amd.synth.uniform_int(805, 0, len(qv), 1, 4))
    tables = [Table.from_rows(a["rows"]), Table.from_rows(b["rows"], 1024, True, codec=11),
              Table.from_sets(rp, vals)]
    ranges = [a["ranges"], b["ranges"], sr]
    nq = [20, 14, 11]
    counts = [lambda r: O.count_exhaustive_single(a["rows"], r), lambda r: O.count_exhaustive_single(b["rows"], r),
              lambda r: O.count_exhaustive_biject(rp, vals, r)]
    ref = []
    for r, n, cf in zip(ranges, nq, counts):
        q_of = r["query"]
        ref.append(oracle_topk_batch(lambda q: cf(r[q_of == q]), n, K))
    return tables, ranges, nq, ref


def test_knn_mt_rounds_match_oracle():
    tables, ranges, nq, ref = _cases()
    est = [t.device().plan_estimate(n, len(t.compile(r, n)), K) for t, r, n in zip(tables, ranges, nq)]
    MiB = 1 << 20
    one_each = max(est) + 50 * MiB + 1  # every table fits alone ...
    # ... and no table fits next to the one before it in its round
    assert all(est[i] + est[i + 1] + 50 * MiB >= one_each for i in range(len(est) - 1))
    for budget in (0, one_each):
        out = tables_knn_mt(tables, ranges, nq, K, budget=budget)
        for i, ((ids, cnt), (ri, rc)) in enumerate(zip(out, ref)):
            assert_topk_equal(ids, cnt, ri, rc, K, nq[i], f"table {i}, budget {budget}")
    with pytest.raises(GenieError) as e:
        tables_knn_mt(tables, ranges, nq, K, budget=50 * MiB + 1)
    assert e.value.code == 2 and "MEMORY NOT ENOUGH" in str(e.value)
    # the budget is per call: the next call admits against free memory again
    out = tables_knn_mt(tables[:1], ranges[:1], nq[:1], K)
    assert_topk_equal(out[0][0], out[0][1], ref[0][0], ref[0][1], K, nq[0], "after a refusal")
