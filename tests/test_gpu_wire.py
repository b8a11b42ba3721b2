"""The serving precedent's search path on the GPU: a query document over
several clusters (one table each) gives the reference's merged output --
each named cluster's top-k, ids shifted by the cluster's offset
(CalculateIdOffset, online.cu:80-105), the pairs of every query sorted by
(count, id) descending and cut to k (sorting.cc:20-64, file.cu:42-53) --
checked against the oracle's exhaustive counts on every cluster."""
import json

import numpy as np
import pytest

from oracle import oracle as O

from genie_amd import GenieError, Table, synth, wire, workloads as W

pytestmark = pytest.mark.gpu


def expected(clusters, content, named, radius, k):
    offsets = np.concatenate([[0], np.cumsum([len(c) for c in clusters])[:-1]])
    out = []
    for q, cl in enumerate(named):
        pairs = []
        for c in cl:
            r = W.single_ranges(np.asarray([content[q]], np.int32), radius)
            ids, cnt = O.topk(O.count_exhaustive_single(clusters[c], r), k)
            for i, n in zip(ids, cnt):
                pairs.append((int(n), int(offsets[c] + i) if n > 0 else 0))
        pairs.sort(reverse=True)
        pairs = (pairs + [(0, 0)] * k)[:k]
        out += [(i, n) for n, i in pairs]
    return out


def test_multi_cluster_search_matches_the_merge_rule():
    rows = synth.rows(81, 30_000, 8, synth.ZIPF_PERM, 1.1, 0, 0, 63, 64)
    clusters = [rows[:9_000], rows[9_000:21_000], rows[21_000:]]
    tables = [Table.from_rows(c) for c in clusters]
    qv, _ = synth.queries(82, rows, 7, noise=1, lo=0, hi=63)
    named = [[0], [1, 2], [0, 1, 2], [2], [], [1], [0, 2]]
    k, radius = 20, 1
    query_doc = json.dumps({"topk": k, "queries": [{"content": [int(v) for v in qv[q]], "clusters": named[q]}
                                                   for q in range(7)]})
    config_doc = json.dumps(dict(data_file="unused", dim=8, count_threshold=8, data_type=0, search_type=0,
                                 max_data_size=0, num_of_cluster=3, data_format=0, radius=radius, selectivity=0.0))
    got = wire.search(tables, config_doc, query_doc)
    want = expected(clusters, qv, named, radius, k)
    assert len(got) == 7 * k
    assert got == want


def test_unknown_cluster_is_refused():
    rows = synth.rows(83, 2_000, 4, synth.UNIFORM, 0, 0, 0, 9, 10)
    t = Table.from_rows(rows)
    config_doc = json.dumps(dict(data_file="unused", dim=4, count_threshold=4, data_type=0, search_type=0,
                                 max_data_size=0, num_of_cluster=1, data_format=0, radius=0, selectivity=0.0))
    with pytest.raises(GenieError):
        wire.search([t], config_doc, json.dumps({"topk": 5, "queries": [{"content": [1, 2, 3, 4], "clusters": [1]}]}))
