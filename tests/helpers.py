"""Shared helpers of the tests: reference-style comparisons."""
import numpy as np


def assert_topk_equal(ids, cnt, ref_ids, ref_cnt, k, n_queries, where="", strict=True):
    """Bit-exact counts; ids identical except among ties at the k-th count of
    a query, which are checked by count (the reference picks tied ids
    arbitrarily, heap_count.cu:97-114).  With this build's deterministic tie
    rule (count desc, id asc) both sides normally agree exactly."""
    if strict:
        ids_a, cnt_a = np.asarray(ids), np.asarray(cnt)
        if np.array_equal(ids_a, np.asarray(ref_ids)) and np.array_equal(cnt_a, np.asarray(ref_cnt)):
            return
    if np.ndim(k):  # per-query k: results are concatenated, k[q] slots each
        cuts = np.cumsum(np.asarray(k, np.int64))[:-1]
        ids, cnt, ref_ids, ref_cnt = (np.split(np.asarray(x), cuts) for x in (ids, cnt, ref_ids, ref_cnt))
    else:
        ids = np.asarray(ids).reshape(n_queries, -1)
        cnt = np.asarray(cnt).reshape(n_queries, -1)
        ref_ids = np.asarray(ref_ids).reshape(n_queries, -1)
        ref_cnt = np.asarray(ref_cnt).reshape(n_queries, -1)
    for q in range(n_queries):
        if not np.array_equal(cnt[q], ref_cnt[q]):
            bad = np.nonzero(cnt[q] != ref_cnt[q])[0][:5]
            raise AssertionError(f"{where} query {q}: counts differ at slots {bad}: "
                                 f"{cnt[q][bad]} vs {ref_cnt[q][bad]}")
        kth = ref_cnt[q][-1]
        above = ref_cnt[q] > kth
        if not np.array_equal(ids[q][above], ref_ids[q][above]):
            raise AssertionError(f"{where} query {q}: ids above the k-th count differ")
        if not np.array_equal(ids[q][~above], ref_ids[q][~above]):
            if strict:
                raise AssertionError(f"{where} query {q}: tied ids at count {kth} differ "
                                     f"({ids[q][~above][:8]} vs {ref_ids[q][~above][:8]})")
            # ties at the k-th count may legitimately differ; all must carry that count
            assert np.all(cnt[q][~above] == kth)


def split_rows(rows):
    """CSR (row_ptr, vals) of a list of int lists."""
    lens = np.array([len(r) for r in rows], np.int64)
    row_ptr = np.zeros(len(rows) + 1, np.int64)
    row_ptr[1:] = np.cumsum(lens)
    vals = np.array([v for r in rows for v in r], np.int32)
    return row_ptr, vals


def oracle_topk_batch(counts_fn, n_queries, k):
    from oracle import oracle as O
    ids, cnt = [], []
    for q in range(n_queries):
        kq = int(k[q]) if np.ndim(k) else int(k)
        i, c = O.topk(counts_fn(q), kq)
        ids.append(i)
        cnt.append(c)
    return np.concatenate(ids), np.concatenate(cnt)


def posting_items(export, items):
    """Items of a compressed table (word positions on segment boundaries)
    translated to positions of its uncompressed posting array (the table's
    inv), so that the oracle can count them with codec 0."""
    cpos = np.asarray(export["cpos"], np.int64)
    upos = np.asarray(export["inv_pos"], np.int64)
    out = items.copy()
    out["start_pos"] = upos[np.searchsorted(cpos, items["start_pos"].astype(np.int64))]
    out["end_pos"] = upos[np.searchsorted(cpos, items["end_pos"].astype(np.int64))]
    return out


def oracle_search_sample(words, items, n_objects, k, qs):
    """The oracle's multithreaded index-based search (the CPU-Idx restatement,
    test_compression_query.cu:256-265) over the queries qs of a batch;
    returns their rows concatenated in the order of qs."""
    from oracle import oracle as O
    qs = np.asarray(qs, np.int64)
    n = len(k) if np.ndim(k) else max(int(items["query"].max()) + 1 if len(items) else 1, int(qs.max()) + 1)
    kq = np.broadcast_to(np.asarray(k, np.int64), (n,))
    sel = np.isin(items["query"], qs)
    sub = items[sel].copy()
    remap = np.full(len(kq), -1, np.int64)
    remap[qs] = np.arange(len(qs))
    sub["query"] = remap[sub["query"]]
    sub = sub[np.argsort(sub["query"], kind="stable")]
    ids, cnt, _ = O.search_mt(np.asarray(words, np.uint32), sub, n_objects, len(qs), kq[qs].astype(np.int32))
    return ids, cnt


def rows_of(ids, cnt, k, n_queries, qs):
    """Rows of queries qs out of a batch result (k an int or per query)."""
    kq = np.broadcast_to(np.asarray(k, np.int64), (n_queries,))
    off = np.zeros(n_queries + 1, np.int64)
    off[1:] = np.cumsum(kq)
    return (np.concatenate([ids[off[q]:off[q + 1]] for q in qs]),
            np.concatenate([cnt[off[q]:off[q + 1]] for q in qs]))
