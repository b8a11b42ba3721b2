"""Host side of the drop-in boundary, on the CPU: the C++ host library's
table build, compression and query compilation against the oracle, and the
C-ABI libraries exporting every symbol their headers declare."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_csv
from helpers import split_rows
from oracle import oracle as O

import genie_amd
from genie_amd import Table, workloads as W
from genie_amd import RANGE_DTYPE


def _same_layout(t, ix):
    e = t.export()
    assert np.array_equal(e["inv"], ix.inv)
    assert np.array_equal(e["inv_pos"], ix.inv_pos)
    assert np.array_equal(e["lo"], ix.lo) and np.array_equal(e["hi"], ix.hi)
    assert np.array_equal(e["keyseg"], ix.keyseg)
    return e


@pytest.mark.parametrize("via_lists", [True, False])
@pytest.mark.parametrize("lb,maxlen", [(False, 0), (True, 5), (True, 1024)])
def test_table_layout_matches_oracle(via_lists, lb, maxlen):
    rows = np.array([r[:5] for r in load_csv("sift_20.csv")], np.int32)
    e = _same_layout(Table.from_rows(rows, maxlen, lb, via_lists=via_lists), O.OracleIndex.single(rows, maxlen, lb))
    assert list(e["inv"][:6]) == [8, 9, 7, 0, 2, 4]
    w = W.cfg0(n=500, q=4)
    _same_layout(Table.from_rows(w["rows"], maxlen, lb, via_lists=via_lists),
                 O.OracleIndex.single(w["rows"], maxlen, lb))


@pytest.mark.parametrize("via_lists", [True, False])
def test_biject_layout_matches_oracle(via_lists):
    rp, vals = split_rows(load_csv("tweets_20.csv"))
    e = _same_layout(Table.from_sets(rp, vals, via_lists=via_lists), O.OracleIndex.biject(rp, vals))
    assert list(e["inv"][:6]) == [0, 11, 0, 11, 0, 11]
    rp, vals = genie_amd.synth.sets(5, 3000, 1, 12, 500, 1.05)
    _same_layout(Table.from_sets(rp, vals, 64, True, via_lists=via_lists), O.OracleIndex.biject(rp, vals, 64, True))


@pytest.mark.parametrize("codec", range(1, 14))
def test_compressed_table_matches_oracle(codec):
    w = W.cfg0(n=700, d=16, q=3)
    t = Table.from_rows(w["rows"], 64, True, codec=codec)
    ix = O.OracleIndex.single(w["rows"], 64, True)
    e = _same_layout(t, ix)
    words, cpos = ix.compress(codec, 64)
    assert np.array_equal(e["words"], words) and np.array_equal(e["cpos"], cpos)


def test_compressed_table_requires_short_segments():
    w = W.cfg0(n=3000, d=4, q=1)
    with pytest.raises(genie_amd.GenieError):
        Table.from_rows(w["rows"], 0, False, codec=13)


@pytest.mark.parametrize("lb,max_load", [(False, 0), (True, 37)])
def test_compile_matches_oracle(lb, max_load):
    w = W.cfg0(n=900, d=32, q=6)
    rng = np.random.default_rng(1)
    ranges = w["ranges"].copy()
    ranges["low"] -= rng.integers(0, 3, len(ranges))
    ranges["up"] += rng.integers(0, 3, len(ranges))
    ranges = np.concatenate([ranges, np.array([(2, 40, 0, 10, 1.0, 99), (3, 5, 300, 400, 1.0, 98)], RANGE_DTYPE)])
    t = Table.from_rows(w["rows"])
    ix = O.OracleIndex.single(w["rows"])
    got = t.compile(ranges, w["n_queries"], lb, max_load)
    want = np.concatenate([ix.compile(ranges[ranges["query"] == q], max_load if lb else 0)
                           for q in range(w["n_queries"])])
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_merged_runs_cover_the_same_segments():
    """Merging a range's consecutive segment items keeps exactly the covered
    positions, per (query, order, weight), and never joins different ranges."""
    w = W.cfg0(n=3000, d=8, q=7)
    r = w["ranges"].copy()
    r["low"] -= 3
    r["up"] += 3
    r["weight"][::3] = 2
    t = Table.from_rows(w["rows"], 50, True, codec=13)
    per_seg = t.compile(r, w["n_queries"])
    merged = t.compile(r, w["n_queries"], merge_runs=True)
    assert len(merged) < len(per_seg)

    def cover(items):
        out = set()
        for it in items:
            for p in range(int(it["start_pos"]), int(it["end_pos"])):
                out.add((int(it["query"]), int(it["order"]), float(it["weight"]), p))
        return out
    assert cover(per_seg) == cover(merged)
    # one item per (query, range) that covers any segment
    keys = {(int(it["query"]), int(it["order"])) for it in per_seg if it["end_pos"] > it["start_pos"]}
    assert len([it for it in merged if it["end_pos"] > it["start_pos"]]) == len(keys)


def test_compile_compressed_matches_oracle():
    w = W.cfg0(n=900, d=8, q=5)
    t = Table.from_rows(w["rows"], 50, True, codec=13)
    ix = O.OracleIndex.single(w["rows"], 50, True)
    _, cpos = ix.compress(13, 50)
    got = t.compile(w["ranges"], w["n_queries"])
    want = np.concatenate([ix.compile(w["ranges"][w["ranges"]["query"] == q], compr_pos=cpos)
                           for q in range(w["n_queries"])])
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(genie_\w+|synth_\w+)\s*\(", src)))


@pytest.mark.parametrize("header,lib", [("genie_hip.h", "libgenie_hip.so"), ("genie_table.h", "libgenie.so")])
def test_c_abi_exports_every_declared_symbol(header, lib):
    so = C.CDLL(os.path.join(ROOT, "genie_amd", "lib", lib))
    names = _declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_synth_exports():
    so = C.CDLL(os.path.join(ROOT, "genie_amd", "lib", "libgenie_synth.so"))
    src = open(os.path.join(ROOT, "genie_amd", "csrc", "synth", "synth.h")).read()
    for n in re.findall(r"\b(synth_\w+)\s*\(", src):
        assert hasattr(so, n), n


def test_synth_is_deterministic():
    a = genie_amd.synth.rows(7, 1000, 8, genie_amd.synth.ZIPF_PERM, 1.1, 0, 0, 255, 256)
    b = genie_amd.synth.rows(7, 1000, 8, genie_amd.synth.ZIPF_PERM, 1.1, 0, 0, 255, 256)
    assert np.array_equal(a, b) and a.min() >= 0 and a.max() <= 255
    # Zipf(1.1) over 256 values: the most frequent value holds about a fifth
    top = np.bincount(a[:, 0], minlength=256).max() / 1000
    assert 0.12 < top < 0.3


def test_adaptive_query_range_known_answers():
    """apply_adaptive_query_range (query.cc:196-261) on hand-worked cases;
    the C++ test driver's host-only case."""
    import subprocess
    from conftest import GOLDEN
    exe = os.path.join(ROOT, "genie_amd", "lib", "test_api")
    r = subprocess.run([exe, GOLDEN, "adaptive_range"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0 and "[  OK  ] adaptive_range" in r.stdout, r.stdout


def test_knn_mt_admission_rounds():
    """knn_MT admits tables in order while more than estimate + 50 MiB stays
    free (knn.cu:207-243); a table that does not fit alone is refused with
    gpu_bad_alloc("MEMORY NOT ENOUGH")."""
    from genie_amd import GenieError, knn_mt_admit
    MiB = 1 << 20
    est = [100 * MiB, 100 * MiB, 100 * MiB, 10 * MiB]
    # 300 MiB free: 0 (200 left), 1 (100 left); 2 needs 150
    assert knn_mt_admit(est, 0, 300 * MiB) == 2
    # the next round starts at 2 with the memory of that moment
    assert knn_mt_admit(est, 2, 300 * MiB) == 4
    assert knn_mt_admit(est, 2, 160 * MiB) == 3
    # exactly estimate + 50 MiB is not enough (strictly more is needed)
    assert knn_mt_admit(est, 3, 60 * MiB + 1) == 4
    with pytest.raises(GenieError) as e:
        knn_mt_admit(est, 3, 60 * MiB)
    assert e.value.code == 2 and "MEMORY NOT ENOUGH" in str(e.value)
    with pytest.raises(GenieError):
        knn_mt_admit(est, 0, 150 * MiB)
    # empty query sets cost nothing; the end of the list is a complete round
    assert knn_mt_admit([0, 0], 0, 51 * MiB) == 2
    assert knn_mt_admit(est, 4, 0) == 4
