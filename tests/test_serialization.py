"""Index persistence on the CPU: genie::utility::SaveTable / LoadTable
(src/genie/utility/serialization.cc:18-56; round-trip checks of
test/test_serialization.cc:25-81) in this build's native file format.  A
loaded table has the same layout, the same compiled queries and the same
compressed words as the table that was saved; damaged files are refused."""
import os

import numpy as np
import pytest

import genie_amd
from genie_amd import GenieError, Table, workloads as W


def _same(a, b):
    ea, eb = a.export(), b.export()
    for key in ("inv", "inv_pos", "words", "cpos", "lo", "hi", "keyseg"):
        assert np.array_equal(ea[key], eb[key]), key
    for key in ("n_post", "n_seg", "n_words", "n_dims", "n_objects", "n_keyseg"):
        assert ea[key] == eb[key], key


def _tables():
    w = W.cfg1(n=20_000, d=12, q=8, k=10)
    rp, vals = genie_amd.synth.sets(5, 3000, 2, 30, 500, 1.05)
    return w, [
        ("rows", Table.from_rows(w["rows"])),
        ("rows via lists, load balance 64", Table.from_rows(w["rows"], 64, True, via_lists=True)),
        ("sets", Table.from_sets(rp, vals)),
        ("compressed serial delta bp32", Table.from_rows(w["rows"], 1024, True, codec=11)),
        ("compressed heavyweight", Table.from_rows(w["rows"], 1024, True, codec=13)),
    ]


def test_round_trip_keeps_layout_and_queries(tmp_path):
    w, tables = _tables()
    for name, t in tables:
        path = str(tmp_path / (name.replace(" ", "_") + ".gix"))
        t.save(path)
        u = Table.load(path)
        _same(t, u)
        if name != "sets":
            a = t.compile(w["ranges"], w["n_queries"])
            b = u.compile(w["ranges"], w["n_queries"])
            assert np.array_equal(a, b), name
        # saving the loaded table gives the same bytes
        path2 = path + ".again"
        u.save(path2)
        assert open(path, "rb").read() == open(path2, "rb").read(), name


def test_damaged_files_are_refused(tmp_path):
    w = W.cfg1(n=2000, d=4, q=2, k=5)
    t = Table.from_rows(w["rows"])
    path = str(tmp_path / "t.gix")
    t.save(path)
    good = open(path, "rb").read()

    def refused(data, what):
        p = str(tmp_path / "bad.gix")
        with open(p, "wb") as f:
            f.write(data)
        with pytest.raises(GenieError):
            Table.load(p)

    refused(good[:-1], "truncated")
    refused(good[: len(good) // 2], "cut in half")
    refused(b"NOTGENIE" + good[8:], "magic")
    flipped = bytearray(good)
    flipped[len(good) // 2] ^= 0x10
    refused(bytes(flipped), "one bit flipped (checksum)")
    refused(b"", "empty")
    with pytest.raises(GenieError):
        Table.load(str(tmp_path / "missing.gix"))
