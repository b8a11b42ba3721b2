"""Batched form (lane form: persistent kernel over a (query group, tile) work
queue, queries packed into count lanes, merged spans; dense form: bit-sliced
counting over per-object key codes; the plain form where the autotuner finds
it faster; hipGraph replay, autotuner) against the exhaustive oracle:
identical results."""
import numpy as np
import pytest

from helpers import assert_topk_equal, oracle_topk_batch
from oracle import oracle as O

import genie_amd
from genie_amd import Plan, Table, match_topk, workloads as W

pytestmark = pytest.mark.gpu


def plain(table, items, nq, k):
    return match_topk(table.device(), items, nq, k)


def oracle_rows(rows, ranges, nq, k):
    q_of = ranges["query"]
    return oracle_topk_batch(lambda q: O.count_exhaustive_single(rows, ranges[q_of == q]), nq, k)


def batched(table, items, nq, k, config=None, graph=False, runs=1, dense=False):
    kmax = int(np.max(k))
    p = Plan(table.device(), nq, len(items), kmax, 0, mode=1)
    p.set_dense(dense)
    if config:
        p.set_config(*config)
    p.upload(items, nq, k)
    if graph:
        p.set_graph(True)
    out = None
    for _ in range(runs):
        p.run()
        got = p.download()
        if out is not None:
            assert np.array_equal(got[0], out[0]) and np.array_equal(got[1], out[1])
        out = got
    return out, p


@pytest.mark.parametrize("config", [None, (256, 8192, 16384), (512, 16384, 65536), (1024, 8192, 1024)])
def test_batched_equals_plain_zipf(config):
    w = W.cfg1(n=150_000, d=32, q=37, k=50)  # 37 queries: a partly filled last group
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], w["n_queries"])
    ri, rc = oracle_rows(w["rows"], w["ranges"], w["n_queries"], w["k"])
    (b, _) = batched(t, items, w["n_queries"], w["k"], config)
    assert_topk_equal(b[0], b[1], ri, rc, w["k"], w["n_queries"], f"lane form {config}")


def test_batched_mixed_k_weights_bijectmap():
    rp, vals = genie_amd.synth.sets(51, 60_000, 5, 40, 4000, 1.05)
    qp, qv = genie_amd.synth.sets(52, 70, 1, 64, 4000, 1.05)
    ranges = W.set_ranges(qp, qv, genie_amd.synth.uniform_int(53, 0, len(qv), 1, 4))
    kq = np.array([10, 100, 1000] * 23 + [7], np.uint32)
    t = Table.from_sets(rp, vals)
    items = t.compile(ranges, 70)
    (b, p) = batched(t, items, 70, kq)
    ri, rc = [], []
    for q in range(70):
        i, c = O.topk(O.count_exhaustive_biject(rp, vals, ranges[ranges["query"] == q]), int(kq[q]))
        ri.append(i)
        rc.append(c)
    assert_topk_equal(b[0], b[1], np.concatenate(ri), np.concatenate(rc), kq, 70, "batched biject")


@pytest.mark.parametrize("wmax,lb", [(1, 8), (300, 16), (5000, 32)])
def test_batched_lane_widths(wmax, lb):
    w = W.cfg1(n=40_000, d=24, q=9, k=20)
    r = w["ranges"].copy()
    r["weight"] = np.random.default_rng(wmax).integers(1, wmax + 1, len(r))
    t = Table.from_rows(w["rows"])
    items = t.compile(r, 9)
    ri, rc = oracle_rows(w["rows"], r, 9, 20)
    (b, p) = batched(t, items, 9, 20)
    assert p.config()["lane_bits"] == lb
    assert_topk_equal(b[0], b[1], ri, rc, 20, 9, f"lane width {lb}")


def test_lane_widths_change_between_uploads():
    """One plan, batches of 64 queries at lane widths 8, 16, 32 and 8 again:
    4, 8 and 16 queries share a counter word, so the later batches hold more
    query groups (16, then 32) than the first (8) in the same group table."""
    w = W.cfg1(n=40_000, d=24, q=64, k=20)
    t = Table.from_rows(w["rows"])
    p = Plan(t.device(), 64, 64 * 24, 20, 0, mode=1)
    p.set_dense(False)
    for wmax, lb in [(1, 8), (300, 16), (5000, 32), (1, 8)]:
        r = w["ranges"].copy()
        r["weight"] = np.random.default_rng(wmax).integers(1, wmax + 1, len(r))
        items = t.compile(r, 64)
        p.upload(items, 64, np.full(64, 20, np.uint32))
        p.run()
        b = p.download()
        assert p.config()["lane_bits"] == lb
        a = plain(t, items, 64, 20)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"lane width {lb}"
    p.close()


def test_batched_compressed():
    w = W.cfg4(n=120_000, d=16, q=20)
    comp_t = Table.from_rows(w["rows"], 1024, True, codec=11)
    ri, rc = oracle_rows(w["rows"], w["ranges"], 20, w["k"])
    (b, _) = batched(comp_t, comp_t.compile(w["ranges"], 20), 20, w["k"])
    assert_topk_equal(b[0], b[1], ri, rc, w["k"], 20, "lane form, compressed")


def test_graph_replays_match_eager():
    w = W.cfg1(n=100_000, d=16, q=24, k=32)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 24)
    eager = plain(t, items, 24, 32)
    # replays with eager work and synchronisations between them
    (b, p) = batched(t, items, 24, 32, graph=True, runs=4)
    assert np.array_equal(eager[0], b[0]) and np.array_equal(eager[1], b[1])
    p0 = Plan(t.device(), 24, len(items), 32, 0, mode=0)
    p0.upload(items, 24, 32)
    p0.set_graph(True)
    for _ in range(3):
        p0.run()
        got = p0.download()
        assert np.array_equal(eager[0], got[0]) and np.array_equal(eager[1], got[1])
        plain(t, items[:50], 24, 5)  # eager work in between


def test_autotune_picks_a_valid_config():
    w = W.cfg1(n=200_000, d=32, q=32, k=100)
    t = Table.from_rows(w["rows"])
    items = t.compile(w["ranges"], 32)
    a = plain(t, items, 32, 100)
    p = Plan(t.device(), 32, len(items), 100, 0, mode=1)
    p.set_dense(False)
    p.upload(items, 32, 100)
    ms = p.autotune(reps=1)
    assert len(ms) == 13 and np.all(ms > 0)  # 12 lane configurations, then the plain form
    cfg = p.config()
    assert cfg["form"] == (["lane"] * 12 + ["plain"])[int(np.argmin(ms))]
    if cfg["form"] == "lane":
        assert cfg["block"] in (256, 512, 1024) and cfg["tile_objs"] in (8192, 16384)
    p.run()
    b = p.download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ri, rc = oracle_rows(w["rows"], w["ranges"], 32, 100)
    assert_topk_equal(b[0], b[1], ri, rc, 100, 32, "after autotune")


@pytest.mark.parametrize("form", ["plain", "lane"])
def test_batched_plan_forms_on_weighted_batches(form):
    """Weights 1-4 with the dense form switched off: a batched plan runs the
    lane or the plain form, switched explicitly or chosen by the autotuner,
    with the oracle's results either way; a graph replay of the chosen form
    too."""
    w = W.cfg1(seed=955, n=120_000, d=32, q=40, k=64)
    r = w["ranges"].copy()
    r["weight"] = genie_amd.synth.uniform_int(956, 0, len(r), 1, 4)
    t = Table.from_rows(w["rows"])
    items = t.compile(r, 40)
    ri, rc = oracle_rows(w["rows"], r, 40, 64)
    p = Plan(t.device(), 40, len(items), 64, 0, mode=1)
    p.set_dense(False)
    p.upload(items, 40, np.full(40, 64, np.uint32))
    p.set_form(form)
    assert p.form() == form
    p.run()
    b = p.download()
    assert_topk_equal(b[0], b[1], ri, rc, 64, 40, f"weighted, {form} form")
    ms = p.autotune(reps=1)
    assert len(ms) == 13 and p.form() in ("plain", "lane")
    p.set_graph(True)
    for _ in range(2):
        p.run()
        b = p.download()
        assert_topk_equal(b[0], b[1], ri, rc, 64, 40, f"weighted, autotuned {p.form()} form")
