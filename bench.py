"""Benchmark of the match-count top-k search on one MI355X.

Headline (value): 1M objects x 128 dims, values pi_d(rank-1) with rank ~
Zipf(1.1) over 256 values, 1,024 queries (random row + U{-3..3} noise), range
+-2 per dim (5 keys per dim), unit weights, k = 100; uncompressed posting
lists.  A step is one pass of the search over the batch with the items
already in HBM: device-side item compilation, scan/count/select, merge.
value = queries/s of the fastest form (the batched form; every form is listed
in config.forms).

The other workloads are timed in the same run (config.workloads), each form
against the plain form on the same batch and index, each with its results
compared with the CPU oracle on a fixed sample of queries:
  - the 1M x 128 rows compressed (SERIAL_DELTA_BP32, segments of <= 1,024);
  - 4M x 64 LSH buckets (8,192 per function), 4,096 exact-match queries,
    k = 100, uncompressed and SERIAL_DELTA_COMP_BP32_VARINT;
  - 2M set-valued documents (5-40 tokens, Zipf(1.05) over 1M), 16,384 queries
    of 1-64 tokens with weights 1-4 and k in {10, 100, 1000};
and the serving form (config.serving) on 10M x 128 rows built like the
headline's, resident in HBM compressed with SERIAL_DELTA_BP32: Poisson
arrivals at fixed offered loads, ranges {0, +-2, +-5}, k {10, 100, 1000}.

Prints one JSON line.  --gpus above 1 is refused: every box here has one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
# LDS: 256 CUs x 128 B per clock x 2.4 GHz (MI355X peak engine clock)
LDS_PEAK_GBS = 256 * 128 * 2.4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled(kind, kernel):
    """A per-dispatch figure of `kernel` from the committed rocprofv3
    summaries: lines `<kind> <kernel> <value>` in profiles/*.md (kind
    traffic_bytes_per_dispatch: FETCH_SIZE + WRITE_SIZE, copy-calibrated;
    lds_bytes_per_dispatch: SQ_INSTS_LDS x 64 lanes x 4 B)."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None, None
    for f in sorted(os.listdir(d)):
        if not f.endswith(".md"):
            continue
        with open(os.path.join(d, f)) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == kind and parts[1] == kernel:
                    return int(parts[2]), "profiles/" + f
    return None, None


def time_runs(plan, steps, warmup):
    import genie_amd
    for _ in range(warmup):
        plan.run()
    genie_amd.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    genie_amd.device_synchronize()
    return (time.perf_counter() - t0) * 1000 / steps


def stage_ms(plan, reps, names=("compile", "sample", "scan", "merge")):
    """Per-stage times (HIP events on the plan's stream), mean of reps runs."""
    plan.set_timing(True)
    acc = {k: [] for k in names}
    for _ in range(reps):
        plan.run()
        plan.download()
        for k in names:
            acc[k].append(plan.kernel_ms(k))
    plan.set_timing(False)
    return {k: float(np.mean(v)) for k, v in acc.items()}


def segment_bytes(export, items):
    """Algorithmic bytes of decoding the segments a compressed batch names,
    each once: their words read plus their postings written (4 B each)."""
    cpos = export["cpos"].astype(np.int64)
    upos = export["inv_pos"].astype(np.int64)
    s0 = np.searchsorted(cpos, items["start_pos"].astype(np.int64))
    s1 = np.searchsorted(cpos, items["end_pos"].astype(np.int64))
    named = np.zeros(len(cpos), np.int64)  # +1 at each run's first segment, -1 after its last
    np.add.at(named, s0, 1)
    np.add.at(named, s1, -1)
    seg = np.cumsum(named)[:-1] > 0
    words = int(np.sum((cpos[1:] - cpos[:-1])[seg]))
    posts = int(np.sum((upos[1:] - upos[:-1])[seg]))
    return 4 * (words + posts), int(np.sum(seg))


def kernel_row(kernel, workload, ms, algorithmic, what, key=None):
    """One kernel's live duration against its algorithmic bytes and the HBM
    bytes the committed profile of the same kernel and workload measured
    (profile line key: the kernel, or kernel@workload)."""
    row = dict(kernel=kernel, workload=workload, ms=ms, algorithmic_bytes=algorithmic, what=what,
               algorithmic_GBs=algorithmic / (ms * 1e-3) / 1e9)
    traffic, src = profiled("traffic_bytes_per_dispatch", key or kernel)
    if traffic:
        gbs = traffic / (ms * 1e-3) / 1e9
        row.update(hbm_bytes=traffic, hbm_GBs=gbs, hbm_frac=gbs / HBM_PEAK_GBS, source=src)
    return row


def sample_of(nq, n=256):
    return np.unique(np.linspace(0, nq - 1, min(n, nq)).astype(np.int64))


def rows_of(res, kq, qs):
    """Rows of queries qs from a batch result (per-query k)."""
    off = np.zeros(len(kq) + 1, np.int64)
    off[1:] = np.cumsum(kq)
    ids = np.concatenate([res[0][off[q]:off[q + 1]] for q in qs])
    cnt = np.concatenate([res[1][off[q]:off[q + 1]] for q in qs])
    return ids, cnt


def oracle_sample(words, items, n_objects, kq, qs):
    """The oracle's multithreaded search over queries qs (items of the
    uncompressed index, posting positions)."""
    from oracle import oracle as O
    sel = np.isin(items["query"], qs)
    sub = items[sel].copy()
    remap = np.full(int(items["query"].max()) + 1 if len(items) else 1, -1, np.int64)
    remap[qs] = np.arange(len(qs))
    sub["query"] = remap[sub["query"]]
    sub = sub[np.argsort(sub["query"], kind="stable")]
    ids, cnt, _ = O.search_mt(words, sub, n_objects, len(qs), np.asarray(kq)[qs].astype(np.int32))
    return ids, cnt


def to_posting_items(export, items):
    """Items of a compressed table (word positions on segment boundaries)
    translated to posting positions of its uncompressed lists."""
    cpos = export["cpos"].astype(np.int64)
    upos = export["inv_pos"].astype(np.int64)
    out = items.copy()
    s0 = np.searchsorted(cpos, items["start_pos"].astype(np.int64))
    s1 = np.searchsorted(cpos, items["end_pos"].astype(np.int64))
    out["start_pos"] = upos[s0]
    out["end_pos"] = upos[s1]
    return out


def plain_and_batched(index, items, nq, kq, steps, warmup, label, autotune=True):
    """Plain form (mode 0) and batched form (mode 1, autotuned over its forms
    and the plain form) on one batch; returns (forms, plain result, batched
    result, plain stage times)."""
    import genie_amd
    from genie_amd import Plan
    kmax = int(np.max(kq))
    forms = {}
    p = Plan(index, nq, len(items), kmax, 0, mode=0)
    p.upload(items, nq, kq)
    p.run()
    rp = p.download()  # settles the candidate space before timing
    ms = time_runs(p, steps, warmup)
    forms["plain"] = dict(ms_per_step=ms, qps=nq * 1000 / ms, kernel_ms=stage_ms(p, 3))
    p.close()
    b = Plan(index, nq, len(items), kmax, 0, mode=1)
    b.upload(items, nq, kq)
    if autotune:
        tune = b.autotune(reps=2)
        forms["batched_autotune_ms"] = [round(float(x), 4) for x in tune]
    b.set_graph(True)
    b.run()
    rb = b.download()
    ms = time_runs(b, steps, warmup)
    cfg = b.config()
    forms["batched"] = dict(ms_per_step=ms, qps=nq * 1000 / ms, config=cfg, kernel_ms=stage_ms(b, 3))
    forms["batched_vs_plain"] = forms["batched"]["qps"] / forms["plain"]["qps"]
    same = bool(np.array_equal(rp[0], rb[0]) and np.array_equal(rp[1], rb[1]))
    forms["batched_identical_to_plain"] = same
    b.close()
    genie_amd.device_synchronize()
    log(f"[bench] {label}: plain {forms['plain']['ms_per_step']:.3f} ms, batched {ms:.3f} ms ({cfg}), "
        f"x{forms['batched_vs_plain']:.2f}, identical {same}")
    return forms, rp, rb


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU baseline wall time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sample", type=int, default=256, help="queries of each workload checked against the oracle")
    ap.add_argument("--no-workloads", action="store_true", help="skip the compressed, LSH and set workloads")
    ap.add_argument("--serve-n", type=int, default=10_000_000, help="objects of the serving index")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="seconds of trace per offered load (0 = skip)")
    ap.add_argument("--serve-qps", type=int, nargs="+", default=[2000, 10000, 50000, 100000, 125000, 150000, 200000])
    ap.add_argument("--serve-max-requests", type=int, default=60000,
                    help="requests per offered load at most (a load runs min(qps x seconds, this))")
    ap.add_argument("--serve-executors", type=int, default=4)
    ap.add_argument("--serve-batch", type=int, default=256)
    ap.add_argument("--serve-codec", type=int, default=11,
                    help="codec of the serving index (segments of <= 1,024; default SERIAL_DELTA_BP32)")
    ap.add_argument("--serve-plain-qps", type=int, nargs="*", default=[2000],
                    help="offered loads also served with plain-form executors (posting lists decoded on the device)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device copy of this size (scales FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} refused: the search runs on one GPU here; "
              "an index sharded by object would run one such process per GPU", file=sys.stderr)
        sys.exit(2)

    import genie_amd
    from genie_amd import Plan, Table, workloads as W

    if args.copy_probe_mib:
        from genie_amd._native import copy_probe
        copy_probe(args.copy_probe_mib << 20)

    # ---- headline: 1M x 128, 1,024 queries, k = 100 --------------------------
    t0 = time.time()
    w = W.cfg1(n=args.n, q=args.queries)
    table = Table.from_rows(w["rows"])
    items = table.compile(w["ranges"], w["n_queries"])
    index = table.device()
    log(f"[bench] 1M index built + uploaded in {time.time() - t0:.1f}s: {len(items)} items, "
        f"{index.bytes() / 2**20:.0f} MiB resident")
    nq = w["n_queries"]
    kq = np.full(nq, w["k"], np.uint32)
    forms, res_plain, res_batched = plain_and_batched(index, items, nq, kq, args.steps, args.warmup,
                                                      "1M x 128", autotune=True)
    p = Plan(index, nq, len(items), w["k"])
    p.upload(items, nq, kq)
    logical, compulsory = p.bytes()
    p.close()
    best = "batched" if forms["batched"]["qps"] >= forms["plain"]["qps"] else "plain"
    kern = forms["plain"]["kernel_ms"]

    workloads = {}
    kernels = []  # per-kernel HBM rows
    if not args.no_workloads:
        # ---- the same rows compressed: plain form, segments decoded once per batch
        t0 = time.time()
        ctab = Table.from_rows(w["rows"], 1024, True, codec=11)
        cidx = ctab.device()
        citems = ctab.compile(w["ranges"], nq, merge_runs=True)
        c = Plan(cidx, nq, len(citems), w["k"], 0, mode=0)
        c.upload(citems, nq, kq)
        c.run()
        rc_ = c.download()
        ms = time_runs(c, max(3, args.steps // 2), 1)
        ent = dict(ms_per_step=ms, qps=nq * 1000 / ms,
                   kernel_ms=stage_ms(c, 3, ("compile", "scan", "decode", "merge")), decode_once=c.decode_active(),
                   identical_to_uncompressed=bool(np.array_equal(rc_[0], res_plain[0]) and
                                                  np.array_equal(rc_[1], res_plain[1])))
        c.set_decode(0)
        c.run()
        c.download()
        ms0 = time_runs(c, 2, 0)
        ent["in_kernel_decode_ms_per_step"] = ms0
        ent["vs_uncompressed_plain"] = ms / forms["plain"]["ms_per_step"]
        c.close()
        dec_bytes, dec_segs = segment_bytes(ctab.export(), citems)
        ent["decoded_segments"] = dec_segs
        if ent["kernel_ms"]["decode"] > 0:
            kernels.append(kernel_row("k_decode_jobs", "1M x 128 compressed (SERIAL_DELTA_BP32)",
                                      ent["kernel_ms"]["decode"], dec_bytes,
                                      "the batch's segments decoded once: words read + postings written"))
        workloads["cfg1_compressed_plain"] = dict(
            workload="1M x 128 as above, compressed SERIAL_DELTA_BP32 (segments <= 1,024), plain form",
            index_mib=cidx.bytes() / 2**20, **ent)
        log(f"[bench] compressed 1M plain: {ms:.2f} ms (in-kernel decode {ms0:.2f} ms), identical "
            f"{ent['identical_to_uncompressed']} ({time.time() - t0:.1f}s)")
        del cidx, ctab, citems

    if not args.no_workloads:
        # ---- the same batch with weights 1-4 (the dense form's weighted
        # masks): plain and batched forms
        t0 = time.time()
        wr = w["ranges"].copy()
        wr["weight"] = genie_amd.synth.uniform_int(1011, 0, len(wr), 1, 4)
        witems = table.compile(wr, nq)
        fw, rwp, rwb = plain_and_batched(index, witems, nq, kq, max(5, args.steps // 2), 2, "1M x 128 weighted")
        qsw = sample_of(nq, 64)
        owi, owc = oracle_sample(table.export()["inv"].view(np.uint32), witems, w["rows"].shape[0], kq, qsw)
        fw["bit_exact_vs_oracle_sample"] = bool(all(np.array_equal(a, b) for a, b in
                                                    zip(rows_of(rwb, kq, qsw) + rows_of(rwp, kq, qsw),
                                                        (owi, owc, owi, owc))))
        workloads["cfg1_weighted"] = dict(workload="the headline batch with weights uniform in {1,2,3,4}",
                                          sample=len(qsw), **fw)
        log(f"[bench] weighted: oracle sample exact {fw['bit_exact_vs_oracle_sample']} ({time.time() - t0:.1f}s)")
        del witems

    words1 = table.export()["inv"].view(np.uint32)
    qs1 = sample_of(nq, args.sample)
    oi, oc = oracle_sample(words1, items[np.argsort(items["query"], kind="stable")], w["rows"].shape[0], kq, qs1)
    gi, gc = rows_of(res_batched, kq, qs1)
    exact_sample = bool(np.array_equal(gi, oi) and np.array_equal(gc, oc))

    if not args.no_workloads:
        # ---- LSH: 4M x 64 buckets, 4,096 exact-match queries ----------------
        t0 = time.time()
        w2 = W.cfg2()
        nq2 = w2["n_queries"]
        kq2 = np.full(nq2, w2["k"], np.uint32)
        t2 = Table.from_rows(w2["rows"])
        i2 = t2.compile(w2["ranges"], nq2)
        ix2 = t2.device()
        f2, r2p, r2b = plain_and_batched(ix2, i2, nq2, kq2, max(5, args.steps // 2), 2, "LSH 4M x 64")
        pb = Plan(ix2, nq2, len(i2), w2["k"])
        pb.upload(i2, nq2, kq2)
        _, lsh_compulsory = pb.bytes()
        pb.close()
        qs2 = sample_of(nq2, args.sample)
        oi2, oc2 = oracle_sample(t2.export()["inv"].view(np.uint32), i2, w2["rows"].shape[0], kq2, qs2)
        f2["bit_exact_vs_oracle_sample"] = bool(all(np.array_equal(a, b) for a, b in
                                                    zip(rows_of(r2b, kq2, qs2) + rows_of(r2p, kq2, qs2),
                                                        (oi2, oc2, oi2, oc2))))
        del ix2
        c2 = Table.from_rows(w2["rows"], 1024, True, codec=13)
        ci2 = c2.compile(w2["ranges"], nq2, merge_runs=True)
        cx2 = c2.device()
        cp2 = Plan(cx2, nq2, len(ci2), w2["k"], 0, mode=0)
        cp2.upload(ci2, nq2, kq2)
        cp2.run()
        r2c = cp2.download()
        ms = time_runs(cp2, max(5, args.steps // 2), 2)
        f2["plain_compressed"] = dict(ms_per_step=ms, qps=nq2 * 1000 / ms,
                                      kernel_ms=stage_ms(cp2, 3, ("compile", "scan", "decode", "merge")),
                                      codec="SERIAL_DELTA_COMP_BP32_VARINT", index_mib=cx2.bytes() / 2**20,
                                      identical_to_uncompressed=bool(np.array_equal(r2c[0], r2p[0]) and
                                                                     np.array_equal(r2c[1], r2p[1])))
        cp2.close()
        kernels.append(kernel_row("k_scan_hash", "4M x 64 LSH, 4,096 queries (uncompressed)",
                                  f2["plain"]["kernel_ms"]["scan"], lsh_compulsory,
                                  "distinct postings of the batch x 4 B (each read by one tile)"))
        d2b, _ = segment_bytes(c2.export(), ci2)
        if f2["plain_compressed"]["kernel_ms"]["decode"] > 0:
            kernels.append(kernel_row("k_decode_jobs", "4M x 64 LSH compressed (SERIAL_DELTA_COMP_BP32_VARINT)",
                                      f2["plain_compressed"]["kernel_ms"]["decode"], d2b,
                                      "the batch's segments decoded once: words read + postings written",
                                      key="k_decode_jobs@lsh4Mx64"))
        workloads["cfg2_lsh"] = dict(workload="4M x 64 LSH buckets U[0,8192), 4,096 queries (entries redrawn "
                                              "w.p. 0.3), exact match, k=100", sample=len(qs2), **f2)
        log(f"[bench] LSH compressed plain: {ms:.2f} ms; oracle sample exact {f2['bit_exact_vs_oracle_sample']} "
            f"({time.time() - t0:.1f}s)")
        del cx2, c2, t2, w2, i2, ci2

        # ---- sets: 2M documents, 16,384 weighted queries, mixed k ------------
        t0 = time.time()
        w3 = W.cfg3()
        nq3 = w3["n_queries"]
        kq3 = w3["k"]
        t3 = Table.from_sets(w3["row_ptr"], w3["vals"])
        i3 = t3.compile(w3["ranges"], nq3)
        ix3 = t3.device()
        f3, r3p, r3b = plain_and_batched(ix3, i3, nq3, kq3, max(3, args.steps // 4), 1, "sets 2M")
        qs3 = sample_of(nq3, args.sample)
        oi3, oc3 = oracle_sample(t3.export()["inv"].view(np.uint32), i3, len(w3["row_ptr"]) - 1, kq3, qs3)
        f3["bit_exact_vs_oracle_sample"] = bool(all(np.array_equal(a, b) for a, b in
                                                    zip(rows_of(r3b, kq3, qs3) + rows_of(r3p, kq3, qs3),
                                                        (oi3, oc3, oi3, oc3))))
        if f3["batched"]["config"].get("form") == "set":
            p3 = Plan(ix3, nq3, len(i3), int(np.max(kq3)))
            p3.upload(i3, nq3, kq3)
            _, sets_compulsory = p3.bytes()
            p3.close()
            kernels.append(kernel_row("k_dense_count_sets", "2M sets, 16,384 queries (set form main pass)",
                                      f3["batched"]["kernel_ms"]["scan"], sets_compulsory,
                                      "distinct postings of the batch x 4 B"))
        workloads["cfg3_sets"] = dict(workload="2M sets (5-40 tokens, Zipf(1.05) over 1M), 16,384 queries of "
                                               "1-64 tokens, weights 1-4, k in {10,100,1000}",
                                      sample=len(qs3), **f3)
        log(f"[bench] sets: oracle sample exact {f3['bit_exact_vs_oracle_sample']} ({time.time() - t0:.1f}s)")
        del ix3, t3, w3, i3

    # ---- serving form: 10M x 128 compressed, Poisson arrivals -----------------
    serving = None
    if args.serve_seconds > 0:
        serving = serve(args, log)

    # ---- CPU baseline on the headline batch ------------------------------------
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        srt = items[np.argsort(items["query"], kind="stable")]
        O.search_mt(words1, srt[srt["query"] < 2], w["rows"].shape[0], 2, w["k"])  # warm
        rounds, t0 = 0, time.perf_counter()
        while True:  # whole batches until the minimum wall time
            ids, cnt, threads = O.search_mt(words1, srt, w["rows"].shape[0], nq, w["k"])
            rounds += 1
            cdt = time.perf_counter() - t0
            if cdt >= args.cpu_seconds:
                break
        cpu = dict(value=rounds * nq / cdt, unit="queries/s", cores=threads, kind="port",
                   sample=f"{rounds} x the full {nq}-query batch, oracle/oracle.c orc_search_mt "
                          f"(dense counter array + exact top-k per query), {cdt:.1f} s wall")
        gi, gc = res_batched if best == "batched" else res_plain
        exact_full = bool(np.array_equal(gi, ids) and np.array_equal(gc, cnt))
        log(f"[bench] cpu baseline: {cpu['value']:.1f} queries/s on {threads} threads; GPU == oracle: {exact_full}")
    else:
        exact_full = None

    # roofline of the dominant kernel of the reported form
    scan_ms = forms[best]["kernel_ms"]["scan"]
    if best == "batched":
        dense = forms["batched"]["config"].get("dense")
        scan_kernel = ("k_dense_count (bit-sliced count + candidate emission, main pass)" if dense
                       else "k_scan_batched (count + select)")
    else:
        scan_kernel = "k_scan_tile (count + select)"
    kname = scan_kernel.split()[0]
    traffic, traffic_src = profiled("traffic_bytes_per_dispatch", kname)
    lds_bytes, lds_src = profiled("lds_bytes_per_dispatch", kname)
    achieved = compulsory / (scan_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                "kernel": scan_kernel, "effective_logical_GBs": logical / (scan_ms * 1e-3) / 1e9}
    kernels[:0] = [kernel_row("k_dense_count", "headline batch (dense main pass + check + recount)",
                              forms["batched"]["kernel_ms"]["scan"], compulsory,
                              "distinct postings of the batch x 4 B"),
                   kernel_row("k_scan_tile", "headline batch (plain form)", forms["plain"]["kernel_ms"]["scan"],
                              compulsory, "distinct postings of the batch x 4 B")]
    if lds_bytes:
        # LDS bytes the kernel reads per dispatch (profile) over its live duration
        lds = lds_bytes / (scan_ms * 1e-3) / 1e9
        roofline["lds"] = {"achieved": lds, "peak": LDS_PEAK_GBS, "unit": "GB/s", "frac": lds / LDS_PEAK_GBS,
                           "bytes_per_dispatch": lds_bytes, "source": lds_src}
    out = {
        "metric": "queries/s (1M x 128, k=100) batched vs plain GPU; p50/p99 latency at fixed QPS",
        "value": forms[best]["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": forms[best]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (deterministic generator, seed 1001)",
        "config": {
            "workload": f"{args.n} objects x 128 dims Zipf(1.1) values, {nq} queries, range +-2, "
                        f"k={w['k']}, uncompressed",
            "form": best,
            "forms": forms,
            "kernel_ms": kern,
            "bytes_logical": logical,
            "bytes_compulsory": compulsory,
            "bit_exact_vs_oracle_full_batch": exact_full,
            "bit_exact_vs_oracle_sample": exact_sample,
            "workloads": workloads,
            "kernels": kernels,
            "serving": serving,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def serve(args, log):
    """Serving on a 10M x 128 index (rows built like the headline's),
    resident compressed; returns the config.serving object."""
    import genie_amd
    from genie_amd import Plan, Server, Table, synth, workloads as W
    t0 = time.time()
    n = args.serve_n
    rows = synth.rows(1005, n, 128, synth.ZIPF_PERM, 1.1, 0.0, 0, 255, 256)
    stab = Table.from_rows(rows, 1024, True, codec=args.serve_codec)
    sindex = stab.device()
    log(f"[bench] serving index {n} x 128 built + uploaded in {time.time() - t0:.1f}s, "
        f"{sindex.bytes() / 2**30:.2f} GiB resident")
    nreq = min(int(max(args.serve_qps) * args.serve_seconds), args.serve_max_requests)
    sq, _ = synth.queries(7001, rows, nreq, noise=3, lo=0, hi=255)
    radius = np.array([0, 2, 5], np.int32)[synth.uniform_int(7002, 0, nreq, 0, 2)]
    sk = np.array([10, 100, 1000], np.uint32)[synth.uniform_int(7003, 1, nreq, 0, 2)]
    sranges = W.single_ranges(sq, radius)
    sitems = stab.compile(sranges, nreq, merge_runs=True)
    sptr = np.searchsorted(sitems["query"], np.arange(nreq + 1)).astype(np.uint64)
    serving = {"index": f"{n} objects x 128 dims Zipf(1.1) (built like the headline rows, seed 1005), "
                        f"compressed codec {args.serve_codec} (segments <= 1,024 ids), resident; requests as "
                        f"merged segment runs",
               "index_objects": n, "index_gib": sindex.bytes() / 2**30,
               "requests": "ranges {0,+-2,+-5}, k {10,100,1000}, Poisson arrivals",
               "executors": args.serve_executors, "max_batch": args.serve_batch}

    # the plain per-batch search on the same index and request mix: the first
    # 1,024 requests as one batch (posting lists decoded once per batch)
    nb = min(1024, nreq)
    bi = sitems[: int(sptr[nb])]
    pp = Plan(sindex, nb, len(bi), 1000, 0, mode=0)
    pp.upload(bi, nb, sk[:nb])
    pp.run()
    plain_res = pp.download()
    pms = time_runs(pp, 3, 1)
    serving["plain_same_mix"] = dict(ms_per_batch=pms, qps=nb * 1000 / pms, batch=nb, kernel_ms=stage_ms(pp, 2),
                                     decode_once=pp.decode_active())
    pp.close()
    log(f"[bench] serving: plain form on {nb} requests of the mix: {pms:.1f} ms, {nb * 1000 / pms:.0f} queries/s")

    def ladder(mode, loads, label):
        srv = Server(sindex, executors=args.serve_executors, max_batch=args.serve_batch, max_k=1000, mode=mode)
        nw = min(nreq, 2 * args.serve_batch * args.serve_executors)
        srv.replay(sitems[: int(sptr[nw])], sptr[: nw + 1], sk[:nw], np.zeros(nw))  # untimed warm-up
        srv.slowest(reset=True)
        res, first = [], None
        for qps in loads:
            m = min(int(qps * args.serve_seconds), nreq)
            arr = np.cumsum(synth.exponential(7004 + qps, m, float(qps)))
            lat, ids, cnt, wall = srv.replay(sitems[: int(sptr[m])], sptr[: m + 1], sk[:m], arr)
            if first is None:
                first = (m, ids, cnt)
            e = dict(offered_qps=qps, requests=m, achieved_qps=m / wall, p50_ms=float(np.percentile(lat, 50)),
                     p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
                     slowest_ms={k: v * 1e3 for k, v in srv.slowest(reset=True).items()})
            res.append(e)
            log(f"[bench] serving ({label}) @ {qps} QPS: achieved {e['achieved_qps']:.0f}, p50 {e['p50_ms']:.2f} ms, "
                f"p99 {e['p99_ms']:.2f} ms")
            if e["achieved_qps"] < 0.9 * qps:
                break  # saturated: higher loads only queue
        st = srv.stats()
        srv.close()
        return res, st, first

    loads, st, first = ladder(2, args.serve_qps, "automatic")
    serving["mode"] = "automatic"
    serving["loads"] = loads
    serving["stats"] = st
    kept = [e for e in loads if e["achieved_qps"] >= 0.97 * e["offered_qps"]]
    serving["sustained_qps"] = max((e["achieved_qps"] for e in kept), default=None)
    serving["sustained_vs_plain_same_mix"] = (serving["sustained_qps"] / serving["plain_same_mix"]["qps"]
                                              if serving["sustained_qps"] else None)
    if args.serve_plain_qps:
        ploads, pst, _ = ladder(0, args.serve_plain_qps, "plain executors")
        serving["plain_executors"] = dict(loads=ploads, stats=pst,
                                          note="executors on the plain form: posting lists decoded on the device "
                                               "per batch (no forward codes)")
    # served results of the first load against the oracle, on a sample
    m, ids, cnt = first
    qs = sample_of(m, args.sample)
    ex = stab.export()
    uitems = to_posting_items(ex, sitems[: int(sptr[m])])
    oi, oc = oracle_sample(ex["inv"].view(np.uint32), uitems, n, sk[:m], qs)
    gi, gc = rows_of((ids, cnt), sk[:m], qs)
    serving["bit_exact_vs_oracle_sample"] = bool(np.array_equal(gi, oi) and np.array_equal(gc, oc))
    pq = qs[qs < nb]
    pi, pc = rows_of(plain_res, sk[:nb], pq)
    si, sc = rows_of((ids, cnt), sk[:m], pq)
    serving["served_equals_plain_sample"] = bool(np.array_equal(pi, si) and np.array_equal(pc, sc))
    serving["oracle_sample"] = len(qs)
    log(f"[bench] serving: served == oracle on {len(qs)} requests: {serving['bit_exact_vs_oracle_sample']}; "
        f"served == plain: {serving['served_equals_plain_sample']}")
    del ex, uitems
    genie_amd.device_synchronize()
    return serving


if __name__ == "__main__":
    main()
