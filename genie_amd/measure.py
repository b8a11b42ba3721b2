"""Times the search forms on the synthetic workloads other than the headline
one (bench.py measures 1M x 128, k = 100), with the batch already in HBM:

  zipf-compressed  1M x 128 Zipf(1.1), range +-2, k = 100, SERIAL_DELTA_BP32 segments
  lsh              4M x 64 LSH buckets (8,192 per dim), exact match, 4,096 queries, k = 100
  lsh-compressed   the same index with SERIAL_DELTA_COMP_BP32_VARINT segments
  sets             2M Zipf(1.05) token sets over a 1M vocabulary, 16,384 weighted queries, mixed k
  weighted         the headline batch (1M x 128, 1,024 queries, +-2, k = 100) with weights
                   uniform in {1, 2, 3, 4} (the dense form's weighted masks)

    python genie_amd/measure.py lsh [--reps 5] [--forms plain-hash plain-direct ...] [--merge-runs]

Compressed workloads: the plain forms decode the batch's segments once into a
full-layout scratch (the default); the *-kdecode forms leave the decode to
the kernels.  --merge-runs compiles one item per range (its segments joined).

Every form's results are compared with the first form's.  One line per form:
ms per batch (run + download), queries/s, per-stage kernel times (HIP
events), whether the results are identical.
"""
import argparse
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

# form name -> (plan mode, hash form, sparse select, dense/set form, its threads per workgroup,
#               decode of a compressed index: -1 once per batch, 0 in the kernels)
FORMS = {
    "plain-hash": (0, 1, -1, True, 0, -1),      # every query whose counts fit on the hash form
    "plain-auto": (0, -1, -1, True, 0, -1),     # the plain form's own per-query choice
    "plain-direct": (0, 0, -1, True, 0, -1),    # direct-mapped kernel, its own select choice
    "plain-direct-dense": (0, 0, 0, True, 0, -1),
    "plain-auto-kdecode": (0, -1, -1, True, 0, 0),
    "plain-direct-kdecode": (0, 0, -1, True, 0, 0),
    "batched": (1, -1, -1, True, 0, -1),        # dense (or set) form when the batch allows, else the lane form
    "batched-512": (1, -1, -1, True, 512, -1),  # the same with 512-thread workgroups
    "batched-lanes": (1, -1, -1, False, 0, -1),  # the lane form
    "automatic": (2, -1, -1, True, 0, -1),
}


def workload(name, merge_runs=False):
    from genie_amd import Table, workloads as W
    if name in ("zipf-compressed",):
        w = W.cfg1()
        t = Table.from_rows(w["rows"], 1024, True, codec=11)
    elif name in ("lsh", "lsh-compressed"):
        w = W.cfg2()
        t = Table.from_rows(w["rows"], 1024, True, codec=13) if name == "lsh-compressed" else Table.from_rows(w["rows"])
    elif name == "sets":
        w = W.cfg3()
        t = Table.from_sets(w["row_ptr"], w["vals"])
    elif name == "weighted":
        import genie_amd
        w = dict(W.cfg1())
        w["ranges"] = w["ranges"].copy()
        w["ranges"]["weight"] = genie_amd.synth.uniform_int(1011, 0, len(w["ranges"]), 1, 4)  # as bench.py
        t = Table.from_rows(w["rows"])
    else:
        raise SystemExit(f"unknown workload {name}")
    items = t.compile(w["ranges"], w["n_queries"], merge_runs=merge_runs)
    return t, items, w["n_queries"], w["k"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["zipf-compressed", "lsh", "lsh-compressed", "sets", "weighted"])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--forms", nargs="+", default=["plain-auto", "plain-direct", "automatic"], choices=list(FORMS))
    ap.add_argument("--merge-runs", action="store_true", help="one item per range over a compressed index")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device copy of this size (scales FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    from genie_amd import Plan
    if args.copy_probe_mib:
        from genie_amd._native import copy_probe
        copy_probe(args.copy_probe_mib << 20)
    t0 = time.time()
    t, items, nq, k = workload(args.workload, args.merge_runs)
    print(f"{args.workload}: {nq} queries, {len(items)} items, built in {time.time() - t0:.1f} s", flush=True)
    kq = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
    ref = None
    for f in args.forms:
        mode, hm, sp, dn, nt, dec = FORMS[f]
        p = Plan(t.device(), nq, len(items), int(kq.max()), 0, mode=mode)
        p.set_hash(hm)
        p.set_sparse(sp)
        p.set_decode(dec)
        if mode:
            p.set_dense(dn)
            p.set_dense_threads(nt)
        p.upload(items, nq, kq)
        p.run()
        got = p.download()
        if ref is None:
            ref = got
        same = bool(np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]))
        p.set_timing(True)
        stages = {s: [] for s in ("compile", "sample", "decode", "scan", "merge")}
        t1 = time.perf_counter()
        for _ in range(args.reps):
            p.run()
            p.download()
            for s in stages:
                stages[s].append(p.kernel_ms(s))
        dt = (time.perf_counter() - t1) / args.reps
        ks = {s: round(float(np.mean(v)), 3) for s, v in stages.items()}
        extra = (f"form={p.form()}" if mode else
                 f"hash_queries={p.hash_queries()} sparse={p.sparse_active()} decode_once={p.decode_active()}")
        print(f"{args.workload} {f}: {dt * 1e3:.2f} ms/batch, {nq / dt:.0f} queries/s, {extra}, kernels {ks}, "
              f"identical={same}", flush=True)


if __name__ == "__main__":
    main()
