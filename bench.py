"""Benchmark: batched banded Smith-Waterman on one MI355X.

Workload (synthetic, seed 1): 1,000,000 candidate regions, each a (query
window, subject window, centre diagonal) triple; window lengths lognormal
(median 150, sigma 0.6, clamped 20..2000); band half-widths uniform in
{8, 16, 32, 64}; 40 % planted homologues; BLOSUM62, gap open 11 / extend 1;
no composition bias.

A step is one scoring pass over the whole batch with every input already
resident in HBM: all band kernels, the exact int32 rescore of saturated jobs
and the output clear. The work unit is the DP cell (banded_score's cell
count); value = cells / step time in GCUPS.

The printed JSON line also carries
  roofline      the dominant kernel variant: 10 integer ops per cell over its
                mean launch time (HIP events on the library's stream) against
                the packed-int16 VALU peak of MI355X;
  cpu_baseline  the CPU oracle (the reference's scalar banded kernel restated
                in C) on all host threads over a bounded sample of the same
                batch.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X: 256 CUs x 64 lanes x 2.4 GHz; packed int16 VALU ops count twice.
VALU_I32_TOPS = 256 * 64 * 2.4e9 / 1e12
VALU_I16X2_TOPS = 2 * VALU_I32_TOPS
OPS_PER_CELL = 10


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--regions", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=3000, help="jobs checked against the oracle")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this "
              "single-GPU build (see DESIGN.md for how the path shards)", file=sys.stderr)
        return 2

    import saber_amd as S
    from saber_amd import BLOSUM62, workload

    t0 = time.time()
    d = workload.band_regions(seed=args.seed, n=args.regions)
    gen_s = time.time() - t0
    sch = S.Scheme(BLOSUM62, 11, 1)
    ctx = S.Context(0)
    t0 = time.time()
    batch = ctx.band_batch(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch)
    plan_s = time.time() - t0
    print(f"# workload {args.regions} regions generated in {gen_s:.1f}s, planned+uploaded in "
          f"{plan_s:.1f}s", file=sys.stderr, flush=True)

    ms, per, cells = batch.time(warmup=args.warmup, iters=args.steps)
    score, q_end, s_end, cell_arr = batch.read()

    variants = S.band_variants()
    dom = max(per, key=lambda v: per[v][0])
    dom_ms, dom_cells = per[dom]
    L, Cc, K, B = variants[dom]
    achieved = OPS_PER_CELL * dom_cells / (dom_ms * 1e-3) / 1e12
    gcups = cells / (ms * 1e-3) / 1e9
    print(f"# step {ms:.3f} ms, {cells} cells, {gcups:.1f} GCUPS; dominant variant "
          f"band_pair_kernel<{L},{Cc},{K},{'true' if B else 'false'}> {dom_ms:.3f} ms",
          file=sys.stderr, flush=True)

    # Parity spot check against the oracle on a deterministic sample.
    from oracle import oracle as O
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(args.regions, size=min(args.check, args.regions), replace=False))
    sub = d["jobs"][idx]
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], sub, BLOSUM62, 11, 1)
    got = (score[idx], q_end[idx], s_end[idx], cell_arr[idx])
    mismatches = int(sum(int((np.asarray(g).astype(np.int64) != np.asarray(w).astype(np.int64)).sum())
                         for g, w in zip(got, want)))
    if mismatches:
        print(f"# PARITY FAILURE: {mismatches} mismatching fields in {len(idx)} sampled jobs",
              file=sys.stderr)

    cpu = None
    if not args.no_cpu:
        nthreads = os.cpu_count() or 1
        # Calibrate on a small prefix, then time a sample sized to ~cpu-seconds.
        probe = d["jobs"][:2000]
        t0 = time.perf_counter()
        pc = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], probe, BLOSUM62, 11, 1,
                           nthreads=nthreads)[3].sum()
        dt = time.perf_counter() - t0
        rate = pc / max(dt, 1e-6)
        ncells = rate * args.cpu_seconds
        mean_cells = cells / args.regions
        nsample = int(min(args.regions, max(2000, ncells / mean_cells)))
        sample = d["jobs"][:nsample]
        t0 = time.perf_counter()
        sc = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], sample, BLOSUM62, 11, 1,
                           nthreads=nthreads)[3].sum()
        dt = time.perf_counter() - t0
        cpu = {"value": round(sc / dt / 1e9, 3), "unit": "GCUPS", "cores": nthreads, "kind": "port",
               "sample": f"first {nsample} regions of the same batch ({int(sc)} cells, {dt:.1f} s), "
                         "scalar banded kernel (banded.rs:134-224) restated in C, one job per thread"}

    out = {
        "metric": "banded-SW GCUPS",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i16x2 (exact int32 rescore on saturation)",
        "data": "synthetic",
        "config": {
            "workload": "batched banded SW: 1e6 regions (query window, subject window, centre "
                        "diagonal), lognormal lengths median 150 sigma 0.6 [20,2000], half-widths "
                        "{8,16,32,64}, 40% planted homologues, BLOSUM62 11/1, no bias",
            "regions": args.regions,
            "cells_per_step": int(cells),
            "parity_sample": int(len(idx)),
            "parity_mismatches": mismatches,
        },
        "roofline": {
            "bound": "valu",
            "kernel": f"band_pair_kernel<{L}, {Cc}, {K}, {'true' if B else 'false'}>",
            "achieved": round(achieved, 3),
            "peak": round(VALU_I16X2_TOPS, 2),
            "unit": "Tops/s (int16 ops, 10 per cell)",
            "frac": round(achieved / VALU_I16X2_TOPS, 4),
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    ctx.close()
    return 0 if mismatches == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
