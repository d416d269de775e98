"""Benchmark: batched banded Smith-Waterman on one MI355X.

Workload (synthetic, seed 1): 1,000,000 candidate regions, each a (query
window, subject window, centre diagonal) triple, 1000 regions per query; window lengths lognormal
(median 150, sigma 0.6, clamped 20..2000); band half-widths uniform in
{8, 16, 32, 64}; 40 % planted homologues; BLOSUM62, gap open 11 / extend 1;
no composition bias.

A step is one scoring pass over the whole batch with every input already
resident in HBM, computing what score_banded_many returns (one score per
job): the packed band kernel, the exact int32 rescore of saturated jobs and
the output clear. The work unit is the DP cell (banded_score's cell count);
value = cells / step time in GCUPS. The build that also finds banded_score's
end cell is timed beside it (extra.band_with_end_cells).

The printed JSON line also carries
  roofline      the packed band kernel (one launch per step, every shape):
                10 integer ops per band cell it scores over its mean launch
                time (HIP events on the library's stream) against the
                packed-int16 VALU peak of MI355X;
  cpu_baseline  the AVX2 restatement of the reference's score_banded_many
                (oracle/cpu_avx2.c) on all host threads over a bounded sample
                of whole queries of the same batch;
  extra.xdrop   X-drop extension of 4 seed hits per region (steps/s), with
                the AVX2 restatement of xdrop_avx2 (oracle/cpu_avx2.c) timed
                on all host threads over the first 2 M hits beside it;
  extra.traceback  banded traceback of the best-scoring regions, stopping at
                their exact score (alignments/s, the oracle on one core), and
                the engine's widening loop over them as hits (wall time);
  extra.full_sw exhaustive full-matrix SW (queries x subjects, GCUPS; scores
                only, as score_many, and with the end cell; shapes changed by
                the autotuner);
  extra.search  the blastp workloads end to end on the GPU: 100 x 5000
                (hit lists against the oracle's plain CPU search, also
                timed) and 11,206 queries against ~148 M residues
                (queries/s, stage times, counters, size-independent checks);
  extra.prefilter  the learned region prefilter on the at-scale workload:
                trained on ~22 M regions the plain search labels, reported
                hits lost, regions kept, queries/s plain vs filtered, MLP
                forward throughput against the bf16 MFMA peak.
"""
import argparse
import gc
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
ROOFLINE_KERNEL = "band_multi_kernel<false, false>"


def measured_traffic(args):
    """HBM bytes per launch of the band kernel for this workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised under profiles/ (each
    counter scaled by what it reads for a plain copy of known size), or None
    when no measurement of this exact workload is recorded."""
    path = os.path.join(ROOT, "profiles", "band_multi_kernel_1M_regions.traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    wl = rec.get("workload", {})
    if (wl.get("regions"), wl.get("seed"), wl.get("per_query")) != (args.regions, args.seed, args.per_query):
        return None
    if rec.get("kernel") != ROOFLINE_KERNEL:
        return None
    return int(rec["traffic_bytes"])


def prefilter_leg(ctx, args, sch):
    """Learned region prefilter (64 -> 256 -> 256 -> 1, bf16 MFMA kernels)
    on the at-scale blastp workload (config 4 on config 3's workload).

    Training: every capped region of the plain GPU search over two workloads
    of the at-scale shape (family sets, seeds 11 and 12: ~22 M regions),
    labelled by the plain search itself (its exact score passes E <= 10);
    features computed on the device, standardised with the training
    statistics, resident in HBM; Adam on minibatches of 65,536. Threshold:
    at most --pf-loss of the regions the plain search reported a hit from
    are pruned on a third workload (seed 13). Test: the at-scale leg's own
    workload (seed 0), plain vs filtered search: reported (query, subject)
    hits lost, regions kept, queries/s end to end and for the
    candidate-scoring stage (regions to hits: plan, prefilter, scores,
    traceback and E-values), band cells scored; MLP forward throughput over
    the test regions against the bf16 MFMA peak."""
    from saber_amd import prefilter as P, search, stats, workload
    k = stats.gapped_protein(62, 11, 1)
    nq, nd = args.search_queries, args.search_decoys
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    t0 = time.time()
    train = [workload.family_set(seed=s, nq=nq, ndecoy=nd) for s in (11, 12)]
    valid = workload.family_set(seed=13, nq=nq, ndecoy=nd)
    pf = search.train_prefilter(ctx, train, valid, sch, k, loss_target=args.pf_loss, epochs=args.pf_epochs,
                                batch=65536, seed=args.seed, log=log)
    train_s = time.time() - t0
    del train, valid
    test = workload.family_set(seed=0, nq=nq, ndecoy=nd)
    ctx.upload_db(test["sres"], test["soff"])
    search.run(ctx, test["qres"], test["qoff"], sch, k)     # the database's seed-scan data
    search.run(ctx, test["qres"], test["qoff"], sch, k, prefilter=search.prefilter_arg(pf))   # and the filter's buffers
    # both timed searches return the hits only (no region records); Python's
    # cyclic garbage collector is paused around them (the engine has none)
    gc.collect()
    gc.disable()
    t1 = time.perf_counter()
    plain = search.run(ctx, test["qres"], test["qoff"], sch, k)
    tp = time.perf_counter() - t1
    t1 = time.perf_counter()
    filt = search.run(ctx, test["qres"], test["qoff"], sch, k, prefilter=search.prefilter_arg(pf))
    tf = time.perf_counter() - t1
    gc.enable()
    # the plain search again, keeping its regions for their features (untimed)
    kept = search.run(ctx, test["qres"], test["qoff"], sch, k, keep_handle=True)
    pp, fp = search.hit_pairs(plain), search.hit_pairs(filt)
    lost = len(pp - fp)
    m = pf["model"]
    X = search.features(kept, pf["mean"], pf["std"])
    search.close(kept)
    m.upload(X)
    f_ms, _ = m.time(len(X), warmup=2, iters=5)
    _, s_ms = m.time(65536, warmup=2, iters=5)
    m.close()
    stage = lambda r: sum(r.timing[x] for x in ("plan", "prefilter", "score", "finish"))
    sp, sf = stage(plain), stage(filt)
    flop = 164352 * len(X)
    out = {"workload": "at-scale blastp workload (search leg), default preset",
           "train_regions": pf["train_regions"], "train_positives": pf["positives"], "train_steps": pf["steps"],
           "epochs": args.pf_epochs, "minibatch": 65536, "pos_weight": round(pf["pos_weight"], 2),
           "train_s": round(train_s, 1), "threshold": round(pf["threshold"], 4),
           "valid_regions": pf["valid_regions"], "loss_target_valid": args.pf_loss,
           "test_regions": int(filt.counters["regions"]), "regions_pruned": int(filt.counters["regions_pruned"]),
           "kept_fraction": round(1 - filt.counters["regions_pruned"] / max(1, filt.counters["regions"]), 4),
           "plain_hits": len(pp), "hits_lost": lost, "hit_loss": round(lost / max(1, len(pp)), 5),
           "queries_per_s_plain": round(nq / tp, 1), "queries_per_s_filtered": round(nq / tf, 1),
           "end_to_end_speedup": round(tp / tf, 3),
           "candidate_scoring_ms_plain": round(sp, 1), "candidate_scoring_ms_filtered": round(sf, 1),
           "candidate_scoring_speedup": round(sp / max(sf, 1e-9), 3),
           "banded_cells_plain": int(plain.counters["banded_cells"]),
           "banded_cells_filtered": int(filt.counters["banded_cells"]),
           "prefilter_stage_ms": round(filt.timing["prefilter"], 2),
           "stages_ms_plain": {a: round(b, 1) for a, b in plain.timing.items()},
           "stages_ms_filtered": {a: round(b, 1) for a, b in filt.timing.items()},
           "wall_ms_plain": round(tp * 1e3, 1), "wall_ms_filtered": round(tf * 1e3, 1),
           "mlp_forward_ms": round(f_ms, 3), "mlp_forward_tflops": round(flop / (f_ms * 1e-3) / 1e12, 1),
           "mlp_forward_mfma_frac": round(flop / (f_ms * 1e-3) / 2.5e15, 4),
           "train_step_ms_65536": round(s_ms, 3),
           "prunable_bound": "gradient-boosted trees on the same features, fitted offline to a sample of one "
                             "at-scale search (23 k positives, 300 k negatives), prune at most ~33 % of regions at "
                             "1 % reported-hit loss: the weakest reported hits' regions carry noise-level X-drop "
                             "evidence"}
    log(f"# prefilter: {pf['train_regions']} training regions, kept {out['kept_fraction']}, hit loss "
        f"{out['hit_loss']} ({lost}/{len(pp)}), queries/s {out['queries_per_s_plain']} -> "
        f"{out['queries_per_s_filtered']}, candidate scoring {sp:.0f} -> {sf:.0f} ms, MLP forward "
        f"{out['mlp_forward_tflops']} TFLOP/s")
    return out


def cpu_info():
    """CPU model and physical core count of the host (lscpu), as the baseline states them."""
    model, sockets, cores, threads = None, 1, None, os.cpu_count()
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {k.strip(): v.strip() for k, v in (l.split(":", 1) for l in out.splitlines() if ":" in l)}
        model = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "1"))
        cores = int(kv.get("Core(s) per socket", "0")) * sockets or None
    except Exception:
        pass
    return {"model": model, "physical_cores": cores, "logical_cpus": threads}


def search_leg(ctx, args, sch):
    """The blastp workloads end to end on the GPU (saber_hip_search_run).

    Small (100 queries x 5000 subjects): hit lists compared query by query
    with the oracle's plain CPU search, which is also timed (one host core).
    At scale (11,206 queries in ~2,000 planted families against themselves
    plus 320,000 reversed decoys, ~148 M residues): queries/s end to end and
    per stage, with size-independent checks -- every query's self hit is
    reported with at least the score of the best ungapped segment of its main
    diagonal, a sample of hits rescores to its score from its CIGAR, and a
    repeated call gives the same hits.
    """
    from oracle import search as OS
    from saber_amd import BLOSUM62, search, stats, workload
    k = stats.gapped_protein(62, 11, 1)
    out = {}
    # small workload
    d = workload.search_set(seed=0)
    ns = len(d["soff"]) - 1
    ctx.upload_db(d["sres"], d["soff"])
    search.run(ctx, d["qres"], d["qoff"], sch, k)          # builds the database's seed-scan data
    t0 = time.perf_counter()
    g = search.run(ctx, d["qres"], d["qoff"], sch, k)
    gw = time.perf_counter() - t0
    st = stats.SearchStatistics(k, int(d["soff"][-1]), ns, stats.PER_SUBJECT)
    sd = search.ScanSeeds.default()
    t0 = time.perf_counter()
    want, wctr = OS.search(d["qres"], d["qoff"], d["sres"], d["soff"], BLOSUM62, 11, 1, st, sd.alpha_map,
                           sd.alpha_size, sd.shapes)
    cw = time.perf_counter() - t0
    bad = sum(int([h.key() for h in a] != [OS.hit_key(h) for h in b]) for a, b in zip(g.hits, want))
    nq = len(d["qoff"]) - 1
    out["small"] = {"workload": "100 queries x 5000 subjects, lognormal median 170 sigma 0.6 [30,1000], 30% planted "
                                "homologues at 25-60% identity, default preset (Murphy-10, 16 weight-7 seeds, "
                                "window filter 27), composition bias on, masking off, E <= 10",
                    "queries_per_s": round(nq / gw, 1), "wall_ms": round(gw * 1e3, 2),
                    "hits": int(sum(len(h) for h in g.hits)), "parity_queries": nq, "parity_mismatches": bad,
                    "counters_equal": all(g.counters[c] == wctr[c] for c in wctr if c in g.counters),
                    "cpu_plain_queries_per_s": round(nq / cw, 2), "cpu_plain": "oracle/search.py (scan schedule "
                    "restated in C, scoring and traceback by the oracle), 1 host core"}
    print(f"# search small: GPU {gw * 1e3:.1f} ms ({nq / gw:.0f} queries/s), oracle {cw:.1f} s, "
          f"mismatching queries {bad}/{nq}", file=sys.stderr, flush=True)
    if args.search_queries <= 0:
        return out
    # at scale
    t0 = time.perf_counter()
    d = workload.family_set(seed=0, nq=args.search_queries, ndecoy=args.search_decoys)
    gen = time.perf_counter() - t0
    nq, ns = len(d["qoff"]) - 1, len(d["soff"]) - 1
    ctx.upload_db(d["sres"], d["soff"])
    t0 = time.perf_counter()
    g0 = search.run(ctx, d["qres"], d["qoff"], sch, k)
    first = time.perf_counter() - t0
    gc.collect()
    gc.disable()
    t0 = time.perf_counter()
    g = search.run(ctx, d["qres"], d["qoff"], sch, k)
    wall = time.perf_counter() - t0
    gc.enable()
    fam = d["family"]
    # every query finds itself, with at least the best ungapped segment of
    # its main diagonal (a lower bound of the local alignment score)
    self_ok = 0
    for qi in range(nq):
        q = d["qres"][int(d["qoff"][qi]):int(d["qoff"][qi + 1])]
        b = search.composition_bias(q, BLOSUM62)
        best = run_ = 0
        for c, bi in zip(q.tolist(), b.tolist()):
            run_ = max(0, run_ + BLOSUM62[c][c] + int(bi))
            best = max(best, run_)
        h = [x for x in g.hits[qi] if x.subject == qi]
        self_ok += int(bool(h) and h[0].score >= best)
    rng = np.random.default_rng(3)
    flat = [(qi, x) for qi, hl in enumerate(g.hits) for x in hl]
    rs_bad = 0
    for i in rng.choice(len(flat), size=min(300, len(flat)), replace=False).tolist():
        qi, x = flat[i]
        q = d["qres"][int(d["qoff"][qi]):int(d["qoff"][qi + 1])]
        s = d["sres"][int(d["soff"][x.subject]):int(d["soff"][x.subject + 1])]
        b = search.composition_bias(q, BLOSUM62)
        sc, a, j = 0, x.q_start - 1, x.s_start - 1
        for op, l in x.ops:
            if op == ord("M"):
                for _ in range(l):
                    sc += BLOSUM62[int(q[a])][int(s[j])] + int(b[a])
                    a += 1
                    j += 1
            else:
                sc -= 11 + (l - 1)
                a, j = (a + l, j) if op == ord("I") else (a, j + l)
        rs_bad += int(sc != x.score or a != x.q_end or j != x.s_end)
    same = sum(1 for qi, hl in enumerate(g.hits) for x in hl
               if x.subject < nq and x.subject != qi and fam[x.subject] == fam[qi])
    t = g.timing
    scoring = sum(t[s_] for s_ in ("sort", "xdrop", "regions", "plan", "prefilter", "score", "finish"))
    out["scale"] = {"workload": f"{nq} queries (lognormal median 170 sigma 0.6 [30,1000], {int(fam.max()) + 1} planted "
                                f"families of 2-30 members at 20-50% identity) vs themselves + {args.search_decoys} "
                                f"reversed decoys (lognormal median 380 sigma 0.6 [30,3000]), {ns} subjects, "
                                f"{int(d['sres'].size)} residues; default preset, masking off",
                    "queries_per_s": round(nq / wall, 1), "wall_s": round(wall, 3),
                    "first_call_wall_s": round(first, 3), "candidate_scoring_ms": round(scoring, 1),
                    "candidate_scoring_queries_per_s": round(nq / (scoring * 1e-3), 1),
                    "stages_ms": {a: round(v, 1) for a, v in t.items()}, "counters": g.counters,
                    "hits": len(flat), "same_family_hits": same, "self_hits": self_ok,
                    "rescore_sample": min(300, len(flat)), "rescore_mismatches": rs_bad,
                    "repeat_identical": [[x.key() for x in a] for a in g0.hits] == [[x.key() for x in a] for a in g.hits],
                    "generate_s": round(gen, 1),
                    "reference_published": "132 queries/s: 11,206 SCOP40 queries vs 146.8 M residues, 2 threads "
                                           "Xeon @2.80 GHz (README.md:64, bench/results/scaling/times.jsonl)"}
    print(f"# search at scale: {nq} x {ns}: {wall:.2f} s ({nq / wall:.0f} queries/s), first call {first:.2f} s; "
          f"self hits {self_ok}/{nq}, rescore mismatches {rs_bad}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--regions", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--per-query", type=int, default=1000, help="regions sharing one query")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target duration of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the planner's modelled shape choice instead of measuring it first")
    ap.add_argument("--xdrop-hits", type=int, default=4_000_000, help="seed hits for the X-drop leg")
    ap.add_argument("--traceback", type=int, default=20_000,
                    help="best-scoring regions aligned with traceback (0: skip the leg)")
    ap.add_argument("--check", type=int, default=3000, help="jobs checked against the oracle")
    ap.add_argument("--full-queries", type=int, default=1000,
                    help="queries of the exhaustive-mode leg (x 100,000 subjects); 0 skips it")
    ap.add_argument("--full-subjects", type=int, default=100_000)
    ap.add_argument("--pf-loss", type=float, default=0.008,
                    help="prefilter threshold: largest share of reported-hit regions pruned on validation")
    ap.add_argument("--no-prefilter", action="store_true", help="skip the learned-prefilter leg")
    ap.add_argument("--pf-epochs", type=int, default=8)
    ap.add_argument("--search-queries", type=int, default=11206,
                    help="queries of the at-scale blastp leg (0: small workload only)")
    ap.add_argument("--search-decoys", type=int, default=320000)
    ap.add_argument("--no-search", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only time a plain device copy of this many bytes and exit")
    args = ap.parse_args()

    if args.copy_probe:
        import saber_amd as S
        ctx = S.Context(0)
        ms = ctx.copy_probe(args.copy_probe, iters=5)
        print(json.dumps({"copy_bytes": args.copy_probe, "ms": ms,
                          "GBps": round(2 * args.copy_probe / (ms * 1e-3) / 1e9, 1)}))
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this "
              "single-GPU build (see DESIGN.md for how the path shards)", file=sys.stderr)
        return 2

    import saber_amd as S
    from saber_amd import BLOSUM62, workload

    t0 = time.time()
    d = workload.band_regions(seed=args.seed, n=args.regions, per_query=args.per_query)
    gen_s = time.time() - t0
    sch = S.Scheme(BLOSUM62, 11, 1)
    ctx = S.Context(0)
    tune = None
    if not args.no_autotune:
        # measured shape per class of band widths, on a sample of this batch
        t0 = time.time()
        tune = ctx.autotune(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, target_cells=1 << 62)
        print(f"# autotune {time.time() - t0:.1f}s: " + ", ".join(
            f"w {e['w_lo']}-{e['w_hi']} -> {e['shape']}{' (changed)' if e['changed'] else ''}" for e in tune)
            + (f"; sample {tune[-1]['gcups']} GCUPS" if tune else ""), file=sys.stderr, flush=True)
    t0 = time.time()
    # score_banded_many returns scores only: the step is the scores-only build;
    # the end-cell build (banded_score's outputs) is timed beside it.
    batch = ctx.band_batch(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, ends=False)
    plan_s = time.time() - t0
    print(f"# workload {args.regions} regions generated in {gen_s:.1f}s, planned+uploaded in "
          f"{plan_s:.1f}s", file=sys.stderr, flush=True)

    ms, band_ms, per, cells = batch.time(warmup=args.warmup, iters=args.steps)
    score, _, _, cell_arr = batch.read()
    eb = ctx.band_batch(d["qres"], d["qoff"], d["sres"], d["soff"], d["jobs"], sch, ends=True)
    ems, eband_ms, _, _ = eb.time(warmup=1, iters=max(3, args.steps // 2))
    escore, q_end, s_end, _ = eb.read()
    eb.close()

    variants = S.band_variants()
    # The packed kernel scores every job of the step in one launch (all
    # shapes); jobs it flags and jobs too wide for it go to the int32 kernel.
    band_cells = sum(per.values())
    achieved = OPS_PER_CELL * band_cells / (band_ms * 1e-3) / 1e12
    gcups = cells / (ms * 1e-3) / 1e9
    for v in sorted(per, key=lambda v: -per[v]):
        vL, vC, vK, vB = variants[v]
        print(f"#   shape ({vL},{vC},{vK}) bias {vB}: {per[v]} cells", file=sys.stderr)
    print(f"# step {ms:.3f} ms, {cells} cells, {gcups:.1f} GCUPS; band_multi_kernel launch "
          f"{band_ms:.3f} ms for {band_cells} cells (with end cells: step {ems:.3f} ms, "
          f"{cells / (ems * 1e-3) / 1e9:.1f} GCUPS)", file=sys.stderr, flush=True)

    # Parity spot check against the oracle on a deterministic sample.
    from oracle import oracle as O
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(args.regions, size=min(args.check, args.regions), replace=False))
    sub = d["jobs"][idx]
    want = O.banded_many(d["qres"], d["qoff"], d["sres"], d["soff"], sub, BLOSUM62, 11, 1)
    got = (score[idx], q_end[idx], s_end[idx], cell_arr[idx])
    mismatches = int(sum(int((np.asarray(g).astype(np.int64) != np.asarray(w).astype(np.int64)).sum())
                         for g, w in zip(got, want)))
    # the two builds agree on every job's score
    mismatches += int((escore != score).sum())
    if mismatches:
        print(f"# PARITY FAILURE: {mismatches} mismatching fields in {len(idx)} sampled jobs",
              file=sys.stderr)

    # Secondary: banded traceback (banded_align_until, the consumer of the band
    # scores) of the best-scoring regions, as the engine runs it for reported
    # hits: stop at the region's exact score (engine.rs:1360). Kernel time by
    # HIP events; the oracle's scalar restatement timed on one core beside it.
    tb = None
    if args.traceback > 0:
        ntb = min(args.traceback, args.regions)
        top = np.sort(np.argsort(-score.astype(np.int64), kind="stable")[:ntb])
        tjobs = d["jobs"][top]
        tstop = score[top].astype(np.int32)
        res, tms = ctx.band_align(d["qres"], d["qoff"], d["sres"], d["soff"], tjobs, sch, stop_at=tstop,
                                  time_iters=3)
        tcells = sum(r[6] for r in res if isinstance(r, tuple))
        ti = rng.choice(ntb, size=min(300, ntb), replace=False)
        tbad = 0
        t0 = time.time()
        for k in ti.tolist():
            J = tjobs[k]
            q = d["qres"][d["qoff"][J["query"]]:d["qoff"][J["query"] + 1]]
            s0 = d["soff"][J["subject"]]
            w = O.band_align(q, d["sres"][s0 + J["win_lo"]:s0 + J["win_hi"]], BLOSUM62, None, 11, 1,
                             int(J["band_lo"]), int(J["band_hi"]), int(tstop[k]), 32 << 20)
            tbad += int(w != res[k])
        cpu_per = (time.time() - t0) / len(ti)
        tb = {"alignments": int(ntb), "kernel_ms": round(tms, 3),
              "alignments_per_s": round(ntb / (tms * 1e-3), 1), "cells": int(tcells),
              "gcups": round(tcells / (tms * 1e-3) / 1e9, 2),
              "cpu_oracle_alignments_per_s_1core": round(1.0 / cpu_per, 1),
              "parity_sample": int(len(ti)), "parity_mismatches": tbad}
        print(f"# traceback {ntb} alignments: kernel {tms:.3f} ms ({tb['alignments_per_s']:.0f}/s, "
              f"{tb['gcups']} GCUPS), oracle 1 core {tb['cpu_oracle_alignments_per_s_1core']}/s, "
              f"mismatches {tbad}/{len(ti)}", file=sys.stderr, flush=True)
        # The consumer: Engine::traceback's widening loop over the same hits
        # (saber_hip_traceback), region diagonals at the band centre, the
        # scoring band as the limit, want = the exact band score, the
        # engine's default band_pad 16 (engine.rs:179). Every hit must come
        # back exact (the limit band reproduces its score). As the engine
        # calls it: the database uploaded once beforehand (timed apart) and
        # the call given no subject set. Wall time of the library call (host
        # rounds and copies included; not the Python conversion of its
        # results).
        hj = np.zeros(ntb, S.TRACEBACK_JOB)
        for f in ("query", "subject", "win_lo", "win_hi"):
            hj[f] = tjobs[f]
        centre = (tjobs["band_lo"].astype(np.int64) + tjobs["band_hi"].astype(np.int64)) // 2
        hj["d_min"] = centre + tjobs["win_lo"].astype(np.int64)
        hj["d_max"] = hj["d_min"]
        hj["best"] = 1
        hj["want"] = tstop
        hj["has_limit"] = 1
        hj["limit_lo"] = tjobs["band_lo"]
        hj["limit_hi"] = tjobs["band_hi"]
        t0 = time.perf_counter()
        ctx.upload_db(d["sres"], d["soff"])
        up_ms = (time.perf_counter() - t0) * 1e3
        ctx.traceback(d["qres"], d["qoff"], None, None, hj[:64], sch, band_pad=16)   # warm-up
        hres, hcnt, hms = ctx.traceback(d["qres"], d["qoff"], None, None, hj, sch, band_pad=16, timed=True)
        hbad = sum(1 for r, w in zip(hres, tstop.tolist()) if r is None or not r[1] or r[0][0] != w)
        tb["consumer"] = {"hits": int(ntb), "wall_ms": round(hms, 2), "hits_per_s": round(ntb / (hms * 1e-3), 1),
                          "wall_over_kernel": round(hms / tms, 2), "db_upload_ms": round(up_ms, 1),
                          "counters": hcnt, "not_exact": hbad}
        print(f"# traceback consumer {ntb} hits: {hms:.1f} ms wall ({tb['consumer']['hits_per_s']:.0f}/s, "
              f"{hms / tms:.1f}x the kernel; DB upload {up_ms:.0f} ms apart), {hcnt['tracebacks']} attempts, "
              f"{hcnt['band_widenings']} widenings, not exact {hbad}", file=sys.stderr, flush=True)

    # Secondary: X-drop ungapped extension over seed hits of the same batch
    # (work unit: extension steps, xdrop_steps of scan.rs:902).
    xd = None
    if args.xdrop_hits > 0:
        hits = workload.xdrop_hits(d, seed=args.seed, per_region=max(1, args.xdrop_hits // args.regions))
        xb = ctx.xdrop_batch(d["qres"], d["qoff"], d["sres"], d["soff"], hits, sch, 20)
        xms, xsteps = xb.time(warmup=2, iters=max(3, args.steps // 2))
        xs, xlo, xhi, xst = xb.read()
        xb.close()
        xi = np.sort(rng.choice(len(hits), size=min(2000, len(hits)), replace=False))
        xbad = 0
        for k in xi.tolist():
            qi, si, qp, sp = hits[k].tolist()
            w = O.xdrop(d["qres"][d["qoff"][qi]:d["qoff"][qi + 1]], d["sres"][d["soff"][si]:d["soff"][si + 1]],
                        None, BLOSUM62, qp, sp, 20)
            xbad += int((xs[k], xlo[k], xhi[k], xst[k]) != w)
        xd = {"hits": int(len(hits)), "ms": round(xms, 4), "steps": int(xsteps),
              "gsteps_per_s": round(xsteps / (xms * 1e-3) / 1e9, 2),
              "hits_per_s": round(len(hits) / (xms * 1e-3), 1), "parity_sample": int(len(xi)),
              "parity_mismatches": xbad}
        if not args.no_cpu:
            # CPU baseline: the AVX2 restatement of xdrop_avx2 on every host
            # thread over the batch's first hits (a bounded sample), checked
            # equal to the GPU's outputs on that sample.
            nth = os.cpu_count() or 1
            ncpu = min(len(hits), 2_000_000)
            t0 = time.time()
            cs, clo, chi, cst = O.xdrop_many_avx2(d["qres"], d["qoff"], d["sres"], d["soff"], hits[:ncpu], BLOSUM62,
                                                  20, nthreads=nth)
            cpu_s = time.time() - t0
            cbad = int(np.sum((cs != xs[:ncpu]) | (clo != xlo[:ncpu]) | (chi != xhi[:ncpu]) | (cst != xst[:ncpu])))
            xd["cpu_baseline"] = {"kind": "port", "impl": "AVX2 restatement of xdrop_avx2 (scan.rs:934-1031)",
                                  "hits": ncpu, "seconds": round(cpu_s, 3), "threads_used": nth,
                                  "gsteps_per_s": round(float(cst.sum()) / cpu_s / 1e9, 3),
                                  "hits_per_s": round(ncpu / cpu_s, 1), "mismatches_vs_gpu": cbad,
                                  "host": cpu_info(), "sample": f"first {ncpu} hits of the batch"}
        print(f"# xdrop {len(hits)} hits: {xms:.3f} ms, {xsteps} steps, "
              f"{xd['gsteps_per_s']} G steps/s, mismatches {xbad}/{len(xi)}"
              + (f"; CPU AVX2 {xd['cpu_baseline']['gsteps_per_s']} G steps/s on {xd['cpu_baseline']['threads_used']}"
                 f" threads ({xd['cpu_baseline']['mismatches_vs_gpu']} differ)" if "cpu_baseline" in xd else ""),
              file=sys.stderr, flush=True)

    # Secondary: exhaustive full-matrix SW, queries x subjects drawn like the
    # exhaustive workload (lognormal median 250, sigma 0.6, clamped 30..3000).
    fl = None
    if args.full_queries > 0:
        # sigma 0.4 gives the stated ~27 M subject residues (median 250, [30, 3000])
        fq, fqo = workload.sequences(args.seed + 100, args.full_queries, 250, 0.4, 30, 3000)
        fs, fso = workload.sequences(args.seed + 200, args.full_subjects, 250, 0.4, 30, 3000)
        # score_many returns scores only: the headline figure is the scores-only
        # kernel; the end-cell kernel (sw_reference's first optimal cell) is
        # timed beside it. The shape per subject-length group is measured
        # first (both flavours) unless --no-autotune.
        ftune = None
        if not args.no_autotune:
            t0 = time.time()
            ftune = {"scores": ctx.full_autotune(fq, fqo, fs, fso, sch, ends=False),
                     "ends": ctx.full_autotune(fq, fqo, fs, fso, sch, ends=True)}
            print(f"# full-SW autotune {time.time() - t0:.1f}s: " + ", ".join(
                f"{k} {e['len_lo']}-{e['len_hi']}->{e['shape']}" for k, v in ftune.items() for e in v
                if e["changed"]), file=sys.stderr, flush=True)
        fb = ctx.full_batch(fq, fqo, fs, fso, sch, ends=False)
        fms, fcells, fresc = fb.time(warmup=1, iters=2)
        fsc = fb.read()[0]
        fb.close()
        fb = ctx.full_batch(fq, fqo, fs, fso, sch, ends=True)
        fems, _, _ = fb.time(warmup=1, iters=2)
        esc, fqe, fse = fb.read()
        fb.close()
        fbad = 0
        for k in range(40):
            a, b = int(rng.integers(args.full_queries)), int(rng.integers(args.full_subjects))
            w = O.sw_reference(fq[fqo[a]:fqo[a + 1]], fs[fso[b]:fso[b + 1]], BLOSUM62, None, 11, 1)
            fbad += int(int(fsc[a, b]) != w[0]) + int((int(esc[a, b]), int(fqe[a, b]), int(fse[a, b])) != w)
        # CPU: the AVX2 restatement of score_many on all host threads over the
        # first k queries x every subject (k sized to ~--cpu-seconds),
        # extrapolated to the whole query set by cells
        fcpu = None
        if not args.no_cpu:
            nth = os.cpu_count() or 1
            qlens = (fqo[1:] - fqo[:-1]).astype(np.int64)
            sres_cells = int(fso[-1])
            t0 = time.perf_counter()
            probe = O.score_many_avx2(fq, fqo[:2], fs, fso, BLOSUM62, 11, 1, nthreads=1)
            rate1 = qlens[0] * sres_cells / max(time.perf_counter() - t0, 1e-6)
            kq = 1
            while kq < args.full_queries and qlens[:kq + 1].sum() * sres_cells / (rate1 * nth) < args.cpu_seconds:
                kq += 1
            kq = max(kq, min(nth, args.full_queries))
            t0 = time.perf_counter()
            cs = O.score_many_avx2(fq, fqo[:kq + 1], fs, fso, BLOSUM62, 11, 1, nthreads=nth)
            dt = time.perf_counter() - t0
            ccells = int(qlens[:kq].sum()) * sres_cells
            same = bool(np.array_equal(cs, np.asarray(fsc[:kq])))
            fcpu = {"gcups": round(ccells / dt / 1e9, 2), "threads": nth, "host": cpu_info(),
                    "sample": f"first {kq} queries x all {args.full_subjects} subjects, {ccells} cells in {dt:.1f} s",
                    "scores_equal_gpu": same,
                    "full_run_extrapolated_s": round(fcells / (ccells / dt), 1),
                    "kind": "port (AVX2 restatement of score_many, sw_simd.rs:95-153, 157-427)"}
            _ = probe
        fl = {"queries": args.full_queries, "subjects": args.full_subjects, "ms": round(fms, 3),
              "cells": int(fcells), "gcups": round(fcells / (fms * 1e-3) / 1e9, 1),
              "ms_with_ends": round(fems, 3), "gcups_with_ends": round(fcells / (fems * 1e-3) / 1e9, 1),
              "headline": "gcups_with_ends (the exhaustive mode's check is best score and end coordinates)",
              "cpu_baseline": fcpu,
              "int32_rescored_pairs": int(fresc), "parity_sample": 40, "parity_mismatches": fbad,
              "autotune": None if ftune is None else {
                  k: [{"len": [e["len_lo"], e["len_hi"]], "shape": list(e["shape"])} for e in v if e["changed"]]
                  for k, v in ftune.items()}}
        print(f"# full SW {args.full_queries} x {args.full_subjects}: {fms:.2f} ms, "
              f"{fl['gcups']} GCUPS (with end cells {fems:.2f} ms, {fl['gcups_with_ends']} GCUPS), "
              f"mismatches {fbad}/40", file=sys.stderr, flush=True)

    pf = None if (args.no_prefilter or args.search_queries <= 0) else prefilter_leg(ctx, args, sch)
    srch = None if args.no_search else search_leg(ctx, args, sch)

    cpu = None
    if not args.no_cpu:
        nthreads = os.cpu_count() or 1
        # AVX2 restatement of score_banded_many on all host threads, over a
        # prefix of the batch (whole queries) sized to about --cpu-seconds.
        probe = d["jobs"][: 20 * args.per_query]
        t0 = time.perf_counter()
        O.banded_many_avx2(d["qres"], d["qoff"], d["sres"], d["soff"], probe, BLOSUM62, 11, 1,
                           nthreads=nthreads)
        dt = time.perf_counter() - t0
        pc = int(cell_arr[: len(probe)].sum())
        rate = pc / max(dt, 1e-6)
        mean_cells = cells / args.regions
        nsample = int(min(args.regions, max(len(probe), rate * args.cpu_seconds / mean_cells)))
        nsample = max(args.per_query, nsample // args.per_query * args.per_query)
        sample = d["jobs"][:nsample]
        # Repeat the sample until about --cpu-seconds of work has been timed.
        reps = 0
        t0 = time.perf_counter()
        while True:
            cpu_scores = O.banded_many_avx2(d["qres"], d["qoff"], d["sres"], d["soff"], sample,
                                            BLOSUM62, 11, 1, nthreads=nthreads)
            reps += 1
            if time.perf_counter() - t0 >= args.cpu_seconds or reps >= 100:
                break
        dt = time.perf_counter() - t0
        sc = int(cell_arr[:nsample].sum()) * reps
        cpu_ok = bool(np.array_equal(cpu_scores, score[:nsample]))
        # bands wider than 126 (the u8 kernel's limit, band_simd8.rs:38) take
        # the int16 path in the reference's CPU code
        sj = d["jobs"][:nsample]
        m_ = (d["qoff"][sj["query"].astype(np.int64) + 1] - d["qoff"][sj["query"].astype(np.int64)]).astype(np.int64)
        n_ = (sj["win_hi"].astype(np.int64) - sj["win_lo"].astype(np.int64))
        w_ = np.minimum(sj["band_hi"].astype(np.int64), n_ - 1) - np.maximum(sj["band_lo"].astype(np.int64), -(m_ - 1)) + 1
        i16_share = float(cell_arr[:nsample][w_ > 126].sum() / max(1, cell_arr[:nsample].sum()))
        cpu = {"value": round(sc / dt / 1e9, 3), "unit": "GCUPS", "cores": nthreads, "kind": "port",
               "host": cpu_info(), "threads_used": nthreads, "int16_path_cell_share": round(i16_share, 4),
               "sample": f"first {nsample} regions ({nsample // args.per_query} whole queries) of the "
                         f"same batch x {reps} passes, {sc} cells in {dt:.1f} s; AVX2 restatement of score_banded_many "
                         f"(u8 32-lane batches + int16 banded kernel, band_simd8.rs:78-388, "
                         f"banded_simd.rs:393-599), one thread per query group; scores equal GPU: {cpu_ok}"}

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
                        "diagonal), 1000 regions per query, lognormal lengths median 150 sigma 0.6 "
                        "[20,2000], half-widths {8,16,32,64}, 40% planted homologues, BLOSUM62 11/1, no bias",
            "regions": args.regions,
            "cells_per_step": int(cells),
            "parity_sample": int(len(idx)),
            "parity_mismatches": mismatches,
        },
        "roofline": {
            "bound": "valu",
            "kernel": ROOFLINE_KERNEL,
            "achieved": round(achieved, 3),
            "peak": round(VALU_I16X2_TOPS, 2),
            "unit": "Tops/s (int16 ops, 10 per cell)",
            "frac": round(achieved / VALU_I16X2_TOPS, 4),
            "traffic": measured_traffic(args),
        },
        "cpu_baseline": cpu,
        "extra": {"autotune": tune,
                  "band_with_end_cells": {"ms_per_step": round(ems, 4), "launch_ms": round(eband_ms, 4),
                                          "gcups": round(cells / (ems * 1e-3) / 1e9, 2)},
                  "xdrop": xd, "traceback": tb, "full_sw": fl, "prefilter": pf, "search": srch},
    }
    print(json.dumps(out), flush=True)
    batch.close()
    ctx.close()
    ok = mismatches == 0 and (xd is None or xd["parity_mismatches"] == 0) and (fl is None or fl["parity_mismatches"] == 0)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
