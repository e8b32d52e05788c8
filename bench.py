"""Benchmark of the ColBERT late-interaction search path on one MI355X.

Workload (default, --config c2): 1,048,576 docs with lognormal token counts
(median 64, sigma 0.29, clipped to [8, 300]; ~70M tokens), 65,536 mixture
components, 65,536 k-means centroids (4 Lloyd iterations on a 2,097,152-token
sample), 2-bit residuals; 4,096 queries x 32 tokens; nprobe 8, probe_topk
(rerank) 1,000, k 100.  The corpus and index are built on the device
(synthetic data, seeded); a step is one search of the 4,096-query batch with
inputs resident in HBM, in BF16 MaxSim with an FP32 rescore of the best 200
(--precision, --rescore).

Prints one JSON line: queries/sec (value), ms per step, p50/p99 latency of a
device pass (--batch queries, default the whole 4,096-query batch), recall@100
against FP64 exhaustive MaxSim over the original embeddings (and whether the
metric's 0.99 condition holds), the plain FP32 pipelines on the same batch,
unique candidates per query against the 8,192-doc cap, the rerank kernel's
roofline figures, and a CPU baseline (the C oracle of the same path on the
host cores, on a bounded query sample, with its answers compared to the
GPU's).  --config northstar prints one line per pipeline (FP32, FP32 MFMA,
BF16/FP8 with and without the rescore, persistent + graph, learned probing)
and a recall-QPS frontier.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "multi-vector-retrieval-gpu_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

BASELINE_METRIC = "queries/sec at recall@100\u22650.99 vs FP64 exact MaxSim; p50/p99 latency, 1 GPU"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "fp32_mfma": 157.3, "bf16": 2500.0, "fp8": 5000.0}

CONFIGS = {
    "c2": dict(n_doc=1 << 20, n_comp=65536, n_c=65536, lognormal=True, n_q=4096, P=8, R=1000, k=100,
               kmeans_sample=1 << 21, n_iter=4),
    "c1": dict(n_doc=8192, n_comp=1024, n_c=1024, lognormal=False, n_q=64, P=4, R=256, k=10,
               kmeans_sample=None, n_iter=10),
    # uneven batches on the C2 corpus: 16,384 queries, U{8..32} tokens, nprobe
    # uniform over {2, 4, 8, 16, 32}
    # learned probe, as specified: config-1 corpus (mixture recipe), 262,144
    # training queries labelled by recall@100 >= 0.99 against exhaustive MaxSim
    "c5": dict(n_doc=8192, n_comp=1024, n_c=1024, lognormal=False, n_q=4096, P=4, R=256, k=100,
               kmeans_sample=None, n_iter=10, learned=True, n_train=262144, target=0.99),
    # learned probe on a topic-structured corpus where query difficulty varies
    "c5t": dict(n_doc=8192, n_comp=1024, n_c=1024, lognormal=False, n_q=2048, P=4, R=2048, k=10,
                kmeans_sample=None, n_iter=10, learned=True, n_train=65536, target=0.9, topic=True),
    "c4": dict(n_doc=1 << 20, n_comp=65536, n_c=65536, lognormal=True, n_q=16384, P=32, R=1000, k=100,
               kmeans_sample=1 << 21, n_iter=4, ragged=True),
    # the PLAID path (ColBERTv2 IndexScorer.rank, CPU-branch semantics) on a
    # PLAID index of the c2 corpus, searcher defaults for k
    "plaid": dict(n_doc=1 << 20, n_comp=65536, n_c=65536, lognormal=True, n_q=4096, k=10,
                  kmeans_sample=1 << 21, n_iter=4, plaid=True),
    "plaid100": dict(n_doc=1 << 20, n_comp=65536, n_c=65536, lognormal=True, n_q=4096, k=100,
                     kmeans_sample=1 << 21, n_iter=4, plaid=True),
    # the north_star target run on the c2 corpus: one line per pipeline, then
    # a recall-QPS frontier (see bench_northstar)
    "northstar": dict(n_doc=1 << 20, n_comp=65536, n_c=65536, lognormal=True, n_q=4096, P=8, R=1000, k=100,
                      kmeans_sample=1 << 21, n_iter=4, n_train=16384),
    # the groundtruth module (BruteForceProgressive): all-pairs MaxSim of 1,024
    # queries over 65,536 docs of U{16..128} tokens, exact FMA chains and FP32 MFMA
    "exhaustive": dict(n_doc=65536, n_comp=4096, n_q=1024, exhaustive=True),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="bf16", choices=["fp32", "fp32_mfma", "bf16", "fp8"])
    ap.add_argument("--rescore", type=int, default=200)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--recall-queries", type=int, default=256)
    ap.add_argument("--cpu-queries", type=int, default=640)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the plain FP32 pipelines and the candidate counts (profiling runs)")
    ap.add_argument("--persistent", action="store_true",
                    help="MaxSim + top-k as one persistent work-queue kernel per pass")
    ap.add_argument("--graph", action="store_true", help="replay each device pass from a hipGraph")
    ap.add_argument("--maxsim-order", default="auto", choices=["auto", "query", "doc"],
                    help="BF16 MaxSim loop order (same scores in every order)")
    ap.add_argument("--batch", type=int, default=4096, help="queries per device pass")
    ap.add_argument("--residual-on-host", action="store_true",
                    help="codes and residuals in pinned host memory, staged per device pass")
    ap.add_argument("--epochs", type=int, default=10, help="training epochs of the learned-probe configs")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path is query "
              f"data-parallel and would shard by replicating the index per GPU.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return

    import torch
    import mvr
    from mvr import synth

    cfg = CONFIGS[args.config]
    if cfg.get("learned"):
        bench_learned(torch, args, cfg)
        return
    if cfg.get("plaid"):
        bench_plaid(torch, args, cfg)
        return
    if cfg.get("exhaustive"):
        bench_exhaustive(torch, args, cfg)
        return
    t_setup = time.time()
    if cfg["lognormal"]:
        doclens = synth.doclens_lognormal(args.seed, cfg["n_doc"], 64.0, 0.29, 8, 300)
    else:
        doclens = synth.doclens_uniform(args.seed, cfg["n_doc"], 16, 128)
    log(f"[bench] config {args.config}: {cfg['n_doc']} docs, {int(doclens.sum())} tokens")
    corpus = synth.build_corpus(torch, args.seed, cfg["n_doc"], cfg["n_comp"], cfg["n_c"], doclens,
                                n_iter=cfg["n_iter"], kmeans_sample=cfg["kmeans_sample"],
                                log=lambda m: log("[bench]", m))
    ix = corpus.index
    if args.residual_on_host:
        ix = mvr.Index.from_device(torch.from_numpy(doclens.astype(np.int32)).cuda(), corpus.C, corpus.codes,
                                   corpus.residual, corpus.W, corpus.ivf, residual_on_host=True)
    ix.set_maxsim_order(args.maxsim_order)
    n_q, P, R, k = cfg["n_q"], cfg["P"], cfg["R"], cfg["k"]
    Q = synth.make_queries(torch, args.seed, corpus.centers, n_q)
    q_nt = q_np = None
    if cfg.get("ragged"):
        q_nt, q_np = synth.query_shapes(args.seed, n_q)
        log(f"[bench] ragged queries: tokens {q_nt.min()}..{q_nt.max()}, nprobe {sorted(set(q_np.tolist()))}")
    torch.cuda.synchronize()
    log(f"[bench] setup {time.time() - t_setup:.1f}s; max IVF list {ix.max_list}")
    if args.config == "northstar":
        bench_northstar(torch, args, cfg, corpus, Q, doclens)
        return
    rescore = args.rescore if args.precision in ("bf16", "fp8") else 0
    batch = args.batch
    mode = dict(precision=args.precision, rescore=rescore, P=P, R=R, persistent=args.persistent,
                graph=args.graph)
    run = run_pipeline(torch, ix, Q, k, batch, mode, args.steps, args.warmup, q_nt, q_np, keep=True)
    stages = run["stages"]
    ms_per_step, qps, tm = run["ms_per_step"], run["qps"], run["tm"]
    stage_keys = tuple(stages)
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, {qps:.0f} QPS")
    if args.residual_on_host:
        log(f"[bench] host residual staging: {tm['staging_ms']:.2f} ms, {tm['staged_docs']} docs / "
            f"{tm['staged_tokens']} tokens per step ({36.0 * tm['staged_tokens'] / tm['staging_ms'] / 1e6:.1f} GB/s "
            f"over the staging time)")
    log(f"[bench] stages (ms/step): " + ", ".join(f"{f}={stages[f]:.2f}" for f in stage_keys))
    out_ids = run["out_ids"]
    stream = torch.cuda.current_stream().cuda_stream

    # per-pass latency: single device passes bracketed by events on the search stream
    o_i = torch.empty((n_q, k), dtype=torch.int32, device="cuda")
    o_s = torch.empty((n_q, k), dtype=torch.float32, device="cuda")
    ev = []
    for _ in range(max(2, min(args.steps, 8))):
        for b0 in range(0, n_q, batch):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            ix.search_device(Q[b0:b0 + batch], o_i[b0:b0 + batch], o_s[b0:b0 + batch], P, R, k,
                             args.precision, rescore, batch=batch, stream=stream,
                             q_ntok=None if q_nt is None else q_nt[b0:b0 + batch],
                             q_nprobe=None if q_np is None else q_np[b0:b0 + batch],
                             persistent=args.persistent, graph=args.graph)
            e1.record()
            ev.append((e0, e1))
    torch.cuda.synchronize()
    lat = np.array([a.elapsed_time(b) for a, b in ev])

    # the same search through host buffers (PCIe copies of queries and results
    # included), as the reference's executables use it
    Qh = Q.cpu().numpy()
    t_h = time.perf_counter()
    ix.search(Qh, P, R, k, args.precision, rescore, q_ntok=q_nt, q_nprobe=q_np, batch=batch,
              persistent=args.persistent, graph=args.graph)
    qps_host = n_q / (time.perf_counter() - t_h)

    # roofline of the MaxSim kernel: algorithmic flops / bytes per launch over
    # its average launch duration in the timed region
    # (query tokens x doc tokens) pairs: 2 x 128 flops each
    n_launch = run["n_launch"]
    tok_per_launch = tm["rerank_tokens"] / tm["n_pass"]
    avg_launch_ms = run["acc"]["rerank_ms"] / n_launch
    flops = 2.0 * 128 * tm["rerank_pairs"] / tm["n_pass"]
    bytes_alg = 36.0 * tok_per_launch
    peak = PEAK_TFLOPS[args.precision]
    # a replayed graph has no stage boundaries: no per-kernel figure then
    achieved_tf = flops / (avg_launch_ms * 1e-3) / 1e12 if avg_launch_ms > 0 else None
    kname = {"fp32": "k_rerank_f32", "fp32_mfma": "k_rerank_f32mfma"}.get(args.precision,
                                                                       f"k_rerank_{args.precision}")
    # the library's rule for the BF16/FP8 loop order (mvr_index_set_maxsim_order:
    # doc-major from a quarter as many pairs per pass as docs); persistent:
    # the autotuner's choice of the first phase
    if args.precision in ("bf16", "fp8"):
        if args.persistent:
            t = ix.get_tuning_ex(args.precision)
            if args.maxsim_order == "doc" or (args.maxsim_order == "auto" and t and t["order"] == "doc"):
                kname = f"k_rerank_{args.precision}_dm"
            else:
                kname = "k_rerank_persist"
        elif args.maxsim_order == "doc" or (args.maxsim_order == "auto" and min(batch, n_q) * R * 4 >= cfg["n_doc"]):
            kname = f"k_rerank_{args.precision}_dm"
    traffic = measured_traffic(kname, args.config, args.precision)

    # the plain FP32 pipelines on the same batch (the reference of the
    # north_star's speed-up), and their answers for the recall and CPU checks
    others = {}
    if q_nt is None and not args.residual_on_host and not args.no_others:
        for prec in ("fp32", "fp32_mfma"):
            if prec == args.precision:
                continue
            others[prec] = run_pipeline(torch, ix, Q, k, batch, dict(precision=prec, rescore=0, P=P, R=R),
                                        max(2, min(args.steps, 5)), 1, keep=True)
            log(f"[bench] {prec} pipeline: {others[prec]['qps']:.0f} QPS")
    answers = {args.precision: out_ids, **{p: r["out_ids"] for p, r in others.items()}}

    # the exhaustive 2-bit ranking needs the device-resident index
    recall = recall_report(torch, args, cfg, corpus, corpus.index, Q, answers, k, doclens, q_nt)
    cands = (candidate_counts(ix, Q, P, R, k, q_nt, q_np, min(n_q, 1024))
             if not (args.residual_on_host or args.no_others) else None)
    probe8 = (probe_agreement(ix, Q, P, R, k, min(n_q, 256))
              if args.precision == "fp8" and q_nt is None and not args.residual_on_host else None)
    if probe8:
        log(f"[bench] fp8 probe vs exact FP32 probe: {probe8}")
    cpu = None
    if not args.no_cpu and args.cpu_queries > 0 and q_nt is None:
        ref = others.get("fp32") or (run if args.precision == "fp32" else None)
        mref = others.get("fp32_mfma") or (run if args.precision == "fp32_mfma" else None)
        cpu = cpu_baseline(torch, corpus, Q, P, R, k, args.cpu_queries, ref, mref)
        if cpu is not None and corpus.X is not None:
            cpu["plaid_cpu_path"] = plaid_cpu_baseline(torch, corpus, Q, k, args.cpu_queries)
    rec_main = recall["by_pipeline"][args.precision]["search_vs_fp64"] if recall else None
    line = {
        "metric": BASELINE_METRIC,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision.startswith("fp32") else args.precision,
        "data": "synthetic (seeded mixture corpus and queries, index built on the device)",
        "config": {"workload": f"{args.config}: {cfg['n_doc']} docs / {int(doclens.sum())} tokens, "
                               f"{cfg['n_c']} centroids, 2-bit residuals; {n_q} queries x "
                               + ("U{8..32} tokens, nprobe U{2,4,8,16,32}" if q_nt is not None
                                  else f"32 tokens; nprobe {P}") + f", probe_topk {R}, k {k}",
                   "persistent": args.persistent, "graph": args.graph,
                   "tuning": ({p: ix.get_tuning_ex(p) for p in (args.precision, "fp32_mfma")}
                              if args.persistent else None),
                   "maxsim_order": args.maxsim_order,
                   "batch_per_pass": batch, "precision": args.precision, "rescore": rescore,
                   "residual_on_host": args.residual_on_host},
        # the metric's recall clause, as measured: this value is the QPS at the
        # measured recall, and whether the 0.99 condition holds is stated here
        "recall_condition": {"target_recall_at_k_vs_fp64": 0.99, "measured": rec_main,
                             "met": (rec_main is not None and rec_main >= 0.99),
                             "ceiling_exhaustive_2bit_vs_fp64": recall.get("exhaustive_2bit_vs_fp64") if recall else None},
        "qps_host_buffers": round(qps_host, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "unit": f"per {batch}-query device pass"},
        "recall_at_k": rec_main,
        "recall": recall,
        "pipelines": {p: {"qps": round(r["qps"], 1), "ms_per_step": round(r["ms_per_step"], 3),
                          "speedup_of_headline": round(qps / r["qps"], 3),
                          "stages_ms_per_step": {f: round(v, 3) for f, v in r["stages"].items()},
                          # the same figures for this pipeline's MaxSim kernel
                          "roofline": pipeline_roofline(r, p, args.config)}
                      for p, r in others.items()},
        "stages_ms_per_step": {f: round(stages[f], 3) for f in stage_keys},
        "candidates": cands,
        "fp8_probe": probe8,
        "host_staging": ({"ms_per_step": round(tm["staging_ms"], 3), "docs_per_step": tm["staged_docs"],
                          "tokens_per_step": tm["staged_tokens"], "bytes_per_token": 36}
                         if args.residual_on_host else None),
        "roofline": roofline(kname, achieved_tf, peak, bytes_alg, avg_launch_ms, n_launch, tok_per_launch,
                             flops, traffic),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


STAGE_KEYS = ("centroid_ms", "filter_ms", "select_ms", "rerank_ms", "rescore_ms", "topk_ms", "total_ms")


def run_pipeline(torch, ix, Q, k, batch, mode, steps, warmup, q_nt=None, q_np=None, keep=False):
    """Times `steps` whole-batch searches of one pipeline (after `warmup`
    untimed ones), bracketed by device synchronisations.  Each call records
    HIP events around every stage of every device pass on the search stream
    (read back once per call), so per-stage and per-kernel durations come from
    the timed work itself.  Returns qps, ms/step, mean stage times, the last
    call's counters and (keep) the answers of the last step on the host."""
    import ctypes as C
    import mvr
    from mvr import _capi
    n_q, t_max = Q.shape[0], Q.shape[1]
    stream = torch.cuda.current_stream().cuda_stream
    out_ids = torch.empty((n_q, k), dtype=torch.int32, device="cuda")
    out_s = torch.empty((n_q, k), dtype=torch.float32, device="cuda")
    prm = mvr.Index._params(mode["P"], mode["R"], k, mode["precision"], mode.get("rescore", 0), False, batch,
                            mode.get("learned", False), mode.get("persistent", False), mode.get("graph", False))
    qt = None if q_nt is None else q_nt.ctypes.data_as(C.c_void_p)
    qp = None if q_np is None else q_np.ctypes.data_as(C.c_void_p)

    def step():
        tm = _capi.Timings()
        _capi.check(_capi.lib().mvr_search(ix._h, Q.data_ptr(), n_q, t_max, qt, qp, C.byref(prm),
                                           out_ids.data_ptr(), out_s.data_ptr(), 1, C.c_void_p(stream),
                                           None, C.byref(tm)))
        return {f: getattr(tm, f) for f, _ in _capi.Timings._fields_}

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    acc = {f: 0.0 for f in STAGE_KEYS}
    n_launch = 0
    tm = None
    t0 = time.perf_counter()
    for _ in range(steps):
        tm = step()
        for f in STAGE_KEYS:
            acc[f] += tm[f]
        n_launch += tm["n_pass"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    out = {"qps": n_q * steps / elapsed, "ms_per_step": elapsed * 1e3 / steps,
           "stages": {f: acc[f] / steps for f in STAGE_KEYS}, "acc": acc, "n_launch": n_launch, "tm": tm}
    if keep:
        out["out_ids"] = out_ids.cpu().numpy().view(np.uint32)
        out["out_s"] = out_s.cpu().numpy()
    return out


def candidate_counts(ix, Q, P, R, k, q_nt, q_np, n):
    """Unique filter candidates per query (the stage output of an FP32 search
    of the first n queries), and how often the 8,192-doc candidate cap of the
    1M-doc configuration binds (the search keeps every candidate; the cap is
    a reported bound)."""
    Qh = Q[:n].cpu().numpy()
    _, _, st = ix.search(Qh, P, R, k, "fp32", q_ntok=None if q_nt is None else q_nt[:n],
                         q_nprobe=None if q_np is None else q_np[:n], stages=True, batch=n)
    c = st["cand_n"].astype(np.int64)
    return {"queries": int(n), "mean": round(float(c.mean()), 1), "p50": int(np.percentile(c, 50)),
            "p99": int(np.percentile(c, 99)), "max": int(c.max()), "cap": 8192,
            "cap_binding_frac": round(float((c > 8192).mean()), 4)}


def plaid_cpu_baseline(torch, corpus, Q, k, n):
    """The reference's CPU search path (PLAID's CPU branch, the restatement in
    oracle/plaid_oracle.c) timed on the same corpus and queries: a PLAID index
    of the corpus (float16 centroids, binarized residuals, pid IVF), the
    searcher's defaults for k, OpenMP over the first n queries on every host
    core.  A baseline of the reference's CPU speed, not a check of this
    pipeline's answers (PLAID ranks with its own codec)."""
    from mvr import plaid, synth
    from oracle import plaid as OP
    pa = synth.plaid_from_corpus(torch, corpus, log=lambda m: log("[bench]", m))
    ncells, thr, ndocs = plaid.searcher_defaults(k)
    cores = len(os.sched_getaffinity(0))
    oix = OP.PlaidIndex(**pa.index_kwargs())
    Qh = Q[:n].cpu().numpy()
    oix.search(Qh[:16], k, ncells, thr, ndocs, n_threads=cores)
    t0 = time.time()
    oix.search(Qh, k, ncells, thr, ndocs, n_threads=cores)
    wall = time.time() - t0
    log(f"[bench] plaid cpu path: {n / wall:.1f} QPS on {cores} cores (k {k}, ncells {ncells}, ndocs {ndocs})")
    return {"value": round(n / wall, 2), "unit": "queries/s", "cores": cores, "kind": "port",
            "sample": f"first {n} queries, PLAID index of the same corpus, searcher defaults for k={k} "
                      f"(ncells {ncells}, threshold {thr}, ndocs {ndocs})",
            "build": "oracle/plaid_oracle.c: the restatement of PLAID's CPU search (filter_pids, "
                     "decompress_residuals, segmented_maxsim), gcc -O3 -mavx2 -mfma -fopenmp",
            "seconds": round(wall, 2)}


def probe_agreement(ix, Q, P, R, k, n):
    """FP8 probe against the exact FP32 probe on the first n queries: the
    fraction of query-token rows whose probed centroid set is the exact top-P
    (ids in order), and the fraction of probed ids that belong to it."""
    Qh = Q[:n].cpu().numpy()
    _, _, s8 = ix.search(Qh, P, R, k, "fp8", stages=True, batch=n)
    _, _, s32 = ix.search(Qh, P, R, k, "fp32", stages=True, batch=n)
    a, b = s8["top_ids"].reshape(-1, P), s32["top_ids"].reshape(-1, P)
    same = float((a == b).all(1).mean())
    inter = float(np.mean([len(set(x.tolist()) & set(y.tolist())) / P for x, y in zip(a, b)]))
    return {"queries": int(n), "rows_exact": round(same, 4), "ids_shared": round(inter, 4)}


def score_errors(torch, corpus, Q, ids, sc, doclens, nr):
    """Errors of returned scores on the first nr queries: against FP64 MaxSim
    over the decompressed (2-bit) embeddings of the same docs (the arithmetic
    error of the pipeline) and against FP64 MaxSim over the original
    embeddings (quantisation and arithmetic together)."""
    dev = torch.device("cuda")
    dl = torch.from_numpy(doclens.astype(np.int64)).to(dev)
    off = torch.cumsum(dl, 0) - dl
    C64 = corpus.C.double()
    W64 = corpus.W.reshape(128, 4).double()
    shifts = torch.tensor([6 - 2 * (j % 4) for j in range(128)], device=dev, dtype=torch.int32)
    byte_of = torch.tensor([j // 4 for j in range(128)], device=dev, dtype=torch.int64)
    e_dec, e_org = [], []
    for q in range(nr):
        docs = torch.from_numpy(np.asarray(ids[q], np.int64)).to(dev)
        L = dl[docs]
        seg = torch.repeat_interleave(torch.arange(docs.numel(), device=dev), L)
        start = torch.repeat_interleave(off[docs], L)
        pos = torch.arange(int(L.sum()), device=dev) - torch.repeat_interleave(torch.cumsum(L, 0) - L, L)
        tok = start + pos
        res = corpus.residual[tok].to(torch.int32)
        buck = (res[:, byte_of] >> shifts) & 3
        Xd = C64[corpus.codes[tok].long()] + W64[torch.arange(128, device=dev)[None, :], buck.long()]
        Xo = corpus.X[tok].double()
        qv = Q[q].double()
        outs = []
        for X in (Xd, Xo):
            S = qv @ X.T                                               # [T, n_tok]
            M = torch.full((qv.shape[0], docs.numel()), -float("inf"), dtype=torch.float64, device=dev)
            M.scatter_reduce_(1, seg[None, :].expand_as(S), S, "amax", include_self=True)
            outs.append(M.sum(0))
        got = torch.from_numpy(np.asarray(sc[q], np.float64)).to(dev)
        e_dec.append((got - outs[0]).abs())
        e_org.append((got - outs[1]).abs())
    e_dec, e_org = torch.cat(e_dec), torch.cat(e_org)
    return {"queries": nr, "max_abs_vs_fp64_decompressed": float(e_dec.max()),
            "mean_abs_vs_fp64_decompressed": float(e_dec.mean()),
            "max_abs_vs_fp64_original": float(e_org.max()), "mean_abs_vs_fp64_original": float(e_org.mean())}


def bench_northstar(torch, args, cfg, corpus, Q, doclens):
    """The north_star target on the 1M-doc corpus, one JSON line per mode:
    the plain FP32 pipelines, BF16 and FP8 MaxSim with an FP32 rescore of the
    best 200, the persistent work-queue kernel with hipGraph replay, and
    learned probing; then one line with a recall-QPS frontier over nprobe and
    probe_topk.  Each line: queries/s over the 4,096-query batch, recall@k
    against FP64 exhaustive MaxSim over the original embeddings and against
    exhaustive MaxSim over the 2-bit index (256 queries), the C3 score errors,
    and the speed-up over the plain FP32 pipeline."""
    from mvr import probe
    ix = corpus.index
    n_q, P, R, k = cfg["n_q"], cfg["P"], cfg["R"], cfg["k"]
    nr = min(args.recall_queries, n_q)
    steps, warm = max(2, min(args.steps, 5)), 1
    t0 = time.time()
    gnd, dec = ground_truth(torch, corpus, ix, Q, k, doclens, nr)
    ceiling = overlap(dec, gnd, k)
    log(f"[bench] ground truth for {nr} queries in {time.time() - t0:.1f}s; exhaustive 2-bit vs FP64 {ceiling:.4f}")
    workload = (f"northstar: {cfg['n_doc']} docs / {int(doclens.sum())} tokens, {cfg['n_c']} centroids, 2-bit "
                f"residuals; {n_q} queries x 32 tokens; k {k}")
    base_qps = None

    def emit(name, mode, run, extra=None):
        nonlocal base_qps
        ids, sc = run["out_ids"], run["out_s"]
        rec = {"search_vs_fp64": overlap(ids[:nr], gnd, k), "search_vs_exhaustive_2bit": overlap(ids[:nr], dec, k),
               "exhaustive_2bit_vs_fp64": ceiling, "queries": nr}
        if name == "fp32":
            base_qps = run["qps"]
        errs = score_errors(torch, corpus, Q, ids, sc, doclens, nr)
        line = {"metric": BASELINE_METRIC, "mode": name, "value": round(run["qps"], 1), "unit": "queries/s",
                "n_gpus": 1, "steps": steps, "warmup": warm, "ms_per_step": round(run["ms_per_step"], 3),
                "higher_is_better": True, "dtype": mode["precision"] if not mode["precision"].startswith("fp32") else "fp32",
                "data": "synthetic (seeded mixture corpus and queries, index built on the device)",
                "config": {"workload": workload, "nprobe": mode["P"], "probe_topk": mode["R"], **{
                    kk: mode.get(kk) for kk in ("precision", "rescore", "persistent", "graph", "learned")}},
                "speedup_vs_plain_fp32": round(run["qps"] / base_qps, 3) if base_qps else None,
                "recall_at_k": rec, "score_error": errs,
                "stages_ms_per_step": {f: round(v, 3) for f, v in run["stages"].items()}}
        if extra:
            line.update(extra)
        print(json.dumps(line), flush=True)
        log(f"[bench] {name}: {run['qps']:.0f} QPS, recall vs FP64 {rec['search_vs_fp64']:.4f}, vs 2-bit "
            f"{rec['search_vs_exhaustive_2bit']:.4f}")

    modes = [("fp32", dict(precision="fp32", rescore=0)),
             ("fp32_mfma", dict(precision="fp32_mfma", rescore=0)),
             ("bf16_rescore200", dict(precision="bf16", rescore=200)),
             ("bf16_rescore200_persistent_graph", dict(precision="bf16", rescore=200, persistent=True, graph=True)),
             ("fp8_rescore200", dict(precision="fp8", rescore=200)),
             ("bf16", dict(precision="bf16", rescore=0)),
             ("fp8", dict(precision="fp8", rescore=0))]
    for name, m in modes:
        mode = dict(P=P, R=R, **m)
        emit(name, mode, run_pipeline(torch, ix, Q, k, args.batch, mode, steps, warm, keep=True))

    # learned probing: the MLP is trained on the device on queries labelled by
    # the smallest nprobe that reproduces the deepest fixed probe's top-k
    # (nprobe 32, probe_topk 4,096) to recall >= 0.9, then searches with its
    # per-query nprobe and rerank depth in BF16 (FP32 rescore of 200)
    n_train = cfg.get("n_train", 16384)
    t1 = time.time()
    Qtr = synth_queries(torch, args.seed + 1, corpus, n_train)
    deep = run_pipeline(torch, ix, Qtr, k, 4096, dict(precision="fp32_mfma", rescore=0, P=32, R=4096), 1, 0,
                        keep=True)["out_ids"].astype(np.int64)
    feats, cls, logr, rec_tab, _ = probe.make_labels(torch, ix, Qtr, k, 4096, cfg["n_doc"], target=0.9,
                                                     log=lambda m: log("[bench]", m), gnd=deep)
    model = probe.ProbeModel(seed=1)
    losses = probe.train(torch, model, feats, cls, logr, epochs=args.epochs, log=lambda m: log("[bench]", m))
    model.attach(ix, conf=0.6, fallback=16, r_min=k)
    log(f"[bench] learned probe: {n_train} training queries labelled and trained in {time.time() - t1:.1f}s; "
        f"class counts {np.bincount(cls, minlength=5).tolist()}")
    mode = dict(P=32, R=4096, precision="bf16", rescore=200, learned=True)
    run = run_pipeline(torch, ix, Q, k, args.batch, mode, steps, warm, keep=True)
    deep_ev = run_pipeline(torch, ix, Q, k, 4096, dict(precision="fp32_mfma", rescore=0, P=32, R=4096), 1, 0,
                           keep=True)["out_ids"]
    _, _, st = ix.search(Q[:1024].cpu().numpy(), 32, 4096, k, "bf16", 200, stages=True, learned=True, batch=1024)
    emit("bf16_rescore200_learned", mode, run,
         {"learned_probe": {"training_queries": n_train, "label": "smallest nprobe in {2,4,8,16,32} whose top-100 "
                            "holds >= 90% of the nprobe-32 / probe_topk-4096 top-100",
                            "class_counts": np.bincount(cls, minlength=5).tolist(),
                            "losses": [round(float(x), 4) for x in losses],
                            "mean_nprobe": float(st["nprobe"].mean()), "mean_rerank": float(st["rerank_n"].mean()),
                            "recall_vs_deepest_probe": overlap(run["out_ids"], deep_ev, k)}})
    from mvr._capi import lib
    lib().mvr_index_set_probe(ix._h, None, 0.6, 16, 0)

    # recall-QPS frontier (BF16, FP32 rescore of 200)
    pts = []
    for p_ in (8, 16, 32):
        for r_ in (1000, 4096):
            mode = dict(P=p_, R=r_, precision="bf16", rescore=200)
            run = run_pipeline(torch, ix, Q, k, args.batch, mode, 2, 1, keep=True)
            pts.append({"nprobe": p_, "probe_topk": r_, "qps": round(run["qps"], 1),
                        "search_vs_fp64": overlap(run["out_ids"][:nr], gnd, k),
                        "search_vs_exhaustive_2bit": overlap(run["out_ids"][:nr], dec, k)})
            log(f"[bench] frontier nprobe {p_} probe_topk {r_}: {pts[-1]}")
    print(json.dumps({"metric": "recall-QPS frontier (BF16 MaxSim, FP32 rescore of 200)", "mode": "frontier",
                      "unit": "queries/s", "n_gpus": 1, "config": {"workload": workload},
                      "exhaustive_2bit_vs_fp64": ceiling, "points": pts}), flush=True)


def synth_queries(torch, seed, corpus, n):
    from mvr import synth
    return synth.make_queries(torch, seed, corpus.centers, n)


def bench_learned(torch, args, cfg):
    """Learned probing against fixed nprobe (recall@k vs exhaustive MaxSim over
    the index, queries/s, mean nprobe and rerank depth); one JSON line."""
    from mvr import probe, synth
    t0 = time.time()
    doclens = synth.doclens_uniform(args.seed, cfg["n_doc"], 16, 128)
    if cfg.get("topic"):
        corpus = synth.topic_corpus(torch, args.seed, cfg["n_doc"], cfg["n_comp"], cfg["n_c"], doclens,
                                    n_topics=128, comps_per_topic=16, p_topic=0.5, log=lambda m: log("[bench]", m))
        kw = dict(noise_lo=0.05, noise_hi=0.6, p_lo=0.3, p_hi=1.0)
        Qtr = synth.topic_queries(torch, corpus, args.seed + 1, cfg["n_train"], **kw)
        Qev = synth.topic_queries(torch, corpus, args.seed + 2, cfg["n_q"], **kw)
    else:
        corpus = synth.build_corpus(torch, args.seed, cfg["n_doc"], cfg["n_comp"], cfg["n_c"], doclens,
                                    n_iter=cfg["n_iter"], log=lambda m: log("[bench]", m))
        Qtr = synth.make_queries(torch, args.seed + 1, corpus.centers, cfg["n_train"])
        Qev = synth.make_queries(torch, args.seed + 2, corpus.centers, cfg["n_q"])
    log(f"[bench] corpus and {cfg['n_train']} + {cfg['n_q']} queries in {time.time() - t0:.1f}s")
    out = probe.learned_vs_fixed(torch, corpus.index, Qtr, Qev, cfg["k"], cfg["R"], cfg["n_doc"],
                                 target=cfg["target"], epochs=args.epochs, log=lambda m: log("[bench]", m))
    out.pop("model")
    learned = out["eval_full_depth"][-1]
    line = {"metric": f"learned-probe recall@{cfg['k']} and queries/s against fixed nprobe",
            "value": round(learned["qps"], 1), "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "bf16", "data": "synthetic" + (" topic-structured" if cfg.get("topic") else " mixture"),
            "config": {"workload": f"{args.config}: {cfg['n_doc']} docs, {cfg['n_c']} centroids, "
                                   f"{cfg['n_train']} training / {cfg['n_q']} evaluation queries, "
                                   f"labels: smallest nprobe with recall@{cfg['k']} >= {cfg['target']} at "
                                   f"rerank depth {cfg['R']}; MLP 1024-512-256-6, Adam lr 1e-3, batch 1024, "
                                   f"{args.epochs} epochs, fallback nprobe 16 below softmax 0.6"},
            **out}
    print(json.dumps(line), flush=True)


def bench_exhaustive(torch, args, cfg):
    """All-pairs MaxSim through the groundtruth module's C ABI
    (mvr_maxsim_scores_host_ex, host arrays as BruteForceProgressive passes
    them): seconds per call, whole-call TFLOP/s (2 x 128 flops per query
    token x doc token, PCIe copies included) for the exact FMA chains and for
    FP32 MFMA, and the largest score difference between the two."""
    from mvr import synth
    from mvr._capi import lib
    doclens = synth.doclens_uniform(args.seed, cfg["n_doc"], 16, 128)
    n_tok = int(doclens.sum())
    dev = torch.device("cuda")
    centers = synth.gen_centers(torch, args.seed, cfg["n_comp"])
    X = torch.empty((n_tok, 128), dtype=torch.float32, device=dev)
    synth.synth_tokens(torch, args.seed, synth.STREAM_TOKCOMP, synth.STREAM_TOKNOISE, centers, n_tok, out=X)
    Q = synth.make_queries(torch, args.seed, centers, cfg["n_q"])
    Xh, Qh = X.cpu().numpy(), Q.cpu().numpy()
    del X
    lens = doclens.astype(np.uint32)
    n_q, T = Qh.shape[0], Qh.shape[1]
    flop = 2.0 * 128 * T * n_q * n_tok
    out, res = {}, {}
    for name, prec in (("fp32", 0), ("fp32_mfma", 3)):
        sc = np.zeros((n_q, lens.size), np.float32)
        args_c = (Qh.ctypes.data, n_q, T, 128, Xh.ctypes.data, lens.ctypes.data, lens.size, sc.ctypes.data, 0, prec)
        assert lib().mvr_maxsim_scores_host_ex(*args_c) == 0, lib().mvr_maxsim_last_error()  # warm-up
        ts = []
        for _ in range(max(1, min(args.steps, 3))):
            t0 = time.perf_counter()
            assert lib().mvr_maxsim_scores_host_ex(*args_c) == 0, lib().mvr_maxsim_last_error()
            ts.append(time.perf_counter() - t0)
        sec = float(np.median(ts))
        out[name] = sc
        res[name] = {"seconds_per_call": round(sec, 4), "tflops_whole_call": round(flop / sec / 1e12, 2),
                     "frac_of_fp32_mfma_peak": round(flop / sec / 1e12 / 157.3, 4)}
        log(f"[bench] exhaustive {name}: {sec:.3f} s per call, {flop / sec / 1e12:.1f} TFLOP/s")
    d = np.abs(out["fp32_mfma"] - out["fp32"]) / np.maximum(np.abs(out["fp32"]), 1e-6)
    line = {"metric": "exhaustive MaxSim TFLOP/s (groundtruth module)", "value": res["fp32_mfma"]["tflops_whole_call"],
            "unit": "TFLOP/s", "n_gpus": 1, "higher_is_better": True, "dtype": "fp32",
            "data": "synthetic (seeded mixture tokens and queries)",
            "config": {"workload": f"exhaustive: {n_q} queries x {T} tokens against {lens.size} docs / {n_tok} tokens "
                                   f"(U{{16..128}})"},
            "by_precision": res, "max_rel_score_diff_mfma_vs_fp32": float(d.max())}
    print(json.dumps(line), flush=True)


def bench_plaid(torch, args, cfg):
    """PLAID search (two-stage centroid filter, FP32 decompression, L2
    normalisation, zero-initialised MaxSim) on a PLAID index of the c2
    corpus; the C oracle of the same search on the host cores beside it,
    with the GPU results checked bit for bit on that sample.  One JSON line."""
    from mvr import plaid, synth
    t0 = time.time()
    doclens = synth.doclens_lognormal(args.seed, cfg["n_doc"], 64.0, 0.29, 8, 300)
    corpus = synth.build_corpus(torch, args.seed, cfg["n_doc"], cfg["n_comp"], cfg["n_c"], doclens,
                                n_iter=cfg["n_iter"], kmeans_sample=cfg["kmeans_sample"],
                                log=lambda m: log("[bench]", m))
    pa = synth.plaid_from_corpus(torch, corpus, log=lambda m: log("[bench]", m))
    n_q, k = cfg["n_q"], cfg["k"]
    Q = synth.make_queries(torch, args.seed, corpus.centers, n_q)
    corpus.index.close()
    corpus = None
    torch.cuda.empty_cache()
    g = plaid.PlaidIndex(**pa.index_kwargs())
    ncells, thr, ndocs = plaid.searcher_defaults(k)
    log(f"[bench] plaid setup {time.time() - t0:.1f}s: {int(doclens.sum())} tokens, ncells {ncells}, "
        f"threshold {thr}, ndocs {ndocs}")
    for _ in range(args.warmup):
        g.search(Q, k=k, batch=args.batch)
    torch.cuda.synchronize()
    keys = ("candidates_ms", "filter_ms", "refine_ms", "total_ms", "n_refine_ivf", "n_refine_filter",
            "n_vec_score_refine", "scores_ms", "n_pass")
    acc = {f: 0.0 for f in keys}
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ids, sc, tm = g.search(Q, k=k, timings=True, batch=args.batch)
        for f in keys:
            acc[f] += tm[f]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    qps = n_q * args.steps / elapsed
    st = {f: acc[f] / args.steps for f in keys}
    log(f"[bench] plaid {args.steps} steps: {elapsed * 1e3 / args.steps:.2f} ms/step, {qps:.0f} QPS; "
        + ", ".join(f"{f}={st[f]:.2f}" for f in keys[:4]))
    cpu = None
    if not args.no_cpu and args.cpu_queries > 0:
        from oracle import plaid as OP
        cores = len(os.sched_getaffinity(0))
        n_s = min(args.cpu_queries, n_q)
        oix = OP.PlaidIndex(**pa.index_kwargs())
        Qh = Q[:n_s].cpu().numpy()
        oix.search(Qh[:16], k, ncells, thr, ndocs, n_threads=cores)
        t2 = time.time()
        rid, rsc, _ = oix.search(Qh, k, ncells, thr, ndocs, n_threads=cores)
        wall = time.time() - t2
        same = bool(np.array_equal(rid, ids[:n_s]) and np.array_equal(rsc, sc[:n_s]))
        cpu = {"value": round(n_s / wall, 2), "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"first {n_s} queries, same PLAID index and settings",
               "seconds": round(wall, 2), "gpu_results_bitexact": same}
        log(f"[bench] plaid cpu baseline {n_s / wall:.1f} QPS on {cores} cores; GPU results bit-exact: {same}")
    # roofline of the dominant kernel, the exact FP32 centroid-score GEMM
    # (k_pl_sgemm): 2 x 128 x 32 x n_c flops per query, over its launch
    # durations (HIP events around it on the search stream)
    n_c = cfg["n_c"]
    flops_launch = 2.0 * 128 * 32 * n_c * n_q / st["n_pass"]
    avg_ms = st["scores_ms"] / st["n_pass"]
    ach = flops_launch / (avg_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": round(ach, 2), "peak": 157.3, "unit": "TFLOP/s",
            "frac": round(ach / 157.3, 3), "traffic": None, "kernel": "k_pl_sgemm",
            "note": "FP32 FMA chains on the vector ALUs (exact order); peak is MI355X's FP32 vector figure",
            "avg_launch_ms": round(avg_ms, 3), "launches_per_step": st["n_pass"]}
    line = {
        "metric": f"PLAID search queries/sec (searcher defaults for k={k}), 1 GPU",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (c2 corpus as a PLAID index, built on the device)",
        "config": {"workload": f"{args.config}: {cfg['n_doc']} docs / {int(doclens.sum())} tokens, "
                               f"{cfg['n_c']} centroids, 2-bit residuals; {n_q} queries x 32 tokens; "
                               f"ncells {ncells}, centroid_score_threshold {thr}, ndocs {ndocs}, k {k}"},
        "stages_ms_per_step": {f: round(st[f], 3) for f in keys[:4]},
        "roofline": roof,
        "per_query": {"n_refine_ivf": st["n_refine_ivf"] / n_q, "n_refine_filter": st["n_refine_filter"] / n_q,
                      "n_vec_score_refine": st["n_vec_score_refine"] / n_q},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    g.close()


def roofline(kname, achieved_tf, peak_tf, bytes_alg, avg_launch_ms, n_launch, tok_per_launch, flops, traffic):
    """Both roofs of the MaxSim kernel -- matrix-unit FLOP/s and HBM bytes/s
    (algorithmic: 36 B of codes per doc token) -- with the binding one (the
    larger fraction) as the headline."""
    if achieved_tf is None:
        return {"kernel": kname, "bound": None, "achieved": None, "peak": None, "unit": None, "frac": None,
                "traffic": traffic, "note": "graph replay: no per-kernel timing"}
    gbs = bytes_alg / (avg_launch_ms * 1e-3) / 1e9
    mfma = {"bound": "mfma", "achieved": round(achieved_tf, 3), "peak": peak_tf, "unit": "TFLOP/s",
            "frac": round(achieved_tf / peak_tf, 4)}
    hbm = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4)}
    main, other = (mfma, hbm) if mfma["frac"] >= hbm["frac"] else (hbm, mfma)
    return {"kernel": kname, **main, "traffic": traffic, "other_roof": other,
            "avg_launch_ms": round(avg_launch_ms, 4), "launches": n_launch,
            "doc_tokens_per_launch": int(tok_per_launch), "flop_per_launch": flops,
            "flop_per_doc_token_and_query_token": 2 * 128, "alg_bytes_per_doc_token": 36,
            "measured_GBs": round(traffic / (avg_launch_ms * 1e-3) / 1e9, 1) if traffic else None}


def pipeline_roofline(r, precision, config):
    """roofline() of a pipeline's MaxSim kernel from its own timed run."""
    tm, n_launch = r["tm"], r["n_launch"]
    if not n_launch or not tm or not tm["n_pass"]:
        return None
    avg_ms = r["acc"]["rerank_ms"] / n_launch
    tok = tm["rerank_tokens"] / tm["n_pass"]
    flops = 2.0 * 128 * tm["rerank_pairs"] / tm["n_pass"]
    kname = {"fp32": "k_rerank_f32", "fp32_mfma": "k_rerank_f32mfma"}[precision]
    achieved = flops / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else None
    return roofline(kname, achieved, PEAK_TFLOPS[precision], 36.0 * tok, avg_ms, n_launch, tok, flops,
                    measured_traffic(kname, config, precision))


def measured_traffic(kname, config, precision):
    """HBM bytes per launch of `kname` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when that workload has not been profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rows = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rows:
        if r.get("kernel") == kname and r.get("config") == config and r.get("precision") == precision:
            return r.get("bytes_per_launch")
    return None


def recall_report(torch, args, cfg, corpus, ix, Q, answers, k, doclens, q_nt=None):
    """recall@k of each pipeline's answers (answers: {name: ids [n_q, k]})
    against FP64 exhaustive MaxSim over the original embeddings, and the two
    numbers that bound it: exhaustive FP32 MaxSim over the decompressed
    (2-bit) index against the same FP64 truth (the ceiling any search over
    this index can reach), and each search against that exhaustive
    decompressed ranking (what the probing loses)."""
    nr = min(args.recall_queries, Q.shape[0])
    if nr <= 0 or corpus.X is None:
        return None
    t1 = time.time()
    gnd, dec = ground_truth(torch, corpus, ix, Q, k, doclens, nr, q_nt)
    rep = {"queries": nr, "k": k, "exhaustive_2bit_vs_fp64": overlap(dec, gnd, k), "by_pipeline": {}}
    for name, ids in answers.items():
        rep["by_pipeline"][name] = {"search_vs_fp64": overlap(ids[:nr], gnd, k),
                                    "search_vs_exhaustive_2bit": overlap(ids[:nr], dec, k)}
        log(f"[bench] recall@{k} of {name} on {nr} queries: vs FP64 exhaustive "
            f"{rep['by_pipeline'][name]['search_vs_fp64']:.4f}, vs exhaustive 2-bit "
            f"{rep['by_pipeline'][name]['search_vs_exhaustive_2bit']:.4f}")
    log(f"[bench] exhaustive over the 2-bit index vs FP64: {rep['exhaustive_2bit_vs_fp64']:.4f} "
        f"({time.time() - t1:.1f}s)")
    return rep


def overlap(a, b, k):
    return float(np.mean([len(set(np.asarray(x).tolist()) & set(np.asarray(y).tolist())) / k for x, y in zip(a, b)]))


def ground_truth(torch, corpus, ix, Q, k, doclens, nr, q_nt=None):
    """(FP64 exhaustive top-k over the original embeddings, exhaustive FP32
    top-k over the decompressed 2-bit index) of the first nr queries."""
    from mvr import _capi
    import ctypes as C
    n_doc = doclens.size
    dl = torch.from_numpy(doclens.astype(np.int32)).cuda()
    s64 = torch.empty((nr, n_doc), dtype=torch.float64, device="cuda")
    qt_d = None if q_nt is None else torch.from_numpy(q_nt[:nr].astype(np.int32)).cuda()
    _capi.check(_capi.lib().mvr_exhaustive_f64(Q.data_ptr(), nr, Q.shape[1],
                                               None if qt_d is None else qt_d.data_ptr(),
                                               corpus.X.data_ptr(), dl.data_ptr(), n_doc, s64.data_ptr(), None))
    gnd = torch.topk(s64, k, dim=1).indices.cpu().numpy()
    del s64
    docs = torch.arange(n_doc, dtype=torch.int32, device="cuda")
    dec = np.empty((nr, k), np.int64)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for q0 in range(0, nr, 64):
        nb = min(64, nr - q0)
        s32 = torch.empty((nb, n_doc), dtype=torch.float32, device="cuda")
        qt_h = None if q_nt is None else np.ascontiguousarray(q_nt[q0:q0 + nb], np.uint32)
        _capi.check(_capi.lib().mvr_rerank_docs(ix._h, Q[q0:q0 + nb].contiguous().data_ptr(), nb, Q.shape[1],
                                                None if qt_h is None else qt_h.ctypes.data_as(C.c_void_p),
                                                docs.expand(nb, n_doc).contiguous().data_ptr(), None, n_doc,
                                                0, s32.data_ptr(), stream))
        dec[q0:q0 + nb] = torch.topk(s32, k, dim=1).indices.cpu().numpy()
    return gnd, dec


def copy_probe(mib):
    """A plain device copy of known size, the calibration run for FETCH_SIZE /
    WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}), flush=True)


def cpu_baseline(torch, corpus, Q, P, R, k, n_sample, ref=None, mref=None):
    """The C oracle of the same search (OpenMP over queries) on the host cores,
    on the first n_sample queries of the batch, and its answers checked
    against the GPU's: bit for bit against the FP32 pipeline (ref), and up to
    ties within 1e-5 against FP32 MFMA (mref)."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        log(f"[bench] cpu baseline unavailable: {e}")
        return None
    cores = len(os.sched_getaffinity(0))
    t0 = time.time()
    ivf = (corpus.ivf[0].cpu().numpy().astype(np.uint64), corpus.ivf[1].cpu().numpy().astype(np.uint32),
           corpus.ivf[2].cpu().numpy().astype(np.uint32))
    ix = O.Index(corpus.doclens, corpus.codes.cpu().numpy().astype(np.uint32), corpus.residual.cpu().numpy(),
                 corpus.C.cpu().numpy(), corpus.W.cpu().numpy(), 2, ivf)
    Qh = Q[:n_sample].cpu().numpy()
    log(f"[bench] cpu baseline: index on host in {time.time() - t0:.1f}s, {cores} cores")
    ix.search(Qh[:16], P, R, k, n_threads=cores)  # warm the host index pages
    t1 = time.time()
    ids, s, secs = ix.search(Qh, P, R, k, n_threads=cores, timing=True)
    wall = time.time() - t1
    out = {"value": round(n_sample / wall, 2), "unit": "queries/s", "cores": cores, "kind": "port",
           "sample": f"first {n_sample} queries of the batch, same index and parameters; "
                     f"p50 {np.percentile(secs, 50) * 1e3:.1f} ms/query on one core",
           "seconds": round(wall, 2),
           "build": "oracle/mvr_oracle.c (the correctness oracle): gcc -O3 -mavx2 -mfma -fopenmp, scalar "
                    "fixed-order FMA chains; built off the box, so no -march=native"}
    if ref is not None:
        gi, gs = ref["out_ids"][:n_sample], ref["out_s"][:n_sample]
        out["gpu_fp32_vs_oracle"] = {"queries_ids_equal": round(float(np.mean(np.all(gi == ids, axis=1))), 4),
                                     "queries_scores_bitequal": round(float(np.mean(np.all(
                                         gs.view(np.uint32) == s.view(np.uint32), axis=1))), 4)}
    if mref is not None:
        gi, gs = mref["out_ids"][:n_sample], mref["out_s"][:n_sample]
        tie_ok = []
        for q in range(n_sample):
            # equal up to ties: position-wise scores within 1e-5, and every id in
            # only one of the two lists scores within 1e-5 of that list's k-th
            oq, gq = ids[q].tolist(), gi[q].tolist()
            tol = 1e-5 * max(1.0, abs(float(s[q][-1])))
            ok = bool(np.allclose(gs[q], s[q], rtol=1e-5, atol=1e-6))
            ok = ok and all(abs(float(s[q][oq.index(d)]) - float(s[q][-1])) <= tol for d in set(oq) - set(gq))
            ok = ok and all(abs(float(gs[q][gq.index(d)]) - float(gs[q][-1])) <= tol for d in set(gq) - set(oq))
            tie_ok.append(ok)
        out["gpu_fp32_mfma_vs_oracle"] = {"queries_equal_up_to_ties": round(float(np.mean(tie_ok)), 4),
                                          "max_abs_score_diff": float(np.abs(gs - s).max())}
    log(f"[bench] cpu baseline {out['value']} QPS; " + ", ".join(f"{kk}: {v}" for kk, v in out.items()
                                                                 if kk.startswith("gpu_")))
    return out


if __name__ == "__main__":
    main()
