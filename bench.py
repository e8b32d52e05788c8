"""Benchmark of the ColBERT late-interaction search path on one MI355X.

Workload (default, --config c2): 1,048,576 docs with lognormal token counts
(median 64, sigma 0.29, clipped to [8, 300]; ~70M tokens), 65,536 mixture
components, 65,536 k-means centroids, 2-bit residuals; 4,096 queries x 32
tokens; nprobe 8, probe_topk (rerank) 1,000, k 100.  The corpus and index are
built on the device (synthetic data, seeded); a step is one search of the
4,096-query batch with inputs resident in HBM.

Prints one JSON line: queries/sec (value), ms per step, p50/p99 latency of a
1,024-query device pass, recall@100 against FP64 exhaustive MaxSim over the
original embeddings, the rerank kernel's roofline figures, and a CPU baseline
(the C oracle of the same path on the host cores, on a bounded query sample).
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

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0}

CONFIGS = {
    "c2": dict(n_doc=1 << 20, n_comp=65536, n_c=65536, lognormal=True, n_q=4096, P=8, R=1000, k=100,
               kmeans_sample=1 << 21, n_iter=4),
    "c1": dict(n_doc=8192, n_comp=1024, n_c=1024, lognormal=False, n_q=64, P=4, R=256, k=10,
               kmeans_sample=None, n_iter=10),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp8"])
    ap.add_argument("--rescore", type=int, default=200)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--recall-queries", type=int, default=256)
    ap.add_argument("--cpu-queries", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path is query "
              f"data-parallel and would shard by replicating the index per GPU.", file=sys.stderr)
        sys.exit(2)

    import torch
    import mvr
    from mvr import synth

    cfg = CONFIGS[args.config]
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
    n_q, P, R, k = cfg["n_q"], cfg["P"], cfg["R"], cfg["k"]
    Q = synth.make_queries(torch, args.seed, corpus.centers, n_q)
    out_ids = torch.empty((n_q, k), dtype=torch.int32, device="cuda")
    out_s = torch.empty((n_q, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    log(f"[bench] setup {time.time() - t_setup:.1f}s; max IVF list {ix.max_list}")
    stream = torch.cuda.current_stream().cuda_stream
    rescore = args.rescore if args.precision != "fp32" else 0
    batch = 1024

    def step(timings=False):
        if timings:
            import ctypes as C
            from mvr import _capi
            tm = _capi.Timings()
            prm = mvr.Index._params(P, R, k, args.precision, rescore, False, batch)
            _capi.check(_capi.lib().mvr_search(ix._h, Q.data_ptr(), n_q, 32, None, None, C.byref(prm),
                                               out_ids.data_ptr(), out_s.data_ptr(), 1, C.c_void_p(stream),
                                               None, C.byref(tm)))
            return {f: getattr(tm, f) for f, _ in _capi.Timings._fields_}
        ix.search_device(Q, out_ids, out_s, P, R, k, args.precision, rescore, batch=batch, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # timed region: K whole-batch searches, per-pass latency from HIP events
    ev = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for b0 in range(0, n_q, batch):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            ix.search_device(Q[b0:b0 + batch], out_ids[b0:b0 + batch], out_s[b0:b0 + batch], P, R, k,
                             args.precision, rescore, batch=batch, stream=stream)
            e1.record()
            ev.append((e0, e1))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    qps = n_q * args.steps / elapsed
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, {qps:.0f} QPS")

    # one instrumented step: per-stage device time and work counters
    tm = step(timings=True)
    torch.cuda.synchronize()
    rr_ms = tm["rerank_ms"]
    n_pass = (n_q + batch - 1) // batch
    flops = 2.0 * 32 * tm["rerank_tokens"] * 128
    bytes_alg = 36.0 * tm["rerank_tokens"]
    peak = PEAK_TFLOPS[args.precision]
    achieved_tf = flops / (rr_ms / n_pass * 1e-3) / n_pass / 1e12 if rr_ms > 0 else 0.0
    log(f"[bench] stages (ms/step): " + ", ".join(f"{f}={tm[f]:.2f}" for f in
                                                 ("centroid_ms", "filter_ms", "select_ms", "rerank_ms", "topk_ms")))

    # recall@k against FP64 exhaustive MaxSim over the original embeddings
    nr = min(args.recall_queries, n_q)
    recall = None
    if nr > 0 and corpus.X is not None:
        ids_h = out_ids[:nr].cpu().numpy()
        dl = torch.from_numpy(doclens.astype(np.int32)).cuda()
        scores = torch.empty((nr, cfg["n_doc"]), dtype=torch.float64, device="cuda")
        t1 = time.time()
        from mvr import _capi
        _capi.check(_capi.lib().mvr_exhaustive_f64(Q.data_ptr(), nr, 32, None, corpus.X.data_ptr(),
                                                   dl.data_ptr(), cfg["n_doc"], scores.data_ptr(), None))
        gnd = torch.topk(scores, k, dim=1).indices.cpu().numpy()
        recall = float(np.mean([len(set(a.tolist()) & set(b.tolist())) / k for a, b in zip(ids_h, gnd)]))
        log(f"[bench] recall@{k} vs FP64 exhaustive on {nr} queries: {recall:.4f} ({time.time() - t1:.1f}s)")

    cpu = None
    if not args.no_cpu and args.cpu_queries > 0:
        cpu = cpu_baseline(torch, corpus, Q, P, R, k, args.cpu_queries)

    line = {
        "metric": "queries/sec at recall@100>=0.99 vs FP64 exact MaxSim; p50/p99 latency, 1 GPU",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded mixture corpus and queries, index built on the device)",
        "config": {"workload": f"{args.config}: {cfg['n_doc']} docs / {int(doclens.sum())} tokens, "
                               f"{cfg['n_c']} centroids, 2-bit residuals; {n_q} queries x 32 tokens; "
                               f"nprobe {P}, probe_topk {R}, k {k}",
                   "batch_per_pass": batch, "precision": args.precision, "rescore": rescore},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "unit": f"per {batch}-query device pass"},
        "recall_at_k": recall,
        "stages_ms_per_step": {f: round(tm[f], 3) for f in
                               ("centroid_ms", "filter_ms", "select_ms", "rerank_ms", "topk_ms", "total_ms")},
        "roofline": {"kernel": "k_rerank_f32" if args.precision == "fp32" else f"k_rerank_{args.precision}",
                     "bound": "mfma", "achieved": round(achieved_tf, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved_tf / peak, 4), "traffic": None,
                     "hbm_achieved_GBs": round(bytes_alg / (rr_ms * 1e-3) / 1e9, 1) if rr_ms > 0 else None,
                     "rerank_tokens_per_step": int(tm["rerank_tokens"])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(torch, corpus, Q, P, R, k, n_sample):
    """The C oracle of the same search (OpenMP over queries) on the host cores,
    on the first n_sample queries of the batch."""
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
    ids, s, secs = ix.search(Qh, P, R, k, n_threads=cores, timing=True)
    t1 = time.time()
    ids, s, secs = ix.search(Qh, P, R, k, n_threads=cores, timing=True)
    wall = time.time() - t1
    return {"value": round(n_sample / wall, 2), "unit": "queries/s", "cores": cores, "kind": "port",
            "sample": f"first {n_sample} queries of the batch, same index and parameters; "
                      f"p50 {np.percentile(secs, 50) * 1e3:.1f} ms/query on one core",
            "seconds": round(wall, 2)}


if __name__ == "__main__":
    main()
