"""Learned probing: a small MLP that picks nprobe (and probe_topk) per query.

Features: each query token's top-32 centroid scores (S' - 1, sorted), 32 x 32
= 1,024 values.  Labels: the smallest nprobe in {2, 4, 8, 16, 32} whose
search reaches the target recall@k against exhaustive MaxSim over the
decompressed index (32 when none does), and log of the smallest probe_topk
that reaches it at that nprobe.  The network trains and runs on libmvr's HIP
kernels (mvr_mlp_*); torch only holds device memory.  The plain fixed-nprobe
pipeline is the fallback when the model is unsure (top class probability
below `conf`) and the yardstick of its recall loss.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._capi import check, check_mlp, lib

CLASSES = (2, 4, 8, 16, 32)
DIMS = (1024, 512, 256, 6)


def init_params(seed: int, dims=DIMS) -> np.ndarray:
    """Flat W1 b1 W2 b2 W3 b3; W ~ U(+-1/sqrt(fan_in)) (seeded), b = 0."""
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(len(dims) - 1):
        bound = 1.0 / math.sqrt(dims[i])
        parts.append(rng.uniform(-bound, bound, (dims[i + 1], dims[i])).astype(np.float32).ravel())
        parts.append(np.zeros(dims[i + 1], np.float32))
    return np.concatenate(parts)


def split_params(flat, dims=DIMS):
    out, o = [], 0
    for i in range(len(dims) - 1):
        n = dims[i + 1] * dims[i]
        out.append(flat[o:o + n].reshape(dims[i + 1], dims[i]))
        o += n
        out.append(flat[o:o + dims[i + 1]])
        o += dims[i + 1]
    return out


class ProbeModel:
    """Handle to an mvr_mlp on the device."""

    def __init__(self, dims=DIMS, device: int = 0, seed: int = 0):
        self.dims = tuple(dims)
        d = (C.c_uint32 * 4)(*dims)
        self._h = C.c_void_p()
        check_mlp(lib().mvr_mlp_create(d, device, C.byref(self._h)))
        self.n_params = int(lib().mvr_mlp_num_params(self._h))
        self.set_params(init_params(seed, dims))

    def close(self):
        if self._h:
            lib().mvr_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == self.n_params
        check_mlp(lib().mvr_mlp_set_params(self._h, flat.ctypes.data))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check_mlp(lib().mvr_mlp_get_params(self._h, out.ctypes.data))
        return out

    def forward(self, torch, x):
        out = torch.empty((x.shape[0], 6), dtype=torch.float32, device=x.device)
        check_mlp(lib().mvr_mlp_forward(self._h, x.data_ptr(), x.shape[0], out.data_ptr(),
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out

    def grads(self, x, y, r):
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        check_mlp(lib().mvr_mlp_grads(self._h, x.data_ptr(), y.data_ptr(), r.data_ptr(), x.shape[0],
                                      g.ctypes.data, C.byref(loss)))
        return float(loss.value), g

    def train_step(self, torch, x, y, r, lr=1e-3, want_loss=False):
        loss = C.c_float()
        check_mlp(lib().mvr_mlp_train_step(self._h, x.data_ptr(), y.data_ptr(), r.data_ptr(), x.shape[0],
                                           lr, C.byref(loss) if want_loss else None,
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return float(loss.value) if want_loss else None

    def attach(self, index, conf=0.6, fallback=16, r_min=0):
        check(lib().mvr_index_set_probe(index._h, self._h, conf, fallback, r_min))


def features_from_top_s(top_s: np.ndarray, q_ntok=None) -> np.ndarray:
    """[n_q, t_max, 32] stage scores (S') -> [n_q, 1024] features (S' - 1,
    zero rows past the query's token count), the layout the search computes."""
    n_q, t_max, P = top_s.shape
    assert P == 32 and t_max <= 32
    f = np.zeros((n_q, 32, 32), np.float32)
    f[:, :t_max] = top_s - 1.0
    if q_ntok is not None:
        for q, T in enumerate(q_ntok):
            f[q, T:] = 0.0
    return f.reshape(n_q, 1024)


def exhaustive_topk(torch, index, Q, k, n_doc, batch=512, rescore=64):
    """Top-k by exhaustive MaxSim over the decompressed index: BF16 MaxSim over
    every doc, then FP32 (bit-exact with the search) on the best k + rescore."""
    dev = Q.device
    n_q = Q.shape[0]
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    all_docs = torch.arange(n_doc, dtype=torch.int32, device=dev)
    out = torch.empty((n_q, k), dtype=torch.int64, device=dev)
    kk = min(n_doc, k + rescore)
    for b0 in range(0, n_q, batch):
        nb = min(batch, n_q - b0)
        Qb = Q[b0:b0 + nb].contiguous()
        docs = all_docs.expand(nb, n_doc).contiguous()
        sc = torch.empty((nb, n_doc), dtype=torch.float32, device=dev)
        check(lib().mvr_rerank_docs(index._h, Qb.data_ptr(), nb, Qb.shape[1], None, docs.data_ptr(), None,
                                    n_doc, 1, sc.data_ptr(), stream))
        cand = torch.topk(sc, kk, dim=1).indices.to(torch.int32).contiguous()
        sc2 = torch.empty((nb, kk), dtype=torch.float32, device=dev)
        check(lib().mvr_rerank_docs(index._h, Qb.data_ptr(), nb, Qb.shape[1], None, cand.data_ptr(), None,
                                    kk, 0, sc2.data_ptr(), stream))
        # order by (score desc, id asc) like the search's final selection:
        # stable sort by id, then stable sort by descending score
        o1 = torch.argsort(cand, dim=1, stable=True)
        c1, s1 = torch.gather(cand, 1, o1), torch.gather(sc2, 1, o1)
        o2 = torch.argsort(-s1, dim=1, stable=True)
        out[b0:b0 + nb] = torch.gather(c1, 1, o2)[:, :k].long()
    return out


def make_labels(torch, index, Q, k, R_max, n_doc, target=0.99, batch=1024, log=None, gnd=None):
    """Features and labels for training queries Q (torch [n, 32, 128]).
    The reference ranking is exhaustive MaxSim over the index, or the given
    top-k ids gnd [n, k].  Returns (features [n, 1024], class index [n],
    log R needed [n], recall table [n, 5] at R_max, reference top-k ids
    [n, k])."""
    dev = Q.device
    n_q = Q.shape[0]
    if gnd is None:
        gnd = exhaustive_topk(torch, index, Q, k, n_doc)
        if log:
            log(f"labels: exhaustive top-{k} of {n_q} queries done")
    else:
        gnd = torch.as_tensor(np.asarray(gnd, np.int64), device=dev)
    feats = np.empty((n_q, 1024), np.float32)
    rec = torch.zeros((n_q, len(CLASSES)), dtype=torch.float64, device=dev)
    r_need = torch.full((n_q, len(CLASSES)), float(R_max), dtype=torch.float64, device=dev)
    need = int(math.ceil(target * k))
    Qh = Q.cpu().numpy()
    for ci, P in enumerate(CLASSES):
        for b0 in range(0, n_q, batch):
            nb = min(batch, n_q - b0)
            _, _, st = index.search(Qh[b0:b0 + nb], P, R_max, k, stages=True, batch=batch)
            if P == 32:
                feats[b0:b0 + nb] = features_from_top_s(st["top_s"])
            sel = torch.from_numpy(st["sel_ids"].astype(np.int64)).to(dev)          # [nb, R]
            g = gnd[b0:b0 + nb]                                                      # [nb, k]
            eq = sel[:, :, None] == g[:, None, :]                                    # [nb, R, k]
            found = eq.any(dim=1)
            pos = torch.where(found, eq.float().argmax(dim=1), torch.full_like(g, R_max))
            pos, _ = torch.sort(pos, dim=1)
            hit = found.sum(dim=1)
            rec[b0:b0 + nb, ci] = hit.double() / k
            ok = hit >= need
            r_need[b0:b0 + nb, ci] = torch.where(ok, torch.clamp(pos[:, need - 1] + 1, min=k).double(),
                                                 torch.full_like(r_need[b0:b0 + nb, ci], float(R_max)))
        if log:
            log(f"labels: nprobe {P} mean recall@{k} {rec[:, ci].mean().item():.4f}")
    ok = rec >= target
    cls = torch.where(ok.any(dim=1), ok.double().argmax(dim=1), torch.full((n_q,), len(CLASSES) - 1,
                                                                           device=dev, dtype=torch.int64))
    logr = torch.log(r_need[torch.arange(n_q, device=dev), cls])
    return (feats, cls.cpu().numpy().astype(np.int32), logr.cpu().numpy().astype(np.float32),
            rec.cpu().numpy(), gnd.cpu().numpy())


def train(torch, model, feats, cls, logr, epochs=8, batch=1024, lr=1e-3, seed=0, log=None):
    """Adam on shuffled mini-batches; returns the mean loss of each epoch."""
    dev = torch.device("cuda")
    X = torch.from_numpy(np.ascontiguousarray(feats)).to(dev)
    Y = torch.from_numpy(np.ascontiguousarray(cls, dtype=np.int32)).to(dev)
    Rt = torch.from_numpy(np.ascontiguousarray(logr, dtype=np.float32)).to(dev)
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    losses = []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).to(dev)
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            l = model.train_step(torch, X[idx].contiguous(), Y[idx].contiguous(), Rt[idx].contiguous(),
                                 lr, want_loss=(b0 // batch) % 16 == 0)
            if l is not None:
                tot += l
                cnt += 1
        losses.append(tot / max(cnt, 1))
        if log:
            log(f"probe epoch {ep}: loss {losses[-1]:.4f}")
    return losses


def evaluate(torch, index, Q, k, R, gnd, nprobes=CLASSES, learned=False, precision="bf16", rescore=0,
             batch=1024, reps=3):
    """recall@k against `gnd` [n_q, k] and queries/s of fixed-nprobe searches
    (and of learned probing when a model is attached).  Returns a list of
    dicts: mode, nprobe (mean), rerank (mean docs reranked), recall, qps."""
    import time as _time
    n_q = Q.shape[0]
    Qh = Q.cpu().numpy()
    dev = Q.device
    oi = torch.empty((n_q, k), dtype=torch.int32, device=dev)
    os_ = torch.empty((n_q, k), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    modes = [("fixed", P) for P in nprobes] + ([("learned", None)] if learned else [])
    out = []
    for mode, P in modes:
        lp = mode == "learned"
        nprobe = 32 if lp else P
        ids, _, st = index.search(Qh, nprobe, R, k, precision=precision, rescore=rescore, batch=batch,
                                  stages=True, learned=lp)
        rec = float(np.mean([len(set(a.tolist()) & set(b.tolist())) / k for a, b in zip(ids, gnd)]))
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = _time.perf_counter()
            index.search_device(Q, oi, os_, nprobe, R, k, precision=precision, rescore=rescore, batch=batch,
                                stream=stream, learned=lp)
            torch.cuda.synchronize()
            dt = _time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        out.append({"mode": mode, "nprobe": float(st["nprobe"].mean()), "rerank": float(st["rerank_n"].mean()),
                    "recall": rec, "qps": n_q / best})
    return out


def learned_vs_fixed(torch, index, Q_train, Q_eval, k, R_max, n_doc, target=0.9, epochs=20, batch=1024,
                     lr=1e-3, conf=0.6, fallback=16, r_margin=0.0, seed=1, log=None):
    """The learned-probe experiment end to end: labels from exhaustive MaxSim
    over the index for the training queries, Adam training on the device,
    then recall@k (against exhaustive MaxSim) of learned probing next to fixed
    nprobe at the full rerank depth R_max and at the learned mean depth.
    r_margin is added to the log-R labels (a safety factor on the cutoff)."""
    feats, cls, logr, rec, _ = make_labels(torch, index, Q_train, k, R_max, n_doc, target=target, log=log)
    model = ProbeModel(seed=seed)
    losses = train(torch, model, feats, cls, logr + r_margin, epochs=epochs, batch=batch, lr=lr, log=log)
    _, _, _, _, gnd = make_labels(torch, index, Q_eval, k, R_max, n_doc, target=target)
    model.attach(index, conf=conf, fallback=fallback, r_min=k)
    res = evaluate(torch, index, Q_eval, k, R_max, gnd, learned=True)
    learned = res[-1]
    r_eq = max(k, int(round(learned["rerank"])))
    res_eq = evaluate(torch, index, Q_eval, k, r_eq, gnd, learned=False)
    index_detach = getattr(index, "_h", None)
    if index_detach is not None:
        check(lib().mvr_index_set_probe(index._h, None, conf, fallback, 0))
    return {"class_counts": np.bincount(cls, minlength=len(CLASSES)).tolist(),
            "train_recall_by_class": rec.mean(axis=0).round(4).tolist(),
            "losses": [round(float(l), 4) for l in losses],
            "eval_full_depth": res, "rerank_depth_of_learned": r_eq, "eval_same_depth": res_eq,
            "model": model}
