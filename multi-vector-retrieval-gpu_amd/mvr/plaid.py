"""PLAID (ColBERTv2) search path and residual codec ops on the GPU.

Drop-ins for the ops ColBERT loads as torch extensions, with the reference's
names, argument order and meaning (tensors in, tensors out), all computed by
HIP kernels in libmvr (include/mvr_plaid.h):

  decompress_residuals_cuda  indexing/codecs/decompress_residuals.cu:42-75
                             (ResidualCodec.decompress_residuals, residual.py:98-115)
  packbits_cuda              indexing/codecs/packbits.cu:39-57 (ResidualCodec.packbits)
  binarize                   ResidualCodec.binarize, residual.py:189-224
  decompress_residuals       search/decompress_residuals.cpp:87-155, the CPU
                             signature (IndexScorer.decompress_residuals)
  filter_pids                search/filter_pids.cpp:147-193 (IndexScorer.filter_pids)
  segmented_maxsim           modeling/segmented_maxsim.cpp:47-93 (ColBERT.segmented_maxsim)
  segmented_lookup           search/segmented_lookup.cpp:52-148 (StridedTensor.segmented_lookup)

the float16 functions of the reference's GPU branch:

  colbert_score_packed       modeling/colbert.py:181-205 (GPU branch: half scores, no zero clip)
  colbert_score_reduce       modeling/colbert.py:133-155 (-9999 padding in a half tensor)
  PlaidIndex.lookup_pids     residual_embeddings_strided.py:30-37 + ResidualCodec.decompress
                             (residual.py:243-279): half decompression and normalisation

and `PlaidIndex`, the device-resident IndexScorer: `rank` for one query
(index_storage.py:82-98) and `search` for a batch, running either branch of
score_pids: branch="cpu" (two-stage filter_pids, FP32 decompression, L2
normalisation, zero-initialised MaxSim; index_storage.py:154-190) or
branch="gpu" (float16 centroid scores, both filter stages through
colbert_score_reduce, half decompression and normalisation,
colbert_score_packed; index_storage.py:112-153,160-161,185-187).  Inputs on
the CPU are moved to the current GPU and results come back on the input's
device.  There is no CPU fallback: without the library every call raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as _c


def _torch():
    import torch
    return torch


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t, dtype):
    torch = _torch()
    if not torch.is_tensor(t):
        t = torch.as_tensor(np.asarray(t))
    return t.to(device="cuda", dtype=dtype).contiguous()


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _back(out, like):
    return out if like.is_cuda else out.cpu()


def _lib():
    return _c.lib()


# ------------------------------------------------------------ codec ops
def decompress_residuals_cuda(binary_residuals, bucket_weights, reversed_bit_map, bucket_weight_combinations,
                              codes, centroids, dim: int, nbits: int):
    """float16 [n, dim]: half(bucket_weights[lut[rbm[byte]][k]] + centroids[code])."""
    torch = _torch()
    res = _dev(binary_residuals, torch.uint8)
    bw = _dev(bucket_weights, torch.float16)
    rbm = _dev(reversed_bit_map, torch.uint8)
    lut = _dev(bucket_weight_combinations, torch.uint8)
    cd = _dev(codes, torch.int32)
    cent = _dev(centroids, torch.float16)
    n = res.shape[0]
    out = torch.empty((n, dim), dtype=torch.float16, device="cuda")
    _c.check_plaid(_lib().mvr_plaid_decompress_residuals(_ptr(res), n, _ptr(bw), _ptr(rbm), _ptr(lut), _ptr(cd),
                                                         _ptr(cent), dim, nbits, _ptr(out), _stream(torch)))
    return out


def packbits_cuda(residuals):
    """uint8 [m / 8]: numpy.packbits of the (nonzero -> 1) bytes; m % 32 == 0."""
    torch = _torch()
    r = _dev(residuals, torch.uint8).reshape(-1)
    out = torch.empty(r.numel() // 8, dtype=torch.uint8, device="cuda")
    _c.check_plaid(_lib().mvr_plaid_packbits(_ptr(r), r.numel(), _ptr(out), _stream(torch)))
    return out


def binarize(residuals, bucket_cutoffs, nbits: int):
    """uint8 [n, dim * nbits / 8]: bucketize(right=False) + bits LSB first + packbits."""
    torch = _torch()
    e = _dev(residuals, torch.float32)
    cut = _dev(bucket_cutoffs, torch.float32).reshape(-1)
    n, dim = e.shape
    out = torch.empty((n, dim * nbits // 8), dtype=torch.uint8, device="cuda")
    _c.check_plaid(_lib().mvr_plaid_binarize(_ptr(e), n, dim, _ptr(cut), cut.numel(), nbits, _ptr(out),
                                             _stream(torch)))
    return out


# ------------------------------------------------------------ search ops
def _cumsum(torch, pids, lengths):
    n = pids.numel() if pids is not None else lengths.numel()
    cum = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    total = C.c_int64(0)
    _c.check_plaid(_lib().mvr_plaid_lengths_cumsum(_ptr(pids) if pids is not None else None, n, _ptr(lengths),
                                                   _ptr(cum), C.byref(total), _stream(torch)))
    return cum, int(total.value)


def decompress_residuals(pids, lengths, offsets, bucket_weights, reversed_bit_map, bucket_weight_combinations,
                         binary_residuals, codes, centroids, dim: int, nbits: int):
    """float32 [sum(lengths[pids]), dim]: the tokens of pids in order,
    bucket_weights[lut[rbm[byte]][l]] + centroids[code] (CPU-side signature)."""
    torch = _torch()
    p = _dev(pids, torch.int32)
    ln = _dev(lengths, torch.int64)
    if p.numel() == 0:
        return _back(torch.zeros((0, dim), dtype=torch.float32, device="cuda"), pids if torch.is_tensor(pids)
                     else p)
    cum, total = _cumsum(torch, p, ln)
    out = torch.empty((total, dim), dtype=torch.float32, device="cuda")
    args = [_dev(offsets, torch.int64), _dev(bucket_weights, torch.float32), _dev(reversed_bit_map, torch.uint8),
            _dev(bucket_weight_combinations, torch.uint8), _dev(binary_residuals, torch.uint8),
            _dev(codes, torch.int32), _dev(centroids, torch.float32)]
    _c.check_plaid(_lib().mvr_plaid_decompress_pids(_ptr(p), p.numel(), _ptr(cum), _ptr(ln), *map(_ptr, args), dim,
                                                    nbits, _ptr(out), _stream(torch)))
    return _back(out, pids if torch.is_tensor(pids) else p)


def segmented_maxsim(scores, lengths):
    """float32 [ndocs]: sum_t max(0, max over each doc's rows of scores[:, t])."""
    torch = _torch()
    s = _dev(scores, torch.float32)
    ln = _dev(lengths, torch.int64)
    out = torch.empty(ln.numel(), dtype=torch.float32, device="cuda")
    nq = s.shape[1] if s.dim() == 2 else 0
    _c.check_plaid(_lib().mvr_plaid_segmented_maxsim(_ptr(s), nq, _ptr(ln), ln.numel(), _ptr(out),
                                                     _stream(torch)))
    return _back(out, scores if torch.is_tensor(scores) else s)


def segmented_lookup(input, pids, lengths, offsets):
    """segmented_lookup_cpp (strided_tensor.py:93, segmented_lookup.cpp:52-148):
    the rows input[offsets[i] : offsets[i] + lengths[i]] for i < len(pids),
    concatenated; input 1-D or 2-D, uint8 / int32 / int64 / float32 / float16.
    Runs on the device; the result is on the device when input is."""
    torch = _torch()
    src = input if torch.is_tensor(input) else torch.as_tensor(np.asarray(input))
    if src.dtype not in (torch.uint8, torch.int32, torch.int64, torch.float32, torch.float16):
        raise _c.MvrError(1, f"segmented_lookup: dtype {src.dtype} unsupported")
    if src.dim() not in (1, 2):
        raise _c.MvrError(1, "segmented_lookup: input must be 1-D or 2-D")
    x = src.to("cuda").contiguous()
    n = int(pids.shape[0]) if hasattr(pids, "shape") else len(pids)
    ln = _dev(lengths, torch.int64)[:n].contiguous()
    of = _dev(offsets, torch.int64)[:n].contiguous()
    if ln.numel() < n or of.numel() < n:
        raise _c.MvrError(1, "segmented_lookup: fewer lengths/offsets than pids")
    total = int(ln.sum().item()) if n else 0
    dim = 1 if x.dim() == 1 else int(x.shape[1])
    out = torch.empty((total,) if x.dim() == 1 else (total, dim), dtype=x.dtype, device="cuda")
    _c.check_plaid(_lib().mvr_plaid_segmented_lookup(_ptr(x), x.shape[0], dim * x.element_size(), _ptr(ln),
                                                     _ptr(of), n, _ptr(out), total, _stream(torch)))
    return _back(out, src)


def filter_pids(pids, centroid_scores, codes, doclens, offsets, idx, nfiltered_docs: int):
    """(int32 pids [<= nfiltered_docs // 4] by (score, pid) descending,
    n_vec_score_refine): the two-stage centroid filter."""
    torch = _torch()
    p = _dev(pids, torch.int32)
    S = _dev(centroid_scores, torch.float32)
    n_c, nq = S.shape
    ix = _dev(idx, torch.bool).view(torch.uint8)
    out = torch.empty(max(nfiltered_docs // 4, 1), dtype=torch.int32, device="cuda")
    n = C.c_uint32(0)
    nv = C.c_uint64(0)
    # keep every device copy alive until the op has run
    cd, dl, of = _dev(codes, torch.int32), _dev(doclens, torch.int64), _dev(offsets, torch.int64)
    _c.check_plaid(_lib().mvr_plaid_filter_pids(
        _ptr(p), p.numel(), _ptr(S), n_c, nq, _ptr(cd), _ptr(dl), _ptr(of), _ptr(ix), nfiltered_docs, _ptr(out),
        C.byref(n), C.byref(nv), _stream(torch)))
    res = out[: n.value].clone()
    return _back(res, pids if torch.is_tensor(pids) else p), int(nv.value)


def colbert_score_packed(Q, D_packed, D_lengths):
    """float16 [ndocs]: colbert_score_packed of the GPU branch for one query
    Q [1, nq, dim] or [nq, dim] (rounded to half) against the half rows
    D_packed [sum(D_lengths), dim]: per doc the sum over query tokens of the
    max over its rows of half(D . Q_t)."""
    torch = _torch()
    Qt = Q if torch.is_tensor(Q) else torch.as_tensor(np.asarray(Q))
    Qt = Qt.reshape(-1, Qt.shape[-1])
    q = Qt.to(device="cuda", dtype=torch.float16).contiguous()
    D = _dev(D_packed, torch.float16)
    ln = _dev(D_lengths, torch.int64)
    if q.shape[1] != 128 or (D.numel() and D.shape[-1] != 128):
        raise _c.MvrError(4, "colbert_score_packed: dim must be 128")
    out = torch.empty(ln.numel(), dtype=torch.float16, device="cuda")
    _c.check_plaid(_lib().mvr_plaid_colbert_score_packed_f16(_ptr(q), q.shape[0], _ptr(D),
                                                             D.shape[0] if D.dim() == 2 else 0, _ptr(ln),
                                                             ln.numel(), _ptr(out), _stream(torch)))
    return _back(out, D_packed if torch.is_tensor(D_packed) else D)


def colbert_score_reduce(scores_padded, D_mask):
    """float16 [n]: colbert_score_reduce of the GPU branch on its padded
    input: scores_padded [n, L, nq] (half), D_mask [n, L] (or any shape with
    n * L entries); masked-out entries read as half(-9999) = -10000."""
    torch = _torch()
    sp = _dev(scores_padded, torch.float16)
    n, L, nq = sp.shape
    m = _dev(D_mask, torch.bool).reshape(n, L).contiguous().view(torch.uint8)
    out = torch.empty(n, dtype=torch.float16, device="cuda")
    _c.check_plaid(_lib().mvr_plaid_colbert_score_reduce_f16(_ptr(sp), _ptr(m), n, L, nq, _ptr(out),
                                                             _stream(torch)))
    return _back(out, scores_padded if torch.is_tensor(scores_padded) else sp)


# ------------------------------------------------------- batched search
def searcher_defaults(k: int):
    """(ncells, centroid_score_threshold, ndocs) by k (searcher.py:291-312)."""
    if k <= 10:
        return 1, 0.5, 256
    if k <= 100:
        return 2, 0.45, 1024
    return 4, 0.4, max(k * 4, 4096)


class PlaidIndex:
    """Device-resident PLAID index (the arrays IndexScorer and ResidualCodec
    load: doclens, codes, residuals, centroids widened to float32, bucket
    weights, the pid IVF of ivf.pid.pt)."""

    def __init__(self, doclens, codes, residuals, centroids, bucket_weights, ivf, ivf_lengths, nbits: int = 2,
                 device: int = 0):
        def a(x, dt):
            if hasattr(x, "detach"):
                x = x.detach().cpu().numpy()
            return np.ascontiguousarray(np.asarray(x), dtype=dt)

        self._keep = [a(doclens, np.int64), a(codes, np.int32), a(residuals, np.uint8),
                      a(centroids, np.float32), a(bucket_weights, np.float32), a(ivf, np.int32),
                      a(ivf_lengths, np.int64)]
        dl, cd, rs, ce, bw, iv, il = self._keep
        self.n_doc, self.n_centroid, self.dim = dl.size, ce.shape[0], ce.shape[1]
        desc = _c.PlaidIndexDesc(n_doc=dl.size, n_centroid=ce.shape[0], dim=ce.shape[1], nbits=nbits,
                                 n_tok=cd.size, n_ivf=iv.size, doclens=dl.ctypes.data, codes=cd.ctypes.data,
                                 residuals=rs.ctypes.data, centroids=ce.ctypes.data, bucket_weights=bw.ctypes.data,
                                 ivf=iv.ctypes.data, ivf_lengths=il.ctypes.data)
        h = C.c_void_p()
        _c.check_plaid(_lib().mvr_plaid_index_create(C.byref(desc), device, C.byref(h)))
        self._h = h
        self._doclens_host = dl
        self._keep = None

    def close(self):
        if getattr(self, "_h", None):
            _lib().mvr_plaid_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lookup_pids(self, passage_ids):
        """(D_packed float16 [sum of doclens, dim], D_lengths int64): the GPU
        branch's lookup_pids, normalised half rows of the pids' tokens in
        order, on the current GPU."""
        torch = _torch()
        p = _dev(passage_ids, torch.int32).reshape(-1)
        ph = p.cpu().numpy().astype(np.int64)
        if ph.size and (ph.min() < 0 or ph.max() >= self.n_doc):
            raise _c.MvrError(1, f"lookup_pids: a pid lies outside [0, {self.n_doc})")
        ln = torch.as_tensor(self._doclens_host[ph] if ph.size else np.zeros(0, np.int64), device="cuda")
        rows = int(ln.sum().item()) if p.numel() else 0
        out = torch.empty((rows, self.dim), dtype=torch.float16, device="cuda")
        got = C.c_uint64(0)
        _c.check_plaid(_lib().mvr_plaid_lookup_pids_f16(self._h, _ptr(p), p.numel(), _ptr(out), rows,
                                                        C.byref(got), _stream(torch)))
        return out, ln

    def search(self, Q, k: int = 10, ncells=None, centroid_score_threshold=None, ndocs=None, batch: int = 1024,
               return_stats: bool = False, timings: bool = False, branch: str = "cpu"):
        """Q [n_q, nq, dim] (numpy or torch, host or device).  Returns
        (pids int32 [n_q, k], scores float32 [n_q, k]) as numpy, -1 / -inf
        where fewer than k docs survive; plus stats [n_q, 3] (candidate pids,
        filtered pids, n_vec_score_refine) and/or a timings dict.
        branch: "cpu" (the reference's CPU branch, FP32) or "gpu" (its GPU
        branch, float16; scores are half values)."""
        if branch not in ("cpu", "gpu"):
            raise ValueError(f"branch must be 'cpu' or 'gpu', not {branch!r}")
        d_nc, d_thr, d_nd = searcher_defaults(k)
        p = _c.PlaidParams(k=k, ncells=ncells or d_nc, ndocs=ndocs or d_nd, batch=batch,
                           centroid_score_threshold=d_thr if centroid_score_threshold is None
                           else centroid_score_threshold, gpu_branch=1 if branch == "gpu" else 0)
        on_dev = hasattr(Q, "is_cuda") and Q.is_cuda
        if on_dev:
            Qc = Q.float().contiguous()
            qp = C.c_void_p(Qc.data_ptr())
        else:
            Qc = np.ascontiguousarray(Q.cpu().numpy() if hasattr(Q, "cpu") else Q, dtype=np.float32)
            qp = C.c_void_p(Qc.ctypes.data)
        n_q, nq = Qc.shape[0], Qc.shape[1]
        ids = np.empty((n_q, k), np.int32)
        sc = np.empty((n_q, k), np.float32)
        stats = np.zeros((n_q, 3), np.uint64)
        t = _c.PlaidTimings()
        if on_dev:
            _torch().cuda.synchronize()
        _c.check_plaid(_lib().mvr_plaid_search(self._h, qp, n_q, nq, C.byref(p), ids.ctypes.data, sc.ctypes.data,
                                               1 if on_dev else 0, stats.ctypes.data, C.byref(t)))
        out = [ids, sc]
        if return_stats:
            out.append(stats)
        if timings:
            out.append({f: getattr(t, f) for f, _ in t._fields_})
        return tuple(out)

    def rank(self, Q, k: int = 10, **kw):
        """IndexScorer.rank for one query Q [1, nq, dim] or [nq, dim]:
        (pids list, scores list) best first."""
        Q = Q if Q.ndim == 3 else Q[None]
        ids, sc = self.search(Q, k=k, **kw)[:2]
        keep = ids[0] >= 0
        return ids[0][keep].tolist(), sc[0][keep].tolist()
