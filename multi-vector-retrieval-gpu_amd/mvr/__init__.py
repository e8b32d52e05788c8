"""mvr -- MI355X ColBERT late-interaction search (host side of libmvr).

`Index` wraps an mvr_index handle (include/mvr.h).  `IGPGPUPP` mirrors the
reference's method object (src/impl/IGPGPUPP.hpp:28-333: constructor with the
doc-length file, buildIndex from the index files, search returning per-query
top-k ids and scores) so code written against it ports directly.
`build_index_module()` returns the GPGPUBuildIndexPackGPU pybind module.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
import sysconfig
import time
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import MVR_BF16, MVR_FP8, MVR_FP32, MVR_FP32_MFMA, MvrError, check, lib

__all__ = ["Index", "IGPGPUPP", "IGPGPUPPResult", "MvrError", "MVR_FP32", "MVR_BF16", "MVR_FP8", "MVR_FP32_MFMA",
           "build_index_module", "device_count"]

_PRECISION = {"fp32": MVR_FP32, "bf16": MVR_BF16, "fp8": MVR_FP8, "fp32_mfma": MVR_FP32_MFMA}


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().mvr_device_count(C.byref(n)))
    return n.value


def _ivf_from_codes(codes, doclens, n_c):
    """IVF lists (docs ascending) from codes, src/include/struct/IVFIndex.hpp:92-166."""
    codes = _c(codes, np.uint32)
    doclens = _c(doclens, np.uint32)
    doc_of_tok = np.repeat(np.arange(doclens.size, dtype=np.uint64), doclens)
    key = np.unique(codes.astype(np.uint64) * np.uint64(doclens.size) + doc_of_tok)
    cent = (key // np.uint64(doclens.size)).astype(np.uint32)
    ids = (key % np.uint64(doclens.size)).astype(np.uint32)
    sizes = np.bincount(cent, minlength=n_c).astype(np.uint32)
    off = np.zeros(n_c, np.uint64)
    off[1:] = np.cumsum(sizes[:-1], dtype=np.uint64)
    return off, sizes, ids


def _n_bit(n_weight: int) -> int:
    """Bits per residual value of a decode table with n_weight entries per dim
    (weight_l is [2^n_bit] or [dim, 2^n_bit])."""
    for b in (1, 2, 4, 8):
        if n_weight == 1 << b:
            return b
    raise MvrError(1, f"weights hold {n_weight} values per dim (expected 2, 4, 16 or 256)")


def _flags(residual_on_host: bool) -> int:
    return _capi.INDEX_RESIDUAL_ON_HOST if residual_on_host else 0


class Index:
    """IGP index on one device (centroids, codes, residuals, IVF); with
    residual_on_host the codes and residuals stay in host memory."""

    def __init__(self, handle: C.c_void_p, device: int = 0):
        self._h = handle
        self.device = device
        nd, nt, nc, ml, md = C.c_uint32(), C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().mvr_index_info(self._h, C.byref(nd), C.byref(nt), C.byref(nc), C.byref(ml), C.byref(md)))
        self.n_doc, self.n_tok, self.n_centroid = nd.value, nt.value, nc.value
        self.max_list, self.max_doclen = ml.value, md.value

    # ---------------------------------------------------------------- create
    @classmethod
    def from_arrays(cls, doclens, centroids, codes, residual, weights, ivf=None, device: int = 0,
                    residual_on_host: bool = False):
        """residual_on_host keeps codes and residual bytes in (pinned) host
        memory; each device pass then stages its selected docs."""
        doclens = _c(doclens, np.uint32)
        centroids = _c(centroids, np.float32)
        codes = _c(codes, np.uint32)
        residual = _c(residual, np.uint8)
        weights = _c(weights, np.float32)
        if ivf is None:
            ivf = _ivf_from_codes(codes, doclens, centroids.shape[0])
        off, size, ids = (_c(ivf[0], np.uint64), _c(ivf[1], np.uint32), _c(ivf[2], np.uint32))
        d = _capi.IndexDesc()
        d.n_doc, d.n_tok = doclens.size, codes.size
        d.n_centroid, d.dim = centroids.shape
        d.n_bit = _n_bit(weights.shape[-1])
        d.doclens, d.centroids, d.codes = _ptr(doclens), _ptr(centroids), _ptr(codes)
        d.residual, d.weights = _ptr(residual), _ptr(weights)
        d.weights_per_dim = 1 if weights.ndim == 2 else 0
        d.ivf_offsets, d.ivf_sizes, d.ivf_ids = _ptr(off), _ptr(size), _ptr(ids)
        d.ivf_n_ele = ids.size
        d.on_device = 0
        h = C.c_void_p()
        check(lib().mvr_index_create_ex(C.byref(d), device, _flags(residual_on_host), C.byref(h)))
        return cls(h, device).set_doclens(doclens)

    @classmethod
    def from_device(cls, doclens, centroids, codes, residual, weights, ivf, device: int = 0,
                    residual_on_host: bool = False):
        """Arguments are torch tensors on `device` (copied by the library)."""
        d = _capi.IndexDesc()
        d.n_doc, d.n_tok = doclens.numel(), codes.numel()
        d.n_centroid, d.dim = centroids.shape
        d.n_bit = _n_bit(weights.shape[-1])
        d.doclens, d.centroids = doclens.data_ptr(), centroids.data_ptr()
        d.codes, d.residual, d.weights = codes.data_ptr(), residual.data_ptr(), weights.data_ptr()
        d.weights_per_dim = 1 if weights.dim() == 2 else 0
        d.ivf_offsets, d.ivf_sizes, d.ivf_ids = ivf[0].data_ptr(), ivf[1].data_ptr(), ivf[2].data_ptr()
        d.ivf_n_ele = ivf[2].numel()
        d.on_device = 1
        h = C.c_void_p()
        check(lib().mvr_index_create_ex(C.byref(d), device, _flags(residual_on_host), C.byref(h)))
        return cls(h, device).set_doclens(doclens.cpu().numpy())

    @classmethod
    def load(cls, index_dir: str, device: int = 0, residual_on_host: bool = False):
        h = C.c_void_p()
        check(lib().mvr_index_load_ex(index_dir.encode(), device, _flags(residual_on_host), C.byref(h)))
        return cls(h, device).set_doclens(np.load(os.path.join(index_dir, "item_n_vec_l.npy")))

    def close(self):
        if self._h:
            lib().mvr_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- search
    @staticmethod
    def _params(nprobe, probe_topk, topk, precision, rescore, clip_zero, batch, learned=False,
                persistent=False, graph=False):
        p = _capi.SearchParams()
        p.nprobe, p.probe_topk, p.topk = nprobe, probe_topk, topk
        p.precision = _PRECISION.get(precision, precision) if isinstance(precision, str) else precision
        p.rescore = rescore
        p.clip_zero = int(clip_zero)
        p.batch = batch
        p.learned = int(learned)
        p.persistent = int(persistent)
        p.graph = int(graph)
        return p

    def set_tuning(self, precision, wgs_per_cu, tiles):
        """Fix the work-queue MaxSim settings of a precision."""
        prec = _PRECISION.get(precision, precision) if isinstance(precision, str) else precision
        check(lib().mvr_index_set_tuning(self._h, prec, wgs_per_cu, tiles))

    def set_maxsim_order(self, order: str):
        """BF16 MaxSim loop order of search(): "auto", "query" or "doc" (same
        scores in every order)."""
        check(lib().mvr_index_set_maxsim_order(self._h, {"auto": 0, "query": 1, "doc": 2}[order]))

    def get_tuning(self, precision):
        """(workgroups per CU, tiles in flight) or None before tuning."""
        prec = _PRECISION.get(precision, precision) if isinstance(precision, str) else precision
        out = (C.c_uint32 * 2)()
        if lib().mvr_index_get_tuning(self._h, prec, out) != 0:
            return None
        return tuple(int(v) for v in out)

    def get_tuning_ex(self, precision):
        """{"order": "query"|"doc", "wgs_per_cu": w, "tiles": t} (the first
        phase's loop order the autotuner chose, and the per-query kernel's
        setting) or None before tuning."""
        prec = _PRECISION.get(precision, precision) if isinstance(precision, str) else precision
        out = (C.c_uint32 * 3)()
        if lib().mvr_index_get_tuning_ex(self._h, prec, out) != 0:
            return None
        return {"order": {1: "query", 2: "doc"}[int(out[0])], "wgs_per_cu": int(out[1]), "tiles": int(out[2])}

    def search(self, Q, nprobe: int, probe_topk: int, topk: int, precision="fp32", rescore: int = 0,
               q_ntok=None, q_nprobe=None, clip_zero=False, batch: int = 0, stages=False,
               timings=False, learned=False, persistent=False, graph=False):
        """Host-buffer search.  Q: [n_q, t_max, 128] float32.  Returns
        (ids [n_q, topk] uint32, scores [n_q, topk] float32[, stages][, timings])."""
        Q = _c(Q, np.float32)
        n_q, t_max, _ = Q.shape
        ids = np.empty((n_q, topk), np.uint32)
        sc = np.empty((n_q, topk), np.float32)
        qt = None if q_ntok is None else _c(q_ntok, np.uint32)
        qp = None if q_nprobe is None else _c(q_nprobe, np.uint32)
        prm = self._params(nprobe, probe_topk, topk, precision, rescore, clip_zero, batch, learned,
                           persistent, graph)
        st = None
        so = None
        if stages:
            pmax = 32 if learned else max(nprobe, int(qp.max()) if qp is not None else 0)
            pmax = min(pmax, self.n_centroid)
            st = {"top_ids": np.empty((n_q, t_max, pmax), np.uint32),
                  "top_s": np.empty((n_q, t_max, pmax), np.float32),
                  "cand_n": np.empty(n_q, np.uint32),
                  "sel_ids": np.empty((n_q, probe_topk), np.uint32),
                  "sel_F": np.empty((n_q, probe_topk), np.float32),
                  "sel_score": np.empty((n_q, probe_topk), np.float32),
                  "nprobe": np.empty(n_q, np.uint32),
                  "rerank_n": np.empty(n_q, np.uint32)}
            so = _capi.StageOut(*[st[k].ctypes.data for k in
                                  ("top_ids", "top_s", "cand_n", "sel_ids", "sel_F", "sel_score", "nprobe",
                                   "rerank_n")])
        tm = _capi.Timings() if timings else None
        check(lib().mvr_search(self._h, _ptr(Q), n_q, t_max, _ptr(qt), _ptr(qp), C.byref(prm),
                               _ptr(ids), _ptr(sc), 0, None, C.byref(so) if so else None,
                               C.byref(tm) if tm else None))
        out = [ids, sc]
        if stages:
            out.append(st)
        if timings:
            out.append({k: getattr(tm, k) for k, _ in _capi.Timings._fields_})
        return tuple(out)

    def search_device(self, Q, out_ids, out_scores, nprobe, probe_topk, topk, precision="fp32",
                      rescore=0, q_ntok=None, q_nprobe=None, clip_zero=False, batch=0, stream=0,
                      learned=False, persistent=False, graph=False):
        """Device-buffer search (torch tensors); queued on `stream`, no host sync."""
        n_q, t_max, _ = Q.shape
        qt = None if q_ntok is None else _c(q_ntok, np.uint32)
        qp = None if q_nprobe is None else _c(q_nprobe, np.uint32)
        prm = self._params(nprobe, probe_topk, topk, precision, rescore, clip_zero, batch, learned,
                           persistent, graph)
        check(lib().mvr_search(self._h, Q.data_ptr(), n_q, t_max, _ptr(qt), _ptr(qp), C.byref(prm),
                               out_ids.data_ptr(), out_scores.data_ptr(), 1, stream or None, None, None))

    def decompress(self, doc_ids):
        """FP32 embeddings of the given docs, stacked [sum(L), 128]."""
        doc_ids = _c(doc_ids, np.uint32)
        raise_if = doc_ids.size and int(doc_ids.max()) >= self.n_doc
        if raise_if:
            raise MvrError(1, "doc id out of range")
        lens = self._doclens()[doc_ids]
        out = np.empty((int(lens.sum()), 128), np.float32)
        check(lib().mvr_decompress(self._h, _ptr(doc_ids), doc_ids.size, _ptr(out)))
        return out

    def _doclens(self):
        if not hasattr(self, "_dl"):
            raise MvrError(1, "doc lengths unknown for this handle; pass them to set_doclens()")
        return self._dl

    def set_doclens(self, doclens):
        self._dl = _c(doclens, np.uint32)
        return self


@dataclass
class IGPGPUPPResult:
    """src/impl/IGPGPUPP.hpp:19-26."""
    result_score_l: np.ndarray
    result_ID_l: np.ndarray
    compute_time_l: np.ndarray
    batch_query_time: float
    n_query: int
    topk: int


class IGPGPUPP:
    """Method object with the reference's call sequence:
    IGPGPUPP(item_n_vec_l_fname, n_thread) -> buildIndex(...) -> search(...)."""

    def __init__(self, item_n_vec_l_fname: str, n_cpu_thread: int = 1, device: int = 0,
                 residual_on_host: bool = False):
        self._item_n_vec_l = np.load(item_n_vec_l_fname).astype(np.uint32)
        self.n_cpu_thread = n_cpu_thread
        self.device = device
        self.residual_on_host = residual_on_host
        self.index: Index | None = None

    def buildIndex(self, centroid_l_fname, vq_code_l_fname, ivf_index_fname, weight_l_fname,
                   residual_code_l_fname, n_centroid: int, n_bit: int) -> bool:
        dirs = {os.path.dirname(os.path.abspath(f)) for f in
                (centroid_l_fname, vq_code_l_fname, ivf_index_fname, weight_l_fname, residual_code_l_fname)}
        if len(dirs) != 1:
            raise MvrError(1, "index files must share one directory")
        self.index = Index.load(dirs.pop(), self.device, self.residual_on_host).set_doclens(self._item_n_vec_l)
        ix_bit = _n_bit(np.load(weight_l_fname, mmap_mode="r").shape[-1])
        if self.index.n_centroid != n_centroid or n_bit != ix_bit:
            raise MvrError(1, f"index has {self.index.n_centroid} centroids/{ix_bit} bits, "
                              f"asked {n_centroid}/{n_bit}")
        return True

    def search(self, query_l, n_query: int, query_n_vec: int, topk: int, nprobe: int,
               probe_topk: int, precision="fp32", rescore: int = 0, batch: int = 0) -> IGPGPUPPResult:
        if probe_topk < topk:
            raise MvrError(1, "the number of refined topk is smaller than the number of returned topk")
        Q = np.asarray(query_l, np.float32).reshape(n_query, query_n_vec, -1)
        t0 = time.perf_counter()
        ids, sc = self.index.search(Q, nprobe, probe_topk, topk, precision, rescore, batch=batch)
        dt = time.perf_counter() - t0
        return IGPGPUPPResult(sc.ravel(), ids.ravel(), np.full(n_query, dt), dt, n_query, topk)


def bruteforce_module():
    """Import the BruteForceProgressive pybind module (exhaustive MaxSim,
    DocRetrieval) from the built lib/."""
    path = os.path.join(_capi.LIB_DIR, "BruteForceProgressive" + sysconfig.get_config_var("EXT_SUFFIX"))
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built")
    spec = importlib.util.spec_from_file_location("BruteForceProgressive", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def build_index_module():
    """Import the GPGPUBuildIndexPackGPU pybind module from the built lib/."""
    path = os.path.join(_capi.LIB_DIR, "GPGPUBuildIndexPackGPU" + sysconfig.get_config_var("EXT_SUFFIX"))
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built")
    spec = importlib.util.spec_from_file_location("GPGPUBuildIndexPackGPU", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
