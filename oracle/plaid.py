"""ctypes front end of the PLAID part of the CPU oracle (oracle/plaid_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.  The C source cites
the reference file:line each function restates (ColBERT's PLAID side tree:
filter_pids.cpp, search/decompress_residuals.cpp, segmented_maxsim.cpp,
codecs/decompress_residuals.cu, codecs/packbits.cu, residual.py,
candidate_generation.py, index_storage.py).
"""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np

from . import oracle as _O

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")

_sig_done = False


def lib():
    global _sig_done
    L = _O.lib()
    if _sig_done:
        return L
    u32, u64 = C.c_uint32, C.c_uint64
    sig = {
        "orc_f32_to_f16_array": (None, [_f32p, u64, _u16p]),
        "orc_f16_to_f32_array": (None, [_u16p, u64, _f32p]),
        "orc_plaid_decompress_f16": (None, [_u8p, u64, _u16p, _u8p, _u8p, _i32p, _u16p, u32, u32, _u16p]),
        "orc_plaid_packbits": (None, [_u8p, u64, _u8p]),
        "orc_plaid_binarize": (None, [_f32p, u64, u32, _f32p, u32, u32, _u8p]),
        "orc_plaid_decompress_pids": (u64, [_i32p, u32, _i64p, _i64p, _f32p, _u8p, _u8p, _u8p, _i32p,
                                            _f32p, u32, u32, _f32p]),
        "orc_plaid_segmented_maxsim": (None, [_f32p, _i64p, u32, u32, _f32p]),
        "orc_plaid_filter_pids": (u32, [_i32p, u32, _f32p, u32, u32, _i32p, _i64p, _i64p, _u8p, u32,
                                        _i32p, C.c_void_p, C.POINTER(C.c_uint64)]),
        "orc_plaid_centroid_scores": (None, [_f32p, u32, _f32p, u32, u32, _f32p]),
        "orc_plaid_score_pids": (None, [_f32p, u32, _i32p, u32, _i64p, _i64p, _i32p, _u8p, _f32p, _f32p,
                                        _u8p, _u8p, u32, u32, _f32p]),
        "orc_plaid_search_batch": (None, [u32, u32, u32, u32, _i64p, _i64p, _i32p, _u8p, _f32p, _f32p,
                                          _u8p, _u8p, _i32p, _i64p, _i64p, _f32p, u32, u32, u32,
                                          C.c_float, u32, u32, C.c_int, _i32p, _f32p, C.c_void_p]),
        "orc_plaid_search_batch_gpu": (None, [u32, u32, u32, u32, _i64p, _i64p, _i32p, _u8p, _f32p, _f32p,
                                              _u8p, _u8p, _i32p, _i64p, _i64p, _f32p, u32, u32, u32,
                                              C.c_float, u32, u32, C.c_int, _i32p, _f32p, C.c_void_p]),
        "orc_plaid_lookup_pids_f16": (u64, [_i32p, u32, _i64p, _i64p, _i32p, _u8p, _f32p, _f32p, _u8p, _u8p,
                                            u32, u32, _f32p]),
        "orc_plaid_score_reduce_f16": (None, [_f32p, _i64p, u32, u32, _f32p]),
        "orc_plaid_score_packed_f16": (None, [_f32p, u32, _f32p, _i64p, u32, u32, _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _sig_done = True
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------- codec tables
def reversed_bit_map(nbits: int = 2) -> np.ndarray:
    """residual.py:51-72: reverse the bit order inside every nbits-wide group
    of a byte (written here as a closed form over the groups)."""
    out = np.zeros(256, np.uint8)
    groups = 8 // nbits
    for i in range(256):
        z = 0
        for g in range(groups):
            v = (i >> (g * nbits)) & ((1 << nbits) - 1)
            r = int(format(v, f"0{nbits}b")[::-1], 2)
            z |= r << (g * nbits)
        out[i] = z
    return out


def decompression_lookup_table(nbits: int = 2) -> np.ndarray:
    """residual.py:76-91: itertools.product(range(2^nbits), repeat=8/nbits)."""
    return np.array(list(itertools.product(range(1 << nbits), repeat=8 // nbits)), np.uint8)


# ------------------------------------------------------------------ halves
def f16_round(x) -> np.ndarray:
    """float32 -> binary16 bits (RNE) by the oracle's C conversion."""
    x = _c(x, np.float32)
    out = np.empty(x.shape, np.uint16)
    lib().orc_f32_to_f16_array(x.reshape(-1), x.size, out.reshape(-1))
    return out


def f16_widen(h) -> np.ndarray:
    h = _c(h, np.uint16)
    out = np.empty(h.shape, np.float32)
    lib().orc_f16_to_f32_array(h.reshape(-1), h.size, out.reshape(-1))
    return out


# ------------------------------------------------------------------- ops
def decompress_f16(binary_residuals, bucket_weights_f16, rbm, lut, codes, centroids_f16,
                   dim: int = 128, nbits: int = 2) -> np.ndarray:
    """decompress_residuals_cuda (decompress_residuals.cu:42-75); halves as uint16 bits."""
    res = _c(binary_residuals, np.uint8)
    n = res.shape[0]
    out = np.empty((n, dim), np.uint16)
    lib().orc_plaid_decompress_f16(res.reshape(-1), n, _c(bucket_weights_f16, np.uint16), _c(rbm, np.uint8),
                                   _c(lut, np.uint8).reshape(-1), _c(codes, np.int32),
                                   _c(centroids_f16, np.uint16).reshape(-1), dim, nbits, out)
    return out


def packbits(bits) -> np.ndarray:
    bits = _c(bits, np.uint8).reshape(-1)
    out = np.empty(bits.size // 8, np.uint8)
    lib().orc_plaid_packbits(bits, bits.size, out)
    return out


def binarize(e, cutoffs, nbits: int = 2) -> np.ndarray:
    e = _c(e, np.float32)
    n, dim = e.shape
    out = np.empty((n, dim * nbits // 8), np.uint8)
    cut = _c(cutoffs, np.float32)
    lib().orc_plaid_binarize(e.reshape(-1), n, dim, cut, cut.size, nbits, out)
    return out


def decompress_pids(pids, lengths, offsets, bucket_weights, rbm, lut, residuals, codes, centroids,
                    dim: int = 128, nbits: int = 2) -> np.ndarray:
    """search/decompress_residuals.cpp:87-155 (CPU signature)."""
    pids = _c(pids, np.int32)
    lengths = _c(lengths, np.int64)
    n = int(lengths[pids.astype(np.int64)].sum()) if pids.size else 0
    out = np.empty((n, dim), np.float32)
    got = lib().orc_plaid_decompress_pids(pids, pids.size, lengths, _c(offsets, np.int64),
                                          _c(bucket_weights, np.float32), _c(rbm, np.uint8),
                                          _c(lut, np.uint8).reshape(-1), _c(residuals, np.uint8).reshape(-1),
                                          _c(codes, np.int32), _c(centroids, np.float32).reshape(-1),
                                          dim, nbits, out.reshape(-1) if n else np.empty(1, np.float32))
    assert got == n
    return out


def segmented_maxsim(scores, lengths) -> np.ndarray:
    scores = _c(scores, np.float32)
    lengths = _c(lengths, np.int64)
    out = np.empty(lengths.size, np.float32)
    nq = scores.shape[1] if scores.ndim == 2 else 0
    lib().orc_plaid_segmented_maxsim(scores.reshape(-1) if scores.size else np.empty(1, np.float32),
                                     lengths, lengths.size, nq, out)
    return out


def segmented_lookup(input, pids, lengths, offsets) -> np.ndarray:
    """segmented_lookup.cpp:52-108 (the lookup workers, :25-49): for
    i < len(pids), the rows input[offsets[i] : offsets[i] + lengths[i]],
    concatenated in order (the output starts zero-filled and every row is
    written, so only the copies matter); 1-D input is rows of one element."""
    x = np.asarray(input)
    n = len(pids)
    ln = np.asarray(lengths, np.int64)[:n]
    of = np.asarray(offsets, np.int64)[:n]
    parts = [x[int(o):int(o) + int(l)] for o, l in zip(of, ln)]
    if not parts:
        return np.zeros((0,) + x.shape[1:], x.dtype)
    return np.concatenate(parts, axis=0)


def filter_pids(pids, centroid_scores, codes, doclens, offsets, idx, ndocs: int):
    """filter_pids.cpp:147-193.  Returns (pids int32 [min(ndocs//4, ...)],
    stage-2 scores, n_vec_score_refine)."""
    pids = _c(pids, np.int32)
    S = _c(centroid_scores, np.float32)
    n_c, nq = S.shape
    out = np.empty(max(ndocs // 4, 1), np.int32)
    out_s = np.empty(max(ndocs // 4, 1), np.float32)
    nv = C.c_uint64(0)
    n = lib().orc_plaid_filter_pids(pids if pids.size else np.empty(1, np.int32), pids.size, S.reshape(-1),
                                    n_c, nq, _c(codes, np.int32), _c(doclens, np.int64), _c(offsets, np.int64),
                                    _c(idx, np.uint8), ndocs, out, out_s.ctypes.data_as(C.c_void_p),
                                    C.byref(nv))
    return out[:n].copy(), out_s[:n].copy(), int(nv.value)


def centroid_scores(centroids, Q) -> np.ndarray:
    """centroids @ Q.T as [n_c][nq] (candidate_generation.py:14)."""
    Cf = _c(centroids, np.float32)
    Qf = _c(Q, np.float32)
    S = np.empty((Cf.shape[0], Qf.shape[0]), np.float32)
    lib().orc_plaid_centroid_scores(Cf.reshape(-1), Cf.shape[0], Qf.reshape(-1), Qf.shape[0], Cf.shape[1],
                                    S.reshape(-1))
    return S


class PlaidIndex:
    """A PLAID index in host memory: the arrays IndexScorer loads
    (index_storage.py, residual.py:133-150), centroids and bucket weights
    widened to float32 as the CPU branch does (residual.py:27-45)."""

    def __init__(self, doclens, codes, residuals, centroids, bucket_weights, ivf, ivf_lengths,
                 nbits: int = 2):
        self.doclens = _c(doclens, np.int64)
        self.offsets = np.zeros(self.doclens.size, np.int64)
        if self.doclens.size > 1:
            self.offsets[1:] = np.cumsum(self.doclens[:-1])
        self.codes = _c(codes, np.int32)
        self.residuals = _c(residuals, np.uint8)
        self.centroids = _c(centroids, np.float32)
        self.bucket_weights = _c(bucket_weights, np.float32)
        self.ivf = _c(ivf, np.int32)
        self.ivf_lengths = _c(ivf_lengths, np.int64)
        self.ivf_off = np.zeros(self.ivf_lengths.size, np.int64)
        if self.ivf_lengths.size > 1:
            self.ivf_off[1:] = np.cumsum(self.ivf_lengths[:-1])
        self.nbits = nbits
        self.dim = self.centroids.shape[1]
        self.rbm = reversed_bit_map(nbits)
        self.lut = decompression_lookup_table(nbits)

    @property
    def n_doc(self):
        return self.doclens.size

    @property
    def n_c(self):
        return self.centroids.shape[0]

    def score_pids(self, Q, pids) -> np.ndarray:
        """The FP32 refine of index_storage.py:164-190 for one query."""
        Q = _c(Q, np.float32)
        pids = _c(pids, np.int32)
        out = np.empty(max(pids.size, 1), np.float32)
        lib().orc_plaid_score_pids(Q.reshape(-1), Q.shape[0], pids if pids.size else np.empty(1, np.int32),
                                   pids.size, self.doclens, self.offsets, self.codes, self.residuals.reshape(-1),
                                   self.centroids.reshape(-1), self.bucket_weights, self.rbm,
                                   self.lut.reshape(-1), self.dim, self.nbits, out)
        return out[:pids.size]

    def lookup_pids_f16(self, pids):
        """lookup_pids of the GPU branch (residual_embeddings_strided.py:30-37,
        residual.py:243-279): (normalised float16 rows of the pids' tokens,
        their doclens)."""
        pids = _c(pids, np.int32)
        ln = self.doclens[pids.astype(np.int64)] if pids.size else np.zeros(0, np.int64)
        rows = int(ln.sum())
        out = np.empty((max(rows, 1), self.dim), np.float32)
        got = lib().orc_plaid_lookup_pids_f16(pids if pids.size else np.empty(1, np.int32), pids.size,
                                              self.doclens, self.offsets, self.codes,
                                              self.residuals.reshape(-1), self.centroids.reshape(-1),
                                              self.bucket_weights, self.rbm, self.lut.reshape(-1), self.dim,
                                              self.nbits, out.reshape(-1))
        assert got == rows
        return out[:rows].astype(np.float16), ln.astype(np.int64)

    def search(self, Q, k: int, ncells: int, threshold: float, ndocs: int, n_threads: int = 0,
               branch: str = "cpu"):
        """IndexScorer.rank for a batch Q [n_q][nq][dim], CPU branch (FP32)
        or GPU branch (float16, branch="gpu").  Returns (pids int32 [n_q][k],
        scores [n_q][k], stats uint64 [n_q][3])."""
        Q = _c(Q, np.float32)
        n_q, nq, dim = Q.shape
        ids = np.empty((n_q, k), np.int32)
        sc = np.empty((n_q, k), np.float32)
        stats = np.zeros((n_q, 3), np.uint64)
        fn = {"cpu": lib().orc_plaid_search_batch, "gpu": lib().orc_plaid_search_batch_gpu}[branch]
        fn(self.n_doc, self.n_c, dim, self.nbits, self.doclens, self.offsets,
                                     self.codes, self.residuals.reshape(-1), self.centroids.reshape(-1),
                                     self.bucket_weights, self.rbm, self.lut.reshape(-1), self.ivf,
                                     self.ivf_off, self.ivf_lengths, Q.reshape(-1), n_q, nq, ncells,
                                     float(threshold), ndocs, k, n_threads, ids, sc,
                                     stats.ctypes.data_as(C.c_void_p))
        return ids, sc, stats


def score_reduce_f16(scores, lengths) -> np.ndarray:
    """colbert_score_reduce of the GPU branch (colbert.py:133-155) over a
    packed half score matrix [sum(lengths)][nq]: float16 [ndocs]."""
    s = _c(np.asarray(scores, np.float16).astype(np.float32), np.float32)
    ln = _c(lengths, np.int64)
    nq = s.shape[1] if s.ndim == 2 else 0
    out = np.empty(max(ln.size, 1), np.float32)
    lib().orc_plaid_score_reduce_f16(s.reshape(-1) if s.size else np.empty(1, np.float32), ln, ln.size, nq, out)
    return out[:ln.size].astype(np.float16)


def score_packed_f16(Q, D_packed, lengths) -> np.ndarray:
    """colbert_score_packed of the GPU branch (colbert.py:181-205), one query
    Q [nq][dim] against half rows D_packed [sum(lengths)][dim]: float16 [ndocs]."""
    Qf = _c(Q, np.float32)
    D = _c(np.asarray(D_packed, np.float16).astype(np.float32), np.float32)
    ln = _c(lengths, np.int64)
    out = np.empty(max(ln.size, 1), np.float32)
    lib().orc_plaid_score_packed_f16(Qf.reshape(-1), Qf.shape[0], D.reshape(-1) if D.size else np.empty(1, np.float32),
                                     ln, ln.size, Qf.shape[1], out)
    return out[:ln.size].astype(np.float16)


def searcher_defaults(k: int):
    """(ncells, centroid_score_threshold, ndocs) by k (searcher.py:291-312)."""
    if k <= 10:
        return 1, 0.5, 256
    if k <= 100:
        return 2, 0.45, 1024
    return 4, 0.4, max(k * 4, 4096)
