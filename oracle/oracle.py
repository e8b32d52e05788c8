"""ctypes front end of the CPU oracle (oracle/mvr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.  Every function
restates a piece of the reference's IGP-PP search path or index format; the C
source cites the reference file:line of each one.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmvr_oracle.so")
_lib = None

STREAM_CENTER, STREAM_DOCLEN, STREAM_TOKCOMP, STREAM_TOKNOISE = 1, 2, 3, 4
STREAM_QCOMP, STREAM_QNOISE, STREAM_QLEN, STREAM_QPROBE = 5, 6, 7, 8

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (fast; no-op when up to date)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    sig = {
        "orc_rand_u64": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_normal": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_gen_centers": (None, [C.c_uint64, C.c_uint32, C.c_uint32, _f32p]),
        "orc_gen_doclens_uniform": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _u32p]),
        "orc_gen_doclens_lognormal": (None, [C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                                             C.c_uint32, C.c_uint32, _u32p]),
        "orc_gen_tokens": (None, [C.c_uint64, C.c_uint64, C.c_uint64, _f32p, C.c_uint32, C.c_uint32,
                                  C.c_double, C.c_uint64, C.c_uint64, _f32p]),
        "orc_dot": (C.c_float, [_f32p, _f32p, C.c_uint32]),
        "orc_assign_ip": (None, [_f32p, C.c_uint64, C.c_uint32, _f32p, C.c_uint32, _u32p]),
        "orc_assign_bias": (None, [_f32p, C.c_uint64, C.c_uint32, _f32p, C.c_uint32, _f32p, _u32p]),
        "orc_kmeans": (None, [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]),
        "orc_residuals": (None, [_f32p, _u32p, C.c_uint64, C.c_uint32, _f32p, _f32p]),
        "orc_quantize_pooled": (None, [_f32p, C.c_uint64, C.c_uint32, _f32p, _f32p]),
        "orc_quantize_perdim": (None, [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, _f32p, _f32p]),
        "orc_bucket_of": (C.c_uint32, [_f32p, C.c_uint32, C.c_float]),
        "orc_compress": (None, [_f32p, _u32p, C.c_uint64, C.c_uint32, _f32p, _f32p, C.c_uint32,
                                C.c_uint32, _u8p]),
        "orc_decode_table": (None, [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]),
        "orc_decompress_tokens": (None, [_u64p, C.c_uint64, _u32p, _u8p, _f32p, _f32p, C.c_uint32,
                                         C.c_uint32, _f32p]),
        "orc_ivf_build": (C.c_uint64, [_u32p, _u32p, C.c_uint32, C.c_uint32, _u64p, _u32p, C.c_void_p]),
        "orc_centroid_topk": (None, [_f32p, C.c_uint64, _f32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                     _u32p, _f32p]),
        "orc_filter_all": (C.c_uint32, [_u32p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, _u64p, _u32p,
                                        _u32p, C.c_uint32, C.c_int, _u32p, _f32p]),
        "orc_select_top": (C.c_uint32, [_u32p, _f32p, C.c_uint32, C.c_uint32, _u32p, _f32p]),
        "orc_rerank": (None, [_f32p, C.c_uint32, C.c_uint32, _u32p, C.c_uint32, _u32p, _u64p, _u32p,
                              _u8p, _f32p, _f32p, C.c_uint32, C.c_int, _f32p]),
        "orc_search_batch": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                    _u32p, _u64p, _u32p, _u8p, _f32p, _f32p, _u64p, _u32p, _u32p,
                                    _f32p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                    C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                    _u32p, _f32p, C.c_void_p]),
        "orc_exhaustive_f64": (None, [_f32p, C.c_uint32, C.c_uint32, C.c_void_p, _f32p, _u32p, _u64p,
                                      C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _u32p, _f64p]),
        "orc_maxsim_scores_f32": (None, [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _f32p, _u32p, C.c_uint32,
                                         C.c_int, _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr_or_null(a, dt):
    if a is None:
        return None, None
    a = _c(a, dt)
    return a, a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- generator
def gen_centers(seed: int, n_comp: int, dim: int = 128) -> np.ndarray:
    out = np.empty((n_comp, dim), np.float32)
    lib().orc_gen_centers(seed, n_comp, dim, out)
    return out


def gen_doclens_uniform(seed: int, n_doc: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(n_doc, np.uint32)
    lib().orc_gen_doclens_uniform(seed, n_doc, lo, hi, out)
    return out


def gen_doclens_lognormal(seed: int, n_doc: int, median: float, sigma: float, lo: int, hi: int):
    out = np.empty(n_doc, np.uint32)
    lib().orc_gen_doclens_lognormal(seed, n_doc, median, sigma, lo, hi, out)
    return out


def gen_tokens(seed: int, centers: np.ndarray, n_tok: int, noise_std: float = 0.15,
               tok_begin: int = 0, query: bool = False) -> np.ndarray:
    centers = _c(centers, np.float32)
    n_comp, dim = centers.shape
    out = np.empty((n_tok, dim), np.float32)
    sc, sn = (STREAM_QCOMP, STREAM_QNOISE) if query else (STREAM_TOKCOMP, STREAM_TOKNOISE)
    lib().orc_gen_tokens(seed, sc, sn, centers, n_comp, dim, noise_std, tok_begin, n_tok, out)
    return out


def normal(seed: int, stream: int, i: int) -> float:
    return lib().orc_normal(seed, stream, i)


def uniform(seed: int, stream: int, i: int) -> float:
    return lib().orc_uniform(seed, stream, i)


# ------------------------------------------------------------------- codec
def kmeans(x: np.ndarray, n_c: int, n_iter: int) -> np.ndarray:
    x = _c(x, np.float32)
    out = np.empty((n_c, x.shape[1]), np.float32)
    lib().orc_kmeans(x, x.shape[0], x.shape[1], n_c, n_iter, out)
    return out


def assign_ip(x: np.ndarray, C_: np.ndarray) -> np.ndarray:
    x, C_ = _c(x, np.float32), _c(C_, np.float32)
    out = np.empty(x.shape[0], np.uint32)
    lib().orc_assign_ip(x, x.shape[0], x.shape[1], C_, C_.shape[0], out)
    return out


def assign_bias(x: np.ndarray, C_: np.ndarray, bias: np.ndarray) -> np.ndarray:
    """argmax_c (<x, C_c> - bias[c]), lowest id on ties (the Lloyd step)."""
    x, C_ = _c(x, np.float32), _c(C_, np.float32)
    out = np.empty(x.shape[0], np.uint32)
    lib().orc_assign_bias(x, x.shape[0], x.shape[1], C_, C_.shape[0], _c(bias, np.float32), out)
    return out


def residuals(x, codes, C_):
    x, C_ = _c(x, np.float32), _c(C_, np.float32)
    e = np.empty_like(x)
    lib().orc_residuals(x, _c(codes, np.uint32), x.shape[0], x.shape[1], C_, e)
    return e


def quantize_pooled(e: np.ndarray, n_bit: int):
    e = _c(e, np.float32).ravel()
    cut = np.empty((1 << n_bit) - 1, np.float32)
    w = np.empty(1 << n_bit, np.float32)
    lib().orc_quantize_pooled(e, e.size, n_bit, cut, w)
    return cut, w


def quantize_perdim(e: np.ndarray, n_bit: int):
    e = _c(e, np.float32)
    n, dim = e.shape
    cut = np.empty((dim, (1 << n_bit) - 1), np.float32)
    w = np.empty((dim, 1 << n_bit), np.float32)
    lib().orc_quantize_perdim(e, n, dim, n_bit, cut, w)
    return cut, w


def compress(x, codes, C_, cutoffs, n_bit: int) -> np.ndarray:
    """Packed residual codes [n, dim*n_bit/8]; cutoffs [2^b-1] or [dim, 2^b-1]."""
    x, C_ = _c(x, np.float32), _c(C_, np.float32)
    cutoffs = _c(cutoffs, np.float32)
    stride = 0 if cutoffs.ndim == 1 else cutoffs.shape[1]
    n, dim = x.shape
    vpb = 8 // n_bit
    out = np.empty((n, (dim + vpb - 1) // vpb), np.uint8)
    lib().orc_compress(x, _c(codes, np.uint32), n, dim, C_, cutoffs.ravel(), stride, n_bit, out)
    return out


def decode_table(weights, dim: int, n_bit: int) -> np.ndarray:
    weights = _c(weights, np.float32)
    stride = 0 if weights.ndim == 1 else weights.shape[1]
    W = np.empty((dim, 1 << n_bit), np.float32)
    lib().orc_decode_table(weights.ravel(), stride, dim, n_bit, W)
    return W


def decompress_tokens(tok_ids, codes, residual, C_, W, n_bit: int) -> np.ndarray:
    tok_ids = _c(tok_ids, np.uint64)
    C_ = _c(C_, np.float32)
    out = np.empty((tok_ids.size, C_.shape[1]), np.float32)
    lib().orc_decompress_tokens(tok_ids, tok_ids.size, _c(codes, np.uint32), _c(residual, np.uint8).ravel(),
                                C_, _c(W, np.float32), C_.shape[1], n_bit, out)
    return out


def ivf_build(codes, doclens, n_c: int):
    codes, doclens = _c(codes, np.uint32), _c(doclens, np.uint32)
    off = np.empty(n_c, np.uint64)
    size = np.empty(n_c, np.uint32)
    n = lib().orc_ivf_build(codes, doclens, doclens.size, n_c, off, size, None)
    ids = np.empty(n, np.uint32)
    lib().orc_ivf_build(codes, doclens, doclens.size, n_c, off, size, ids.ctypes.data_as(C.c_void_p))
    return off, size, ids


# ------------------------------------------------------------------ search
def centroid_topk(Q: np.ndarray, C_: np.ndarray, P: int):
    Q, C_ = _c(Q, np.float32), _c(C_, np.float32)
    rows = Q.reshape(-1, Q.shape[-1])
    ids = np.empty((rows.shape[0], P), np.uint32)
    s = np.empty((rows.shape[0], P), np.float32)
    lib().orc_centroid_topk(rows, rows.shape[0], C_, C_.shape[0], C_.shape[1], P, ids, s)
    return ids, s


def filter_all(top_ids, top_s, P: int, ivf, n_doc: int, basic: bool = False):
    """PP filter (per-token maxima max(0, .)); basic=True: the basic
    variant's first-hit scores."""
    top_ids, top_s = _c(top_ids, np.uint32), _c(top_s, np.float32)
    T, Pmax = top_ids.shape
    off, size, ids = ivf
    cand = np.empty(n_doc, np.uint32)
    F = np.empty(n_doc, np.float32)
    n = lib().orc_filter_all(top_ids, top_s, T, Pmax, P, _c(off, np.uint64), _c(size, np.uint32),
                             _c(ids, np.uint32), n_doc, int(basic), cand, F)
    return cand[:n].copy(), F[:n].copy()


def select_top(ids, s, R: int):
    ids, s = _c(ids, np.uint32), _c(s, np.float32)
    oi = np.empty(max(R, 1), np.uint32)
    os_ = np.empty(max(R, 1), np.float32)
    n = lib().orc_select_top(ids, s, ids.size, R, oi, os_)
    return oi[:n].copy(), os_[:n].copy()


class Index:
    """Host-side IGP index in the reference layout plus the decode table."""

    def __init__(self, doclens, codes, residual, centroids, weights, n_bit=2, ivf=None):
        self.doclens = _c(doclens, np.uint32)
        self.docoff = np.zeros(self.doclens.size, np.uint64)
        if self.doclens.size > 1:
            self.docoff[1:] = np.cumsum(self.doclens[:-1], dtype=np.uint64)
        self.codes = _c(codes, np.uint32)
        self.residual = _c(residual, np.uint8).reshape(self.codes.size, -1)
        self.C = _c(centroids, np.float32)
        self.weights = _c(weights, np.float32)
        self.n_bit = n_bit
        self.dim = self.C.shape[1]
        self.W = decode_table(self.weights, self.dim, n_bit)
        self.ivf = ivf if ivf is not None else ivf_build(self.codes, self.doclens, self.C.shape[0])

    @property
    def n_doc(self):
        return self.doclens.size

    @property
    def n_c(self):
        return self.C.shape[0]

    def rerank(self, Q, docs, clip_zero=False):
        Q = _c(Q, np.float32)
        docs = _c(docs, np.uint32)
        out = np.empty(docs.size, np.float32)
        lib().orc_rerank(Q, Q.shape[0], self.dim, docs, docs.size, self.doclens, self.docoff, self.codes,
                         self.residual.ravel(), self.C, self.W, self.n_bit, int(clip_zero), out)
        return out

    def decompress_doc(self, d: int):
        tok = np.arange(self.docoff[d], self.docoff[d] + self.doclens[d], dtype=np.uint64)
        return decompress_tokens(tok, self.codes, self.residual, self.C, self.W, self.n_bit)

    def search(self, Q, P, R, k, q_ntok=None, q_nprobe=None, clip_zero=False, n_threads=0,
               timing=False):
        Q = _c(Q, np.float32)
        n_q, Tmax, dim = Q.shape
        assert Tmax <= 32 and dim == self.dim
        ids = np.empty((n_q, k), np.uint32)
        s = np.empty((n_q, k), np.float32)
        qt, qt_p = _ptr_or_null(q_ntok, np.uint32)
        qp, qp_p = _ptr_or_null(q_nprobe, np.uint32)
        secs = np.zeros(n_q, np.float64)
        off, size, ivf_ids = self.ivf
        lib().orc_search_batch(self.n_doc, self.n_c, dim, self.n_bit, self.codes.size, self.doclens,
                               self.docoff, self.codes, self.residual.ravel(), self.C, self.W,
                               _c(off, np.uint64), _c(size, np.uint32), _c(ivf_ids, np.uint32),
                               Q, n_q, Tmax, qt_p, qp_p, P, R, k, int(clip_zero), int(n_threads),
                               ids, s, secs.ctypes.data_as(C.c_void_p))
        return (ids, s, secs) if timing else (ids, s)


def exhaustive_f64(Q, X, doclens, k: int, q_ntok=None, n_threads=0):
    """FP64 exhaustive MaxSim top-k over embeddings X [n_tok, dim]."""
    Q = _c(Q, np.float32)
    n_q, Tmax, dim = Q.shape
    X = _c(X, np.float32)
    doclens = _c(doclens, np.uint32)
    docoff = np.zeros(doclens.size, np.uint64)
    if doclens.size > 1:
        docoff[1:] = np.cumsum(doclens[:-1], dtype=np.uint64)
    ids = np.empty((n_q, k), np.uint32)
    s = np.empty((n_q, k), np.float64)
    qt, qt_p = _ptr_or_null(q_ntok, np.uint32)
    lib().orc_exhaustive_f64(Q, n_q, Tmax, qt_p, X, doclens, docoff, doclens.size, dim, k,
                             int(n_threads), ids, s)
    return ids, s


# ----------------------------------------------------- synthetic workloads
def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest even), returned as float32.  NaN
    stays NaN (quietened); finite values past the largest bfloat16 round to
    +-inf, as the hardware conversion does."""
    x32 = np.ascontiguousarray(x, np.float32)
    b = x32.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16) & 0xFFFFFFFF
    out = b.astype(np.uint32)
    nan = np.isnan(x32)
    out[nan] = (x32.view(np.uint32)[nan] & 0xFFFF0000) | 0x00400000
    return out.view(np.float32)


def e4m3_round(x: np.ndarray) -> np.ndarray:
    """Round float32 to OCP FP8 E4M3 (3 mantissa bits, bias 7, subnormals down
    to 2^-9, saturating at +-448), nearest even; returned as float32.  Model
    of the FP8 precision mode (no counterpart in the reference)."""
    x = np.asarray(x, np.float32).astype(np.float64)
    a = np.minimum(np.abs(x), 448.0)
    _, e = np.frexp(a)                      # a = m 2^e, m in [0.5, 1)
    step = np.exp2(np.maximum(e - 4, -9).astype(np.float64))
    q = np.round(a / step) * step           # numpy rounds half to even
    return (np.sign(x) * np.minimum(q, 448.0)).astype(np.float32)


def fp8_scaled(x: np.ndarray, axis=-1):
    """Per-row fp8 encoding as the FP8 mode does it: s = max|x| / 448 and
    e4m3(x * (448 / max|x|)) in float32 arithmetic; returns (codes as float32
    values, s)."""
    x = np.asarray(x, np.float32)
    amax = np.max(np.abs(x), axis=axis, keepdims=True).astype(np.float32)
    with np.errstate(divide="ignore"):
        inv = np.where(amax > 0, np.float32(448.0) / amax, np.float32(0.0)).astype(np.float32)
    return e4m3_round((x * inv).astype(np.float32)), (amax / np.float32(448.0)).astype(np.float32)


def fp8_tile_maxima(Q: np.ndarray, C_: np.ndarray) -> np.ndarray:
    """FP8 centroid tile maxima as the FP8 mode defines them (no counterpart
    in the reference, whose centroid scores are FP32 GEMMs): query tokens and
    centroids encoded per row (fp8_scaled), score(t, c) = s_t s_c q8.c8 in
    float64, maximum over each tile of 32 consecutive centroids.  Returns
    [rows, ceil(n_c / 32)] float64."""
    rows = np.asarray(Q, np.float32).reshape(-1, np.shape(C_)[1])
    q8, sq = fp8_scaled(rows)
    c8, sc = fp8_scaled(C_)
    S = (q8.astype(np.float64) @ (c8.astype(np.float64) * sc.astype(np.float64)).T) * sq.astype(np.float64)
    n_c = S.shape[1]
    n_t = (n_c + 31) // 32
    pad = np.full((S.shape[0], n_t * 32 - n_c), -np.inf)
    return np.concatenate([S, pad], 1).reshape(S.shape[0], n_t, 32).max(2)


def centroid_topk_fp8(Q: np.ndarray, C_: np.ndarray, P: int, probe_tiles: int = 2, rel: float = 1e-5):
    """The FP8 mode's probe: the exact FP32 top-P (centroid_topk's chain and
    order) among the centroids of the P' = min(n_tiles, probe_tiles P, 256)
    tiles with the highest FP8 maxima (every tile reaching the P'-th largest).
    Returns (ids, scores, ambiguous): ambiguous rows hold a tile maximum
    within `rel` of the P'-th largest on either side of it, where FP32
    summation in the matrix unit can order the tiles differently."""
    C_ = np.asarray(C_, np.float32)
    rows = np.asarray(Q, np.float32).reshape(-1, C_.shape[1])
    tm = fp8_tile_maxima(rows, C_)
    n_t = tm.shape[1]
    Pt = min(n_t, probe_tiles * P, max(P, 256))
    ids = np.empty((rows.shape[0], P), np.uint32)
    s = np.empty((rows.shape[0], P), np.float32)
    amb = np.zeros(rows.shape[0], bool)
    for r in range(rows.shape[0]):
        a = np.sort(tm[r])[::-1][Pt - 1]
        sel = np.nonzero(tm[r] >= a)[0]
        near = np.abs(tm[r] - a) <= rel * max(abs(a), 1e-30)
        amb[r] = Pt < n_t and near.sum() > 1
        cand = (sel[:, None] * 32 + np.arange(32)[None, :]).ravel()
        cand = cand[cand < C_.shape[0]]
        ci, cs = centroid_topk(rows[r:r + 1], C_[cand], min(P, cand.size))
        ids[r, :] = 0xFFFFFFFF
        s[r, :] = -np.inf
        ids[r, :ci.shape[1]] = cand[ci[0]]
        s[r, :ci.shape[1]] = cs[0]
    return ids, s, amb


def recall_at_k(est: np.ndarray, gnd: np.ndarray) -> float:
    """|gnd ∩ est| / k averaged over queries (script/evaluation/performance_metric.py:259-272)."""
    k = gnd.shape[1]
    tot = 0.0
    for e, g in zip(est, gnd):
        tot += len(set(e[:k].tolist()) & set(g.tolist())) / k
    return tot / len(gnd)


def maxsim_scores_f32(Q, X, lens, n_threads=0) -> np.ndarray:
    """All-pairs FP32 MaxSim scores [n_q, n_item] of the exhaustive module
    (GPUComputeScore.hpp:196-246): FMA-chain dots, max per item and query
    token, left-to-right sum over query tokens."""
    Q = _c(Q, np.float32)
    X = _c(X, np.float32)
    lens = _c(lens, np.uint32)
    n_q, T, dim = Q.shape
    out = np.empty((n_q, lens.size), np.float32)
    lib().orc_maxsim_scores_f32(Q, n_q, T, dim, X, lens, lens.size, n_threads, out)
    return out
