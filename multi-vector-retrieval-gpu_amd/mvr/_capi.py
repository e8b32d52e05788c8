"""ctypes binding of libmvr (include/mvr.h, include/mvr_plaid.h).  No fallback: if the library is
missing or fails to load, every call raises."""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(_PKG, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmvr.so")

MVR_OK = 0
MVR_FP32, MVR_BF16, MVR_FP8, MVR_FP32_MFMA = 0, 1, 2, 3
INDEX_RESIDUAL_ON_HOST = 1  # MVR_INDEX_RESIDUAL_ON_HOST


class MvrError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"libmvr status {status}: {msg}")
        self.status = status


class IndexDesc(C.Structure):
    _fields_ = [("n_doc", C.c_uint32), ("n_tok", C.c_uint64), ("n_centroid", C.c_uint32),
                ("dim", C.c_uint32), ("n_bit", C.c_uint32), ("doclens", C.c_void_p),
                ("centroids", C.c_void_p), ("codes", C.c_void_p), ("residual", C.c_void_p),
                ("weights", C.c_void_p), ("weights_per_dim", C.c_uint32),
                ("ivf_offsets", C.c_void_p), ("ivf_sizes", C.c_void_p), ("ivf_ids", C.c_void_p),
                ("ivf_n_ele", C.c_uint64), ("on_device", C.c_int)]


class SearchParams(C.Structure):
    _fields_ = [("nprobe", C.c_uint32), ("probe_topk", C.c_uint32), ("topk", C.c_uint32),
                ("precision", C.c_uint32), ("rescore", C.c_uint32), ("clip_zero", C.c_uint32),
                ("batch", C.c_uint32), ("learned", C.c_uint32), ("persistent", C.c_uint32),
                ("graph", C.c_uint32)]


class StageOut(C.Structure):
    _fields_ = [("top_ids", C.c_void_p), ("top_s", C.c_void_p), ("cand_n", C.c_void_p),
                ("sel_ids", C.c_void_p), ("sel_F", C.c_void_p), ("sel_score", C.c_void_p),
                ("nprobe", C.c_void_p), ("rerank_n", C.c_void_p)]


class Timings(C.Structure):
    _fields_ = [("centroid_ms", C.c_double), ("filter_ms", C.c_double), ("select_ms", C.c_double),
                ("rerank_ms", C.c_double), ("rescore_ms", C.c_double), ("topk_ms", C.c_double),
                ("total_ms", C.c_double), ("rerank_tokens", C.c_uint64), ("candidates", C.c_uint64),
                ("n_pass", C.c_uint32), ("rerank_pairs", C.c_uint64), ("staging_ms", C.c_double),
                ("staged_docs", C.c_uint64), ("staged_tokens", C.c_uint64)]


class PlaidIndexDesc(C.Structure):
    _fields_ = [("n_doc", C.c_uint32), ("n_centroid", C.c_uint32), ("dim", C.c_uint32), ("nbits", C.c_uint32),
                ("n_tok", C.c_uint64), ("n_ivf", C.c_uint64), ("doclens", C.c_void_p), ("codes", C.c_void_p),
                ("residuals", C.c_void_p), ("centroids", C.c_void_p), ("bucket_weights", C.c_void_p),
                ("ivf", C.c_void_p), ("ivf_lengths", C.c_void_p)]


class PlaidParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("ncells", C.c_uint32), ("ndocs", C.c_uint32), ("batch", C.c_uint32),
                ("centroid_score_threshold", C.c_float), ("gpu_branch", C.c_uint32)]


class PlaidTimings(C.Structure):
    _fields_ = [("candidates_ms", C.c_double), ("filter_ms", C.c_double), ("refine_ms", C.c_double),
                ("total_ms", C.c_double), ("n_refine_ivf", C.c_uint64), ("n_refine_filter", C.c_uint64),
                ("n_vec_score_refine", C.c_uint64), ("scores_ms", C.c_double), ("n_pass", C.c_uint32)]


_vp, _u32, _u64 = C.c_void_p, C.c_uint32, C.c_uint64

# Every symbol declared in include/mvr.h and include/mvr_plaid.h, with its
# ctypes signature.
SIGNATURES = {
    "mvr_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mvr_last_error": (C.c_char_p, []),
    "mvr_index_create": (C.c_int, [C.POINTER(IndexDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "mvr_index_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    "mvr_index_create_ex": (C.c_int, [C.POINTER(IndexDesc), C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    "mvr_index_load_ex": (C.c_int, [C.c_char_p, C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    "mvr_index_free": (None, [C.c_void_p]),
    "mvr_index_set_maxsim_order": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mvr_index_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "mvr_search": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                             C.POINTER(SearchParams), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                             C.POINTER(StageOut), C.POINTER(Timings)]),
    "mvr_decompress": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "mvr_index_set_tuning": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "mvr_index_get_tuning": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "mvr_index_get_tuning_ex": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "mvr_rerank_docs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mvr_rerank_docs_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mvr_exhaustive_f64": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mvr_mlp_last_error": (C.c_char_p, []),
    "mvr_mlp_create": (C.c_int, [C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_void_p)]),
    "mvr_mlp_free": (None, [C.c_void_p]),
    "mvr_mlp_num_params": (C.c_uint64, [C.c_void_p]),
    "mvr_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mvr_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mvr_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "mvr_mlp_grads": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                C.c_void_p, C.POINTER(C.c_float)]),
    "mvr_mlp_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                     C.c_float, C.POINTER(C.c_float), C.c_void_p]),
    "mvr_index_set_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_uint32, C.c_uint32]),
    "mvr_synth_tokens": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint32,
                                   C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, C.c_void_p,
                                   C.c_void_p]),
    "mvr_assign": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                             C.c_void_p]),
    "mvr_assign_bias": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint32,
                                  C.c_uint32, C.c_void_p, C.c_void_p]),
    "mvr_compress": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32,
                               C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "mvr_compress_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.c_uint32,
                                    C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "mvr_maxsim_scores_host": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_uint32, C.c_void_p, C.c_int]),
    "mvr_maxsim_scores_host_ex": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_uint32, C.c_void_p, C.c_int, C.c_uint32]),
    "mvr_maxsim_last_error": (C.c_char_p, []),
    # include/mvr_plaid.h
    "mvr_plaid_last_error": (C.c_char_p, []),
    "mvr_plaid_decompress_residuals": (C.c_int, [_vp, _u64, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp]),
    "mvr_plaid_packbits": (C.c_int, [_vp, _u64, _vp, _vp]),
    "mvr_plaid_binarize": (C.c_int, [_vp, _u64, _u32, _vp, _u32, _u32, _vp, _vp]),
    "mvr_plaid_lengths_cumsum": (C.c_int, [_vp, _u32, _vp, _vp, C.POINTER(C.c_int64), _vp]),
    "mvr_plaid_decompress_pids": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32,
                                            _vp, _vp]),
    "mvr_plaid_segmented_maxsim": (C.c_int, [_vp, _u32, _vp, _u32, _vp, _vp]),
    "mvr_plaid_segmented_lookup": (C.c_int, [_vp, C.c_uint64, _u32, _vp, _vp, _u32, _vp, C.c_uint64, _vp]),
    "mvr_plaid_colbert_score_packed_f16": (C.c_int, [_vp, _u32, _vp, C.c_uint64, _vp, _u32, _vp, _vp]),
    "mvr_plaid_colbert_score_reduce_f16": (C.c_int, [_vp, _vp, _u32, _u32, _u32, _vp, _vp]),
    "mvr_plaid_lookup_pids_f16": (C.c_int, [_vp, _vp, _u32, _vp, C.c_uint64, C.POINTER(C.c_uint64), _vp]),
    "mvr_plaid_filter_pids": (C.c_int, [_vp, _u32, _vp, _u32, _u32, _vp, _vp, _vp, _vp, _u32, _vp,
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), _vp]),
    "mvr_plaid_index_create": (C.c_int, [C.POINTER(PlaidIndexDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "mvr_plaid_index_free": (None, [_vp]),
    "mvr_plaid_search": (C.c_int, [_vp, _vp, _u32, _u32, C.POINTER(PlaidParams), _vp, _vp, C.c_int, _vp,
                                   C.POINTER(PlaidTimings)]),
}

_lib = None


def lib():
    """Load libmvr.so once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                           f"python multi-vector-retrieval-gpu_amd/build.py lib")
    # libmvr links libamdhip64; torch ships its own copy under the same
    # soname.  Load torch first so both share one HIP runtime (a second
    # runtime would see no devices once torch owns them).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int):
    if status != MVR_OK:
        raise MvrError(status, lib().mvr_last_error().decode())


def check_plaid(status: int):
    if status != MVR_OK:
        raise MvrError(status, lib().mvr_plaid_last_error().decode())


def check_mlp(status: int):
    if status != MVR_OK:
        raise MvrError(status, lib().mvr_mlp_last_error().decode())
