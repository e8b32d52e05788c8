"""Synthetic workloads built on the CPU with the oracle (test infrastructure only).

The corpus recipe: mixture centres are L2-normalised N(0, I) draws; each token
picks a centre uniformly, adds N(0, noise^2) per dimension and is
L2-normalised.  The index follows the reference's build pipeline
(script/evaluation/technique/vq_sq_pack.py:91-109): k-means centroids
(L2-normalised), codes by argmax inner product, 2-bit residual codes, IVF of
each document's unique centroids (src/include/struct/IVFIndex.hpp:92-166).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import oracle as O


@dataclass
class Corpus:
    seed: int
    centers: np.ndarray
    doclens: np.ndarray
    X: np.ndarray            # [n_tok, dim] original embeddings (float32)


def make_corpus(seed: int, n_doc: int, n_comp: int, lo: int = 16, hi: int = 128, dim: int = 128,
                noise: float = 0.15) -> Corpus:
    centers = O.gen_centers(seed, n_comp, dim)
    doclens = O.gen_doclens_uniform(seed, n_doc, lo, hi)
    X = O.gen_tokens(seed, centers, int(doclens.sum()), noise)
    return Corpus(seed, centers, doclens, X)


def make_queries(seed: int, centers: np.ndarray, n_q: int, T: int = 32, noise: float = 0.15):
    Q = O.gen_tokens(seed, centers, n_q * T, noise, query=True)
    return Q.reshape(n_q, T, -1)


def build_index(corpus: Corpus, n_c: int, n_iter: int = 10, n_bit: int = 2, per_dim: bool = True,
                kmeans_sample: int | None = None):
    """k-means + codes + codec + IVF.  per_dim selects per-dimension quantile
    cutoffs with bucket-mean weights; otherwise the reference's pooled rule."""
    X = corpus.X
    sample = X if kmeans_sample is None else X[:: max(1, X.shape[0] // kmeans_sample)]
    C_ = O.kmeans(sample, n_c, n_iter)
    codes = O.assign_ip(X, C_)
    e = O.residuals(X, codes, C_)
    if per_dim:
        cut, w = O.quantize_perdim(e, n_bit)
    else:
        cut, w = O.quantize_pooled(e, n_bit)
    res = O.compress(X, codes, C_, cut, n_bit)
    ix = O.Index(corpus.doclens, codes, res, C_, w, n_bit)
    ix.cutoffs = cut
    return ix


_C1_CACHE = {}


def config1(seed: int = 20261203, n_doc: int = 8192, n_comp: int = 1024, n_c: int = 1024,
            n_q: int = 64, n_iter: int = 10):
    """Config 1: 8,192 docs with U{16..128} tokens, 1,024 components and
    centroids, 64 queries x 32 tokens (search with P=4, R=256, k=10)."""
    key = (seed, n_doc, n_comp, n_c, n_q, n_iter)
    if key not in _C1_CACHE:
        corpus = make_corpus(seed, n_doc, n_comp)
        ix = build_index(corpus, n_c, n_iter)
        Q = make_queries(seed, corpus.centers, n_q)
        _C1_CACHE[key] = (corpus, ix, Q)
    return _C1_CACHE[key]
