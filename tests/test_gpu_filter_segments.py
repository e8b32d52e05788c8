"""The filter across many 4,096-doc segments: 70,000 docs make 18 segments,
so a workgroup of the grouped filter walks 8 of them (two full groups and a
group of two, the last segment partial).  Lists are short (1,024 centroids:
runs of about 16 entries per segment) or long (32 centroids: runs of about
500, far beyond the entries loaded ahead).  nprobe 8, 12 and 16 take the
kernel with one probe rank per lane, 32 (the largest nprobe the library
takes) the one with two; a segment's ranks go in one to four batches of
eight.  Per-query nprobe and token counts are ragged.  Every stage is
compared bit-exact with the CPU oracle (precision fp32), through the C ABI.
"""
import numpy as np
import pytest

from test_gpu_parity import _check, _oracle_stages

pytestmark = pytest.mark.gpu

N_DOC = 70_000


def _wide_index(n_c, seed):
    """Random doc lengths (2..6 tokens), random code assignment and residual
    bytes: the filter stage depends only on the lists and the centroids."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    doclens = rng.integers(2, 7, N_DOC).astype(np.uint32)
    n_tok = int(doclens.sum())
    C = rng.standard_normal((n_c, 128)).astype(np.float32)
    C /= np.linalg.norm(C, axis=1, keepdims=True)
    codes = rng.integers(0, n_c, n_tok).astype(np.uint32)
    residual = rng.integers(0, 256, (n_tok, 32), dtype=np.uint8)
    weights = np.array([-0.03, -0.01, 0.01, 0.03], np.float32)
    return O.Index(doclens, codes, residual, C, weights)


_CACHE = {}


def _case(n_c):
    if n_c not in _CACHE:
        ix = _wide_index(n_c, 100 + n_c)
        rng = np.random.default_rng(n_c)
        Q = rng.standard_normal((8, 32, 128)).astype(np.float32)
        Q /= np.linalg.norm(Q, axis=2, keepdims=True)
        _CACHE[n_c] = (ix, Q)
    return _CACHE[n_c]


@pytest.mark.parametrize("n_c", [1024, 32])
@pytest.mark.parametrize("P", [8, 12, 16, 32])
def test_filter_many_segments_matches_oracle(n_c, P):
    import mvr
    ix, Q = _case(n_c)
    rng = np.random.default_rng(P)
    qt = rng.integers(1, 33, Q.shape[0]).astype(np.uint32)
    qt[0] = 32
    qp = rng.integers(1, P + 1, Q.shape[0]).astype(np.uint32)
    qp[0] = P
    g = mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)
    for kw in (dict(), dict(q_ntok=qt, q_nprobe=qp)):
        ids, sc, st = g.search(Q, P, 64, 10, stages=True, **kw)
        ref = _oracle_stages(ix, Q, P, 64, 10, kw.get("q_ntok"), kw.get("q_nprobe"))
        if not kw:
            assert max(r["cand_n"] for r in ref) > 4096 * 8  # candidates in every group
        _check(st, ids, sc, ref)
    g.close()
