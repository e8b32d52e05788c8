"""nprobe above 32, the values the reference's own drivers pass
(script/evaluation/eval_igp_gpu_pp.py:243,266,283 use -nprobe 64; the grid
there lists up to 256), and the PP filter's zero clamp.

* nprobe 33..64 take the token-major filter with four ranks per lane (64-rank
  run tables); 65..256 the rank-group filter.  The centroid top-P keeps one,
  two or four list registers per lane (64 elements each).
* Every stage is bit-exact against the CPU oracle in precision fp32, with
  uniform and ragged per-query nprobe and token counts; BF16 centroid top-P is
  exact on bf16-rounded inputs; FP32 MFMA agrees within 1e-5 up to ties.
* PP filter: per-(doc, query token) maxima start at zero (FilterPP.hpp:136-141
  with the memset at :257-259), so lists whose centroid score S' = dot + 1 is
  negative add nothing; the basic variant adds its first hit as it is.
"""
import numpy as np
import pytest

from test_gpu_parity import _check, _check_mfma, _oracle_stages

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wide():
    """2,048 docs, 512 centroids (so nprobe 256 is half the centroids), 8 queries."""
    from oracle import synth
    return synth.config1(n_doc=2048, n_comp=1024, n_c=512, n_q=8, n_iter=4)


def _index(ix):
    import mvr
    return mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)


@pytest.mark.parametrize("P", [33, 48, 64, 65, 100, 128, 129, 200, 256])
def test_wide_nprobe_fp32_bit_exact(wide, P):
    corpus, ix, Q = wide
    g = _index(ix)
    ids, sc, st = g.search(Q, P, 256, 10, stages=True)
    _check(st, ids, sc, _oracle_stages(ix, Q, P, 256, 10))
    g.close()


def test_wide_nprobe_ragged_per_query(wide):
    corpus, ix, Q = wide
    rng = np.random.default_rng(64)
    qt = rng.integers(1, 33, Q.shape[0]).astype(np.uint32)
    qp = rng.choice([1, 31, 40, 64, 77, 150, 256], Q.shape[0]).astype(np.uint32)
    qp[0] = 256
    g = _index(ix)
    ids, sc, st = g.search(Q, 8, 300, 20, q_ntok=qt, q_nprobe=qp, stages=True)
    _check(st, ids, sc, _oracle_stages(ix, Q, 8, 300, 20, qt, qp))
    g.close()


@pytest.mark.parametrize("P", [64, 256])
def test_wide_nprobe_bf16_centroids_and_fp32_mfma(wide, P):
    from oracle import oracle as O
    corpus, ix, Q = wide
    g = _index(ix)
    _, _, st = g.search(Q, P, 64, 10, precision="bf16", stages=True)
    ti, ts = O.centroid_topk(O.bf16_round(Q), O.bf16_round(ix.C), P)
    np.testing.assert_array_equal(st["top_ids"], ti.reshape(Q.shape[0], Q.shape[1], P))
    np.testing.assert_array_equal(st["top_s"], ts.reshape(Q.shape[0], Q.shape[1], P))
    ids, sc, st = g.search(Q, P, 128, 10, precision="fp32_mfma", stages=True)
    _check_mfma(st, ids, sc, _oracle_stages(ix, Q, P, 128, 10), ix, Q)
    g.close()


def test_nprobe_above_limit_is_unsupported(wide):
    import mvr
    corpus, ix, Q = wide
    g = _index(ix)
    with pytest.raises(mvr.MvrError) as e:
        g.search(Q, 257, 64, 10)
    assert e.value.status == 4  # MVR_ERR_UNSUPPORTED
    g.close()


def _cone_index(seed=5):
    """Centroids in a narrow cone around a direction v: queries along -2v score
    S' = dot + 1 < 0 against every centroid."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(128)
    v /= np.linalg.norm(v)
    n_c, n_doc = 64, 600
    C_ = v[None, :] + 0.05 * rng.standard_normal((n_c, 128))
    C_ = (C_ / np.linalg.norm(C_, axis=1, keepdims=True)).astype(np.float32)
    doclens = rng.integers(3, 12, n_doc).astype(np.uint32)
    n_tok = int(doclens.sum())
    codes = rng.integers(0, n_c, n_tok).astype(np.uint32)
    residual = rng.integers(0, 256, (n_tok, 32), dtype=np.uint8)
    weights = np.array([-0.02, -0.005, 0.005, 0.02], np.float32)
    ix = O.Index(doclens, codes, residual, C_, weights)
    Q = np.empty((4, 32, 128), np.float32)
    for q in range(4):
        # half the tokens opposite the cone (every S' negative), half along it
        sgn = np.where(np.arange(32) % 2 == 0, -2.0, 1.0)[:, None]
        Q[q] = (sgn * v[None, :] + 0.05 * rng.standard_normal((32, 128))).astype(np.float32)
    return ix, Q


@pytest.mark.parametrize("P", [4, 40, 64])
def test_pp_filter_clamps_negative_centroid_scores(P):
    from oracle import oracle as O
    ix, Q = _cone_index()
    g = _index(ix)
    # the probed scores of the even tokens are all negative
    _, ts = O.centroid_topk(Q[0], ix.C, P)
    assert (ts[0::2] < 0).all() and (ts[1::2] > 0).all()
    for clip in (False, True):
        ids, sc, st = g.search(Q, P, 128, 10, clip_zero=clip, stages=True)
        ref = _oracle_stages(ix, Q, P, 128, 10, clip_zero=clip)
        _check(st, ids, sc, ref)
    # the clamp is what separates them: PP filter scores are never negative,
    # and every doc hit by a negative token scores lower in the basic variant
    ti, ts = O.centroid_topk(Q[0], ix.C, P)
    _, F_pp = O.filter_all(ti, ts, P, ix.ivf, ix.n_doc)
    _, F_basic = O.filter_all(ti, ts, P, ix.ivf, ix.n_doc, basic=True)
    assert (F_pp >= 0).all() and (F_basic < F_pp).any()
    _, _, pp = g.search(Q, P, 128, 10, stages=True)
    n = int((pp["sel_ids"][0] != 0xFFFFFFFF).sum())
    assert (pp["sel_F"][0, :n] >= 0).all()
    g.close()
