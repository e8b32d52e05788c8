"""MaxSim loop orders: doc-major (each selected doc's tiles loaded once per
device pass for every query that selected it) against query-major (the
candidates of each query in turn), for the BF16 and the FP8 MaxSim.  Both
orders must give bit-identical scores at every stage, on docs that fit one
register chunk (<= 128 tokens), docs that take several chunks, and docs
selected by more queries than one pair group holds (16); the scores also stay
within the tolerance of the float64 model of each precision.  Calls go through
the C ABI (libmvr) via the mvr package."""
import numpy as np
import pytest

from test_gpu_parity import _bf16_maxsim, _fp8_maxsim, _index

pytestmark = pytest.mark.gpu

PRECISIONS = ["bf16", "fp8"]


def _both_orders(g, Q, P, R, k, precision="bf16", **kw):
    out = {}
    for order in ("query", "doc"):
        g.set_maxsim_order(order)
        out[order] = g.search(Q, P, R, k, precision=precision, stages=True, **kw)
    g.set_maxsim_order("auto")
    return out


def _check_model(ix, Qq, docs, got, precision, clip_zero=False):
    """The BF16 scores within 2e-4 of their float64 model, the FP8 ones within
    1e-4 relative (+1e-4) of theirs."""
    if precision == "bf16":
        np.testing.assert_allclose(got, _bf16_maxsim(ix, Qq, docs, clip_zero=clip_zero), rtol=0, atol=2e-4)
    else:
        np.testing.assert_allclose(got, _fp8_maxsim(ix, Qq, docs, clip_zero=clip_zero), rtol=1e-4, atol=1e-4)


def _assert_same(out):
    (iq, sq, stq), (idd, sd, std) = out["query"], out["doc"]
    np.testing.assert_array_equal(stq["sel_ids"], std["sel_ids"])
    np.testing.assert_array_equal(stq["sel_score"].view(np.uint32), std["sel_score"].view(np.uint32))
    np.testing.assert_array_equal(iq, idd)
    np.testing.assert_array_equal(sq.view(np.uint32), sd.view(np.uint32))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("clip_zero", [False, True])
def test_doc_order_bit_identical_config1(c1, clip_zero, precision):
    """Config 1 (U{16..128}-token docs: one and two BF16 register chunks, one
    FP8 chunk)."""
    corpus, ix, Q = c1
    g = _index(ix)
    out = _both_orders(g, Q, 4, 256, 10, precision=precision, rescore=0, clip_zero=clip_zero)
    _assert_same(out)
    st = out["doc"][2]
    for q in range(0, Q.shape[0], 8):
        n = int((st["sel_ids"][q] != 0xFFFFFFFF).sum())
        _check_model(ix, Q[q], st["sel_ids"][q, :n], st["sel_score"][q, :n], precision, clip_zero)
    g.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_doc_order_ragged_tokens_and_rescore(c1, precision):
    """Uneven query token counts and a FP32 rescore of the best 64."""
    corpus, ix, Q = c1
    rng = np.random.default_rng(5)
    q_ntok = rng.integers(1, 33, Q.shape[0]).astype(np.uint32)
    q_np = rng.choice([2, 4, 8], Q.shape[0]).astype(np.uint32)
    g = _index(ix)
    out = _both_orders(g, Q, 8, 256, 10, precision=precision, rescore=64, q_ntok=q_ntok, q_nprobe=q_np)
    _assert_same(out)
    g.close()


@pytest.fixture(scope="module")
def long_docs():
    from oracle import synth
    corpus = synth.make_corpus(77, 300, 64, lo=8, hi=300)
    ix = synth.build_index(corpus, 64, n_iter=4)
    Q = synth.make_queries(78, corpus.centers, 96)
    return ix, Q


@pytest.mark.parametrize("precision", PRECISIONS)
def test_doc_order_long_docs_many_queries(long_docs, precision):
    """300 docs of 8..300 tokens (up to three register chunks), 96 queries x
    probe_topk 200: about 64 queries per doc, so each doc's pairs span four
    or more groups of 16."""
    ix, Q = long_docs
    assert ix.doclens.max() > 2 * 128
    g = _index(ix)
    out = _both_orders(g, Q, 8, 200, 20, precision=precision, rescore=0)
    _assert_same(out)
    st = out["doc"][2]
    for q in (0, 31, 95):
        n = int((st["sel_ids"][q] != 0xFFFFFFFF).sum())
        _check_model(ix, Q[q], st["sel_ids"][q, :n], st["sel_score"][q, :n], precision)
    g.close()


def test_doc_order_small_batches_and_batch_split(small):
    """Forced doc order on passes with fewer pairs than docs, and a search
    split into several passes (the pair buffers reused pass after pass)."""
    corpus, ix, Q = small
    g = _index(ix)
    for precision in PRECISIONS:
        for batch in (1, 5, 0):
            out = {}
            for order in ("query", "doc"):
                g.set_maxsim_order(order)
                out[order] = g.search(Q, 4, 64, 10, precision=precision, rescore=0, batch=batch, stages=True)
            _assert_same(out)
    g.set_maxsim_order("auto")
    g.close()


def test_maxsim_order_rejects_unknown():
    import mvr
    from mvr._capi import lib
    from oracle import synth
    corpus, ix, Q = synth.config1(n_doc=256, n_comp=64, n_c=64, n_q=2, n_iter=2)
    g = _index(ix)
    assert lib().mvr_index_set_maxsim_order(g._h, 3) != 0
    g.close()
