"""BF16 MaxSim loop orders: doc-major (each selected doc's tiles loaded once per
device pass for every query that selected it) against query-major (the
candidates of each query in turn).  Both orders must give bit-identical
scores at every stage, on docs that fit one register chunk (<= 96 tokens),
docs that take several chunks, and docs selected by more queries than one
pair group holds (16); the scores also stay within the BF16 model's
tolerance.  Calls go through the C ABI (libmvr) via the mvr package."""
import numpy as np
import pytest

from test_gpu_parity import _bf16_maxsim, _index

pytestmark = pytest.mark.gpu


def _both_orders(g, Q, P, R, k, **kw):
    out = {}
    for order in ("query", "doc"):
        g.set_maxsim_order(order)
        out[order] = g.search(Q, P, R, k, precision="bf16", stages=True, **kw)
    g.set_maxsim_order("auto")
    return out


def _assert_same(out):
    (iq, sq, stq), (idd, sd, std) = out["query"], out["doc"]
    np.testing.assert_array_equal(stq["sel_ids"], std["sel_ids"])
    np.testing.assert_array_equal(stq["sel_score"].view(np.uint32), std["sel_score"].view(np.uint32))
    np.testing.assert_array_equal(iq, idd)
    np.testing.assert_array_equal(sq.view(np.uint32), sd.view(np.uint32))


@pytest.mark.parametrize("clip_zero", [False, True])
def test_doc_order_bit_identical_config1(c1, clip_zero):
    """Config 1 (U{16..128}-token docs: one and two register chunks)."""
    corpus, ix, Q = c1
    g = _index(ix)
    out = _both_orders(g, Q, 4, 256, 10, rescore=0, clip_zero=clip_zero)
    _assert_same(out)
    st = out["doc"][2]
    for q in range(0, Q.shape[0], 8):
        n = int((st["sel_ids"][q] != 0xFFFFFFFF).sum())
        ref = _bf16_maxsim(ix, Q[q], st["sel_ids"][q, :n], clip_zero=clip_zero)
        np.testing.assert_allclose(st["sel_score"][q, :n], ref, rtol=0, atol=2e-4)
    g.close()


def test_doc_order_ragged_tokens_and_rescore(c1):
    """Uneven query token counts and a FP32 rescore of the best 64."""
    corpus, ix, Q = c1
    rng = np.random.default_rng(5)
    q_ntok = rng.integers(1, 33, Q.shape[0]).astype(np.uint32)
    q_np = rng.choice([2, 4, 8], Q.shape[0]).astype(np.uint32)
    g = _index(ix)
    out = _both_orders(g, Q, 8, 256, 10, rescore=64, q_ntok=q_ntok, q_nprobe=q_np)
    _assert_same(out)
    g.close()


def test_doc_order_long_docs_many_queries():
    """300 docs of 8..300 tokens (up to four register chunks), 96 queries x
    probe_topk 200: about 64 queries per doc, so each doc's pairs span four
    or more groups of 16."""
    from oracle import synth
    corpus = synth.make_corpus(77, 300, 64, lo=8, hi=300)
    ix = synth.build_index(corpus, 64, n_iter=4)
    Q = synth.make_queries(78, corpus.centers, 96)
    assert ix.doclens.max() > 3 * 96
    g = _index(ix)
    out = _both_orders(g, Q, 8, 200, 20, rescore=0)
    _assert_same(out)
    st = out["doc"][2]
    for q in (0, 31, 95):
        n = int((st["sel_ids"][q] != 0xFFFFFFFF).sum())
        ref = _bf16_maxsim(ix, Q[q], st["sel_ids"][q, :n])
        np.testing.assert_allclose(st["sel_score"][q, :n], ref, rtol=0, atol=2e-4)
    g.close()


def test_doc_order_small_batches_and_batch_split(small):
    """Forced doc order on passes with fewer pairs than docs, and a search
    split into several passes (the pair buffers reused pass after pass)."""
    corpus, ix, Q = small
    g = _index(ix)
    for batch in (1, 5, 0):
        out = {}
        for order in ("query", "doc"):
            g.set_maxsim_order(order)
            out[order] = g.search(Q, 4, 64, 10, precision="bf16", rescore=0, batch=batch, stages=True)
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
