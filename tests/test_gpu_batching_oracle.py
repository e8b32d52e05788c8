"""The batching workload's execution modes checked against the CPU oracle
directly (not only against the launch path): the persistent work-queue
MaxSim + top-k kernel, with and without hipGraph replay, on 2,048 queries of
the config-1 corpus with U{8..32} tokens and nprobe uniform over
{2, 4, 8, 16, 32}, k = 100, in 1,024-query device passes.

* FP32 MFMA: every query's top-100 is the oracle's up to ties within 1e-5,
  with scores within 1e-5 of the oracle's exact scores.
* BF16 with an FP32 rescore of the best 200: the candidates are the oracle's
  filter selection on bf16-rounded centroid scores (exact), the returned
  scores are the exact FP32 MaxSim of the returned docs, and the returned docs
  are the FP32 top-100 of the BF16 top-200 -- judged against a float64 model
  of the BF16 MaxSim, with docs inside the model's tolerance band around the
  200th score free to go either way.
"""
import numpy as np
import pytest

from test_gpu_parity import RTOL, _bf16_maxsim, _check_topk_up_to_ties

pytestmark = pytest.mark.gpu

N_Q, R, K, R2 = 2048, 256, 100, 200


@pytest.fixture(scope="module")
def workload(c1):
    from mvr import synth as dsynth
    from oracle import synth
    corpus, ix, Q = c1
    Qb = synth.make_queries(4242, corpus.centers, N_Q)
    qt, qp = dsynth.query_shapes(4242, N_Q)
    oi, os_ = ix.search(Qb, 32, R, K, q_ntok=qt, q_nprobe=qp)
    return Qb, qt, qp, oi, os_


def _index(ix):
    import mvr
    return mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)


@pytest.mark.parametrize("graph", [False, True])
def test_persistent_fp32_mfma_against_oracle(c1, workload, graph):
    corpus, ix, _ = c1
    Qb, qt, qp, oi, os_ = workload
    assert sorted(set(qp.tolist())) == [2, 4, 8, 16, 32] and qt.min() == 8 and qt.max() == 32
    g = _index(ix)
    for rep in range(2 if graph else 1):  # a second round replays the captured passes
        ids, sc = g.search(Qb, 32, R, K, precision="fp32_mfma", q_ntok=qt, q_nprobe=qp, batch=1024,
                           persistent=True, graph=graph)
        for q in range(N_Q):
            _check_topk_up_to_ties(ix, Qb[q, :qt[q]], ids[q], sc[q], oi[q], os_[q], msg=f"rep{rep} q{q}")
    g.close()


def _bf16_answer_ok(ix, Qq, Pq, ids, sc, msg):
    from oracle import oracle as O
    ti, ts = O.centroid_topk(O.bf16_round(Qq), O.bf16_round(ix.C), Pq)
    cand, F = O.filter_all(ti, ts, Pq, ix.ivf, ix.n_doc)
    sel, _ = O.select_top(cand, F, R)
    exact = ix.rerank(Qq, sel)
    ex_of = dict(zip(sel.tolist(), exact.tolist()))
    got = ids[:K]
    assert set(got.tolist()) <= set(sel.tolist()), msg
    np.testing.assert_allclose(sc[:K], [ex_of[int(d)] for d in got], rtol=RTOL, atol=1e-6, err_msg=msg)
    order = np.lexsort((got, -sc[:K].astype(np.float64)))
    np.testing.assert_array_equal(order, np.arange(K), err_msg=msg)
    if sel.size <= R2:
        allowed, required = set(sel.tolist()), set(sel.tolist())
    else:
        model = _bf16_maxsim(ix, Qq, sel)
        s200 = np.sort(model)[::-1][R2 - 1]
        band = 4e-4
        allowed = set(sel[model >= s200 - band].tolist())
        required = set(sel[model > s200 + band].tolist())
    assert set(got.tolist()) <= allowed, msg
    kth = float(sc[K - 1])
    tol = RTOL * max(1.0, abs(kth)) * 2
    for d in required - set(got.tolist()):
        assert ex_of[d] <= kth + tol, f"{msg}: doc {d} ({ex_of[d]}) missing above the k-th score {kth}"


@pytest.mark.parametrize("graph", [False, True])
def test_persistent_bf16_rescore_against_oracle(c1, workload, graph):
    corpus, ix, _ = c1
    Qb, qt, qp, _, _ = workload
    g = _index(ix)
    ids, sc = g.search(Qb, 32, R, K, precision="bf16", rescore=R2, q_ntok=qt, q_nprobe=qp, batch=1024,
                       persistent=True, graph=graph)
    for q in range(0, N_Q, 32):  # 64 queries through the float64 BF16 model
        _bf16_answer_ok(ix, Qb[q, :qt[q]], int(qp[q]), ids[q], sc[q], f"q{q}")
    g.close()
