"""Host-resident residual codes (MVR_INDEX_RESIDUAL_ON_HOST): codes and
residual bytes stay in host memory and every device pass stages its selected
docs over PCIe (the reference's RerankPP::rerank / transmitSQCode scenario,
src/include/alg/igp_gpu_pp/RerankPP.hpp:289-361).  The search must return
exactly what the device-resident index returns, which the parity tests tie to
the CPU oracle; the oracle is also checked directly here.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _index(ix, host):
    import mvr
    return mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf,
                                 residual_on_host=host)


def _ragged(Q, seed):
    rng = np.random.default_rng(seed)
    qt = rng.integers(8, Q.shape[1] + 1, Q.shape[0]).astype(np.uint32)
    qp = rng.choice(np.array([2, 4, 8, 16, 32], np.uint32), Q.shape[0])
    return qt, qp


@pytest.mark.parametrize("prec,rescore", [("fp32", 0), ("fp32_mfma", 0), ("bf16", 64), ("fp8", 64),
                                          ("bf16", 0)])
def test_host_residual_matches_device_index(c1, prec, rescore):
    corpus, ix, Q = c1
    dev, host = _index(ix, False), _index(ix, True)
    qt, qp = _ragged(Q, 11)
    kw = dict(precision=prec, rescore=rescore, q_ntok=qt, q_nprobe=qp, batch=24)
    ref_ids, ref_sc = dev.search(Q, 8, 256, 10, **kw)
    ids, sc = host.search(Q, 8, 256, 10, **kw)
    np.testing.assert_array_equal(ids, ref_ids)
    np.testing.assert_array_equal(sc, ref_sc)
    dev.close()
    host.close()


def test_host_residual_matches_oracle_and_stages(small):
    corpus, ix, Q = small
    g = _index(ix, True)
    ids, sc, st = g.search(Q, 4, 64, 10, stages=True)
    oi, os_ = ix.search(Q, 4, 64, 10)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(sc, os_)
    # stage outputs report original doc ids, not the staged pass's compact ids
    d = _index(ix, False)
    _, _, st_d = d.search(Q, 4, 64, 10, stages=True)
    for key in ("sel_ids", "sel_score", "rerank_n", "cand_n"):
        np.testing.assert_array_equal(st[key], st_d[key], err_msg=key)


@pytest.mark.parametrize("persistent,graph", [(True, False), (False, True), (True, True)])
def test_host_residual_persistent_and_graph_flags(c1, persistent, graph):
    """The work-queue MaxSim runs on the staged pass; graph replay is not used
    for such an index (results equal the eager device-resident search)."""
    corpus, ix, Q = c1
    dev, host = _index(ix, False), _index(ix, True)
    qt, qp = _ragged(Q, 12)
    kw = dict(precision="bf16", rescore=64, q_ntok=qt, q_nprobe=qp, batch=16)
    ref_ids, ref_sc = dev.search(Q, 8, 256, 10, **kw)
    for _ in range(2):
        ids, sc = host.search(Q, 8, 256, 10, persistent=persistent, graph=graph, **kw)
        np.testing.assert_array_equal(ids, ref_ids)
        np.testing.assert_array_equal(sc, ref_sc)


def test_host_residual_device_buffers_and_short_lists(small):
    """On-device queries/outputs; one-token queries with topk above their
    candidate count (the UINT32_MAX fill must not be remapped); then a second
    search with fewer queries reusing the staging buffers."""
    import torch
    corpus, ix, Q = small
    dev, host = _index(ix, False), _index(ix, True)
    for n_q, nprobe, R, k, one_tok in ((Q.shape[0], 1, 512, 512, True), (3, 2, 64, 10, False)):
        qt = np.ones(n_q, np.uint32) if one_tok else None
        ref_ids, ref_sc = dev.search(Q[:n_q], nprobe, R, k, precision="fp32_mfma", q_ntok=qt)
        if one_tok:
            assert (ref_ids == 0xFFFFFFFF).any()
        Qd = torch.from_numpy(np.ascontiguousarray(Q[:n_q])).cuda()
        oi = torch.empty((n_q, k), dtype=torch.int32, device="cuda")
        os_ = torch.empty((n_q, k), dtype=torch.float32, device="cuda")
        host.search_device(Qd, oi, os_, nprobe, R, k, precision="fp32_mfma", q_ntok=qt, batch=4)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(oi.cpu().numpy().view(np.uint32), ref_ids)
        np.testing.assert_array_equal(os_.cpu().numpy(), ref_sc)


def test_host_residual_timings_count_original_docs(c1):
    corpus, ix, Q = c1
    dev, host = _index(ix, False), _index(ix, True)
    _, _, td = dev.search(Q, 8, 256, 10, precision="bf16", batch=32, timings=True)
    _, _, th = host.search(Q, 8, 256, 10, precision="bf16", batch=32, timings=True)
    for key in ("rerank_tokens", "candidates", "rerank_pairs", "n_pass"):
        assert td[key] == th[key], key
    assert td["staged_docs"] == 0 and td["staging_ms"] == 0.0
    # each pass stages its distinct docs once: at most one per reranked (query, doc)
    assert 0 < th["staged_docs"] <= th["candidates"]
    assert th["staged_tokens"] <= th["rerank_tokens"] and th["staging_ms"] > 0.0


def test_host_residual_unsupported_ops(small):
    import mvr
    corpus, ix, Q = small
    g = _index(ix, True)
    with pytest.raises(mvr.MvrError, match="device-resident"):
        g.decompress(np.array([0, 1], np.uint32))


def test_load_ex_residual_on_host(tmp_path, small):
    import mvr
    from test_cli import write_dataset
    corpus, ix, Q = small
    write_dataset(str(tmp_path), ix, Q, list(range(Q.shape[0])), ix.n_c)
    d = os.path.join(str(tmp_path), "Index", "ds", f"igp-n_centroid_{ix.n_c}-n_bit_2")
    g = mvr.Index.load(d, residual_on_host=True)
    ids, sc = g.search(Q, 4, 64, 10)
    oi, os_ = ix.search(Q, 4, 64, 10)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(sc, os_)


def test_host_residual_with_learned_probing(c1):
    """The probe model's per-query nprobe and rerank depth are decided in the
    pass's front half; with host-resident residuals (pipelined passes) the
    answers equal the device-resident index's."""
    from mvr.probe import ProbeModel
    corpus, ix, Q = c1
    dev, host = _index(ix, False), _index(ix, True)
    models = [ProbeModel(seed=6), ProbeModel(seed=6)]  # attached models must outlive the searches
    for m, g in zip(models, (dev, host)):
        m.attach(g, conf=0.6, fallback=16, r_min=32)
    kw = dict(precision="fp32_mfma", learned=True, batch=16)
    ref_ids, ref_sc = dev.search(Q, 4, 256, 10, **kw)
    ids, sc = host.search(Q, 4, 256, 10, **kw)
    np.testing.assert_array_equal(ids, ref_ids)
    np.testing.assert_array_equal(sc, ref_sc)
    dev.close()
    host.close()
    del models
