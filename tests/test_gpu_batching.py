"""Batched execution on the GPU: the persistent work-queue MaxSim + top-k
kernel, its tuning settings and hipGraph replay of device passes.  Every
variant must return exactly what the per-query launch path returns (same
MaxSim arithmetic, same (score desc, id asc) order), which the parity tests
tie to the CPU oracle.
"""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from mvr._capi import MvrError  # noqa: E402


def _index(ix):
    import mvr
    return mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)


def _ragged(Q, seed):
    rng = np.random.default_rng(seed)
    qt = rng.integers(8, Q.shape[1] + 1, Q.shape[0]).astype(np.uint32)
    qp = rng.choice(np.array([2, 4, 8, 16, 32], np.uint32), Q.shape[0])
    return qt, qp


@pytest.mark.parametrize("prec,rescore", [("fp32_mfma", 0), ("bf16", 0), ("bf16", 64), ("fp8", 64)])
def test_persistent_matches_launch_path_every_setting(c1, prec, rescore):
    corpus, ix, Q = c1
    g = _index(ix)
    qt, qp = _ragged(Q, 1)
    ref_ids, ref_sc = g.search(Q, 8, 256, 10, precision=prec, rescore=rescore, q_ntok=qt, q_nprobe=qp,
                               batch=24)
    n_set = 0
    for wpc, tps in itertools.product((1, 2, 3, 4), (1, 2)):
        try:
            g.set_tuning(prec, wpc, tps)
        except MvrError:
            # one workgroup per CU is always resident; the FP32 kernel has one tile setting
            assert wpc > 1 or (prec == "fp32_mfma" and tps == 2)
            continue  # more workgroups per CU than that kernel keeps resident
        n_set += 1
        ids, sc = g.search(Q, 8, 256, 10, precision=prec, rescore=rescore, q_ntok=qt, q_nprobe=qp,
                           batch=24, persistent=True)
        np.testing.assert_array_equal(ids, ref_ids, err_msg=f"{prec} wpc={wpc} tps={tps}")
        np.testing.assert_array_equal(sc, ref_sc, err_msg=f"{prec} wpc={wpc} tps={tps}")
    assert n_set >= 2
    with pytest.raises(MvrError):
        g.set_tuning(prec, 9, 1)


def test_autotuner_picks_a_valid_setting(small):
    corpus, ix, Q = small
    g = _index(ix)
    g.set_maxsim_order("query")  # the per-query persistent kernel in the first phase
    assert g.get_tuning("bf16") is None
    ref = g.search(Q, 4, 64, 10, precision="bf16")
    got = g.search(Q, 4, 64, 10, precision="bf16", persistent=True)
    wpc, tps = g.get_tuning("bf16")
    assert 1 <= wpc <= 8 and tps in (1, 2)
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
    # with a rescore the FP32 phase gets a setting of its own
    assert g.get_tuning("fp32_mfma") is None
    g2 = _index(ix)
    g2.set_maxsim_order("query")
    ref = g2.search(Q, 4, 64, 10, precision="bf16", rescore=32)
    got = g2.search(Q, 4, 64, 10, precision="bf16", rescore=32, persistent=True)
    assert g2.get_tuning("bf16") is not None and g2.get_tuning("fp32_mfma") is not None
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
@pytest.mark.parametrize("order", ["doc", "auto"])
@pytest.mark.parametrize("rescore", [0, 64])
def test_persistent_doc_major_first_phase(c1, rescore, order, prec):
    """The persistent BF16/FP8 phase doc-major (a work queue of docs, then the
    per-query top-R2 or top-k), forced or as the autotuner's choice: same
    results as the launch path, eagerly and replayed from a graph.  Forced,
    only the FP32 phase has a persistent setting to tune; automatic, the
    tuner records its choice of order."""
    corpus, ix, Q = c1
    g = _index(ix)
    g.set_maxsim_order(order)
    qt, qp = _ragged(Q, 3)
    kw = dict(precision=prec, rescore=rescore, q_ntok=qt, q_nprobe=qp)
    ref_ids, ref_sc = g.search(Q, 8, 256, 10, **kw)
    for graph in (False, True, True, True):
        ids, sc = g.search(Q, 8, 256, 10, persistent=True, graph=graph, **kw)
        np.testing.assert_array_equal(ids, ref_ids, err_msg=f"graph={graph}")
        np.testing.assert_array_equal(sc, ref_sc, err_msg=f"graph={graph}")
    if order == "doc":
        assert g.get_tuning(prec) is None
    else:
        t = g.get_tuning_ex(prec)
        assert t["order"] in ("query", "doc") and t["wgs_per_cu"] >= 1 and t["tiles"] in (1, 2)
    assert (g.get_tuning("fp32_mfma") is not None) == (rescore > 0)
    # the query order uses the per-query kernel whatever the tuner chose
    g.set_maxsim_order("query")
    ids, sc = g.search(Q, 8, 256, 10, persistent=True, **kw)
    np.testing.assert_array_equal(ids, ref_ids)
    np.testing.assert_array_equal(sc, ref_sc)
    g.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_graph_replays_match_eager(c1, persistent):
    """Each pass captured once and replayed: three rounds of replays with
    eager searches and synchronisations in between, every replay checked."""
    corpus, ix, Q = c1
    g = _index(ix)
    qt, qp = _ragged(Q, 2)
    kw = dict(precision="bf16", rescore=64, q_ntok=qt, q_nprobe=qp, batch=16)
    ref_ids, ref_sc = g.search(Q, 8, 256, 10, **kw)
    for _ in range(3):
        ids, sc = g.search(Q, 8, 256, 10, graph=True, persistent=persistent, **kw)
        np.testing.assert_array_equal(ids, ref_ids)
        np.testing.assert_array_equal(sc, ref_sc)
        e_ids, e_sc = g.search(Q[:8], 4, 64, 10, precision="fp32_mfma")  # eager work in between
        assert (e_ids != 0xFFFFFFFF).all()


def test_graph_on_device_buffers(c1):
    import torch
    corpus, ix, Q = c1
    g = _index(ix)
    ref_ids, ref_sc = g.search(Q, 4, 256, 10, precision="fp32_mfma")
    Qd = torch.from_numpy(np.ascontiguousarray(Q)).cuda()
    for _ in range(3):
        oi = torch.empty((Q.shape[0], 10), dtype=torch.int32, device="cuda")
        os_ = torch.empty((Q.shape[0], 10), dtype=torch.float32, device="cuda")
        g.search_device(Qd, oi, os_, 4, 256, 10, precision="fp32_mfma", batch=32, graph=True,
                        persistent=True, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(oi.cpu().numpy().view(np.uint32), ref_ids)
        np.testing.assert_array_equal(os_.cpu().numpy(), ref_sc)


def test_persistent_rejects_unsupported_settings(small):
    """The persistent kernel has no exact-FP32 variant and selects at most 256:
    such requests are errors, not a silent switch to per-pass launches."""
    from mvr._capi import MvrError
    corpus, ix, Q = small
    g = _index(ix)
    for kw in (dict(precision="fp32"), dict(precision="fp32_mfma", k=300),
               dict(precision="bf16", rescore=300, k=10)):
        k = kw.pop("k", 10)
        with pytest.raises(MvrError):
            g.search(Q, 4, 512, k, persistent=True, **kw)
    g.close()
