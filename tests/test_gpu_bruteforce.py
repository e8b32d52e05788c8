"""BruteForceProgressive.DocRetrieval on the GPU (libmvr's
mvr_maxsim_scores_host): all-pairs FP32 MaxSim scores bit-exact against the
oracle (oracle.maxsim_scores_f32), searchKNN order and fill, and the
groundtruth driver's batch-and-merge loop (script/data/groundtruth.py:40-109)
equal to one exhaustive pass and to the FP64 exhaustive oracle's ids; the
FP32 MFMA precision (mvr_maxsim_scores_host_ex) within 1e-5 of the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mod():
    import mvr
    return mvr.bruteforce_module()


def _items(rng, lens):
    return [rng.standard_normal((int(L), 128)).astype(np.float32) for L in lens]


@pytest.mark.parametrize("n_q,T", [(5, 32), (9, 20), (4, 1)])
def test_scores_bitexact(mod, n_q, T):
    rng = np.random.default_rng(n_q * 100 + T)
    # short items (many per column block), long items (several 128-token tiles), length 1
    lens = np.concatenate([rng.integers(1, 12, 150), [700, 1, 513, 129], rng.integers(20, 200, 60)]).astype(np.uint32)
    items = _items(rng, lens)
    Q = rng.standard_normal((n_q, T, 128)).astype(np.float32)
    r = mod.DocRetrieval(query_l=Q, vec_dim=128)
    r.computeScore(item_l=items, itemID_l=list(range(len(items))))
    k = 17
    dist, ids = r.searchKNN(topk=k)
    ref = O.maxsim_scores_f32(Q, np.concatenate(items), lens)
    for q in range(n_q):
        order = sorted(range(len(items)), key=lambda i: (-ref[q, i], i))[:k][::-1]
        np.testing.assert_array_equal(ids[q], np.array(order, np.uint32))
        np.testing.assert_array_equal(dist[q], ref[q, order])
    r.finish_compute()


def test_searchknn_fill_when_topk_exceeds_items(mod):
    rng = np.random.default_rng(3)
    items = _items(rng, [3, 4, 5])
    r = mod.DocRetrieval(rng.standard_normal((2, 8, 128)).astype(np.float32), 128)
    r.computeScore(items, [70, 71, 72])
    dist, ids = r.searchKNN(5)
    assert np.all(ids[:, 3:] == 0xFFFFFFFF) and np.all(dist[:, 3:] == -np.finfo(np.float32).max)
    assert set(ids[0, :3].tolist()) == {70, 71, 72} and np.all(np.diff(dist[:, :3], axis=1) >= 0)


def test_groundtruth_loop_equals_one_pass_and_fp64(mod):
    """Batches of items scored one after another and merged, as the
    groundtruth script does, give the same top-k as one exhaustive pass."""
    rng = np.random.default_rng(9)
    lens = rng.integers(5, 60, 900).astype(np.uint32)
    items = _items(rng, lens)
    Q = rng.standard_normal((6, 32, 128)).astype(np.float32)
    k = 10
    r = mod.DocRetrieval(query_l=Q, vec_dim=128)
    final_d, final_i = np.array([]), np.array([])
    for b0 in range(0, len(items), 250):
        b1 = min(len(items), b0 + 250)
        r.computeScore(items[b0:b1], list(range(b0, b1)))
        d, i = r.searchKNN(k)
        final_d, final_i = mod.DocRetrieval.merge_result(final_d, final_i, d, i)
    r.finish_compute()
    one = mod.DocRetrieval(query_l=Q, vec_dim=128)
    one.computeScore(items, list(range(len(items))))
    d1, i1 = one.searchKNN(k)
    np.testing.assert_array_equal(final_i, i1)
    np.testing.assert_array_equal(final_d, d1)
    g_ids, g_s = O.exhaustive_f64(Q, np.concatenate(items), lens, k)
    np.testing.assert_array_equal(i1[:, ::-1], g_ids.astype(np.uint32))     # random data: no near ties
    np.testing.assert_allclose(d1[:, ::-1], g_s, rtol=1e-5)


@pytest.mark.parametrize("n_q,T", [(5, 32), (9, 20), (4, 1)])
def test_mfma_scores_within_tolerance(mod, n_q, T):
    """precision="fp32_mfma": the dots on FP32 MFMA, each score within 1e-5
    relative of the FMA-chain oracle (items spanning tiles and column blocks,
    length-1 items), and the same top-k up to ties inside that tolerance."""
    rng = np.random.default_rng(n_q * 100 + T + 7)
    lens = np.concatenate([rng.integers(1, 12, 150), [700, 1, 513, 129], rng.integers(20, 200, 60)]).astype(np.uint32)
    items = _items(rng, lens)
    Q = rng.standard_normal((n_q, T, 128)).astype(np.float32)
    r = mod.DocRetrieval(query_l=Q, vec_dim=128, precision="fp32_mfma")
    r.computeScore(item_l=items, itemID_l=list(range(len(items))))
    k = 17
    dist, ids = r.searchKNN(topk=k)
    ref = O.maxsim_scores_f32(Q, np.concatenate(items), lens)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    for q in range(n_q):
        got = ref[q, ids[q]]
        np.testing.assert_allclose(dist[q], got, rtol=1e-5, atol=1e-5 * scale[q, 0])
        kth = np.sort(ref[q])[::-1][k - 1]
        assert (ref[q, ids[q]] >= kth - 2e-5 * scale[q, 0]).all()
    r.finish_compute()


def test_capi_exhaustive_precisions():
    """mvr_maxsim_scores_host_ex through the C ABI: MVR_FP32 equals
    mvr_maxsim_scores_host bit for bit, MVR_FP32_MFMA is within 1e-5, and
    other precisions are refused."""
    import ctypes as C
    from mvr._capi import lib
    rng = np.random.default_rng(21)
    lens = rng.integers(1, 90, 300).astype(np.uint32)
    X = rng.standard_normal((int(lens.sum()), 128)).astype(np.float32)
    Q = rng.standard_normal((3, 32, 128)).astype(np.float32)
    out = {}
    for name, prec in (("plain", None), ("fp32", 0), ("mfma", 3)):
        s = np.zeros((3, lens.size), np.float32)
        args = (Q.ctypes.data, 3, 32, 128, X.ctypes.data, lens.ctypes.data, lens.size, s.ctypes.data, 0)
        st = lib().mvr_maxsim_scores_host(*args) if prec is None else lib().mvr_maxsim_scores_host_ex(*args, prec)
        assert st == 0, lib().mvr_maxsim_last_error()
        out[name] = s
    np.testing.assert_array_equal(out["plain"].view(np.uint32), out["fp32"].view(np.uint32))
    np.testing.assert_allclose(out["mfma"], out["fp32"], rtol=1e-5, atol=1e-5 * np.abs(out["fp32"]).max())
    s = np.zeros((3, lens.size), np.float32)
    assert lib().mvr_maxsim_scores_host_ex(Q.ctypes.data, 3, 32, 128, X.ctypes.data, lens.ctypes.data, lens.size,
                                           s.ctypes.data, 0, 1) != 0
