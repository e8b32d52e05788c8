"""BruteForceProgressive.DocRetrieval on the GPU (libmvr's
mvr_maxsim_scores_host): all-pairs FP32 MaxSim scores bit-exact against the
oracle (oracle.maxsim_scores_f32), searchKNN order and fill, and the
groundtruth driver's batch-and-merge loop (script/data/groundtruth.py:40-109)
equal to one exhaustive pass and to the FP64 exhaustive oracle's ids."""
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
