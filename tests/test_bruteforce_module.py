"""The BruteForceProgressive module's host-side surface
(src/app/BruteForceProgressive.cpp:16-33): constructor checks and the static
merge_result against a Python model of the reference's min-heap merge
(src/impl/BruteForceProgressive.hpp:258-346).  computeScore needs a GPU
(tests/test_gpu_bruteforce.py)."""
import heapq

import numpy as np
import pytest


@pytest.fixture(scope="module")
def mod():
    import mvr
    return mvr.bruteforce_module()


def _heap_merge(fd, fi, d, i):
    """The reference's merge: a min-heap seeded with the earlier list, a new
    entry replaces the minimum only when strictly greater; pops ascending."""
    out_d, out_i = [], []
    for q in range(fd.shape[0]):
        h = [(float(a), k, int(b)) for k, (a, b) in enumerate(zip(fd[q], fi[q]))]
        heapq.heapify(h)
        for k, (a, b) in enumerate(zip(d[q], i[q])):
            if a > h[0][0]:
                heapq.heapreplace(h, (float(a), fd.shape[1] + k, int(b)))
        srt = sorted(h, key=lambda e: (e[0], -e[1]))
        out_d.append([e[0] for e in srt])
        out_i.append([e[2] for e in srt])
    return np.array(out_d, np.float32), np.array(out_i, np.uint32)


def test_merge_result_matches_heap_model(mod):
    rng = np.random.default_rng(0)
    for topk in (1, 5, 10):
        fd = np.sort(rng.standard_normal((7, topk)).astype(np.float32), 1)
        d = np.sort(rng.standard_normal((7, topk)).astype(np.float32), 1)
        fi = rng.integers(0, 1000, (7, topk)).astype(np.uint32)
        i = rng.integers(1000, 2000, (7, topk)).astype(np.uint32)
        got_d, got_i = mod.DocRetrieval.merge_result(fd, fi, d, i)
        ref_d, ref_i = _heap_merge(fd, fi, d, i)
        np.testing.assert_array_equal(got_d, ref_d)
        np.testing.assert_array_equal(got_i, ref_i)
        assert np.all(np.diff(got_d, axis=1) >= 0)       # ascending, as the heap pops
        # equal scores in both lists: the reference heap leaves the choice among
        # them open; the kept score values are fixed
        d2 = fd.copy()
        got_d, _ = mod.DocRetrieval.merge_result(fd, fi, d2, i)
        ref_d, _ = _heap_merge(fd, fi, d2, i)
        np.testing.assert_array_equal(got_d, ref_d)


def test_merge_result_first_call_passes_through(mod):
    d = np.ones((2, 3), np.float32)
    i = np.arange(6, dtype=np.uint32).reshape(2, 3)
    od, oi = mod.DocRetrieval.merge_result(np.array([]), np.array([]), d, i)
    np.testing.assert_array_equal(od, d)
    np.testing.assert_array_equal(oi, i)


def test_constructor_checks(mod):
    with pytest.raises(RuntimeError, match="dimension should be 3"):
        mod.DocRetrieval(query_l=np.zeros((4, 128), np.float32), vec_dim=128)
    with pytest.raises(RuntimeError, match="pre-assignment"):
        mod.DocRetrieval(query_l=np.zeros((2, 4, 64), np.float32), vec_dim=128)
    r = mod.DocRetrieval(query_l=np.zeros((2, 4, 128), np.float32), vec_dim=128)
    with pytest.raises(RuntimeError, match="externalID"):
        r.computeScore(item_l=[np.zeros((3, 128), np.float32)], itemID_l=[0, 1])
