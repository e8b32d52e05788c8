"""Pins of oracle.maxsim_scores_f32 (the exhaustive module's all-pairs FP32
MaxSim, GPUComputeScore.hpp:196-246): float64 brute force, a dyadic known
answer where FP32 is exact, and agreement with the FP64 exhaustive top-k."""
import numpy as np

from oracle import oracle as O


def test_maxsim_scores_float64_brute_force():
    rng = np.random.default_rng(1)
    lens = np.array([1, 5, 130, 2, 9], np.uint32)
    X = rng.standard_normal((int(lens.sum()), 128)).astype(np.float32)
    Q = rng.standard_normal((3, 7, 128)).astype(np.float32)
    got = O.maxsim_scores_f32(Q, X, lens)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    for q in range(3):
        for i in range(lens.size):
            ref = (Q[q].astype(np.float64) @ X[off[i]:off[i + 1]].T.astype(np.float64)).max(1).sum()
            assert abs(got[q, i] - ref) <= 1e-5 * max(1.0, abs(ref))


def test_maxsim_scores_known_answer():
    # one-hot vectors: every dot is exact, the per-token max picks the match
    Q = np.zeros((1, 3, 128), np.float32)
    Q[0, 0, 0], Q[0, 1, 1], Q[0, 2, 2] = 1.0, 2.0, -1.0
    X = np.zeros((4, 128), np.float32)
    X[0, 0], X[1, 1], X[2, 2], X[3, 3] = 0.5, 0.25, 4.0, 1.0
    got = O.maxsim_scores_f32(Q, X, np.array([2, 2], np.uint32))
    # item 0 = tokens 0,1: maxima (0.5, 0.5, 0) ; item 1 = tokens 2,3: maxima (0, 0, -0) -> -4 vs 0
    np.testing.assert_array_equal(got, np.array([[1.0, 0.0]], np.float32))


def test_maxsim_scores_top_k_matches_fp64_exhaustive():
    rng = np.random.default_rng(2)
    lens = rng.integers(3, 40, 300).astype(np.uint32)
    X = rng.standard_normal((int(lens.sum()), 128)).astype(np.float32)
    Q = rng.standard_normal((4, 32, 128)).astype(np.float32)
    s = O.maxsim_scores_f32(Q, X, lens)
    ids, _ = O.exhaustive_f64(Q, X, lens, 10)
    top = np.argsort(-s, axis=1, kind="stable")[:, :10]
    np.testing.assert_array_equal(top, ids.astype(np.int64))
