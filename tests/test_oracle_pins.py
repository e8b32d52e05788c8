"""Pins of oracle functions that no reference artefact fixes directly.

* bf16_round: the checker of BF16 centroid parity; checked against torch's
  float32 -> bfloat16 -> float32 round trip (round to nearest even) on random
  values over the whole exponent range, exact ties, subnormals, infinities and
  NaN.
* assign_ip / kmeans: the index-build helpers (argmax inner product with the
  lowest id on ties, Lloyd iterations with argmin-L2 assignment, means in
  double, empty clusters kept, L2 normalisation at the end); checked against a
  float64 numpy brute force on small inputs whose assignments have clear gaps.
* fp8_tile_maxima / centroid_topk_fp8: the FP8 probe's model; tile maxima
  against a per-element loop over decoded e4m3 values, and the probe against
  the exact top-P when its tiles cover every centroid.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def _torch_bf16(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def test_bf16_round_matches_torch_on_random_bits():
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2**32, 200_000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    got, ref = O.bf16_round(x), _torch_bf16(x)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_bf16_round_ties_to_even_known_answers():
    cases = {
        0x3F808000: 0x3F800000,  # 1 + 2^-8 exactly halfway, even neighbour below
        0x3F818000: 0x3F820000,  # halfway, even neighbour above
        0x3F808001: 0x3F810000,  # just above halfway
        0x3F807FFF: 0x3F800000,  # just below halfway
        0xBF808000: 0xBF800000,  # negative tie
        0x7F7FFFFF: 0x7F800000,  # largest float rounds to +inf
        0x00008000: 0x00000000,  # subnormal tie to even (zero)
        0x00018000: 0x00020000,  # subnormal tie to even (above)
        0x00000001: 0x00000000,  # smallest subnormal
        0x80000001: 0x80000000,  # keeps the sign of zero
    }
    x = np.array(list(cases), np.uint32).view(np.float32)
    want = np.array(list(cases.values()), np.uint32)
    np.testing.assert_array_equal(O.bf16_round(x).view(np.uint32), want)
    np.testing.assert_array_equal(_torch_bf16(x).view(np.uint32), want)


def test_bf16_round_specials():
    x = np.array([np.inf, -np.inf, np.nan, -np.nan, 0.0, -0.0], np.float32)
    nan_payload = np.array([0x7F800001, 0xFF800001, 0x7FFFFFFF], np.uint32).view(np.float32)
    got = O.bf16_round(np.concatenate([x, nan_payload]))
    assert got[0] == np.inf and got[1] == -np.inf
    assert np.isnan(got[2]) and np.isnan(got[3]) and np.isnan(got[6:]).all()
    assert got[4] == 0.0 and not np.signbit(got[4]) and got[5] == 0.0 and np.signbit(got[5])
    # every result is representable in bfloat16 (low 16 bits clear)
    assert (got.view(np.uint32) & 0xFFFF == 0).all()


def test_bf16_round_subnormal_range_matches_torch():
    bits = np.arange(0, 0x00800000, 977, dtype=np.uint32)
    x = np.concatenate([bits, bits | 0x80000000]).astype(np.uint32).view(np.float32)
    np.testing.assert_array_equal(O.bf16_round(x).view(np.uint32), _torch_bf16(x).view(np.uint32))


def _separated(rng, n_c, per, dim=16, spread=0.02):
    centres = rng.standard_normal((n_c, dim))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    x = np.repeat(centres, per, axis=0) + spread * rng.standard_normal((n_c * per, dim))
    return x.astype(np.float32), centres


def test_assign_ip_matches_float64_argmax():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((500, 32)).astype(np.float32)
    C_ = rng.standard_normal((40, 32)).astype(np.float32)
    got = O.assign_ip(x, C_)
    s = x.astype(np.float64) @ C_.astype(np.float64).T
    srt = np.sort(s, axis=1)
    clear = srt[:, -1] - srt[:, -2] > 1e-4  # away from near-ties
    assert clear.mean() > 0.95
    np.testing.assert_array_equal(got[clear], np.argmax(s, axis=1)[clear])
    # the chosen centroid is always within rounding of the maximum
    assert np.all(s[np.arange(len(x)), got] >= srt[:, -1] - 1e-5)


def test_assign_ip_ties_take_lowest_index():
    rng = np.random.default_rng(4)
    C_ = rng.standard_normal((6, 8)).astype(np.float32)
    C_[4] = C_[1]  # duplicate centroid: exact tie, lower id wins
    C_[5] = C_[1]
    x = np.stack([C_[1] * 3.0, C_[0], C_[4]]).astype(np.float32)
    got = O.assign_ip(x, C_)
    assert got[0] == 1 and got[2] == 1


def _kmeans_f64(x, n_c, n_iter):
    """Lloyd in float64 with the oracle's initialisation and empty-cluster rule."""
    n = x.shape[0]
    C_ = x[[(j * n) // n_c for j in range(n_c)]].astype(np.float64)
    X = x.astype(np.float64)
    for _ in range(n_iter):
        d2 = ((X[:, None, :] - C_[None, :, :]) ** 2).sum(-1)
        a = np.argmin(d2, axis=1)
        for c in range(n_c):
            m = a == c
            if m.any():
                C_[c] = X[m].mean(0).astype(np.float32)  # the oracle stores float32 means
    C_ = C_ / np.linalg.norm(C_, axis=1, keepdims=True)
    return C_


@pytest.mark.parametrize("n_iter", [1, 3, 10])
def test_kmeans_matches_float64_lloyd(n_iter):
    rng = np.random.default_rng(5 + n_iter)
    x, _ = _separated(rng, n_c=12, per=20)
    perm = rng.permutation(x.shape[0])
    x = x[perm]
    got = O.kmeans(x, 12, n_iter)
    ref = _kmeans_f64(x, 12, n_iter)
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6)
    np.testing.assert_allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-6)


def test_kmeans_known_answer_two_clusters_and_empty_cluster():
    # points on two tight groups around (1, 0) and (0, 1); initial centroids are
    # points 0 and 2 (floor(j n / n_c)), one from each group
    x = np.array([[1.0, 0.1], [1.0, -0.1], [0.1, 1.0], [-0.1, 1.0]], np.float32)
    got = O.kmeans(x, 2, 5)
    np.testing.assert_allclose(got, [[1.0, 0.0], [0.0, 1.0]], atol=1e-7)
    # three centroids on two groups: initial points 0, 1, 2; centroid 1 (a
    # copy of point 1) keeps its point, so no cluster is empty; with four
    # identical points every later centroid stays empty and keeps its start
    y = np.array([[2.0, 0.0]] * 4, np.float32)
    got = O.kmeans(y, 3, 4)
    np.testing.assert_allclose(got, [[1.0, 0.0]] * 3, atol=1e-7)


def test_assign_bias_matches_float64_argmin_l2():
    """With bias |C_c|^2 / 2 the biased argmax is the L2 nearest centroid."""
    rng = np.random.default_rng(8)
    x = rng.standard_normal((400, 24)).astype(np.float32)
    C_ = (rng.standard_normal((30, 24)) * rng.uniform(0.5, 2.0, (30, 1))).astype(np.float32)
    bias = (0.5 * (C_.astype(np.float64) ** 2).sum(1)).astype(np.float32)
    got = O.assign_bias(x, C_, bias)
    d2 = ((x.astype(np.float64)[:, None, :] - C_.astype(np.float64)[None]) ** 2).sum(-1)
    srt = np.sort(d2, axis=1)
    clear = srt[:, 1] - srt[:, 0] > 1e-3
    assert clear.mean() > 0.95
    np.testing.assert_array_equal(got[clear], np.argmin(d2, axis=1)[clear])
    # a zero bias is the plain inner-product assignment
    np.testing.assert_array_equal(O.assign_bias(x, C_, np.zeros(30, np.float32)), O.assign_ip(x, C_))


def test_fp8_tile_maxima_brute_force():
    """Tile maxima of the FP8 probe against a loop over every (row, centroid)
    with e4m3 values decoded from their bit patterns (a partial last tile)."""
    rng = np.random.default_rng(11)
    Q = rng.standard_normal((3, 128)).astype(np.float32)
    C = rng.standard_normal((70, 128)).astype(np.float32)
    tm = O.fp8_tile_maxima(Q, C)
    assert tm.shape == (3, 3)

    def enc(x):
        amax = np.float32(np.abs(x).max())
        inv = np.float32(448.0) / amax
        return [float(O.e4m3_round(np.array([np.float32(v) * inv], np.float32))[0]) for v in x], \
            float(amax / np.float32(448.0))
    for r in range(3):
        q8, sq = enc(Q[r])
        best = [-np.inf] * 3
        for c in range(70):
            c8, sc = enc(C[c])
            v = sq * sc * sum(a * b for a, b in zip(q8, c8))
            best[c // 32] = max(best[c // 32], v)
        np.testing.assert_allclose(tm[r], best, rtol=1e-12)


def test_fp8_probe_with_every_tile_is_the_exact_topk():
    """With P' covering every tile the FP8 probe is the exact FP32 top-P, and
    its scores are the chain's bits; with fewer tiles each id lies in one of
    the P' tiles of highest FP8 maximum."""
    rng = np.random.default_rng(12)
    C = rng.standard_normal((200, 128)).astype(np.float32)
    C /= np.linalg.norm(C, axis=1, keepdims=True)
    Q = rng.standard_normal((8, 128)).astype(np.float32)
    ids, s, amb = O.centroid_topk_fp8(Q, C, 4, probe_tiles=2)      # 7 tiles <= 8: all
    ei, es = O.centroid_topk(Q, C, 4)
    np.testing.assert_array_equal(ids, ei)
    np.testing.assert_array_equal(s, es)
    assert not amb.any()
    ids, s, _ = O.centroid_topk_fp8(Q, C, 2, probe_tiles=1)         # 2 of 7 tiles
    tm = O.fp8_tile_maxima(Q, C)
    for r in range(8):
        top2 = set(np.argsort(-tm[r])[:2].tolist())
        assert {int(i) // 32 for i in ids[r]} <= top2
