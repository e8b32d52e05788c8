"""Pins of the oracle's search stages against brute force in float64 and
hand-built cases.

Reference: centroid scoring GenerateVQScoreCandidatePP.hpp:137-174 (S' = Q C^T
+ 1, top-P), filter FilterPP.hpp:99-142,366-394 (max over probes per (doc,
token), sum over tokens, top-R), MaxSim RerankPP.hpp:115-161,458-471 and
Distance.hpp:200-218, final sort IGPGPURetrievalPP.hpp:153-161.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _rand_unit(rng, n, dim=128):
    x = rng.standard_normal((n, dim))
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


def test_sequential_fma_dot_matches_float64():
    rng = np.random.default_rng(0)
    a, b = _rand_unit(rng, 2)
    got = O.lib().orc_dot(a, b, 128)
    exact = float(a.astype(np.float64) @ b.astype(np.float64))
    assert abs(got - exact) <= 128 * np.finfo(np.float32).eps  # forward error bound of 128 FMAs


def test_centroid_topk_against_float64_bruteforce():
    rng = np.random.default_rng(1)
    Q = _rand_unit(rng, 40)
    C_ = _rand_unit(rng, 300)
    P = 8
    ids, s = O.centroid_topk(Q, C_, P)
    S64 = Q.astype(np.float64) @ C_.astype(np.float64).T + 1.0
    for r in range(Q.shape[0]):
        order = np.lexsort((np.arange(300), -S64[r]))
        assert (np.diff(s[r]) <= 0).all()
        np.testing.assert_allclose(s[r], S64[r, ids[r]], atol=1e-6)
        gap = S64[r, order[P - 1]] - S64[r, order[P]]
        if gap > 1e-5:
            assert set(ids[r].tolist()) == set(order[:P].tolist())


def test_centroid_topk_ties_lower_id_first():
    C_ = np.zeros((6, 128), np.float32)
    C_[[1, 3, 4], 0] = 1.0   # three equal best scores
    C_[5, 0] = 0.5
    Q = np.zeros((1, 128), np.float32)
    Q[0, 0] = 1.0
    ids, s = O.centroid_topk(Q, C_, 4)
    np.testing.assert_array_equal(ids[0], [1, 3, 4, 5])
    np.testing.assert_array_equal(s[0], [2.0, 2.0, 2.0, 1.5])


def _tiny_ivf(lists, n_c):
    sizes = np.array([len(lists.get(c, [])) for c in range(n_c)], np.uint32)
    off = np.zeros(n_c, np.uint64)
    off[1:] = np.cumsum(sizes[:-1])
    ids = np.array([d for c in range(n_c) for d in sorted(lists.get(c, []))], np.uint32)
    return off, sizes, ids


def test_filter_max_over_probes_sum_over_tokens_bruteforce():
    # 3 query tokens x 2 probes; lists overlap so docs are hit several times
    lists = {0: [0, 2, 5], 1: [2, 3], 2: [5, 6, 7], 3: [0, 7], 4: [9]}
    ivf = _tiny_ivf(lists, 5)
    top_ids = np.array([[0, 1], [2, 0], [3, 4]], np.uint32)
    top_s = np.array([[1.75, 1.5], [1.625, 1.25], [1.4, 1.1]], np.float32)
    cand, F = O.filter_all(top_ids, top_s, 2, ivf, 10)
    expect = {}
    for t in range(3):
        best = {}
        for j in range(2):
            for d in lists[int(top_ids[t, j])]:
                best[d] = max(best.get(d, -np.inf), float(top_s[t, j]))
        for d, v in best.items():
            expect[d] = expect.get(d, 0.0) + v
    assert sorted(cand.tolist()) == sorted(expect)
    for d, f in zip(cand, F):
        assert f == np.float32(expect[int(d)])   # dyadic values: the sums are exact
    # doc 0: token0 max(1.75 via list0) + token1 1.25 (list0) + token2 1.4 (list3)
    assert F[cand.tolist().index(0)] == np.float32(1.75 + 1.25 + 1.4)


def test_filter_respects_probe_count_and_fixed_point():
    lists = {0: [1], 1: [1, 2]}
    ivf = _tiny_ivf(lists, 2)
    top_ids = np.array([[0, 1]], np.uint32)
    top_s = np.array([[1.3, 1.2]], np.float32)
    cand, F = O.filter_all(top_ids, top_s, 1, ivf, 3)      # only the first probe
    assert cand.tolist() == [1]
    fx = np.floor(np.float64(np.float32(1.3)) * 2**32) / 2**32
    assert F[0] == np.float32(fx)


def test_filter_pp_clamps_token_maxima_at_zero_basic_keeps_first_hit():
    """PP: maxima start at 0 and take atomicMax of every hit, so a token whose
    probed scores are all negative adds 0 (FilterPP.hpp:136-141, :257-259);
    the basic variant adds the first hit's score as it is (Filter.hpp:44-59).
    Candidates are the same either way."""
    lists = {0: [1, 2], 1: [2, 3], 2: [3]}
    ivf = _tiny_ivf(lists, 3)
    # token 0: both probes negative; token 1: a positive and a negative probe
    top_ids = np.array([[0, 1], [2, 0]], np.uint32)
    top_s = np.array([[-0.5, -0.75], [0.25, -0.125]], np.float32)
    cand, F = O.filter_all(top_ids, top_s, 2, ivf, 4)
    assert cand.tolist() == [1, 2, 3]
    # doc1: t0 max(0,-0.5)=0, t1 max(0,-0.125)=0; doc2: 0 + 0; doc3: 0 + 0.25
    np.testing.assert_array_equal(F, np.array([0.0, 0.0, 0.25], np.float32))
    cb, Fb = O.filter_all(top_ids, top_s, 2, ivf, 4, basic=True)
    assert cb.tolist() == [1, 2, 3]
    # doc1: -0.5 + -0.125; doc2: -0.5 (rank 0 of t0) + -0.125; doc3: -0.75 + 0.25
    np.testing.assert_array_equal(Fb, np.array([-0.625, -0.625, -0.5], np.float32))


def test_select_top_orders_by_score_then_id():
    ids = np.array([7, 3, 9, 1, 4], np.uint32)
    s = np.array([0.5, 0.9, 0.5, 0.9, 0.1], np.float32)
    oi, os_ = O.select_top(ids, s, 4)
    np.testing.assert_array_equal(oi, [1, 3, 7, 9])
    np.testing.assert_array_equal(os_, np.float32([0.9, 0.9, 0.5, 0.5]))
    oi, _ = O.select_top(ids, s, 10)
    assert oi.size == 5


def _tiny_index(rng, n_doc=12, n_c=8):
    dim = 128
    doclens = rng.integers(1, 9, n_doc).astype(np.uint32)
    n_tok = int(doclens.sum())
    C_ = _rand_unit(rng, n_c)
    codes = rng.integers(0, n_c, n_tok).astype(np.uint32)
    res = rng.integers(0, 256, (n_tok, 32)).astype(np.uint8)
    w = np.array([-0.05, -0.01, 0.01, 0.05], np.float32)
    return O.Index(doclens, codes, res, C_, w)


def test_maxsim_against_float64_bruteforce():
    rng = np.random.default_rng(3)
    ix = _tiny_index(rng)
    Q = _rand_unit(rng, 32)
    docs = np.arange(ix.n_doc, dtype=np.uint32)
    got = ix.rerank(Q, docs)
    for d in docs:
        X = ix.decompress_doc(int(d)).astype(np.float64)
        ref = (Q.astype(np.float64) @ X.T).max(axis=1).sum()
        assert abs(got[d] - ref) <= 1e-5 * abs(ref) + 1e-5
    clipped = ix.rerank(-Q, docs, clip_zero=True)
    for d in docs:
        X = ix.decompress_doc(int(d)).astype(np.float64)
        ref = np.maximum((-Q.astype(np.float64) @ X.T).max(axis=1), 0).sum()
        assert abs(clipped[d] - ref) <= 1e-5 * abs(ref) + 1e-5


def test_search_end_to_end_exhaustive_when_probing_everything():
    """With every centroid probed and R = n_doc, the pipeline is exhaustive
    MaxSim over the decompressed embeddings."""
    rng = np.random.default_rng(4)
    ix = _tiny_index(rng, n_doc=30, n_c=6)
    Q = _rand_unit(rng, 3 * 8).reshape(3, 8, 128)
    ids, s = ix.search(Q, 6, 30, 5)
    X = np.concatenate([ix.decompress_doc(d) for d in range(ix.n_doc)])
    gi, gs = O.exhaustive_f64(Q, X, ix.doclens, 5)
    for q in range(3):
        np.testing.assert_allclose(s[q], gs[q], rtol=1e-5)
        # equal up to float32-level ties
        assert set(ids[q].tolist()) == set(gi[q].tolist()) or np.ptp(gs[q][-2:]) < 1e-5


def test_exhaustive_f64_bruteforce_python():
    rng = np.random.default_rng(5)
    doclens = np.array([3, 1, 4, 2], np.uint32)
    X = _rand_unit(rng, int(doclens.sum()))
    Q = _rand_unit(rng, 2 * 4).reshape(2, 4, 128)
    ids, s = O.exhaustive_f64(Q, X, doclens, 4)
    off = np.concatenate([[0], np.cumsum(doclens)]).astype(int)
    for q in range(2):
        sc = [max_dot_sum(Q[q], X[off[d]:off[d + 1]]) for d in range(4)]
        order = sorted(range(4), key=lambda d: (-sc[d], d))
        assert ids[q].tolist() == order
        np.testing.assert_allclose(s[q], [sc[d] for d in order], rtol=1e-12)


def max_dot_sum(Qq, Xd):
    return sum(max(sum(float(a) * float(b) for a, b in zip(qt, x)) for x in Xd) for qt in Qq)


def test_ivf_build_bruteforce_sets():
    rng = np.random.default_rng(6)
    doclens = rng.integers(1, 20, 50).astype(np.uint32)
    codes = rng.integers(0, 16, int(doclens.sum())).astype(np.uint32)
    off, size, ids = O.ivf_build(codes, doclens, 16)
    start = np.concatenate([[0], np.cumsum(doclens)]).astype(int)
    for c in range(16):
        lst = ids[off[c]:off[c] + size[c]]
        expect = [d for d in range(50) if c in set(codes[start[d]:start[d + 1]].tolist())]
        assert lst.tolist() == expect
    assert size.sum() == ids.size


def test_generator_moments_and_determinism():
    z = np.array([O.normal(7, 4, i) for i in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    cen = O.gen_centers(7, 16)
    np.testing.assert_allclose(np.linalg.norm(cen, axis=1), 1, atol=1e-6)
    t1 = O.gen_tokens(7, cen, 100)
    t2 = O.gen_tokens(7, cen, 50, tok_begin=50)
    np.testing.assert_array_equal(t1[50:], t2)
    np.testing.assert_allclose(np.linalg.norm(t1, axis=1), 1, atol=1e-6)
    lens = O.gen_doclens_uniform(7, 5000, 16, 128)
    assert lens.min() == 16 and lens.max() == 128
    ln = O.gen_doclens_lognormal(7, 20000, 64, 0.29, 8, 300)
    assert 62 <= np.median(ln) <= 66 and ln.min() >= 8 and ln.max() <= 300
