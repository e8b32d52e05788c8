"""Pins of the PLAID part of the CPU oracle (oracle/plaid_oracle.c) against
independent references: numpy's float16 conversions and np.packbits,
torch.bucketize (the op residual.py:190 calls), the lotte-500 PLAID index
fixture, a Python simulation of filter_pids.cpp's per-thread priority queues,
and float64 brute force.
"""
import heapq
import os

import numpy as np
import pytest

from oracle import plaid as P

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lotte500.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


@pytest.fixture(scope="module")
def pix(fx):
    return P.PlaidIndex(fx["doclens"], fx["codes"], fx["residuals"], fx["centroids"].astype(np.float32),
                        fx["bucket_weights"].astype(np.float32), fx["ivf"], fx["ivf_lengths"])


# ------------------------------------------------------------------ halves
def test_f16_widen_all_halves_match_numpy():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    got = P.f16_widen(h)
    ref = h.view(np.float16).astype(np.float32)
    nan = np.isnan(ref)
    np.testing.assert_array_equal(np.isnan(got), nan)
    np.testing.assert_array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def _half_midpoints():
    h = np.arange(0, 0x7C00, dtype=np.uint32).astype(np.uint16).view(np.float16).astype(np.float64)
    return ((h[:-1] + h[1:]) / 2).astype(np.float32)  # 12 significant bits: exact in float32


def test_f16_round_matches_numpy_including_ties_subnormals_overflow():
    rng = np.random.default_rng(5)
    parts = [
        (rng.standard_normal(200000) * np.exp2(rng.integers(-30, 18, 200000))).astype(np.float32),
        # exact midpoints between neighbouring finite halves (ties to even)
        _half_midpoints(),
        np.array([65504, 65519.99, 65520, 70000, 2.0 ** -24, 2.0 ** -25, 2.0 ** -25 * 1.0000001,
                  3 * 2.0 ** -26, 2.0 ** -14, 2.0 ** -14 * (1 - 2 ** -12), 0.0, -0.0, np.inf, -np.inf],
                 np.float32),
    ]
    x = np.concatenate(parts)
    x = np.concatenate([x, -x])
    got = P.f16_round(x)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got, ref)


# ------------------------------------------------------------- codec ops
def test_packbits_matches_numpy_packbits():
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2, 32 * 1000).astype(np.uint8)
    bits[rng.random(bits.size) < 0.1] = 7  # nonzero bytes count as 1 (packbits.cu:27 ballot)
    np.testing.assert_array_equal(P.packbits(bits), np.packbits(bits != 0))
    assert P.packbits(np.array([1, 0, 0, 0, 0, 0, 0, 1] + [0] * 24, np.uint8))[0] == 0b10000001


def test_binarize_matches_torch_bucketize_bits_and_packbits():
    import torch
    rng = np.random.default_rng(2)
    cut = np.array([-0.03, 0.0, 0.03], np.float32)
    e = (rng.standard_normal((300, 128)) * 0.04).astype(np.float32)
    e[0, :6] = [-0.03, 0.0, 0.03, -0.0300001, 0.0300001, -1e-9]  # values on and next to cutoffs
    got = P.binarize(e, cut, 2)
    b = torch.bucketize(torch.from_numpy(e), torch.from_numpy(cut)).to(torch.uint8)
    bits = (b.unsqueeze(-1) >> torch.arange(2, dtype=torch.uint8)) & 1     # residual.py:191-193
    ref = np.packbits(bits.numpy().reshape(-1)).reshape(300, 32)
    np.testing.assert_array_equal(got, ref)


def test_binarize_then_decode_recovers_buckets():
    """compress -> decompress round trip: the PLAID decode
    (lut[reversed_bit_map[byte]]) of binarize's bytes gives back each
    dim's bucket."""
    rng = np.random.default_rng(3)
    cut = np.array([-0.5, 0.0, 0.5], np.float32)
    e = rng.standard_normal((64, 128)).astype(np.float32)
    packed = P.binarize(e, cut, 2)
    bucket = (e[..., None] > cut).sum(-1)
    rbm, lut = P.reversed_bit_map(2), P.decompression_lookup_table(2)
    dec = lut[rbm[packed]].reshape(64, 128)
    np.testing.assert_array_equal(dec, bucket)


def test_decompress_f16_on_fixture_matches_numpy_half_arithmetic(fx):
    rng = np.random.default_rng(4)
    tok = rng.choice(fx["codes"].size, 3000, replace=False)
    res, codes = fx["residuals"][tok], fx["codes"][tok]
    rbm, lut = P.reversed_bit_map(2), P.decompression_lookup_table(2)
    got = P.decompress_f16(res, fx["bucket_weights"].view(np.uint16), rbm, lut, codes,
                           fx["centroids"].view(np.uint16))
    w = fx["bucket_weights"][lut[rbm[res]].reshape(-1, 128)].astype(np.float32)
    ref = (w + fx["centroids"][codes].astype(np.float32)).astype(np.float16)
    np.testing.assert_array_equal(got, ref.view(np.uint16))


def test_decompress_pids_on_fixture(fx, pix):
    pids = np.array([7, 0, 499, 123, 7], np.int32)  # any order, repeats allowed
    got = P.decompress_pids(pids, pix.doclens, pix.offsets, pix.bucket_weights, pix.rbm, pix.lut,
                            pix.residuals, pix.codes, pix.centroids)
    rows = np.concatenate([np.arange(pix.offsets[p], pix.offsets[p] + pix.doclens[p]) for p in pids])
    w = pix.bucket_weights[pix.lut[pix.rbm[pix.residuals[rows]]].reshape(-1, 128)]
    ref = pix.centroids[pix.codes[rows]] + w
    np.testing.assert_array_equal(got, ref)
    # FP16 decompress of the same rows equals the FP32 one rounded at the end
    f16 = P.decompress_f16(pix.residuals[rows], fx["bucket_weights"].view(np.uint16), pix.rbm, pix.lut,
                           pix.codes[rows], fx["centroids"].view(np.uint16))
    np.testing.assert_array_equal(f16, got.astype(np.float16).view(np.uint16))


def test_segmented_maxsim_zero_init_and_brute_force():
    rng = np.random.default_rng(6)
    lengths = np.array([3, 0, 5, 1, 7], np.int64)
    s = rng.standard_normal((lengths.sum(), 32)).astype(np.float32)
    s[3:8] = -np.abs(s[3:8])                  # doc 2 all negative -> 0
    got = P.segmented_maxsim(s, lengths)
    off = np.concatenate([[0], np.cumsum(lengths)])
    for d in range(lengths.size):
        m = np.zeros(32) if lengths[d] == 0 else np.maximum(0, s[off[d]:off[d + 1]].max(0))
        assert abs(got[d] - m.sum()) <= 1e-5 * max(1.0, abs(m.sum()))
    assert got[1] == 0.0 and got[2] == 0.0


# ------------------------------------------------------------ filter_pids
def _filter_pids_sim(pids, S, codes, doclens, offsets, idx, ndocs, nthreads):
    """Python simulation of filter_pids.cpp: each thread scores a contiguous
    slice of pids into a max-heap of (score, pid); the merge pops each
    thread's best ndocs into a global heap and keeps its best ndocs."""
    nq = S.shape[1]

    def stage(plist, mask, keep):
        per = -(-len(plist) // nthreads)
        heaps, nv = [], 0
        for t in range(nthreads):
            h = []
            for pid in plist[t * per:(t + 1) * per]:
                m = [np.float32(-9999.0)] * nq
                seen = set()
                for j in range(doclens[pid]):
                    c = int(codes[offsets[pid] + j])
                    if mask[c] and c not in seen:
                        m = [max(a, b) for a, b in zip(m, S[c])]
                        seen.add(c)
                        nv += nq
                sc = np.float32(0.0)
                for v in m:
                    sc = np.float32(sc + v)
                heapq.heappush(h, (-sc, -pid))
            heaps.append(h)
        g = []
        for h in heaps:
            for _ in range(keep):
                if not h:
                    break
                heapq.heappush(g, heapq.heappop(h))
        out = [-heapq.heappop(g)[1] for _ in range(min(keep, len(g)))]
        return out, nv

    p1, nv1 = stage(list(pids), idx, ndocs)
    p2, nv2 = stage(p1, np.ones_like(idx), ndocs // 4)
    return np.array(p2, np.int32), nv1 + nv2


@pytest.mark.parametrize("seed,ties", [(0, False), (1, True), (2, True)])
def test_filter_pids_matches_priority_queue_simulation(seed, ties):
    rng = np.random.default_rng(seed)
    n_c, nq, n_doc = 64, 8, 90
    doclens = rng.integers(0, 12, n_doc).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(doclens)[:-1]]).astype(np.int64)
    codes = rng.integers(0, n_c, int(doclens.sum())).astype(np.int32)
    if ties:   # few distinct values: many equal scores, the pid order decides
        S = rng.integers(0, 3, (n_c, nq)).astype(np.float32) * 0.25
    else:
        S = rng.standard_normal((n_c, nq)).astype(np.float32)
    idx = (rng.random(n_c) < 0.5).astype(np.uint8)
    pids = np.sort(rng.choice(n_doc, 70, replace=False)).astype(np.int32)
    for ndocs in (8, 40, 70):
        got, _, nv = P.filter_pids(pids, S, codes, doclens, offsets, idx, ndocs)
        for nthreads in (1, 3):
            ref, nv_ref = _filter_pids_sim(pids, S, codes, doclens, offsets, idx, ndocs, nthreads)
            np.testing.assert_array_equal(got, ref)
            assert nv == nv_ref


def test_filter_pids_scores_are_descending_pid_desc_on_ties():
    rng = np.random.default_rng(9)
    n_c, nq, n_doc = 16, 4, 200
    doclens = np.full(n_doc, 2, np.int64)
    offsets = np.arange(n_doc, dtype=np.int64) * 2
    codes = rng.integers(0, n_c, 2 * n_doc).astype(np.int32)
    S = rng.integers(0, 2, (n_c, nq)).astype(np.float32)
    pids = np.arange(n_doc, dtype=np.int32)
    got, s2, _ = P.filter_pids(pids, S, codes, doclens, offsets, np.ones(n_c, np.uint8), 400)
    assert got.size == 100
    keys = list(zip(s2.tolist(), got.tolist()))
    assert keys == sorted(keys, reverse=True)


# ------------------------------------------------------ whole CPU branch
def test_centroid_scores_and_refine_against_float64(fx, pix):
    Q = fx["queries"][0]
    S = P.centroid_scores(pix.centroids, Q)
    ref = pix.centroids.astype(np.float64) @ Q.astype(np.float64).T
    assert np.max(np.abs(S - ref)) < 1e-5
    pids = np.array([3, 17, 250], np.int32)
    got = pix.score_pids(Q, pids)
    X = P.decompress_pids(pids, pix.doclens, pix.offsets, pix.bucket_weights, pix.rbm, pix.lut,
                          pix.residuals, pix.codes, pix.centroids).astype(np.float64)
    X /= np.maximum(np.linalg.norm(X, axis=1, keepdims=True), 1e-12)
    off = 0
    for i, p in enumerate(pids):
        L = pix.doclens[p]
        s = np.maximum(0, (X[off:off + L] @ Q.astype(np.float64).T).max(0)).sum()
        off += L
        assert abs(got[i] - s) <= 1e-5 * abs(s)


def test_search_pipeline_composes_the_ops(fx, pix):
    """orc_plaid_search_batch == candidate generation + filter_pids +
    score_pids + sort, composed here from numpy and the single ops."""
    Q = fx["queries"][:6]
    for k in (10, 100):
        ncells, thr, ndocs = P.searcher_defaults(k)
        ids, sc, stats = pix.search(Q, k, ncells, thr, ndocs)
        for q in range(Q.shape[0]):
            S = P.centroid_scores(pix.centroids, Q[q])
            cells = set()
            for t in range(S.shape[1]):
                order = sorted(range(S.shape[0]), key=lambda c: (-S[c, t], c))[:ncells]
                cells.update(order)
            pids = np.unique(np.concatenate([pix.ivf[pix.ivf_off[c]:pix.ivf_off[c] + pix.ivf_lengths[c]]
                                             for c in sorted(cells)])).astype(np.int32)
            assert stats[q, 0] == pids.size
            idx = (S.max(1) >= thr).astype(np.uint8)
            fp, _, nv = P.filter_pids(pids, S, pix.codes, pix.doclens, pix.offsets, idx, ndocs)
            assert stats[q, 1] == fp.size and stats[q, 2] == nv
            s = pix.score_pids(Q[q], fp)
            order = sorted(range(fp.size), key=lambda i: (-s[i], fp[i]))[:k]
            n = len(order)
            np.testing.assert_array_equal(ids[q, :n], fp[order])
            np.testing.assert_array_equal(sc[q, :n], s[order])
            assert np.all(ids[q, n:] == -1)


def test_search_recall_against_fixture_groundtruth(fx, pix):
    """The fixture's top-10 is brute force over the original embeddings; PLAID
    with the searcher's k=10 defaults finds most of it (weak sanity check)."""
    ids, _, _ = pix.search(fx["queries"], 10, *P.searcher_defaults(10))
    hits = np.mean([len(set(a.tolist()) & set(b.tolist())) / 10 for a, b in zip(ids, fx["gnd_ids"])])
    assert hits >= 0.6, hits


def test_segmented_lookup_known_answers():
    from oracle import plaid as P
    x = np.arange(10, dtype=np.int32)
    got = P.segmented_lookup(x, [7, 8, 9], [2, 0, 3], [1, 5, 7])
    np.testing.assert_array_equal(got, [1, 2, 7, 8, 9])
    assert got.dtype == np.int32
    X = np.arange(24, dtype=np.float32).reshape(8, 3)
    got = P.segmented_lookup(X, np.zeros(2), [1, 2], [6, 0])
    np.testing.assert_array_equal(got, [[18, 19, 20], [0, 1, 2], [3, 4, 5]])
    # only the first len(pids) lengths/offsets are read, like ndocs = pids.size(0)
    np.testing.assert_array_equal(P.segmented_lookup(x, [0], [1, 5], [9, 0]), [9])
    assert P.segmented_lookup(X, [], [], []).shape == (0, 3)
