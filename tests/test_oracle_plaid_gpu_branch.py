"""Pins of the oracle's GPU-branch PLAID functions (oracle/plaid_oracle.c:
orc_plaid_score_reduce_f16, orc_plaid_lookup_pids_f16,
orc_plaid_score_packed_f16, orc_plaid_search_batch_gpu) against torch's own
float16 operators on the CPU (the tensor expressions of colbert.py:133-155 and
residual.py:273-274 written out here), float64 brute force, and a composition
of the single ops on the lotte-500 fixture."""
import os

import numpy as np
import pytest

from oracle import plaid as P

torch = pytest.importorskip("torch")

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lotte500.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


@pytest.fixture(scope="module")
def pix(fx):
    return P.PlaidIndex(fx["doclens"], fx["codes"], fx["residuals"], fx["centroids"].astype(np.float32),
                        fx["bucket_weights"].astype(np.float32), fx["ivf"], fx["ivf_lengths"])


def _torch_reduce(scores16, lengths):
    """colbert_score_reduce on a torch half tensor: pad to [ndocs, maxlen, nq],
    write -9999 into the padding, max over rows, sum over query tokens."""
    n, nq = len(lengths), scores16.shape[1]
    L = int(max(lengths)) if n else 0
    padded = torch.zeros((n, max(L, 1), nq), dtype=torch.float16)
    mask = torch.zeros((n, max(L, 1)), dtype=torch.bool)
    off = 0
    for i, ln in enumerate(lengths):
        padded[i, :ln] = scores16[off:off + ln]
        mask[i, :ln] = True
        off += ln
    padded[~mask] = -9999
    return padded.max(1).values.sum(-1)


@pytest.mark.parametrize("nq", [32, 3])
def test_score_reduce_f16_matches_torch_half(nq):
    rng = np.random.default_rng(11 + nq)
    lengths = rng.integers(0, 9, size=200)
    lengths[:3] = 0  # docs without rows: -10000 per token, -inf once the sum leaves half's range
    rows = int(lengths.sum())
    # multiples of 2^-10 below 2: every partial sum of 32 is exact in FP32, so
    # torch's reduction order cannot matter
    s = (rng.integers(-2048, 2048, size=(rows, nq)) / 1024.0).astype(np.float16)
    got = P.score_reduce_f16(s, lengths)
    ref = _torch_reduce(torch.from_numpy(s), lengths.tolist()).numpy()
    np.testing.assert_array_equal(got.view(np.uint16), ref.view(np.uint16))
    if nq == 32:
        assert np.all(np.isneginf(got[:3].astype(np.float32)))
    else:
        assert np.all(got[:3].astype(np.float32) == -30000.0)


def test_lookup_pids_f16_matches_torch_normalize(fx, pix):
    pids = np.array([7, 0, 499, 123, 7, 250, 31], np.int32)
    got, ln = pix.lookup_pids_f16(pids)
    np.testing.assert_array_equal(ln, pix.doclens[pids])
    rows = np.concatenate([np.arange(pix.offsets[p], pix.offsets[p] + pix.doclens[p]) for p in pids])
    x16 = P.decompress_f16(pix.residuals[rows], fx["bucket_weights"].view(np.uint16), pix.rbm, pix.lut,
                           pix.codes[rows], fx["centroids"].view(np.uint16)).view(np.float16)
    ref = torch.nn.functional.normalize(torch.from_numpy(x16), p=2, dim=-1).numpy()
    assert ref.dtype == np.float16
    diff = np.abs(got.view(np.uint16).astype(np.int32) - ref.view(np.uint16).astype(np.int32))
    # torch's FP32 sum of squares may run in another order: at most one half
    # ulp apart, and on all but a handful of rows identical
    assert diff.max() <= 1
    assert np.mean(np.all(diff == 0, axis=1)) >= 0.99
    # unit norm in float64 to half precision
    n = np.linalg.norm(got.astype(np.float64), axis=1)
    assert np.all(np.abs(n - 1) < 2e-3)


def test_score_packed_f16_against_rounded_float64(fx, pix):
    pids = np.array([3, 17, 250, 499, 0], np.int32)
    D, ln = pix.lookup_pids_f16(pids)
    Q = fx["queries"][1]
    got = P.score_packed_f16(Q, D, ln)
    Q16 = Q.astype(np.float16).astype(np.float64)
    s16 = (D.astype(np.float64) @ Q16.T).astype(np.float16)  # each dot rounded once to half
    ref = _torch_reduce(torch.from_numpy(s16), ln.tolist()).numpy()
    # the FP32 chain and the exact dot round to the same half except at a
    # rounding boundary, so scores agree to a half ulp of their magnitude
    g, r = got.astype(np.float64), ref.astype(np.float64)
    assert np.all(np.abs(g - r) <= np.spacing(np.abs(r).astype(np.float16)).astype(np.float64) + 1e-12)
    # no zero clip: a query orthogonal to every doc row scores its negative maxima
    Qn = -np.abs(np.asarray(D[:1], np.float32))
    neg = P.score_packed_f16(np.repeat(Qn, 2, 0), D[:1], np.array([1]))
    assert neg[0] < 0


def test_gpu_branch_threshold_compares_in_half():
    # half(0.45) = 0.44995 and half(0.4) = 0.39990: torch compares the half
    # tensor with the threshold rounded to half, and so does the oracle
    h = torch.tensor([0.449951171875, 0.39990234375], dtype=torch.float16)
    assert (h >= 0.45).tolist() == [True, False]
    assert (h >= 0.4).tolist() == [True, True]
    assert float(np.float16(0.45)) == 0.449951171875
    # the oracle's half rounding of the threshold gives the same values
    assert P.f16_widen(P.f16_round(np.array([0.45, 0.4], np.float32))).tolist() == [0.449951171875,
                                                                                    0.39990234375]


def _compose_gpu_branch(pix, Q, k, ncells, thr, ndocs):
    """IndexScorer.rank, GPU branch, from the single ops: half centroid scores,
    cells, unique pids, idx in half, colbert_score_reduce for both filter stages
    (torch.topk ties by pid descending), lookup_pids + colbert_score_packed,
    then (score desc, pid asc)."""
    C16 = pix.centroids.astype(np.float16).astype(np.float32)
    Q16 = Q.astype(np.float16).astype(np.float32)
    S = P.centroid_scores(C16, Q16).astype(np.float16)
    Sf = S.astype(np.float32)
    cells = set()
    for t in range(S.shape[1]):
        cells.update(sorted(range(S.shape[0]), key=lambda c: (-Sf[c, t], c))[:ncells])
    pids = np.unique(np.concatenate([pix.ivf[pix.ivf_off[c]:pix.ivf_off[c] + pix.ivf_lengths[c]]
                                     for c in sorted(cells)])).astype(np.int32)
    idx = (torch.from_numpy(S).max(1).values >= thr).numpy()

    def approx(ps, use_idx):
        sc = []
        for p in ps:
            cd = pix.codes[pix.offsets[p]:pix.offsets[p] + pix.doclens[p]]
            if use_idx:
                cd = cd[idx[cd]]
            sc.append(P.score_reduce_f16(S[cd], np.array([cd.size]))[0])
        return np.array(sc, np.float16).astype(np.float32)

    s1 = approx(pids, True)
    if not np.any([idx[pix.codes[pix.offsets[p]:pix.offsets[p] + pix.doclens[p]]].any() for p in pids]):
        s1[:] = 0.0
    keep = sorted(range(pids.size), key=lambda i: (-s1[i], -pids[i]))[:ndocs]
    p1 = pids[keep]
    s2 = approx(p1, False)
    keep = sorted(range(p1.size), key=lambda i: (-s2[i], -p1[i]))[:ndocs // 4]
    p2 = p1[keep]
    D, ln = pix.lookup_pids_f16(p2)
    s = P.score_packed_f16(Q, D, ln).astype(np.float32)
    order = sorted(range(p2.size), key=lambda i: (-s[i], p2[i]))[:k]
    return pids.size, p2.size, p2[order], s[order]


def test_gpu_branch_search_composes_the_ops(fx, pix):
    Q = fx["queries"][:5]
    for k in (10, 100):
        ncells, thr, ndocs = P.searcher_defaults(k)
        ids, sc, stats = pix.search(Q, k, ncells, thr, ndocs, branch="gpu")
        for q in range(Q.shape[0]):
            n_cand, n_fin, ref_ids, ref_s = _compose_gpu_branch(pix, Q[q], k, ncells, thr, ndocs)
            assert stats[q, 0] == n_cand and stats[q, 1] == n_fin
            n = ref_ids.size
            np.testing.assert_array_equal(ids[q, :n], ref_ids)
            np.testing.assert_array_equal(sc[q, :n], ref_s)
            assert np.all(ids[q, n:] == -1)
            # every returned score is a half value
            assert np.array_equal(sc[q, :n], sc[q, :n].astype(np.float16).astype(np.float32))


def test_gpu_branch_recall_close_to_cpu_branch(fx, pix):
    """Same candidates and nearly the same ranking as the FP32 CPU branch; the
    fixture's brute-force top-10 is found as often (weak sanity check)."""
    d = P.searcher_defaults(10)
    g, _, _ = pix.search(fx["queries"], 10, *d, branch="gpu")
    c, _, _ = pix.search(fx["queries"], 10, *d)
    hits = np.mean([len(set(a.tolist()) & set(b.tolist())) / 10 for a, b in zip(g, fx["gnd_ids"])])
    same = np.mean([len(set(a.tolist()) & set(b.tolist())) / 10 for a, b in zip(g, c)])
    assert hits >= 0.6 and same >= 0.9, (hits, same)
