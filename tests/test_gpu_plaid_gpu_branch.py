"""GPU parity of the PLAID GPU branch (the float16 path eval_plaid_gpu.py runs:
index_storage.py:112-153,160-161,185-187, colbert_score_reduce /
colbert_score_packed, half lookup_pids) against the CPU oracle's restatement
(oracle/plaid_oracle.c, orc_plaid_*_f16 and orc_plaid_search_batch_gpu).

Both sides sum in the same fixed FP32 orders and round to half at the same
points, so the bar is bit equality: decompressed half embeddings, half
scores, pids, order and per-query counts."""
import os

import numpy as np
import pytest

from test_gpu_plaid import _synthetic_plaid

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lotte500.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


@pytest.fixture(scope="module")
def opix(fx):
    from oracle import plaid as P
    return P.PlaidIndex(fx["doclens"], fx["codes"], fx["residuals"], fx["centroids"].astype(np.float32),
                        fx["bucket_weights"].astype(np.float32), fx["ivf"], fx["ivf_lengths"])


@pytest.fixture(scope="module")
def gpix(fx):
    from mvr import plaid
    ix = plaid.PlaidIndex(fx["doclens"], fx["codes"], fx["residuals"], fx["centroids"].astype(np.float32),
                          fx["bucket_weights"].astype(np.float32), fx["ivf"], fx["ivf_lengths"])
    yield ix
    ix.close()


@pytest.fixture(scope="module")
def synth():
    """3,000 docs, 512 FP32 centroids and weights (not half values, so the
    branch's rounding to half is exercised), noisy doc-token queries."""
    from mvr import plaid
    from oracle import plaid as P
    d = _synthetic_plaid(5, 3000, 512)
    X = d.pop("X")
    rng = np.random.default_rng(9)
    Q = X[rng.choice(X.shape[0], 70 * 32)].reshape(70, 32, 128)
    Q = Q + 0.3 * rng.standard_normal(Q.shape).astype(np.float32)
    Q = (Q / np.linalg.norm(Q, axis=2, keepdims=True)).astype(np.float32)
    g = plaid.PlaidIndex(**d)
    yield P.PlaidIndex(**d), g, Q
    g.close()


@pytest.mark.parametrize("k", [10, 100, 1000])
def test_gpu_branch_search_bitexact_on_fixture(fx, opix, gpix, k):
    from mvr import plaid
    ncells, thr, ndocs = plaid.searcher_defaults(k)
    Q = fx["queries"]
    print(f"[gpu-branch fixture] k={k} n_q={Q.shape[0]}", flush=True)
    ids, sc, st = gpix.search(Q, k=k, return_stats=True, branch="gpu")
    rid, rsc, rst = opix.search(Q, k, ncells, thr, ndocs, branch="gpu")
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_array_equal(sc, rsc)
    np.testing.assert_array_equal(st, rst)
    # returned scores are halves
    ok = ids >= 0
    assert np.array_equal(sc[ok], sc[ok].astype(np.float16).astype(np.float32))


def test_gpu_branch_search_bitexact_on_synthetic_index(synth):
    o, g, Q = synth
    for k, ncells, thr, ndocs, nq in ((10, 1, 0.5, 256, 32), (100, 2, 0.45, 1024, 32), (10, 4, 0.2, 4096, 20),
                                      (5, 3, 0.9, 64, 3), (1000, 4, 0.4, 4096, 32)):
        print(f"[gpu-branch synthetic] k={k} ncells={ncells} thr={thr} ndocs={ndocs} nq={nq}", flush=True)
        Qn = np.ascontiguousarray(Q[:, :nq])
        ids, sc, st = g.search(Qn, k=k, ncells=ncells, centroid_score_threshold=thr, ndocs=ndocs, batch=32,
                               return_stats=True, branch="gpu")
        rid, rsc, rst = o.search(Qn, k, ncells, thr, ndocs, branch="gpu")
        np.testing.assert_array_equal(ids, rid)
        np.testing.assert_array_equal(sc, rsc)
        np.testing.assert_array_equal(st, rst)


def test_gpu_branch_ndocs_of_the_reference_drivers():
    """ndocs 8,000 (eval_plaid_gpu.py:79) on a 20,000-doc index, GPU branch."""
    from mvr import plaid
    from oracle import plaid as P
    d = _synthetic_plaid(7, 20000, 256, lo=4, hi=40)
    X = d.pop("X")
    rng = np.random.default_rng(3)
    Q = X[rng.choice(X.shape[0], 6 * 32)].reshape(6, 32, 128)
    g = plaid.PlaidIndex(**d)
    o = P.PlaidIndex(**d)
    for k, ndocs in ((100, 8000), (1000, 14000)):
        print(f"[gpu-branch ndocs] k={k} ndocs={ndocs}", flush=True)
        ids, sc, st = g.search(Q, k=k, ncells=4, centroid_score_threshold=0.3, ndocs=ndocs, return_stats=True,
                               branch="gpu")
        rid, rsc, rst = o.search(Q, k, 4, 0.3, ndocs, branch="gpu")
        np.testing.assert_array_equal(ids, rid)
        np.testing.assert_array_equal(sc, rsc)
        np.testing.assert_array_equal(st, rst)
    g.close()


def test_lookup_pids_bitexact(fx, opix, gpix, synth):
    import torch
    from mvr import _capi
    pids = np.array([7, 0, 499, 123, 7, 250, 31], np.int32)
    D, ln = gpix.lookup_pids(torch.from_numpy(pids))
    ref, rln = opix.lookup_pids_f16(pids)
    np.testing.assert_array_equal(ln.cpu().numpy(), rln)
    np.testing.assert_array_equal(D.cpu().numpy().view(np.uint16), ref.view(np.uint16))
    o, g, _ = synth
    pids = np.arange(0, 3000, 7, dtype=np.int32)
    D, ln = g.lookup_pids(pids)
    ref, rln = o.lookup_pids_f16(pids)
    np.testing.assert_array_equal(D.cpu().numpy().view(np.uint16), ref.view(np.uint16))
    assert g.lookup_pids(np.zeros(0, np.int32))[0].shape == (0, 128)
    with pytest.raises(_capi.MvrError):
        g.lookup_pids(np.array([3, 3000], np.int32))


def test_colbert_score_packed_bitexact(fx, opix, gpix):
    import torch
    from mvr import plaid
    from oracle import plaid as P
    pids = np.array([3, 17, 250, 499, 0, 42], np.int32)
    D, ln = gpix.lookup_pids(pids)
    for q in range(4):
        Q = fx["queries"][q]
        got = plaid.colbert_score_packed(torch.from_numpy(Q[None]).cuda(), D, ln).cpu().numpy()
        ref = P.score_packed_f16(Q, D.cpu().numpy(), ln.cpu().numpy())
        np.testing.assert_array_equal(got.view(np.uint16), ref.view(np.uint16))
    # 20 query tokens, and a doc without rows (max stays -10000: -inf after the sum)
    Q = fx["queries"][5][:20]
    ln2 = torch.cat([ln, torch.zeros(1, dtype=torch.int64, device=ln.device)])
    got = plaid.colbert_score_packed(Q, D, ln2).cpu().numpy()
    ref = P.score_packed_f16(Q, D.cpu().numpy(), ln2.cpu().numpy())
    np.testing.assert_array_equal(got.view(np.uint16), ref.view(np.uint16))
    assert np.isneginf(got[-1].astype(np.float32))


def test_colbert_score_reduce_padded_matches_torch_and_oracle():
    import torch
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(21)
    for nq in (32, 5):
        n, L = 300, 12
        lengths = rng.integers(0, L + 1, n)
        s = (rng.integers(-2048, 2048, size=(n, L, nq)) / 1024.0).astype(np.float16)
        mask = np.arange(L)[None, :] < lengths[:, None]
        got = plaid.colbert_score_reduce(torch.from_numpy(s).cuda(), torch.from_numpy(mask).cuda()).cpu().numpy()
        t = torch.from_numpy(s.copy())
        t[~torch.from_numpy(mask)] = -9999
        ref = t.max(1).values.sum(-1).numpy()  # the reference's expression on CPU torch
        np.testing.assert_array_equal(got.view(np.uint16), ref.view(np.uint16))
        packed = np.concatenate([s[i, :lengths[i]] for i in range(n)])
        np.testing.assert_array_equal(got.view(np.uint16), P.score_reduce_f16(packed, lengths).view(np.uint16))


def test_gpu_branch_close_to_cpu_branch(fx, gpix):
    """The two branches keep nearly the same top-10 on the fixture."""
    g, _ = gpix.search(fx["queries"], k=10, branch="gpu")
    c, _ = gpix.search(fx["queries"], k=10)
    same = np.mean([len(set(a.tolist()) & set(b.tolist())) / 10 for a, b in zip(g, c)])
    assert same >= 0.9, same
