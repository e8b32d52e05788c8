"""GPU parity of the PLAID path (mvr.plaid, libmvr's mvr_plaid_* C ABI)
against the CPU oracle (oracle/plaid_oracle.c), on the lotte-500 PLAID index
fixture and on synthetic PLAID indices.

Every op is integer/byte work or a fixed-order FP32 computation, so the bar
is bit equality: FP16 decompression, packbits, binarize, FP32 decompression,
segmented MaxSim, the two-stage filter (pids, order, n_vec_score_refine) and
the batched search (pids, scores, per-query counts).
"""
import os

import numpy as np
import pytest

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


def _synthetic_plaid(seed, n_doc, n_c, lo=4, hi=90):
    """A PLAID index over random unit vectors: codes by inner-product argmax,
    residual bytes by the oracle's binarize with pooled quantile cutoffs."""
    from oracle import oracle as O, plaid as P
    rng = np.random.default_rng(seed)
    doclens = rng.integers(lo, hi, n_doc).astype(np.int64)
    n_tok = int(doclens.sum())
    X = rng.standard_normal((n_tok, 128)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    C_ = X[rng.choice(n_tok, n_c, replace=False)].copy()
    codes = O.assign_ip(X, C_).astype(np.int32)
    E = X - C_[codes]
    cut = np.quantile(E, [0.25, 0.5, 0.75]).astype(np.float32)
    w = np.quantile(E, [0.125, 0.375, 0.625, 0.875]).astype(np.float32)
    res = P.binarize(E, cut, 2)
    ivf_sizes = []
    ivf = []
    pid_of = np.repeat(np.arange(n_doc), doclens)
    for c in range(n_c):
        p = np.unique(pid_of[codes == c])
        ivf.append(p)
        ivf_sizes.append(p.size)
    ivf = np.concatenate(ivf).astype(np.int32)
    return dict(doclens=doclens, codes=codes, residuals=res, centroids=C_, bucket_weights=w, ivf=ivf,
                ivf_lengths=np.array(ivf_sizes, np.int64), X=X)


# ------------------------------------------------------------- codec ops
def test_decompress_residuals_cuda_bitexact_on_fixture(fx):
    import torch
    from mvr import plaid
    from oracle import plaid as P
    rbm, lut = P.reversed_bit_map(2), P.decompression_lookup_table(2)
    got = plaid.decompress_residuals_cuda(torch.from_numpy(fx["residuals"]).cuda(),
                                          torch.from_numpy(fx["bucket_weights"]).cuda(),
                                          torch.from_numpy(rbm).cuda(), torch.from_numpy(lut).cuda(),
                                          torch.from_numpy(fx["codes"]).cuda(),
                                          torch.from_numpy(fx["centroids"]).cuda(), 128, 2)
    torch.cuda.synchronize()
    ref = P.decompress_f16(fx["residuals"], fx["bucket_weights"].view(np.uint16), rbm, lut, fx["codes"],
                           fx["centroids"].view(np.uint16))
    np.testing.assert_array_equal(got.cpu().numpy().view(np.uint16), ref)


@pytest.mark.parametrize("nbits", [1, 4])
def test_decompress_residuals_cuda_other_widths(nbits):
    import torch
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(nbits)
    n, dim, n_c = 777, 64, 50
    res = rng.integers(0, 256, (n, dim * nbits // 8)).astype(np.uint8)
    bw = rng.standard_normal(1 << nbits).astype(np.float16)
    C16 = rng.standard_normal((n_c, dim)).astype(np.float16)
    codes = rng.integers(0, n_c, n).astype(np.int32)
    rbm, lut = P.reversed_bit_map(nbits), P.decompression_lookup_table(nbits)
    got = plaid.decompress_residuals_cuda(res, bw, rbm, lut, codes, C16, dim, nbits)
    ref = P.decompress_f16(res, bw.view(np.uint16), rbm, lut, codes, C16.view(np.uint16), dim, nbits)
    np.testing.assert_array_equal(got.cpu().numpy().view(np.uint16), ref)


def test_packbits_cuda_matches_numpy():
    from mvr import plaid
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2, 32 * 12345).astype(np.uint8)
    bits[rng.random(bits.size) < 0.05] = 200
    got = plaid.packbits_cuda(bits).cpu().numpy()
    np.testing.assert_array_equal(got, np.packbits(bits != 0))


@pytest.mark.parametrize("nbits", [1, 2, 4])
def test_binarize_bitexact(nbits):
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(10 + nbits)
    n_cut = (1 << nbits) - 1
    cut = np.sort(rng.standard_normal(n_cut).astype(np.float32) * 0.05)
    e = (rng.standard_normal((2000, 128)) * 0.05).astype(np.float32)
    e[0, :n_cut] = cut                       # exactly on the cutoffs (bucketize right=False)
    got = plaid.binarize(e, cut, nbits).cpu().numpy()
    np.testing.assert_array_equal(got, P.binarize(e, cut, nbits))


def test_decompress_pids_bitexact_on_fixture(opix):
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(3)
    pids = rng.integers(0, opix.n_doc, 300).astype(np.int32)   # any order, repeats
    got = plaid.decompress_residuals(pids, opix.doclens, opix.offsets, opix.bucket_weights, opix.rbm, opix.lut,
                                     opix.residuals, opix.codes, opix.centroids, 128, 2).cpu().numpy()
    ref = P.decompress_pids(pids, opix.doclens, opix.offsets, opix.bucket_weights, opix.rbm, opix.lut,
                            opix.residuals, opix.codes, opix.centroids)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("nq", [32, 20, 64])
def test_segmented_maxsim_bitexact(nq):
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(nq)
    lengths = rng.integers(0, 40, 500).astype(np.int64)
    lengths[:3] = [0, 1, 0]
    s = rng.standard_normal((int(lengths.sum()), nq)).astype(np.float32)
    got = plaid.segmented_maxsim(s, lengths).cpu().numpy()
    np.testing.assert_array_equal(got, P.segmented_maxsim(s, lengths))


# ------------------------------------------------------- segmented_lookup
@pytest.mark.parametrize("dtype,dim", [(np.uint8, 1), (np.uint8, 32), (np.int32, 1), (np.int64, 3),
                                       (np.float32, 128), (np.float16, 5), (np.float16, 128)])
def test_segmented_lookup_matches_oracle(dtype, dim):
    """IVF-style lookups (segments of any length, zero-length ones,
    overlapping and repeated segments, the last rows) for every dtype the
    reference dispatches, 1-D and 2-D, host and device inputs."""
    import torch
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(dim * 7 + np.dtype(dtype).itemsize)
    n_rows = 5000
    shape = (n_rows,) if dim == 1 else (n_rows, dim)
    x = rng.integers(0, 250, shape).astype(dtype) if np.dtype(dtype).kind in "iu" else \
        rng.standard_normal(shape).astype(dtype)
    n = 700
    lengths = rng.integers(0, 40, n).astype(np.int64)
    lengths[:5] = [0, 0, 1, 39, 0]
    offsets = (rng.random(n) * (n_rows - lengths)).astype(np.int64)
    offsets[5], lengths[5] = n_rows - 7, 7            # the last rows
    offsets[6:8], lengths[6:8] = offsets[3], lengths[3]  # a repeated segment
    pids = np.arange(n, dtype=np.int64)
    ref = P.segmented_lookup(x, pids, lengths, offsets)
    got = plaid.segmented_lookup(x, pids, lengths, offsets)
    assert not torch.is_tensor(got) or got.device.type == "cpu"
    got = got.numpy() if torch.is_tensor(got) else got
    assert got.dtype == ref.dtype and got.shape == ref.shape
    np.testing.assert_array_equal(got.view(np.uint8), ref.view(np.uint8))
    # device input gives a device result; pids shorter than lengths read only the first entries
    xd = torch.from_numpy(x).cuda()
    gd = plaid.segmented_lookup(xd, torch.from_numpy(pids[:300]), torch.from_numpy(lengths),
                                torch.from_numpy(offsets))
    assert gd.is_cuda
    np.testing.assert_array_equal(gd.cpu().numpy().view(np.uint8),
                                  P.segmented_lookup(x, pids[:300], lengths, offsets).view(np.uint8))


def test_segmented_lookup_empty_and_out_of_range():
    import torch
    from mvr import plaid, _capi
    x = np.arange(100, dtype=np.int32)
    assert plaid.segmented_lookup(x, np.zeros(0), np.zeros(0, np.int64), np.zeros(0, np.int64)).shape == (0,)
    assert plaid.segmented_lookup(x, [1, 2], [0, 0], [3, 99]).shape == (0,)
    with pytest.raises(_capi.MvrError):
        plaid.segmented_lookup(x, [1], [5], [98])       # past the end: the reference reads out of bounds
    with pytest.raises(_capi.MvrError):
        plaid.segmented_lookup(x, [1], [2], [-1])


# ------------------------------------------------------------ filter_pids
def test_filter_pids_bitexact_on_fixture_queries(fx, opix):
    from mvr import plaid
    from oracle import plaid as P
    for qi in range(4):
        S = P.centroid_scores(opix.centroids, fx["queries"][qi])
        for thr, ndocs in ((0.5, 256), (0.45, 1024), (0.3, 64), (0.5, 4)):
            idx = (S.max(1) >= thr).astype(np.uint8)
            pids = np.arange(opix.n_doc, dtype=np.int32)
            got, nv = plaid.filter_pids(pids, S, opix.codes, opix.doclens, opix.offsets, idx.astype(bool), ndocs)
            ref, _, nv_ref = P.filter_pids(pids, S, opix.codes, opix.doclens, opix.offsets, idx, ndocs)
            np.testing.assert_array_equal(got.cpu().numpy(), ref)
            assert nv == nv_ref


def test_filter_pids_ties_and_short_input():
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(11)
    n_c, nq, n_doc = 64, 8, 3000
    doclens = rng.integers(0, 12, n_doc).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(doclens)[:-1]]).astype(np.int64)
    codes = rng.integers(0, n_c, int(doclens.sum())).astype(np.int32)
    S = (rng.integers(0, 3, (n_c, nq)) * 0.25).astype(np.float32)     # many equal scores
    idx = (rng.random(n_c) < 0.6).astype(np.uint8)
    for npids, ndocs in ((2500, 1000), (100, 1000), (2999, 4096), (9, 8)):
        pids = np.sort(rng.choice(n_doc, npids, replace=False)).astype(np.int32)
        got, nv = plaid.filter_pids(pids, S, codes, doclens, offsets, idx.astype(bool), ndocs)
        ref, _, nv_ref = P.filter_pids(pids, S, codes, doclens, offsets, idx, ndocs)
        np.testing.assert_array_equal(got.cpu().numpy(), ref)
        assert nv == nv_ref


def test_filter_pids_ndocs_above_4096():
    """Stage-1 selections of 8,000 and 14,000 pids (the large select
    variant), ties included, bit-exact against the oracle."""
    from mvr import plaid
    from oracle import plaid as P
    rng = np.random.default_rng(12)
    n_c, nq, n_doc = 64, 8, 40000
    doclens = rng.integers(0, 12, n_doc).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(doclens)[:-1]]).astype(np.int64)
    codes = rng.integers(0, n_c, int(doclens.sum())).astype(np.int32)
    S = (rng.integers(0, 5, (n_c, nq)) * 0.125).astype(np.float32)
    idx = (rng.random(n_c) < 0.7).astype(np.uint8)
    for npids, ndocs in ((30000, 8000), (39000, 14000), (20000, 16384), (12000, 14000)):
        pids = np.sort(rng.choice(n_doc, npids, replace=False)).astype(np.int32)
        got, nv = plaid.filter_pids(pids, S, codes, doclens, offsets, idx.astype(bool), ndocs)
        ref, _, nv_ref = P.filter_pids(pids, S, codes, doclens, offsets, idx, ndocs)
        np.testing.assert_array_equal(got.cpu().numpy(), ref, err_msg=f"{npids} {ndocs}")
        assert nv == nv_ref


# -------------------------------------------------------- batched search
@pytest.mark.parametrize("k", [10, 100, 1000])
def test_search_bitexact_on_fixture(fx, opix, gpix, k):
    from oracle import plaid as P
    Q = fx["queries"]
    ncells, thr, ndocs = P.searcher_defaults(k)
    ids, sc, st = gpix.search(Q, k=k, return_stats=True)
    rid, rsc, rst = opix.search(Q, k, ncells, thr, ndocs)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_array_equal(sc, rsc)
    np.testing.assert_array_equal(st, rst)


def test_search_batches_and_nq_below_32(fx, opix, gpix):
    from oracle import plaid as P
    Q = np.ascontiguousarray(fx["queries"][:, :20])
    a = gpix.search(Q, k=10, batch=3, return_stats=True)
    b = gpix.search(Q, k=10, batch=1024, return_stats=True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    rid, rsc, rst = opix.search(Q, 10, *P.searcher_defaults(10))
    np.testing.assert_array_equal(a[0], rid)
    np.testing.assert_array_equal(a[1], rsc)
    np.testing.assert_array_equal(a[2], rst)


def test_search_bitexact_on_synthetic_index():
    from mvr import plaid
    from oracle import plaid as P
    d = _synthetic_plaid(5, 3000, 512)
    X = d.pop("X")
    o = P.PlaidIndex(**d)
    g = plaid.PlaidIndex(**d)
    rng = np.random.default_rng(9)
    # queries: noisy copies of doc tokens so the filter has real work
    Q = X[rng.choice(X.shape[0], 70 * 32)].reshape(70, 32, 128)
    Q = Q + 0.3 * rng.standard_normal(Q.shape).astype(np.float32)
    Q = (Q / np.linalg.norm(Q, axis=2, keepdims=True)).astype(np.float32)
    for k, ncells, thr, ndocs in ((10, 1, 0.5, 256), (100, 2, 0.45, 1024), (10, 4, 0.2, 4096), (5, 3, 0.9, 64)):
        ids, sc, st = g.search(Q, k=k, ncells=ncells, centroid_score_threshold=thr, ndocs=ndocs, batch=32,
                               return_stats=True)
        rid, rsc, rst = o.search(Q, k, ncells, thr, ndocs)
        np.testing.assert_array_equal(ids, rid)
        np.testing.assert_array_equal(sc, rsc)
        np.testing.assert_array_equal(st, rst)
    g.close()


def test_search_ndocs_of_the_reference_drivers():
    """ndocs 8,000 and 14,000 (eval_plaid_gpu.py:79,85; searcher.py:310 sets
    ndocs = max(4k, 4096)) on a 20,000-doc synthetic index: the stage-1
    selection keeps up to 16,384 pids (the large select variant), stage 2
    ndocs/4; bit-exact against the oracle."""
    from mvr import plaid
    from oracle import plaid as P
    d = _synthetic_plaid(7, 20000, 256, lo=4, hi=40)
    X = d.pop("X")
    o = P.PlaidIndex(**d)
    g = plaid.PlaidIndex(**d)
    rng = np.random.default_rng(17)
    Q = X[rng.choice(X.shape[0], 12 * 32)].reshape(12, 32, 128)
    Q = Q + 0.5 * rng.standard_normal(Q.shape).astype(np.float32)
    Q = (Q / np.linalg.norm(Q, axis=2, keepdims=True)).astype(np.float32)
    for k, ncells, thr, ndocs in ((100, 8, 0.2, 8000), (1000, 8, 0.1, 14000), (10, 4, 0.3, 16384)):
        ids, sc, st = g.search(Q, k=k, ncells=ncells, centroid_score_threshold=thr, ndocs=ndocs, batch=6,
                               return_stats=True)
        rid, rsc, rst = o.search(Q, k, ncells, thr, ndocs)
        np.testing.assert_array_equal(ids, rid, err_msg=f"ndocs {ndocs}")
        np.testing.assert_array_equal(sc, rsc, err_msg=f"ndocs {ndocs}")
        np.testing.assert_array_equal(st, rst, err_msg=f"ndocs {ndocs}")
        # stage 1 keeps min(ndocs, candidates) pids: more than 4,096 here
        assert (st[:, 0] > 4096).all(), "stage 1 should keep more than 4,096 pids"
    g.close()


def test_rank_single_query_matches_search(fx, gpix):
    import torch
    Q = fx["queries"][:1]
    pids, scores = gpix.rank(torch.from_numpy(Q).cuda(), k=10)
    ids, sc = gpix.search(Q, k=10)
    assert pids == ids[0].tolist() and scores == sc[0].tolist()
    assert len(pids) == 10 and scores == sorted(scores, reverse=True)


def test_candidate_marking_paths_agree(fx, opix, monkeypatch):
    """Candidate generation in one workgroup per query with the pid bitmap in
    LDS (used whenever it fits) and the global-bitmap kernels
    (MVR_PLAID_FUSED_MARK=0, the path for corpora whose bitmap does not fit)
    give the oracle's answers, ncells 1 to 8."""
    from mvr import plaid
    from oracle import plaid as P
    Q = fx["queries"]
    idx = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("MVR_PLAID_FUSED_MARK", mode)
        idx[mode] = plaid.PlaidIndex(fx["doclens"], fx["codes"], fx["residuals"],
                                     fx["centroids"].astype(np.float32), fx["bucket_weights"].astype(np.float32),
                                     fx["ivf"], fx["ivf_lengths"])
    for ncells, thr, ndocs in ((1, 0.5, 256), (8, 0.3, 1024)):
        rid, rsc, rst = opix.search(Q, 10, ncells, thr, ndocs)
        for mode, g in idx.items():
            ids, sc, st = g.search(Q, k=10, ncells=ncells, centroid_score_threshold=thr, ndocs=ndocs,
                                   return_stats=True)
            np.testing.assert_array_equal(ids, rid, err_msg=mode)
            np.testing.assert_array_equal(sc, rsc, err_msg=mode)
            np.testing.assert_array_equal(st, rst, err_msg=mode)
    for g in idx.values():
        g.close()


def test_search_with_exactly_tied_centroid_scores():
    """Centroid rows duplicated in pairs, so every query token sees exact ties
    in its top-ncells list (the float16 GPU branch rounds many scores equal):
    the cell lists keep (score desc, centroid asc) without duplicates, in both
    branches, bit-exact against the oracle."""
    from mvr import plaid
    from oracle import plaid as P
    d = _synthetic_plaid(13, 2000, 256)
    X = d.pop("X")
    C_ = d["centroids"].copy()
    C_[128:] = C_[:128]  # centroid c + 128 duplicates c (its IVF list stays empty)
    d["centroids"] = C_
    rng = np.random.default_rng(17)
    Q = X[rng.choice(X.shape[0], 24 * 32)].reshape(24, 32, 128).astype(np.float32)
    g = plaid.PlaidIndex(**d)
    o = P.PlaidIndex(**d)
    for branch in ("cpu", "gpu"):
        for ncells in (2, 3, 4):
            print(f"[ties] branch={branch} ncells={ncells}", flush=True)
            ids, sc, st = g.search(Q, k=10, ncells=ncells, centroid_score_threshold=0.3, ndocs=256,
                                   return_stats=True, branch=branch)
            rid, rsc, rst = o.search(Q, 10, ncells, 0.3, 256, branch=branch)
            np.testing.assert_array_equal(st, rst)
            np.testing.assert_array_equal(ids, rid)
            np.testing.assert_array_equal(sc, rsc)
    g.close()
