"""Residual codes of 1, 4 and 8 bits per value (and 2, the default, as the
control) on the exact FP32 path: mvr_search, mvr_rerank_docs(_noalloc) and
mvr_decompress bit-identical to the oracle, whose MSB-first unpack for 8/b
values per byte follows DataStructPP.hpp:184-200 and RerankPP.hpp:100-108.
Indexes of other widths load from files (the width read from weight_l.npy)
and answer like the same arrays; every mode whose kernels decode 2-bit codes
only (BF16, FP8, FP32-MFMA, persistent, host-resident residuals) is refused
with MVR_ERR_UNSUPPORTED.  Calls go through the C ABI (libmvr)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WIDTHS = [1, 2, 4, 8]


@pytest.fixture(scope="module")
def corpus():
    from oracle import synth
    cp = synth.make_corpus(4242, 600, 96, lo=4, hi=140)
    Q = synth.make_queries(4343, cp.centers, 12)
    return cp, Q


_IX = {}


def _oracle_index(corpus, n_bit, per_dim=True):
    from oracle import synth
    key = (n_bit, per_dim)
    if key not in _IX:
        _IX[key] = synth.build_index(corpus[0], 64, n_iter=3, n_bit=n_bit, per_dim=per_dim)
    return _IX[key]


def _lib_index(ix, **kw):
    import mvr
    return mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf, **kw)


@pytest.mark.parametrize("n_bit", WIDTHS)
def test_search_fp32_bit_exact(corpus, n_bit):
    ix = _oracle_index(corpus, n_bit)
    assert ix.residual.shape[1] == 16 * n_bit
    Q = corpus[1]
    g = _lib_index(ix)
    ids, sc = g.search(Q, 4, 96, 10)
    oi, os_ = ix.search(Q, 4, 96, 10)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(sc.view(np.uint32), os_.view(np.uint32))
    g.close()


@pytest.mark.parametrize("n_bit", [1, 4, 8])
def test_search_fp32_pooled_weights_ragged_clip_zero(corpus, n_bit):
    """One decode table for every dim (the reference's pooled weight_l),
    uneven query token and probe counts, the PP filter's clip at zero."""
    ix = _oracle_index(corpus, n_bit, per_dim=False)
    assert ix.weights.ndim == 1 and ix.weights.size == 1 << n_bit
    Q = corpus[1]
    rng = np.random.default_rng(n_bit)
    q_ntok = rng.integers(1, 33, Q.shape[0]).astype(np.uint32)
    q_np = rng.choice([1, 3, 8], Q.shape[0]).astype(np.uint32)
    g = _lib_index(ix)
    for clip in (False, True):
        ids, sc = g.search(Q, 8, 64, 7, q_ntok=q_ntok, q_nprobe=q_np, clip_zero=clip, batch=5)
        oi, os_ = ix.search(Q, 8, 64, 7, q_ntok=q_ntok, q_nprobe=q_np, clip_zero=clip)
        np.testing.assert_array_equal(ids, oi)
        np.testing.assert_array_equal(sc.view(np.uint32), os_.view(np.uint32))
    g.close()


@pytest.mark.parametrize("n_bit", WIDTHS)
def test_decompress_bit_exact(corpus, n_bit):
    ix = _oracle_index(corpus, n_bit)
    g = _lib_index(ix)
    docs = np.array([0, 5, 17, 599, 300, 5], np.uint32)
    got = g.decompress(docs)
    ref = np.concatenate([ix.decompress_doc(int(d)) for d in docs])
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    g.close()


@pytest.mark.parametrize("n_bit", [1, 4, 8])
def test_rerank_docs_fp32_bit_exact(corpus, n_bit):
    """mvr_rerank_docs and mvr_rerank_docs_noalloc over every doc (4 to 140
    tokens), ragged doc counts per query."""
    import torch
    from mvr import _capi
    ix = _oracle_index(corpus, n_bit)
    Q = corpus[1]
    g = _lib_index(ix)
    nq = 3
    docs = np.stack([np.roll(np.arange(ix.n_doc, dtype=np.uint32), 37 * q) for q in range(nq)])
    n_docs = np.array([ix.n_doc, 1, 250], np.uint32)
    Qd = torch.from_numpy(np.ascontiguousarray(Q[:nq])).cuda()
    Dd = torch.from_numpy(docs.view(np.int32)).cuda()
    for fn in ("mvr_rerank_docs", "mvr_rerank_docs_noalloc"):
        out = torch.full((nq, ix.n_doc), -7.0, dtype=torch.float32, device="cuda")
        _capi.check(getattr(_capi.lib(), fn)(g._h, Qd.data_ptr(), nq, Q.shape[1], None, Dd.data_ptr(),
                                             n_docs.ctypes.data_as(C.c_void_p), ix.n_doc, 0, out.data_ptr(),
                                             C.c_void_p(0)))
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        for q in range(nq):
            n = int(n_docs[q])
            ref = ix.rerank(Q[q], docs[q, :n])
            np.testing.assert_array_equal(got[q, :n].view(np.uint32), ref.view(np.uint32), err_msg=fn)
            assert (got[q, n:] == -7.0).all()
    g.close()


@pytest.mark.parametrize("n_bit", [1, 4, 8])
def test_other_widths_refuse_2bit_only_modes(corpus, n_bit):
    import torch
    import mvr
    from mvr import _capi
    ix = _oracle_index(corpus, n_bit)
    Q = corpus[1]
    g = _lib_index(ix)
    for prec, rescore in (("bf16", 0), ("bf16", 32), ("fp8", 0), ("fp32_mfma", 0)):
        with pytest.raises(mvr.MvrError, match="MVR_FP32 only"):
            g.search(Q, 4, 64, 10, precision=prec, rescore=rescore)
    with pytest.raises(mvr.MvrError, match="MVR_FP32 only"):
        g.search(Q, 4, 64, 10, precision="fp32", persistent=True)
    Qd = torch.from_numpy(np.ascontiguousarray(Q[:2])).cuda()
    Dd = torch.zeros((2, 4), dtype=torch.int32, device="cuda")
    out = torch.empty((2, 4), dtype=torch.float32, device="cuda")
    for fn in ("mvr_rerank_docs", "mvr_rerank_docs_noalloc"):
        for prec in (1, 2, 3):
            rc = getattr(_capi.lib(), fn)(g._h, Qd.data_ptr(), 2, Q.shape[1], None, Dd.data_ptr(), None, 4,
                                          prec, out.data_ptr(), C.c_void_p(0))
            assert rc == 4, (fn, prec)  # MVR_ERR_UNSUPPORTED
    # the FP32 search still answers after the refusals
    np.testing.assert_array_equal(g.search(Q, 4, 64, 10)[0], ix.search(Q, 4, 64, 10)[0])
    g.close()
    with pytest.raises(mvr.MvrError, match="2-bit"):
        _lib_index(ix, residual_on_host=True)


def test_bad_widths_are_errors(corpus):
    import mvr
    from mvr import _capi
    ix = _oracle_index(corpus, 2)
    d = _capi.IndexDesc()
    d.n_doc, d.n_tok = ix.doclens.size, ix.codes.size
    d.n_centroid, d.dim = ix.C.shape
    d.n_bit = 3
    keep = [np.ascontiguousarray(a) for a in (ix.doclens, ix.C, ix.codes, ix.residual, ix.weights,
                                               ix.ivf[0].astype(np.uint64), ix.ivf[1].astype(np.uint32),
                                               ix.ivf[2].astype(np.uint32))]
    (d.doclens, d.centroids, d.codes, d.residual, d.weights, d.ivf_offsets, d.ivf_sizes,
     d.ivf_ids) = [a.ctypes.data for a in keep]
    d.weights_per_dim = 1
    d.ivf_n_ele = keep[-1].size
    h = C.c_void_p()
    assert _capi.lib().mvr_index_create_ex(C.byref(d), 0, 0, C.byref(h)) == 1  # MVR_ERR_INVALID
    assert b"must be 1, 2, 4 or 8" in _capi.lib().mvr_last_error()
    with pytest.raises(mvr.MvrError, match="expected 2, 4, 16 or 256"):
        mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights[:, :3], ix.ivf)


@pytest.mark.parametrize("n_bit", [1, 8])
def test_load_from_files(tmp_path, corpus, n_bit):
    """Index files of another width: the loader reads the width from
    weight_l.npy and checks residual_code_l.npy's size against it; the loaded
    index answers like the arrays."""
    import mvr
    from test_cli import write_dataset
    ix = _oracle_index(corpus, n_bit)
    Q = corpus[1]
    write_dataset(str(tmp_path), ix, Q, list(range(Q.shape[0])), ix.n_c, n_bit=n_bit)
    d = os.path.join(str(tmp_path), "Index", "ds", f"igp-n_centroid_{ix.n_c}-n_bit_{n_bit}")
    g = mvr.Index.load(d)
    ids, sc = g.search(Q, 4, 96, 10)
    oi, os_ = ix.search(Q, 4, 96, 10)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(sc.view(np.uint32), os_.view(np.uint32))
    docs = np.arange(0, ix.n_doc, 53, dtype=np.uint32)
    ref = np.concatenate([ix.decompress_doc(int(x)) for x in docs])
    np.testing.assert_array_equal(g.decompress(docs), ref)
    g.close()
    # a residual file whose size does not match the width is an error
    res = np.load(os.path.join(d, "residual_code_l.npy"))
    np.save(os.path.join(d, "residual_code_l.npy"), res[:-1])
    with pytest.raises(mvr.MvrError, match="residual_code_l.npy has"):
        mvr.Index.load(d)


@pytest.mark.parametrize("n_bit", [1, 4])
def test_cli_other_width(tmp_path, corpus, n_bit):
    """IGPGPUPP -num-bit 1/4 over index files of that width, default FP32."""
    import subprocess
    from test_cli import EXE, FLAGS, write_dataset
    ix = _oracle_index(corpus, n_bit)
    Q = corpus[1]
    qids = list(range(Q.shape[0]))
    write_dataset(str(tmp_path), ix, Q, qids, ix.n_c, n_bit=n_bit)
    args = list(FLAGS)
    args[args.index("-num-centroid") + 1] = str(ix.n_c)
    args[args.index("-num-bit") + 1] = str(n_bit)
    env = dict(os.environ, MVR_DATASET_ROOT=str(tmp_path))
    r = subprocess.run([EXE, *args], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    ans = os.path.join(str(tmp_path), "Result", "answer",
                       f"ds-igp-top10-n_centroid_{ix.n_c}-n_bit_{n_bit}-nprobe_4-probe_topk_64-n_thread_1.tsv")
    rows = [l.split("\t") for l in open(ans).read().splitlines()]
    oi, os_ = ix.search(Q, 4, 64, 10)
    assert len(rows) == Q.shape[0] * 10
    for n, (qid, doc, rank, score) in enumerate(rows):
        q, k = divmod(n, 10)
        assert int(doc) == oi[q, k] and int(rank) == k + 1
        assert abs(float(score) - os_[q, k]) <= 1e-5 * abs(os_[q, k])
