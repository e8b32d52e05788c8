"""GPU parity of the search path against the CPU oracle.

Every stage of the HIP pipeline is compared with the oracle on the same
seed-generated index: probed centroid ids and scores, the number of filter
candidates, the filter's top-R ids and scores, the reranked MaxSim scores and
the final top-k.  Precision "fp32" is bit-exact at every stage; "fp32_mfma"
is bit-exact up to the filter and within 1e-5 relative for MaxSim scores, with
top-k lists identical up to ties within that tolerance.  Calls go through the
C ABI (libmvr) via the mvr package.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _index(ix):
    import mvr
    return mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)


def _oracle_stages(ix, Q, P, R, k, q_ntok=None, q_nprobe=None, clip_zero=False):
    from oracle import oracle as O
    n_q, t_max, _ = Q.shape
    out = []
    for q in range(n_q):
        T = t_max if q_ntok is None else int(q_ntok[q])
        Pq = P if q_nprobe is None else int(q_nprobe[q])
        top_ids, top_s = O.centroid_topk(Q[q, :T], ix.C, Pq)
        cand, F = O.filter_all(top_ids, top_s, Pq, ix.ivf, ix.n_doc, basic=clip_zero)
        sel, selF = O.select_top(cand, F, R)
        sc = ix.rerank(Q[q, :T], sel, clip_zero=clip_zero)
        fid, fs = O.select_top(sel, sc, k)
        out.append(dict(top_ids=top_ids, top_s=top_s, cand_n=cand.size, sel=sel, selF=selF, sc=sc,
                        fid=fid, fs=fs))
    return out


def _check(st, ids, sc, ref, q_ntok=None, q_nprobe=None, P=None):
    for q, r in enumerate(ref):
        T = r["top_ids"].shape[0]
        Pq = r["top_ids"].shape[1]
        np.testing.assert_array_equal(st["top_ids"][q, :T, :Pq], r["top_ids"], err_msg=f"q{q} top ids")
        np.testing.assert_array_equal(st["top_s"][q, :T, :Pq], r["top_s"], err_msg=f"q{q} top scores")
        assert st["cand_n"][q] == r["cand_n"], f"q{q} candidates"
        n = r["sel"].size
        np.testing.assert_array_equal(st["sel_ids"][q, :n], r["sel"], err_msg=f"q{q} filter ids")
        np.testing.assert_array_equal(st["sel_F"][q, :n], r["selF"], err_msg=f"q{q} filter scores")
        np.testing.assert_array_equal(st["sel_score"][q, :n], r["sc"], err_msg=f"q{q} MaxSim")
        kk = r["fid"].size
        np.testing.assert_array_equal(ids[q, :kk], r["fid"], err_msg=f"q{q} top-k ids")
        np.testing.assert_array_equal(sc[q, :kk], r["fs"], err_msg=f"q{q} top-k scores")
        assert (ids[q, kk:] == 0xFFFFFFFF).all()


RTOL = 1e-5  # FP32 MaxSim on MFMA vs the sequential-FMA oracle


def _check_topk_up_to_ties(ix, Qq, got_ids, got_sc, ref_ids, ref_sc, clip_zero=False, rtol=RTOL, msg=""):
    """got is a valid top-k: its scores are the oracle scores of its ids within
    rtol, and it differs from the oracle's top-k only by docs whose scores are
    within rtol of the k-th score."""
    k = ref_ids.size
    got = got_ids[:k]
    assert (got != 0xFFFFFFFF).all(), msg
    exact = ix.rerank(Qq, got, clip_zero=clip_zero)
    np.testing.assert_allclose(got_sc[:k], exact, rtol=rtol, atol=1e-6, err_msg=msg)
    kth = float(ref_sc[k - 1])
    tol = rtol * max(1.0, abs(kth)) * 2
    missing = set(ref_ids.tolist()) - set(got.tolist())
    extra = set(got.tolist()) - set(ref_ids.tolist())
    assert len(missing) == len(extra), msg
    sc_of = dict(zip(got.tolist(), exact.tolist()))
    for d in extra:
        assert sc_of[d] >= kth - tol, f"{msg}: doc {d} score {sc_of[d]} below k-th {kth}"
    ref_of = dict(zip(ref_ids.tolist(), ref_sc.tolist()))
    for d in missing:
        assert ref_of[d] <= kth + tol, f"{msg}: dropped doc {d} score {ref_of[d]} above k-th {kth}"


def _check_mfma(st, ids, sc, ref, ix, Q, clip_zero=False):
    for q, r in enumerate(ref):
        T = r["top_ids"].shape[0]
        Pq = r["top_ids"].shape[1]
        np.testing.assert_array_equal(st["top_ids"][q, :T, :Pq], r["top_ids"], err_msg=f"q{q} top ids")
        np.testing.assert_array_equal(st["top_s"][q, :T, :Pq], r["top_s"], err_msg=f"q{q} top scores")
        assert st["cand_n"][q] == r["cand_n"], f"q{q} candidates"
        n = r["sel"].size
        np.testing.assert_array_equal(st["sel_ids"][q, :n], r["sel"], err_msg=f"q{q} filter ids")
        np.testing.assert_array_equal(st["sel_F"][q, :n], r["selF"], err_msg=f"q{q} filter scores")
        np.testing.assert_allclose(st["sel_score"][q, :n], r["sc"], rtol=RTOL, atol=1e-6,
                                   err_msg=f"q{q} MaxSim")
        _check_topk_up_to_ties(ix, Q[q, :T], ids[q], sc[q], r["fid"], r["fs"], clip_zero, msg=f"q{q} top-k")


def test_small_fp32_bit_exact(small):
    corpus, ix, Q = small
    g = _index(ix)
    ids, sc, st = g.search(Q, 4, 64, 10, stages=True)
    _check(st, ids, sc, _oracle_stages(ix, Q, 4, 64, 10))


def test_config1_fp32_bit_exact(c1):
    """Config 1: 8,192 docs, 1,024 centroids, 64 queries, P=4, R=256, k=10."""
    corpus, ix, Q = c1
    g = _index(ix)
    ids, sc, st = g.search(Q, 4, 256, 10, stages=True)
    _check(st, ids, sc, _oracle_stages(ix, Q, 4, 256, 10))


def test_config1_probe_topk_above_4096_bit_exact(c1):
    """probe_topk 6,000 of the 8,192 docs (the 16,384-key select and top-k
    variants, all candidates sorted in LDS), fp32 bit-exact at every stage."""
    corpus, ix, Q = c1
    g = _index(ix)
    ids, sc, st = g.search(Q[:8], 16, 6000, 100, stages=True)
    _check(st, ids, sc, _oracle_stages(ix, Q[:8], 16, 6000, 100))
    assert (st["cand_n"] > 4096).all()


def test_config1_search_batch_matches_oracle_batch(c1):
    corpus, ix, Q = c1
    g = _index(ix)
    ids, sc = g.search(Q, 8, 512, 100, batch=16)
    oi, os_ = ix.search(Q, 8, 512, 100)
    np.testing.assert_array_equal(ids, oi)
    np.testing.assert_array_equal(sc, os_)


def test_ragged_tokens_and_per_query_nprobe(small):
    """Uneven token counts and per-query nprobe (the batching workload's shape)."""
    corpus, ix, Q = small
    rng = np.random.default_rng(7)
    q_ntok = rng.integers(8, 33, Q.shape[0]).astype(np.uint32)
    q_np = rng.choice([2, 4, 8, 16, 32], Q.shape[0]).astype(np.uint32)
    g = _index(ix)
    ids, sc, st = g.search(Q, 2, 100, 20, q_ntok=q_ntok, q_nprobe=q_np, stages=True)
    ref = _oracle_stages(ix, Q, 2, 100, 20, q_ntok, q_np)
    _check(st, ids, sc, ref)


def test_basic_variant_clip_zero(small):
    """IGPGPUBasic semantics: per-token max clipped at 0 (zero padding)."""
    corpus, ix, Q = small
    g = _index(ix)
    Qn = -Q  # negated queries make many per-token maxima negative
    ids, sc, st = g.search(Qn, 4, 64, 10, clip_zero=True, stages=True)
    _check(st, ids, sc, _oracle_stages(ix, Qn, 4, 64, 10, clip_zero=True))


def test_fewer_candidates_than_probe_topk(small):
    """R larger than the candidate count: outputs padded with (UINT32_MAX, -inf)."""
    corpus, ix, Q = small
    g = _index(ix)
    Q1 = np.ascontiguousarray(Q[:4, :1])     # one token, one probe: a few hundred candidates
    ids, sc, st = g.search(Q1, 1, 1024, 1024, stages=True)
    ref = _oracle_stages(ix, Q1, 1, 1024, 1024)
    _check(st, ids, sc, ref)
    for q, r in enumerate(ref):
        assert r["cand_n"] < 1024
        assert np.isneginf(sc[q, r["cand_n"]:]).all()


def test_decompress_bit_exact(c1):
    corpus, ix, Q = c1
    g = _index(ix)
    docs = np.array([0, 5, 8191, 17, 17], np.uint32)
    out = g.decompress(docs)
    ref = np.concatenate([ix.decompress_doc(int(d)) for d in docs])
    np.testing.assert_array_equal(out, ref)


def test_invalid_arguments_raise(small):
    import mvr
    corpus, ix, Q = small
    g = _index(ix)
    with pytest.raises(mvr.MvrError):
        g.search(Q, 4, 5, 10)            # probe_topk < topk
    with pytest.raises(mvr.MvrError):
        g.search(Q, 200, 64, 10)         # nprobe > n_centroid
    with pytest.raises(mvr.MvrError):
        g.search(np.zeros((1, 33, 128), np.float32), 4, 64, 10)


def _bf16_maxsim(ix, Qq, docs, clip_zero=False):
    """MaxSim with bf16 inputs as the BF16 kernel defines it: query, centroid
    and residual values rounded to bf16 (nearest even), products and sums in
    float64: s = bf16(q).(bf16(C[code]) + bf16(w[d][bucket]))."""
    from oracle import oracle as O
    qb = O.bf16_round(Qq).astype(np.float64)
    Cb = O.bf16_round(ix.C).astype(np.float64)
    Wb = O.bf16_round(ix.W).astype(np.float64)
    out = []
    for d in docs:
        off, L = int(ix.docoff[d]), int(ix.doclens[d])
        res = ix.residual[off:off + L]
        b = np.stack([(res[:, j // 4] >> (6 - 2 * (j % 4))) & 3 for j in range(128)], 1)
        X = Cb[ix.codes[off:off + L]] + Wb[np.arange(128)[None, :], b]
        m = (qb @ X.T).max(axis=1)
        if clip_zero:
            m = np.maximum(m, 0)
        out.append(m.sum())
    return np.array(out)


def _fp8_maxsim(ix, Qq, docs, clip_zero=False):
    """MaxSim as the FP8 kernel defines it, in float64: query tokens and
    centroid rows encoded per row (s = max|.|/448, e4m3 of x*448/max|.|),
    residual weights with one global scale, score = s_q (s_c C8.q8 + s_r R8.q8)."""
    from oracle import oracle as O
    q8, sq = O.fp8_scaled(Qq)
    c8, sc = O.fp8_scaled(ix.C)
    wmax = np.float32(np.abs(ix.W).max())
    sr = np.float32(wmax / np.float32(448.0))
    w8 = O.e4m3_round((ix.W * (np.float32(1.0) / sr)).astype(np.float32)).astype(np.float64)
    q8, sq, c8, sc = q8.astype(np.float64), sq.astype(np.float64), c8.astype(np.float64), sc.astype(np.float64)
    out = []
    for d in docs:
        off, L = int(ix.docoff[d]), int(ix.doclens[d])
        res = ix.residual[off:off + L]
        b = np.stack([(res[:, j // 4] >> (6 - 2 * (j % 4))) & 3 for j in range(128)], 1)
        codes = ix.codes[off:off + L]
        SC = (q8 @ c8[codes].T) * sc[codes, 0][None, :]
        SR = (q8 @ w8[np.arange(128)[None, :], b].T) * float(sr)
        m = (sq * (SC + SR)).max(axis=1)
        if clip_zero:
            m = np.maximum(m, 0)
        out.append(m.sum())
    return np.array(out)


def test_fp8_rerank_matches_fp8_model(small):
    """FP8 MaxSim against its float64 definition on e4m3-encoded inputs (the
    encoding modelled bit for bit on the CPU); differences come only from FP32
    accumulation in the matrix unit."""
    corpus, ix, Q = small
    g = _index(ix)
    ids, sc, st = g.search(Q, 4, 64, 10, precision="fp8", rescore=0, stages=True)
    for q in range(Q.shape[0]):
        n = int((st["sel_ids"][q] != 0xFFFFFFFF).sum())
        ref = _fp8_maxsim(ix, Q[q], st["sel_ids"][q, :n])
        np.testing.assert_allclose(st["sel_score"][q, :n], ref, rtol=1e-4, atol=1e-4)


def test_fp8_with_fp32_rescore(c1):
    """FP8 MaxSim over R=256 then FP32 rescoring of the best 64: the answer is
    the FP32 top-10 of the FP8 path's best 64 (ties within 1e-5), the FP8
    scores stay close to FP32, and the centroid stage is the FP8 probe (its
    model: exact FP32 top-P within the best 2P tiles by FP8 maxima)."""
    from oracle import oracle as O
    corpus, ix, Q = c1
    g = _index(ix)
    ids, sc, st = g.search(Q, 4, 256, 10, precision="fp8", rescore=64, stages=True)
    ti, ts, amb = O.centroid_topk_fp8(Q, ix.C, 4)
    ok = ~amb.reshape(Q.shape[0], Q.shape[1])
    np.testing.assert_array_equal(st["top_ids"][ok], ti.reshape(Q.shape[0], Q.shape[1], 4)[ok])
    for q in range(Q.shape[0]):
        n = min(int(st["cand_n"][q]), 256)
        cid = st["sel_ids"][q, :n].astype(np.int64)
        cs = st["sel_score"][q, :n]
        exact = ix.rerank(Q[q], cid.astype(np.uint32))
        assert np.median(np.abs(cs - exact) / np.abs(exact)) < 0.02
        best = np.lexsort((cid, -cs.astype(np.float64)))[:64]
        rid = cid[best]
        rs = exact[best]
        top = np.lexsort((rid, -rs.astype(np.float64)))[:10]
        _check_topk_up_to_ties(ix, Q[q], ids[q], sc[q], rid[top].astype(np.uint32), rs[top], msg=f"q{q}")


def test_bf16_rerank_scores_within_tolerance(small):
    corpus, ix, Q = small
    g = _index(ix)
    ids, sc, st = g.search(Q, 4, 64, 10, precision="bf16", rescore=0, stages=True)
    for q in range(Q.shape[0]):
        n = int((st["sel_ids"][q] != 0xFFFFFFFF).sum())
        ref = _bf16_maxsim(ix, Q[q], st["sel_ids"][q, :n])
        np.testing.assert_allclose(st["sel_score"][q, :n], ref, rtol=0, atol=2e-4)


def test_fp32_mfma_config1_within_tolerance(c1):
    corpus, ix, Q = c1
    g = _index(ix)
    ids, sc, st = g.search(Q, 4, 256, 10, precision="fp32_mfma", stages=True)
    _check_mfma(st, ids, sc, _oracle_stages(ix, Q, 4, 256, 10), ix, Q)


def test_fp32_mfma_ragged_clip_zero(small):
    """Ragged query token counts, per-query nprobe and the basic variant's
    clip at zero on the MFMA kernel."""
    corpus, ix, Q = small
    g = _index(ix)
    rng = np.random.default_rng(7)
    qt = rng.integers(1, Q.shape[1] + 1, Q.shape[0]).astype(np.uint32)
    qp = rng.integers(1, 9, Q.shape[0]).astype(np.uint32)
    for clip in (False, True):
        ids, sc, st = g.search(Q, 8, 64, 10, precision="fp32_mfma", q_ntok=qt, q_nprobe=qp, clip_zero=clip,
                               stages=True)
        ref = _oracle_stages(ix, Q, 8, 64, 10, q_ntok=qt, q_nprobe=qp, clip_zero=clip)
        _check_mfma(st, ids, sc, ref, ix, Q, clip_zero=clip)


def test_fp32_mfma_rerank_docs_long_docs(c1):
    """mvr_rerank_docs on MFMA over every doc of the corpus (token counts 16 to
    128, so 1 to 4 tiles of 32 with ragged tails) against the oracle."""
    import torch
    from mvr import _capi
    import ctypes as C
    corpus, ix, Q = c1
    g = _index(ix)
    nq = 4
    docs = np.tile(np.arange(ix.n_doc, dtype=np.uint32), (nq, 1))
    Qd = torch.from_numpy(np.ascontiguousarray(Q[:nq])).cuda()
    Dd = torch.from_numpy(docs.view(np.int32)).cuda()
    out = torch.empty((nq, ix.n_doc), dtype=torch.float32, device="cuda")
    _capi.check(_capi.lib().mvr_rerank_docs(g._h, Qd.data_ptr(), nq, Q.shape[1], None, Dd.data_ptr(), None,
                                            ix.n_doc, 3, out.data_ptr(), C.c_void_p(0)))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for q in range(nq):
        np.testing.assert_allclose(got[q], ix.rerank(Q[q], docs[q]), rtol=RTOL, atol=1e-6)


def test_bf16_with_fp32_rescore_returns_fp32_scores(c1):
    """BF16 MaxSim over R=256, FP32 rescoring (MFMA) of the best 64, top-10.
    Check of the rescore logic: from the BF16 path's own candidates and BF16
    scores (stage outputs) take the best 64 by (score desc, id asc), score them
    with the FP32 oracle, keep the best 10 -- that must be the answer up to
    ties within 1e-5, with FP32 scores.  Also: the answer agrees with the plain
    FP32 pipeline's top-10."""
    corpus, ix, Q = c1
    g = _index(ix)
    ids, sc, st = g.search(Q, 4, 256, 10, precision="bf16", rescore=64, stages=True)
    oi, os_ = ix.search(Q, 4, 256, 10)
    for q in range(Q.shape[0]):
        n = min(int(st["cand_n"][q]), 256)
        cid = st["sel_ids"][q, :n].astype(np.int64)
        cs = st["sel_score"][q, :n]
        best = np.lexsort((cid, -cs.astype(np.float64)))[:64]
        rid = cid[best]
        rs = ix.rerank(Q[q], rid.astype(np.uint32))
        top = np.lexsort((rid, -rs.astype(np.float64)))[:10]
        _check_topk_up_to_ties(ix, Q[q], ids[q], sc[q], rid[top].astype(np.uint32), rs[top], msg=f"q{q}")
    same = np.mean([len(set(a.tolist()) & set(b.tolist())) / 10 for a, b in zip(ids, oi)])
    assert same >= 0.8, same


def test_bf16_centroid_topk_matches_rounded_input_oracle(c1):
    """BF16 centroid scoring: exact top-P of S' = fmaf-chain(bf16(q), bf16(C)) + 1
    (the oracle's FP32 scoring applied to bf16-rounded inputs); the MFMA tile
    maxima only choose which tiles are rescanned, within a proven error
    margin, so the result is exact."""
    from oracle import oracle as O
    corpus, ix, Q = c1
    g = _index(ix)
    for P in (1, 8, 32):
        ids, sc, st = g.search(Q, P, 64, 10, precision="bf16", stages=True)
        ti, ts = O.centroid_topk(O.bf16_round(Q), O.bf16_round(ix.C), P)
        np.testing.assert_array_equal(st["top_ids"], ti.reshape(Q.shape[0], Q.shape[1], P))
        np.testing.assert_array_equal(st["top_s"], ts.reshape(Q.shape[0], Q.shape[1], P))


def test_centroid_topk_partial_tiles_and_tiny_nprobe():
    """Centroid count not a multiple of the 32-centroid tile (100 centroids,
    4 tiles, the last one partial) and nprobe from 1 to 32: FP32 and
    FP32-MFMA top-P bit-exact with the oracle, BF16 exact on rounded inputs."""
    from oracle import oracle as O
    from oracle import synth
    corpus, ix, Q = synth.config1(n_doc=512, n_comp=128, n_c=100, n_q=8, n_iter=3)
    g = _index(ix)
    for P in (1, 3, 32):
        for prec in ("fp32", "fp32_mfma"):
            _, _, st = g.search(Q, P, 32, 10, precision=prec, stages=True)
            ti, ts = O.centroid_topk(Q, ix.C, P)
            np.testing.assert_array_equal(st["top_ids"], ti.reshape(Q.shape[0], Q.shape[1], P))
            np.testing.assert_array_equal(st["top_s"], ts.reshape(Q.shape[0], Q.shape[1], P))
        _, _, st = g.search(Q, P, 32, 10, precision="bf16", stages=True)
        ti, ts = O.centroid_topk(O.bf16_round(Q), O.bf16_round(ix.C), P)
        np.testing.assert_array_equal(st["top_ids"], ti.reshape(Q.shape[0], Q.shape[1], P))
        np.testing.assert_array_equal(st["top_s"], ts.reshape(Q.shape[0], Q.shape[1], P))


def test_timings_work_counts_exact(c1):
    """The work counters of a timed search equal sums over its stage outputs:
    reranked doc tokens, filter candidates, and query tokens x reranked doc
    tokens (ragged token counts, several passes)."""
    corpus, ix, Q = c1
    g = _index(ix)
    rng = np.random.default_rng(21)
    qt = rng.integers(1, 33, Q.shape[0]).astype(np.uint32)
    ids, sc, st, tm = g.search(Q, 8, 300, 10, q_ntok=qt, batch=24, stages=True, timings=True)
    toks = np.array([int(ix.doclens[st["sel_ids"][q, :st["rerank_n"][q]]].astype(np.uint64).sum())
                     for q in range(Q.shape[0])], np.uint64)
    assert tm["rerank_tokens"] == int(toks.sum())
    assert tm["candidates"] == int(st["cand_n"].astype(np.uint64).sum())
    assert tm["rerank_pairs"] == int((toks * qt.astype(np.uint64)).sum())
    assert tm["n_pass"] == 3


@pytest.mark.parametrize("R,k", [(4096, 4096), (4096, 100), (1, 1)])
def test_largest_and_smallest_result_sizes(c1, R, k):
    """probe_topk at its maximum (4,096: the radix select with a 4,096-key
    output, the final top-k over 4,096 scores) and at its minimum, nprobe 32
    so that the candidates exceed 4,096; bit-exact against the oracle."""
    corpus, ix, Q = c1
    g = _index(ix)
    Qs = Q[:4]
    ids, sc, st = g.search(Qs, 32, R, k, stages=True)
    assert (st["cand_n"] > 4096).all()
    _check(st, ids, sc, _oracle_stages(ix, Qs, 32, R, k))


def test_zero_queries(small):
    """An empty batch returns without work, for host and device buffers."""
    import torch
    corpus, ix, Q = small
    g = _index(ix)
    ids, sc = g.search(Q[:0], 4, 64, 10)
    assert ids.shape == (0, 10) and sc.shape == (0, 10)
    Qd = torch.zeros((0, Q.shape[1], Q.shape[2]), dtype=torch.float32, device="cuda")
    oi = torch.empty((0, 10), dtype=torch.int32, device="cuda")
    os_ = torch.empty((0, 10), dtype=torch.float32, device="cuda")
    g.search_device(Qd, oi, os_, 4, 64, 10)
    torch.cuda.synchronize()


def test_index_create_rejects_bad_codes_and_lists(small):
    """Codes must index a centroid and IVF lists must be strictly ascending
    doc ids below n_doc (the kernels index centroid rows by code and the
    filter's segment runs assume sorted lists); both are checked at creation."""
    import mvr
    corpus, ix, Q = small
    codes = ix.codes.copy()
    codes[17] = ix.n_c
    with pytest.raises(mvr.MvrError, match="codes are >= n_centroid"):
        mvr.Index.from_arrays(ix.doclens, ix.C, codes, ix.residual, ix.weights, ix.ivf)
    off, size, ids = ix.ivf
    c = int(np.argmax(size))
    for bad in ("swap", "range"):
        ids2 = ids.copy()
        o = int(off[c])
        if bad == "swap":
            ids2[o], ids2[o + 1] = ids2[o + 1], ids2[o]
        else:
            ids2[o + int(size[c]) - 1] = ix.n_doc
        with pytest.raises(mvr.MvrError, match="IVF entries"):
            mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, (off, size, ids2))
    g = _index(ix)  # the unmodified index is accepted
    g.close()
