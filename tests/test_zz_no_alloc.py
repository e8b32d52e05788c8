"""mvr_rerank_docs_noalloc: the MaxSim of given candidates with its per-query
metadata in the library's static device scratch instead of an allocated
workspace.  Its scores must equal mvr_rerank_docs' bit for bit (same kernels)
and, in precision fp32, the CPU oracle's; two calls in a row, the second with
fewer queries and a shorter candidate list, so stale scratch from the first
call would show."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(fn, g, Qd, nq, t_max, qt, Dd, nd, R, prec, out):
    from mvr import _capi
    qt_p = None if qt is None else qt.ctypes.data_as(C.c_void_p)
    nd_p = None if nd is None else nd.ctypes.data_as(C.c_void_p)
    _capi.check(fn(g._h, Qd.data_ptr(), nq, t_max, qt_p, Dd.data_ptr(), nd_p, R, prec, out.data_ptr(),
                   C.c_void_p(0)))


@pytest.mark.parametrize("prec", [0, 3, 1])
def test_noalloc_rerank_equals_rerank_docs_twice(small, prec):
    import torch
    import mvr
    from mvr import _capi
    corpus, ix, Q = small
    g = mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)
    L = _capi.lib()
    rng = np.random.default_rng(prec)
    for call, (nq, R) in enumerate(((8, 96), (3, 40))):
        docs = rng.integers(0, ix.n_doc, (nq, R)).astype(np.uint32)
        qt = rng.integers(1, Q.shape[1] + 1, nq).astype(np.uint32)
        nd = rng.integers(R // 2, R + 1, nq).astype(np.uint32)
        nd[0] = R
        Qd = torch.from_numpy(np.ascontiguousarray(Q[:nq])).cuda()
        Dd = torch.from_numpy(docs.view(np.int32)).cuda()
        ref = torch.full((nq, R), -7.0, dtype=torch.float32, device="cuda")
        got = torch.full((nq, R), -7.0, dtype=torch.float32, device="cuda")
        _run(L.mvr_rerank_docs, g, Qd, nq, Q.shape[1], qt, Dd, nd, R, prec, ref)
        _run(L.mvr_rerank_docs_noalloc, g, Qd, nq, Q.shape[1], qt, Dd, nd, R, prec, got)
        torch.cuda.synchronize()
        r, o = ref.cpu().numpy(), got.cpu().numpy()
        np.testing.assert_array_equal(o, r, err_msg=f"call {call}")
        # slots past n_docs[q] are untouched by both
        for q in range(nq):
            assert (o[q, nd[q]:] == -7.0).all()
            if prec == 0:
                np.testing.assert_array_equal(o[q, :nd[q]], ix.rerank(Q[q, :qt[q]], docs[q, :nd[q]]))
        # defaults (all tokens, R docs per query)
        got2 = torch.empty((nq, R), dtype=torch.float32, device="cuda")
        ref2 = torch.empty((nq, R), dtype=torch.float32, device="cuda")
        _run(L.mvr_rerank_docs, g, Qd, nq, Q.shape[1], None, Dd, None, R, prec, ref2)
        _run(L.mvr_rerank_docs_noalloc, g, Qd, nq, Q.shape[1], None, Dd, None, R, prec, got2)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got2.cpu().numpy(), ref2.cpu().numpy())
    g.close()


def test_noalloc_rejects_more_queries_than_scratch(small):
    import torch
    import mvr
    from mvr import _capi
    corpus, ix, Q = small
    g = mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)
    nq = 16385
    Qd = torch.zeros((1, 32, 128), dtype=torch.float32, device="cuda")
    Dd = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    out = torch.zeros((1, 1), dtype=torch.float32, device="cuda")
    st = _capi.lib().mvr_rerank_docs_noalloc(g._h, Qd.data_ptr(), nq, 32, None, Dd.data_ptr(), None, 1, 0,
                                             out.data_ptr(), C.c_void_p(0))
    assert st == 1  # MVR_ERR_INVALID
    assert b"16384" in _capi.lib().mvr_last_error()
    g.close()
