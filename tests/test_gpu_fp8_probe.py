"""FP8 centroid scoring (the FP8 mode's probe): e4m3 MFMA tile maxima with
per-token and per-centroid scales, then the exact FP32 top-P among the
centroids of the 2P tiles with the highest FP8 maxima.  Checked against the
oracle's model of the same probe (oracle.centroid_topk_fp8: float64 tile
maxima of the e4m3 encodings, the exact FP32 chain for the scores): ids and
scores bit-exact on every row whose P'-th tile maximum has no near tie, and
against the exact FP32 top-P, which it reaches on most rows.  Calls go
through the C ABI via the mvr package."""
import numpy as np
import pytest

from test_gpu_parity import _index

pytestmark = pytest.mark.gpu


def _check_probe(ix, Q, P, min_exact):
    from oracle import oracle as O
    g = _index(ix)
    _, _, st = g.search(Q, P, 64, 10, precision="fp8", rescore=0, stages=True)
    g.close()
    n_q, T = Q.shape[0], Q.shape[1]
    mi, ms, amb = O.centroid_topk_fp8(Q, ix.C, P)
    gi = st["top_ids"].reshape(n_q * T, -1)[:, :P]
    gs = st["top_s"].reshape(n_q * T, -1)[:, :P]
    ok = ~amb
    assert ok.mean() > 0.95
    np.testing.assert_array_equal(gi[ok], mi[ok])
    np.testing.assert_array_equal(gs[ok].view(np.uint32), ms[ok].view(np.uint32))
    ei, es = O.centroid_topk(Q, ix.C, P)
    exact = (gi == ei).all(1).mean()
    assert exact >= min_exact, exact
    return exact


def test_fp8_probe_config1(c1):
    corpus, ix, Q = c1
    _check_probe(ix, Q, 4, 0.9)


@pytest.mark.parametrize("P", [1, 8, 16])
def test_fp8_probe_partial_tiles_and_widths(P):
    """1,000 centroids (a partial last tile), P' = 2, 16 and 32 (every tile)."""
    from oracle import synth
    corpus = synth.make_corpus(31, 1024, 256)
    ix = synth.build_index(corpus, 1000, n_iter=3)
    Q = synth.make_queries(32, corpus.centers, 16)
    exact = _check_probe(ix, Q, P, 0.8)
    if 2 * P >= 32:
        assert exact == 1.0
