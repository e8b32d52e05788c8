"""mvr_index_load(_ex) reads vq_code_l.npy and residual_code_l.npy as streams
(chunks through two pinned buffers and hipMemcpyAsync for a device-resident
index, straight into the pinned arrays for host-resident residuals), checking
codes per chunk.  The loaded index must answer exactly like one created from
the same arrays; corrupt files must fail with an error, not a fault.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dir(tmp_path, ix, Q):
    from test_cli import write_dataset
    write_dataset(str(tmp_path), ix, Q, list(range(Q.shape[0])), ix.n_c)
    return os.path.join(str(tmp_path), "Index", "ds", f"igp-n_centroid_{ix.n_c}-n_bit_2")


@pytest.fixture
def small_chunks(monkeypatch):
    # 4 KiB reads: the small index's 1 MB of codes and 4 MB of residuals take
    # hundreds of chunks, so both staging buffers are reused many times
    monkeypatch.setenv("MVR_LOAD_CHUNK_BYTES", "4096")


@pytest.mark.parametrize("host", [False, True])
def test_streamed_load_matches_arrays(tmp_path, small, small_chunks, host):
    import mvr
    corpus, ix, Q = small
    d = _dir(tmp_path, ix, Q)
    ref = mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)
    g = mvr.Index.load(d, residual_on_host=host)
    for prec, rescore in (("fp32", 0), ("fp32_mfma", 0), ("bf16", 32)):
        r_ids, r_sc = ref.search(Q, 4, 64, 10, precision=prec, rescore=rescore)
        ids, sc = g.search(Q, 4, 64, 10, precision=prec, rescore=rescore)
        np.testing.assert_array_equal(ids, r_ids, err_msg=prec)
        np.testing.assert_array_equal(sc, r_sc, err_msg=prec)
    if not host:
        docs = np.arange(0, ix.n_doc, 97, dtype=np.uint32)
        np.testing.assert_array_equal(g.decompress(docs), ref.decompress(docs))
    g.close()
    ref.close()


def test_streamed_load_uneven_last_chunk(tmp_path, small, monkeypatch):
    """A chunk size that divides neither array: the last chunk is partial."""
    import mvr
    monkeypatch.setenv("MVR_LOAD_CHUNK_BYTES", str(4096 * 3 + 4))
    corpus, ix, Q = small
    d = _dir(tmp_path, ix, Q)
    ref = mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)
    g = mvr.Index.load(d)
    np.testing.assert_array_equal(g.search(Q, 4, 64, 10)[0], ref.search(Q, 4, 64, 10)[0])


def test_code_out_of_range_is_an_error(tmp_path, small, small_chunks):
    import mvr
    corpus, ix, Q = small
    d = _dir(tmp_path, ix, Q)
    codes = ix.codes.astype(np.int32).copy()
    codes[len(codes) // 2 + 3] = ix.n_c  # in a later chunk
    np.save(os.path.join(d, "vq_code_l.npy"), codes)
    for host in (False, True):
        with pytest.raises(mvr.MvrError):
            mvr.Index.load(d, residual_on_host=host)


def test_truncated_residuals_are_an_error(tmp_path, small, small_chunks):
    import mvr
    corpus, ix, Q = small
    d = _dir(tmp_path, ix, Q)
    p = os.path.join(d, "residual_code_l.npy")
    with open(p, "rb") as f:
        raw = f.read()
    with open(p, "wb") as f:
        f.write(raw[:len(raw) - 100])  # header still claims the full size
    for host in (False, True):
        with pytest.raises(mvr.MvrError, match="truncated"):
            mvr.Index.load(d, residual_on_host=host)
