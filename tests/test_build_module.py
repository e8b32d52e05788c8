"""The GPGPUBuildIndexPackGPU module (product C++) against the oracle, and its
reference surface (src/app/build_index/GPGPUBuildIndexPackGPU.cpp:14-42)."""
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def mod():
    import mvr
    return mvr.build_index_module()


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(11)
    C_ = rng.standard_normal((32, 128)).astype(np.float32)
    codes = rng.integers(0, 32, 3000).astype(np.uint32)
    x = (C_[codes] + 0.1 * rng.standard_normal((3000, 128))).astype(np.float32)
    return x, C_, codes


def test_compute_quantized_scalar_matches_oracle(mod, data):
    x, C_, codes = data
    cut, w = mod.compute_quantized_scalar(item_vec_l=x, centroid_l=C_, code_l=codes, n_bit=2)
    oc, ow = O.quantize_pooled(O.residuals(x, codes, C_), 2)
    np.testing.assert_array_equal(np.float32(cut), oc)
    np.testing.assert_array_equal(np.float32(w), ow)


def test_per_dim_quantizer_matches_oracle(mod, data):
    x, C_, codes = data
    cut, w = mod.compute_quantized_scalar_per_dim(x, C_, codes, 2)
    oc, ow = O.quantize_perdim(O.residuals(x, codes, C_), 2)
    np.testing.assert_array_equal(cut, oc)
    np.testing.assert_array_equal(w, ow)


@pytest.mark.parametrize("per_dim", [False, True])
def test_compress_residual_code_matches_oracle(mod, data, per_dim):
    x, C_, codes = data
    e = O.residuals(x, codes, C_)
    cut, w = O.quantize_perdim(e, 2) if per_dim else O.quantize_pooled(e, 2)
    sq = mod.CompressResidualCode(centroid_l=C_, cutoff_l=cut, weight_l=w, n_bit=2)
    assert sq.n_val_per_vec == 32
    (res,) = sq.compute_residual_code(vec_l=x, code_l=codes)
    np.testing.assert_array_equal(np.array(res, np.uint8).reshape(-1, 32), O.compress(x, codes, C_, cut, 2))


def test_ivf_index_file_format(mod, tmp_path):
    rng = np.random.default_rng(12)
    lens = rng.integers(1, 30, 200).astype(np.uint32)
    codes = rng.integers(0, 40, int(lens.sum())).astype(np.uint32)
    ivf = mod.IVFIndex(vq_code_l=codes.tolist(), item_n_vec_l=lens.tolist(), n_vec=int(lens.sum()),
                       n_item=200, n_centroid=40)
    ivf.build()
    path = str(tmp_path / "ivf.index")
    ivf.save(path)
    raw = open(path, "rb").read()
    n_c, n_ele = struct.unpack_from("<IQ", raw, 0)
    off = np.frombuffer(raw, np.uint64, n_c, 12)
    size = np.frombuffer(raw, np.uint32, n_c, 12 + 8 * n_c)
    ids = np.frombuffer(raw, np.uint32, n_ele, 12 + 12 * n_c)
    assert len(raw) == 12 + 12 * n_c + 4 * n_ele
    oo, os_, oi = O.ivf_build(codes, lens, 40)
    np.testing.assert_array_equal(off, oo)
    np.testing.assert_array_equal(size, os_)
    np.testing.assert_array_equal(ids, oi)


def test_invalid_arguments_raise(mod, data):
    x, C_, codes = data
    with pytest.raises(ValueError):
        mod.compute_quantized_scalar(x, C_, codes, 3)
    with pytest.raises(ValueError):
        mod.CompressResidualCode(C_, np.zeros(2, np.float32), np.zeros(4, np.float32), 2)
    bad = codes.copy()
    bad[0] = 99
    with pytest.raises(ValueError):
        mod.compute_quantized_scalar(x, C_, bad, 2)
