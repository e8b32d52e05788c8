"""GPGPUBuildIndexPackGPU.CompressResidualCode.compute_residual_code on the
GPU (libmvr's mvr_compress_host behind the module, as the reference's
compressSQCode, ResidualScalarQuantizationPackGPU.hpp:30-80): the same bytes
as the oracle and as the module's host path."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mod():
    import mvr
    return mvr.build_index_module()


@pytest.mark.parametrize("per_dim", [False, True])
def test_gpu_residual_coder_matches_oracle_and_host(mod, per_dim, monkeypatch):
    rng = np.random.default_rng(21 + per_dim)
    C_ = rng.standard_normal((64, 128)).astype(np.float32)
    codes = rng.integers(0, 64, 20000).astype(np.uint32)
    x = (C_[codes] + 0.1 * rng.standard_normal((20000, 128))).astype(np.float32)
    e = O.residuals(x, codes, C_)
    cut, w = O.quantize_perdim(e, 2) if per_dim else O.quantize_pooled(e, 2)
    # residuals exactly on a cutoff: the bucket is the first cutoff strictly above
    c0 = cut[0] if not per_dim else cut[0, 0]
    x[0, 0] = C_[codes[0], 0] + c0
    sq = mod.CompressResidualCode(centroid_l=C_, cutoff_l=cut, weight_l=w, n_bit=2)
    monkeypatch.delenv("MVR_BUILD_DEVICE", raising=False)
    (res,) = sq.compute_residual_code(vec_l=x, code_l=codes)
    assert sq.device == "gpu"
    got = np.array(res, np.uint8).reshape(-1, 32)
    np.testing.assert_array_equal(got, O.compress(x, codes, C_, cut, 2))
    monkeypatch.setenv("MVR_BUILD_DEVICE", "cpu")
    (res_h,) = sq.compute_residual_code(vec_l=x, code_l=codes)
    assert sq.device == "cpu"
    np.testing.assert_array_equal(np.array(res_h, np.uint8).reshape(-1, 32), got)


def test_gpu_residual_coder_across_chunks():
    """More vectors than one device chunk (4,194,304), through the C ABI
    entry the module calls, against the oracle."""
    import ctypes as C
    import mvr._capi as capi
    rng = np.random.default_rng(5)
    n = (1 << 22) + 3000
    C_ = rng.standard_normal((256, 128)).astype(np.float32)
    codes = rng.integers(0, 256, n).astype(np.uint32)
    x = C_[codes]
    x += rng.standard_normal((n, 128), dtype=np.float32) * np.float32(0.1)
    cut = np.array([-0.07, 0.0, 0.07], np.float32)
    out = np.empty((n, 32), np.uint8)
    st = capi.lib().mvr_compress_host(x.ctypes.data, codes.ctypes.data, n, C_.ctypes.data, 256, 128,
                                      cut.ctypes.data, 0, out.ctypes.data, 0)
    assert st == 0, capi.lib().mvr_last_error()
    np.testing.assert_array_equal(out, O.compress(x, codes, C_, cut, 2))


def test_gpu_mode_rejects_unsupported_shape(mod, monkeypatch):
    C_ = np.zeros((4, 64), np.float32)
    sq = mod.CompressResidualCode(C_, np.array([0.0], np.float32), np.array([-1.0, 1.0], np.float32), 1)
    monkeypatch.setenv("MVR_BUILD_DEVICE", "gpu")
    with pytest.raises(RuntimeError, match="n_bit 2 and dim 128"):
        sq.compute_residual_code(np.zeros((3, 64), np.float32), np.zeros(3, np.uint32))
