"""In-bounds check of the MFMA lane layouts and the FP8 encoding the bf16/fp8
kernels assume on gfx950 (one wave, exact small-integer inputs).

Assumed (v_mfma_f32_32x32x16_{bf16,fp8_fp8}):
  A: lane l holds A[l % 32][8*(l // 32) + e], e = 0..7
  B: lane l holds B[8*(l // 32) + e][l % 32]
  D: lane l, element j holds D[8*(j // 4) + 4*(l // 32) + j % 4][l % 32]
Assumed (v_mfma_f32_16x16x32_bf16, the BF16 MaxSim):
  A: lane l holds A[l % 16][8*(l // 16) + e], e = 0..7
  B: lane l holds B[8*(l // 16) + e][l % 16]
  D: lane l, element j holds D[4*(l // 16) + j][l % 16]
Assumed (v_mfma_f32_32x32x64_f8f6f4 with e4m3 operands, the FP8 centroid
tile maxima): lane l holds A[l % 32][32*(l // 32) + e] and
B[32*(l // 32) + e][l % 32], e = 0..31; D as for 32x32x16.
FP8: OCP e4m3 (bias 7, 1.0 = 0x38).
global_load_lds (16 bytes): lane l's piece lands at the LDS destination + the
immediate offset + 16 l, the offset also being added to the global address
(the MaxSim kernels stage centroid rows this way).
FP32 sqrtf and '/' in device code are correctly rounded (the PLAID refine
kernel's normalisation relies on it; __fsqrt_rn is not, on this toolchain).
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe():
    lib = ctypes.CDLL(os.path.join(HERE, "probes", "libmvr_probe.so"))
    vp = ctypes.c_void_p
    lib.probe_mfma.argtypes = [ctypes.c_int, vp, vp, vp]
    lib.probe_mfma.restype = ctypes.c_int
    lib.probe_cvt_fp8.argtypes = [vp, vp, ctypes.c_uint32]
    lib.probe_cvt_fp8.restype = ctypes.c_int
    lib.probe_divsqrt.argtypes = [vp, vp, vp, ctypes.c_uint32]
    lib.probe_divsqrt.restype = ctypes.c_int
    lib.probe_gll.argtypes = [vp, vp, vp]
    lib.probe_gll.restype = ctypes.c_int
    return lib


def bf16_bits(x):
    return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def pack_operands(A, B):
    a = np.zeros((64, 8), A.dtype)
    b = np.zeros((64, 8), B.dtype)
    for l in range(64):
        for e in range(8):
            a[l, e] = A[l % 32, 8 * (l // 32) + e]
            b[l, e] = B[8 * (l // 32) + e, l % 32]
    return a, b


def unpack_d(d):
    D = np.zeros((32, 32), np.float32)
    for l in range(64):
        for j in range(16):
            D[8 * (j // 4) + 4 * (l // 32) + j % 4, l % 32] = d[l, j]
    return D


def test_bf16_32x32x16_layout(probe):
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    B = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    a, b = pack_operands(bf16_bits(A), bf16_bits(B))
    d = np.zeros((64, 16), np.float32)
    assert probe.probe_mfma(0, a.ctypes.data, b.ctypes.data, d.ctypes.data) == 0
    np.testing.assert_array_equal(unpack_d(d), A @ B)


def test_bf16_16x16x32_layout(probe):
    rng = np.random.default_rng(2)
    A = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    a = np.zeros((64, 8), np.uint16)
    b = np.zeros((64, 8), np.uint16)
    for l in range(64):
        for e in range(8):
            a[l, e] = bf16_bits(A[l % 16, 8 * (l // 16) + e])
            b[l, e] = bf16_bits(B[8 * (l // 16) + e, l % 16])
    d = np.zeros((64, 16), np.float32)
    assert probe.probe_mfma(2, a.ctypes.data, b.ctypes.data, d.ctypes.data) == 0
    D = np.zeros((16, 16), np.float32)
    for l in range(64):
        for j in range(4):
            D[4 * (l // 16) + j, l % 16] = d[l, j]
    np.testing.assert_array_equal(D, A @ B)
    assert not d[:, 4:].any()


def ocp_e4m3(v):
    """OCP FP8 E4M3 encoding of exactly representable values."""
    if v == 0:
        return 0
    s = 0x80 if v < 0 else 0
    v = abs(v)
    e = int(np.floor(np.log2(v)))
    m = v / 2.0 ** e - 1.0
    return s | ((e + 7) << 3) | int(round(m * 8))


def test_fp8_conversion_is_ocp_e4m3(probe):
    vals = np.array([1.0, 2.0, -3.0, 0.5, 448.0, -0.015625, 1.125, 6.0], np.float32)
    out = np.zeros(vals.size // 2, np.uint32)
    assert probe.probe_cvt_fp8(vals.ctypes.data, out.ctypes.data, vals.size // 2) == 0
    got = [(int(o) >> (8 * k)) & 0xFF for o in out for k in range(2)]
    assert got == [ocp_e4m3(float(v)) for v in vals]


def test_fp8_32x32x16_layout(probe):
    rng = np.random.default_rng(1)
    A = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    B = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    enc = np.vectorize(ocp_e4m3)
    a, b = pack_operands(enc(A).astype(np.uint8), enc(B).astype(np.uint8))
    d = np.zeros((64, 16), np.float32)
    assert probe.probe_mfma(1, a.ctypes.data, b.ctypes.data, d.ctypes.data) == 0
    np.testing.assert_array_equal(unpack_d(d), A @ B)


def test_fp32_sqrt_and_division_correctly_rounded(probe):
    rng = np.random.default_rng(0)
    n = 1 << 18
    a = (np.abs(rng.standard_normal(n)) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)
    b = ((rng.standard_normal(n) + 0.01) * np.exp2(rng.integers(-10, 10, n))).astype(np.float32)
    a[:4] = [0.0, 1e-40, 2.0, 3.4e38]        # zero, subnormal, exact, large
    out = np.empty((n, 4), np.float32)
    assert probe.probe_divsqrt(a.ctypes.data, b.ctypes.data, out.ctypes.data, n) == 0
    np.testing.assert_array_equal(out[:, 2].view(np.uint32), np.sqrt(a).view(np.uint32))
    np.testing.assert_array_equal(out[:, 3].view(np.uint32), (a / b).view(np.uint32))


def test_global_load_lds_lane_layout(probe):
    rng = np.random.default_rng(3)
    src = rng.integers(0, 2**32, (128, 4), dtype=np.uint64).astype(np.uint32)
    perm = rng.permutation(64).astype(np.uint32)
    out = np.zeros((192, 4), np.uint32)
    assert probe.probe_gll(src.ctypes.data, perm.ctypes.data, out.ctypes.data) == 0
    want = np.full((192, 4), 0xDEAD, np.uint32)
    want[32:96] = src[perm]            # destination buf + 32 pieces, offset 0
    want[96:160] = src[perm + 64]      # offset 1024 bytes = 64 pieces, on both addresses
    np.testing.assert_array_equal(out, want)


def test_fp8_32x32x64_layout(probe):
    rng = np.random.default_rng(4)
    A = rng.integers(-4, 5, (32, 64)).astype(np.float32)
    B = rng.integers(-4, 5, (64, 32)).astype(np.float32)
    enc = np.vectorize(ocp_e4m3)
    Ae, Be = enc(A).astype(np.uint8), enc(B).astype(np.uint8)
    a = np.zeros((64, 32), np.uint8)
    b = np.zeros((64, 32), np.uint8)
    for l in range(64):
        a[l] = Ae[l % 32, 32 * (l // 32):32 * (l // 32) + 32]
        b[l] = Be[32 * (l // 32):32 * (l // 32) + 32, l % 32]
    d = np.zeros((64, 16), np.float32)
    assert probe.probe_mfma(3, a.ctypes.data, b.ctypes.data, d.ctypes.data) == 0
    np.testing.assert_array_equal(unpack_d(d), A @ B)
