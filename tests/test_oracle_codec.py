"""Pins of the oracle's codec against known answers and closed forms.

Reference: ResidualScalarQuantizationPack.hpp:17-237 (quantiles, bucket rule,
MSB-first packing), DataStructPP.hpp:147-205 (decode table), RerankPP.hpp:65-113
(x = C[code] + w[bucket]).
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_unpack_known_answer_msb_first():
    # byte 0b00011011 holds buckets (0, 1, 2, 3) for dims 4j..4j+3
    dim = 128
    res = np.zeros((1, 32), np.uint8)
    res[0, 5] = 0b00011011
    C = np.zeros((1, dim), np.float32)
    W = O.decode_table(np.array([10.0, 11.0, 12.0, 13.0], np.float32), dim, 2)
    x = O.decompress_tokens(np.array([0], np.uint64), np.array([0], np.uint32), res, C, W, 2)[0]
    np.testing.assert_array_equal(x[20:24], [10.0, 11.0, 12.0, 13.0])
    assert (x[:20] == 10.0).all() and (x[24:] == 10.0).all()


def test_bucket_rule_first_cutoff_strictly_greater():
    cut = np.array([-1.0, 0.0, 1.0], np.float32)
    cases = {-2.0: 0, -1.0: 1, -0.5: 1, 0.0: 2, 0.5: 2, 1.0: 3, 2.0: 3}
    for e, b in cases.items():
        assert O.lib().orc_bucket_of(cut, 3, e) == b, e


def test_pooled_quantiles_known_answer():
    # 8 values 0..7: cutoffs at ranks floor(i/4*8) = 2,4,6; weights at floor((i+.5)/4*8) = 1,3,5,7
    v = np.array([5, 2, 7, 0, 3, 6, 1, 4], np.float32)
    cut, w = O.quantize_pooled(v, 2)
    np.testing.assert_array_equal(cut, [2, 4, 6])
    np.testing.assert_array_equal(w, [1, 3, 5, 7])


def test_pooled_quantiles_fraction_below_cutoff():
    rng = np.random.default_rng(0)
    v = rng.standard_normal(10007).astype(np.float32)
    cut, w = O.quantize_pooled(v, 2)
    for i, c in enumerate(cut, start=1):
        assert (v < c).sum() == int(np.floor(i / 4 * v.size))  # distinct values: rank = count below
    assert (np.diff(cut) > 0).all() and (np.diff(w) > 0).all()
    assert w[0] < cut[0] < w[1] < cut[1] < w[2] < cut[2] < w[3]


def test_perdim_quantiles_and_bucket_means():
    rng = np.random.default_rng(1)
    n, dim = 4001, 128
    e = (rng.standard_normal((n, dim)) * np.linspace(0.01, 0.2, dim)).astype(np.float32)
    cut, w = O.quantize_perdim(e, 2)
    for d in (0, 63, 127):
        col = e[:, d]
        for i in range(3):
            assert (col < cut[d, i]).sum() == int(np.floor((i + 1) / 4 * n))
        b = np.searchsorted(cut[d], col, side="right")   # count of cutoffs <= value
        for k in range(4):
            np.testing.assert_allclose(w[d, k], col[b == k].astype(np.float64).mean(), rtol=1e-6)


def test_compress_known_answer():
    dim = 128
    C = np.zeros((2, dim), np.float32)
    C[1] = 0.5
    cut = np.array([-0.1, 0.0, 0.1], np.float32)
    # buckets cycle 3,2,1,0 over dims -> byte 0b11100100 = 0xE4
    e = np.tile(np.array([0.2, 0.05, -0.05, -0.2], np.float32), dim // 4)
    x = np.stack([C[0] + e, C[1] + e])
    codes = np.array([0, 1], np.uint32)
    out = O.compress(x, codes, C, cut, 2)
    assert out.shape == (2, 32)
    assert (out == 0xE4).all()
    # per-dimension cutoffs: dim 0 moved so that 0.2 falls in bucket 2
    cutd = np.tile(cut, (dim, 1))
    cutd[0] = [-0.1, 0.0, 0.3]
    out2 = O.compress(x, codes, C, cutd, 2)
    assert out2[0, 0] == 0b10100100 and (out2[:, 1:] == 0xE4).all()


def test_encode_decode_roundtrip_error_bounded():
    """Decoding an encoded vector returns C + weight of its bucket; with
    bucket-mean weights the error per dim is bounded by the bucket width."""
    rng = np.random.default_rng(2)
    dim, n = 128, 500
    C = rng.standard_normal((4, dim)).astype(np.float32)
    codes = rng.integers(0, 4, n).astype(np.uint32)
    e = (rng.standard_normal((n, dim)) * 0.05).astype(np.float32)
    x = C[codes] + e
    cut, w = O.quantize_perdim(e, 2)
    res = O.compress(x, codes, C, cut, 2)
    W = O.decode_table(w, dim, 2)
    xd = O.decompress_tokens(np.arange(n, dtype=np.uint64), codes, res, C, W, 2)
    b = np.stack([np.searchsorted(cut[d], (x[:, d] - C[codes, d]), side="right") for d in range(dim)], 1)
    np.testing.assert_array_equal(xd, (C[codes] + W[np.arange(dim)[None, :], b]).astype(np.float32))
    inner = (b == 1) | (b == 2)
    assert np.abs(xd - x)[inner].max() <= (cut[:, 2] - cut[:, 0]).max()


def test_e4m3_round_known_values():
    """OCP E4M3: exact values survive; halfway cases round to even; tiny
    values go to subnormal steps of 2^-9; large values saturate at 448."""
    from oracle import oracle as O
    exact = np.array([1.0, 1.125, -3.0, 448.0, 2.0 ** -9, 0.015625, 240.0], np.float32)
    np.testing.assert_array_equal(O.e4m3_round(exact), exact)
    # 1 + 1/16 is halfway between 1 and 1.125 -> even mantissa (1.0);
    # 1 + 3/16 is halfway between 1.125 and 1.25 -> 1.25
    np.testing.assert_array_equal(O.e4m3_round(np.array([1.0625, 1.1875], np.float32)), [1.0, 1.25])
    np.testing.assert_array_equal(O.e4m3_round(np.array([3 * 2.0 ** -10, 1000.0, -500.0], np.float32)),
                                  np.array([2 * 2.0 ** -9, 448.0, -448.0], np.float32))
    # every representable positive value is a fixed point, and the grid has 126 finite positives
    vals = []
    for e in range(-6, 9):
        for m in range(8):
            v = (1 + m / 8) * 2.0 ** e
            if v <= 448:
                vals.append(v)
    vals += [m * 2.0 ** -9 for m in range(1, 8)]
    v = np.array(sorted(vals), np.float32)
    assert v.size == 126
    np.testing.assert_array_equal(O.e4m3_round(v), v)
    mids = ((v[:-1].astype(np.float64) + v[1:]) / 2 * 0.999).astype(np.float32)
    np.testing.assert_array_equal(O.e4m3_round(mids), v[:-1])


# ---- other widths: 1, 4 and 8 bits per value (8/b values per byte) ----

def _unpack_np(res, n_bit, dim):
    """MSB-first unpack by a different route than the oracle's shifts: the
    row's bits in order (np.unpackbits, most significant first), n_bit at a time."""
    bits = np.unpackbits(res, axis=1)[:, : dim * n_bit].reshape(res.shape[0], dim, n_bit).astype(np.int64)
    return (bits << np.arange(n_bit - 1, -1, -1)).sum(-1)


def _pack_np(b, n_bit):
    """Buckets [n, dim] -> bytes by np.packbits over each value's n_bit bits."""
    bits = (b[..., None] >> np.arange(n_bit - 1, -1, -1)) & 1
    return np.packbits(bits.reshape(b.shape[0], -1).astype(np.uint8), axis=1)


@pytest.mark.parametrize("n_bit,byte,vals", [
    (1, 0b10110001, [1, 0, 1, 1, 0, 0, 0, 1]),
    (2, 0b00011011, [0, 1, 2, 3]),
    (4, 0xA5, [0xA, 0x5]),
    (8, 0xC3, [0xC3]),
])
def test_unpack_known_answers_every_width(n_bit, byte, vals):
    """Value i of a byte is (byte >> (8 - n_bit*(i+1))) & mask
    (DataStructPP.hpp:184-200); weights 10*b make the decoded value show b."""
    dim = 128
    res = np.full((1, dim * n_bit // 8), byte, np.uint8)
    W = O.decode_table(np.arange(1 << n_bit, dtype=np.float32) * 10, dim, n_bit)
    x = O.decompress_tokens(np.array([0], np.uint64), np.array([0], np.uint32), res,
                            np.zeros((1, dim), np.float32), W, n_bit)[0]
    np.testing.assert_array_equal(x, np.tile(np.array(vals, np.float32) * 10, dim // len(vals)))
    np.testing.assert_array_equal(_unpack_np(res, n_bit, dim)[0], np.tile(vals, dim // len(vals)))


@pytest.mark.parametrize("n_bit", [1, 2, 4])
def test_pooled_quantiles_known_answer_every_width(n_bit):
    # 2*nq distinct values 0..2nq-1: cutoffs at ranks floor(i/nq*2nq) = 2i,
    # weights at floor((i+.5)/nq*2nq) = 2i+1
    nq = 1 << n_bit
    v = np.random.default_rng(n_bit).permutation(2 * nq).astype(np.float32)
    cut, w = O.quantize_pooled(v, n_bit)
    np.testing.assert_array_equal(cut, 2 * np.arange(1, nq))
    np.testing.assert_array_equal(w, 2 * np.arange(nq) + 1)


@pytest.mark.parametrize("n_bit", [1, 4, 8])
def test_perdim_quantiles_every_width(n_bit):
    rng = np.random.default_rng(30 + n_bit)
    n, dim, nq = 4001, 128, 1 << n_bit
    e = (rng.standard_normal((n, dim)) * np.linspace(0.01, 0.2, dim)).astype(np.float32)
    cut, w = O.quantize_perdim(e, n_bit)
    assert cut.shape == (dim, nq - 1) and w.shape == (dim, nq)
    for d in (0, 77, 127):
        col = e[:, d]
        for i in range(nq - 1):
            assert (col < cut[d, i]).sum() == int(np.floor((i + 1) / nq * n))
        b = np.searchsorted(cut[d], col, side="right")
        for k in range(nq):
            np.testing.assert_allclose(w[d, k], col[b == k].astype(np.float64).mean(), rtol=1e-5)


@pytest.mark.parametrize("n_bit", [1, 2, 4, 8])
@pytest.mark.parametrize("per_dim", [True, False])
def test_codec_roundtrip_every_width(n_bit, per_dim):
    """compress == np.packbits of the searchsorted buckets; decompress ==
    C[code] + W[d][bucket]; the oracle's MaxSim over those tokens == float64
    MaxSim of the decoded vectors."""
    rng = np.random.default_rng(40 + n_bit)
    dim, n = 128, 600
    C = rng.standard_normal((5, dim)).astype(np.float32)
    codes = rng.integers(0, 5, n).astype(np.uint32)
    x = (C[codes] + rng.standard_normal((n, dim)) * 0.05).astype(np.float32)
    e = x - C[codes]
    cut, w = O.quantize_perdim(e, n_bit) if per_dim else O.quantize_pooled(e, n_bit)
    res = O.compress(x, codes, C, cut, n_bit)
    cutd = cut if per_dim else np.tile(cut, (dim, 1))
    b = np.stack([np.searchsorted(cutd[d], e[:, d], side="right") for d in range(dim)], 1)
    assert res.shape == (n, dim * n_bit // 8)
    np.testing.assert_array_equal(res, _pack_np(b, n_bit))
    np.testing.assert_array_equal(_unpack_np(res, n_bit, dim), b)
    W = O.decode_table(w, dim, n_bit)
    wd = w if per_dim else np.tile(w, (dim, 1))
    np.testing.assert_array_equal(W, wd)
    xd = O.decompress_tokens(np.arange(n, dtype=np.uint64), codes, res, C, W, n_bit)
    np.testing.assert_array_equal(xd, (C[codes] + wd[np.arange(dim)[None, :], b]).astype(np.float32))
    # MaxSim: 6 docs of 100 tokens
    doclens = np.full(6, 100, np.uint32)
    ix = O.Index(doclens, codes, res, C, w, n_bit)
    Q = rng.standard_normal((7, dim)).astype(np.float32)
    got = ix.rerank(Q, np.arange(6, dtype=np.uint32))
    ref = [(Q.astype(np.float64) @ xd[100 * k:100 * (k + 1)].astype(np.float64).T).max(1).sum() for k in range(6)]
    np.testing.assert_allclose(got, ref, rtol=1e-5)
