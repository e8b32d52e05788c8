"""Checks against the reference's own data fixture (lotte-500-gnd PLAID index,
tests/golden/lotte500.npz written by tests/golden/make_lotte_fixture.py).

Pins: the 2-bit layout shared by PLAID and IGP (residual.py:51-91,208-224 vs
DataStructPP.hpp:184-200), the IVF build against the index's own ivf.pid.pt
(IVFIndex.hpp:92-166), decompression of real codes, and an end-to-end recall
sanity check against the fixture's brute-force groundtruth.
"""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lotte500.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


def plaid_reversed_bit_map(nbits=2):
    """The byte map of residual.py:51-72: reverse the bits inside every
    nbits-wide group of a byte."""
    out = []
    mask = (1 << nbits) - 1
    for i in range(256):
        z = 0
        for j in range(8, 0, -nbits):
            x = (i >> (j - nbits)) & mask
            y = 0
            for k in range(nbits - 1, -1, -1):
                y += ((x >> (nbits - k - 1)) & 1) * (2 ** k)
            z |= y
            if j > nbits:
                z <<= nbits
        out.append(z)
    return np.array(out, np.uint8)


def plaid_pack(buckets):
    """binarize + packbits of residual.py:208-224 for one byte (4 dims)."""
    bits = []
    for b in buckets:
        bits += [(b >> 0) & 1, (b >> 1) & 1]
    return int(np.packbits(np.array(bits, np.uint8))[0])


def test_plaid_and_igp_bit_layouts_agree():
    rbm = plaid_reversed_bit_map()
    lut = np.array(list(itertools.product(range(4), repeat=4)), np.uint8)  # residual.py:76-91
    for buckets in itertools.product(range(4), repeat=4):
        p = plaid_pack(buckets)
        assert tuple(lut[rbm[p]]) == buckets                      # PLAID decode
        igp = [(int(rbm[p]) >> (6 - 2 * k)) & 3 for k in range(4)]  # IGP MSB-first decode
        assert tuple(igp) == buckets
    # the IGP packing of (0,1,2,3) is 0b00011011
    assert int(rbm[plaid_pack((0, 1, 2, 3))]) == 0b00011011


def _igp_index(fx):
    rbm = plaid_reversed_bit_map()
    res_igp = rbm[fx["residuals"]]
    C = fx["centroids"].astype(np.float32)
    w = fx["bucket_weights"].astype(np.float32)
    return O.Index(fx["doclens"].astype(np.uint32), fx["codes"].astype(np.uint32), res_igp, C, w)


def test_decompress_fixture_equals_plaid_decode(fx):
    ix = _igp_index(fx)
    rbm = plaid_reversed_bit_map()
    lut = np.array(list(itertools.product(range(4), repeat=4)), np.uint8)
    tok = np.arange(0, fx["codes"].size, 97, dtype=np.uint64)
    got = O.decompress_tokens(tok, ix.codes, ix.residual, ix.C, ix.W, 2)
    res = fx["residuals"][tok.astype(np.int64)]
    idx = lut[rbm[res]].reshape(tok.size, 128)
    ref = ix.C[fx["codes"][tok.astype(np.int64)]] + fx["bucket_weights"].astype(np.float32)[idx]
    np.testing.assert_array_equal(got, ref.astype(np.float32))
    n = np.linalg.norm(got, axis=1)
    assert 0.8 < n.mean() < 1.2   # decompressed ColBERT vectors are near unit length


def test_ivf_matches_fixture_ivf_pid(fx):
    off, size, ids = O.ivf_build(fx["codes"].astype(np.uint32), fx["doclens"].astype(np.uint32), 2048)
    np.testing.assert_array_equal(size, fx["ivf_lengths"])
    np.testing.assert_array_equal(ids, fx["ivf"].astype(np.uint32))


def test_end_to_end_recall_sanity_against_groundtruth(fx):
    """Weak check: the fixture's top-10 came from the original (missing)
    embeddings; exhaustive MaxSim over L2-normalised decompressed embeddings
    and the IGP pipeline with wide probing should recover most of it."""
    ix = _igp_index(fx)
    X = np.concatenate([ix.decompress_doc(d) for d in range(ix.n_doc)])
    Xn = (X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32)
    Q = fx["queries"]
    gi, _ = O.exhaustive_f64(Q, Xn, ix.doclens, 10)
    r_exh = O.recall_at_k(gi, fx["gnd_ids"])
    ids, _ = ix.search(Q, 32, 256, 10)
    r_igp = O.recall_at_k(ids, fx["gnd_ids"])
    # measured here: exhaustive 0.77, IGP (P=32, R=256) 0.77 over the 20 queries
    assert r_exh >= 0.7, r_exh
    assert r_igp >= 0.7, (r_igp, r_exh)
