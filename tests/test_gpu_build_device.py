"""The device index-build path that the benchmark's 1M-doc index is made
with (multi-vector-retrieval-gpu_amd/mvr/synth.py), stage by stage against
the CPU oracle: the counter-based token generator (mvr_synth_tokens vs
orc_gen_tokens), the argmax-inner-product code assignment (mvr_assign vs
orc_assign_ip, bit-exact including ties), the Lloyd assignment step
(mvr_assign_bias vs orc_assign_bias), the k-means loop, the per-dimension
codec and the IVF build (torch plumbing over those kernels vs
orc_quantize_perdim / orc_ivf_build)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_synth_tokens_match_oracle_generator():
    import torch
    from mvr import synth
    seed = 77
    cen = synth.gen_centers(torch, seed, 300)
    np.testing.assert_allclose(cen.cpu().numpy(), O.gen_centers(seed, 300), rtol=0, atol=2e-7)
    x = synth.synth_tokens(torch, seed, synth.STREAM_TOKCOMP, synth.STREAM_TOKNOISE, cen, 5000,
                           tok_begin=12345).cpu().numpy()
    ref = O.gen_tokens(seed, cen.cpu().numpy(), 5000, tok_begin=12345)
    # same components and normals; double-precision sums in another order
    # change at most the last float bit
    np.testing.assert_allclose(x, ref, rtol=0, atol=2e-7)
    assert (x == ref).mean() > 0.99
    q = synth.make_queries(torch, seed, cen, 4).cpu().numpy()
    np.testing.assert_allclose(q.reshape(-1, 128), O.gen_tokens(seed, cen.cpu().numpy(), 128, query=True),
                               rtol=0, atol=2e-7)


def test_assign_and_assign_bias_bit_exact():
    import torch
    from mvr import synth
    rng = np.random.default_rng(9)
    C_ = rng.standard_normal((700, 128)).astype(np.float32)
    C_[650] = C_[3]  # an exact tie: lower id
    x = rng.standard_normal((20000, 128)).astype(np.float32)
    x[:50] = C_[3]
    xd, Cd = torch.from_numpy(x).cuda(), torch.from_numpy(C_).cuda()
    got = synth.assign(torch, xd, Cd).cpu().numpy().astype(np.uint32)
    np.testing.assert_array_equal(got, O.assign_ip(x, C_))
    assert (got[:50] == 3).all()
    hn = (0.5 * (C_.astype(np.float64) ** 2).sum(1)).astype(np.float32)
    gotb = synth.assign(torch, xd, Cd, torch.from_numpy(hn).cuda()).cpu().numpy().astype(np.uint32)
    np.testing.assert_array_equal(gotb, O.assign_bias(x, C_, hn))


def test_kmeans_codec_and_ivf_match_oracle():
    import torch
    from mvr import synth
    rng = np.random.default_rng(10)
    centres = rng.standard_normal((40, 128))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    x = (np.repeat(centres, 50, 0) + 0.03 * rng.standard_normal((2000, 128))).astype(np.float32)
    x = x[rng.permutation(2000)]
    xd = torch.from_numpy(x).cuda()
    C_ = synth.kmeans(torch, xd, 40, 5)
    # well-separated clusters: the same assignments every iteration, means
    # summed in double in another order
    np.testing.assert_allclose(C_.cpu().numpy(), O.kmeans(x, 40, 5), rtol=0, atol=1e-6)
    Ch = O.kmeans(x, 40, 5)
    codes = O.assign_ip(x, Ch)
    e = O.residuals(x, codes, Ch)
    cut, w = synth.perdim_codec(torch, torch.from_numpy(e).cuda())
    cut_o, w_o = O.quantize_perdim(e, 2)
    np.testing.assert_array_equal(cut.cpu().numpy(), cut_o)
    np.testing.assert_allclose(w.cpu().numpy(), w_o, rtol=2e-7, atol=0)
    doclens = rng.integers(1, 40, 300).astype(np.uint32)
    n_tok = int(doclens.sum())
    tc = rng.integers(0, 97, n_tok).astype(np.uint32)
    off, size, ids = synth.ivf_build(torch, torch.from_numpy(tc.astype(np.int32)).cuda(),
                                     torch.from_numpy(doclens.astype(np.int32)).cuda(), 97)
    o2, s2, i2 = O.ivf_build(tc, doclens, 97)
    np.testing.assert_array_equal(off.cpu().numpy().astype(np.uint64), o2)
    np.testing.assert_array_equal(size.cpu().numpy().astype(np.uint32), s2)
    np.testing.assert_array_equal(ids.cpu().numpy().astype(np.uint32), i2)


def test_device_compress_matches_oracle():
    import torch
    from mvr import synth
    rng = np.random.default_rng(11)
    C_ = rng.standard_normal((64, 128)).astype(np.float32)
    codes = rng.integers(0, 64, 30000).astype(np.uint32)
    x = (C_[codes] + 0.1 * rng.standard_normal((30000, 128))).astype(np.float32)
    cut, _ = O.quantize_perdim(O.residuals(x, codes, C_), 2)
    got = synth.compress(torch, torch.from_numpy(x).cuda(), torch.from_numpy(codes.astype(np.int32)).cuda(),
                         torch.from_numpy(C_).cuda(), torch.from_numpy(cut).cuda())
    np.testing.assert_array_equal(got.cpu().numpy(), O.compress(x, codes, C_, cut, 2))
