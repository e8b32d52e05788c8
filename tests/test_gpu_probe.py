"""Learned-probe MLP on the GPU against the float64 CPU oracle (oracle/mlp.py):
forward outputs, loss and gradients of every parameter, one Adam step; then
the learned search path end to end on config 1.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    return T


def _data(rng, B):
    x = (rng.standard_normal((B, 1024)) * 0.3).astype(np.float32)
    y = rng.integers(0, 5, B).astype(np.int32)
    r = rng.uniform(2.0, 7.0, B).astype(np.float32)
    return x, y, r


def test_forward_loss_and_gradients_match_oracle(torch):
    from mvr.probe import ProbeModel, split_params
    from oracle import mlp as M
    rng = np.random.default_rng(0)
    model = ProbeModel(seed=1)
    flat = model.get_params()
    flat[np.arange(flat.size) % 97 == 0] += 0.01  # nonzero biases too
    model.set_params(flat)
    params = split_params(flat)
    x, y, r = _data(rng, 64)
    xd, yd, rd = (torch.from_numpy(a).cuda() for a in (x, y, r))
    out = model.forward(torch, xd).cpu().numpy()
    ref, _ = M.forward(params, x)
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-5)
    loss, g = model.grads(xd, yd, rd)
    rl, rg = M.loss_and_grads(params, x, y, r)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    for got, exp in zip(split_params(g), rg):
        np.testing.assert_allclose(got, exp, rtol=2e-3, atol=2e-6 * max(1.0, np.abs(exp).max()))


def test_partial_tiles_match_oracle(torch):
    """Widths and batch that fill no 64x64 output tile and no 16-deep K chunk
    (1000 -> 100 -> 37 -> 6, batch 77): the matrix-core GEMMs' zero padding
    past M, N and K, forward and every gradient."""
    from mvr.probe import ProbeModel, split_params
    from oracle import mlp as M
    dims = (1000, 100, 37, 6)
    rng = np.random.default_rng(9)
    model = ProbeModel(dims=dims, seed=2)
    flat = model.get_params()
    flat[np.arange(flat.size) % 31 == 0] += 0.02
    model.set_params(flat)
    params = split_params(flat, dims)
    x = (rng.standard_normal((77, 1000)) * 0.3).astype(np.float32)
    y = rng.integers(0, 5, 77).astype(np.int32)
    r = rng.uniform(2.0, 7.0, 77).astype(np.float32)
    xd, yd, rd = (torch.from_numpy(a).cuda() for a in (x, y, r))
    ref, _ = M.forward(params, x)
    np.testing.assert_allclose(model.forward(torch, xd).cpu().numpy(), ref, rtol=1e-4, atol=1e-5)
    loss, g = model.grads(xd, yd, rd)
    rl, rg = M.loss_and_grads(params, x, y, r)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    for got, exp in zip(split_params(g, dims), rg):
        np.testing.assert_allclose(got, exp, rtol=2e-3, atol=2e-6 * max(1.0, np.abs(exp).max()))
    model.close()


def test_adam_step_matches_oracle(torch):
    from mvr.probe import ProbeModel, split_params
    from oracle import mlp as M
    rng = np.random.default_rng(2)
    model = ProbeModel(seed=3)
    params = split_params(model.get_params())
    x, y, r = _data(rng, 128)
    xd, yd, rd = (torch.from_numpy(a).cuda() for a in (x, y, r))
    model.train_step(torch, xd, yd, rd, lr=1e-3)
    torch.cuda.synchronize()
    _, grads = M.loss_and_grads(params, x, y, r)
    zeros = [np.zeros_like(p, dtype=np.float64) for p in params]
    exp, _, _ = M.adam_step([p.astype(np.float64) for p in params], grads, zeros, [z.copy() for z in zeros], 1)
    got = split_params(model.get_params())
    for g_, e_ in zip(got, exp):
        # first Adam step moves each weight by ~lr*sign(grad); allow grads near 0 to differ
        np.testing.assert_allclose(g_, e_, atol=2.1e-3)
        assert np.mean(np.abs(g_ - e_) < 1e-5) > 0.98


def test_training_reduces_loss(torch):
    from mvr.probe import ProbeModel, train
    rng = np.random.default_rng(4)
    n = 4096
    x = (rng.standard_normal((n, 1024)) * 0.3).astype(np.float32)
    w = rng.standard_normal((1024, 5)).astype(np.float32)
    y = (x @ w).argmax(axis=1).astype(np.int32)             # learnable labels
    r = (x[:, :8].sum(axis=1) + 4.0).astype(np.float32)
    model = ProbeModel(seed=5)
    losses = train(torch, model, x, y, r, epochs=12, batch=512, lr=1e-3)
    assert losses[-1] < 0.5 * losses[0], losses


def test_learned_search_runs_and_respects_decisions(torch, c1):
    """Attach a model and search with learned probing: results are valid top-k
    lists whose scores equal the FP32 oracle scores of the returned docs."""
    import mvr
    from mvr.probe import ProbeModel
    corpus, ix, Q = c1
    g = mvr.Index.from_arrays(ix.doclens, ix.C, ix.codes, ix.residual, ix.weights, ix.ivf)
    model = ProbeModel(seed=6)
    model.attach(g, conf=0.6, fallback=16, r_min=32)
    ids, sc = g.search(Q[:16], 4, 256, 10, learned=True)
    for q in range(16):
        assert (ids[q] != 0xFFFFFFFF).all()
        np.testing.assert_array_equal(sc[q], ix.rerank(Q[q], ids[q]))


def test_learned_probe_gains_at_equal_rerank_depth(torch):
    """The trained network against the method without it.  Workload with
    query-dependent difficulty (topic-structured corpus, per-query noise and
    topicality); labels from exhaustive MaxSim over the index (smallest nprobe
    reaching recall@10 >= 0.9 at rerank depth 2,048, and the depth that
    needs).  At the learned probe's mean rerank depth, every fixed nprobe has
    lower recall@10 than learned probing: the model spends depth where a
    query needs it."""
    from mvr import probe, synth
    seed = 20261203
    doclens = synth.doclens_uniform(seed, 8192, 16, 128)
    corpus = synth.topic_corpus(torch, seed, 8192, 1024, 1024, doclens, n_topics=128, comps_per_topic=16,
                                p_topic=0.5, log=lambda *a: None)
    kw = dict(noise_lo=0.05, noise_hi=0.6, p_lo=0.3, p_hi=1.0)
    Qtr = synth.topic_queries(torch, corpus, seed + 1, 16384, **kw)
    Qev = synth.topic_queries(torch, corpus, seed + 2, 1024, **kw)
    out = probe.learned_vs_fixed(torch, corpus.index, Qtr, Qev, 10, 2048, 8192, target=0.9, epochs=20)
    counts = out["class_counts"]
    assert sum(c > 0 for c in counts) >= 4, counts          # the labels carry information
    assert out["losses"][-1] < 0.5 * out["losses"][1], out["losses"]
    learned = out["eval_full_depth"][-1]
    best_fixed = max(r["recall"] for r in out["eval_same_depth"])
    assert learned["recall"] > best_fixed + 0.01, (learned, out["eval_same_depth"])
