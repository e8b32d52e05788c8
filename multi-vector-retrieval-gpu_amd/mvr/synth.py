"""Synthetic corpora and IGP indexes built on the device (for benchmarks at
sizes the CPU cannot build).

Recipe: mixture centres are L2-normalised N(0, I) draws; a token picks a
centre uniformly, adds N(0, noise^2) per dimension and is L2-normalised
(counter-based generator: splitmix64 mixing + Box-Muller, the same streams as
the CPU generator).  The index follows the reference's build pipeline
(script/evaluation/technique/vq_sq_pack.py:91-109): k-means on a sample
(Lloyd, L2 assignment, centroids L2-normalised at the end), codes by argmax
inner product over all tokens, 2-bit residual codes with per-dimension
quantile cutoffs and bucket-mean weights, and the IVF of each document's
unique centroids with documents ascending (src/include/struct/IVFIndex.hpp:92-166).

Heavy math runs in libmvr kernels (token generation, assignment, packing);
torch only holds device memory and does the sort/scan plumbing of the build.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import Index
from ._capi import check, lib

STREAM_CENTER, STREAM_DOCLEN, STREAM_TOKCOMP, STREAM_TOKNOISE = 1, 2, 3, 4
STREAM_QCOMP, STREAM_QNOISE, STREAM_QLEN, STREAM_QPROBE = 5, 6, 7, 8
_M64 = (1 << 64) - 1


def _mix64_np(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15))
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def uniform_np(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        key = _mix64_np(np.array([(seed ^ ((stream * 0xD1B54A32D192ED03) & _M64)) & _M64], np.uint64))[0]
        h = _mix64_np(key + idx.astype(np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal_np(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    idx = idx.astype(np.uint64)
    pair = idx >> np.uint64(1)
    u1 = uniform_np(seed, stream, 2 * pair)
    u2 = uniform_np(seed, stream, 2 * pair + np.uint64(1))
    r = np.sqrt(-2.0 * np.log(1.0 - u1))
    a = 6.283185307179586 * u2
    return np.where(idx & np.uint64(1), r * np.sin(a), r * np.cos(a))


def doclens_lognormal(seed, n_doc, median=64.0, sigma=0.29, lo=8, hi=300) -> np.ndarray:
    z = normal_np(seed, STREAM_DOCLEN, np.arange(n_doc, dtype=np.uint64))
    v = np.floor(np.exp(np.log(median) + sigma * z) + 0.5)
    return np.clip(v, lo, hi).astype(np.uint32)


def doclens_uniform(seed, n_doc, lo=16, hi=128) -> np.ndarray:
    u = uniform_np(seed, STREAM_DOCLEN, np.arange(n_doc, dtype=np.uint64))
    return np.minimum(lo + np.floor(u * (hi - lo + 1)), hi).astype(np.uint32)


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def synth_tokens(torch, seed, stream_comp, stream_noise, centers, n_tok, noise=0.15, tok_begin=0,
                 out=None):
    if out is None:
        out = torch.empty((n_tok, 128), dtype=torch.float32, device=centers.device)
    check(lib().mvr_synth_tokens(seed, stream_comp, stream_noise, centers.data_ptr(), centers.shape[0],
                                 128, noise, tok_begin, n_tok, out.data_ptr(), _stream(torch)))
    return out


def gen_centers(torch, seed, n_comp, device="cuda"):
    """normalise(N(0, I)): the token recipe with a zero centre and unit noise."""
    zero = torch.zeros((1, 128), dtype=torch.float32, device=device)
    return synth_tokens(torch, seed, STREAM_TOKCOMP, STREAM_CENTER, zero, n_comp, noise=1.0)


def assign(torch, x, C_, half_norm=None, out=None):
    if out is None:
        out = torch.empty(x.shape[0], dtype=torch.int32, device=x.device)
    check(lib().mvr_assign_bias(x.data_ptr(), x.shape[0], C_.data_ptr(),
                                None if half_norm is None else half_norm.data_ptr(), C_.shape[0], 128,
                                out.data_ptr(), _stream(torch)))
    return out


def kmeans(torch, sample, n_c, n_iter):
    """Lloyd on the sample: init from points floor(j*n/n_c), argmin L2 via
    argmax(<x,c> - |c|^2/2), empty clusters keep their centroid; L2-normalised."""
    n = sample.shape[0]
    init = (torch.arange(n_c, device=sample.device, dtype=torch.int64) * n) // n_c
    cent = sample[init].clone()
    for _ in range(n_iter):
        hn = 0.5 * (cent.double() ** 2).sum(1).float()
        a = assign(torch, sample, cent, hn).long()
        sums = torch.zeros((n_c, 128), dtype=torch.float64, device=sample.device)
        sums.index_add_(0, a, sample.double())
        cnt = torch.bincount(a, minlength=n_c)
        nz = cnt > 0
        cent[nz] = (sums[nz] / cnt[nz, None]).float()
    return torch.nn.functional.normalize(cent, dim=1)


def perdim_codec(torch, residual_sample):
    """Per-dimension cutoffs at ranks floor(i/4 n), bucket-mean weights."""
    n = residual_sample.shape[0]
    srt, _ = torch.sort(residual_sample, dim=0)
    ranks = [int(np.floor(i / 4 * n)) for i in (1, 2, 3)]
    cut = torch.stack([srt[r] for r in ranks], dim=1).contiguous()          # [128, 3]
    b = (residual_sample[:, :, None] >= cut[None, :, :]).sum(2)              # count of cutoffs <= e
    w = torch.empty((128, 4), dtype=torch.float32, device=residual_sample.device)
    for k in range(4):
        m = (b == k)
        s = torch.where(m, residual_sample.double(), torch.zeros((), dtype=torch.float64,
                                                                   device=m.device)).sum(0)
        c = m.sum(0)
        w[:, k] = torch.where(c > 0, s / c.clamp(min=1), torch.zeros_like(s)).float()
    return cut, w


def compress(torch, x, codes, C_, cut, out=None):
    if out is None:
        out = torch.empty((x.shape[0], 32), dtype=torch.uint8, device=x.device)
    check(lib().mvr_compress(x.data_ptr(), codes.data_ptr(), x.shape[0], C_.data_ptr(), 128,
                             cut.data_ptr(), 1, out.data_ptr(), _stream(torch)))
    return out


def ivf_build(torch, codes, doclens_t, n_c):
    n_doc = doclens_t.numel()
    doc_of_tok = torch.repeat_interleave(torch.arange(n_doc, device=codes.device, dtype=torch.int64),
                                         doclens_t.long())
    key = torch.unique(codes.long() * n_doc + doc_of_tok)       # sorted: centroid, then doc
    cent = key // n_doc
    ids = (key % n_doc).to(torch.int32)
    sizes = torch.bincount(cent, minlength=n_c).to(torch.int32)
    off = torch.zeros(n_c, dtype=torch.int64, device=codes.device)
    off[1:] = torch.cumsum(sizes[:-1].long(), 0)
    return off, sizes, ids


@dataclass
class DeviceCorpus:
    seed: int
    doclens: np.ndarray
    centers: object       # torch [n_comp, 128]
    X: object             # torch [n_tok, 128] original embeddings (or None)
    index: Index
    C: object
    codes: object
    residual: object
    W: object
    ivf: tuple


def index_from_tokens(torch, X, doclens, n_c, n_iter=10, kmeans_sample=None, chunk=1 << 22, log=print):
    """IGP index of token embeddings X [n_tok, 128] (device): k-means on a
    sample, codes, per-dimension 2-bit codec, residuals and IVF."""
    dev = X.device
    n_tok = X.shape[0]
    if kmeans_sample and kmeans_sample < n_tok:
        sel = (torch.arange(kmeans_sample, device=dev, dtype=torch.int64) * n_tok) // kmeans_sample
        sample = X[sel].contiguous()
    else:
        sample = X
    C_ = kmeans(torch, sample, n_c, n_iter).contiguous()
    torch.cuda.synchronize()
    log(f"kmeans: {n_c} centroids on {sample.shape[0]} tokens, {n_iter} iterations")
    codes = torch.empty(n_tok, dtype=torch.int32, device=dev)
    for i, t0 in enumerate(range(0, n_tok, chunk)):
        n = min(chunk, n_tok - t0)
        assign(torch, X[t0:t0 + n], C_, None, out=codes[t0:t0 + n])
        if i % 4 == 3:
            torch.cuda.synchronize()
            log(f"assign: {t0 + n}/{n_tok} tokens")
    torch.cuda.synchronize()
    log("assign: done")
    cs = sample if sample is not X else X[:: max(1, n_tok // (1 << 20))]
    cs_codes = assign(torch, cs.contiguous(), C_)
    cut, W = perdim_codec(torch, cs - C_[cs_codes.long()])
    residual = torch.empty((n_tok, 32), dtype=torch.uint8, device=dev)
    for t0 in range(0, n_tok, chunk):
        n = min(chunk, n_tok - t0)
        compress(torch, X[t0:t0 + n], codes[t0:t0 + n], C_, cut, out=residual[t0:t0 + n])
    dl = torch.from_numpy(doclens.astype(np.int32)).to(dev)
    ivf = ivf_build(torch, codes, dl, n_c)
    torch.cuda.synchronize()
    log(f"codec + ivf: {ivf[2].numel()} IVF entries")
    index = Index.from_device(dl, C_, codes, residual, W, ivf)
    return index, C_, codes, residual, W, ivf


def build_corpus(torch, seed, n_doc, n_comp, n_c, doclens, n_iter=10, kmeans_sample=None,
                 keep_embeddings=True, chunk=1 << 22, log=print):
    dev = torch.device("cuda")
    n_tok = int(doclens.sum())
    centers = gen_centers(torch, seed, n_comp)
    X = torch.empty((n_tok, 128), dtype=torch.float32, device=dev)
    for t0 in range(0, n_tok, chunk):
        n = min(chunk, n_tok - t0)
        synth_tokens(torch, seed, STREAM_TOKCOMP, STREAM_TOKNOISE, centers, n, tok_begin=t0, out=X[t0:t0 + n])
    torch.cuda.synchronize()
    log(f"synth: {n_tok} tokens")
    index, C_, codes, residual, W, ivf = index_from_tokens(torch, X, doclens, n_c, n_iter, kmeans_sample,
                                                           chunk, log)
    return DeviceCorpus(seed, doclens, centers, X if keep_embeddings else None, index, C_, codes,
                        residual, W, ivf)


def query_shapes(seed, n_q, t_lo=8, t_hi=32, probes=(2, 4, 8, 16, 32)):
    """Per-query token counts U{t_lo..t_hi} and nprobe uniform over `probes`
    (the uneven batching workload)."""
    idx = np.arange(n_q, dtype=np.uint64)
    u = uniform_np(seed, STREAM_QLEN, idx)
    nt = np.minimum(t_lo + np.floor(u * (t_hi - t_lo + 1)), t_hi).astype(np.uint32)
    v = uniform_np(seed, STREAM_QPROBE, idx)
    pr = np.asarray(probes, np.uint32)
    npr = pr[np.minimum(np.floor(v * len(pr)).astype(np.int64), len(pr) - 1)]
    return nt, npr


def topic_corpus(torch, seed, n_doc, n_comp, n_c, doclens, n_topics=128, comps_per_topic=16, noise=0.15,
                 p_topic=1.0, n_iter=10, log=print):
    """A corpus with topical structure (for the learned-probe evaluation, where
    query difficulty has to vary): topic t owns `comps_per_topic` mixture
    components drawn at random (topics overlap); a doc picks a topic and each
    of its tokens one of the topic's components (with probability p_topic,
    else any component), plus N(0, noise^2) per dim, L2-normalised.  Seeded
    torch generator on the device."""
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    centers = gen_centers(torch, seed, n_comp)
    topics = torch.randint(0, n_comp, (n_topics, comps_per_topic), generator=gen, device=dev)
    dl = torch.from_numpy(doclens.astype(np.int64)).to(dev)
    doc_topic = torch.randint(0, n_topics, (n_doc,), generator=gen, device=dev)
    tok_topic = torch.repeat_interleave(doc_topic, dl)
    n_tok = tok_topic.numel()
    pick = torch.randint(0, comps_per_topic, (n_tok,), generator=gen, device=dev)
    comp = topics[tok_topic, pick]
    if p_topic < 1.0:
        off = torch.rand((n_tok,), generator=gen, device=dev) >= p_topic
        comp = torch.where(off, torch.randint(0, n_comp, (n_tok,), generator=gen, device=dev), comp)
    X = centers[comp] + noise * torch.randn((n_tok, 128), generator=gen, device=dev)
    X = torch.nn.functional.normalize(X, dim=1).contiguous()
    log(f"topic corpus: {n_tok} tokens, {n_topics} topics x {comps_per_topic} components")
    index, C_, codes, residual, W, ivf = index_from_tokens(torch, X, doclens, n_c, n_iter, None, 1 << 22, log)
    corpus = DeviceCorpus(seed, doclens, centers, X, index, C_, codes, residual, W, ivf)
    corpus.topics = topics
    corpus.doc_topic = doc_topic
    return corpus


def topic_queries(torch, corpus, seed, n_q, T=32, noise_lo=0.05, noise_hi=0.4, p_lo=1.0, p_hi=1.0):
    """Queries of the topic corpus: a topic, T tokens from its components and
    a per-query noise level uniform in [noise_lo, noise_hi]; each token comes
    from the topic with a per-query probability uniform in [p_lo, p_hi], else
    from any component.  Noisier, less topical queries have flatter
    centroid-score profiles and need more probes."""
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    n_topics, cpt = corpus.topics.shape
    qt = torch.randint(0, n_topics, (n_q,), generator=gen, device=dev)
    pick = torch.randint(0, cpt, (n_q, T), generator=gen, device=dev)
    comp = corpus.topics[qt[:, None], pick]
    pq = p_lo + (p_hi - p_lo) * torch.rand((n_q, 1), generator=gen, device=dev)
    off = torch.rand((n_q, T), generator=gen, device=dev) >= pq
    comp = torch.where(off, torch.randint(0, corpus.centers.shape[0], (n_q, T), generator=gen, device=dev), comp)
    sig = noise_lo + (noise_hi - noise_lo) * torch.rand((n_q, 1, 1), generator=gen, device=dev)
    Q = corpus.centers[comp] + sig * torch.randn((n_q, T, 128), generator=gen, device=dev)
    return torch.nn.functional.normalize(Q, dim=2).contiguous()


def make_queries(torch, seed, centers, n_q, T=32, noise=0.15):
    Q = synth_tokens(torch, seed, STREAM_QCOMP, STREAM_QNOISE, centers, n_q * T, noise)
    return Q.view(n_q, T, 128)


@dataclass
class PlaidArrays:
    """Host arrays of a PLAID index (the layout IndexScorer loads)."""
    doclens: np.ndarray         # int64 [n_doc]
    codes: np.ndarray           # int32 [n_tok]
    residuals: np.ndarray       # uint8 [n_tok, 32], PLAID bit order
    centroids: np.ndarray       # float32 [n_c, 128] (float16 values, as centroids.pt stores them)
    bucket_cutoffs: np.ndarray  # float32 [3]
    bucket_weights: np.ndarray  # float32 [4] (float16 values, as buckets.pt stores them)
    ivf: np.ndarray             # int32 [n_ivf]
    ivf_lengths: np.ndarray     # int64 [n_c]

    def index_kwargs(self):
        return dict(doclens=self.doclens, codes=self.codes, residuals=self.residuals, centroids=self.centroids,
                    bucket_weights=self.bucket_weights, ivf=self.ivf, ivf_lengths=self.ivf_lengths)


def plaid_from_corpus(torch, corpus: DeviceCorpus, sample=1 << 17, chunk=1 << 22, log=print) -> PlaidArrays:
    """A PLAID index of a device corpus's embeddings with its centroids and
    codes: centroids rounded to float16 (centroids.pt), residuals against
    them binarized on the GPU with pooled cutoffs / weights at the quantiles
    i/4 and (i + 1/2)/4 of a residual sample (collection_indexer.py:320-345),
    the pid IVF of ivf.pid.pt (indexing/utils.py:8-53)."""
    from . import plaid
    X, n_tok = corpus.X, corpus.X.shape[0]
    C16 = corpus.C.half().float().contiguous()
    codes = corpus.codes
    sel = (torch.arange(min(sample, n_tok), device=X.device, dtype=torch.int64) * n_tok) // min(sample, n_tok)
    E = (X[sel] - C16[codes[sel].long()]).reshape(-1)
    E = E[torch.randperm(E.numel(), device=E.device, generator=torch.Generator(device=E.device).manual_seed(0))
          [: 1 << 23]]
    qs = torch.arange(0, 4, device=E.device, dtype=torch.float32) / 4
    cut = E.quantile(qs[1:]).float()
    w = E.quantile(qs + 0.125).half().float()
    residuals = torch.empty((n_tok, 32), dtype=torch.uint8, device=X.device)
    for t0 in range(0, n_tok, chunk):
        n = min(chunk, n_tok - t0)
        e = X[t0:t0 + n] - C16[codes[t0:t0 + n].long()]
        residuals[t0:t0 + n] = plaid.binarize(e, cut, 2)
    torch.cuda.synchronize()
    off, sizes, ids = corpus.ivf
    log(f"plaid: {n_tok} tokens binarized, cutoffs {cut.tolist()}, weights {w.tolist()}")
    return PlaidArrays(doclens=corpus.doclens.astype(np.int64), codes=codes.cpu().numpy().astype(np.int32),
                       residuals=residuals.cpu().numpy(), centroids=C16.cpu().numpy(),
                       bucket_cutoffs=cut.cpu().numpy(), bucket_weights=w.cpu().numpy(),
                       ivf=ids.cpu().numpy().astype(np.int32), ivf_lengths=sizes.cpu().numpy().astype(np.int64))
