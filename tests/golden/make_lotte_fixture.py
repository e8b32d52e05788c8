"""Writes tests/golden/lotte500.npz from the reference's own data fixture
multi-vector-retrieval-data/{Index,Embedding,RawData}/lotte-500-gnd (a PLAID
index of 500 LoTTE docs: 2,048 centroids, 2-bit residuals, 56,658 tokens;
100 x 32 query embeddings; brute-force top-10 groundtruth).

Data only: tensors are read with torch.load(weights_only=True) / numpy, and
no reference code is imported or run.  Run once in a container that has the
reference mounted:  python tests/golden/make_lotte_fixture.py
"""
import os
import sys

import numpy as np

REF = "/root/reference/multi-vector-retrieval-data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lotte500.npz")
N_QUERY = 20


def main():
    import torch

    idx = os.path.join(REF, "Index/lotte-500-gnd/plaid")
    emb = os.path.join(REF, "Embedding/lotte-500-gnd")

    def load(name):
        return torch.load(os.path.join(idx, name), map_location="cpu", weights_only=True)

    codes = load("0.codes.pt").numpy().astype(np.int32)
    residuals = load("0.residuals.pt").numpy().astype(np.uint8)
    centroids = load("centroids.pt").numpy().astype(np.float16)
    cutoffs, weights = load("buckets.pt")
    ivf, ivf_lengths = load("ivf.pid.pt")
    doclens = np.load(os.path.join(emb, "doclens.npy")).astype(np.int32)
    queries = np.load(os.path.join(emb, "query_embedding.npy"))[:N_QUERY].astype(np.float32)
    qids = []
    with open(os.path.join(REF, "RawData/lotte-500-gnd/document/queries.dev.tsv")) as f:
        for line in f:
            if line.strip():
                qids.append(int(line.split("\t")[0]))
    qids = np.array(qids[:N_QUERY], np.int64)
    gnd = {}
    with open(os.path.join(emb, "lotte-500-gnd-groundtruth-top10--.tsv")) as f:
        for line in f:
            q, d, r, s = line.rstrip("\n").split("\t")
            gnd.setdefault(int(q), []).append((int(r), int(d), float(s)))
    gnd_ids = np.array([[d for _, d, _ in sorted(gnd[int(q)])] for q in qids], np.int32)
    gnd_scores = np.array([[s for _, _, s in sorted(gnd[int(q)])] for q in qids], np.float32)
    np.savez_compressed(OUT, codes=codes, residuals=residuals, centroids=centroids,
                        bucket_cutoffs=cutoffs.numpy().astype(np.float32),
                        bucket_weights=weights.numpy().astype(np.float16),
                        ivf=ivf.numpy().astype(np.int32), ivf_lengths=ivf_lengths.numpy().astype(np.int64),
                        doclens=doclens, queries=queries, qids=qids, gnd_ids=gnd_ids,
                        gnd_scores=gnd_scores)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    sys.exit(main())
