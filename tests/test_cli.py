"""IGPGPUPP / IGPGPUBasic executables: flag grammar and error exits (CPU), and
an end-to-end run over a small dataset directory against the oracle (GPU).

Reference: src/app/IGPGPUPP.cpp:15-217, deps/cmd_line_parser/include/parser.hpp:20-60.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

EXE = os.path.join(PKG, "lib", "IGPGPUPP")
EXE_BASIC = os.path.join(PKG, "lib", "IGPGPUBasic")
FLAGS = ["-username", "u", "-dataset", "ds", "-method-name", "igp", "-num-centroid", "128",
         "-num-bit", "2", "-topk", "10", "-nprobe", "4", "-probe_topk", "64", "-n_thread", "1",
         "-is_profile", "false"]


def _run(exe, args, env=None):
    return subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=120)


def test_unknown_flag_exits_1():
    r = _run(EXE, FLAGS + ["-bogus", "1"])
    assert r.returncode == 1
    assert "shorthand '-bogus' not found" in r.stderr


def test_missing_flag_exits_1():
    r = _run(EXE, FLAGS[:-2])
    assert r.returncode == 1


def test_duplicate_flag_exits_1():
    r = _run(EXE, FLAGS + ["-topk", "5"])
    assert r.returncode == 1 and "already parsed" in r.stderr


def test_basic_has_no_n_thread_flag():
    i = FLAGS.index("-n_thread")
    r = _run(EXE_BASIC, FLAGS)
    assert r.returncode == 1
    r = _run(EXE_BASIC, FLAGS[:i] + FLAGS[i + 2:] + ["-bogus", "x"])
    assert r.returncode == 1


def test_missing_index_is_an_error(tmp_path):
    env = dict(os.environ, MVR_DATASET_ROOT=str(tmp_path))
    r = _run(EXE, FLAGS, env)
    assert r.returncode == 1 and "cannot open" in r.stderr


def write_dataset(root, ix, Q, qids, n_c, n_bit=2):
    """Index/Embedding/RawData layout of src/app/IGPGPUPP.cpp:33-73."""
    import mvr
    d = os.path.join(root, "Index", "ds", f"igp-n_centroid_{n_c}-n_bit_{n_bit}")
    os.makedirs(d)
    np.save(os.path.join(d, "item_n_vec_l.npy"), ix.doclens.astype(np.uint32))
    np.save(os.path.join(d, "centroid_l.npy"), ix.C.astype(np.float32))
    np.save(os.path.join(d, "vq_code_l.npy"), ix.codes.astype(np.int32))
    np.save(os.path.join(d, "weight_l.npy"), ix.weights.astype(np.float32))
    np.save(os.path.join(d, "residual_code_l.npy"), ix.residual.reshape(-1))
    m = mvr.build_index_module()
    ivf = m.IVFIndex(ix.codes.tolist(), ix.doclens.tolist(), int(ix.codes.size), int(ix.n_doc), n_c)
    ivf.build()
    ivf.save(os.path.join(d, "ivf.index"))
    e = os.path.join(root, "Embedding", "ds")
    os.makedirs(e)
    np.save(os.path.join(e, "query_embedding.npy"), Q.astype(np.float32))
    r = os.path.join(root, "RawData", "ds", "document")
    os.makedirs(r)
    with open(os.path.join(r, "queries.dev.tsv"), "w") as f:
        for q in qids:
            f.write(f"{q}\tquery text {q}\n")
    os.makedirs(os.path.join(root, "Result", "answer"))


@pytest.mark.gpu
@pytest.mark.parametrize("exe,clip,host,nprobe", [(EXE, False, False, 4), (EXE_BASIC, True, False, 4),
                                                  (EXE, False, True, 4), (EXE, False, False, 64),
                                                  (EXE_BASIC, True, False, 64)])
def test_cli_answer_matches_oracle(tmp_path, small, exe, clip, host, nprobe):
    """nprobe 64 is what eval_igp_gpu_pp.py:243,266,283 pass."""
    corpus, ix, Q = small
    qids = [100 + 3 * i for i in range(Q.shape[0])]
    write_dataset(str(tmp_path), ix, Q, qids, ix.n_c)
    env = dict(os.environ, MVR_DATASET_ROOT=str(tmp_path))
    flags = list(FLAGS)
    flags[flags.index("-nprobe") + 1] = str(nprobe)
    args = flags if exe == EXE else [a for i, a in enumerate(flags)
                                     if a != "-n_thread" and (i == 0 or flags[i - 1] != "-n_thread")]
    r = _run(exe, args + ["-batch", "5"] + (["-residual_on_host", "true"] if host else []), env)
    assert r.returncode == 0, r.stderr
    ans = os.path.join(tmp_path, "Result", "answer")
    suffix = f"nprobe_{nprobe}-probe_topk_64" + ("-n_thread_1" if exe == EXE else "")
    rows = [l.split("\t") for l in open(os.path.join(
        ans, f"ds-igp-top10-n_centroid_{ix.n_c}-n_bit_2-{suffix}.tsv")).read().splitlines()]
    oi, os_ = ix.search(Q, nprobe, 64, 10, clip_zero=clip)
    assert len(rows) == Q.shape[0] * 10
    for n, (qid, doc, rank, score) in enumerate(rows):
        q, k = divmod(n, 10)
        assert int(qid) == qids[q] and int(rank) == k + 1
        assert int(doc) == oi[q, k]
        assert abs(float(score) - os_[q, k]) <= 1e-5 * abs(os_[q, k])
    perf = open(os.path.join(ans, f"ds-igp-performance-top10-n_centroid_{ix.n_c}-n_bit_2-{suffix}.txt")).read()
    assert "batch_query_time(s)" in perf and f"n_query\t{Q.shape[0]}" in perf


@pytest.mark.gpu
def test_cli_probe_topk_below_topk_exits_255(tmp_path, small):
    corpus, ix, Q = small
    write_dataset(str(tmp_path), ix, Q, list(range(Q.shape[0])), ix.n_c)
    env = dict(os.environ, MVR_DATASET_ROOT=str(tmp_path))
    args = list(FLAGS)
    args[args.index("-probe_topk") + 1] = "5"
    r = _run(EXE, args, env)
    assert r.returncode == 255
