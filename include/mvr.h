/*
 * mvr.h -- C ABI of the MI355X ColBERT late-interaction search path (libmvr).
 *
 * Plain pointers and sizes only.  Every function returns an mvr_status; the
 * library never calls exit().  mvr_last_error() gives the message of the last
 * failure on the calling thread.  One index handle may be used by one host
 * thread at a time; concurrency happens on the device.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference repository root):
 *   mvr_index_create / mvr_index_load (and the _ex variants)
 *       IGPGPUPP ctor + IGPGPUPP::buildIndex  src/impl/IGPGPUPP.hpp:67-251
 *       (npy loads :124-173, readIVFIndex src/include/struct/IVFIndex.hpp:27-59,
 *        device upload src/include/alg/igp_gpu_pp/DataStructPP.hpp:27-205)
 *   mvr_search
 *       IGPGPUPP::search  src/impl/IGPGPUPP.hpp:253-332, i.e. per query
 *       GPGPURetrievalPP::search  src/include/alg/igp_gpu_pp/IGPGPURetrievalPP.hpp:114-176
 *       (centroid scoring GenerateVQScoreCandidatePP.hpp:118-179, filter
 *        FilterPP.hpp:267-398, rerank RerankPP.hpp:271-475, host top-k
 *        IGPGPURetrievalPP.hpp:153-161).  With clip_zero = 1 it is the basic
 *       variant's search, src/impl/IGPGPUBasic.hpp:190-310.
 *   mvr_decompress
 *       ResidualCodePack::Decode  src/include/alg/refine/ResidualScalarQuantizationPack.hpp:330-378
 *       and decompressAndAddResidual  src/include/alg/igp_gpu_pp/RerankPP.hpp:65-113
 */
#ifndef MVR_H
#define MVR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MVR_OK = 0,
  MVR_ERR_INVALID = 1,   /* bad argument or shape */
  MVR_ERR_IO = 2,        /* missing or malformed file */
  MVR_ERR_DEVICE = 3,    /* HIP runtime error */
  MVR_ERR_UNSUPPORTED = 4
} mvr_status;

/* MVR_FP32: MaxSim as sequential FP32 FMA chains, bit-identical to the CPU
 * restatement of the reference.  MVR_FP32_MFMA: the same FP32 decompressed
 * embeddings and FP32 queries on v_mfma_f32_32x32x2_f32 (sums in the matrix
 * unit's order; within 1e-5 relative of MVR_FP32).  MVR_BF16: centroid scoring
 * (exact on bf16-rounded inputs) and MaxSim on BF16 MFMA.  MVR_FP8: MaxSim on
 * FP8 (OCP e4m3) MFMA with per-centroid, per-query-token and residual scales;
 * centroid top-P exact in FP32.  Low-precision modes optionally rescore their
 * best `rescore` docs in FP32. */
typedef enum { MVR_FP32 = 0, MVR_BF16 = 1, MVR_FP8 = 2, MVR_FP32_MFMA = 3 } mvr_precision;

typedef struct mvr_index mvr_index;

/* Index in the reference layout (src/impl/IGPGPUPP.hpp:30-55).  Pointers are
 * host pointers unless on_device != 0, in which case they are device pointers
 * on `device`; either way the library copies them. */
typedef struct {
  uint32_t n_doc;
  uint64_t n_tok;
  uint32_t n_centroid;
  uint32_t dim;              /* must be 128 */
  uint32_t n_bit;            /* 1, 2, 4 or 8 bits per residual value; other
                              * than 2: device-resident, searched and reranked
                              * with MVR_FP32 only (the MFMA, persistent and
                              * host-staging kernels decode 2-bit codes; the
                              * reference's decode table for 8/b values per
                              * byte is DataStructPP.hpp:184-200) */
  const uint32_t *doclens;   /* [n_doc]                         item_n_vec_l.npy */
  const float *centroids;    /* [n_centroid][dim]               centroid_l.npy */
  const uint32_t *codes;     /* [n_tok]                         vq_code_l.npy */
  const uint8_t *residual;   /* [n_tok][dim*n_bit/8], MSB-first residual_code_l.npy */
  const float *weights;      /* [2^n_bit] or [dim][2^n_bit]     weight_l.npy */
  uint32_t weights_per_dim;  /* 0: one array for all dims, 1: per dimension */
  const uint64_t *ivf_offsets; /* [n_centroid]                  ivf.index */
  const uint32_t *ivf_sizes;   /* [n_centroid] */
  const uint32_t *ivf_ids;     /* [ivf_n_ele], docs ascending per list */
  uint64_t ivf_n_ele;
  int on_device;
} mvr_index_desc;

typedef struct {
  uint32_t nprobe;      /* centroids probed per query token (-nprobe), at most
                           min(n_centroid, 256); larger is MVR_ERR_UNSUPPORTED */
  uint32_t probe_topk;  /* docs kept by the filter and reranked (-probe_topk), <= 16384 */
  uint32_t topk;        /* results per query (-topk) */
  uint32_t precision;   /* mvr_precision of centroid scoring and MaxSim */
  uint32_t rescore;     /* BF16/FP8: FP32 rescoring of the best `rescore` docs (0 = off) */
  uint32_t clip_zero;   /* 0: IGP-PP semantics: the filter's per-token maxima are
                           max(0, .) (FilterPP.hpp:136-141,257-259), MaxSim per token
                           has identity -FLT_MAX.  1: the basic variant: the filter
                           adds each token's first hit as it is (igp_gpu_basic/
                           Filter.hpp:44-59), MaxSim per token is max(0, .) */
  uint32_t batch;       /* queries per device pass (0 = 1024) */
  uint32_t learned;     /* 1: per-query nprobe/probe_topk from the attached probe model */
  uint32_t persistent;  /* 1: MaxSim (and the rescore) + top-k as one persistent work-queue
                           kernel per pass (MVR_FP32_MFMA / MVR_BF16 / MVR_FP8, topk and
                           rescore <= 256; other settings are MVR_ERR_UNSUPPORTED) */
  uint32_t graph;       /* 1: each device pass is replayed from a captured hipGraph
                           (cached per pass shape; not with learned probing) */
} mvr_search_params;

/* Optional per-stage outputs (host buffers, any may be NULL), for tests. */
typedef struct {
  uint32_t *top_ids;   /* [n_q][t_max][nprobe_max] probed centroid ids */
  float *top_s;        /* [n_q][t_max][nprobe_max] centroid scores + 1 */
  uint32_t *cand_n;    /* [n_q] unique filter candidates */
  uint32_t *sel_ids;   /* [n_q][probe_topk] filter top docs, best first */
  float *sel_F;        /* [n_q][probe_topk] their filter scores */
  float *sel_score;    /* [n_q][probe_topk] their MaxSim scores (final precision) */
  uint32_t *nprobe;    /* [n_q] nprobe used (the model's choice with learned probing) */
  uint32_t *rerank_n;  /* [n_q] docs reranked (min(candidates, per-query probe_topk)) */
} mvr_stage_out;

/* Per-call device time per stage in milliseconds (HIP events on the call's
 * stream, summed over passes) and the work done: reranked doc tokens (sum of
 * the lengths of every reranked doc) and filter candidates, summed over queries.
 * rerank_ms is the MaxSim kernel alone; rescore_ms the FP32 rescore of a
 * BF16/FP8 search (its top-N selection and FP32 MaxSim; 0 otherwise).  A pass
 * replayed from a hipGraph only adds to total_ms.  With host-resident
 * residuals, rerank_ms also holds staging_ms, the time from the end of the
 * selection to the staged docs being on the device, and staged_docs /
 * staged_tokens count the docs copied from host memory (each once per pass). */
typedef struct {
  double centroid_ms, filter_ms, select_ms, rerank_ms, rescore_ms, topk_ms, total_ms;
  uint64_t rerank_tokens, candidates;
  uint32_t n_pass;        /* device passes the call made */
  uint64_t rerank_pairs;  /* sum over queries of (query tokens x reranked doc tokens) */
  double staging_ms;
  uint64_t staged_docs, staged_tokens;
} mvr_timings;

int mvr_device_count(int *count);
const char *mvr_last_error(void);

int mvr_index_create(const mvr_index_desc *desc, int device, mvr_index **out);
/* Loads centroid_l.npy, vq_code_l.npy, weight_l.npy, residual_code_l.npy,
 * item_n_vec_l.npy and ivf.index from index_dir (IGPGPUPP.cpp:33-73).  The
 * bit width comes from weight_l.npy's last dimension (2^n_bit). */
int mvr_index_load(const char *index_dir, int device, mvr_index **out);

/* Index flags for the _ex variants.
 * MVR_INDEX_RESIDUAL_ON_HOST: codes and residual bytes stay in host memory
 * (the residuals pinned), for corpora whose 36 bytes per token do not fit in
 * HBM; only centroids, doc lengths and the IVF lists go to the device.  It is
 * the reference's host-resident-codes rerank (RerankPP::rerank and
 * transmitSQCode, src/include/alg/igp_gpu_pp/RerankPP.hpp:289-361): each
 * device pass of mvr_search reads back its selected doc ids, gathers those
 * docs' codes and residual rows on the host and uploads them once for the
 * MaxSim stages.  Results are identical to a device-resident index.  Graph
 * replay is not used for such an index, and mvr_rerank_docs and
 * mvr_decompress return MVR_ERR_UNSUPPORTED. */
#define MVR_INDEX_RESIDUAL_ON_HOST 1u
int mvr_index_create_ex(const mvr_index_desc *desc, int device, uint32_t flags, mvr_index **out);
int mvr_index_load_ex(const char *index_dir, int device, uint32_t flags, mvr_index **out);
void mvr_index_free(mvr_index *ix);
/* Loop order of the BF16 MaxSim in mvr_search (no counterpart in the
 * reference, whose rerank is per query: RerankPP.hpp:416-471).  Both orders
 * give bit-identical scores.  MVR_MAXSIM_QUERY: per query, its candidate docs
 * in turn.  MVR_MAXSIM_DOC: the pass's (query, doc) pairs grouped by doc, each
 * doc's tokens loaded once for all the queries that selected it.
 * MVR_MAXSIM_AUTO (default): doc order when a device pass holds at least a
 * quarter as many pairs as the index has docs.  In a persistent search the doc order
 * makes the BF16 phase a work queue of docs followed by a per-query top-R2
 * (or top-k) launch; the query order runs the per-query persistent kernel
 * with its fused top-k; the automatic order is the autotuner's choice of the
 * two (mvr_index_get_tuning_ex).  The FP32 rescore phase is always per
 * query. */
enum { MVR_MAXSIM_AUTO = 0, MVR_MAXSIM_QUERY = 1, MVR_MAXSIM_DOC = 2 };
int mvr_index_set_maxsim_order(mvr_index *ix, uint32_t order);
int mvr_index_info(const mvr_index *ix, uint32_t *n_doc, uint64_t *n_tok, uint32_t *n_centroid,
                   uint32_t *max_list, uint32_t *max_doclen);

/* Batched search.  Q is [n_q][t_max][dim] (rows t >= q_ntok[q] are ignored);
 * q_ntok (NULL: all t_max) and q_nprobe (NULL: params->nprobe) are host
 * arrays.  With on_device = 0, Q and the outputs are host buffers and the call
 * returns when results are on the host.  With on_device != 0 they are device
 * buffers and work is queued on `stream` (a hipStream_t, NULL = default)
 * without host synchronisation.  Outputs are [n_q][topk], best first;
 * missing results are (UINT32_MAX, -inf). */
int mvr_search(mvr_index *ix, const float *Q, uint32_t n_q, uint32_t t_max, const uint32_t *q_ntok,
               const uint32_t *q_nprobe, const mvr_search_params *params, uint32_t *out_ids,
               float *out_scores, int on_device, void *stream, mvr_stage_out *stages,
               mvr_timings *timings);

/* MaxSim of given candidates (the rerank stage alone, RerankPP.hpp:368-475):
 * scores_dev[q][r] = sum_t max_i <Q[q][t], x_i> over the decompressed tokens of
 * doc_ids_dev[q][r] for r < n_docs[q] (host array; NULL: R for every query),
 * in `precision` (MVR_FP32 bit-exact with mvr_search; MVR_FP32_MFMA,
 * MVR_BF16 and MVR_FP8 on MFMA).
 * Device buffers Q [n_q][t_max][128], doc_ids [n_q][R], scores [n_q][R];
 * q_ntok is a host array (NULL: t_max).  Queued on `stream`. */
int mvr_rerank_docs(mvr_index *ix, const float *Q_dev, uint32_t n_q, uint32_t t_max,
                    const uint32_t *q_ntok, const uint32_t *doc_ids_dev, const uint32_t *n_docs,
                    uint32_t R, uint32_t precision, float *scores_dev, void *stream);

/* mvr_rerank_docs without allocating.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under hipGraph capture,
 * serving loops and harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results as mvr_rerank_docs; the per-query token and doc counts are
 * copied from the caller's host arrays into a static device array of the
 * library (2 x 16,384 u32 per device, held until the process ends) instead of
 * a workspace the index allocates.  n_q must be <= 16,384 (MVR_ERR_INVALID
 * beyond).  Every call on one device shares that array: calls that overlap on
 * one device (e.g. queued on two streams) race, so the caller serialises them. */
int mvr_rerank_docs_noalloc(mvr_index *ix, const float *Q_dev, uint32_t n_q, uint32_t t_max,
                            const uint32_t *q_ntok, const uint32_t *doc_ids_dev, const uint32_t *n_docs,
                            uint32_t R, uint32_t precision, float *scores_dev, void *stream);

/* Work-queue MaxSim tuning per precision (no counterpart in the reference,
 * which launches one kernel per query): workgroups per CU of the persistent
 * grid and doc tiles in flight per wave (16-token tiles in BF16 and FP8: 1 or
 * 2; the FP32 MFMA kernel streams one packed 32-row tile: 1).  The first
 * persistent search in a precision times every candidate setting on its first
 * pass and keeps the fastest, and with a rescore also tunes the FP32 MFMA
 * phase; set_tuning fixes a setting instead (wgs_per_cu from 1 to the number
 * of that kernel's workgroups a CU holds, MVR_ERR_INVALID beyond).
 * get_tuning writes (wgs_per_cu, tiles) and fails with MVR_ERR_INVALID until
 * a setting exists. */
int mvr_index_set_tuning(mvr_index *ix, uint32_t precision, uint32_t wgs_per_cu, uint32_t tiles);
int mvr_index_get_tuning(const mvr_index *ix, uint32_t precision, uint32_t out[2]);
/* The same with the first phase's loop order: out = (MVR_MAXSIM_QUERY or
 * MVR_MAXSIM_DOC, wgs_per_cu, tiles).  With the automatic MaxSim order
 * (mvr_index_set_maxsim_order) the autotuner of a persistent BF16 search also
 * times the doc-major first phase (a work queue of docs, then the per-query
 * top-R2) against the best per-query setting and keeps the faster; wgs_per_cu
 * and tiles are then still the best per-query setting, which searches in the
 * query order use.  No counterpart in the reference. */
int mvr_index_get_tuning_ex(const mvr_index *ix, uint32_t precision, uint32_t out[3]);

/* FP32 decompression of whole documents: out receives sum(doclens[doc_ids])
 * rows of dim floats, documents in the given order (host buffers). */
int mvr_decompress(mvr_index *ix, const uint32_t *doc_ids, uint32_t n, float *out);

/* FP64 exhaustive MaxSim over raw FP32 embeddings (the accuracy oracle recall
 * is measured against; Distance.hpp:200-218 vectorSetDistance summed over
 * query tokens): out[q][doc] = sum_{t < q_ntok[q]} max_i <Q[q][t], X_i> in
 * double.  Device buffers: Q [n_q][t_max][128] f32, q_ntok [n_q] (NULL: t_max),
 * X [sum(doclens)][128] f32, doclens [n_doc], out [n_q][n_doc] f64. */
int mvr_exhaustive_f64(const float *Q_dev, uint32_t n_q, uint32_t t_max, const uint32_t *q_ntok_dev,
                       const float *X_dev, const uint32_t *doclens_dev, uint32_t n_doc,
                       double *out_dev, void *stream);

/* ---- learned probe (no counterpart in the reference, which probes a fixed
 * nprobe: GenerateVQScoreCandidatePP.hpp:166-174) ----
 * A 3-layer MLP (dims {1024, d1, d2, 6}, GELU) reads a query's top-32
 * centroid scores per token and predicts its nprobe class over {2,4,8,16,32}
 * (outputs 0..4) and log(probe_topk) (output 5).  Trained with Adam on the
 * device; parameters are one flat array W1 b1 W2 b2 W3 b3 (row-major W[out][in]). */
typedef struct mvr_mlp mvr_mlp;
const char *mvr_mlp_last_error(void);
int mvr_mlp_create(const uint32_t *dims, int device, mvr_mlp **out);
void mvr_mlp_free(mvr_mlp *mlp);
uint64_t mvr_mlp_num_params(const mvr_mlp *mlp);
int mvr_mlp_set_params(mvr_mlp *mlp, const float *flat_host);      /* also resets Adam state */
int mvr_mlp_get_params(const mvr_mlp *mlp, float *flat_host);
/* out_dev [B][6] = network(x_dev [B][dims[0]]) */
int mvr_mlp_forward(mvr_mlp *mlp, const float *x_dev, uint32_t B, float *out_dev, void *stream);
/* loss (mean CE + mean squared error) and gradients for labels y (class
 * index) and targets r, without updating; grads_host has num_params floats. */
int mvr_mlp_grads(mvr_mlp *mlp, const float *x_dev, const int32_t *y_dev, const float *r_dev,
                  uint32_t B, float *grads_host, float *loss_host);
/* one Adam step (lr; betas 0.9/0.999, eps 1e-8); loss_host may be NULL */
int mvr_mlp_train_step(mvr_mlp *mlp, const float *x_dev, const int32_t *y_dev, const float *r_dev,
                       uint32_t B, float lr, float *loss_host, void *stream);
/* Attach a probe model to an index: searches with params->learned = 1 score
 * the top-32 centroids per token, run the model, and probe the predicted
 * nprobe (fallback_nprobe when the top class probability is below conf) with
 * probe_topk min(params->probe_topk, max(r_min, round(exp(output 5)))).
 * The index keeps the pointer; pass NULL to detach. */
int mvr_index_set_probe(mvr_index *ix, mvr_mlp *mlp, float conf, uint32_t fallback_nprobe,
                        uint32_t r_min);

/* ---- exhaustive MaxSim (the groundtruth module) ---- */

/* All-pairs FP32 MaxSim, replacing GPUComputeScore::computeItemScore
 * (src/include/compute_score/GPUComputeScore.hpp:196-246) behind
 * BruteForceProgressive.DocRetrieval.computeScore
 * (src/app/BruteForceProgressive.cpp:16-33): scores[q][i] = left-to-right sum
 * over the T query tokens of max over item i's vectors of dot(Q_q,t, x), each
 * dot one FMA chain over d.  Host arrays: Q f32 [n_q][T][dim] (T <= 32,
 * dim = 128), X f32 [sum lens][dim] (items back to back), lens [n_item]
 * (every item >= 1 vector), scores f32 [n_q][n_item].  Runs on `device`;
 * errors are described by mvr_maxsim_last_error(). */
int mvr_maxsim_scores_host(const float *Q, uint32_t n_q, uint32_t T, uint32_t dim, const float *X,
                           const uint32_t *lens, uint32_t n_item, float *scores, int device);
/* The same with precision MVR_FP32 (as above) or MVR_FP32_MFMA: the dot
 * products on FP32 MFMA, summed in the matrix unit's order (within 1e-5
 * relative of MVR_FP32), as the reference's cuBLAS SGEMM leaves its order to
 * the library (GPUComputeScore.hpp:196-246). */
int mvr_maxsim_scores_host_ex(const float *Q, uint32_t n_q, uint32_t T, uint32_t dim, const float *X,
                              const uint32_t *lens, uint32_t n_item, float *scores, int device,
                              uint32_t precision);
const char *mvr_maxsim_last_error(void);

/* ---- synthetic corpus and index build on the device (bench and tests) ---- */

/* Mixture tokens: row k of out (device, [n_tok][dim]) is token tok_begin+k of
 * the stream pair (stream_comp, stream_noise): normalise(centers[comp] +
 * noise_std * z) with comp and z from the counter-based generator. */
int mvr_synth_tokens(uint64_t seed, uint64_t stream_comp, uint64_t stream_noise,
                     const float *centers_dev, uint32_t n_comp, uint32_t dim, double noise_std,
                     uint64_t tok_begin, uint64_t n_tok, float *out_dev, void *stream);
/* argmax_c dot(x_i, C_c) (lowest c on ties), FP32, device buffers. */
int mvr_assign(const float *x_dev, uint64_t n, const float *C_dev, uint32_t n_c, uint32_t dim,
               uint32_t *codes_dev, void *stream);
/* argmax_c (dot(x_i, C_c) - half_norm[c]) for Lloyd iterations (half_norm may be NULL). */
int mvr_assign_bias(const float *x_dev, uint64_t n, const float *C_dev, const float *half_norm_dev,
                    uint32_t n_c, uint32_t dim, uint32_t *codes_dev, void *stream);
/* Packed 2-bit residual codes of x against C[codes] with cutoffs [dim][3]
 * (per_dim = 1) or [3]; device buffers. */
int mvr_compress(const float *x_dev, const uint32_t *codes_dev, uint64_t n, const float *C_dev,
                 uint32_t dim, const float *cutoffs_dev, int per_dim, uint8_t *out_dev, void *stream);
/* The same on host arrays, run on `device` in chunks of 4M vectors: the GPU
 * residual coder behind CompressResidualCode.compute_residual_code
 * (compressSQCode, src/include/alg/refine/ResidualScalarQuantizationPackGPU.hpp:30-80;
 * bound at src/app/build_index/GPGPUBuildIndexPackGPU.cpp:14-42).  Codes are
 * checked against n_c; dim must be 128 (2-bit codes, 32 bytes per vector). */
int mvr_compress_host(const float *x, const uint32_t *codes, uint64_t n, const float *C, uint32_t n_c,
                      uint32_t dim, const float *cutoffs, int per_dim, uint8_t *out, int device);

#ifdef __cplusplus
}
#endif
#endif /* MVR_H */
