/*
 * mvr_plaid.h -- C ABI of the PLAID (ColBERTv2) search path and residual
 * codec ops on MI355X (gfx950).  Part of libmvr.so.
 *
 * Plain pointers and sizes only; every function returns an mvr_status
 * (include/mvr.h) and mvr_plaid_last_error() gives the message of the last
 * failure on the calling thread.  The library never calls exit().
 *
 * The single ops take DEVICE pointers and run on `stream` (a hipStream_t, or
 * NULL for the null stream).  Dtypes are those of the reference's tensors:
 * int32 pids and codes, int64 lengths and offsets, float16 passed as its
 * uint16 bits, bool as uint8.  The two-stage filter and the segmented MaxSim
 * return host-visible counts and therefore synchronise `stream`.
 *
 * Arithmetic (shared with the CPU oracle, bit for bit): every FP32 dot
 * product and squared norm is a sequential FMA chain over d = 0..dim-1; the
 * filter score is the left-to-right FP32 sum over query tokens from 0.0f; the
 * refine score is the left-to-right FP32 sum of max(0, max_i s) over query
 * tokens; the filter keeps pids by (score, pid) descending and the final
 * ranking orders by (score desc, pid asc).
 *
 * The GPU branch of the reference (index_storage.py:112-153,160-161,185-187,
 * the path eval_plaid_gpu.py runs) computes in float16: halves for
 * centroids, bucket weights and queries; every matmul output and every
 * reduction result rounded to half once, with FP32 accumulation in the
 * orders above; the padding value of colbert_score_reduce is half(-9999) =
 * -10000, and MaxSim has no zero clip.  mvr_plaid_search with gpu_branch = 1
 * and the *_f16 ops below follow it.
 */
#ifndef MVR_PLAID_H
#define MVR_PLAID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *mvr_plaid_last_error(void);

/* ---- residual codec ops (ColBERT indexing/codecs) ---- */

/* decompress_residuals_cpp, CUDA signature
 * (baseline/ColBERT/colbert/indexing/codecs/decompress_residuals.cu:42-75;
 * bound at residual.py:98-115, called from ResidualCodec.decompress
 * residual.py:243-279):
 *   x = reversed_bit_map[binary_residuals[i][j]];
 *   out[i][j*8/nbits + k] = half(float(bucket_weights[lut[x][k]]) + float(centroids[codes[i]][.]))
 * binary_residuals [n][dim*nbits/8] u8, bucket_weights f16 [2^nbits],
 * reversed_bit_map [256], lut [256][8/nbits], codes [n], centroids f16
 * [n_centroid][dim], out f16 [n][dim].  nbits in {1, 2, 4, 8}, dim a multiple
 * of 8. */
int mvr_plaid_decompress_residuals(const uint8_t *binary_residuals, uint64_t n,
                                   const uint16_t *bucket_weights, const uint8_t *reversed_bit_map,
                                   const uint8_t *bucket_weight_combinations, const int32_t *codes,
                                   const uint16_t *centroids, uint32_t dim, uint32_t nbits,
                                   uint16_t *out, void *stream);

/* packbits_cpp (indexing/codecs/packbits.cu:39-57; residual.py:117-130):
 * out[j] = bits 8j..8j+7 (nonzero byte = 1) with bit 8j in the MSB, the
 * numpy.packbits order.  m must be a multiple of 32 (packbits.cu:47). */
int mvr_plaid_packbits(const uint8_t *residuals, uint64_t m, uint8_t *packed, void *stream);

/* ResidualCodec.binarize (residual.py:189-224) fused: bucket =
 * bucketize(e, cutoffs, right=False) (cutoffs strictly below e), its nbits
 * bits LSB first, packed MSB first.  residuals f32 [n][dim], cutoffs
 * [n_cutoffs] (n_cutoffs < 2^nbits), out [n][dim*nbits/8]. */
int mvr_plaid_binarize(const float *residuals, uint64_t n, uint32_t dim, const float *bucket_cutoffs,
                       uint32_t n_cutoffs, uint32_t nbits, uint8_t *out, void *stream);

/* ---- search-side ops (ColBERT search/ and modeling/) ---- */

/* cum[0] = 0, cum[i+1] = cum[i] + lengths[pids ? pids[i] : i] for i < n, and
 * *total_host (nullable; synchronises) = cum[n].  The cumulative lengths of
 * decompress_residuals.cpp:104-110 and segmented_maxsim.cpp:64-66. */
int mvr_plaid_lengths_cumsum(const int32_t *pids, uint32_t n, const int64_t *lengths, int64_t *cum,
                             int64_t *total_host, void *stream);

/* decompress_residuals_cpp, CPU signature
 * (baseline/ColBERT/colbert/search/decompress_residuals.cpp:87-155; bound at
 * index_storage.py:63, called at :169-181): the tokens of pids[0..npids) in
 * that order; row cum[i] + j of out (f32 [cum[npids]][dim]) is token j of
 * pids[i]:  out = bucket_weights[lut[rbm[byte]][l]] + centroids[code][.] in
 * FP32.  cum from mvr_plaid_lengths_cumsum(pids, npids, lengths). */
int mvr_plaid_decompress_pids(const int32_t *pids, uint32_t npids, const int64_t *cum,
                              const int64_t *lengths, const int64_t *offsets,
                              const float *bucket_weights, const uint8_t *reversed_bit_map,
                              const uint8_t *bucket_weight_combinations,
                              const uint8_t *binary_residuals, const int32_t *codes,
                              const float *centroids, uint32_t dim, uint32_t nbits, float *out,
                              void *stream);

/* segmented_maxsim_cpp (baseline/ColBERT/colbert/modeling/segmented_maxsim.cpp:47-93;
 * called by colbert_score_packed, colbert.py:181-205): out[d] = sum over t
 * (left to right) of max(0, max over doc d's rows of scores[row][t]).
 * scores f32 [sum(lengths)][nq] (nq <= 64), lengths [ndocs]. */
int mvr_plaid_segmented_maxsim(const float *scores, uint32_t nq, const int64_t *lengths,
                               uint32_t ndocs, float *out, void *stream);

/* colbert_score_packed, GPU branch (baseline/ColBERT/colbert/modeling/colbert.py:181-205,
 * called at index_storage.py:185-187 after lookup_pids): one query's half
 * rows Q16 [nq][128] (nq <= 32) against the half rows D_packed
 * [D_rows][128] (D_rows >= sum(D_lengths), lengths >= 0, else
 * MVR_ERR_INVALID): s = half(FMA chain of D . Q_t), per doc M_t = max
 * over its rows with the padding value -10000 as identity (exact whenever
 * the scores are >= -10000, as dot products of normalised rows are), out[d] =
 * half(left-to-right FP32 sum over t).  Device pointers, halves as uint16
 * bits; synchronises `stream` (row offsets). */
int mvr_plaid_colbert_score_packed_f16(const uint16_t *Q16, uint32_t nq, const uint16_t *D_packed, uint64_t D_rows,
                                       const int64_t *D_lengths, uint32_t ndocs, uint16_t *out, void *stream);

/* colbert_score_reduce, GPU branch, on its padded input
 * (baseline/ColBERT/colbert/modeling/colbert.py:133-155; called at
 * index_storage.py:138,149 with the centroid scores of each doc's codes):
 * scores_padded f16 [n][L][nq] (nq <= 64), D_mask u8 [n][L]; the entries
 * whose mask is 0 read as -10000 (half(-9999)), out[i] = half(left-to-right
 * FP32 sum over t of the max over the L positions).  L >= 1. */
int mvr_plaid_colbert_score_reduce_f16(const uint16_t *scores_padded, const uint8_t *D_mask, uint32_t n, uint32_t L,
                                       uint32_t nq, uint16_t *out, void *stream);

/* segmented_lookup_cpp (baseline/ColBERT/colbert/search/segmented_lookup.cpp:52-148,
 * bound at strided_tensor.py:20-38 and called at :93 for the IVF lookups of
 * candidate_generation.py:26,36): out = the rows input[offsets[i] ..
 * offsets[i] + lengths[i]) for i = 0 .. n-1, concatenated.  A row is
 * row_bytes bytes (dim x the element size; the reference's u8, i32, i64, f32
 * and f16 inputs are all byte copies).  input holds n_rows rows; out must hold
 * out_rows = sum(lengths) rows (checked).  Device pointers; returns after the
 * copy on `stream`.  MVR_ERR_INVALID if a segment lies outside the input (the
 * reference copies out of bounds). */
int mvr_plaid_segmented_lookup(const void *input, uint64_t n_rows, uint32_t row_bytes, const int64_t *lengths,
                               const int64_t *offsets, uint32_t n, void *out, uint64_t out_rows, void *stream);

/* filter_pids_cpp (baseline/ColBERT/colbert/search/filter_pids.cpp:147-193;
 * bound at index_storage.py:46, called at :154-159): stage 1 scores every pid
 * over its codes c with idx[c] (per token the max of centroid_scores[c][t],
 * identity -9999, summed over t) and keeps the best ndocs by (score, pid)
 * descending; stage 2 rescores those over all their codes and keeps ndocs/4,
 * written to out_pids in that order.  centroid_scores f32 [n_centroid][nq]
 * (nq <= 32), idx u8 [n_centroid], ndocs <= 16384.  *out_n (host) =
 * min(ndocs/4, min(ndocs, npids)) (the reference requires npids >= ndocs);
 * *n_vec_score_refine (host, nullable) = nq x the number of distinct codes
 * visited by both stages.  Synchronises stream. */
int mvr_plaid_filter_pids(const int32_t *pids, uint32_t npids, const float *centroid_scores,
                          uint32_t n_centroid, uint32_t nq, const int32_t *codes,
                          const int64_t *doclens, const int64_t *offsets, const uint8_t *idx,
                          uint32_t ndocs, int32_t *out_pids, uint32_t *out_n,
                          uint64_t *n_vec_score_refine, void *stream);

/* ---- batched search: IndexScorer.rank, CPU branch, on the device ---- */

typedef struct mvr_plaid_index mvr_plaid_index;

/* A PLAID index (the arrays IndexScorer / ResidualCodec load:
 * index_storage.py:23-30, residual.py:133-150, ivf.pid.pt), host pointers,
 * copied to the device.  dim = 128 and nbits = 2 for the search kernels. */
typedef struct {
  uint32_t n_doc, n_centroid, dim, nbits;
  uint64_t n_tok;                /* sum of doclens */
  uint64_t n_ivf;                /* sum of ivf_lengths */
  const int64_t *doclens;        /* [n_doc] */
  const int32_t *codes;          /* [n_tok] */
  const uint8_t *residuals;      /* [n_tok][dim*nbits/8], PLAID bit order */
  const float *centroids;        /* [n_centroid][dim] (the CPU branch widens f16 to f32) */
  const float *bucket_weights;   /* [2^nbits] */
  const int32_t *ivf;            /* [n_ivf] per centroid: exactly the docs holding it, ascending */
  const int64_t *ivf_lengths;    /* [n_centroid] */
} mvr_plaid_index_desc;

/* Checks the arrays first (codes in range, doclens summing to n_tok, every
 * IVF list equal to the docs that hold its centroid; MVR_ERR_INVALID
 * otherwise), then uploads them to `device`. */
int mvr_plaid_index_create(const mvr_plaid_index_desc *desc, int device, mvr_plaid_index **out);
void mvr_plaid_index_free(mvr_plaid_index *ix);

/* Searcher.dense_search settings (searcher.py:291-312): ncells centroids per
 * query token, centroid_score_threshold (idx = max_t S >= threshold), ndocs
 * (<= 16384; stage 2 keeps ndocs/4), k results, queries per device pass.
 * gpu_branch: 0 = the CPU branch's semantics (FP32, filter_pids,
 * zero-initialised MaxSim), 1 = the GPU branch's (float16 centroid scores,
 * idx against half(threshold), both filter stages through
 * colbert_score_reduce with -10000 padding, half decompression and
 * normalisation, colbert_score_packed without zero clip; scores are halves). */
typedef struct {
  uint32_t k, ncells, ndocs, batch;
  float centroid_score_threshold;
  uint32_t gpu_branch;
} mvr_plaid_params;

typedef struct {
  double candidates_ms, filter_ms, refine_ms, total_ms;
  uint64_t n_refine_ivf, n_refine_filter, n_vec_score_refine; /* sums over the queries */
  double scores_ms;  /* the centroid-score kernel alone (part of candidates_ms) */
  uint32_t n_pass;   /* device passes */
} mvr_plaid_timings;

/* IndexScorer.rank (index_storage.py:82-198, CPU or GPU branch by
 * params->gpu_branch) for n_q queries of
 * nq tokens (nq <= 32): Q f32 [n_q][nq][dim], host memory unless
 * q_on_device.  Writes out_ids (int32, -1 when fewer than k) and out_scores
 * [n_q][k]; stats (nullable, host) [n_q][3] = (pids after candidate
 * generation, pids after the filter, n_vec_score_refine); timings nullable. */
int mvr_plaid_search(mvr_plaid_index *ix, const float *Q, uint32_t n_q, uint32_t nq,
                     const mvr_plaid_params *params, int32_t *out_ids, float *out_scores,
                     int q_on_device, uint64_t *stats, mvr_plaid_timings *timings);

/* ResidualEmbeddingsStrided.lookup_pids + ResidualCodec.decompress, GPU
 * branch (baseline/ColBERT/colbert/indexing/codecs/residual_embeddings_strided.py:30-37,
 * residual.py:243-279; called at index_storage.py:160-161): the tokens of
 * pids[0..npids) (device int32) in order as normalised half rows, out f16
 * [out_rows][128]: x_d = half(w + c) with half bucket weights and centroids,
 * y_d = half(x_d / half(sqrtf(FP32 sum of x_d^2))).  *n_rows (host) =
 * sum of the pids' doclens; MVR_ERR_INVALID when it exceeds out_rows or a
 * pid lies outside the index.  Synchronises `stream`. */
int mvr_plaid_lookup_pids_f16(mvr_plaid_index *ix, const int32_t *pids, uint32_t npids, uint16_t *out,
                              uint64_t out_rows, uint64_t *n_rows, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MVR_PLAID_H */
