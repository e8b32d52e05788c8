// Internal declarations shared by the libmvr translation units (kernel
// launchers and the C-ABI layer).  All kernels are written for gfx950
// (wave64, 160 KB LDS per CU) and specialise dim = 128, 2-bit residuals.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace mvr {

constexpr uint32_t kDim = 128;             // embedding width
constexpr uint32_t kResBytes = 32;         // packed 2-bit residual bytes per token
constexpr uint32_t kMaxQTok = 32;          // query tokens (bit mask width in the filter)
constexpr uint32_t kMaxProbe = 256;        // largest nprobe
constexpr uint32_t kProbeFeat = 32;        // centroid scores per token read by the learned probe (its largest nprobe)
constexpr uint32_t kSeg = 4096;            // docs per filter segment (LDS state)
constexpr uint64_t kPoolBudget = 16ull << 30;  // candidate pool bytes per device pass (default)
constexpr uint64_t kPoolBudgetMin = 2ull << 30, kPoolBudgetMax = 64ull << 30;
constexpr uint32_t kMaxSelect = 16384;     // largest probe_topk (and PLAID ndocs)
constexpr double kFxScale = 4294967296.0;  // filter fixed point: 2^32

// Device-resident index (layout mirrors the reference files).
struct DevIndex {
  uint32_t n_doc = 0, n_c = 0, n_seg = 0, max_list = 0, max_doclen = 0;
  uint64_t n_tok = 0, n_ele = 0;
  float c_norm_max = 0.0f;     // max_c ||C_c||_2 (bounds the centroid-score error)
  uint32_t n_cu = 0;           // compute units of the device
  float *C = nullptr;          // [n_c][128] f32
  uint16_t *C_bf16 = nullptr;  // [n_c][128] bf16 (RNE of C)
  uint16_t *C_bf16m = nullptr; // [n_c][128] the same values in MaxSim order: piece 4s + g = dims 32g + 8s .. +7
  uint8_t *C_fp8 = nullptr;    // [n_c][128] e4m3 of C / scale
  float *C_fp8_scale = nullptr;// [n_c]
  uint32_t *codes = nullptr;   // [n_tok]
  uint32_t n_bit = 2;          // residual bits per dim: 1, 2, 4 or 8 (only 2 outside the exact FP32 path)
  uint8_t *residual = nullptr; // [n_tok][16 * n_bit], values packed MSB-first
  float *W = nullptr;          // [128][2^n_bit] decode table
  uint32_t *doclens = nullptr; // [n_doc]
  uint64_t *docoff = nullptr;  // [n_doc]
  uint64_t *ivf_off = nullptr; // [n_c]
  uint32_t *ivf_size = nullptr;// [n_c]
  uint32_t *ivf_ids = nullptr; // [n_ele]
  uint32_t *seg_off = nullptr; // [n_c][n_seg+1] list index where each segment starts
  uint32_t *lut_fp8 = nullptr; // [32][256]: 4 e4m3 residual values / fp8_res_scale
  float fp8_res_scale = 0.0f;  // max |W| / 448
};

// ---- launchers (return hipError_t of the launch) ----
hipError_t launch_seg_offsets(const DevIndex &ix, hipStream_t s);
// bad[0] = codes >= n_c (ix.codes may be null: skipped), bad[1] = IVF entries
// >= n_doc or not strictly ascending in their list
hipError_t launch_validate_index(const DevIndex &ix, unsigned long long *bad, hipStream_t s);
hipError_t launch_convert_centroids(const DevIndex &ix, hipStream_t s);

// centroid scoring + top-P.  rows: n_rows query-token rows; row_map (nullable)
// maps row -> index into Q rows and into the output rows.
hipError_t launch_centroid_topk(const float *Q, const uint32_t *row_map, uint32_t n_rows,
                                uint32_t n_out_rows, const DevIndex &ix, uint32_t P, int precision,
                                uint32_t n_split, uint32_t *part_ids, float *part_s,
                                uint32_t *top_ids, float *top_s, const uint16_t *Qbf,
                                hipStream_t s);

// BF16 centroid scoring: MFMA tile maxima into bmax [n_q*t_max][ceil(n_c/32)],
// then per row exact top-P over the best P+2 tiles (bf16 inputs, FP32 FMA chain).
// exact_f32: the top-P is exact on FP32 inputs (BF16 MFMA tile maxima only
// choose which tiles are rescanned); else exact on bf16-rounded inputs
hipError_t launch_centroid_topk_bf16(const float *Q, uint32_t n_q, uint32_t t_max,
                                     const uint32_t *row_map, uint32_t n_rows, const DevIndex &ix,
                                     uint32_t P, float *bmax, uint32_t *top_ids, float *top_s,
                                     bool exact_f32, hipStream_t s);

// FP8 centroid scoring (the FP8 mode's probe, approximate): e4m3 MFMA tile
// maxima with per-token and per-centroid scales, then the exact FP32 top-P
// over the centroids of the kFp8ProbeTiles * P tiles with the highest maxima
constexpr uint32_t kFp8ProbeTiles = 2;
hipError_t launch_centroid_topk_fp8(const float *Q, uint32_t n_q, uint32_t t_max, const uint32_t *row_map,
                                    uint32_t n_rows, const DevIndex &ix, uint32_t P, float *bmax,
                                    uint32_t *top_ids, float *top_s, hipStream_t s);

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);

// clamp0: per-(doc, token) maxima are max(0, .) (the PP filter); else the
// first hit's score as it is (the basic variant)
// Candidate keys go to cand_keys[q * keys_per_q ..], q_fill[q] of them (q_fill:
// n_q counters, zeroed by the launch); keys_per_q must be at least the number
// of distinct docs a query's probed lists can hold (candidate_keys_per_query).
hipError_t launch_filter(const uint32_t *top_ids, const float *top_s, uint32_t t_max, uint32_t P,
                         const uint32_t *q_ntok, const uint32_t *q_np, uint32_t n_q,
                         const DevIndex &ix, int clamp0, uint32_t *q_fill, uint64_t keys_per_q,
                         uint64_t *cand_keys, hipStream_t s);
// min(docs in the index rounded up to whole segments, t_max * P * longest IVF
// list): a query's candidates are distinct docs of its t_max * P probed lists
inline uint64_t candidate_keys_per_query(const DevIndex &ix, uint32_t t_max, uint32_t P) {
  const uint64_t all = (uint64_t)ix.n_seg * kSeg;
  const uint64_t lists = (uint64_t)t_max * P * (ix.max_list ? ix.max_list : 1u);
  return lists < all ? lists : all;
}

// top-R per query (q_r, nullable: per-query R <= R; R is the output row stride)
// over each query's candidate keys (launch_filter's layout); cand_tot[q]
// receives the query's candidate count
hipError_t launch_select(const uint64_t *cand_keys, const uint32_t *q_fill, uint64_t keys_per_q,
                         const uint32_t *q_r, uint32_t n_q, uint32_t R, uint32_t *sel_ids,
                         float *sel_F, uint32_t *sel_n, uint32_t *cand_tot, hipStream_t s);
// The same over per-(query, segment) regions: seg_cnt[q][sg] keys at
// cand_keys + (q * n_seg + sg) * kSeg (the PLAID filter's layout).
hipError_t launch_select_segments(const uint64_t *cand_keys, const uint32_t *seg_cnt, uint32_t n_seg,
                                  const uint32_t *q_r, uint32_t n_q, uint32_t R, uint32_t *sel_ids,
                                  float *sel_F, uint32_t *sel_n, uint32_t *cand_tot, hipStream_t s);

// persistent work-queue MaxSim + top-kk (kk <= 256) over whole queries;
// precision 1 (BF16) or 3 (FP32 MFMA).  queue [1] is scratch.
struct PersistCfg {
  uint32_t n_wg = 0;   // workgroups in the grid
  uint32_t tps = 1;    // doc tiles in flight per wave
  bool dm = false;     // BF16 first phase doc-major (a work queue of docs), not per query
};
constexpr uint32_t kPersistTopkMax = 256;
// resident workgroups per CU of the persistent kernel for (precision, tiles
// in flight): the largest useful grid is this times the CU count; 0 for a
// combination without a kernel
uint32_t persist_max_wgs(int precision, uint32_t tps);
hipError_t launch_persist(const float *Q, uint32_t t_max, const uint32_t *q_ntok,
                          const uint32_t *in_ids, const uint32_t *in_n, uint32_t R_in, uint32_t nb,
                          const DevIndex &ix, int precision, int clip_zero, const PersistCfg &cfg,
                          uint32_t *queue, float *score, uint32_t kk, uint32_t *out_ids, float *out_s,
                          uint32_t *out_n, uint32_t out_stride, hipStream_t s);

// learned probe: features [n_q][1024] from the top-32 scores, and decisions
hipError_t launch_probe_features(const float *top_s, uint32_t t_max, uint32_t P,
                                 const uint32_t *q_ntok, uint32_t n_q, float *X, hipStream_t s);
hipError_t launch_probe_decide(const float *O, uint32_t n_q, float conf, uint32_t fallback,
                               uint32_t r_min, uint32_t r_max, uint32_t *q_np, uint32_t *q_r,
                               hipStream_t s);

hipError_t launch_rerank(const float *Q, uint32_t t_max, const uint32_t *q_ntok, uint32_t n_q,
                         const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R,
                         const DevIndex &ix, int precision, int clip_zero, float *sel_score,
                         hipStream_t s);

// Doc-major BF16 MaxSim: the pass's (query, slot) pairs grouped by doc, each
// doc's tiles loaded once for all of its queries.  Buffers (device):
struct DocMajorBufs {
  uint32_t *cnt = nullptr;    // [n_doc] pairs per doc, then the scatter cursor, then each doc's end
  uint2 *work = nullptr;      // [n_doc] (doc, first pair) of every doc with pairs, by doc id
  uint32_t *plist = nullptr;  // [n_q * R] pair ids q * R + slot, grouped by doc
  uint32_t *b_p = nullptr;    // [ceil(n_doc / 4096)] pairs per block of docs, then their prefix
  uint32_t *b_d = nullptr;    // [ceil(n_doc / 4096)] docs with pairs per block, then their prefix
  uint32_t *tot = nullptr;    // [3] pairs, work items, work-queue counter
  uint4 *Qm = nullptr;        // [n_q][512] bf16 query operands in MFMA lane order, or the
                              // FP8 ones ([n_q][256]) followed by their token scales
};
constexpr uint32_t kDmDocsPerBlock = 4096;
// Same scores, bit for bit, as launch_rerank with precision 1 (BF16).
hipError_t launch_rerank_bf16_dm(const float *Q, uint32_t t_max, const uint32_t *q_ntok, uint32_t n_q,
                                 const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R,
                                 const DevIndex &ix, int clip_zero, const DocMajorBufs &b,
                                 float *sel_score, hipStream_t s);

// Doc-major FP8 MaxSim: same scores, bit for bit, as launch_rerank with
// precision 2 (FP8); b.Qm must hold n_q * (4096 + 128) bytes.
hipError_t launch_rerank_fp8_dm(const float *Q, uint32_t t_max, const uint32_t *q_ntok, uint32_t n_q,
                                const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R,
                                const DevIndex &ix, int clip_zero, const DocMajorBufs &b,
                                float *sel_score, hipStream_t s);

// Per-pass compaction buffers of a host-resident-residual index (device).
struct HostStage {
  uint32_t *bm = nullptr;           // [ceil(n_doc/32)] selected-doc bitmap
  uint32_t *b_n = nullptr;          // [blocks of 256 words] docs per block, then their exclusive prefix
  unsigned long long *b_t = nullptr;// [blocks] tokens per block, then prefix
  uint32_t *wbase = nullptr;        // [ceil(n_doc/32)] compact id of the word's first doc
  uint32_t *docs = nullptr;         // [U] compact id -> doc id (ascending)
  uint32_t *lens = nullptr;         // [U]
  uint64_t *offs = nullptr;         // [U] token offset in the pass buffers
  uint32_t *cids = nullptr;         // [nb][R] compact id of each selected doc
  unsigned long long *cnt = nullptr;// [2] U, T
};

// marks the selected docs and writes the compact lists, cids and cnt
hipError_t launch_hs_compact(const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R, uint32_t nb,
                             const uint32_t *doclens, uint32_t n_doc, const HostStage &hs, hipStream_t s);
// copies the U compact docs' codes and residual rows from host-mapped memory
hipError_t launch_hs_gather(const HostStage &hs, uint32_t U, const uint64_t *docoff, const uint32_t *h_codes,
                            const uint8_t *h_res, uint32_t *codes, uint8_t *res, hipStream_t s);

// ids[i] = map[ids[i]] for ids[i] < n_map, other ids unchanged
hipError_t launch_map_ids(uint32_t *ids, uint32_t n, const uint32_t *map, uint32_t n_map, hipStream_t s);

// best k of n_in[q] (score, id) pairs; out_n (nullable) receives min(n_in, k)
hipError_t launch_topk(const uint32_t *ids, const float *score, const uint32_t *n_in, uint32_t R,
                       uint32_t n_q, uint32_t k, uint32_t *out_ids, float *out_s, uint32_t *out_n,
                       hipStream_t s);

// counters[0] += sum of doclens over the selected docs, counters[1] += candidates
hipError_t launch_count_work(const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R, uint32_t n_q,
                             const DevIndex &ix, const uint32_t *cand_cnt, const uint32_t *q_ntok,
                             unsigned long long *counters, hipStream_t s);

hipError_t launch_exhaustive_f64(const float *Q, uint32_t n_q, uint32_t t_max,
                                 const uint32_t *q_ntok, const float *X, const uint32_t *doclens,
                                 const uint64_t *docoff, uint32_t n_doc, double *out, hipStream_t s);

hipError_t launch_decompress(const uint64_t *tok_ids, uint64_t n, const DevIndex &ix, float *out,
                             hipStream_t s);

hipError_t launch_synth_tokens(uint64_t seed, uint64_t sc, uint64_t sn, const float *centers,
                               uint32_t n_comp, double noise, uint64_t tok_begin, uint64_t n_tok,
                               float *out, hipStream_t s);
hipError_t launch_assign(const float *x, uint64_t n, const float *C, const float *bias,
                         uint32_t n_c, uint32_t *codes, hipStream_t s);
hipError_t launch_compress(const float *x, const uint32_t *codes, uint64_t n, const float *C,
                           const float *cut, int per_dim, uint8_t *out, hipStream_t s);

}  // namespace mvr
