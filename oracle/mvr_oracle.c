/*
 * mvr_oracle.c -- CPU oracle for the ColBERT late-interaction search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libmvr, the IGPGPUPP/IGPGPUBasic executables, the mvr Python
 * package) never links or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's IGP-PP search path
 * and its index format:
 *
 *   centroid scoring   src/include/alg/igp_gpu_pp/GenerateVQScoreCandidatePP.hpp:137-174
 *   IVF probe/filter   src/include/alg/igp_gpu_pp/FilterPP.hpp:267-398 (max_qvec_doc_score :99-142)
 *   decompress         src/include/alg/igp_gpu_pp/RerankPP.hpp:65-113, DataStructPP.hpp:147-205
 *   MaxSim             src/include/alg/igp_gpu_pp/RerankPP.hpp:115-161,416-471
 *   final top-k        src/include/alg/igp_gpu_pp/IGPGPURetrievalPP.hpp:153-161
 *   codec              src/include/alg/refine/ResidualScalarQuantizationPack.hpp:17-237
 *   IVF build / file   src/include/struct/IVFIndex.hpp:27-183
 *   brute force        src/include/alg/Distance.hpp:200-218 (vectorSetDistance)
 *
 * Where the reference leaves an order unpinned (cuBLAS accumulation order, raft
 * select_k tie order, raft reduce summation order) this file fixes one, and the
 * GPU kernels follow the same definition:
 *   - every FP32 dot product is a sequential fused multiply-add chain over
 *     d = 0..dim-1 starting from 0.0f;
 *   - the filter score F[d] is the correctly rounded FP32 value of the exact
 *     sum of per-token maxima, each truncated to a multiple of 2^-32
 *     (int64 fixed point), so it does not depend on summation order;
 *   - the rerank score is the left-to-right FP32 sum over query tokens;
 *   - every top-n selection orders by (score descending, id ascending).
 *
 * Build: see oracle/Makefile (gcc -O2 -mfma -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <stdio.h>
#include <immintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_FX_SCALE 4294967296.0          /* 2^32 */
#define ORC_FX_INV (1.0 / 4294967296.0)

/* ------------------------------------------------------------------------ */
/* Counter-based generator (splitmix64 mixing, Box-Muller normals).          */
/* ------------------------------------------------------------------------ */

static inline uint64_t orc_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

uint64_t orc_rand_u64(uint64_t seed, uint64_t stream, uint64_t i) {
  return orc_mix64(orc_mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) + i);
}

double orc_uniform(uint64_t seed, uint64_t stream, uint64_t i) {
  return (double)(orc_rand_u64(seed, stream, i) >> 11) * (1.0 / 9007199254740992.0);
}

/* Standard normal number i of a stream: Box-Muller on uniforms 2*(i/2) and
 * 2*(i/2)+1; even i takes the cosine branch, odd i the sine branch. */
double orc_normal(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t pair = i >> 1;
  const double u1 = orc_uniform(seed, stream, 2 * pair);
  const double u2 = orc_uniform(seed, stream, 2 * pair + 1);
  const double r = sqrt(-2.0 * log(1.0 - u1));
  const double a = 6.283185307179586 * u2;
  return (i & 1) ? r * sin(a) : r * cos(a);
}

enum {
  ORC_STREAM_CENTER = 1,
  ORC_STREAM_DOCLEN = 2,
  ORC_STREAM_TOKCOMP = 3,
  ORC_STREAM_TOKNOISE = 4,
  ORC_STREAM_QCOMP = 5,
  ORC_STREAM_QNOISE = 6,
  ORC_STREAM_QLEN = 7,
  ORC_STREAM_QPROBE = 8,
};

/* Mixture centres: L2-normalised N(0, I) draws, stored as float. */
void orc_gen_centers(uint64_t seed, uint32_t n_comp, uint32_t dim, float *out) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)n_comp; c++) {
    double buf[1024];
    double nrm = 0.0;
    for (uint32_t d = 0; d < dim; d++) {
      buf[d] = orc_normal(seed, ORC_STREAM_CENTER, (uint64_t)c * dim + d);
      nrm += buf[d] * buf[d];
    }
    nrm = sqrt(nrm);
    for (uint32_t d = 0; d < dim; d++) out[(size_t)c * dim + d] = (float)(buf[d] / nrm);
  }
}

/* Document lengths uniform over {lo..hi}. */
void orc_gen_doclens_uniform(uint64_t seed, uint32_t n_doc, uint32_t lo, uint32_t hi, uint32_t *out) {
  for (uint32_t d = 0; d < n_doc; d++) {
    const double u = orc_uniform(seed, ORC_STREAM_DOCLEN, d);
    uint32_t v = lo + (uint32_t)(u * (double)(hi - lo + 1));
    if (v > hi) v = hi;
    out[d] = v;
  }
}

/* Document lengths round(exp(ln(median) + sigma*z)) clipped to [lo, hi]. */
void orc_gen_doclens_lognormal(uint64_t seed, uint32_t n_doc, double median, double sigma,
                               uint32_t lo, uint32_t hi, uint32_t *out) {
  for (uint32_t d = 0; d < n_doc; d++) {
    const double z = orc_normal(seed, ORC_STREAM_DOCLEN, d);
    double v = floor(exp(log(median) + sigma * z) + 0.5);
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    out[d] = (uint32_t)v;
  }
}

/* Tokens i in [tok_begin, tok_begin+n_tok): pick a component uniformly, add
 * noise_std * N(0, I), L2-normalise.  stream_comp/stream_noise select the
 * corpus (3, 4) or query (5, 6) streams. */
void orc_gen_tokens(uint64_t seed, uint64_t stream_comp, uint64_t stream_noise,
                    const float *centers, uint32_t n_comp, uint32_t dim, double noise_std,
                    uint64_t tok_begin, uint64_t n_tok, float *out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)n_tok; k++) {
    const uint64_t i = tok_begin + (uint64_t)k;
    double buf[1024];
    uint32_t comp = (uint32_t)(orc_uniform(seed, stream_comp, i) * (double)n_comp);
    if (comp >= n_comp) comp = n_comp - 1;
    const float *c = centers + (size_t)comp * dim;
    double nrm = 0.0;
    for (uint32_t d = 0; d < dim; d++) {
      buf[d] = (double)c[d] + noise_std * orc_normal(seed, stream_noise, i * dim + d);
      nrm += buf[d] * buf[d];
    }
    nrm = sqrt(nrm);
    for (uint32_t d = 0; d < dim; d++) out[(size_t)k * dim + d] = (float)(buf[d] / nrm);
  }
}

/* ------------------------------------------------------------------------ */
/* Dot products: the fixed FP32 order shared with the GPU kernels.           */
/* ------------------------------------------------------------------------ */

static inline float orc_dot_f32(const float *a, const float *b, uint32_t dim) {
  float acc = 0.0f;
  for (uint32_t d = 0; d < dim; d++) acc = fmaf(a[d], b[d], acc);
  return acc;
}

float orc_dot(const float *a, const float *b, uint32_t dim) { return orc_dot_f32(a, b, dim); }

/* Dot products of x with NB consecutive columns of a transposed matrix
 * MT[d][col] (row stride `stride`): every column keeps its own sequential FMA
 * chain over d, so the result equals orc_dot_f32 bit for bit; the block form
 * only lets the compiler run the NB chains in SIMD lanes. */
#define ORC_NB 16
static inline void orc_dots_block(const float *x, const float *MT, size_t stride, uint32_t dim,
                                  float *out) {
  /* _mm256_fmadd_ps is fmaf() in each lane: one rounding per step */
  __m256 a0 = _mm256_setzero_ps(), a1 = _mm256_setzero_ps();
  for (uint32_t d = 0; d < dim; d++) {
    const __m256 xd = _mm256_set1_ps(x[d]);
    const float *m = MT + (size_t)d * stride;
    a0 = _mm256_fmadd_ps(xd, _mm256_loadu_ps(m), a0);
    a1 = _mm256_fmadd_ps(xd, _mm256_loadu_ps(m + 8), a1);
  }
  _mm256_storeu_ps(out, a0);
  _mm256_storeu_ps(out + 8, a1);
}

/* Transpose rows [n][dim] into [dim][n_pad] with n_pad a multiple of ORC_NB
 * (padding columns are zero). */
static float *orc_transpose_pad(const float *M, uint32_t n, uint32_t dim, uint32_t *n_pad) {
  const uint32_t np = (n + ORC_NB - 1) / ORC_NB * ORC_NB;
  float *T = (float *)calloc((size_t)np * dim, sizeof(float));
  for (uint32_t r = 0; r < n; r++)
    for (uint32_t d = 0; d < dim; d++) T[(size_t)d * np + r] = M[(size_t)r * dim + d];
  *n_pad = np;
  return T;
}


/* Pack rows [n][dim] into column blocks: block b holds rows b*16..b*16+15 as
 * [dim][16] (zero padded), contiguous, so one block stays in L1. */
static float *orc_pack_blocks(const float *M, uint32_t n, uint32_t dim, uint32_t *n_blk) {
  const uint32_t nb = (n + ORC_NB - 1) / ORC_NB;
  float *P = (float *)calloc((size_t)nb * dim * ORC_NB, sizeof(float));
  for (uint32_t r = 0; r < n; r++)
    for (uint32_t d = 0; d < dim; d++) P[((size_t)(r / ORC_NB) * dim + d) * ORC_NB + r % ORC_NB] = M[(size_t)r * dim + d];
  *n_blk = nb;
  return P;
}

/* Dots of 4 rows x0..x3 with the 16 columns of one packed block; each of the
 * 64 outputs is its own sequential FMA chain over d (same bits as orc_dot_f32). */
static inline void orc_dots_4x16(const float *x0, const float *x1, const float *x2, const float *x3,
                                 const float *blk, uint32_t dim, float out[4][ORC_NB]) {
  __m256 a00 = _mm256_setzero_ps(), a01 = _mm256_setzero_ps();
  __m256 a10 = _mm256_setzero_ps(), a11 = _mm256_setzero_ps();
  __m256 a20 = _mm256_setzero_ps(), a21 = _mm256_setzero_ps();
  __m256 a30 = _mm256_setzero_ps(), a31 = _mm256_setzero_ps();
  for (uint32_t d = 0; d < dim; d++) {
    const __m256 m0 = _mm256_loadu_ps(blk + (size_t)d * ORC_NB);
    const __m256 m1 = _mm256_loadu_ps(blk + (size_t)d * ORC_NB + 8);
    __m256 v = _mm256_set1_ps(x0[d]);
    a00 = _mm256_fmadd_ps(v, m0, a00); a01 = _mm256_fmadd_ps(v, m1, a01);
    v = _mm256_set1_ps(x1[d]);
    a10 = _mm256_fmadd_ps(v, m0, a10); a11 = _mm256_fmadd_ps(v, m1, a11);
    v = _mm256_set1_ps(x2[d]);
    a20 = _mm256_fmadd_ps(v, m0, a20); a21 = _mm256_fmadd_ps(v, m1, a21);
    v = _mm256_set1_ps(x3[d]);
    a30 = _mm256_fmadd_ps(v, m0, a30); a31 = _mm256_fmadd_ps(v, m1, a31);
  }
  _mm256_storeu_ps(out[0], a00); _mm256_storeu_ps(out[0] + 8, a01);
  _mm256_storeu_ps(out[1], a10); _mm256_storeu_ps(out[1] + 8, a11);
  _mm256_storeu_ps(out[2], a20); _mm256_storeu_ps(out[2] + 8, a21);
  _mm256_storeu_ps(out[3], a30); _mm256_storeu_ps(out[3] + 8, a31);
}

/* argmax over columns of (dot(x_i, C_c) - bias[c]) for every row i; lowest
 * column wins ties; bias may be NULL. */
static void orc_argmax_rows(const float *x, uint64_t n, uint32_t dim, const float *C, uint32_t n_c,
                            const float *bias, uint32_t *out) {
  uint32_t nb;
  float *Pk = orc_pack_blocks(C, n_c, dim, &nb);
  const int64_t n4 = (int64_t)((n + 3) / 4);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n4; g++) {
    const float *xr[4];
    for (int r = 0; r < 4; r++) {
      uint64_t i = (uint64_t)g * 4 + r;
      if (i >= n) i = n - 1;
      xr[r] = x + i * dim;
    }
    float best[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    uint32_t arg[4] = {0, 0, 0, 0};
    float s[4][ORC_NB];
    for (uint32_t b = 0; b < nb; b++) {
      orc_dots_4x16(xr[0], xr[1], xr[2], xr[3], Pk + (size_t)b * dim * ORC_NB, dim, s);
      for (uint32_t j = 0; j < ORC_NB; j++) {
        const uint32_t c = b * ORC_NB + j;
        if (c >= n_c) break;
        for (int r = 0; r < 4; r++) {
          const float v = bias ? s[r][j] - bias[c] : s[r][j];
          if (v > best[r]) { best[r] = v; arg[r] = c; }
        }
      }
    }
    for (int r = 0; r < 4; r++) {
      const uint64_t i = (uint64_t)g * 4 + r;
      if (i < n) out[i] = arg[r];
    }
  }
  free(Pk);
}

/* ------------------------------------------------------------------------ */
/* k-means (index build) and code assignment.                                */
/* ------------------------------------------------------------------------ */

/* argmax_c <x, C_c> with the sequential FMA dot product; lowest index wins ties.
 * Reference: script/evaluation/technique/vq.py compress_into_codes (argmax IP). */
void orc_assign_ip(const float *x, uint64_t n, uint32_t dim, const float *C, uint32_t n_c,
                   uint32_t *codes) {
  orc_argmax_rows(x, n, dim, C, n_c, NULL, codes);
}

/* argmax_c (<x, C_c> - bias[c]) in FP32 (sequential FMA dot, then one
 * subtraction), lowest index on ties: the assignment step of the Lloyd
 * iterations below with bias = |C_c|^2 / 2 (argmin L2). */
void orc_assign_bias(const float *x, uint64_t n, uint32_t dim, const float *C, uint32_t n_c,
                     const float *bias, uint32_t *codes) {
  orc_argmax_rows(x, n, dim, C, n_c, bias, codes);
}

/* Lloyd's k-means: initial centroid j is point floor(j*n/n_c); each iteration
 * assigns by argmax (<x,c> - |c|^2/2) (= argmin L2, lowest index on ties, FP32
 * sequential FMA), then recomputes means in double; an empty cluster keeps its
 * previous centroid.  Centroids are L2-normalised at the end
 * (script/evaluation/technique/vq_sq_pack.py:76-88 normalises faiss output). */
void orc_kmeans(const float *x, uint64_t n, uint32_t dim, uint32_t n_c, uint32_t n_iter,
                float *centroids) {
  for (uint32_t j = 0; j < n_c; j++) {
    const uint64_t idx = ((uint64_t)j * n) / n_c;
    memcpy(centroids + (size_t)j * dim, x + idx * dim, sizeof(float) * dim);
  }
  uint32_t *assign = (uint32_t *)malloc(sizeof(uint32_t) * n);
  float *half_norm = (float *)malloc(sizeof(float) * n_c);
  double *sums = (double *)malloc(sizeof(double) * (size_t)n_c * dim);
  uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * n_c);
  for (uint32_t it = 0; it < n_iter; it++) {
    for (uint32_t c = 0; c < n_c; c++)
      half_norm[c] = 0.5f * orc_dot_f32(centroids + (size_t)c * dim, centroids + (size_t)c * dim, dim);
    orc_argmax_rows(x, n, dim, centroids, n_c, half_norm, assign);
    memset(sums, 0, sizeof(double) * (size_t)n_c * dim);
    memset(cnt, 0, sizeof(uint64_t) * n_c);
    for (uint64_t i = 0; i < n; i++) {
      const uint32_t a = assign[i];
      cnt[a]++;
      for (uint32_t d = 0; d < dim; d++) sums[(size_t)a * dim + d] += (double)x[i * dim + d];
    }
    for (uint32_t c = 0; c < n_c; c++) {
      if (cnt[c] == 0) continue;
      for (uint32_t d = 0; d < dim; d++)
        centroids[(size_t)c * dim + d] = (float)(sums[(size_t)c * dim + d] / (double)cnt[c]);
    }
  }
  for (uint32_t c = 0; c < n_c; c++) {
    double nrm = 0.0;
    for (uint32_t d = 0; d < dim; d++) nrm += (double)centroids[(size_t)c * dim + d] * centroids[(size_t)c * dim + d];
    nrm = sqrt(nrm);
    if (nrm > 0.0)
      for (uint32_t d = 0; d < dim; d++) centroids[(size_t)c * dim + d] = (float)(centroids[(size_t)c * dim + d] / nrm);
  }
  free(assign); free(half_norm); free(sums); free(cnt);
}

/* ------------------------------------------------------------------------ */
/* Residual codec.                                                           */
/* ------------------------------------------------------------------------ */

static int orc_cmp_float(const void *a, const void *b) {
  const float x = *(const float *)a, y = *(const float *)b;
  return (x > y) - (x < y);
}

/* e[i][d] = x[i][d] - C[code_i][d] in FP32 (ResidualScalarQuantizationPack.hpp:31-45). */
void orc_residuals(const float *x, const uint32_t *codes, uint64_t n, uint32_t dim,
                   const float *C, float *e) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++)
    for (uint32_t d = 0; d < dim; d++)
      e[(size_t)i * dim + d] = x[(size_t)i * dim + d] - C[(size_t)codes[i] * dim + d];
}

/* Pooled quantiles over every residual value, exactly as the reference:
 * cutoff i (i=1..2^b-1) is the element of rank floor(i/2^b * N) and weight i
 * (i=0..2^b-1) the element of rank floor((i+0.5)/2^b * N) of the sorted values
 * (ResidualScalarQuantizationPack.hpp:47-79; nth_element returns that rank). */
void orc_quantize_pooled(const float *e, uint64_t n_val, uint32_t n_bit, float *cutoffs, float *weights) {
  float *s = (float *)malloc(sizeof(float) * n_val);
  memcpy(s, e, sizeof(float) * n_val);
  qsort(s, n_val, sizeof(float), orc_cmp_float);
  const uint32_t nq = 1u << n_bit;
  for (uint32_t i = 1; i < nq; i++) {
    const double q = (double)i / nq;
    cutoffs[i - 1] = s[(uint64_t)floor(q * (double)n_val)];
  }
  for (uint32_t i = 0; i < nq; i++) {
    const double q = (double)i / nq + 0.5 / nq;
    weights[i] = s[(uint64_t)floor(q * (double)n_val)];
  }
  free(s);
}

/* Bucket of a residual value: index of the first cutoff strictly greater than
 * e, i.e. the number of cutoffs <= e (ResidualScalarQuantizationPack.hpp:186-191
 * std::upper_bound; ResidualScalarQuantizationPackGPU.hpp:66-71). */
static inline uint32_t orc_bucket(const float *cut, uint32_t n_cut, float e) {
  uint32_t b = n_cut;
  for (uint32_t k = 0; k < n_cut; k++)
    if (cut[k] > e) { b = k; break; }
  return b;
}

uint32_t orc_bucket_of(const float *cut, uint32_t n_cut, float e) { return orc_bucket(cut, n_cut, e); }

/* Per-dimension quantiles (the synthetic configs' codec): for each dimension,
 * cutoffs at ranks floor(i/2^b * n) of that dimension's sorted residuals and
 * weights equal to the mean of the residuals that fall in each bucket (double
 * accumulation, rounded to float).  An empty bucket gets the midpoint of its
 * neighbouring cutoffs. */
void orc_quantize_perdim(const float *e, uint64_t n, uint32_t dim, uint32_t n_bit,
                         float *cutoffs /* [dim][2^b-1] */, float *weights /* [dim][2^b] */) {
  const uint32_t nq = 1u << n_bit, ncut = nq - 1;
#pragma omp parallel for schedule(dynamic)
  for (int64_t d = 0; d < (int64_t)dim; d++) {
    float *s = (float *)malloc(sizeof(float) * n);
    for (uint64_t i = 0; i < n; i++) s[i] = e[i * dim + d];
    qsort(s, n, sizeof(float), orc_cmp_float);
    float *cut = cutoffs + (size_t)d * ncut;
    for (uint32_t i = 1; i < nq; i++) cut[i - 1] = s[(uint64_t)floor((double)i / nq * (double)n)];
    double sum[256] = {0};
    uint64_t cnt[256] = {0};
    for (uint64_t i = 0; i < n; i++) {
      const float v = e[i * dim + d];
      const uint32_t b = orc_bucket(cut, ncut, v);
      sum[b] += v;
      cnt[b]++;
    }
    for (uint32_t b = 0; b < nq; b++) {
      float w;
      if (cnt[b]) w = (float)(sum[b] / (double)cnt[b]);
      else if (b == 0) w = cut[0];
      else if (b == ncut) w = cut[ncut - 1];
      else w = 0.5f * (cut[b - 1] + cut[b]);
      weights[(size_t)d * nq + b] = w;
    }
    free(s);
  }
}

/* Pack buckets MSB-first: value k of a byte sits at bits [8-n_bit(k+1), 8-n_bit*k)
 * (ResidualScalarQuantizationPack.hpp:226-236).  cut_stride = 0 means one
 * pooled cutoff array for every dimension, otherwise cutoffs[d*cut_stride..]. */
void orc_compress(const float *x, const uint32_t *codes, uint64_t n, uint32_t dim, const float *C,
                  const float *cutoffs, uint32_t cut_stride, uint32_t n_bit, uint8_t *out) {
  const uint32_t vpb = 8 / n_bit, ncut = (1u << n_bit) - 1;
  const uint32_t bpv = (dim + vpb - 1) / vpb;
  const uint8_t mask = (uint8_t)((1u << n_bit) - 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *xi = x + (size_t)i * dim;
    const float *ci = C + (size_t)codes[i] * dim;
    for (uint32_t p = 0; p < bpv; p++) {
      uint8_t byte = 0;
      for (uint32_t v = 0; v < vpb; v++) {
        const uint32_t d = p * vpb + v;
        uint32_t b = 0;
        if (d < dim) {
          const float err = xi[d] - ci[d];
          b = orc_bucket(cutoffs + (size_t)(cut_stride ? d * cut_stride : 0), ncut, err);
        }
        byte = (uint8_t)((byte << n_bit) | (b & mask));
      }
      out[(size_t)i * bpv + p] = byte;
    }
  }
}

/* Decode table W[d][b]: weights given as one pooled array (w_stride = 0) or per
 * dimension (w_stride = 2^b).  DataStructPP.hpp:160-170 builds the same
 * mapping as a 256 x (8/b) table indexed by the packed byte. */
void orc_decode_table(const float *weights, uint32_t w_stride, uint32_t dim, uint32_t n_bit, float *W) {
  const uint32_t nq = 1u << n_bit;
  for (uint32_t d = 0; d < dim; d++)
    for (uint32_t b = 0; b < nq; b++) W[d * nq + b] = weights[(w_stride ? d * w_stride : 0) + b];
}

/* Bucket of dimension d in a packed residual row (MSB-first). */
static inline uint32_t orc_unpack(const uint8_t *row, uint32_t d, uint32_t n_bit) {
  const uint32_t vpb = 8 / n_bit;
  const uint32_t shift = 8 - n_bit * (d % vpb + 1);
  return (row[d / vpb] >> shift) & ((1u << n_bit) - 1);
}

/* x[d] = C[code][d] + W[d][bucket] in FP32, no normalisation
 * (RerankPP.hpp:100-112; ResidualScalarQuantizationPack.hpp:330-378). */
void orc_decompress_tokens(const uint64_t *tok_ids, uint64_t n, const uint32_t *codes,
                           const uint8_t *residual, const float *C, const float *W,
                           uint32_t dim, uint32_t n_bit, float *out) {
  const uint32_t bpv = (dim + 8 / n_bit - 1) / (8 / n_bit), nq = 1u << n_bit;
  for (uint64_t k = 0; k < n; k++) {
    const uint64_t i = tok_ids[k];
    const uint8_t *row = residual + i * bpv;
    const float *c = C + (size_t)codes[i] * dim;
    for (uint32_t d = 0; d < dim; d++) out[k * dim + d] = c[d] + W[d * nq + orc_unpack(row, d, n_bit)];
  }
}

/* ------------------------------------------------------------------------ */
/* IVF index (IVFIndex.hpp:92-183).                                          */
/* ------------------------------------------------------------------------ */

/* Count pass: sizes[c] = number of docs whose token set contains centroid c. */
static void orc_doc_unique(const uint32_t *codes, uint64_t off, uint32_t len, uint32_t *scratch,
                           uint32_t *n_out) {
  /* insertion into a small sorted array (docs are <= a few hundred tokens) */
  uint32_t m = 0;
  for (uint32_t k = 0; k < len; k++) {
    const uint32_t c = codes[off + k];
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (scratch[mid] < c) lo = mid + 1; else hi = mid;
    }
    if (lo < m && scratch[lo] == c) continue;
    memmove(scratch + lo + 1, scratch + lo, sizeof(uint32_t) * (m - lo));
    scratch[lo] = c;
    m++;
  }
  *n_out = m;
}

/* Returns n_ele; ids may be NULL for a size-only call.  Lists hold doc ids in
 * ascending order; offsets are exclusive prefix sums of sizes. */
uint64_t orc_ivf_build(const uint32_t *codes, const uint32_t *doclens, uint32_t n_doc, uint32_t n_c,
                       uint64_t *offsets, uint32_t *sizes, uint32_t *ids) {
  uint32_t *scratch = (uint32_t *)malloc(sizeof(uint32_t) * 65536);
  memset(sizes, 0, sizeof(uint32_t) * n_c);
  uint64_t off = 0;
  for (uint32_t d = 0; d < n_doc; d++) {
    uint32_t m;
    orc_doc_unique(codes, off, doclens[d], scratch, &m);
    for (uint32_t k = 0; k < m; k++) sizes[scratch[k]]++;
    off += doclens[d];
  }
  uint64_t n_ele = 0;
  for (uint32_t c = 0; c < n_c; c++) { offsets[c] = n_ele; n_ele += sizes[c]; }
  if (ids) {
    uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * n_c);
    memcpy(fill, offsets, sizeof(uint64_t) * n_c);
    off = 0;
    for (uint32_t d = 0; d < n_doc; d++) {
      uint32_t m;
      orc_doc_unique(codes, off, doclens[d], scratch, &m);
      for (uint32_t k = 0; k < m; k++) ids[fill[scratch[k]]++] = d;
      off += doclens[d];
    }
    free(fill);
  }
  free(scratch);
  return n_ele;
}

/* ------------------------------------------------------------------------ */
/* Search path (IGP-PP).                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  float s;
  uint32_t id;
} orc_pair;

static inline int orc_better(float s1, uint32_t i1, float s2, uint32_t i2) {
  return (s1 > s2) || (s1 == s2 && i1 < i2);
}

static int orc_cmp_pair(const void *a, const void *b) {
  const orc_pair *x = (const orc_pair *)a, *y = (const orc_pair *)b;
  if (orc_better(x->s, x->id, y->s, y->id)) return -1;
  if (orc_better(y->s, y->id, x->s, x->id)) return 1;
  return 0;
}

/* Centroid scoring and top-P for each of n_row query tokens:
 * S'[t][c] = dot(Q_t, C_c) + 1.0f (GenerateVQScoreCandidatePP.hpp:137-144), then
 * the P best by (S' desc, c asc) (:166-174), written sorted. */
void orc_centroid_topk(const float *Q, uint64_t n_row, const float *C, uint32_t n_c, uint32_t dim,
                       uint32_t P, uint32_t *top_ids, float *top_s) {
  uint32_t np;
  float *CT = orc_transpose_pad(C, n_c, dim, &np);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t r = 0; r < (int64_t)n_row; r++) {
    orc_pair *best = (orc_pair *)malloc(sizeof(orc_pair) * (P + 1));
    uint32_t m = 0;
    const float *q = Q + (size_t)r * dim;
    float blk[ORC_NB];
    for (uint32_t c = 0; c < n_c; c++) {
      if (c % ORC_NB == 0) orc_dots_block(q, CT + c, np, dim, blk);
      const float s = blk[c % ORC_NB] + 1.0f;
      if (m == P && !orc_better(s, c, best[P - 1].s, best[P - 1].id)) continue;
      uint32_t pos = (m < P) ? m : P - 1;
      while (pos > 0 && orc_better(s, c, best[pos - 1].s, best[pos - 1].id)) {
        best[pos] = best[pos - 1];
        pos--;
      }
      best[pos].s = s;
      best[pos].id = c;
      if (m < P) m++;
    }
    for (uint32_t j = 0; j < P; j++) {
      top_ids[(size_t)r * P + j] = j < m ? best[j].id : UINT32_MAX;
      top_s[(size_t)r * P + j] = j < m ? best[j].s : -INFINITY;
    }
    free(best);
  }
  free(CT);
}

static inline int64_t orc_fx(float s) { return (int64_t)((double)s * ORC_FX_SCALE); }

/* Filter of one query (FilterPP.hpp:267-398):
 *   candidates = union over t < T, j < P of IVF[top_ids[t][j]];
 *   m[d][t]    = max(0, max_j { top_s[t][j] : d in IVF[top_ids[t][j]] }), 0 if
 *                none: the maxima start zero-filled (memset at :257-259) and
 *                take atomicMax of every hit (max_qvec_doc_score :136-141);
 *   F[d]       = sum_t m[d][t] (reduce :366-379), here as the rounded exact
 *                fixed-point sum (see file header).
 * basic != 0 restates the basic variant instead (igp_gpu_basic/Filter.hpp:44-59):
 * m[d][t] is the score of the first probe rank whose list holds d, as it is
 * (it may be negative), and 0 if none.
 * Writes every candidate (doc ascending) to cand_ids/cand_F and returns the
 * count.  top_ids/top_s rows are Pmax wide and sorted (score desc). */
uint32_t orc_filter_all(const uint32_t *top_ids, const float *top_s, uint32_t T, uint32_t Pmax,
                        uint32_t P, const uint64_t *ivf_off, const uint32_t *ivf_size,
                        const uint32_t *ivf_ids, uint32_t n_doc, int basic, uint32_t *cand_ids,
                        float *cand_F) {
  uint32_t *mask = (uint32_t *)calloc(n_doc, sizeof(uint32_t));
  int64_t *acc = (int64_t *)calloc(n_doc, sizeof(int64_t));
  uint8_t *seen = (uint8_t *)calloc(n_doc, 1);
  /* probe ranks in order: the first list of token t that contains d carries
   * the largest score of t for d, because top_s[t][.] is sorted descending */
  for (uint32_t j = 0; j < P; j++)
    for (uint32_t t = 0; t < T; t++) {
      const uint32_t c = top_ids[(size_t)t * Pmax + j];
      if (c == UINT32_MAX) continue;
      const float s = top_s[(size_t)t * Pmax + j];
      const int64_t fx = (!basic && orc_fx(s) < 0) ? 0 : orc_fx(s);
      const uint32_t *lst = ivf_ids + ivf_off[c];
      for (uint32_t e = 0; e < ivf_size[c]; e++) {
        const uint32_t d = lst[e];
        seen[d] = 1;
        const uint32_t bit = 1u << t;
        if (mask[d] & bit) continue;
        mask[d] |= bit;
        acc[d] += fx;
      }
    }
  uint32_t n = 0;
  for (uint32_t d = 0; d < n_doc; d++)
    if (seen[d]) {
      cand_ids[n] = d;
      cand_F[n] = (float)((double)acc[d] * ORC_FX_INV);
      n++;
    }
  free(mask); free(acc); free(seen);
  return n;
}

/* Top-R of the candidates by (F desc, doc asc) (FilterPP.hpp:386-394).
 * Returns min(R, n). */
uint32_t orc_select_top(const uint32_t *ids, const float *s, uint32_t n, uint32_t R,
                        uint32_t *out_ids, float *out_s) {
  orc_pair *p = (orc_pair *)malloc(sizeof(orc_pair) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) { p[i].s = s[i]; p[i].id = ids[i]; }
  qsort(p, n, sizeof(orc_pair), orc_cmp_pair);
  const uint32_t m = n < R ? n : R;
  for (uint32_t i = 0; i < m; i++) { out_ids[i] = p[i].id; out_s[i] = p[i].s; }
  free(p);
  return m;
}

/* MaxSim of one query against n docs (RerankPP.hpp:368-475):
 * s[t][i] = dot(Q_t, x_i) with x_i decompressed in FP32; M[t] = max_i s[t][i]
 * (identity -FLT_MAX, RerankPP.hpp:141-157), or max(0, .) when clip_zero
 * (the basic variant's zero padding, igp_gpu_basic/Rerank.hpp:293-306);
 * score = ((0 + M[0]) + M[1]) + ... over t < T in FP32. */
void orc_rerank(const float *Q, uint32_t T, uint32_t dim, const uint32_t *docs, uint32_t n,
                const uint32_t *doclens, const uint64_t *docoff, const uint32_t *codes,
                const uint8_t *residual, const float *C, const float *W, uint32_t n_bit,
                int clip_zero, float *scores) {
  const uint32_t nq = 1u << n_bit, bpv = (dim + 8 / n_bit - 1) / (8 / n_bit);
  float *x = (float *)malloc(sizeof(float) * dim);
  float *M = (float *)malloc(sizeof(float) * T);
  uint32_t tp;
  float *QT = orc_transpose_pad(Q, T, dim, &tp);
  float *sv = (float *)malloc(sizeof(float) * tp);
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t d = docs[k];
    if (d == UINT32_MAX) { scores[k] = -INFINITY; continue; }
    for (uint32_t t = 0; t < T; t++) M[t] = clip_zero ? 0.0f : -FLT_MAX;
    for (uint32_t i = 0; i < doclens[d]; i++) {
      const uint64_t tok = docoff[d] + i;
      const float *c = C + (size_t)codes[tok] * dim;
      const uint8_t *row = residual + tok * bpv;
      for (uint32_t e = 0; e < dim; e++) x[e] = c[e] + W[e * nq + orc_unpack(row, e, n_bit)];
      for (uint32_t t0 = 0; t0 < tp; t0 += ORC_NB) orc_dots_block(x, QT + t0, tp, dim, sv + t0);
      for (uint32_t t = 0; t < T; t++) M[t] = fmaxf(M[t], sv[t]);
    }
    float sc = 0.0f;
    for (uint32_t t = 0; t < T; t++) sc += M[t];
    scores[k] = sc;
  }
  free(x); free(M); free(QT); free(sv);
}

typedef struct {
  uint32_t n_doc, n_c, dim, n_bit;
  uint64_t n_tok;
  const uint32_t *doclens;
  const uint64_t *docoff;
  const uint32_t *codes;
  const uint8_t *residual;
  const float *C;
  const float *W; /* [dim][2^b] */
  const uint64_t *ivf_off;
  const uint32_t *ivf_size;
  const uint32_t *ivf_ids;
} orc_index;

/* Whole search of one query: centroid top-P, filter, top-R, rerank, top-k
 * (IGPGPURetrievalPP.hpp:114-176).  Missing results are (UINT32_MAX, -inf). */
static void orc_search_one(const orc_index *ix, const float *Q, uint32_t T, uint32_t P, uint32_t R,
                           uint32_t k, int clip_zero, uint32_t *out_ids, float *out_s,
                           uint32_t *cand_buf, float *candF_buf, uint32_t *top_ids, float *top_s,
                           uint32_t *sel_ids, float *sel_F, float *sel_score) {
  orc_centroid_topk(Q, T, ix->C, ix->n_c, ix->dim, P, top_ids, top_s);
  const uint32_t nc = orc_filter_all(top_ids, top_s, T, P, P, ix->ivf_off, ix->ivf_size, ix->ivf_ids,
                                     ix->n_doc, clip_zero, cand_buf, candF_buf);
  const uint32_t nr = orc_select_top(cand_buf, candF_buf, nc, R, sel_ids, sel_F);
  orc_rerank(Q, T, ix->dim, sel_ids, nr, ix->doclens, ix->docoff, ix->codes, ix->residual, ix->C,
             ix->W, ix->n_bit, clip_zero, sel_score);
  const uint32_t nk = orc_select_top(sel_ids, sel_score, nr, k, out_ids, out_s);
  for (uint32_t i = nk; i < k; i++) { out_ids[i] = UINT32_MAX; out_s[i] = -INFINITY; }
}

/* Batched search over n_q queries stored [n_q][Tmax][dim]; q_ntok (may be NULL:
 * all Tmax) and q_nprobe (may be NULL: P) give per-query token counts and
 * probe counts.  OpenMP over queries with n_threads threads (0 = default). */
void orc_search_batch(uint32_t n_doc, uint32_t n_c, uint32_t dim, uint32_t n_bit, uint64_t n_tok,
                      const uint32_t *doclens, const uint64_t *docoff, const uint32_t *codes,
                      const uint8_t *residual, const float *C, const float *W,
                      const uint64_t *ivf_off, const uint32_t *ivf_size, const uint32_t *ivf_ids,
                      const float *Qall, uint32_t n_q, uint32_t Tmax, const uint32_t *q_ntok,
                      const uint32_t *q_nprobe, uint32_t P, uint32_t R, uint32_t k, int clip_zero,
                      int n_threads, uint32_t *out_ids, float *out_s, double *q_seconds) {
  orc_index ix = {n_doc, n_c, dim, n_bit, n_tok, doclens, docoff, codes, residual, C, W,
                  ivf_off, ivf_size, ivf_ids};
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  uint32_t Pmax = P;
  if (q_nprobe)
    for (uint32_t q = 0; q < n_q; q++) if (q_nprobe[q] > Pmax) Pmax = q_nprobe[q];
#pragma omp parallel
  {
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * n_doc);
    float *candF = (float *)malloc(sizeof(float) * n_doc);
    uint32_t *top_ids = (uint32_t *)malloc(sizeof(uint32_t) * Tmax * Pmax);
    float *top_s = (float *)malloc(sizeof(float) * Tmax * Pmax);
    uint32_t *sel = (uint32_t *)malloc(sizeof(uint32_t) * (R + 1));
    float *selF = (float *)malloc(sizeof(float) * (R + 1));
    float *selS = (float *)malloc(sizeof(float) * (R + 1));
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)n_q; q++) {
#ifdef _OPENMP
      const double t0 = omp_get_wtime();
#endif
      const uint32_t T = q_ntok ? q_ntok[q] : Tmax;
      const uint32_t Pq = q_nprobe ? q_nprobe[q] : P;
      orc_search_one(&ix, Qall + (size_t)q * Tmax * dim, T, Pq, R, k, clip_zero,
                     out_ids + (size_t)q * k, out_s + (size_t)q * k, cand, candF, top_ids, top_s,
                     sel, selF, selS);
#ifdef _OPENMP
      if (q_seconds) q_seconds[q] = omp_get_wtime() - t0;
#endif
    }
    free(cand); free(candF); free(top_ids); free(top_s); free(sel); free(selF); free(selS);
  }
}

/* ------------------------------------------------------------------------ */
/* FP64 exhaustive MaxSim (the accuracy oracle; Distance.hpp:200-218).       */
/* ------------------------------------------------------------------------ */

/* For each query: score[d] = sum_t max_i <Q_t, X_i> in double over the
 * embeddings X [n_tok][dim] (float), top-k by (score desc, doc asc). */
void orc_exhaustive_f64(const float *Qall, uint32_t n_q, uint32_t Tmax, const uint32_t *q_ntok,
                        const float *X, const uint32_t *doclens, const uint64_t *docoff,
                        uint32_t n_doc, uint32_t dim, uint32_t k, int n_threads,
                        uint32_t *out_ids, double *out_s) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
#pragma omp parallel
  {
    double *sc = (double *)malloc(sizeof(double) * n_doc);
    double *M = (double *)malloc(sizeof(double) * Tmax);
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)n_q; q++) {
      const uint32_t T = q_ntok ? q_ntok[q] : Tmax;
      const float *Q = Qall + (size_t)q * Tmax * dim;
      for (uint32_t d = 0; d < n_doc; d++) {
        for (uint32_t t = 0; t < T; t++) M[t] = -INFINITY;
        for (uint32_t i = 0; i < doclens[d]; i++) {
          const float *x = X + (docoff[d] + i) * dim;
          for (uint32_t t = 0; t < T; t++) {
            double s = 0.0;
            for (uint32_t e = 0; e < dim; e++) s += (double)Q[(size_t)t * dim + e] * (double)x[e];
            if (s > M[t]) M[t] = s;
          }
        }
        double tot = 0.0;
        for (uint32_t t = 0; t < T; t++) tot += M[t];
        sc[d] = tot;
      }
      /* top-k by partial selection */
      for (uint32_t r = 0; r < k; r++) {
        double bs = -INFINITY;
        uint32_t bi = UINT32_MAX;
        for (uint32_t d = 0; d < n_doc; d++) {
          if (sc[d] > bs || (sc[d] == bs && d < bi)) { bs = sc[d]; bi = d; }
        }
        out_ids[(size_t)q * k + r] = bi;
        out_s[(size_t)q * k + r] = bs;
        if (bi != UINT32_MAX) sc[bi] = -INFINITY;
      }
    }
    free(sc); free(M);
  }
}

/* ------------------------------------------------------------------------ */
/* All-pairs FP32 MaxSim of the exhaustive module (BruteForceProgressive):    */
/* src/include/compute_score/GPUComputeScore.hpp:196-246 (SGEMM of the query  */
/* tokens against every item token, select_max_functor :63-89 per item and    */
/* query token, thrust::plus over query tokens from 0.0f).                    */
/* ------------------------------------------------------------------------ */

/* scores[q][i] = ((0 + M[0]) + M[1]) + ... over t < T, M[t] = max over item
 * i's tokens x of dot(Q_q,t, x), each dot one FMA chain over d.  Items are
 * laid out back to back in X [sum lens][dim]; every item has >= 1 token. */
void orc_maxsim_scores_f32(const float *Qall, uint32_t n_q, uint32_t T, uint32_t dim, const float *X,
                           const uint32_t *lens, uint32_t n_item, int n_threads, float *scores) {
  uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (n_item + 1));
  off[0] = 0;
  for (uint32_t i = 0; i < n_item; i++) off[i + 1] = off[i] + lens[i];
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
#pragma omp parallel
  {
    float *M = (float *)malloc(sizeof(float) * (T ? T : 1));
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)n_q; q++) {
      const float *Q = Qall + (size_t)q * T * dim;
      for (uint32_t i = 0; i < n_item; i++) {
        for (uint32_t t = 0; t < T; t++) M[t] = -INFINITY;
        for (uint64_t r = off[i]; r < off[i + 1]; r++) {
          const float *x = X + r * dim;
          for (uint32_t t = 0; t < T; t++) {
            const float s = orc_dot_f32(Q + (size_t)t * dim, x, dim);
            if (s > M[t]) M[t] = s;
          }
        }
        float sc = 0.0f;
        for (uint32_t t = 0; t < T; t++) sc += M[t];
        scores[(size_t)q * n_item + i] = sc;
      }
    }
    free(M);
  }
  free(off);
}
