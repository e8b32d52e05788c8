/*
 * plaid_oracle.c -- CPU oracle for the PLAID (ColBERTv2) search path and its
 * residual codec ops.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path never links or calls it.
 *
 * It restates in plain C what the reference's PLAID side tree computes:
 *
 *   decompress (GPU, FP16)  baseline/ColBERT/colbert/indexing/codecs/decompress_residuals.cu:8-75
 *   packbits (GPU)          baseline/ColBERT/colbert/indexing/codecs/packbits.cu:10-57
 *   binarize                baseline/ColBERT/colbert/indexing/codecs/residual.py:189-224
 *   decompress (CPU, FP32)  baseline/ColBERT/colbert/search/decompress_residuals.cpp:27-155
 *   two-stage pid filter    baseline/ColBERT/colbert/search/filter_pids.cpp:30-193
 *   segmented MaxSim        baseline/ColBERT/colbert/modeling/segmented_maxsim.cpp:22-93
 *   candidate generation    baseline/ColBERT/colbert/search/candidate_generation.py:13-66
 *   score_pids, CPU branch  baseline/ColBERT/colbert/search/index_storage.py:101-198
 *   normalisation           residual.py:273-274 (torch.nn.functional.normalize, p=2, eps=1e-12)
 *   score_pids, GPU branch  index_storage.py:112-153,160-161,185-187 with colbert_score_reduce /
 *                           colbert_score_packed (modeling/colbert.py:133-155,181-205) and the
 *                           half-precision decompress + normalise of residual.py:255-274
 *
 * Orders the reference leaves to MKL / torch reductions are fixed here, and the
 * GPU kernels follow the same definition:
 *   - every FP32 dot product and squared norm is a sequential fused
 *     multiply-add chain over d = 0..dim-1 from 0.0f;
 *   - the filter score is the left-to-right FP32 sum over query tokens from
 *     0.0f (the loop of filter_pids.cpp:71-76, kept as is);
 *   - the MaxSim score is the left-to-right FP32 sum over query tokens of
 *     max(0, max_i s) (segmented_maxsim.cpp:62 zero-initialised maxima);
 *   - the filter keeps the best pids by (score, pid) descending, the order of
 *     std::priority_queue<std::pair<float, int>> (filter_pids.cpp:27,136-145);
 *   - the final ranking orders by (score descending, pid ascending) and
 *     top-ncells per token by (score descending, centroid ascending), which is
 *     argmax's first-maximum rule for ncells = 1.
 * The GPU branch computes in float16: every matmul output and every reduction
 * result is rounded to half once (FP32 accumulation inside, in the orders
 * above), and torch.topk's open tie order is fixed as in the CPU branch.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <immintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* IEEE binary16 conversions (exact widening; round-to-nearest-even         */
/* narrowing with subnormals, overflow to infinity, NaN kept quiet).         */
/* ------------------------------------------------------------------------ */

float orc_f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t exp = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
  uint32_t bits;
  if (exp == 0) {
    if (man == 0) {
      bits = sign;
    } else { /* subnormal: value = man * 2^-24 */
      float v = (float)man * 0x1p-24f;
      uint32_t vb;
      memcpy(&vb, &v, 4);
      bits = sign | vb;
    }
  } else if (exp == 31) {
    bits = sign | 0x7F800000u | (man << 13) | (man ? 0x400000u : 0u);
  } else {
    bits = sign | ((exp + 112u) << 23) | (man << 13);
  }
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

uint16_t orc_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax > 0x7F800000u) return sign | 0x7E00u;           /* NaN */
  if (ax >= 0x477FF000u) return sign | 0x7C00u;          /* >= 65520: rounds to inf */
  if (ax < 0x38800000u) {                                /* below 2^-14: subnormal or 0 */
    /* value / 2^-24 rounded to nearest even integer, exact in double */
    float a;
    memcpy(&a, &ax, 4);
    const double m = (double)a * 16777216.0; /* 2^24 */
    double r = nearbyint(m);                 /* default mode: ties to even */
    return sign | (uint16_t)r;               /* r <= 1024 -> 0x0400 is the smallest normal */
  }
  /* normal range: keep 10 mantissa bits, round to nearest even */
  const uint32_t e = (ax >> 23) - 112u; /* rebias 127 -> 15 */
  uint32_t m = ax & 0x7FFFFFu;
  uint32_t h = (e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u; /* carry may bump the exponent */
  return sign | (uint16_t)h;
}

void orc_f32_to_f16_array(const float *x, uint64_t n, uint16_t *out) {
  for (uint64_t i = 0; i < n; i++) out[i] = orc_f32_to_f16(x[i]);
}

void orc_f16_to_f32_array(const uint16_t *x, uint64_t n, float *out) {
  for (uint64_t i = 0; i < n; i++) out[i] = orc_f16_to_f32(x[i]);
}

/* ------------------------------------------------------------------------ */
/* Residual codec ops.                                                       */
/* ------------------------------------------------------------------------ */

/* decompress_residuals.cu:8-40: for every packed byte x of row i,
 *   x = reversed_bit_map[x]; for k < 8/nbits:
 *     out[i][j*8/nbits + k] = half(half(w[lut[x][k]]) + half(C[code_i][.]))
 * where the add is at::Half's (widen both to float, add, round to half). */
void orc_plaid_decompress_f16(const uint8_t *res, uint64_t n, const uint16_t *bw,
                              const uint8_t *rbm, const uint8_t *lut, const int32_t *codes,
                              const uint16_t *C, uint32_t dim, uint32_t nbits, uint16_t *out) {
  const uint32_t vpb = 8 / nbits, pk = dim / vpb;
  for (uint64_t i = 0; i < n; i++) {
    const uint16_t *c = C + (size_t)codes[i] * dim;
    for (uint32_t j = 0; j < pk; j++) {
      const uint8_t x = rbm[res[i * pk + j]];
      for (uint32_t k = 0; k < vpb; k++) {
        const uint32_t o = j * vpb + k;
        const float v = orc_f16_to_f32(bw[lut[(uint32_t)x * vpb + k]]) + orc_f16_to_f32(c[o]);
        out[i * dim + o] = orc_f32_to_f16(v);
      }
    }
  }
}

/* packbits.cu:10-37 (and numpy.packbits): bit b of the stream is 1 iff input
 * byte b is nonzero; byte j holds bits 8j..8j+7 with bit 8j in the MSB. */
void orc_plaid_packbits(const uint8_t *bits, uint64_t m, uint8_t *out) {
  for (uint64_t j = 0; j < m / 8; j++) {
    uint8_t v = 0;
    for (uint32_t b = 0; b < 8; b++) v = (uint8_t)((v << 1) | (bits[8 * j + b] != 0));
    out[j] = v;
  }
}

/* residual.py:189-224: bucket = torch.bucketize(e, cutoffs, right=False), the
 * number of cutoffs strictly below e; its nbits bits go out LSB first
 * (residuals >> arange_bits & 1) and are packed MSB first (packbits). */
void orc_plaid_binarize(const float *e, uint64_t n, uint32_t dim, const float *cutoffs,
                        uint32_t n_cut, uint32_t nbits, uint8_t *out) {
  const uint32_t row_bytes = dim * nbits / 8;
  for (uint64_t i = 0; i < n; i++) {
    uint8_t *o = out + i * row_bytes;
    memset(o, 0, row_bytes);
    for (uint32_t d = 0; d < dim; d++) {
      uint32_t b = 0;
      while (b < n_cut && cutoffs[b] < e[i * dim + d]) b++;
      for (uint32_t k = 0; k < nbits; k++) {
        const uint32_t pos = d * nbits + k; /* position in the row's bit stream */
        if ((b >> k) & 1u) o[pos >> 3] |= (uint8_t)(0x80u >> (pos & 7u));
      }
    }
  }
}

/* search/decompress_residuals.cpp:27-155 (CPU signature): the tokens of
 * pids[0..npids) in that order, row (cum[i] + j) of out = token j of pids[i],
 *   out[.][k*vpb + l] = w[lut[rbm[byte k]][l]] + C[code][k*vpb + l]   (FP32).
 * Returns the number of rows written. */
uint64_t orc_plaid_decompress_pids(const int32_t *pids, uint32_t npids, const int64_t *lengths,
                                   const int64_t *offsets, const float *bw, const uint8_t *rbm,
                                   const uint8_t *lut, const uint8_t *res, const int32_t *codes,
                                   const float *C, uint32_t dim, uint32_t nbits, float *out) {
  const uint32_t vpb = 8 / nbits, pk = dim / vpb;
  uint64_t row = 0;
  for (uint32_t i = 0; i < npids; i++) {
    const int32_t pid = pids[i];
    const int64_t off = offsets[pid];
    for (int64_t j = 0; j < lengths[pid]; j++, row++) {
      const int32_t code = codes[off + j];
      for (uint32_t k = 0; k < pk; k++) {
        const uint8_t x = rbm[res[(uint64_t)(off + j) * pk + k]];
        for (uint32_t l = 0; l < vpb; l++) {
          const uint32_t o = k * vpb + l;
          out[row * dim + o] = bw[lut[(uint32_t)x * vpb + l]] + C[(size_t)code * dim + o];
        }
      }
    }
  }
  return row;
}

/* segmented_maxsim.cpp:22-93: per doc, max over its rows of scores[row][t]
 * with zero-initialised maxima, then the left-to-right sum over t. */
void orc_plaid_segmented_maxsim(const float *scores, const int64_t *lengths, uint32_t ndocs,
                                uint32_t nq, float *out) {
  const float *p = scores;
  float *m = (float *)malloc(sizeof(float) * (nq ? nq : 1));
  for (uint32_t d = 0; d < ndocs; d++) {
    for (uint32_t t = 0; t < nq; t++) m[t] = 0.0f;
    for (int64_t j = 0; j < lengths[d]; j++, p += nq)
      for (uint32_t t = 0; t < nq; t++) m[t] = (m[t] < p[t]) ? p[t] : m[t]; /* std::max */
    float s = 0.0f;
    for (uint32_t t = 0; t < nq; t++) s += m[t];
    out[d] = s;
  }
  free(m);
}

/* ------------------------------------------------------------------------ */
/* Two-stage pid filter (filter_pids.cpp).                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
  float s;
  int32_t pid;
} orc_sp;

/* (score, pid) descending: std::priority_queue<std::pair<float,int>>::top order */
static int orc_cmp_sp_desc(const void *a, const void *b) {
  const orc_sp *x = (const orc_sp *)a, *y = (const orc_sp *)b;
  if (x->s != y->s) return (x->s > y->s) ? -1 : 1;
  return (x->pid > y->pid) ? -1 : (x->pid < y->pid);
}

/* filter_pids.cpp:30-80, one pid: m[k] = max over the doc's codes c with
 * idx[c] of S[c][k] (identity -9999), score = ((0 + m[0]) + m[1]) + ...
 * *n_unique receives the number of distinct codes that passed idx. */
static float orc_plaid_approx(int32_t pid, const float *S, uint32_t nq, const int32_t *codes,
                              const int64_t *doclens, const int64_t *offsets, const uint8_t *idx,
                              uint8_t *seen, float *m, uint64_t *n_unique) {
  for (uint32_t k = 0; k < nq; k++) m[k] = -9999.0f;
  uint64_t nu = 0;
  const int64_t off = offsets[pid];
  for (int64_t j = 0; j < doclens[pid]; j++) {
    const int32_t code = codes[off + j];
    if ((!idx || idx[code]) && !seen[code]) {
      for (uint32_t k = 0; k < nq; k++) {
        const float v = S[(size_t)code * nq + k];
        m[k] = (m[k] < v) ? v : m[k];
      }
      seen[code] = 1;
      nu++;
    }
  }
  for (int64_t j = 0; j < doclens[pid]; j++) seen[codes[off + j]] = 0;
  float score = 0.0f;
  for (uint32_t k = 0; k < nq; k++) score += m[k];
  *n_unique += nu;
  return score;
}

/* filter_pids.cpp:82-145 (helper): keep the best n_keep of pids by (score,
 * pid) descending, written in that order.  Returns min(n_keep, npids) (the
 * reference pops an empty heap when npids < n_keep). */
static uint32_t orc_plaid_filter_stage(const int32_t *pids, uint32_t npids, const float *S,
                                       uint32_t n_c, uint32_t nq, const int32_t *codes,
                                       const int64_t *doclens, const int64_t *offsets,
                                       const uint8_t *idx, uint32_t n_keep, int32_t *out,
                                       float *out_s, uint64_t *n_vec) {
  uint8_t *seen = (uint8_t *)calloc(n_c ? n_c : 1, 1);
  float *m = (float *)malloc(sizeof(float) * (nq ? nq : 1));
  orc_sp *sp = (orc_sp *)malloc(sizeof(orc_sp) * (npids ? npids : 1));
  uint64_t nu = 0;
  for (uint32_t i = 0; i < npids; i++) {
    sp[i].pid = pids[i];
    sp[i].s = orc_plaid_approx(pids[i], S, nq, codes, doclens, offsets, idx, seen, m, &nu);
  }
  qsort(sp, npids, sizeof(orc_sp), orc_cmp_sp_desc);
  const uint32_t n_out = npids < n_keep ? npids : n_keep;
  for (uint32_t i = 0; i < n_out; i++) {
    out[i] = sp[i].pid;
    if (out_s) out_s[i] = sp[i].s;
  }
  *n_vec += nu * nq; /* filter_pids.cpp:67 counts nquery_vectors per distinct code */
  free(seen); free(m); free(sp);
  return n_out;
}

/* filter_pids.cpp:147-193: stage 1 keeps ndocs pids scored over the codes
 * that pass idx, stage 2 rescores those over all codes and keeps ndocs / 4.
 * S is centroid_scores [n_c][nq].  Returns the number of pids written to out
 * (min(ndocs / 4, min(ndocs, npids))); *n_vec_score_refine gets both stages'
 * counts.  out_s (nullable) receives the stage-2 scores. */
uint32_t orc_plaid_filter_pids(const int32_t *pids, uint32_t npids, const float *S, uint32_t n_c,
                               uint32_t nq, const int32_t *codes, const int64_t *doclens,
                               const int64_t *offsets, const uint8_t *idx, uint32_t ndocs,
                               int32_t *out, float *out_s, uint64_t *n_vec_score_refine) {
  int32_t *p1 = (int32_t *)malloc(sizeof(int32_t) * (ndocs ? ndocs : 1));
  uint64_t nv = 0;
  const uint32_t n1 = orc_plaid_filter_stage(pids, npids, S, n_c, nq, codes, doclens, offsets, idx,
                                             ndocs, p1, NULL, &nv);
  const uint32_t n2 = orc_plaid_filter_stage(p1, n1, S, n_c, nq, codes, doclens, offsets, NULL,
                                             ndocs / 4, out, out_s, &nv);
  free(p1);
  if (n_vec_score_refine) *n_vec_score_refine = nv;
  return n2;
}

/* ------------------------------------------------------------------------ */
/* Whole CPU branch of IndexScorer.rank for a batch of queries.              */
/* ------------------------------------------------------------------------ */

/* dots of x [dim] with 8 columns of QT [dim][ld]: eight sequential FMA chains */
static inline __m256 orc_p_dot8(const float *x, const float *QT, size_t ld, uint32_t dim) {
  __m256 a = _mm256_setzero_ps();
  for (uint32_t d = 0; d < dim; d++)
    a = _mm256_fmadd_ps(_mm256_set1_ps(x[d]), _mm256_loadu_ps(QT + (size_t)d * ld), a);
  return a;
}

/* centroid_scores = C @ Q^T (candidate_generation.py:14): S[c][t], one FMA
 * chain over d per entry.  Q is [nq][dim]. */
void orc_plaid_centroid_scores(const float *C, uint32_t n_c, const float *Q, uint32_t nq,
                               uint32_t dim, float *S) {
  const uint32_t tp = (nq + 7) & ~7u;
  float *QT = (float *)calloc((size_t)dim * tp, sizeof(float));
  float *row = (float *)malloc(sizeof(float) * tp);
  for (uint32_t t = 0; t < nq; t++)
    for (uint32_t d = 0; d < dim; d++) QT[(size_t)d * tp + t] = Q[(size_t)t * dim + d];
  for (uint32_t c = 0; c < n_c; c++) {
    for (uint32_t t0 = 0; t0 < tp; t0 += 8)
      _mm256_storeu_ps(row + t0, orc_p_dot8(C + (size_t)c * dim, QT + t0, tp, dim));
    memcpy(S + (size_t)c * nq, row, sizeof(float) * nq);
  }
  free(QT); free(row);
}

/* score_pids refine (index_storage.py:164-190, CPU branch) for one query:
 * decompress in FP32, y = x / max(||x||, 1e-12) with ||x|| = sqrtf(FMA chain
 * of x_d^2), s_t = FMA chain of y_d * Q_t,d, then segmented MaxSim. */
void orc_plaid_score_pids(const float *Q, uint32_t nq, const int32_t *pids, uint32_t npids,
                          const int64_t *doclens, const int64_t *offsets, const int32_t *codes,
                          const uint8_t *res, const float *C, const float *bw,
                          const uint8_t *rbm, const uint8_t *lut, uint32_t dim, uint32_t nbits,
                          float *scores) {
  const uint32_t vpb = 8 / nbits, pk = dim / vpb;
  const uint32_t tp = (nq + 7) & ~7u;
  float *QT = (float *)calloc((size_t)dim * tp, sizeof(float));
  float *x = (float *)malloc(sizeof(float) * dim);
  float *sv = (float *)malloc(sizeof(float) * tp);
  float *M = (float *)malloc(sizeof(float) * tp);
  for (uint32_t t = 0; t < nq; t++)
    for (uint32_t d = 0; d < dim; d++) QT[(size_t)d * tp + t] = Q[(size_t)t * dim + d];
  for (uint32_t i = 0; i < npids; i++) {
    const int32_t pid = pids[i];
    const int64_t off = offsets[pid];
    for (uint32_t t = 0; t < nq; t++) M[t] = 0.0f;
    for (int64_t j = 0; j < doclens[pid]; j++) {
      const uint64_t tok = (uint64_t)(off + j);
      const float *c = C + (size_t)codes[tok] * dim;
      for (uint32_t k = 0; k < pk; k++) {
        const uint8_t b = rbm[res[tok * pk + k]];
        for (uint32_t l = 0; l < vpb; l++) x[k * vpb + l] = bw[lut[(uint32_t)b * vpb + l]] + c[k * vpb + l];
      }
      float n2 = 0.0f;
      for (uint32_t d = 0; d < dim; d++) n2 = fmaf(x[d], x[d], n2);
      const float den = fmaxf(sqrtf(n2), 1e-12f);
      for (uint32_t d = 0; d < dim; d++) x[d] = x[d] / den;
      for (uint32_t t0 = 0; t0 < tp; t0 += 8) _mm256_storeu_ps(sv + t0, orc_p_dot8(x, QT + t0, tp, dim));
      for (uint32_t t = 0; t < nq; t++) M[t] = (M[t] < sv[t]) ? sv[t] : M[t];
    }
    float s = 0.0f;
    for (uint32_t t = 0; t < nq; t++) s += M[t];
    scores[i] = s;
  }
  free(QT); free(x); free(sv); free(M);
}

typedef struct {
  float s;
  int32_t id;
} orc_si;

/* (score desc, id asc) */
static int orc_cmp_si(const void *a, const void *b) {
  const orc_si *x = (const orc_si *)a, *y = (const orc_si *)b;
  if (x->s != y->s) return (x->s > y->s) ? -1 : 1;
  return (x->id < y->id) ? -1 : (x->id > y->id);
}

/* get_cells + generate_candidate_pids (candidate_generation.py:13-66) for one
 * query: top-ncells centroids per token of S [n_c][nq], their union, then the
 * sorted unique pids of those IVF lists.  Returns the pid count. */
static uint32_t orc_plaid_candidates(const float *S, uint32_t n_c, uint32_t nq, uint32_t ncells,
                                     const int32_t *ivf, const int64_t *ivf_off,
                                     const int64_t *ivf_len, uint32_t n_doc, uint8_t *cell_mark,
                                     uint8_t *pid_mark, int32_t *pids) {
  orc_si *best = (orc_si *)malloc(sizeof(orc_si) * (ncells ? ncells : 1));
  memset(cell_mark, 0, n_c);
  for (uint32_t t = 0; t < nq; t++) {
    uint32_t nb = 0;
    for (uint32_t c = 0; c < n_c; c++) {
      const orc_si e = {S[(size_t)c * nq + t], (int32_t)c};
      if (nb < ncells) {
        best[nb++] = e;
      } else if (orc_cmp_si(&e, &best[nb - 1]) < 0) {
        best[nb - 1] = e;
      } else {
        continue;
      }
      for (uint32_t i = nb - 1; i > 0 && orc_cmp_si(&best[i], &best[i - 1]) < 0; i--) {
        const orc_si tmp = best[i]; best[i] = best[i - 1]; best[i - 1] = tmp;
      }
    }
    for (uint32_t i = 0; i < nb; i++) cell_mark[best[i].id] = 1;
  }
  memset(pid_mark, 0, n_doc);
  for (uint32_t c = 0; c < n_c; c++)
    if (cell_mark[c])
      for (int64_t j = 0; j < ivf_len[c]; j++) pid_mark[ivf[ivf_off[c] + j]] = 1;
  uint32_t n = 0;
  for (uint32_t d = 0; d < n_doc; d++)
    if (pid_mark[d]) pids[n++] = (int32_t)d;
  free(best);
  return n;
}

/* ------------------------------------------------------------------------ */
/* GPU branch of IndexScorer.score_pids (index_storage.py:112-153, 160-161,  */
/* 185-187) in float16, as the reference computes it on a CUDA device:       */
/* centroids and bucket weights are halves (residual.py:30-45), so are the   */
/* queries (candidate_generation.py:56, colbert.py:198), and so is every     */
/* matmul output and reduction result.                                       */
/* ------------------------------------------------------------------------ */

static inline float orc_r16(float x) { return orc_f16_to_f32(orc_f32_to_f16(x)); }

/* colbert_score_reduce writes -9999 into the half tensor's padding
 * (colbert.py:135): half(-9999) = -10000 */
#define ORC_PAD16 (-10000.0f)

/* One token of ResidualCodec.decompress, GPU branch (residual.py:255-274):
 * x_d = half(w16[bucket] + c16_d) (decompress_residuals.cu, at::Half add),
 * then torch.nn.functional.normalize on the half row: n = half(sqrtf(sum of
 * x_d^2 in FP32)) (the squares of halves are exact; the sum is a sequential
 * FMA chain), eps = 1e-12 rounds to 0 in half so the denominator is n, and
 * y_d = half(x_d / n) with the FP32 quotient.  c16, w16: half values. */
static void orc_p_token16(const float *c16, const uint8_t *rb, const float *w16,
                          const uint8_t *rbm, const uint8_t *lut, uint32_t dim, uint32_t nbits,
                          float *y) {
  const uint32_t vpb = 8 / nbits, pk = dim / vpb;
  for (uint32_t k = 0; k < pk; k++) {
    const uint8_t b = rbm[rb[k]];
    for (uint32_t l = 0; l < vpb; l++)
      y[k * vpb + l] = orc_r16(w16[lut[(uint32_t)b * vpb + l]] + c16[k * vpb + l]);
  }
  float n2 = 0.0f;
  for (uint32_t d = 0; d < dim; d++) n2 = fmaf(y[d], y[d], n2);
  const float n = orc_r16(sqrtf(n2));
  for (uint32_t d = 0; d < dim; d++) y[d] = orc_r16(y[d] / n);
}

/* ResidualEmbeddingsStrided.lookup_pids + ResidualCodec.decompress, GPU
 * branch (residual_embeddings_strided.py:30-37, residual.py:243-279): the
 * normalised half rows of the tokens of pids[0..npids) in order, returned as
 * float (half values).  C and bw are rounded to half first.  Returns the
 * number of rows. */
uint64_t orc_plaid_lookup_pids_f16(const int32_t *pids, uint32_t npids, const int64_t *doclens,
                                   const int64_t *offsets, const int32_t *codes, const uint8_t *res,
                                   const float *C, const float *bw, const uint8_t *rbm,
                                   const uint8_t *lut, uint32_t dim, uint32_t nbits, float *out) {
  const uint32_t pk = dim * nbits / 8;
  float w16[256], *c16 = (float *)malloc(sizeof(float) * dim);
  for (uint32_t b = 0; b < (1u << nbits); b++) w16[b] = orc_r16(bw[b]);
  uint64_t row = 0;
  for (uint32_t i = 0; i < npids; i++) {
    const int64_t off = offsets[pids[i]];
    for (int64_t j = 0; j < doclens[pids[i]]; j++, row++) {
      const uint64_t tok = (uint64_t)(off + j);
      for (uint32_t d = 0; d < dim; d++) c16[d] = orc_r16(C[(size_t)codes[tok] * dim + d]);
      orc_p_token16(c16, res + tok * pk, w16, rbm, lut, dim, nbits, out + row * dim);
    }
  }
  free(c16);
  return row;
}

/* colbert_score_reduce, GPU branch (colbert.py:133-155) over a packed half
 * score matrix: per doc, M_t = max over its rows of s[row][t] with the
 * padding value -10000 as identity (a doc without rows keeps it), and
 * score = half(left-to-right FP32 sum over t of M_t). */
void orc_plaid_score_reduce_f16(const float *s, const int64_t *lengths, uint32_t ndocs,
                                uint32_t nq, float *out) {
  const float *p = s;
  float *m = (float *)malloc(sizeof(float) * (nq ? nq : 1));
  for (uint32_t d = 0; d < ndocs; d++) {
    for (uint32_t t = 0; t < nq; t++) m[t] = ORC_PAD16;
    for (int64_t j = 0; j < lengths[d]; j++, p += nq)
      for (uint32_t t = 0; t < nq; t++) m[t] = (m[t] < p[t]) ? p[t] : m[t];
    float sum = 0.0f;
    for (uint32_t t = 0; t < nq; t++) sum += m[t];
    out[d] = orc_r16(sum);
  }
  free(m);
}

/* colbert_score_packed, GPU branch (colbert.py:181-205), one query: D (half
 * values) [sum(lengths)][dim] times Q^T rounded to half (Q is rounded to half
 * first, colbert.py:198), each entry a sequential FMA chain over d rounded to
 * half, then colbert_score_reduce. */
void orc_plaid_score_packed_f16(const float *Q, uint32_t nq, const float *D, const int64_t *lengths,
                                uint32_t ndocs, uint32_t dim, float *out) {
  uint64_t rows = 0;
  for (uint32_t d = 0; d < ndocs; d++) rows += (uint64_t)lengths[d];
  const uint32_t tp = (nq + 7) & ~7u;
  float *QT = (float *)calloc((size_t)dim * tp, sizeof(float));
  float *s = (float *)malloc(sizeof(float) * (rows ? rows : 1) * (nq ? nq : 1));
  float *sv = (float *)malloc(sizeof(float) * tp);
  for (uint32_t t = 0; t < nq; t++)
    for (uint32_t d = 0; d < dim; d++) QT[(size_t)d * tp + t] = orc_r16(Q[(size_t)t * dim + d]);
  for (uint64_t r = 0; r < rows; r++) {
    for (uint32_t t0 = 0; t0 < tp; t0 += 8) _mm256_storeu_ps(sv + t0, orc_p_dot8(D + r * dim, QT + t0, tp, dim));
    for (uint32_t t = 0; t < nq; t++) s[r * nq + t] = orc_r16(sv[t]);
  }
  orc_plaid_score_reduce_f16(s, lengths, ndocs, nq, out);
  free(QT); free(s); free(sv);
}

/* index_storage.py:120-138 (stage 1) or :143-149 (stage 2, idx NULL) for one
 * pid: colbert_score_reduce of the S16 rows of the doc's codes that pass idx
 * (repeated codes change no maximum, so each distinct code is visited once).
 * *n_unique counts the distinct passing codes; *any is set when one passed. */
static float orc_plaid_approx16(int32_t pid, const float *S, uint32_t nq, const int32_t *codes,
                                const int64_t *doclens, const int64_t *offsets, const uint8_t *idx,
                                uint8_t *seen, float *m, uint64_t *n_unique, int *any) {
  for (uint32_t k = 0; k < nq; k++) m[k] = ORC_PAD16;
  uint64_t nu = 0;
  const int64_t off = offsets[pid];
  for (int64_t j = 0; j < doclens[pid]; j++) {
    const int32_t code = codes[off + j];
    if ((!idx || idx[code]) && !seen[code]) {
      for (uint32_t k = 0; k < nq; k++) {
        const float v = S[(size_t)code * nq + k];
        m[k] = (m[k] < v) ? v : m[k];
      }
      seen[code] = 1;
      nu++;
    }
  }
  for (int64_t j = 0; j < doclens[pid]; j++) seen[codes[off + j]] = 0;
  float score = 0.0f;
  for (uint32_t k = 0; k < nq; k++) score += m[k];
  *n_unique += nu;
  if (nu) *any = 1;
  return orc_r16(score);
}

/* The two filter stages of the GPU branch for one query (index_storage.py:
 * 120-153): stage 1 keeps the best ndocs of the candidates (torch.topk, only
 * when there are more), stage 2 rescores those over all their codes and keeps
 * ndocs / 4.  A 2^20-pid batch in which no candidate holds a passing code
 * scores 0 throughout (index_storage.py:133-135).  Both stages keep (score,
 * pid) descending.  Returns the number kept; *n_vec counts distinct visited
 * codes x nq over both stages. */
static uint32_t orc_plaid_filter16(const int32_t *pids, uint32_t npids, const float *S, uint32_t n_c,
                                   uint32_t nq, const int32_t *codes, const int64_t *doclens,
                                   const int64_t *offsets, const uint8_t *idx, uint32_t ndocs,
                                   int32_t *out, uint64_t *n_vec) {
  uint8_t *seen = (uint8_t *)calloc(n_c ? n_c : 1, 1);
  float *m = (float *)malloc(sizeof(float) * (nq ? nq : 1));
  orc_sp *sp = (orc_sp *)malloc(sizeof(orc_sp) * (npids ? npids : 1));
  uint64_t nu = 0;
  const uint32_t batch = 1u << 20;
  for (uint32_t b0 = 0; b0 < npids; b0 += batch) {
    const uint32_t b1 = (npids - b0 > batch) ? b0 + batch : npids;
    int any = 0;
    for (uint32_t i = b0; i < b1; i++) {
      sp[i].pid = pids[i];
      sp[i].s = orc_plaid_approx16(pids[i], S, nq, codes, doclens, offsets, idx, seen, m, &nu, &any);
    }
    if (!any)
      for (uint32_t i = b0; i < b1; i++) sp[i].s = 0.0f;
  }
  qsort(sp, npids, sizeof(orc_sp), orc_cmp_sp_desc);
  const uint32_t n1 = npids < ndocs ? npids : ndocs;
  int dummy = 0;
  for (uint32_t i = 0; i < n1; i++)
    sp[i].s = orc_plaid_approx16(sp[i].pid, S, nq, codes, doclens, offsets, NULL, seen, m, &nu, &dummy);
  qsort(sp, n1, sizeof(orc_sp), orc_cmp_sp_desc);
  const uint32_t n2 = n1 < ndocs / 4 ? n1 : ndocs / 4;
  for (uint32_t i = 0; i < n2; i++) out[i] = sp[i].pid;
  *n_vec += nu * nq;
  free(seen); free(m); free(sp);
  return n2;
}

/* IndexScorer.rank (index_storage.py:82-198) for n_q queries of nq tokens,
 * CPU branch (gpu_branch 0: FP32, filter_pids, zero-initialised MaxSim) or GPU
 * branch (1: the float16 pipeline above, idx = max_t S16 >= half(threshold),
 * the comparison torch makes between a half tensor and a Python float).
 * Candidates, the two-stage filter, the refine, then the best k by (score
 * desc, pid asc).  Missing results are (-1, -inf).  stats [n_q][3]
 * (nullable): pids after candidate generation, pids after the filter,
 * n_vec_score_refine (the GPU branch reports -1 in the reference; here the
 * same count as the CPU branch). */
static void orc_plaid_search_impl(uint32_t n_doc, uint32_t n_c, uint32_t dim, uint32_t nbits,
                                  const int64_t *doclens, const int64_t *offsets, const int32_t *codes,
                                  const uint8_t *res, const float *C, const float *bw,
                                  const uint8_t *rbm, const uint8_t *lut, const int32_t *ivf,
                                  const int64_t *ivf_off, const int64_t *ivf_len, const float *Qall,
                                  uint32_t n_q, uint32_t nq, uint32_t ncells, float threshold,
                                  uint32_t ndocs, uint32_t k, int n_threads, int gpu_branch,
                                  int32_t *out_ids, float *out_s, uint64_t *stats) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  float *C16 = NULL;
  if (gpu_branch) {
    C16 = (float *)malloc(sizeof(float) * (size_t)n_c * dim);
    for (size_t i = 0; i < (size_t)n_c * dim; i++) C16[i] = orc_r16(C[i]);
  }
  const float thr = gpu_branch ? orc_r16(threshold) : threshold;
#pragma omp parallel
  {
    float *S = (float *)malloc(sizeof(float) * (size_t)n_c * nq);
    float *Q16 = (float *)malloc(sizeof(float) * (size_t)nq * dim);
    uint8_t *idx = (uint8_t *)malloc(n_c);
    uint8_t *cell_mark = (uint8_t *)malloc(n_c);
    uint8_t *pid_mark = (uint8_t *)malloc(n_doc ? n_doc : 1);
    int32_t *pids = (int32_t *)malloc(sizeof(int32_t) * (n_doc ? n_doc : 1));
    int32_t *fp = (int32_t *)malloc(sizeof(int32_t) * (ndocs / 4 + 1));
    float *sc = (float *)malloc(sizeof(float) * (ndocs / 4 + 1));
    orc_si *r = (orc_si *)malloc(sizeof(orc_si) * (ndocs / 4 + 1));
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)n_q; q++) {
      const float *Q = Qall + (size_t)q * nq * dim;
      if (gpu_branch) {
        for (size_t i = 0; i < (size_t)nq * dim; i++) Q16[i] = orc_r16(Q[i]);
        orc_plaid_centroid_scores(C16, n_c, Q16, nq, dim, S);
        for (size_t i = 0; i < (size_t)n_c * nq; i++) S[i] = orc_r16(S[i]);
      } else {
        orc_plaid_centroid_scores(C, n_c, Q, nq, dim, S);
      }
      const uint32_t np = orc_plaid_candidates(S, n_c, nq, ncells, ivf, ivf_off, ivf_len, n_doc,
                                               cell_mark, pid_mark, pids);
      for (uint32_t c = 0; c < n_c; c++) {
        float m = S[(size_t)c * nq];
        for (uint32_t t = 1; t < nq; t++) m = fmaxf(m, S[(size_t)c * nq + t]);
        idx[c] = m >= thr;
      }
      uint64_t nv = 0;
      uint32_t nf;
      if (gpu_branch) {
        nf = orc_plaid_filter16(pids, np, S, n_c, nq, codes, doclens, offsets, idx, ndocs, fp, &nv);
        uint64_t rows = 0;
        int64_t *ln = (int64_t *)malloc(sizeof(int64_t) * (nf ? nf : 1));
        for (uint32_t i = 0; i < nf; i++) rows += (uint64_t)(ln[i] = doclens[fp[i]]);
        float *D = (float *)malloc(sizeof(float) * (rows ? rows : 1) * dim);
        orc_plaid_lookup_pids_f16(fp, nf, doclens, offsets, codes, res, C, bw, rbm, lut, dim, nbits, D);
        orc_plaid_score_packed_f16(Q, nq, D, ln, nf, dim, sc);
        free(D); free(ln);
      } else {
        nf = orc_plaid_filter_pids(pids, np, S, n_c, nq, codes, doclens, offsets, idx, ndocs, fp,
                                   NULL, &nv);
        orc_plaid_score_pids(Q, nq, fp, nf, doclens, offsets, codes, res, C, bw, rbm, lut, dim,
                             nbits, sc);
      }
      for (uint32_t i = 0; i < nf; i++) { r[i].s = sc[i]; r[i].id = fp[i]; }
      qsort(r, nf, sizeof(orc_si), orc_cmp_si);
      for (uint32_t i = 0; i < k; i++) {
        out_ids[(size_t)q * k + i] = i < nf ? r[i].id : -1;
        out_s[(size_t)q * k + i] = i < nf ? r[i].s : -INFINITY;
      }
      if (stats) {
        stats[q * 3 + 0] = np;
        stats[q * 3 + 1] = nf;
        stats[q * 3 + 2] = nv;
      }
    }
    free(S); free(Q16); free(idx); free(cell_mark); free(pid_mark); free(pids); free(fp); free(sc);
    free(r);
  }
  free(C16);
}

/* IndexScorer.rank, CPU branch (index_storage.py:82-198). */
void orc_plaid_search_batch(uint32_t n_doc, uint32_t n_c, uint32_t dim, uint32_t nbits,
                            const int64_t *doclens, const int64_t *offsets, const int32_t *codes,
                            const uint8_t *res, const float *C, const float *bw,
                            const uint8_t *rbm, const uint8_t *lut, const int32_t *ivf,
                            const int64_t *ivf_off, const int64_t *ivf_len, const float *Qall,
                            uint32_t n_q, uint32_t nq, uint32_t ncells, float threshold,
                            uint32_t ndocs, uint32_t k, int n_threads, int32_t *out_ids,
                            float *out_s, uint64_t *stats) {
  orc_plaid_search_impl(n_doc, n_c, dim, nbits, doclens, offsets, codes, res, C, bw, rbm, lut, ivf,
                        ivf_off, ivf_len, Qall, n_q, nq, ncells, threshold, ndocs, k, n_threads, 0,
                        out_ids, out_s, stats);
}

/* IndexScorer.rank, GPU branch (index_storage.py:82-198 with use_gpu). */
void orc_plaid_search_batch_gpu(uint32_t n_doc, uint32_t n_c, uint32_t dim, uint32_t nbits,
                                const int64_t *doclens, const int64_t *offsets,
                                const int32_t *codes, const uint8_t *res, const float *C,
                                const float *bw, const uint8_t *rbm, const uint8_t *lut,
                                const int32_t *ivf, const int64_t *ivf_off, const int64_t *ivf_len,
                                const float *Qall, uint32_t n_q, uint32_t nq, uint32_t ncells,
                                float threshold, uint32_t ndocs, uint32_t k, int n_threads,
                                int32_t *out_ids, float *out_s, uint64_t *stats) {
  orc_plaid_search_impl(n_doc, n_c, dim, nbits, doclens, offsets, codes, res, C, bw, rbm, lut, ivf,
                        ivf_off, ivf_len, Qall, n_q, nq, ncells, threshold, ndocs, k, n_threads, 1,
                        out_ids, out_s, stats);
}
