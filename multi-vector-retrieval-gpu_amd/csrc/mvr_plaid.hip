// PLAID (ColBERTv2) search path and residual codec ops for gfx950, with their
// C ABI (include/mvr_plaid.h).
//
// Replaces (reference paths, baseline/ColBERT/colbert/):
//   indexing/codecs/decompress_residuals.cu:8-75  -> k_pl_decompress_f16
//   indexing/codecs/packbits.cu:10-57             -> k_pl_packbits
//   indexing/codecs/residual.py:189-224 (binarize) -> k_pl_binarize
//   search/decompress_residuals.cpp:27-155        -> k_pl_decompress_pids
//   modeling/segmented_maxsim.cpp:22-93           -> k_pl_segmented_maxsim
//   search/filter_pids.cpp:30-193                 -> k_pl_approx (both stages) + k_select_topr
//   search/candidate_generation.py:13-66          -> k_pl_sgemm, k_pl_cell_part/_merge, k_pl_cells,
//                                                    k_pl_mark, k_pl_compact
//   search/index_storage.py:101-198 (CPU branch)  -> the batched pipeline of mvr_plaid_search
//
// Design: the reference runs one query at a time through pthread loops
// (filter, decompress, MaxSim) and a priority_queue merge.  Here a device pass
// takes a batch of queries through seven kernels with no host round trip:
// centroid scores S = C.Q^T for the whole batch, per-token top-ncells and the
// IVF union (a per-query pid bitmap whose first setter appends the pid), the
// two filter stages as one wave per pid (dedup of codes in an LDS bitmap,
// 32 lanes = 32 query tokens reading one 128-byte row of S per code), exact
// top-n selection on 64-bit (score, pid) keys, and a refine kernel that
// decompresses, L2-normalises and scores the surviving docs without writing
// any decompressed embedding to HBM.
#include "../../include/mvr.h"
#include "../../include/mvr_plaid.h"
#include "mvr_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

namespace mvr {
namespace {

constexpr uint32_t kPlMaxCells = 8;      // largest ncells
constexpr uint32_t kPlMaxBmBytes = 65536;  // LDS of one workgroup's dedup bitmaps

__device__ __forceinline__ void pl_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Order-preserving float -> uint32 map (larger float, larger uint).
__device__ __forceinline__ uint32_t pl_f2ord(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Selection key of (score, centroid): score order bits above ~centroid, so
// that larger keys come first and equal scores put the lower centroid first;
// -0 and +0 share a key (they compare equal).  Keys of distinct centroids
// differ, so the top-NC lists below need no tie rule of their own.
__device__ __forceinline__ unsigned long long pl_cell_key(float s, uint32_t c) {
  return ((unsigned long long)pl_f2ord(s == 0.0f ? 0.0f : s) << 32) | (uint32_t)~c;
}

__device__ __forceinline__ float pl_readlane(float v, uint32_t l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// std::max(a, b) of the reference (returns a unless a < b).
__device__ __forceinline__ float pl_max(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ float h2f(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}
// FP32 -> half, round to nearest even.  The empty asm makes the FP32 value
// opaque: without it the compiler may fold the conversion into the FMA that
// produced f (v_fma_mixlo_f16, one rounding of the exact result to half),
// while the reference rounds the FP32 result of its matmul or sum to half
// (two roundings); the two differ when the FP32 value is a half midpoint.
__device__ __forceinline__ uint16_t f2h(float f) {
  asm volatile("" : "+v"(f));
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}
// f rounded to the nearest half (ties to even), as a float
__device__ __forceinline__ float r16(float f) {
  asm volatile("" : "+v"(f));
  return (float)(_Float16)f;
}

// colbert_score_reduce writes -9999 into the padding of a half tensor
// (colbert.py:135): half(-9999) = -10000
constexpr float kPad16 = -10000.0f;

// ---------------------------------------------------------------------------
// Codec ops
// ---------------------------------------------------------------------------
// decompress_residuals.cu:8-40, one thread per packed byte (as the
// reference); the 256 x vpb table bw[lut[rbm[v]][k]] sits in LDS.
__global__ __launch_bounds__(256) void k_pl_decompress_f16(
    const uint8_t *__restrict__ res, uint64_t n_bytes, uint32_t pk, const uint16_t *__restrict__ bw,
    const uint8_t *__restrict__ rbm, const uint8_t *__restrict__ lut, uint32_t vpb,
    const int32_t *__restrict__ codes, const uint16_t *__restrict__ C, uint32_t n_c, uint32_t dim,
    uint16_t *__restrict__ out) {
  __shared__ uint16_t tab[256 * 8];
  for (uint32_t e = threadIdx.x; e < 256 * vpb; e += blockDim.x) {
    const uint32_t v = e / vpb, k = e % vpb;
    tab[e] = bw[lut[(uint32_t)rbm[v] * vpb + k]];
  }
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_bytes; e += stride) {
    const uint64_t i = e / pk;
    const uint32_t j = (uint32_t)(e - i * pk);
    const uint32_t code = (uint32_t)codes[i];
    const uint32_t v = res[e];
    const uint16_t *c = C + (size_t)code * dim + (size_t)j * vpb;
    uint16_t *o = out + i * dim + (size_t)j * vpb;
    for (uint32_t k = 0; k < vpb; ++k) {
      const float cf = code < n_c ? h2f(c[k]) : 0.0f;
      o[k] = f2h(h2f(tab[v * vpb + k]) + cf);
    }
  }
}

// packbits.cu:10-37: output byte j = input bytes 8j..8j+7 (nonzero -> 1), MSB first.
__global__ __launch_bounds__(256) void k_pl_packbits(const uint8_t *__restrict__ in, uint64_t n_out,
                                                     uint8_t *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_out; j += stride) {
    const uint8_t *p = in + 8 * j;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) v = (v << 1) | (p[b] != 0 ? 1u : 0u);
    out[j] = (uint8_t)v;
  }
}

// residual.py:189-224: one thread per output byte.
__global__ __launch_bounds__(256) void k_pl_binarize(const float *__restrict__ e, uint64_t n_out,
                                                     uint32_t row_bytes, uint32_t dim,
                                                     const float *__restrict__ cut, uint32_t n_cut,
                                                     uint32_t nbits, uint8_t *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n_out; o += stride) {
    const uint64_t i = o / row_bytes;
    const uint32_t j = (uint32_t)(o - i * row_bytes);
    uint32_t v = 0, last_d = 0xFFFFFFFFu, b = 0;
    for (uint32_t p = 8 * j; p < 8 * j + 8; ++p) {
      const uint32_t d = p / nbits, k = p - d * nbits;
      if (d != last_d) {
        const float x = e[i * dim + d];
        b = 0;
        for (uint32_t c = 0; c < n_cut; ++c) b += (cut[c] < x) ? 1u : 0u;
        last_d = d;
      }
      v = (v << 1) | ((b >> k) & 1u);
    }
    out[o] = (uint8_t)v;
  }
}

// cum[0] = 0, cum[i+1] = cum[i] + lengths[pids ? pids[i] : i]; one workgroup.
__global__ __launch_bounds__(1024) void k_pl_cumsum(const int32_t *__restrict__ pids, uint32_t n,
                                                    const int64_t *__restrict__ lengths,
                                                    int64_t *__restrict__ cum) {
  __shared__ int64_t part[1024];
  const uint32_t tid = threadIdx.x;
  const uint32_t chunk = (n + 1023) / 1024;
  const uint32_t b = min(n, tid * chunk), e = min(n, b + chunk);
  int64_t s = 0;
  for (uint32_t i = b; i < e; ++i) s += lengths[pids ? pids[i] : i];
  part[tid] = s;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const int64_t add = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  int64_t run = tid ? part[tid - 1] : 0;
  if (tid == 0) cum[0] = 0;
  for (uint32_t i = b; i < e; ++i) {
    run += lengths[pids ? pids[i] : i];
    cum[i + 1] = run;
  }
}

// segmented_lookup.cpp:25-49: segment i's rows, input[offsets[i] ..
// offsets[i] + lengths[i]), copied to out[cum[i] ..] in words of W bytes
// (row_words per row); one workgroup per segment, grid-stride over segments.
// A segment outside the input (negative offset or length, or past n_rows)
// is skipped and flagged in *bad.
template <typename W>
__global__ __launch_bounds__(256) void k_pl_seglookup(const W *__restrict__ in, uint64_t n_rows, uint32_t row_words,
                                                      const int64_t *__restrict__ lengths,
                                                      const int64_t *__restrict__ offsets,
                                                      const int64_t *__restrict__ cum, uint32_t n,
                                                      W *__restrict__ out, uint32_t *__restrict__ bad) {
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int64_t len = lengths[i], off = offsets[i];
    if (len < 0 || off < 0 || (uint64_t)off + (uint64_t)len > n_rows) {
      if (threadIdx.x == 0) atomicOr(bad, 1u);
      continue;
    }
    const uint64_t nw = (uint64_t)len * row_words;
    const W *src = in + (uint64_t)off * row_words;
    W *dst = out + (uint64_t)cum[i] * row_words;
    for (uint64_t w = threadIdx.x; w < nw; w += blockDim.x) dst[w] = src[w];
  }
}

// search/decompress_residuals.cpp:36-75, one workgroup per pid.
__global__ __launch_bounds__(256) void k_pl_decompress_pids(
    const int32_t *__restrict__ pids, const int64_t *__restrict__ cum, const int64_t *__restrict__ lengths,
    const int64_t *__restrict__ offsets, const float *__restrict__ bw, const uint8_t *__restrict__ rbm,
    const uint8_t *__restrict__ lut, uint32_t vpb, const uint8_t *__restrict__ res, uint32_t pk,
    const int32_t *__restrict__ codes, const float *__restrict__ C, uint32_t n_c, uint32_t dim,
    float *__restrict__ out) {
  __shared__ float tab[256 * 8];
  for (uint32_t e = threadIdx.x; e < 256 * vpb; e += blockDim.x) {
    const uint32_t v = e / vpb, k = e % vpb;
    tab[e] = bw[lut[(uint32_t)rbm[v] * vpb + k]];
  }
  __syncthreads();
  const uint32_t i = blockIdx.x;
  const int32_t pid = pids[i];
  const int64_t L = lengths[pid], off = offsets[pid];
  const uint64_t row0 = (uint64_t)cum[i];
  const uint64_t n_e = (uint64_t)L * pk;
  for (uint64_t e = threadIdx.x; e < n_e; e += blockDim.x) {
    const uint64_t j = e / pk;
    const uint32_t kb = (uint32_t)(e - j * pk);
    const uint64_t tok = (uint64_t)off + j;
    const uint32_t code = (uint32_t)codes[tok];
    const uint32_t v = res[tok * pk + kb];
    float *o = out + (row0 + j) * dim + (size_t)kb * vpb;
    const float *c = C + (size_t)code * dim + (size_t)kb * vpb;
    for (uint32_t l = 0; l < vpb; ++l) o[l] = tab[v * vpb + l] + (code < n_c ? c[l] : 0.0f);
  }
}

// segmented_maxsim.cpp:22-46: one wave per doc, lane t = query token.
__global__ __launch_bounds__(256) void k_pl_segmented_maxsim(const float *__restrict__ s, uint32_t nq,
                                                             const int64_t *__restrict__ cum,
                                                             uint32_t ndocs, float *__restrict__ out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t d = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (d >= ndocs) return;
  float m = 0.0f;
  if (lane < nq)
    for (int64_t r = cum[d]; r < cum[d + 1]; ++r) m = pl_max(m, s[(uint64_t)r * nq + lane]);
  float sum = 0.0f;
  for (uint32_t t = 0; t < nq; ++t) sum += pl_readlane(m, t);
  if (lane == 0) out[d] = sum;
}

// ---------------------------------------------------------------------------
// Candidate generation (candidate_generation.py:13-66)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Centroid scores S = C.Q^T (candidate_generation.py:14) for a batch:
// S[q][c][t] = sequential FMA chain over d = 0..127 of C[c][d] * Q[q][t][d],
// idx[q][c] = max_t S >= thr (index_storage.py:117).  Queries come padded to
// 32 rows.  Workgroup tile = 4 queries (128 rows) x 128 centroids; thread
// (tx, ty) of 16 x 16 keeps an 8 x 8 register tile: rows ty*8 .. ty*8+7 (so
// wave w holds all 32 rows of query w) and centroids tx*4+u and 64+tx*4+u
// (u < 4; consecutive lanes read consecutive float4s of the centroid tile).
// K goes through LDS in chunks of 32 dims; per dim a thread reads 4 float4s
// for 64 FMAs.
// ---------------------------------------------------------------------------
constexpr uint32_t kGk = 32;        // dims per LDS chunk
constexpr uint32_t kGt = 128;       // rows / centroids per tile
constexpr uint32_t kGp = kGt + 4;   // padded LDS row

// top1 (nullable): per row the tile's best (score, centroid) as the key
// (score order bits << 32 | ~centroid), written to top1[q][tile][t], so that
// ncells = 1 needs no second pass over S.
__global__ __launch_bounds__(256) void k_pl_sgemm(const float *__restrict__ Q, uint32_t n_q, uint32_t nq,
                                                  const float *__restrict__ C, uint32_t n_c, float thr,
                                                  float *__restrict__ S, uint8_t *__restrict__ idx,
                                                  unsigned long long *__restrict__ top1, int f16) {
  __shared__ __align__(16) float As[kGk][kGp];
  __shared__ __align__(16) float Bs[kGk][kGp];
  const uint32_t tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const uint32_t q0 = blockIdx.y * 4, c0 = blockIdx.x * kGt;
  const uint32_t n_rows = n_q * kMaxQTok;
  float acc[8][8];
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i)
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) acc[i][j] = 0.0f;
  const float4 *Q4 = reinterpret_cast<const float4 *>(Q);
  const float4 *C4 = reinterpret_cast<const float4 *>(C);
#pragma nounroll
  for (uint32_t k0 = 0; k0 < kDim; k0 += kGk) {
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t e = tid + 256 * u, r = e >> 3, d4 = e & 7;
      const uint32_t row = q0 * kMaxQTok + r, c = c0 + r;
      const float4 a = row < n_rows ? Q4[(size_t)row * (kDim / 4) + k0 / 4 + d4] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 b = c < n_c ? C4[(size_t)c * (kDim / 4) + k0 / 4 + d4] : make_float4(0.f, 0.f, 0.f, 0.f);
      As[4 * d4 + 0][r] = a.x; As[4 * d4 + 1][r] = a.y; As[4 * d4 + 2][r] = a.z; As[4 * d4 + 3][r] = a.w;
      Bs[4 * d4 + 0][r] = b.x; Bs[4 * d4 + 1][r] = b.y; Bs[4 * d4 + 2][r] = b.z; Bs[4 * d4 + 3][r] = b.w;
    }
    __syncthreads();
#pragma unroll 4
    for (uint32_t d = 0; d < kGk; ++d) {
      const float4 a0 = *reinterpret_cast<const float4 *>(&As[d][ty * 8]);
      const float4 a1 = *reinterpret_cast<const float4 *>(&As[d][ty * 8 + 4]);
      const float4 b0 = *reinterpret_cast<const float4 *>(&Bs[d][tx * 4]);
      const float4 b1 = *reinterpret_cast<const float4 *>(&Bs[d][64 + tx * 4]);
      const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
      const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
      for (uint32_t i = 0; i < 8; ++i)
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) acc[i][j] = __builtin_fmaf(bv[j], av[i], acc[i][j]);
    }
    __syncthreads();
  }
  const uint32_t qq = ty >> 2, t0 = (ty & 3) * 8, q = q0 + qq;
  const bool qv = q < n_q;
  if (f16) {  // GPU branch: the half matmul's output (FP32 chain rounded once)
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i)
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) acc[i][j] = r16(acc[i][j]);
  }
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    const uint32_t c = c0 + (j < 4 ? tx * 4 + j : 64 + tx * 4 + (j - 4));
    float m = -INFINITY;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i)
      if (t0 + i < nq) m = fmaxf(m, acc[i][j]);
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    if (qv && c < n_c) {
      float *srow = S + ((size_t)q * n_c + c) * nq;
      if (nq == kMaxQTok) {
        reinterpret_cast<float4 *>(srow + t0)[0] = make_float4(acc[0][j], acc[1][j], acc[2][j], acc[3][j]);
        reinterpret_cast<float4 *>(srow + t0)[1] = make_float4(acc[4][j], acc[5][j], acc[6][j], acc[7][j]);
      } else {
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i)
          if (t0 + i < nq) srow[t0 + i] = acc[i][j];
      }
      if ((ty & 3) == 0) idx[(size_t)q * n_c + c] = m >= thr ? 1 : 0;
    }
  }
  if (top1) {
    const uint32_t n_tiles = gridDim.x;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
      unsigned long long best = 0ull;
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t c = c0 + (j < 4 ? tx * 4 + j : 64 + tx * 4 + (j - 4));
        const unsigned long long key = pl_cell_key(acc[i][j], c);
        if (c < n_c && key > best) best = key;
      }
#pragma unroll
      for (uint32_t off = 1; off < 16; off <<= 1) {
        const unsigned long long o = __shfl_xor(best, off);
        best = o > best ? o : best;
      }
      if (tx == 0 && qv && t0 + i < nq) top1[((size_t)q * n_tiles + blockIdx.x) * kMaxQTok + t0 + i] = best;
    }
  }
}

// ncells = 1: per (query, token) the best of the tiles' keys; top_ids[q*32+t].
__global__ __launch_bounds__(256) void k_pl_top1_merge(const unsigned long long *__restrict__ top1, uint32_t n_q,
                                                       uint32_t nq, uint32_t n_tiles, uint32_t *__restrict__ top_ids) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  const uint32_t q = e / kMaxQTok, t = e % kMaxQTok;
  if (q >= n_q || t >= nq) return;
  unsigned long long best = 0ull;
#pragma unroll 8
  for (uint32_t i = 0; i < n_tiles; ++i) {
    const unsigned long long k = top1[((size_t)q * n_tiles + i) * kMaxQTok + t];
    best = k > best ? k : best;
  }
  top_ids[(size_t)q * kMaxQTok + t] = best ? ~(uint32_t)(best & 0xFFFFFFFFull) : 0xFFFFFFFFu;
}

// Inserts key k into the descending list b[0..NC) (the smallest drops out).
template <uint32_t NC>
__device__ __forceinline__ void pl_insert(unsigned long long (&b)[NC], unsigned long long k) {
#pragma unroll
  for (uint32_t j = 0; j < NC; ++j) {
    const unsigned long long hi = k > b[j] ? k : b[j];
    k = k > b[j] ? b[j] : k;
    b[j] = hi;
  }
}

// Per query token the best NC centroids of one chunk of centroids by (score
// desc, centroid asc): one wave per (query, chunk), lane (t, h) scanning the
// chunk's centroids c = h (mod 2) for token t, then the two halves merge.
// Key lists go to part[q][chunk][t][NC] (0 = no entry).
constexpr uint32_t kCellChunk = 1024;

template <uint32_t NC>
__global__ __launch_bounds__(256) void k_pl_cell_part(const float *__restrict__ S, uint32_t n_c, uint32_t nq,
                                                      uint32_t n_chunk, unsigned long long *__restrict__ part) {
  const uint32_t lane = threadIdx.x & 63, t = lane & 31, h = lane >> 5;
  const uint32_t ch = blockIdx.x * 4 + (threadIdx.x >> 6), q = blockIdx.y;
  if (ch >= n_chunk) return;
  const uint32_t c_begin = ch * kCellChunk, c_end = min(n_c, c_begin + kCellChunk);
  unsigned long long b[NC];
#pragma unroll
  for (uint32_t j = 0; j < NC; ++j) b[j] = 0ull;
  if (t < nq) {
    const float *Sq = S + (size_t)q * n_c * nq + t;
#pragma unroll 8
    for (uint32_t c = c_begin + h; c < c_end; c += 2) {
      const unsigned long long k = pl_cell_key(Sq[(size_t)c * nq], c);
      if (k > b[NC - 1]) pl_insert<NC>(b, k);
    }
  }
  // the partner half's list, read whole before any insert
  unsigned long long o[NC];
#pragma unroll
  for (uint32_t j = 0; j < NC; ++j) o[j] = __shfl(b[j], (int)((t + 32) & 63));
  if (h == 0) {
#pragma unroll
    for (uint32_t j = 0; j < NC; ++j) pl_insert<NC>(b, o[j]);
    if (t < nq) {
      const size_t off = (((size_t)q * n_chunk + ch) * kMaxQTok + t) * NC;
#pragma unroll
      for (uint32_t j = 0; j < NC; ++j) part[off + j] = b[j];
    }
  }
}

// Merge of the chunk lists: one thread per (query, token) writes the best NC
// centroids to top_ids[q*32 + t][NC] (0xFFFFFFFF for a missing entry).
template <uint32_t NC>
__global__ __launch_bounds__(256) void k_pl_cell_merge(const unsigned long long *__restrict__ part, uint32_t n_q,
                                                       uint32_t nq, uint32_t n_chunk,
                                                       uint32_t *__restrict__ top_ids) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  const uint32_t q = e / kMaxQTok, t = e % kMaxQTok;
  if (q >= n_q || t >= nq) return;
  unsigned long long b[NC];
#pragma unroll
  for (uint32_t j = 0; j < NC; ++j) b[j] = 0ull;
  for (uint32_t ch = 0; ch < n_chunk; ++ch) {
    const size_t off = (((size_t)q * n_chunk + ch) * kMaxQTok + t) * NC;
    for (uint32_t k = 0; k < NC; ++k) {
      const unsigned long long v = part[off + k];
      if (v <= b[NC - 1]) break;  // the chunk list is sorted
      pl_insert<NC>(b, v);
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < NC; ++j)
    top_ids[((size_t)q * kMaxQTok + t) * NC + j] = b[j] ? ~(uint32_t)(b[j] & 0xFFFFFFFFull) : 0xFFFFFFFFu;
}

// The union of the query's per-token top-ncells lists as an ascending cell list; one wave per query, an LDS bitmap of the
// n_c centroids.
__global__ __launch_bounds__(64) void k_pl_cells(const uint32_t *__restrict__ top_ids, uint32_t P,
                                                 uint32_t nq, uint32_t n_c, uint32_t *__restrict__ cells,
                                                 uint32_t cell_cap, uint32_t *__restrict__ cell_n) {
  extern __shared__ uint32_t bm[];
  const uint32_t lane = threadIdx.x, q = blockIdx.x;
  const uint32_t words = (n_c + 31) / 32;
  for (uint32_t w = lane; w < words; w += 64) bm[w] = 0;
  pl_wave_sync();
  for (uint32_t e = lane; e < nq * P; e += 64) {
    const uint32_t t = e / P, j = e - t * P;
    const uint32_t c = top_ids[((size_t)q * kMaxQTok + t) * P + j];
    if (c < n_c) atomicOr(&bm[c >> 5], 1u << (c & 31));
  }
  pl_wave_sync();
  uint32_t total = 0;
  for (uint32_t w0 = 0; w0 < words; w0 += 64) {
    const uint32_t w = w0 + lane;
    const uint32_t word = w < words ? bm[w] : 0u;
    const uint32_t cnt = __popc(word);
    uint32_t incl = cnt;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t v = (uint32_t)__shfl_up((int)incl, off);
      if (lane >= off) incl += v;
    }
    uint32_t pos = total + incl - cnt;
    uint32_t bits = word;
    while (bits) {
      const uint32_t b = __ffs(bits) - 1;
      bits &= bits - 1;
      if (pos < cell_cap) cells[(size_t)q * cell_cap + pos] = w * 32 + b;
      ++pos;
    }
    total += (uint32_t)__shfl((int)incl, 63);
  }
  if (lane == 0) cell_n[q] = min(total, cell_cap);
}

// Union of the IVF lists of a query's cells, as bits of a per-query pid bitmap.
__global__ __launch_bounds__(256) void k_pl_mark(const uint32_t *__restrict__ cells, uint32_t cell_cap,
                                                 const uint32_t *__restrict__ cell_n, const int32_t *__restrict__ ivf,
                                                 const int64_t *__restrict__ ivf_off,
                                                 const int64_t *__restrict__ ivf_len, uint32_t *__restrict__ pid_bm,
                                                 uint32_t bm_words) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = blockIdx.y;
  const uint32_t nw = gridDim.x * 4;
  const uint32_t n = cell_n[q];
  uint32_t *bmq = pid_bm + (size_t)q * bm_words;
  for (uint32_t ci = blockIdx.x * 4 + wave; ci < n; ci += nw) {
    const uint32_t c = cells[(size_t)q * cell_cap + ci];
    const int64_t off = ivf_off[c], len = ivf_len[c];
    for (int64_t j = lane; j < len; j += 64) {
      const uint32_t pid = (uint32_t)ivf[off + j];
      atomicOr(&bmq[pid >> 5], 1u << (pid & 31));
    }
  }
}

// Exclusive prefix sum over a 1024-thread workgroup; *total gets the sum.
__device__ __forceinline__ uint32_t pl_block_scan(uint32_t v, uint32_t *warp_sums, uint32_t *total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t incl = v;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)incl, off);
    if (lane >= off) incl += u;
  }
  if (lane == 63) warp_sums[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < 16; ++w) {
    const uint32_t x = warp_sums[w];
    before += w < wave ? x : 0u;
    all += x;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// The query's candidates in descending pid order (the set of
// candidate_generation.py:56-64, which sorts ascending), one workgroup of
// 1024 threads per query; the bitmap words are zeroed as they are read.
__global__ __launch_bounds__(1024) void k_pl_compact(uint32_t *__restrict__ pid_bm, uint32_t bm_words,
                                                     uint32_t *__restrict__ cand, uint32_t cand_cap,
                                                     uint32_t *__restrict__ cand_n) {
  __shared__ uint32_t warp_sums[16];
  const uint32_t tid = threadIdx.x, q = blockIdx.x;
  uint32_t *bmq = pid_bm + (size_t)q * bm_words;
  uint32_t *cq = cand + (size_t)q * cand_cap;
  uint32_t total = 0;
  for (uint32_t base = 0; base < bm_words; base += 1024) {
    const uint32_t k = base + tid;
    const uint32_t w = bm_words - 1 - k;  // descending words
    uint32_t word = 0;
    if (k < bm_words) {
      word = bmq[w];
      if (word) bmq[w] = 0;
    }
    uint32_t sum;
    uint32_t pos = total + pl_block_scan((uint32_t)__popc(word), warp_sums, &sum);
    while (word) {
      const uint32_t b = 31u - (uint32_t)__clz(word);
      word &= ~(1u << b);
      if (pos < cand_cap) cq[pos] = w * 32 + b;
      ++pos;
    }
    total += sum;
  }
  if (tid == 0) cand_n[q] = min(total, cand_cap);
}

// k_pl_mark + k_pl_compact in one workgroup per query when the pid bitmap
// fits in LDS: the IVF lists of the query's cells are or'ed into an LDS
// bitmap (list entries spread over the 1024 threads by a prefix of the list
// lengths, a forward list cursor per thread, eight loads in flight), then the
// bitmap is compacted to the same descending pid list.  No global bitmap.
constexpr uint32_t kMcCells = kMaxQTok * 8;  // cells per query: nq x ncells (ncells <= 8)

__global__ __launch_bounds__(1024) void k_pl_markc(const uint32_t *__restrict__ cells, uint32_t cell_cap,
                                                   const uint32_t *__restrict__ cell_n,
                                                   const int32_t *__restrict__ ivf,
                                                   const int64_t *__restrict__ ivf_off,
                                                   const int64_t *__restrict__ ivf_len, uint32_t bm_words,
                                                   uint32_t *__restrict__ cand, uint32_t cand_cap,
                                                   uint32_t *__restrict__ cand_n) {
  extern __shared__ uint32_t bm[];  // [bm_words]
  __shared__ uint32_t cpfx[kMcCells + 1];
  __shared__ int64_t coff[kMcCells];
  __shared__ uint32_t warp_sums[16];
  const uint32_t tid = threadIdx.x, q = blockIdx.x;
  const uint32_t n = min(cell_n[q], kMcCells);
  for (uint32_t i = tid; i < bm_words; i += 1024) bm[i] = 0u;
  uint32_t len = 0;
  if (tid < n) {
    const uint32_t c = cells[(size_t)q * cell_cap + tid];
    len = (uint32_t)ivf_len[c];
    coff[tid] = ivf_off[c];
  }
  uint32_t total = 0;
  const uint32_t ex = pl_block_scan(len, warp_sums, &total);
  if (tid < n) cpfx[tid] = ex;
  if (tid == 0) cpfx[n] = total;
  __syncthreads();
  constexpr uint32_t U = 8;
  uint32_t ci = 0;
  for (uint32_t e0 = tid; e0 < total; e0 += 1024 * U) {
    uint32_t pid[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t e = e0 + u * 1024;
      if (e < total) {
        while (cpfx[ci + 1] <= e) ++ci;
        pid[u] = (uint32_t)ivf[coff[ci] + (e - cpfx[ci])];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u)
      if (e0 + u * 1024 < total) atomicOr(&bm[pid[u] >> 5], 1u << (pid[u] & 31));
  }
  __syncthreads();
  uint32_t *cq = cand + (size_t)q * cand_cap;
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < bm_words; base += 1024) {
    const uint32_t k = base + tid;
    const uint32_t w = bm_words - 1 - k;  // descending words
    uint32_t word = k < bm_words ? bm[w] : 0u;
    uint32_t sum;
    uint32_t pos = cnt + pl_block_scan((uint32_t)__popc(word), warp_sums, &sum);
    while (word) {
      const uint32_t b = 31u - (uint32_t)__clz(word);
      word &= ~(1u << b);
      if (pos < cand_cap) cq[pos] = w * 32 + b;
      ++pos;
    }
    cnt += sum;
  }
  if (tid == 0) cand_n[q] = min(cnt, cand_cap);
}

// seg_cnt[q][sg] = number of keys in segment sg when the query has n keys
// laid out flat (segment sg = keys [sg*kSeg, (sg+1)*kSeg)).
__global__ void k_pl_segcnt(const uint32_t *__restrict__ n_in, uint32_t n_fixed, uint32_t n_seg,
                            uint32_t n_q, uint32_t *__restrict__ seg_cnt) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (uint64_t)n_q * n_seg) return;
  const uint32_t q = (uint32_t)(e / n_seg), sg = (uint32_t)(e % n_seg);
  const uint32_t n = n_in ? n_in[q] : n_fixed;
  const uint32_t lo = sg * kSeg;
  seg_cnt[e] = n > lo ? min(n - lo, kSeg) : 0u;
}

// Segment counts of the split layout: heavy keys in segments 0 .. n_heavy-1
// (heavy_n of them, flat), light keys (light_n <= ndocs of them, flat) in
// the segments from n_heavy on.
__global__ void k_pl_segcnt_split(const uint32_t *__restrict__ heavy_n, const uint32_t *__restrict__ light_n,
                                  uint32_t n_heavy, uint32_t n_seg, uint32_t n_q, uint32_t *__restrict__ seg_cnt) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (uint64_t)n_q * n_seg) return;
  const uint32_t q = (uint32_t)(e / n_seg), sg = (uint32_t)(e % n_seg);
  const bool light = sg >= n_heavy;
  const uint32_t n = light ? light_n[q] : heavy_n[q], lo = (light ? sg - n_heavy : sg) * kSeg;
  seg_cnt[e] = n > lo ? min(n - lo, kSeg) : 0u;
}

// Light-key segments of the split layout: the stage-1 top-ndocs takes at most
// min(ndocs, n_doc) of them.
uint32_t pl_light_segs(uint32_t ndocs, uint32_t n_doc) {
  return std::max(1u, (std::min(ndocs, n_doc) + kSeg - 1) / kSeg);
}

// ---------------------------------------------------------------------------
// Filter stages (filter_pids.cpp:30-80): one wave per pid.  The wave walks
// the doc's codes 64 at a time; a code that passes idx (stage 1) and was not
// seen before in this doc (LDS bitmap, one per wave) is new; new codes are
// consumed two at a time, lanes (t, h) taking max with S[code_h][t].
// key = (score order bits << 32) | pid: larger keys first = (score, pid)
// descending, the priority_queue order.  kInv: input ids are selection
// outputs (~pid).
// ---------------------------------------------------------------------------
template <bool kIdx, bool kInv>
__global__ __launch_bounds__(256) void k_pl_approx(
    const uint32_t *__restrict__ in_ids, const uint32_t *__restrict__ in_n, uint32_t n_fixed,
    uint64_t in_stride, const uint32_t *__restrict__ slots, const uint32_t *__restrict__ slot_n,
    const float *__restrict__ S, uint64_t s_stride, const uint8_t *__restrict__ idx,
    uint64_t idx_stride, uint32_t n_c, uint32_t nq, uint32_t n_doc, const int32_t *__restrict__ codes,
    const int64_t *__restrict__ doclens, const int64_t *__restrict__ offsets, uint32_t *__restrict__ pid_bm,
    uint32_t bm_words, uint32_t waves_per_block, unsigned long long *__restrict__ keys, uint64_t key_stride,
    unsigned long long *__restrict__ nvec, int f16, float empty_score) {
  extern __shared__ uint32_t dyn_bm[];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave >= waves_per_block) return;
  const uint32_t t = lane & 31, h = lane >> 5, q = blockIdx.y;
  const uint32_t words = (n_c + 31) / 32;
  uint32_t *my = dyn_bm + (size_t)wave * words;
  for (uint32_t w = lane; w < words; w += 64) my[w] = 0;
  pl_wave_sync();
  // with slots: visit only the listed input positions (slots[q][j], j < slot_n[q]);
  // the key of the j-th visited input goes to keys[q][j]
  const uint32_t n = slots ? slot_n[q] : (in_n ? in_n[q] : n_fixed);
  const float *Sq = S + (size_t)q * s_stride;
  const uint8_t *iq = idx ? idx + (size_t)q * idx_stride : nullptr;
  unsigned long long nv = 0;
  for (uint32_t jj = blockIdx.x * waves_per_block + wave; jj < n; jj += gridDim.x * waves_per_block) {
    const uint32_t i = slots ? slots[(size_t)q * in_stride + jj] : jj;
    const uint32_t raw = in_ids[(size_t)q * in_stride + i];
    const uint32_t pid = kInv ? ~raw : raw;
    const bool ok = pid < n_doc;
    if (pid_bm && ok && lane == 0) pid_bm[(size_t)q * bm_words + (pid >> 5)] = 0;
    const uint32_t L = ok ? (uint32_t)doclens[pid] : 0u;
    const int64_t off = ok ? offsets[pid] : 0;
    float m = f16 ? kPad16 : -9999.0f;
    uint32_t nu = 0;  // distinct passing codes (wave-uniform)
    for (uint32_t base = 0; base < L; base += 64) {
      const uint32_t j = base + lane;
      uint32_t code = 0;
      bool pass = false;
      if (j < L) {
        code = (uint32_t)codes[off + j];
        pass = code < n_c && (!kIdx || iq[code]);
      }
      bool isnew = false;
      if (pass) {
        const uint32_t bit = 1u << (code & 31);
        isnew = !(atomicOr(&my[code >> 5], bit) & bit);
      }
      uint64_t mask = __ballot(isnew);
      nu += (uint32_t)__popcll(mask);
      // up to 8 new codes per round (lane half h takes every other one), the
      // S-row loads of a round issued before any is used
      while (mask) {
        uint32_t cc[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
          uint32_t l0 = 64, l1 = 64;
          if (mask) {
            l0 = (uint32_t)__ffsll((unsigned long long)mask) - 1;
            mask &= mask - 1;
            l1 = l0;
            if (mask) {
              l1 = (uint32_t)__ffsll((unsigned long long)mask) - 1;
              mask &= mask - 1;
            }
          }
          // an empty slot repeats the first code of the round (max is idempotent)
          const uint32_t c0 = l0 < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)code, l0) : cc[0];
          const uint32_t c1 = l1 < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)code, l1) : cc[0];
          cc[k] = h ? c1 : c0;
        }
        if (t < nq) {
          float v[4];
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) v[k] = Sq[(size_t)cc[k] * nq + t];
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) m = pl_max(m, v[k]);
        }
      }
    }
    // a doc without passing codes keeps every maximum at the identity: its
    // score is empty_score (CPU branch: the exact sum -9999 * nq; GPU branch:
    // half of the FP32 sum of nq x -10000, -inf from nq = 7 on), and no bit
    // was set
    float score = empty_score;
    if (nu) {
      pl_wave_sync();
      for (uint32_t base = 0; base < L; base += 64) {
        const uint32_t j = base + lane;
        if (j < L) {
          const uint32_t code = (uint32_t)codes[off + j];
          if (code < n_c) my[code >> 5] = 0;
        }
      }
      pl_wave_sync();
      m = pl_max(m, __shfl_xor(m, 32));
      score = 0.0f;
      for (uint32_t tt = 0; tt < nq; ++tt) score += pl_readlane(m, tt);
      if (f16) {
        // the half tensor's sum (colbert.py:155); a tiny negative sum rounds
        // to -0, which torch.topk ranks equal to +0: one key for both
        score = r16(score);
        score = score == 0.0f ? 0.0f : score;
      }
    }
    if (lane == 0) {
      keys[(size_t)q * key_stride + jj] = ((unsigned long long)pl_f2ord(score) << 32) | pid;
      nv += nu;
    }
  }
  if (lane == 0 && nv) atomicAdd(&nvec[q], nv * nq);
}

// Docs holding an idx-passing code are exactly the docs in the IVF lists of
// the passing centroids (a doc is in IVF[c] iff one of its codes is c).  One
// workgroup per query lists the passing centroids and, when their lists hold
// fewer entries than the query's candidates hold tokens, marks those docs in
// pass_bm (pmode = 1); otherwise every candidate is scanned (pmode = 0).
constexpr uint32_t kPassCap = 1024;

__global__ __launch_bounds__(256) void k_pl_passmark(
    const uint8_t *__restrict__ idx, uint32_t n_c, const int32_t *__restrict__ ivf,
    const int64_t *__restrict__ ivf_off, const int64_t *__restrict__ ivf_len, const uint32_t *__restrict__ cand_n,
    float avg_len, uint32_t *__restrict__ pass_list, uint32_t *__restrict__ pass_cnt, uint32_t *__restrict__ pmode,
    uint32_t *__restrict__ pass_bm, uint32_t bm_words) {
  __shared__ uint32_t s_list[kPassCap];
  __shared__ uint32_t s_n;
  __shared__ unsigned long long s_len;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, q = blockIdx.x;
  if (tid == 0) { s_n = 0; s_len = 0; }
  __syncthreads();
  const uint8_t *iq = idx + (size_t)q * n_c;
  for (uint32_t c = tid; c < n_c; c += 256) {
    if (iq[c]) {
      const uint32_t slot = atomicAdd(&s_n, 1u);
      if (slot < kPassCap) s_list[slot] = c;
      atomicAdd(&s_len, (unsigned long long)ivf_len[c]);
    }
  }
  __syncthreads();
  const uint32_t n = s_n;
  const bool use = n <= kPassCap && (double)s_len <= (double)cand_n[q] * (double)avg_len;
  if (tid == 0) {
    pmode[q] = use ? 1u : 0u;
    pass_cnt[q] = use ? n : 0u;
  }
  if (!use) return;
  for (uint32_t i = tid; i < n; i += 256) pass_list[(size_t)q * kPassCap + i] = s_list[i];
  uint32_t *bq = pass_bm + (size_t)q * bm_words;
  for (uint32_t ci = wave; ci < n; ci += 4) {
    const uint32_t c = s_list[ci];
    const int64_t off = ivf_off[c], len = ivf_len[c];
    for (int64_t j = lane; j < len; j += 64) {
      const uint32_t pid = (uint32_t)ivf[off + j];
      atomicOr(&bq[pid >> 5], 1u << (pid & 31));
    }
  }
}

// Zeroes the words k_pl_passmark set.
__global__ __launch_bounds__(256) void k_pl_passclear(const uint32_t *__restrict__ pass_list,
                                                      const uint32_t *__restrict__ pass_cnt,
                                                      const int32_t *__restrict__ ivf,
                                                      const int64_t *__restrict__ ivf_off,
                                                      const int64_t *__restrict__ ivf_len,
                                                      uint32_t *__restrict__ pass_bm, uint32_t bm_words) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = blockIdx.x;
  const uint32_t n = pass_cnt[q];
  uint32_t *bq = pass_bm + (size_t)q * bm_words;
  for (uint32_t ci = wave; ci < n; ci += 4) {
    const uint32_t c = pass_list[(size_t)q * kPassCap + ci];
    const int64_t off = ivf_off[c], len = ivf_len[c];
    for (int64_t j = lane; j < len; j += 64) bq[(uint32_t)ivf[off + j] >> 5] = 0;
  }
}

// Stage-1 split over the descending-pid candidate list, one workgroup of
// 1024 threads per query.  Candidates holding a passing code (all of them
// when pmode is 0) are listed in heavy[q] for the wave scan.  The others keep
// every maximum at the identity, so their score is empty_score (below every
// other) and they rank by pid: only the first ndocs of them (the largest
// pids) can reach the stage-1 top-ndocs, and their keys go to
// keys[q][light_base + r].
__global__ __launch_bounds__(1024) void k_pl_split(
    const uint32_t *__restrict__ cand, uint64_t cand_stride, const uint32_t *__restrict__ cand_n,
    const uint32_t *__restrict__ pmode, const uint32_t *__restrict__ pass_bm, uint32_t bm_words, uint32_t nq,
    uint32_t ndocs, unsigned long long *__restrict__ keys, uint64_t key_stride, uint64_t light_base,
    uint32_t *__restrict__ heavy, uint32_t *__restrict__ heavy_n, uint32_t *__restrict__ light_n,
    float empty_score) {
  __shared__ uint32_t warp_sums[16];
  const uint32_t tid = threadIdx.x, q = blockIdx.x;
  const uint32_t n = cand_n[q];
  const bool mark = pmode[q] != 0;
  const unsigned long long ck = (unsigned long long)pl_f2ord(empty_score) << 32;
  const uint32_t *bq = pass_bm + (size_t)q * bm_words;
  uint32_t h_tot = 0, l_tot = 0;
  for (uint32_t base = 0; base < n; base += 1024) {
    const uint32_t i = base + tid;
    const bool v = i < n;
    uint32_t pid = 0;
    bool hv = false;
    if (v) {
      pid = cand[(size_t)q * cand_stride + i];
      hv = !mark || ((bq[pid >> 5] >> (pid & 31)) & 1u);
    }
    // heavy count in the low 16 bits, light count in the high 16 (<= 1024 each)
    uint32_t sum;
    const uint32_t ex = pl_block_scan(v ? (hv ? 1u : 0x10000u) : 0u, warp_sums, &sum);
    if (hv) heavy[(size_t)q * cand_stride + h_tot + (ex & 0xFFFFu)] = i;
    if (v && !hv) {
      const uint32_t r = l_tot + (ex >> 16);
      if (r < ndocs) keys[(size_t)q * key_stride + light_base + r] = ck | pid;
    }
    h_tot += sum & 0xFFFFu;
    l_tot += sum >> 16;
  }
  if (tid == 0) {
    heavy_n[q] = h_tot;
    light_n[q] = min(l_tot, ndocs);
  }
}

__global__ void k_pl_unkey(const uint32_t *__restrict__ sel, const uint32_t *__restrict__ n_sel,
                           int32_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_sel[0]) out[i] = (int32_t)~sel[i];
}

// ---------------------------------------------------------------------------
// Refine (index_storage.py:164-190, CPU branch): per surviving doc, decompress
// (x = w[bucket] + C[code]), y = x / max(||x||, 1e-12), s_t = FMA chain of
// y . Q_t, M_t = max(0, max_i s_t), score = sum over t left to right.
// Wave layout as the IGP rerank: lane (t, h) holds Q_t (128 floats) and the
// wave takes 16 doc tokens per step: every lane decodes its 4 dims of 8
// tokens and divides them by the token's denominator, then runs eight dot
// chains.  The denominator max(||x||, 1e-12) depends on the token alone, so it
// is computed once per index (k_pl_token_den, same FMA chain) and read here.
// tab4[v] = the four weights of raw PLAID byte v (rbm, lut and bw folded).
// ---------------------------------------------------------------------------
constexpr uint32_t kPlTok = 16;

// den[tok] = max(sqrtf(FMA chain over d of x_d^2), 1e-12), x the FP32
// decompressed token (index_storage.py:182, residual.py:273-274).  kF16 (the
// GPU branch, residual.py:255-274 on half tensors): x_d = half(w + c) from
// half weights and centroids, den = half(sqrtf(chain)) (eps 1e-12 rounds to
// 0 in half).
template <bool kF16>
__global__ __launch_bounds__(256) void k_pl_token_den(const int32_t *__restrict__ codes, const uint8_t *__restrict__ res,
                                                      uint64_t n_tok, const float *__restrict__ C,
                                                      const float4 *__restrict__ tab4, float *__restrict__ den) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_tok; g += stride) {
    const float4 *c4 = reinterpret_cast<const float4 *>(C + (size_t)(uint32_t)codes[g] * kDim);
    const uint8_t *rb = res + g * kResBytes;
    float n2 = 0.0f;
#pragma unroll 4
    for (uint32_t j = 0; j < kResBytes; ++j) {
      const float4 w = tab4[rb[j]];
      const float4 cv = c4[j];
      float x0 = w.x + cv.x, x1 = w.y + cv.y, x2 = w.z + cv.z, x3 = w.w + cv.w;
      if (kF16) {
        x0 = r16(x0); x1 = r16(x1); x2 = r16(x2); x3 = r16(x3);
      }
      n2 = __builtin_fmaf(x0, x0, n2);
      n2 = __builtin_fmaf(x1, x1, n2);
      n2 = __builtin_fmaf(x2, x2, n2);
      n2 = __builtin_fmaf(x3, x3, n2);
    }
    // sqrtf and '/' compile to correctly rounded sequences (HIP's default
    // -fhip-fp32-correctly-rounded-divide-sqrt); __fsqrt_rn does not
    den[g] = kF16 ? r16(sqrtf(n2)) : fmaxf(sqrtf(n2), 1e-12f);
  }
}

// kF16: the GPU branch (index_storage.py:160-161,185-187): half tokens
// y_d = half(half(w + c) / den) (lookup_pids, residual.py:255-274), s_t =
// half(FMA chain of y . Q_t) with Q rounded to half (colbert.py:198), M_t =
// max_i s_t with the padding value -10000 as identity (no zero clip,
// colbert_score_reduce), score = half(left-to-right FP32 sum over t).
template <bool kF16>
__global__ __launch_bounds__(256) void k_pl_refine(
    const float *__restrict__ Q, uint32_t nq, const uint32_t *__restrict__ sel_ids,
    const uint32_t *__restrict__ sel_n, uint32_t R, uint32_t n_slice, uint32_t n_doc,
    const int64_t *__restrict__ doclens, const int64_t *__restrict__ offsets,
    const int32_t *__restrict__ codes, const uint8_t *__restrict__ res, const float *__restrict__ C,
    const float4 *__restrict__ tab4, const float *__restrict__ den, uint32_t *__restrict__ out_ids,
    float *__restrict__ out_score) {
  __shared__ __align__(16) float4 xs[4][kPlTok][kDim / 4];
  __shared__ float Ms[4][kMaxQTok];
  __shared__ float4 tab[256];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, t = lane & 31;
  const uint32_t slice = blockIdx.x, q = blockIdx.y;
  tab[tid] = tab4[tid];
  __syncthreads();
  const uint32_t n = sel_n[q];
  float qv[kDim];
  {
    const float4 *q4 = reinterpret_cast<const float4 *>(Q + ((size_t)q * kMaxQTok + (t < nq ? t : 0)) * kDim);
#pragma unroll
    for (uint32_t i = 0; i < kDim / 4; ++i) {
      const float4 v = (t < nq) ? q4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      qv[4 * i] = v.x; qv[4 * i + 1] = v.y; qv[4 * i + 2] = v.z; qv[4 * i + 3] = v.w;
    }
  }
  for (uint32_t r = slice * 4 + wave; r < n; r += n_slice * 4) {
    const uint32_t pid = ~sel_ids[(size_t)q * R + r];
    const uint32_t L = pid < n_doc ? (uint32_t)doclens[pid] : 0u;
    const int64_t off = pid < n_doc ? offsets[pid] : 0;
    float M = kF16 ? kPad16 : 0.0f;
    for (uint32_t i = 0; i < L; i += kPlTok) {
#pragma unroll
      for (uint32_t c = 0; c < kPlTok / 2; ++c) {
        const uint32_t tk = 2 * c + h;
        if (i + tk < L) {
          const uint64_t g = (uint64_t)off + i + tk;
          const uint32_t code = (uint32_t)codes[g];
          const float4 w = tab[res[g * kResBytes + t]];
          const float4 cv = reinterpret_cast<const float4 *>(C + (size_t)code * kDim)[t];
          const float dn = den[g];
          if (kF16)
            xs[wave][tk][t] = make_float4(r16(r16(w.x + cv.x) / dn), r16(r16(w.y + cv.y) / dn),
                                          r16(r16(w.z + cv.z) / dn), r16(r16(w.w + cv.w) / dn));
          else
            xs[wave][tk][t] = make_float4((w.x + cv.x) / dn, (w.y + cv.y) / dn, (w.z + cv.z) / dn, (w.w + cv.w) / dn);
        }
      }
      pl_wave_sync();
      float a[kPlTok / 2];
#pragma unroll
      for (uint32_t c = 0; c < kPlTok / 2; ++c) a[c] = 0.f;
#pragma unroll
      for (uint32_t d4 = 0; d4 < kDim / 4; ++d4) {
#pragma unroll
        for (uint32_t c = 0; c < kPlTok / 2; ++c) {
          const float4 xv = xs[wave][2 * c + h][d4];
          a[c] = __builtin_fmaf(xv.x, qv[4 * d4 + 0], a[c]);
          a[c] = __builtin_fmaf(xv.y, qv[4 * d4 + 1], a[c]);
          a[c] = __builtin_fmaf(xv.z, qv[4 * d4 + 2], a[c]);
          a[c] = __builtin_fmaf(xv.w, qv[4 * d4 + 3], a[c]);
        }
      }
#pragma unroll
      for (uint32_t c = 0; c < kPlTok / 2; ++c)
        if (i + 2 * c + h < L) M = pl_max(M, kF16 ? r16(a[c]) : a[c]);
      pl_wave_sync();
    }
    M = pl_max(M, __shfl_xor(M, 32));
    if (h == 0) Ms[wave][t] = M;
    pl_wave_sync();
    if (lane == 0) {
      float sc = 0.0f;
      for (uint32_t tt = 0; tt < nq; ++tt) sc += Ms[wave][tt];
      // GPU branch: the half sum, with -0 reported as +0 so that the two
      // zeros rank as the tie they are
      if (kF16) {
        sc = r16(sc);
        sc = sc == 0.0f ? 0.0f : sc;
      }
      out_score[(size_t)q * R + r] = sc;
      out_ids[(size_t)q * R + r] = pid;
    }
    pl_wave_sync();
  }
}

// ---------------------------------------------------------------------------
// GPU-branch single ops (float16)
// ---------------------------------------------------------------------------
// out[i] = in[i] rounded to half (in place allowed)
__global__ __launch_bounds__(256) void k_pl_round16(const float *in, uint64_t n, float *out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = r16(in[i]);
}

// *bad = 1 when some length is negative
__global__ __launch_bounds__(256) void k_pl_check_nonneg(const int64_t *__restrict__ v, uint32_t n,
                                                         uint32_t *__restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && v[i] < 0) *bad = 1u;
}

// *bad = 1 when some id lies outside [0, n)
__global__ __launch_bounds__(256) void k_pl_check_ids(const int32_t *__restrict__ ids, uint32_t n_ids, uint32_t n,
                                                      uint32_t *__restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_ids && (ids[i] < 0 || (uint32_t)ids[i] >= n)) *bad = 1u;
}

// lookup_pids of the GPU branch (residual_embeddings_strided.py:30-37 +
// ResidualCodec.decompress, residual.py:243-279): the tokens of pids in
// order, row cum[i] + j = token j of pids[i], each y_d = half(half(w + c) /
// den16) with half weights (tab4h), half centroids (C16) and the token's half
// norm.  Workgroup per pid, 32 lanes per token (four dims each), eight tokens
// per step.
__global__ __launch_bounds__(256) void k_pl_lookup16(const int32_t *__restrict__ pids, uint32_t npids,
                                                     const int64_t *__restrict__ cum,
                                                     const int64_t *__restrict__ doclens,
                                                     const int64_t *__restrict__ offsets,
                                                     const int32_t *__restrict__ codes, const uint8_t *__restrict__ res,
                                                     const float *__restrict__ C16, const float4 *__restrict__ tab4h,
                                                     const float *__restrict__ den16, uint16_t *__restrict__ out) {
  const uint32_t t = threadIdx.x & 31, slot = threadIdx.x >> 5;
  for (uint32_t i = blockIdx.x; i < npids; i += gridDim.x) {
    const int32_t pid = pids[i];
    const int64_t L = doclens[pid], off = offsets[pid], row0 = cum[i];
    for (int64_t j = slot; j < L; j += 8) {
      const uint64_t g = (uint64_t)(off + j);
      const float4 w = tab4h[res[g * kResBytes + t]];
      const float4 cv = reinterpret_cast<const float4 *>(C16 + (size_t)(uint32_t)codes[g] * kDim)[t];
      const float dn = den16[g];
      ushort4 o;
      o.x = f2h(r16(w.x + cv.x) / dn);
      o.y = f2h(r16(w.y + cv.y) / dn);
      o.z = f2h(r16(w.z + cv.z) / dn);
      o.w = f2h(r16(w.w + cv.w) / dn);
      reinterpret_cast<ushort4 *>(out + (size_t)(row0 + j) * kDim)[t] = o;
    }
  }
}

// colbert_score_packed of the GPU branch (colbert.py:181-205): one wave per
// doc; lane (t, h) holds half query row t in registers and scores the doc's
// rows of parity h: s = half(FMA chain over d of D . Q_t), M_t = max with
// the padding value -10000 as identity, score = half(left-to-right FP32 sum
// over t).  D f16 [cum[ndocs]][128], Q16 f16 [nq][128] (nq <= 32).
__global__ __launch_bounds__(256) void k_pl_score_packed16(const uint16_t *__restrict__ Q16, uint32_t nq,
                                                           const uint16_t *__restrict__ D,
                                                           const int64_t *__restrict__ cum, uint32_t ndocs,
                                                           uint16_t *__restrict__ out) {
  __shared__ float Ms[4][kMaxQTok];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, t = lane & 31, h = lane >> 5;
  float qv[kDim];
#pragma unroll
  for (uint32_t d = 0; d < kDim; ++d) qv[d] = t < nq ? h2f(Q16[(size_t)t * kDim + d]) : 0.0f;
  for (uint32_t doc = blockIdx.x * 4 + wave; doc < ndocs; doc += gridDim.x * 4) {
    const int64_t r0 = cum[doc], r1 = cum[doc + 1];
    float M = kPad16;
    for (int64_t r = r0 + h; r < r1; r += 2) {
      const uint4 *row = reinterpret_cast<const uint4 *>(D + (size_t)r * kDim);
      float a = 0.0f;
#pragma unroll
      for (uint32_t k = 0; k < kDim / 8; ++k) {
        const uint4 v = row[k];
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
          a = __builtin_fmaf(h2f((uint16_t)(w4[e] & 0xFFFFu)), qv[8 * k + 2 * e], a);
          a = __builtin_fmaf(h2f((uint16_t)(w4[e] >> 16)), qv[8 * k + 2 * e + 1], a);
        }
      }
      M = pl_max(M, r16(a));
    }
    M = pl_max(M, __shfl_xor(M, 32));
    if (h == 0) Ms[wave][t] = M;
    pl_wave_sync();
    if (lane == 0) {
      float sc = 0.0f;
      for (uint32_t tt = 0; tt < nq; ++tt) sc += Ms[wave][tt];
      out[doc] = f2h(sc);
    }
    pl_wave_sync();
  }
}

// colbert_score_reduce of the GPU branch on the padded form (colbert.py:
// 133-155): scores [n][L][nq] f16, mask [n][L]; the masked-out entries read
// as -10000, M_t = max over the L positions, out = half(left-to-right FP32
// sum over t).  One wave per doc, lane t < nq (nq <= 64).
__global__ __launch_bounds__(256) void k_pl_score_reduce16(const uint16_t *__restrict__ sp, const uint8_t *__restrict__ mask,
                                                           uint32_t n, uint32_t L, uint32_t nq,
                                                           uint16_t *__restrict__ out) {
  __shared__ float Ms[4][64];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t doc = blockIdx.x * 4 + wave; doc < n; doc += gridDim.x * 4) {
    float M = -INFINITY;
    if (lane < nq)
      for (uint32_t j = 0; j < L; ++j) {
        const float v = mask[(size_t)doc * L + j] ? h2f(sp[((size_t)doc * L + j) * nq + lane]) : kPad16;
        M = j ? pl_max(M, v) : v;
      }
    Ms[wave][lane] = M;
    pl_wave_sync();
    if (lane == 0) {
      float sc = 0.0f;
      for (uint32_t tt = 0; tt < nq; ++tt) sc += Ms[wave][tt];
      out[doc] = f2h(sc);
    }
    pl_wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
thread_local std::string g_pl_err;

int pl_fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_pl_err = buf;
  return code;
}

#define PL_TRY(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      return pl_fail(MVR_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                     __FILE__, __LINE__);                                                   \
  } while (0)

struct PBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    const size_t b = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) cap = b;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

inline uint32_t grid_for(uint64_t n, uint32_t block) {
  const uint64_t g = (n + block - 1) / block;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 20));
}

// waves per workgroup of k_pl_approx for a dedup bitmap of n_c bits
inline uint32_t approx_waves(uint32_t n_c) {
  const uint32_t bytes = ((n_c + 31) / 32) * 4;
  return std::max(1u, std::min(4u, kPlMaxBmBytes / std::max(bytes, 4u)));
}

// Scratch of the single ops (per host thread, reallocated on a device change).
struct OpScratch {
  int device = -1;
  PBuf cum, keys, seg, sel1_ids, sel1_F, sel1_n, keys2, sel2_ids, sel2_F, sel2_n, nvec, flag;
  hipError_t bind() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev != device) {
      for (PBuf *b : {&cum, &keys, &seg, &sel1_ids, &sel1_F, &sel1_n, &keys2, &sel2_ids, &sel2_F, &sel2_n,
                      &nvec, &flag})
        b->release();
      device = dev;
    }
    return hipSuccess;
  }
};
thread_local OpScratch g_ops;

// Launch the two filter stages for n_q queries.  Stage-1 input: pids
// (in_ids, in_n / n_fixed, row stride in_stride) with their S, idx rows.
// slots / slot_n / light_n (nullable): stage 1 scans only the listed input
// positions, writing their keys flat from keys1[q][0]; light_n[q] more keys
// already sit in the extra last segment (k_pl_split).
hipError_t launch_filter_stages(const uint32_t *in_ids, const uint32_t *in_n, uint32_t n_fixed,
                                uint64_t in_stride, uint32_t in_cap, const uint32_t *slots,
                                const uint32_t *slot_n, const uint32_t *light_n, const float *S,
                                const uint8_t *idx,
                                uint32_t n_c, uint32_t nq, uint32_t n_doc, const int32_t *codes,
                                const int64_t *doclens, const int64_t *offsets, uint32_t *pid_bm,
                                uint32_t bm_words, uint32_t n_q, uint32_t ndocs, uint64_t *keys1,
                                uint32_t *seg1, uint32_t *sel1_ids, float *sel1_F, uint32_t *sel1_n,
                                uint64_t *keys2, uint32_t *sel2_ids, float *sel2_F, uint32_t *sel2_n,
                                unsigned long long *nvec, int f16, float empty_score, hipStream_t s) {
  // with slots (split layout): heavy keys flat from 0, light keys in the
  // pl_light_segs segments after them
  const uint32_t n_heavy = std::max(1u, (in_cap + kSeg - 1) / kSeg);
  const uint32_t n_seg1 = n_heavy + (slots ? pl_light_segs(ndocs, in_cap) : 0u);
  const uint32_t wpb = approx_waves(n_c);
  const size_t lds = (size_t)wpb * ((n_c + 31) / 32) * 4;
  // about 16k workgroups: one wave per doc is latency-bound, so many in flight
  const uint32_t gx = std::max(1u, std::min(64u, 16384u / std::max(1u, n_q)));
  dim3 grid(gx, n_q);
  k_pl_approx<true, false><<<grid, 256, lds, s>>>(in_ids, in_n, n_fixed, in_stride, slots, slot_n, S,
                                                  (uint64_t)n_c * nq,
                                                  idx, n_c, n_c, nq, n_doc, codes, doclens, offsets, pid_bm,
                                                  bm_words, wpb, (unsigned long long *)keys1,
                                                  (uint64_t)n_seg1 * kSeg, nvec, f16, empty_score);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (slots)
    k_pl_segcnt_split<<<grid_for((uint64_t)n_q * n_seg1, 256), 256, 0, s>>>(slot_n, light_n, n_heavy, n_seg1, n_q,
                                                                             seg1);
  else
    k_pl_segcnt<<<grid_for((uint64_t)n_q * n_seg1, 256), 256, 0, s>>>(in_n, n_fixed, n_seg1, n_q, seg1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_select_segments(keys1, seg1, n_seg1, nullptr, n_q, ndocs, sel1_ids, sel1_F, sel1_n, nullptr,
                                  s)) != hipSuccess)
    return e;
  // stage 2: one key per stage-1 pid, contiguous per query over n_seg2
  // segments (ndocs may exceed one segment); the stage-1 counts buffer
  // (n_seg1 >= n_seg2 per query) is reused for the segment counts
  const uint32_t n_seg2 = std::max(1u, (std::min(ndocs, in_cap) + kSeg - 1) / kSeg);
  k_pl_approx<false, true><<<grid, 256, lds, s>>>(sel1_ids, sel1_n, 0, ndocs, nullptr, nullptr, S,
                                                  (uint64_t)n_c * nq, nullptr,
                                                  0, n_c, nq, n_doc, codes, doclens, offsets, nullptr, 0, wpb,
                                                  (unsigned long long *)keys2, (uint64_t)n_seg2 * kSeg, nvec, f16,
                                                  empty_score);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_pl_segcnt<<<grid_for((uint64_t)n_q * n_seg2, 256), 256, 0, s>>>(sel1_n, 0, n_seg2, n_q, seg1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_select_segments(keys2, seg1, n_seg2, nullptr, n_q, ndocs / 4, sel2_ids, sel2_F, sel2_n, nullptr, s);
}

}  // namespace
}  // namespace mvr

using namespace mvr;

// ---------------------------------------------------------------------------
// Batched search handle
// ---------------------------------------------------------------------------
struct mvr_plaid_index {
  int device = 0;
  uint64_t pool_budget = kPoolBudget;  // device bytes for a pass's per-query workspace
  bool markc = false;  // the pid bitmap fits one workgroup's LDS: k_pl_markc
  uint32_t n_doc = 0, n_c = 0;
  uint64_t n_tok = 0, n_ivf = 0;
  uint32_t bm_words = 0;
  float *C = nullptr;
  int32_t *codes = nullptr;
  uint8_t *res = nullptr;
  int64_t *doclens = nullptr, *offsets = nullptr;
  int32_t *ivf = nullptr;
  int64_t *ivf_off = nullptr, *ivf_len = nullptr;
  float4 *tab4 = nullptr;
  float *den = nullptr;  // [n_tok] per-token normalisation denominators
  // GPU branch (half precision), built on first use: centroids and the byte
  // table rounded to half, per-token half norms
  float *C16 = nullptr;
  float4 *tab4h = nullptr;
  float *den16 = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool bm_dirty = true;  // pid bitmap must be cleared before the next pass
  uint32_t q_nq = 0;     // query rows last copied into the padded Q buffer
  float avg_len = 0.0f;  // mean doc length (tokens)
  PBuf part_keys, top_ids, top1, pass_list, pass_cnt, pmode, pass_bm, heavy, heavy_n, light_n;
  PBuf Q, S, idx, cells, cell_n, pid_bm, cand, cand_n, keys1, seg1, sel1_ids, sel1_F, sel1_n, keys2, sel2_ids,
      sel2_F, sel2_n, ids2, score2, out_ids, out_s, nvec;
  std::vector<PBuf *> bufs() {
    return {&part_keys, &top_ids, &top1, &pass_list, &pass_cnt, &pmode, &pass_bm, &heavy, &heavy_n, &light_n, &Q, &S, &idx, &cells, &cell_n, &pid_bm, &cand, &cand_n, &keys1, &seg1, &sel1_ids, &sel1_F, &sel1_n,
            &keys2, &sel2_ids, &sel2_F, &sel2_n, &ids2, &score2, &out_ids, &out_s, &nvec};
  }
};

namespace {
// half rounding on the host (ties to even)
float host_r16(float f) { return (float)(_Float16)f; }

// The GPU branch's half-precision arrays (ResidualCodec holds centroids and
// bucket weights as halves on a CUDA device, residual.py:30-45): built once
// per index on first use.
int pl_prepare_f16(mvr_plaid_index *ix) {
  if (ix->den16) return MVR_OK;
  hipStream_t s = ix->stream;
  hipError_t e;
  if ((e = hipMalloc((void **)&ix->C16, sizeof(float) * (size_t)ix->n_c * kDim)) != hipSuccess ||
      (e = hipMalloc((void **)&ix->tab4h, sizeof(float4) * 256)) != hipSuccess) {
    for (void **pp : {(void **)&ix->C16, (void **)&ix->tab4h})
      if (*pp) { (void)hipFree(*pp); *pp = nullptr; }
    return pl_fail(MVR_ERR_DEVICE, "half arrays: %s", hipGetErrorString(e));
  }
  k_pl_round16<<<grid_for((uint64_t)ix->n_c * kDim, 256), 256, 0, s>>>(ix->C, (uint64_t)ix->n_c * kDim, ix->C16);
  k_pl_round16<<<4, 256, 0, s>>>(reinterpret_cast<const float *>(ix->tab4), 1024, reinterpret_cast<float *>(ix->tab4h));
  PL_TRY(hipGetLastError());
  float *den16 = nullptr;
  PL_TRY(hipMalloc((void **)&den16, sizeof(float) * std::max<uint64_t>(ix->n_tok, 1)));
  if (ix->n_tok)
    k_pl_token_den<true><<<grid_for(ix->n_tok, 256), 256, 0, s>>>(ix->codes, ix->res, ix->n_tok, ix->C16, ix->tab4h,
                                                                 den16);
  if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess) {
    (void)hipFree(den16);
    return pl_fail(MVR_ERR_DEVICE, "half token norms: %s", hipGetErrorString(e));
  }
  ix->den16 = den16;
  return MVR_OK;
}
}  // namespace

extern "C" {

const char *mvr_plaid_last_error(void) { return g_pl_err.c_str(); }

int mvr_plaid_lookup_pids_f16(mvr_plaid_index *ix, const int32_t *pids, uint32_t npids, uint16_t *out,
                              uint64_t out_rows, uint64_t *n_rows, void *stream) {
  if (!ix || !n_rows) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  *n_rows = 0;
  if (npids == 0) return MVR_OK;
  if (!pids) return pl_fail(MVR_ERR_INVALID, "pids is NULL");
  PL_TRY(hipSetDevice(ix->device));
  int rc = pl_prepare_f16(ix);
  if (rc != MVR_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  OpScratch &w = g_ops;
  PL_TRY(w.bind());
  PL_TRY(w.flag.ensure(sizeof(uint32_t)));
  PL_TRY(w.cum.ensure(sizeof(int64_t) * ((size_t)npids + 1)));
  PL_TRY(launch_fill_u32(w.flag.as<uint32_t>(), 0, 1, s));
  k_pl_check_ids<<<(npids + 255) / 256, 256, 0, s>>>(pids, npids, ix->n_doc, w.flag.as<uint32_t>());
  PL_TRY(hipGetLastError());
  uint32_t bad = 0;
  PL_TRY(hipMemcpyAsync(&bad, w.flag.p, sizeof(bad), hipMemcpyDeviceToHost, s));
  PL_TRY(hipStreamSynchronize(s));
  if (bad) return pl_fail(MVR_ERR_INVALID, "a pid lies outside [0, %u)", ix->n_doc);
  k_pl_cumsum<<<1, 1024, 0, s>>>(pids, npids, ix->doclens, w.cum.as<int64_t>());
  PL_TRY(hipGetLastError());
  int64_t total = 0;
  PL_TRY(hipMemcpyAsync(&total, w.cum.as<int64_t>() + npids, sizeof(total), hipMemcpyDeviceToHost, s));
  PL_TRY(hipStreamSynchronize(s));
  *n_rows = (uint64_t)total;
  if ((uint64_t)total > out_rows)
    return pl_fail(MVR_ERR_INVALID, "out holds %llu rows, the pids have %lld tokens", (unsigned long long)out_rows,
                   (long long)total);
  if (total == 0) return MVR_OK;
  if (!out) return pl_fail(MVR_ERR_INVALID, "out is NULL");
  k_pl_lookup16<<<std::min<uint32_t>(npids, 65536), 256, 0, s>>>(pids, npids, w.cum.as<int64_t>(), ix->doclens,
                                                                  ix->offsets, ix->codes, ix->res, ix->C16, ix->tab4h,
                                                                  ix->den16, out);
  PL_TRY(hipGetLastError());
  return MVR_OK;
}

int mvr_plaid_colbert_score_packed_f16(const uint16_t *Q16, uint32_t nq, const uint16_t *D_packed, uint64_t D_rows,
                                       const int64_t *D_lengths, uint32_t ndocs, uint16_t *out, void *stream) {
  if (ndocs == 0) return MVR_OK;
  if (nq == 0 || nq > kMaxQTok) return pl_fail(MVR_ERR_UNSUPPORTED, "nq %u must be 1..32", nq);
  if (!Q16 || !D_lengths || !out) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  OpScratch &w = g_ops;
  PL_TRY(w.bind());
  PL_TRY(w.cum.ensure(sizeof(int64_t) * ((size_t)ndocs + 1)));
  PL_TRY(w.flag.ensure(sizeof(uint32_t)));
  PL_TRY(launch_fill_u32(w.flag.as<uint32_t>(), 0, 1, s));
  k_pl_check_nonneg<<<(ndocs + 255) / 256, 256, 0, s>>>(D_lengths, ndocs, w.flag.as<uint32_t>());
  k_pl_cumsum<<<1, 1024, 0, s>>>(nullptr, ndocs, D_lengths, w.cum.as<int64_t>());
  PL_TRY(hipGetLastError());
  int64_t total = 0;
  uint32_t bad = 0;
  PL_TRY(hipMemcpyAsync(&total, w.cum.as<int64_t>() + ndocs, sizeof(total), hipMemcpyDeviceToHost, s));
  PL_TRY(hipMemcpyAsync(&bad, w.flag.p, sizeof(bad), hipMemcpyDeviceToHost, s));
  PL_TRY(hipStreamSynchronize(s));
  if (bad) return pl_fail(MVR_ERR_INVALID, "negative doc length");
  if ((uint64_t)total > D_rows)
    return pl_fail(MVR_ERR_INVALID, "D_lengths sum to %lld rows, D_packed holds %llu", (long long)total,
                   (unsigned long long)D_rows);
  if (total > 0 && !D_packed) return pl_fail(MVR_ERR_INVALID, "D_packed is NULL");
  k_pl_score_packed16<<<std::min<uint32_t>((ndocs + 3) / 4, 16384), 256, 0, s>>>(Q16, nq, D_packed,
                                                                                  w.cum.as<int64_t>(), ndocs, out);
  PL_TRY(hipGetLastError());
  return MVR_OK;
}

int mvr_plaid_colbert_score_reduce_f16(const uint16_t *scores_padded, const uint8_t *D_mask, uint32_t n, uint32_t L,
                                       uint32_t nq, uint16_t *out, void *stream) {
  if (n == 0) return MVR_OK;
  if (L == 0) return pl_fail(MVR_ERR_INVALID, "no positions to reduce (L = 0)");
  if (nq == 0 || nq > 64) return pl_fail(MVR_ERR_UNSUPPORTED, "nq %u must be 1..64", nq);
  if (!scores_padded || !D_mask || !out) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  k_pl_score_reduce16<<<std::min<uint32_t>((n + 3) / 4, 16384), 256, 0, (hipStream_t)stream>>>(scores_padded, D_mask,
                                                                                              n, L, nq, out);
  PL_TRY(hipGetLastError());
  return MVR_OK;
}

int mvr_plaid_decompress_residuals(const uint8_t *binary_residuals, uint64_t n, const uint16_t *bucket_weights,
                                   const uint8_t *reversed_bit_map, const uint8_t *bucket_weight_combinations,
                                   const int32_t *codes, const uint16_t *centroids, uint32_t dim,
                                   uint32_t nbits, uint16_t *out, void *stream) {
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8))
    return pl_fail(MVR_ERR_INVALID, "nbits %u must be 1, 2, 4 or 8", nbits);
  if (dim == 0 || dim % 8) return pl_fail(MVR_ERR_INVALID, "dim %u must be a positive multiple of 8", dim);
  if (n == 0) return MVR_OK;
  if (!binary_residuals || !bucket_weights || !reversed_bit_map || !bucket_weight_combinations || !codes ||
      !centroids || !out)
    return pl_fail(MVR_ERR_INVALID, "NULL argument");
  const uint32_t vpb = 8 / nbits, pk = dim / vpb;
  const uint64_t n_bytes = n * pk;
  k_pl_decompress_f16<<<grid_for(n_bytes, 256), 256, 0, (hipStream_t)stream>>>(
      binary_residuals, n_bytes, pk, bucket_weights, reversed_bit_map, bucket_weight_combinations, vpb, codes,
      centroids, 0xFFFFFFFFu, dim, out);
  PL_TRY(hipGetLastError());
  return MVR_OK;
}

int mvr_plaid_packbits(const uint8_t *residuals, uint64_t m, uint8_t *packed, void *stream) {
  if (m % 32) return pl_fail(MVR_ERR_INVALID, "length %llu is not a multiple of 32", (unsigned long long)m);
  if (m == 0) return MVR_OK;
  if (!residuals || !packed) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  k_pl_packbits<<<grid_for(m / 8, 256), 256, 0, (hipStream_t)stream>>>(residuals, m / 8, packed);
  PL_TRY(hipGetLastError());
  return MVR_OK;
}

int mvr_plaid_binarize(const float *residuals, uint64_t n, uint32_t dim, const float *bucket_cutoffs,
                       uint32_t n_cutoffs, uint32_t nbits, uint8_t *out, void *stream) {
  if (nbits == 0 || nbits > 8) return pl_fail(MVR_ERR_INVALID, "nbits %u out of range", nbits);
  if (dim == 0 || (dim * nbits) % 8) return pl_fail(MVR_ERR_INVALID, "dim * nbits must be a multiple of 8");
  if (n_cutoffs >= (1u << nbits)) return pl_fail(MVR_ERR_INVALID, "%u cutoffs for %u bits", n_cutoffs, nbits);
  if (n == 0) return MVR_OK;
  if (!residuals || !out || (n_cutoffs && !bucket_cutoffs)) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  const uint32_t rb = dim * nbits / 8;
  k_pl_binarize<<<grid_for(n * rb, 256), 256, 0, (hipStream_t)stream>>>(residuals, n * rb, rb, dim,
                                                                         bucket_cutoffs, n_cutoffs, nbits, out);
  PL_TRY(hipGetLastError());
  return MVR_OK;
}

int mvr_plaid_lengths_cumsum(const int32_t *pids, uint32_t n, const int64_t *lengths, int64_t *cum,
                             int64_t *total_host, void *stream) {
  if (!lengths || !cum) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  k_pl_cumsum<<<1, 1024, 0, s>>>(pids, n, lengths, cum);
  PL_TRY(hipGetLastError());
  if (total_host) {
    PL_TRY(hipMemcpyAsync(total_host, cum + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PL_TRY(hipStreamSynchronize(s));
  }
  return MVR_OK;
}

int mvr_plaid_decompress_pids(const int32_t *pids, uint32_t npids, const int64_t *cum, const int64_t *lengths,
                              const int64_t *offsets, const float *bucket_weights,
                              const uint8_t *reversed_bit_map, const uint8_t *bucket_weight_combinations,
                              const uint8_t *binary_residuals, const int32_t *codes, const float *centroids,
                              uint32_t dim, uint32_t nbits, float *out, void *stream) {
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8))
    return pl_fail(MVR_ERR_INVALID, "nbits %u must be 1, 2, 4 or 8", nbits);
  if (dim == 0 || dim % 8) return pl_fail(MVR_ERR_INVALID, "dim %u must be a positive multiple of 8", dim);
  if (npids == 0) return MVR_OK;
  if (!pids || !cum || !lengths || !offsets || !bucket_weights || !reversed_bit_map ||
      !bucket_weight_combinations || !binary_residuals || !codes || !centroids || !out)
    return pl_fail(MVR_ERR_INVALID, "NULL argument");
  const uint32_t vpb = 8 / nbits, pk = dim / vpb;
  k_pl_decompress_pids<<<npids, 256, 0, (hipStream_t)stream>>>(
      pids, cum, lengths, offsets, bucket_weights, reversed_bit_map, bucket_weight_combinations, vpb,
      binary_residuals, pk, codes, centroids, 0xFFFFFFFFu, dim, out);
  PL_TRY(hipGetLastError());
  return MVR_OK;
}

int mvr_plaid_segmented_lookup(const void *input, uint64_t n_rows, uint32_t row_bytes, const int64_t *lengths,
                               const int64_t *offsets, uint32_t n, void *out, uint64_t out_rows, void *stream) {
  if (n == 0) {
    if (out_rows != 0) return pl_fail(MVR_ERR_INVALID, "out_rows %llu but no segments", (unsigned long long)out_rows);
    return MVR_OK;
  }
  if (row_bytes == 0) return pl_fail(MVR_ERR_INVALID, "row_bytes is 0");
  if ((!input && n_rows) || !lengths || !offsets || (!out && out_rows)) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  OpScratch &w = g_ops;
  PL_TRY(w.bind());
  PL_TRY(w.cum.ensure(sizeof(int64_t) * ((size_t)n + 1)));
  PL_TRY(w.flag.ensure(sizeof(uint32_t)));
  PL_TRY(launch_fill_u32(w.flag.as<uint32_t>(), 0, 1, s));
  k_pl_cumsum<<<1, 1024, 0, s>>>(nullptr, n, lengths, w.cum.as<int64_t>());
  PL_TRY(hipGetLastError());
  int64_t total = 0;
  PL_TRY(hipMemcpyAsync(&total, w.cum.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  PL_TRY(hipStreamSynchronize(s));
  if (total < 0 || (uint64_t)total != out_rows)
    return pl_fail(MVR_ERR_INVALID, "sum(lengths) = %lld rows, out holds %llu", (long long)total,
                   (unsigned long long)out_rows);
  // widest word that divides the row and both base addresses
  const uintptr_t al = reinterpret_cast<uintptr_t>(input) | reinterpret_cast<uintptr_t>(out);
  uint32_t wb = 16;
  while (wb > 1 && ((row_bytes % wb) != 0 || (al % wb) != 0)) wb >>= 1;
  const uint32_t rw = row_bytes / wb, grid = std::min(n, 16384u);
  uint32_t *bad = w.flag.as<uint32_t>();
  const int64_t *cum = w.cum.as<int64_t>();
#define MVR_SEGL(T)                                                                                              \
  k_pl_seglookup<T><<<grid, 256, 0, s>>>(static_cast<const T *>(input), n_rows, rw, lengths, offsets, cum, n, \
                                         static_cast<T *>(out), bad)
  if (wb == 16) MVR_SEGL(uint4);
  else if (wb == 8) MVR_SEGL(uint2);
  else if (wb == 4) MVR_SEGL(uint32_t);
  else if (wb == 2) MVR_SEGL(uint16_t);
  else MVR_SEGL(uint8_t);
#undef MVR_SEGL
  PL_TRY(hipGetLastError());
  uint32_t h_bad = 0;
  PL_TRY(hipMemcpyAsync(&h_bad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PL_TRY(hipStreamSynchronize(s));
  if (h_bad) return pl_fail(MVR_ERR_INVALID, "a segment lies outside the input's %llu rows", (unsigned long long)n_rows);
  return MVR_OK;
}

int mvr_plaid_segmented_maxsim(const float *scores, uint32_t nq, const int64_t *lengths, uint32_t ndocs,
                               float *out, void *stream) {
  if (nq > 64) return pl_fail(MVR_ERR_UNSUPPORTED, "nq %u > 64", nq);
  if (ndocs == 0) return MVR_OK;
  if (!scores || !lengths || !out) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  PL_TRY(g_ops.bind());
  PL_TRY(g_ops.cum.ensure(sizeof(int64_t) * ((size_t)ndocs + 1)));
  k_pl_cumsum<<<1, 1024, 0, s>>>(nullptr, ndocs, lengths, g_ops.cum.as<int64_t>());
  PL_TRY(hipGetLastError());
  k_pl_segmented_maxsim<<<(ndocs + 3) / 4, 256, 0, s>>>(scores, nq, g_ops.cum.as<int64_t>(), ndocs, out);
  PL_TRY(hipGetLastError());
  PL_TRY(hipStreamSynchronize(s));  // the scratch cum is reused by the next call
  return MVR_OK;
}

int mvr_plaid_filter_pids(const int32_t *pids, uint32_t npids, const float *centroid_scores, uint32_t n_centroid,
                          uint32_t nq, const int32_t *codes, const int64_t *doclens, const int64_t *offsets,
                          const uint8_t *idx, uint32_t ndocs, int32_t *out_pids, uint32_t *out_n,
                          uint64_t *n_vec_score_refine, void *stream) {
  // n_doc bound: pids are trusted to index doclens/offsets as in the reference
  if (!out_n) return pl_fail(MVR_ERR_INVALID, "out_n is NULL");
  *out_n = 0;
  if (n_vec_score_refine) *n_vec_score_refine = 0;
  if (nq == 0 || nq > kMaxQTok) return pl_fail(MVR_ERR_UNSUPPORTED, "nq %u must be 1..32", nq);
  if (ndocs > kMaxSelect) return pl_fail(MVR_ERR_UNSUPPORTED, "ndocs %u > %u", ndocs, kMaxSelect);
  if (n_centroid == 0 || (size_t)((n_centroid + 31) / 32) * 4 > kPlMaxBmBytes)
    return pl_fail(MVR_ERR_UNSUPPORTED, "n_centroid %u out of range (1..%u)", n_centroid, kPlMaxBmBytes * 8);
  if (npids == 0 || ndocs < 4) return MVR_OK;
  if (!pids || !centroid_scores || !codes || !doclens || !offsets || !idx || !out_pids)
    return pl_fail(MVR_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  OpScratch &w = g_ops;
  PL_TRY(w.bind());
  const uint32_t n_seg1 = (npids + kSeg - 1) / kSeg;
  PL_TRY(w.keys.ensure(sizeof(uint64_t) * (size_t)n_seg1 * kSeg));
  PL_TRY(w.seg.ensure(sizeof(uint32_t) * n_seg1));
  PL_TRY(w.sel1_ids.ensure(sizeof(uint32_t) * ndocs));
  PL_TRY(w.sel1_F.ensure(sizeof(float) * ndocs));
  PL_TRY(w.sel1_n.ensure(sizeof(uint32_t)));
  PL_TRY(w.keys2.ensure(sizeof(uint64_t) * (size_t)((ndocs + kSeg - 1) / kSeg) * kSeg));
  PL_TRY(w.sel2_ids.ensure(sizeof(uint32_t) * (ndocs / 4)));
  PL_TRY(w.sel2_F.ensure(sizeof(float) * (ndocs / 4)));
  PL_TRY(w.sel2_n.ensure(sizeof(uint32_t)));
  PL_TRY(w.nvec.ensure(sizeof(unsigned long long)));
  PL_TRY(launch_fill_u32(w.nvec.as<uint32_t>(), 0, 2, s));
  PL_TRY(launch_filter_stages(reinterpret_cast<const uint32_t *>(pids), nullptr, npids, 0, npids, nullptr,
                              nullptr, nullptr, centroid_scores, idx, n_centroid, nq, 0x7FFFFFFFu, codes, doclens, offsets,
                              nullptr, 0, 1, ndocs, w.keys.as<uint64_t>(), w.seg.as<uint32_t>(),
                              w.sel1_ids.as<uint32_t>(), w.sel1_F.as<float>(), w.sel1_n.as<uint32_t>(),
                              w.keys2.as<uint64_t>(), w.sel2_ids.as<uint32_t>(), w.sel2_F.as<float>(),
                              w.sel2_n.as<uint32_t>(), w.nvec.as<unsigned long long>(), 0,
                              -9999.0f * (float)nq, s));
  k_pl_unkey<<<(ndocs / 4 + 255) / 256, 256, 0, s>>>(w.sel2_ids.as<uint32_t>(), w.sel2_n.as<uint32_t>(),
                                                      out_pids);
  PL_TRY(hipGetLastError());
  uint32_t n_out = 0;
  unsigned long long nv = 0;
  PL_TRY(hipMemcpyAsync(&n_out, w.sel2_n.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PL_TRY(hipMemcpyAsync(&nv, w.nvec.p, sizeof(nv), hipMemcpyDeviceToHost, s));
  PL_TRY(hipStreamSynchronize(s));
  *out_n = n_out;
  if (n_vec_score_refine) *n_vec_score_refine = nv;
  return MVR_OK;
}

int mvr_plaid_index_create(const mvr_plaid_index_desc *d, int device, mvr_plaid_index **out) {
  if (!d || !out) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (d->dim != kDim || d->nbits != 2)
    return pl_fail(MVR_ERR_UNSUPPORTED, "search kernels need dim 128 and nbits 2 (got %u, %u)", d->dim, d->nbits);
  if (d->n_doc == 0 || d->n_centroid == 0) return pl_fail(MVR_ERR_INVALID, "empty index");
  if ((size_t)((d->n_centroid + 31) / 32) * 4 > kPlMaxBmBytes)
    return pl_fail(MVR_ERR_UNSUPPORTED, "n_centroid %u > %u", d->n_centroid, kPlMaxBmBytes * 8);
  if (d->n_doc > 0x7FFFFFFFu) return pl_fail(MVR_ERR_UNSUPPORTED, "n_doc too large");
  if (!d->doclens || !d->codes || !d->residuals || !d->centroids || !d->bucket_weights || !d->ivf ||
      !d->ivf_lengths)
    return pl_fail(MVR_ERR_INVALID, "NULL index array");
  std::vector<int64_t> offsets(d->n_doc), ivf_off(d->n_centroid);
  uint64_t tot = 0;
  for (uint32_t i = 0; i < d->n_doc; ++i) {
    if (d->doclens[i] < 0) return pl_fail(MVR_ERR_INVALID, "negative doclen at %u", i);
    offsets[i] = (int64_t)tot;
    tot += (uint64_t)d->doclens[i];
  }
  if (tot != d->n_tok)
    return pl_fail(MVR_ERR_INVALID, "sum of doclens %llu != n_tok %llu", (unsigned long long)tot,
                   (unsigned long long)d->n_tok);
  uint64_t itot = 0;
  for (uint32_t c = 0; c < d->n_centroid; ++c) {
    if (d->ivf_lengths[c] < 0) return pl_fail(MVR_ERR_INVALID, "negative ivf length at %u", c);
    ivf_off[c] = (int64_t)itot;
    itot += (uint64_t)d->ivf_lengths[c];
  }
  if (itot != d->n_ivf) return pl_fail(MVR_ERR_INVALID, "sum of ivf_lengths != n_ivf");
  for (uint64_t i = 0; i < d->n_tok; ++i)
    if (d->codes[i] < 0 || (uint32_t)d->codes[i] >= d->n_centroid)
      return pl_fail(MVR_ERR_INVALID, "code %d at %llu out of range", d->codes[i], (unsigned long long)i);
  // The search relies on IVF[c] being exactly the docs that hold code c, in
  // ascending order (ivf.pid.pt, indexing/utils.py:8-53): check it.
  {
    std::vector<int64_t> last(d->n_centroid, -1), fill(d->n_centroid, 0);
    uint64_t tok = 0;
    for (uint32_t doc = 0; doc < d->n_doc; ++doc) {
      for (int64_t j = 0; j < d->doclens[doc]; ++j, ++tok) {
        const uint32_t c = (uint32_t)d->codes[tok];
        if (last[c] == (int64_t)doc) continue;
        last[c] = doc;
        if (fill[c] >= d->ivf_lengths[c] || d->ivf[ivf_off[c] + fill[c]] != (int32_t)doc)
          return pl_fail(MVR_ERR_INVALID, "ivf list of centroid %u does not hold exactly its docs (doc %u)", c, doc);
        ++fill[c];
      }
    }
    for (uint32_t c = 0; c < d->n_centroid; ++c)
      if (fill[c] != d->ivf_lengths[c])
        return pl_fail(MVR_ERR_INVALID, "ivf list of centroid %u has %lld entries, its docs %lld", c,
                       (long long)d->ivf_lengths[c], (long long)fill[c]);
  }
  // raw PLAID byte v -> weights of its four dims: reversed_bit_map swaps the
  // two bits of each 2-bit group (residual.py:51-72) and the product table
  // reads the groups MSB first (residual.py:76-91)
  std::vector<float> tab(256 * 4);
  for (uint32_t v = 0; v < 256; ++v)
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t g = (v >> (6 - 2 * k)) & 3u;
      const uint32_t b = ((g & 1u) << 1) | (g >> 1);
      tab[v * 4 + k] = d->bucket_weights[b];
    }

  int n_dev = 0;
  PL_TRY(hipGetDeviceCount(&n_dev));
  if (device < 0 || device >= n_dev) return pl_fail(MVR_ERR_INVALID, "device %d out of range (%d)", device, n_dev);
  PL_TRY(hipSetDevice(device));
  auto *ix = new mvr_plaid_index();
  ix->device = device;
  ix->n_doc = d->n_doc;
  ix->n_c = d->n_centroid;
  ix->n_tok = d->n_tok;
  ix->n_ivf = d->n_ivf;
  ix->bm_words = (d->n_doc + 31) / 32;
  ix->avg_len = (float)((double)d->n_tok / (double)d->n_doc);
  auto bail = [&](int code) {
    mvr_plaid_index_free(ix);
    return code;
  };
  auto up = [&](void **dst, const void *src, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) return e;
    return bytes ? hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  hipError_t e = hipSuccess;
  if ((e = up((void **)&ix->C, d->centroids, sizeof(float) * (size_t)d->n_centroid * kDim)) != hipSuccess ||
      (e = up((void **)&ix->codes, d->codes, sizeof(int32_t) * d->n_tok)) != hipSuccess ||
      (e = up((void **)&ix->res, d->residuals, (size_t)kResBytes * d->n_tok)) != hipSuccess ||
      (e = up((void **)&ix->doclens, d->doclens, sizeof(int64_t) * d->n_doc)) != hipSuccess ||
      (e = up((void **)&ix->offsets, offsets.data(), sizeof(int64_t) * d->n_doc)) != hipSuccess ||
      (e = up((void **)&ix->ivf, d->ivf, sizeof(int32_t) * d->n_ivf)) != hipSuccess ||
      (e = up((void **)&ix->ivf_off, ivf_off.data(), sizeof(int64_t) * d->n_centroid)) != hipSuccess ||
      (e = up((void **)&ix->ivf_len, d->ivf_lengths, sizeof(int64_t) * d->n_centroid)) != hipSuccess ||
      (e = up((void **)&ix->tab4, tab.data(), sizeof(float) * tab.size())) != hipSuccess ||
      (e = hipMalloc((void **)&ix->den, sizeof(float) * std::max<uint64_t>(d->n_tok, 1))) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(pl_fail(MVR_ERR_DEVICE, "index upload failed: %s", hipGetErrorString(e)));
  if (d->n_tok) {
    k_pl_token_den<false><<<grid_for(d->n_tok, 256), 256, 0, ix->stream>>>(ix->codes, ix->res, d->n_tok, ix->C,
                                                                          ix->tab4, ix->den);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(ix->stream)) != hipSuccess)
      return bail(pl_fail(MVR_ERR_DEVICE, "token norms failed: %s", hipGetErrorString(e)));
  }
  for (auto &ev : ix->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess)
      return bail(pl_fail(MVR_ERR_DEVICE, "event create failed: %s", hipGetErrorString(e)));
  {
    // fused candidate marking when the bitmap (plus the kernel's static
    // LDS) fits the device's per-workgroup LDS
    // (MVR_PLAID_FUSED_MARK=0 keeps the global-bitmap kernels)
    const char *fm = std::getenv("MVR_PLAID_FUSED_MARK");
    int max_lds = 0;
    const size_t dyn = sizeof(uint32_t) * (size_t)ix->bm_words;
    const size_t stat = sizeof(uint32_t) * (kMcCells + 1) + sizeof(int64_t) * kMcCells + 64 + 256;
    if (!(fm && fm[0] == '0') &&
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
        dyn + stat <= (size_t)max_lds &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(k_pl_markc), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)dyn) == hipSuccess)
      ix->markc = true;
    (void)hipGetLastError();
  }
  {
    // pass workspace: a quarter of the HBM left free by the index, between 2 and 64 GiB
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      ix->pool_budget = std::min<uint64_t>(kPoolBudgetMax, std::max<uint64_t>(kPoolBudgetMin, free_b / 4));
  }
  *out = ix;
  return MVR_OK;
}

void mvr_plaid_index_free(mvr_plaid_index *ix) {
  if (!ix) return;
  (void)hipSetDevice(ix->device);
  if (ix->stream) (void)hipStreamSynchronize(ix->stream);
  for (void *p : {(void *)ix->C, (void *)ix->codes, (void *)ix->res, (void *)ix->doclens, (void *)ix->offsets,
                  (void *)ix->ivf, (void *)ix->ivf_off, (void *)ix->ivf_len, (void *)ix->tab4, (void *)ix->den,
                  (void *)ix->C16, (void *)ix->tab4h, (void *)ix->den16})
    if (p) (void)hipFree(p);
  for (PBuf *b : ix->bufs()) b->release();
  for (auto &ev : ix->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ix->stream) (void)hipStreamDestroy(ix->stream);
  delete ix;
}

int mvr_plaid_search(mvr_plaid_index *ix, const float *Q, uint32_t n_q, uint32_t nq, const mvr_plaid_params *p,
                     int32_t *out_ids, float *out_scores, int q_on_device, uint64_t *stats,
                     mvr_plaid_timings *timings) {
  if (!ix || !p || !out_ids || !out_scores) return pl_fail(MVR_ERR_INVALID, "NULL argument");
  if (timings) *timings = mvr_plaid_timings{};
  if (n_q == 0) return MVR_OK;
  if (!Q) return pl_fail(MVR_ERR_INVALID, "Q is NULL");
  if (nq == 0 || nq > kMaxQTok) return pl_fail(MVR_ERR_UNSUPPORTED, "nq %u must be 1..32", nq);
  if (p->k == 0) return pl_fail(MVR_ERR_INVALID, "k must be positive");
  if (p->ncells == 0 || p->ncells > kPlMaxCells)
    return pl_fail(MVR_ERR_UNSUPPORTED, "ncells %u must be 1..%u", p->ncells, kPlMaxCells);
  if (p->ndocs < 4 || p->ndocs > kMaxSelect)
    return pl_fail(MVR_ERR_UNSUPPORTED, "ndocs %u must be 4..%u", p->ndocs, kMaxSelect);
  if (p->k > kMaxSelect) return pl_fail(MVR_ERR_UNSUPPORTED, "k %u > %u", p->k, kMaxSelect);
  PL_TRY(hipSetDevice(ix->device));
  const bool gb = p->gpu_branch != 0;
  if (gb) {
    const int rc = pl_prepare_f16(ix);
    if (rc != MVR_OK) return rc;
  }
  // GPU branch: idx compares the half scores with the threshold rounded to
  // half (a half tensor against a Python float); a doc whose maxima all keep
  // the identity scores the half of their FP32 sum
  const float thr = gb ? host_r16(p->centroid_score_threshold) : p->centroid_score_threshold;
  float empty_score = 0.0f;
  for (uint32_t t = 0; t < nq; ++t) empty_score += gb ? -10000.0f : -9999.0f;
  if (gb) empty_score = host_r16(empty_score);
  hipStream_t s = ix->stream;
  const uint32_t n_c = ix->n_c, n_doc = ix->n_doc, ndocs = p->ndocs, R2 = ndocs / 4, k = p->k;
  const uint32_t cell_cap = nq * p->ncells;
  const uint32_t n_seg1 = (n_doc + kSeg - 1) / kSeg, n_light = pl_light_segs(ndocs, n_doc);
  // device bytes per query in a pass
  const size_t per_q = sizeof(float) * kDim * kMaxQTok + sizeof(float) * (size_t)n_c * nq + n_c +
                       (((size_t)n_c + kCellChunk - 1) / kCellChunk + 1) * 8 * (size_t)kMaxQTok * p->ncells +
                       8 * (((size_t)n_c + kGt - 1) / kGt) * kMaxQTok +
                       sizeof(uint32_t) * ((size_t)n_doc + ix->bm_words + kPassCap + 3) +
                       sizeof(uint32_t) * (cell_cap + 1) + sizeof(uint32_t) * ix->bm_words +
                       sizeof(uint32_t) * (n_doc + 1) + sizeof(uint64_t) * (size_t)(n_seg1 + n_light) * kSeg +
                       sizeof(uint32_t) * (n_seg1 + n_light + 1) + 12 * (size_t)ndocs +
                       8 * (size_t)((ndocs + kSeg - 1) / kSeg) * kSeg + 16 * (size_t)R2 +
                       8 * (size_t)k + 16;
  uint32_t B = p->batch ? p->batch : 1024;
  B = std::min<uint32_t>(B, n_q);
  B = (uint32_t)std::max<size_t>(1, std::min<size_t>(B, ix->pool_budget / per_q));
  if (ix->Q.cap < sizeof(float) * (size_t)B * kMaxQTok * kDim) {
    PL_TRY(ix->Q.ensure(sizeof(float) * (size_t)B * kMaxQTok * kDim));
    PL_TRY(launch_fill_u32(ix->Q.as<uint32_t>(), 0, ix->Q.cap / 4, s));
    ix->q_nq = kMaxQTok;
  }
  if (nq != ix->q_nq) {  // padding rows of an earlier, wider batch
    PL_TRY(launch_fill_u32(ix->Q.as<uint32_t>(), 0, ix->Q.cap / 4, s));
    ix->q_nq = nq;
  }
  const uint32_t n_chunk = (n_c + kCellChunk - 1) / kCellChunk;
  PL_TRY(ix->part_keys.ensure(sizeof(uint64_t) * (size_t)n_chunk * B * kMaxQTok * p->ncells));
  PL_TRY(ix->top_ids.ensure(sizeof(uint32_t) * (size_t)B * kMaxQTok * p->ncells));
  PL_TRY(ix->top1.ensure(sizeof(unsigned long long) * (size_t)B * ((n_c + kGt - 1) / kGt) * kMaxQTok));
  PL_TRY(ix->S.ensure(sizeof(float) * (size_t)B * n_c * nq));
  PL_TRY(ix->idx.ensure((size_t)B * n_c));
  PL_TRY(ix->cells.ensure(sizeof(uint32_t) * (size_t)B * cell_cap));
  PL_TRY(ix->cell_n.ensure(sizeof(uint32_t) * B));
  if (ix->pid_bm.cap < sizeof(uint32_t) * (size_t)B * ix->bm_words ||
      ix->pass_bm.cap < sizeof(uint32_t) * (size_t)B * ix->bm_words)
    ix->bm_dirty = true;
  PL_TRY(ix->pid_bm.ensure(sizeof(uint32_t) * (size_t)B * ix->bm_words));
  PL_TRY(ix->pass_bm.ensure(sizeof(uint32_t) * (size_t)B * ix->bm_words));
  PL_TRY(ix->pass_list.ensure(sizeof(uint32_t) * (size_t)B * kPassCap));
  PL_TRY(ix->pass_cnt.ensure(sizeof(uint32_t) * B));
  PL_TRY(ix->pmode.ensure(sizeof(uint32_t) * B));
  PL_TRY(ix->heavy.ensure(sizeof(uint32_t) * (size_t)B * n_doc));
  PL_TRY(ix->heavy_n.ensure(sizeof(uint32_t) * B));
  PL_TRY(ix->cand.ensure(sizeof(uint32_t) * (size_t)B * n_doc));
  PL_TRY(ix->cand_n.ensure(sizeof(uint32_t) * B));
  PL_TRY(ix->keys1.ensure(sizeof(uint64_t) * (size_t)B * (n_seg1 + n_light) * kSeg));
  PL_TRY(ix->seg1.ensure(sizeof(uint32_t) * (size_t)B * (n_seg1 + n_light)));
  PL_TRY(ix->light_n.ensure(sizeof(uint32_t) * B));
  PL_TRY(ix->sel1_ids.ensure(sizeof(uint32_t) * (size_t)B * ndocs));
  PL_TRY(ix->sel1_F.ensure(sizeof(float) * (size_t)B * ndocs));
  PL_TRY(ix->sel1_n.ensure(sizeof(uint32_t) * B));
  PL_TRY(ix->keys2.ensure(sizeof(uint64_t) * (size_t)B * ((ndocs + kSeg - 1) / kSeg) * kSeg));
  PL_TRY(ix->sel2_ids.ensure(sizeof(uint32_t) * (size_t)B * R2));
  PL_TRY(ix->sel2_F.ensure(sizeof(float) * (size_t)B * R2));
  PL_TRY(ix->sel2_n.ensure(sizeof(uint32_t) * B));
  PL_TRY(ix->ids2.ensure(sizeof(uint32_t) * (size_t)B * R2));
  PL_TRY(ix->score2.ensure(sizeof(float) * (size_t)B * R2));
  PL_TRY(ix->out_ids.ensure(sizeof(uint32_t) * (size_t)B * k));
  PL_TRY(ix->out_s.ensure(sizeof(float) * (size_t)B * k));
  PL_TRY(ix->nvec.ensure(sizeof(unsigned long long) * B));
  if (ix->bm_dirty) {
    PL_TRY(launch_fill_u32(ix->pid_bm.as<uint32_t>(), 0, ix->pid_bm.cap / 4, s));
    PL_TRY(launch_fill_u32(ix->pass_bm.as<uint32_t>(), 0, ix->pass_bm.cap / 4, s));
    ix->bm_dirty = false;
  }
  std::vector<uint32_t> h_cand(B), h_fin(B);
  std::vector<unsigned long long> h_nv(B);
  double t_cand = 0, t_filt = 0, t_ref = 0, t_sc = 0;
  uint32_t n_pass = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t q0 = 0; q0 < n_q; q0 += B) {
    const uint32_t b = std::min(B, n_q - q0);
    const float *Qb = Q + (size_t)q0 * nq * kDim;
    // queries padded to 32 rows (rows t >= nq stay zero)
    PL_TRY(hipMemcpy2DAsync(ix->Q.p, sizeof(float) * kMaxQTok * kDim, Qb, sizeof(float) * nq * kDim,
                            sizeof(float) * nq * kDim, b,
                            q_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    // a failed pass leaves pid bits set; the next call clears the bitmap
    ix->bm_dirty = true;
    PL_TRY(hipEventRecord(ix->ev[0], s));
    if (gb) {  // queries as halves (candidate_generation.py:56, colbert.py:198)
      k_pl_round16<<<grid_for((uint64_t)b * kMaxQTok * kDim, 256), 256, 0, s>>>(
          ix->Q.as<float>(), (uint64_t)b * kMaxQTok * kDim, ix->Q.as<float>());
      PL_TRY(hipGetLastError());
    }
    const uint32_t n_tiles = (n_c + kGt - 1) / kGt;
    k_pl_sgemm<<<dim3(n_tiles, (b + 3) / 4), 256, 0, s>>>(
        ix->Q.as<float>(), b, nq, gb ? ix->C16 : ix->C, n_c, thr, ix->S.as<float>(), ix->idx.as<uint8_t>(),
        p->ncells == 1 ? ix->top1.as<unsigned long long>() : nullptr, gb ? 1 : 0);
    PL_TRY(hipGetLastError());
    PL_TRY(hipEventRecord(ix->ev[4], s));
    if (p->ncells == 1) {
      k_pl_top1_merge<<<(b * kMaxQTok + 255) / 256, 256, 0, s>>>(ix->top1.as<unsigned long long>(), b, nq, n_tiles,
                                                                 ix->top_ids.as<uint32_t>());
      PL_TRY(hipGetLastError());
    } else {
      const uint32_t n_chunk = (n_c + kCellChunk - 1) / kCellChunk;
      const dim3 gp((n_chunk + 3) / 4, b);
      const uint32_t gm = (b * kMaxQTok + 255) / 256;
      switch (p->ncells) {
#define PL_CELLS(NC)                                                                                          \
  case NC:                                                                                                    \
    k_pl_cell_part<NC><<<gp, 256, 0, s>>>(ix->S.as<float>(), n_c, nq, n_chunk,                                \
                                          ix->part_keys.as<unsigned long long>());                            \
    k_pl_cell_merge<NC><<<gm, 256, 0, s>>>(ix->part_keys.as<unsigned long long>(), b, nq,                     \
                                           n_chunk, ix->top_ids.as<uint32_t>());                              \
    break;
        PL_CELLS(1) PL_CELLS(2) PL_CELLS(3) PL_CELLS(4) PL_CELLS(5) PL_CELLS(6) PL_CELLS(7) PL_CELLS(8)
#undef PL_CELLS
      }
      PL_TRY(hipGetLastError());
    }
    k_pl_cells<<<b, 64, sizeof(uint32_t) * ((n_c + 31) / 32), s>>>(ix->top_ids.as<uint32_t>(), p->ncells, nq, n_c,
                                                                   ix->cells.as<uint32_t>(), cell_cap,
                                                                   ix->cell_n.as<uint32_t>());
    PL_TRY(hipGetLastError());
    PL_TRY(launch_fill_u32(ix->nvec.as<uint32_t>(), 0, 2 * (size_t)b, s));
    if (ix->markc && cell_cap <= kMcCells) {
      k_pl_markc<<<b, 1024, sizeof(uint32_t) * (size_t)ix->bm_words, s>>>(
          ix->cells.as<uint32_t>(), cell_cap, ix->cell_n.as<uint32_t>(), ix->ivf, ix->ivf_off, ix->ivf_len,
          ix->bm_words, ix->cand.as<uint32_t>(), n_doc, ix->cand_n.as<uint32_t>());
      PL_TRY(hipGetLastError());
    } else {
      const uint32_t gx = std::max(1u, std::min(16u, 2048u / b));
      k_pl_mark<<<dim3(gx, b), 256, 0, s>>>(ix->cells.as<uint32_t>(), cell_cap, ix->cell_n.as<uint32_t>(), ix->ivf,
                                            ix->ivf_off, ix->ivf_len, ix->pid_bm.as<uint32_t>(), ix->bm_words);
      PL_TRY(hipGetLastError());
      k_pl_compact<<<b, 1024, 0, s>>>(ix->pid_bm.as<uint32_t>(), ix->bm_words, ix->cand.as<uint32_t>(), n_doc,
                                      ix->cand_n.as<uint32_t>());
      PL_TRY(hipGetLastError());
    }
    PL_TRY(hipEventRecord(ix->ev[1], s));
    k_pl_passmark<<<b, 256, 0, s>>>(ix->idx.as<uint8_t>(), n_c, ix->ivf, ix->ivf_off, ix->ivf_len,
                                    ix->cand_n.as<uint32_t>(), ix->avg_len, ix->pass_list.as<uint32_t>(),
                                    ix->pass_cnt.as<uint32_t>(), ix->pmode.as<uint32_t>(),
                                    ix->pass_bm.as<uint32_t>(), ix->bm_words);
    PL_TRY(hipGetLastError());
    {
      // key layout of launch_filter_stages: heavy segments, then light ones
      const uint32_t n_heavy = (n_doc + kSeg - 1) / kSeg, n_keyseg = n_heavy + pl_light_segs(ndocs, n_doc);
      k_pl_split<<<b, 1024, 0, s>>>(ix->cand.as<uint32_t>(), n_doc, ix->cand_n.as<uint32_t>(), ix->pmode.as<uint32_t>(),
                                    ix->pass_bm.as<uint32_t>(), ix->bm_words, nq, ndocs,
                                    ix->keys1.as<unsigned long long>(), (uint64_t)n_keyseg * kSeg,
                                    (uint64_t)n_heavy * kSeg, ix->heavy.as<uint32_t>(),
                                    ix->heavy_n.as<uint32_t>(), ix->light_n.as<uint32_t>(), empty_score);
      PL_TRY(hipGetLastError());
    }
    PL_TRY(launch_filter_stages(ix->cand.as<uint32_t>(), ix->cand_n.as<uint32_t>(), 0, n_doc, n_doc,
                                ix->heavy.as<uint32_t>(), ix->heavy_n.as<uint32_t>(), ix->light_n.as<uint32_t>(),
                                ix->S.as<float>(), ix->idx.as<uint8_t>(), n_c, nq, n_doc, ix->codes, ix->doclens,
                                ix->offsets, nullptr, ix->bm_words, b, ndocs,
                                ix->keys1.as<uint64_t>(), ix->seg1.as<uint32_t>(), ix->sel1_ids.as<uint32_t>(),
                                ix->sel1_F.as<float>(), ix->sel1_n.as<uint32_t>(), ix->keys2.as<uint64_t>(),
                                ix->sel2_ids.as<uint32_t>(), ix->sel2_F.as<float>(), ix->sel2_n.as<uint32_t>(),
                                ix->nvec.as<unsigned long long>(), gb ? 1 : 0, empty_score, s));
    k_pl_passclear<<<b, 256, 0, s>>>(ix->pass_list.as<uint32_t>(), ix->pass_cnt.as<uint32_t>(), ix->ivf,
                                     ix->ivf_off, ix->ivf_len, ix->pass_bm.as<uint32_t>(), ix->bm_words);
    PL_TRY(hipGetLastError());
    PL_TRY(hipEventRecord(ix->ev[2], s));
    {
      // about 4 docs per wave: each wave holds its query's 32 rows in registers
      const uint32_t n_slice = std::max(1u, std::min((R2 + 15) / 16, std::max(4u, 16384u / b)));
      if (gb)
        k_pl_refine<true><<<dim3(n_slice, b), 256, 0, s>>>(ix->Q.as<float>(), nq, ix->sel2_ids.as<uint32_t>(),
                                                         ix->sel2_n.as<uint32_t>(), R2, n_slice, n_doc, ix->doclens,
                                                         ix->offsets, ix->codes, ix->res, ix->C16, ix->tab4h,
                                                         ix->den16, ix->ids2.as<uint32_t>(), ix->score2.as<float>());
      else
        k_pl_refine<false><<<dim3(n_slice, b), 256, 0, s>>>(ix->Q.as<float>(), nq, ix->sel2_ids.as<uint32_t>(),
                                                          ix->sel2_n.as<uint32_t>(), R2, n_slice, n_doc, ix->doclens,
                                                          ix->offsets, ix->codes, ix->res, ix->C, ix->tab4, ix->den,
                                                          ix->ids2.as<uint32_t>(), ix->score2.as<float>());
      PL_TRY(hipGetLastError());
    }
    PL_TRY(launch_topk(ix->ids2.as<uint32_t>(), ix->score2.as<float>(), ix->sel2_n.as<uint32_t>(), R2, b, k,
                       ix->out_ids.as<uint32_t>(), ix->out_s.as<float>(), nullptr, s));
    PL_TRY(hipEventRecord(ix->ev[3], s));
    PL_TRY(hipMemcpyAsync(out_ids + (size_t)q0 * k, ix->out_ids.p, sizeof(int32_t) * (size_t)b * k,
                          hipMemcpyDeviceToHost, s));
    PL_TRY(hipMemcpyAsync(out_scores + (size_t)q0 * k, ix->out_s.p, sizeof(float) * (size_t)b * k,
                          hipMemcpyDeviceToHost, s));
    PL_TRY(hipMemcpyAsync(h_cand.data(), ix->cand_n.p, sizeof(uint32_t) * b, hipMemcpyDeviceToHost, s));
    PL_TRY(hipMemcpyAsync(h_fin.data(), ix->sel2_n.p, sizeof(uint32_t) * b, hipMemcpyDeviceToHost, s));
    PL_TRY(hipMemcpyAsync(h_nv.data(), ix->nvec.p, sizeof(unsigned long long) * b, hipMemcpyDeviceToHost, s));
    PL_TRY(hipStreamSynchronize(s));
    ix->bm_dirty = false;  // k_pl_compact and k_pl_passclear cleared every word set in the pass
    float ms = 0;
    PL_TRY(hipEventElapsedTime(&ms, ix->ev[0], ix->ev[1]));
    t_cand += ms;
    PL_TRY(hipEventElapsedTime(&ms, ix->ev[1], ix->ev[2]));
    t_filt += ms;
    PL_TRY(hipEventElapsedTime(&ms, ix->ev[2], ix->ev[3]));
    t_ref += ms;
    PL_TRY(hipEventElapsedTime(&ms, ix->ev[0], ix->ev[4]));
    t_sc += ms;
    ++n_pass;
    for (uint32_t i = 0; i < b; ++i) {
      if (stats) {
        stats[(size_t)(q0 + i) * 3 + 0] = h_cand[i];
        stats[(size_t)(q0 + i) * 3 + 1] = h_fin[i];
        stats[(size_t)(q0 + i) * 3 + 2] = h_nv[i];
      }
      if (timings) {
        timings->n_refine_ivf += h_cand[i];
        timings->n_refine_filter += h_fin[i];
        timings->n_vec_score_refine += h_nv[i];
      }
    }
  }
  if (timings) {
    timings->candidates_ms = t_cand;
    timings->filter_ms = t_filt;
    timings->refine_ms = t_ref;
    timings->scores_ms = t_sc;
    timings->n_pass = n_pass;
    timings->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return MVR_OK;
}

}  // extern "C"
