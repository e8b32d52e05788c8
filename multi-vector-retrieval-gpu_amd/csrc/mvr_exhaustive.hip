// All-pairs MaxSim of the exhaustive (groundtruth) module for gfx950, with its
// C ABI entry mvr_maxsim_scores_host (include/mvr.h).
//
// Replaces (reference paths): GPUComputeScore::computeItemScore
// (src/include/compute_score/GPUComputeScore.hpp:196-246: one cuBLAS SGEMM
// per query, then per query token a thrust for_each max over each item's
// tokens, select_max_functor :63-89, and a thrust transform-add over the
// query tokens), reached through BruteForceProgressive::computeQueryItemScore
// (src/impl/BruteForceProgressive.hpp:106-141).
//
// Design: one kernel for all queries of a call.  A workgroup takes 4 queries
// (128 query-token rows) and a column block of whole items (up to 48 items,
// about 512 tokens), walks the block in 128-token tiles with an 8 x 8
// register tile per thread (K through LDS in 32-dim chunks, each dot a
// sequential FMA chain over d), and folds every tile into per-(row, item)
// maxima held in LDS (atomicMax on order-preserving bits); at the end each
// (query, item) score is the left-to-right sum over query tokens.  No
// per-query launches, no score matrix of query tokens x item tokens in HBM.
#include "../../include/mvr.h"
#include "mvr_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

namespace mvr {
namespace {

constexpr uint32_t kXk = 32;        // dims per LDS chunk
constexpr uint32_t kXt = 128;       // rows / item tokens per tile
constexpr uint32_t kXp = kXt + 4;   // padded LDS row
constexpr uint32_t kXItems = 48;    // items per column block (LDS: 64 KB per workgroup in all)
constexpr uint32_t kXTok = 512;     // target tokens per column block

__device__ __forceinline__ uint32_t x_f2ord(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float x_ord2f(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// blocks[b] = (unused, tokens, first item, items) of column block b;
// item_off [n_item + 1] token offsets (relative to X).
__global__ __launch_bounds__(256) void k_exh_maxsim(const float *__restrict__ Q, uint32_t n_q, uint32_t T,
                                                    const float *__restrict__ X,
                                                    const uint4 *__restrict__ blocks,
                                                    const uint64_t *__restrict__ item_off, uint32_t n_item,
                                                    float *__restrict__ scores, uint64_t score_stride) {
  __shared__ __align__(16) float As[kXk][kXp];
  __shared__ __align__(16) float Bs[kXk][kXp];
  __shared__ uint32_t M[kXt][kXItems];
  __shared__ uint32_t slot[kXt];
  __shared__ uint64_t ioff[kXItems + 1];
  const uint32_t tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const uint32_t q0 = blockIdx.y * 4;
  const uint4 blk = blocks[blockIdx.x];
  const uint64_t tok0 = item_off[blk.z];  // first token of the block
  const uint32_t ntok = blk.y, item0 = blk.z, nit = blk.w;
  const uint32_t n_rows = n_q * kMaxQTok;
  for (uint32_t e = tid; e < kXt * kXItems; e += 256) (&M[0][0])[e] = 0x007FFFFFu;  // order bits of -inf
  for (uint32_t e = tid; e <= nit; e += 256) ioff[e] = item_off[item0 + e] - tok0;
  __syncthreads();
  const float4 *Q4 = reinterpret_cast<const float4 *>(Q);
  const float4 *X4 = reinterpret_cast<const float4 *>(X);
  for (uint32_t tile0 = 0; tile0 < ntok; tile0 += kXt) {
    if (tid < kXt) {
      const uint32_t tk = tile0 + tid;
      uint32_t s = 0xFFFFFFFFu;
      if (tk < ntok) {
        s = 0;
        while (s + 1 < nit && ioff[s + 1] <= tk) ++s;
      }
      slot[tid] = s;
    }
    float acc[8][8];
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i)
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) acc[i][j] = 0.0f;
#pragma nounroll
    for (uint32_t k0 = 0; k0 < kDim; k0 += kXk) {
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t e = tid + 256 * u, r = e >> 3, d4 = e & 7;
        const uint32_t row = q0 * kMaxQTok + r, tk = tile0 + r;
        const float4 a =
            row < n_rows ? Q4[(size_t)row * (kDim / 4) + k0 / 4 + d4] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 b =
            tk < ntok ? X4[(tok0 + tk) * (kDim / 4) + k0 / 4 + d4] : make_float4(0.f, 0.f, 0.f, 0.f);
        As[4 * d4 + 0][r] = a.x; As[4 * d4 + 1][r] = a.y; As[4 * d4 + 2][r] = a.z; As[4 * d4 + 3][r] = a.w;
        Bs[4 * d4 + 0][r] = b.x; Bs[4 * d4 + 1][r] = b.y; Bs[4 * d4 + 2][r] = b.z; Bs[4 * d4 + 3][r] = b.w;
      }
      __syncthreads();
#pragma unroll 4
      for (uint32_t d = 0; d < kXk; ++d) {
        const float4 a0 = *reinterpret_cast<const float4 *>(&As[d][ty * 8]);
        const float4 a1 = *reinterpret_cast<const float4 *>(&As[d][ty * 8 + 4]);
        const float4 b0 = *reinterpret_cast<const float4 *>(&Bs[d][tx * 4]);
        const float4 b1 = *reinterpret_cast<const float4 *>(&Bs[d][64 + tx * 4]);
        const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i)
#pragma unroll
          for (uint32_t j = 0; j < 8; ++j) acc[i][j] = __builtin_fmaf(av[i], bv[j], acc[i][j]);
      }
      __syncthreads();
    }
    // fold the tile into the per-(row, item) maxima: runs of 4 consecutive
    // columns usually share an item, so combine them before the atomic
    const uint32_t t0 = (ty & 3) * 8;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
      if (t0 + i >= T) continue;
      const uint32_t row = ty * 8 + i;
#pragma unroll
      for (uint32_t g = 0; g < 2; ++g) {
        uint32_t cur = 0xFFFFFFFFu;
        float m = -INFINITY;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const uint32_t col = g * 64 + tx * 4 + u;
          const uint32_t s = slot[col];
          if (s != cur) {
            if (cur != 0xFFFFFFFFu) atomicMax(&M[row][cur], x_f2ord(m));
            cur = s;
            m = -INFINITY;
          }
          m = fmaxf(m, acc[i][g * 4 + u]);
        }
        if (cur != 0xFFFFFFFFu) atomicMax(&M[row][cur], x_f2ord(m));
      }
    }
    __syncthreads();
  }
  for (uint32_t e = tid; e < 4 * nit; e += 256) {
    const uint32_t qq = e / nit, it = e % nit, q = q0 + qq;
    if (q >= n_q) continue;
    float sc = 0.0f;
    for (uint32_t t = 0; t < T; ++t) sc += x_ord2f(M[qq * kMaxQTok + t][it]);
    scores[(size_t)q * score_stride + item0 + it] = sc;
  }
}

// The same scores with the dot products on FP32 MFMA (v_mfma_f32_32x32x2_f32):
// dots summed in the matrix unit's order instead of one FMA chain (within
// about 1e-6 relative), the max and the sum over query tokens unchanged.
// Wave w of the workgroup owns query q0 + w: its 32 tokens are the B operand
// (columns), held in 64 registers per lane for the whole block; the item
// tokens are the A operand, 32 rows per MFMA tile and four tiles per 128-token
// LDS tile.  In K-chunk c (dims 32c .. +31) MFMA j takes dim 32c + 16h + j
// from lane half h in both operands.  Lane (query token t = l & 31, h) ends a
// tile with the scores of item-token rows 8g + 4h + e (g, e = 0..3), folded
// into the per-(row, item) LDS maxima by runs of the same item.
typedef float x_floatx16 __attribute__((ext_vector_type(16)));
// LDS row of the MFMA kernel: the two lane halves read rows 16 apart, which
// 16 * 130 = 32 (mod 64) puts on disjoint banks (the VALU kernel's 132 would not)
constexpr uint32_t kXpM = kXt + 2;

__global__ __launch_bounds__(256) void k_exh_maxsim_mfma(const float *__restrict__ Q, uint32_t n_q, uint32_t T,
                                                         const float *__restrict__ X,
                                                         const uint4 *__restrict__ blocks,
                                                         const uint64_t *__restrict__ item_off, uint32_t n_item,
                                                         float *__restrict__ scores, uint64_t score_stride) {
  __shared__ float Bs[kXk][kXpM];
  __shared__ uint32_t M[kXt][kXItems];
  __shared__ uint32_t slot[kXt];
  __shared__ uint64_t ioff[kXItems + 1];
  const uint32_t tid = threadIdx.x, lane = tid & 63, t = lane & 31, h = lane >> 5;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t q0 = blockIdx.y * 4, q = q0 + wave;
  const uint4 blk = blocks[blockIdx.x];
  const uint64_t tok0 = item_off[blk.z];
  const uint32_t ntok = blk.y, item0 = blk.z, nit = blk.w;
  for (uint32_t e = tid; e < kXt * kXItems; e += 256) (&M[0][0])[e] = 0x007FFFFFu;  // order bits of -inf
  for (uint32_t e = tid; e <= nit; e += 256) ioff[e] = item_off[item0 + e] - tok0;
  // the wave's query tokens (rows past n_q read as zero; their scores are not stored)
  float qv[4][16];
  {
    const float *qr = Q + ((size_t)(q < n_q ? q : 0) * kMaxQTok + t) * kDim + 16 * h;
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c)
#pragma unroll
      for (uint32_t j4 = 0; j4 < 4; ++j4) {
        const float4 v = reinterpret_cast<const float4 *>(qr + 32 * c)[j4];
        qv[c][4 * j4] = v.x; qv[c][4 * j4 + 1] = v.y; qv[c][4 * j4 + 2] = v.z; qv[c][4 * j4 + 3] = v.w;
      }
  }
  __syncthreads();
  const float4 *X4 = reinterpret_cast<const float4 *>(X);
  // chunk (tile, c) staged through registers one chunk ahead: its global
  // loads are in flight during the previous chunk's MFMAs
  float4 pf[4];
  auto load_chunk = [&](uint32_t tile0, uint32_t c) {
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t e = tid + 256 * u, r = e >> 3, d4 = e & 7, tk = tile0 + r;
      pf[u] = tk < ntok ? X4[(tok0 + tk) * (kDim / 4) + 8 * c + d4] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  load_chunk(0, 0);
  for (uint32_t tile0 = 0; tile0 < ntok; tile0 += kXt) {
    if (tid < kXt) {
      const uint32_t tk = tile0 + tid;
      uint32_t sl = 0xFFFFFFFFu;
      if (tk < ntok) {
        sl = 0;
        while (sl + 1 < nit && ioff[sl + 1] <= tk) ++sl;
      }
      slot[tid] = sl;
    }
    x_floatx16 acc[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) acc[u] = x_floatx16{};
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) {
      __syncthreads();  // the previous chunk's readers are done
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t e = tid + 256 * u, r = e >> 3, d4 = e & 7;
        Bs[4 * d4 + 0][r] = pf[u].x; Bs[4 * d4 + 1][r] = pf[u].y; Bs[4 * d4 + 2][r] = pf[u].z; Bs[4 * d4 + 3][r] = pf[u].w;
      }
      __syncthreads();
      if (c < 3) load_chunk(tile0, c + 1);
      else if (tile0 + kXt < ntok) load_chunk(tile0 + kXt, 0);
#pragma unroll
      for (uint32_t j = 0; j < 16; ++j) {
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
          acc[u] = __builtin_amdgcn_mfma_f32_32x32x2f32(Bs[16 * h + j][32 * u + t], qv[c][j], acc[u], 0, 0, 0);
      }
    }
    // fold: runs of consecutive rows of the same item, one atomic per run
    if (t < T) {
      uint32_t *Mr = &M[wave * kMaxQTok + t][0];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
#pragma unroll
        for (uint32_t g = 0; g < 4; ++g) {
          uint32_t cur = 0xFFFFFFFFu;
          float m = -INFINITY;
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e) {
            const uint32_t sl = slot[32 * u + 8 * g + 4 * h + e];
            if (sl != cur) {
              if (cur != 0xFFFFFFFFu) atomicMax(&Mr[cur], x_f2ord(m));
              cur = sl;
              m = -INFINITY;
            }
            m = fmaxf(m, acc[u][4 * g + e]);
          }
          if (cur != 0xFFFFFFFFu) atomicMax(&Mr[cur], x_f2ord(m));
        }
      }
    }
    __syncthreads();
  }
  for (uint32_t e = tid; e < 4 * nit; e += 256) {
    const uint32_t qq = e / nit, it = e % nit, qo = q0 + qq;
    if (qo >= n_q) continue;
    float sc = 0.0f;
    for (uint32_t tt = 0; tt < T; ++tt) sc += x_ord2f(M[qq * kMaxQTok + tt][it]);
    scores[(size_t)qo * score_stride + item0 + it] = sc;
  }
}

thread_local std::string g_x_err;

int x_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_x_err = buf;
  return code;
}

}  // namespace
}  // namespace mvr

using namespace mvr;

extern "C" {

const char *mvr_maxsim_last_error(void) { return g_x_err.c_str(); }

int mvr_maxsim_scores_host(const float *Q, uint32_t n_q, uint32_t T, uint32_t dim, const float *X,
                           const uint32_t *lens, uint32_t n_item, float *scores, int device) {
  return mvr_maxsim_scores_host_ex(Q, n_q, T, dim, X, lens, n_item, scores, device, MVR_FP32);
}

int mvr_maxsim_scores_host_ex(const float *Q, uint32_t n_q, uint32_t T, uint32_t dim, const float *X,
                              const uint32_t *lens, uint32_t n_item, float *scores, int device,
                              uint32_t precision) {
  if (precision != MVR_FP32 && precision != MVR_FP32_MFMA)
    return x_fail(MVR_ERR_UNSUPPORTED, "precision %u: MVR_FP32 or MVR_FP32_MFMA", precision);
  if (dim != kDim) return x_fail(MVR_ERR_UNSUPPORTED, "dim must be 128");
  if (T == 0 || T > kMaxQTok) return x_fail(MVR_ERR_UNSUPPORTED, "query length %u must be 1..32", T);
  if (n_q == 0 || n_item == 0) return MVR_OK;
  if (!Q || !X || !lens || !scores) return x_fail(MVR_ERR_INVALID, "NULL argument");
  std::vector<uint64_t> off(n_item + 1, 0);
  for (uint32_t i = 0; i < n_item; ++i) {
    if (lens[i] == 0) return x_fail(MVR_ERR_INVALID, "item %u has no vectors", i);
    off[i + 1] = off[i] + lens[i];
  }
  // column blocks of whole items: up to 48 items and about 512 tokens
  std::vector<uint4> blocks;
  for (uint32_t i = 0; i < n_item;) {
    uint32_t j = i;
    uint64_t tok = 0;
    while (j < n_item && j - i < kXItems && (tok == 0 || tok + lens[j] <= kXTok)) tok += lens[j++];
    if (tok > 0xFFFFFFFFull) return x_fail(MVR_ERR_UNSUPPORTED, "item %u is too long", i);
    blocks.push_back(make_uint4(0, (uint32_t)tok, i, j - i));
    i = j;
  }
  int n_dev = 0;
  hipError_t e = hipGetDeviceCount(&n_dev);
  if (e != hipSuccess) return x_fail(MVR_ERR_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  if (device < 0 || device >= n_dev) return x_fail(MVR_ERR_INVALID, "device %d out of range (%d)", device, n_dev);
  if ((e = hipSetDevice(device)) != hipSuccess) return x_fail(MVR_ERR_DEVICE, "%s", hipGetErrorString(e));

  // items go to the device in chunks of whole column blocks (<= 8M tokens)
  const uint64_t kChunkTok = 1ull << 23;
  float *dQ = nullptr, *dX = nullptr, *dS = nullptr;
  uint4 *dB = nullptr;
  uint64_t *dOff = nullptr;
  hipStream_t s = nullptr;
  auto cleanup = [&]() {
    if (s) (void)hipStreamSynchronize(s);
    for (void *p : {(void *)dQ, (void *)dX, (void *)dS, (void *)dB, (void *)dOff})
      if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  };
  auto dev_fail = [&](const char *what) {
    const hipError_t err = e;
    cleanup();
    return x_fail(MVR_ERR_DEVICE, "%s: %s", what, hipGetErrorString(err));
  };
  uint64_t max_tok = 0;
  for (const uint4 &b : blocks) max_tok = std::max<uint64_t>(max_tok, b.y);
  const uint64_t cap_tok = std::max(kChunkTok, max_tok);
  const size_t q_bytes = sizeof(float) * (size_t)n_q * kMaxQTok * kDim;
  if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return dev_fail("stream");
  if ((e = hipMalloc(&dQ, q_bytes)) != hipSuccess || (e = hipMalloc(&dX, sizeof(float) * cap_tok * kDim)) != hipSuccess ||
      (e = hipMalloc(&dB, sizeof(uint4) * blocks.size())) != hipSuccess ||
      (e = hipMalloc(&dOff, sizeof(uint64_t) * (n_item + 1))) != hipSuccess)
    return dev_fail("allocation");
  // queries padded to 32 rows
  if ((e = hipMemsetAsync(dQ, 0, q_bytes, s)) != hipSuccess ||
      (e = hipMemcpy2DAsync(dQ, sizeof(float) * kMaxQTok * kDim, Q, sizeof(float) * T * kDim,
                            sizeof(float) * T * kDim, n_q, hipMemcpyHostToDevice, s)) != hipSuccess)
    return dev_fail("query upload");
  // chunks: runs of whole column blocks with at most cap_tok tokens
  std::vector<std::pair<size_t, size_t>> chunks;
  uint32_t max_cn = 0;
  for (size_t b0 = 0; b0 < blocks.size();) {
    size_t b1 = b0;
    uint64_t ctok = 0;
    while (b1 < blocks.size() && (b1 == b0 || ctok + blocks[b1].y <= cap_tok)) ctok += blocks[b1++].y;
    chunks.emplace_back(b0, b1);
    max_cn = std::max(max_cn, blocks[b1 - 1].z + blocks[b1 - 1].w - blocks[b0].z);
    b0 = b1;
  }
  if ((e = hipMalloc(&dS, sizeof(float) * (size_t)n_q * max_cn)) != hipSuccess) return dev_fail("allocation");
  for (const auto &ch : chunks) {
    const size_t b0 = ch.first, b1 = ch.second;
    const uint32_t it0 = blocks[b0].z, it1 = blocks[b1 - 1].z + blocks[b1 - 1].w, cn = it1 - it0;
    const uint64_t ctok = off[it1] - off[it0];
    std::vector<uint4> cb(blocks.begin() + b0, blocks.begin() + b1);
    for (uint4 &b : cb) b.z -= it0;
    std::vector<uint64_t> coff(cn + 1);
    for (uint32_t i = 0; i <= cn; ++i) coff[i] = off[it0 + i] - off[it0];
    if ((e = hipMemcpyAsync(dX, X + off[it0] * kDim, sizeof(float) * ctok * kDim, hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(dB, cb.data(), sizeof(uint4) * cb.size(), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(dOff, coff.data(), sizeof(uint64_t) * (cn + 1), hipMemcpyHostToDevice, s)) != hipSuccess)
      return dev_fail("item upload");
    for (uint32_t y0 = 0; y0 < (n_q + 3) / 4; y0 += 65535) {
      const uint32_t gy = std::min<uint32_t>(65535, (n_q + 3) / 4 - y0);
      // query groups [y0, y0 + gy): shift Q and the score rows
      const uint32_t qb = y0 * 4, nqb = std::min<uint32_t>(n_q - qb, gy * 4);
      if (precision == MVR_FP32_MFMA)
        k_exh_maxsim_mfma<<<dim3((uint32_t)cb.size(), gy), 256, 0, s>>>(
            dQ + (size_t)qb * kMaxQTok * kDim, nqb, T, dX, dB, dOff, cn, dS + (size_t)qb * cn, cn);
      else
        k_exh_maxsim<<<dim3((uint32_t)cb.size(), gy), 256, 0, s>>>(dQ + (size_t)qb * kMaxQTok * kDim, nqb, T, dX,
                                                                   dB, dOff, cn, dS + (size_t)qb * cn, cn);
      if ((e = hipGetLastError()) != hipSuccess) return dev_fail("launch");
    }
    // cb / coff live until the stream drains (the copies above read them)
    if ((e = hipMemcpy2DAsync(scores + it0, sizeof(float) * n_item, dS, sizeof(float) * cn, sizeof(float) * cn, n_q,
                              hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return dev_fail("score download");
  }
  cleanup();
  return MVR_OK;
}

}  // extern "C"
