// HIP kernels of the ColBERT late-interaction search path for gfx950.
//
// Stage map (reference file:line of the computation each kernel performs):
//   k_centroid_topk_f32  S'[t][c] = Q_t.C_c + 1 and per-row top-P
//                        (GenerateVQScoreCandidatePP.hpp:137-174)
//   k_filter_tok         IVF probe, per-(doc, token) max and per-doc sum
//                        (FilterPP.hpp:99-142, 366-379)
//   k_select_topr        top-probe_topk by filter score (FilterPP.hpp:386-394)
//   k_rerank_f32         fused 2-bit decompress + MaxSim + sum over tokens
//                        (RerankPP.hpp:65-161, 416-471)
//   k_topk_final         best-k by MaxSim score (IGPGPURetrievalPP.hpp:153-161)
//
// Arithmetic contract (shared with the CPU oracle so results match bit for
// bit in FP32 mode): every dot product is a sequential fused multiply-add
// chain over d = 0..127 from 0.0f; filter sums are exact int64 fixed point
// (2^-32) rounded once to FP32; the MaxSim score is the left-to-right FP32 sum
// over query tokens; all selections order by (score desc, id asc).
#include "mvr_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cfloat>
#include <cmath>
#include <type_traits>

namespace mvr {

namespace {

constexpr uint32_t kWave = 64;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bool better(float s1, uint32_t i1, float s2, uint32_t i2) {
  return (s1 > s2) || (s1 == s2 && i1 < i2);
}

// Order-preserving map of a float to uint32 (larger float -> larger uint).
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// Selection key: (score order bits, ~id), so that larger keys come first and
// equal scores put the lower id first.
__device__ __forceinline__ uint64_t make_key(float s, uint32_t id) {
  return ((uint64_t)f2ord(s) << 32) | (uint64_t)(~id);
}

// ---------------------------------------------------------------------------
// Centroid scoring + top-P, FP32 VALU.  Workgroup = 64 query-token rows x one
// slice of centroids, 256 threads.  The 64 rows live in LDS transposed
// ([d][row]); centroids stream through LDS in tiles of 64 (two halves of 64
// dims).  Thread (tx, ty) accumulates a 4x4 register tile, each output a
// sequential FMA chain over d.  After a tile the 64x64 scores go to LDS and
// each wave maintains the top-P lists of 16 rows, one list element per lane
// (lane j holds the j-th best): a candidate beating element P-1 is inserted
// with one ballot + one shuffle.  MODE 0: score = dot + 1; MODE 1: score =
// dot - bias[c] (k-means / code assignment, P = 1).  Each workgroup writes
// sorted partial lists; k_topk_merge combines the centroid slices.
// ---------------------------------------------------------------------------
constexpr uint32_t kRt = 64;   // rows per workgroup
constexpr uint32_t kCt = 64;   // centroids per tile
constexpr uint32_t kCtChunk = kCt;
constexpr uint32_t kHalf = 64; // dims per LDS half tile

__device__ __forceinline__ float readlane_f(float v, uint32_t l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_centroid_topk_f32(
    const float *__restrict__ Q, const uint32_t *__restrict__ row_map, uint32_t n_rows,
    const float *__restrict__ C, const float *__restrict__ bias, uint32_t n_c,
    uint32_t c_per_split, uint32_t P, uint32_t *__restrict__ part_ids,
    float *__restrict__ part_s) {
  __shared__ __align__(16) float As[kDim][kRt + 4];
  __shared__ __align__(16) float Bs[kHalf][kCt + 4];  // reused as the score tile [64][65]
  const uint32_t tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const uint32_t wave = tid >> 6, lane = tid & 63;
  const uint32_t row0 = blockIdx.x * kRt;

#pragma nounroll
  for (uint32_t k = tid; k < kRt * (kDim / 4); k += 256) {
    const uint32_t r = k >> 5, d4 = k & 31, row = row0 + r;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (row < n_rows) {
      const uint64_t qr = row_map ? row_map[row] : row;
      v = reinterpret_cast<const float4 *>(Q + qr * kDim)[d4];
    }
    As[4 * d4 + 0][r] = v.x;
    As[4 * d4 + 1][r] = v.y;
    As[4 * d4 + 2][r] = v.z;
    As[4 * d4 + 3][r] = v.w;
  }

  float ls[16];
  uint32_t li[16];
#pragma unroll
  for (uint32_t rr = 0; rr < 16; ++rr) { ls[rr] = -INFINITY; li[rr] = 0xFFFFFFFFu; }

  const uint32_t c_begin = blockIdx.y * c_per_split;
  const uint32_t c_end = min(n_c, c_begin + c_per_split);
  float *S = &Bs[0][0];
  constexpr uint32_t kSs = kCt + 1;

#pragma nounroll
  for (uint32_t c0 = c_begin; c0 < c_end; c0 += kCt) {
    float acc[4][4];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i)
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) acc[i][j] = 0.f;
#pragma unroll
    for (uint32_t half = 0; half < 2; ++half) {
      __syncthreads();
#pragma nounroll
      for (uint32_t k = tid; k < kCt * (kHalf / 4); k += 256) {
        const uint32_t cc = k >> 4, d4 = k & 15, c = c0 + cc;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (c < c_end) v = reinterpret_cast<const float4 *>(C + (size_t)c * kDim + half * kHalf)[d4];
        Bs[4 * d4 + 0][cc] = v.x;
        Bs[4 * d4 + 1][cc] = v.y;
        Bs[4 * d4 + 2][cc] = v.z;
        Bs[4 * d4 + 3][cc] = v.w;
      }
      __syncthreads();
#pragma unroll 8
      for (uint32_t d = 0; d < kHalf; ++d) {
        const float4 a = *reinterpret_cast<const float4 *>(&As[half * kHalf + d][ty * 4]);
        const float4 b = *reinterpret_cast<const float4 *>(&Bs[d][tx * 4]);
        const float av[4] = {a.x, a.y, a.z, a.w};
        const float bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i)
#pragma unroll
          for (uint32_t j = 0; j < 4; ++j) acc[i][j] = __builtin_fmaf(av[i], bv[j], acc[i][j]);
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i)
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t c = tx * 4 + j, cid = c0 + c;
        float s;
        if (MODE == 0) s = acc[i][j] + 1.0f;
        else s = bias ? acc[i][j] - bias[cid < n_c ? cid : 0] : acc[i][j];
        S[(ty * 4 + i) * kSs + c] = s;
      }
    __syncthreads();
    const uint32_t cid = c0 + lane;
    const bool cvalid = cid < c_end;
#pragma unroll
    for (uint32_t rr = 0; rr < 16; ++rr) {
      const uint32_t r = wave * 16 + rr;
      if (row0 + r >= n_rows) continue;  // uniform per wave
      const float s = S[r * kSs + lane];
      float ts = readlane_f(ls[rr], P - 1);
      uint32_t ti = readlane_u(li[rr], P - 1);
      uint64_t bal = __ballot(cvalid && better(s, cid, ts, ti));
      while (bal) {
        const uint32_t l = (uint32_t)__builtin_ctzll(bal);
        const float cs = readlane_f(s, l);
        const uint32_t ci = c0 + l;
        const uint32_t pos = (uint32_t)__popcll(__ballot(lane < P && better(ls[rr], li[rr], cs, ci)));
        const float us = __shfl_up(ls[rr], 1);
        const uint32_t ui = __shfl_up(li[rr], 1);
        if (lane > pos) { ls[rr] = us; li[rr] = ui; }
        else if (lane == pos) { ls[rr] = cs; li[rr] = ci; }
        bal &= bal - 1;
        ts = readlane_f(ls[rr], P - 1);
        ti = readlane_u(li[rr], P - 1);
        bal &= __ballot(cvalid && better(s, cid, ts, ti));
      }
    }
  }
#pragma unroll
  for (uint32_t rr = 0; rr < 16; ++rr) {
    const uint32_t row = row0 + wave * 16 + rr;
    if (row < n_rows && lane < P) {
      const size_t o = ((size_t)blockIdx.y * n_rows + row) * P + lane;
      part_ids[o] = li[rr];
      part_s[o] = ls[rr];
    }
  }
}

// Merge n_split sorted partial lists per row into the final top-P list.
__global__ __launch_bounds__(256) void k_topk_merge(const uint32_t *__restrict__ part_ids,
                                                    const float *__restrict__ part_s,
                                                    uint32_t n_split, uint32_t n_rows, uint32_t P,
                                                    const uint32_t *__restrict__ row_map,
                                                    uint32_t *__restrict__ out_ids,
                                                    float *__restrict__ out_s) {
  const uint32_t row = blockIdx.x * 256u + threadIdx.x;
  if (row >= n_rows) return;
  const uint64_t orow = row_map ? row_map[row] : row;
  uint32_t head[16];
  for (uint32_t k = 0; k < n_split; ++k) head[k] = 0;
  for (uint32_t j = 0; j < P; ++j) {
    int bk = -1;
    float bs = -INFINITY;
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t k = 0; k < n_split; ++k) {
      if (head[k] >= P) continue;
      const size_t o = ((size_t)k * n_rows + row) * P + head[k];
      const float s = part_s[o];
      const uint32_t i = part_ids[o];
      if (i == 0xFFFFFFFFu) continue;
      if (bk < 0 || better(s, i, bs, bi)) { bk = (int)k; bs = s; bi = i; }
    }
    if (bk >= 0) head[bk]++;
    out_ids[orow * P + j] = bi;
    out_s[orow * P + j] = bs;
  }
}

// ---------------------------------------------------------------------------
// Filter.  A workgroup owns one query and kFiltGroup consecutive 4,096-doc
// segments; per segment it keeps, in LDS, a 32-bit mask of the query tokens
// already counted per doc and an int64 fixed-point accumulator.  Probe ranks
// are applied in order j = 0..P-1 for each token; because each token's probe
// list is sorted by score, the first rank whose IVF list holds the doc carries
// that token's maximum score, so the atomicOr on the mask decides which hit
// adds (max over probes, then an exact sum over tokens).  Each list's
// sub-range inside a segment comes from the precomputed seg_off table.
//
// Token-major work split: the 16 lanes of query token t -- four tokens to a
// wave -- resolve the token's probe ranks j = sub + 16 s once (centroid, list
// start, fixed-point score) and, per segment, write the run [start, len) of
// each of their ranks into an LDS table that only the same 16 lanes read, so
// the tables need no barrier, and neither do the ranks: a token's runs belong
// to lanes of one wave, whose LDS operations execute in program order, so
// rank r's mask bits are set before rank r + 1 tests them.  A segment's ranks
// are applied in batches of eight; the first kFiltHold entries of every run of
// the next batch (of this segment, or the first batch of the next one) are
// loaded into registers while the current batch is applied, and the rest of a
// long run is loaded as it is applied.  The run tables are double-buffered
// across segments, their columns swizzled (token ^ rank) so the per-segment
// writes spread over the banks; the sub-range bounds are loaded one segment
// before their table is written.
//
// XCD-aware order: consecutive workgroup ids go to the 8 XCDs in turn, so the
// 1-D id is remapped to give each XCD a contiguous run of (query, segment
// group) tiles; neighbouring tiles of one query then share the L2 lines of
// the seg_off rows and of the query's probe lists.
//
// Candidates go to the query's region of the candidate pool (keys_per_q
// 64-bit selection keys per query: at most the docs its probed lists hold) as
// one contiguous run per (segment, wave), reserved with one atomic on the
// query's fill counter q_fill[q].  Runs land in any order (the selection is
// by key value).  The emission clears the LDS state it reads, ready for the
// next segment.
// ---------------------------------------------------------------------------
constexpr uint32_t kFiltGroup = 8;                    // segments per workgroup
constexpr uint32_t kFiltHold = 1;                     // list entries per lane and rank loaded ahead
constexpr uint32_t kFtLanes = 16;                     // lanes per query token
constexpr uint32_t kFtThreads = kMaxQTok * kFtLanes;  // threads per workgroup
constexpr uint32_t kFtBatch = 8;                      // ranks per load batch
constexpr uint32_t kFtTokMax = 64;                    // most ranks of the token-major kernel

// Run-table slot of (rank r, token t): rows of 32 tokens, columns swizzled by
// the rank so one rank's writes spread over the banks.
__device__ __forceinline__ uint32_t ft_slot(uint32_t r, uint32_t t) {
  return r * kMaxQTok + (t ^ (r & 31u));
}

// Rank order inside a wave: a token's ranks are applied by lanes of one wave,
// one rank after the other, and the first rank whose list holds a doc must
// set the doc's token bit before the next rank tests it.  The LDS executes a
// wave's operations in issue order; this fence keeps the compiler from
// moving one rank's LDS atomics past the next rank's.
__device__ __forceinline__ void rank_order_fence() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// Fixed-point score of a probed list: S' * 2^32.  The PP filter (clamp0)
// takes max(0, .) per (doc, token), as the reference's zero-filled maxima
// with atomicMax do (FilterPP.hpp:136-141, memset at :257-259); the basic
// variant adds the first hit as it is (igp_gpu_basic/Filter.hpp:44-59).
__device__ __forceinline__ long long filter_fx(float s, int clamp0) {
  const long long fx = (long long)((double)s * kFxScale);
  return (clamp0 && fx < 0) ? 0 : fx;
}

// Emission of one segment: slot k of every lane in turn; a wave's hits
// (slot-major, then lane) are compacted into a list of doc offsets (hl, kSeg
// 16-bit entries per workgroup), so the conversion and key packing run once
// per hit rather than once per slot.  The segment's keys are one contiguous
// run in the query's key region, reserved by one atomic on q_fill[q].  wsum:
// the per-wave counts, then the run's
// start (slot kFtThreads / kWave).  The caller's barrier separates the last
// LDS update of the segment from this; the emission ends with a barrier and
// leaves mask and acc cleared.
__device__ __forceinline__ void filter_emit(uint32_t *mask, unsigned long long *acc, uint32_t *wsum,
                                            uint16_t *hl_all, uint32_t q, uint32_t sg, uint32_t n_doc,
                                            uint32_t *__restrict__ q_fill, uint64_t keys_per_q,
                                            uint64_t *__restrict__ cand_keys) {
  constexpr uint32_t kPer = kSeg / kFtThreads;  // slots per lane
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
  const uint32_t seg_lo = sg * kSeg;
  uint16_t *hl = hl_all + wv * (kSeg / (kFtThreads / kWave));
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  const bool tail = seg_lo + kSeg > n_doc;  // the last segment, partly past the corpus
  uint32_t wcount = 0;  // hits of this wave so far (uniform)
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t dl = lane + kWave * (wv + (kFtThreads / kWave) * k);
    bool hit = false;
    if (mask[dl] != 0u) {
      mask[dl] = 0u;
      hit = !tail || seg_lo + dl < n_doc;
      if (!hit) acc[dl] = 0ull;
    }
    const uint64_t bb = __ballot(hit);
    if (hit) hl[wcount + __popcll(bb & lt_mask)] = (uint16_t)dl;
    wcount += (uint32_t)__popcll(bb);
  }
  if (lane == 0) wsum[wv] = wcount;
  __syncthreads();
  uint32_t pos = 0;
  for (uint32_t w = 0; w < wv; ++w) pos += wsum[w];
  if (tid == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < kFtThreads / kWave; ++w) tot += wsum[w];
    wsum[kFtThreads / kWave] = tot ? atomicAdd(&q_fill[q], tot) : 0u;
  }
  __syncthreads();  // the run's start
  const uint64_t base = (uint64_t)wsum[kFtThreads / kWave] + pos;
  uint64_t *out = cand_keys + (size_t)q * keys_per_q + base;
  for (uint32_t i = lane; i < wcount; i += kWave) {
    const uint32_t dl = hl[i];
    const float F = (float)((double)(long long)acc[dl] * (1.0 / kFxScale));
    acc[dl] = 0ull;
    // (the region holds every doc the query's lists can reach; the check
    // keeps a wrong bound from writing past it)
    if (base + i < keys_per_q) out[i] = make_key(F, seg_lo + dl);
  }
  __syncthreads();  // wsum and the cleared state before the next segment
}

// The XCD-aware workgroup order of the filter kernels: consecutive ids go to
// the 8 XCDs in turn, so id b is remapped to give each XCD a contiguous run of
// (query, segment group) tiles.
__device__ __forceinline__ uint32_t filter_tile(uint32_t n_wg) {
  const uint32_t per_xcd = (n_wg + 7) / 8;
  return (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
}

// Occupancy: two workgroups (4 waves per SIMD) with 32-rank tables; the
// 64-rank tables leave room for one.
template <uint32_t NS>
__global__ __launch_bounds__(kFtThreads) __attribute__((amdgpu_waves_per_eu(NS <= 2 ? 4 : 2))) void k_filter_tok(
    const uint32_t *__restrict__ top_ids, const float *__restrict__ top_s, uint32_t t_max,
    uint32_t P, const uint32_t *__restrict__ q_ntok, const uint32_t *__restrict__ q_np,
    const uint64_t *__restrict__ ivf_off, const uint32_t *__restrict__ ivf_ids,
    const uint32_t *__restrict__ seg_off, uint32_t n_seg, uint32_t n_doc, uint32_t n_q, int clamp0,
    uint32_t *__restrict__ q_fill, uint64_t keys_per_q, uint64_t *__restrict__ cand_keys) {
  // run tables for kR ranks (32, or 64 when a lane holds four ranks)
  constexpr uint32_t kR = NS * kFtLanes > 32 ? NS * kFtLanes : 32;
  static_assert(NS * kFtLanes <= kFtTokMax, "ranks per lane");
  static_assert(kR * kMaxQTok * 8 >= kSeg * 2, "the emission's hit list lives in a run table");
  __shared__ uint32_t mask[kSeg];
  __shared__ unsigned long long acc[kSeg];
  __shared__ uint64_t tb_st[2][kR * kMaxQTok];
  __shared__ uint32_t tb_len[2][kR * kMaxQTok];
  __shared__ long long tb_fx[kR * kMaxQTok];

  const uint32_t n_grp = (n_seg + kFiltGroup - 1) / kFiltGroup;
  const uint32_t n_wg = n_grp * n_q;
  const uint32_t L = filter_tile(n_wg);
  if (L >= n_wg) return;  // uniform for the whole workgroup
  const uint32_t q = L / n_grp, grp = L - q * n_grp;
  const uint32_t sg0 = grp * kFiltGroup, sg1 = min(n_seg, sg0 + kFiltGroup);
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  const uint32_t Pq = q_np ? q_np[q] : P;
  const uint32_t tid = threadIdx.x;
  const uint32_t t = tid / kFtLanes, sub = tid % kFtLanes;
  const uint32_t bit = 1u << t;
  const uint32_t nr = min(Pq, kFtLanes * NS);
  const uint32_t nb = (nr + kFtBatch - 1) / kFtBatch;

  // this lane's ranks j = sub + kFtLanes s: list, score, and the run bounds of the
  // next segment (hi) and the one after it (nhi)
  const uint32_t *so[NS];
  uint64_t base[NS];
  uint32_t hi[NS], nhi[NS];
#pragma unroll
  for (uint32_t s = 0; s < NS; ++s) {
    const uint32_t j = sub + kFtLanes * s;
    so[s] = nullptr;
    base[s] = 0;
    long long fx = 0;
    if (j < Pq && t < T) {
      const size_t o = ((size_t)q * t_max + t) * P + j;
      const uint32_t c = top_ids[o];
      if (c != 0xFFFFFFFFu) {
        so[s] = seg_off + (size_t)c * (n_seg + 1);
        base[s] = ivf_off[c];
        fx = filter_fx(top_s[o], clamp0);
      }
    }
    tb_fx[ft_slot(j, t)] = fx;
    uint32_t lo = 0;
    hi[s] = nhi[s] = 0;
    if (so[s]) {
      lo = so[s][sg0];
      hi[s] = so[s][sg0 + 1];
      nhi[s] = sg0 + 1 < sg1 ? so[s][sg0 + 2] : hi[s];
    }
    tb_st[0][ft_slot(j, t)] = base[s] + lo;
    tb_len[0][ft_slot(j, t)] = hi[s] - lo;
  }

  // first kFiltHold entries of the runs of ranks [rb, rb + 8) from table buffer tb
  auto load_batch = [&](uint32_t tb, uint32_t rb, uint32_t (&ent)[kFtBatch][kFiltHold],
                        uint32_t (&len)[kFtBatch]) {
#pragma unroll
    for (uint32_t u = 0; u < kFtBatch; ++u) {
      const uint32_t r = rb + u;
      uint32_t l = 0;
      uint64_t st = 0;
      if (r < nr) {
        l = tb_len[tb][ft_slot(r, t)];
        st = tb_st[tb][ft_slot(r, t)];
      }
      len[u] = l;
#pragma unroll
      for (uint32_t k = 0; k < kFiltHold; ++k) {
        const uint32_t i = sub + kFtLanes * k;
        ent[u][k] = (i < l) ? ivf_ids[st + i] : 0xFFFFFFFFu;
      }
    }
  };

  uint32_t ent[kFtBatch][kFiltHold], len[kFtBatch];
  load_batch(0, 0, ent, len);
  for (uint32_t i = tid; i < kSeg; i += kFtThreads) { mask[i] = 0u; acc[i] = 0ull; }
  __syncthreads();

  for (uint32_t sg = sg0; sg < sg1; ++sg) {
    const uint32_t b = (sg - sg0) & 1u;
    for (uint32_t ib = 0; ib < nb; ++ib) {
      // the next batch's first entries, in flight while this batch is applied
      uint32_t ent_n[kFtBatch][kFiltHold], len_n[kFtBatch];
      if (ib + 1 < nb) {
        load_batch(b, (ib + 1) * kFtBatch, ent_n, len_n);
      } else if (sg + 1 < sg1) {
        // the next segment's runs into the other table buffer
#pragma unroll
        for (uint32_t s = 0; s < NS; ++s) {
          const uint32_t j = sub + kFtLanes * s;
          tb_st[b ^ 1u][ft_slot(j, t)] = base[s] + hi[s];
          tb_len[b ^ 1u][ft_slot(j, t)] = nhi[s] - hi[s];
          hi[s] = nhi[s];
          if (so[s] && sg + 2 < sg1) nhi[s] = so[s][sg + 3];
        }
        load_batch(b ^ 1u, 0, ent_n, len_n);
      } else {
#pragma unroll
        for (uint32_t u = 0; u < kFtBatch; ++u) {
          len_n[u] = 0;
#pragma unroll
          for (uint32_t k = 0; k < kFiltHold; ++k) ent_n[u][k] = 0xFFFFFFFFu;
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < kFtBatch; ++u) {
        const uint32_t r = ib * kFtBatch + u;
        if (r < nr) {
          const unsigned long long add = (unsigned long long)tb_fx[ft_slot(r, t)];
#pragma unroll
          for (uint32_t k = 0; k < kFiltHold; ++k) {
            if (ent[u][k] != 0xFFFFFFFFu) {
              const uint32_t dl = ent[u][k] & (kSeg - 1);  // segments start at multiples of kSeg
              const uint32_t old = atomicOr(&mask[dl], bit);
              if (!(old & bit)) atomicAdd(&acc[dl], add);
            }
          }
          // long runs: the rest of the entries, loaded here
          if (len[u] > kFtLanes * kFiltHold) {
            const uint64_t st = tb_st[b][ft_slot(r, t)];
            for (uint32_t i = sub + kFtLanes * kFiltHold; i < len[u]; i += kFtLanes) {
              const uint32_t dl = ivf_ids[st + i] & (kSeg - 1);
              const uint32_t old = atomicOr(&mask[dl], bit);
              if (!(old & bit)) atomicAdd(&acc[dl], add);
            }
          }
        }
        rank_order_fence();
      }
#pragma unroll
      for (uint32_t u = 0; u < kFtBatch; ++u) {
        len[u] = len_n[u];
#pragma unroll
        for (uint32_t k = 0; k < kFiltHold; ++k) ent[u][k] = ent_n[u][k];
      }
    }
    __syncthreads();
    // the emission's per-wave counts and hit lists go to this segment's run
    // tables, whose last reader was the loop above and whose next writer runs
    // after the emission's closing barrier
    filter_emit(mask, acc, tb_len[b], reinterpret_cast<uint16_t *>(tb_st[b]), q, sg, n_doc, q_fill, keys_per_q,
                cand_keys);
  }
}

// Filter for nprobe above the token-major kernel's 64 ranks.  Same LDS state,
// grid and emission; the ranks go in groups of 32: per segment and group, the
// 16 lanes of token t resolve two ranks each (centroid, score, the list's run
// in the segment from seg_off) into a run table, then apply the group's ranks
// one after the other, the lanes striding over each run.  Groups and ranks
// keep their order, so the first hit of a doc is still the token's maximum.
__global__ __launch_bounds__(kFtThreads) void k_filter_wide(
    const uint32_t *__restrict__ top_ids, const float *__restrict__ top_s, uint32_t t_max,
    uint32_t P, const uint32_t *__restrict__ q_ntok, const uint32_t *__restrict__ q_np,
    const uint64_t *__restrict__ ivf_off, const uint32_t *__restrict__ ivf_ids,
    const uint32_t *__restrict__ seg_off, uint32_t n_seg, uint32_t n_doc, uint32_t n_q, int clamp0,
    uint32_t *__restrict__ q_fill, uint64_t keys_per_q, uint64_t *__restrict__ cand_keys) {
  constexpr uint32_t kG = 2 * kFtLanes;  // ranks per group
  __shared__ uint32_t wsum[kFtThreads / kWave + 1];  // per-wave counts, then the segment's run start
  __shared__ uint32_t mask[kSeg];
  __shared__ unsigned long long acc[kSeg];
  __shared__ uint64_t tb_st[kG * kMaxQTok];  // also the emission's hit list
  __shared__ uint32_t tb_len[kG * kMaxQTok];
  __shared__ long long tb_fx[kG * kMaxQTok];
  static_assert(kG * kMaxQTok * 8 >= kSeg * 2, "hit list size");

  const uint32_t n_grp = (n_seg + kFiltGroup - 1) / kFiltGroup;
  const uint32_t n_wg = n_grp * n_q;
  const uint32_t L = filter_tile(n_wg);
  if (L >= n_wg) return;
  const uint32_t q = L / n_grp, grp = L - q * n_grp;
  const uint32_t sg0 = grp * kFiltGroup, sg1 = min(n_seg, sg0 + kFiltGroup);
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  const uint32_t nr = min(q_np ? q_np[q] : P, P);
  const uint32_t tid = threadIdx.x;
  const uint32_t t = tid / kFtLanes, sub = tid % kFtLanes;
  const uint32_t bit = 1u << t;

  for (uint32_t i = tid; i < kSeg; i += kFtThreads) { mask[i] = 0u; acc[i] = 0ull; }
  __syncthreads();
  for (uint32_t sg = sg0; sg < sg1; ++sg) {
    for (uint32_t g0 = 0; g0 < nr; g0 += kG) {
#pragma unroll
      for (uint32_t s = 0; s < 2; ++s) {
        const uint32_t r = sub + kFtLanes * s, j = g0 + r;
        uint64_t st = 0;
        uint32_t len = 0;
        long long fx = 0;
        if (j < nr && t < T) {
          const size_t o = ((size_t)q * t_max + t) * P + j;
          const uint32_t c = top_ids[o];
          if (c != 0xFFFFFFFFu) {
            const uint32_t *so = seg_off + (size_t)c * (n_seg + 1);
            const uint32_t lo = so[sg];
            st = ivf_off[c] + lo;
            len = so[sg + 1] - lo;
            fx = filter_fx(top_s[o], clamp0);
          }
        }
        tb_st[ft_slot(r, t)] = st;
        tb_len[ft_slot(r, t)] = len;
        tb_fx[ft_slot(r, t)] = fx;
      }
      wave_lds_sync();  // the token's 16 lanes read each other's table entries
      const uint32_t ng = min(kG, nr - g0);
      for (uint32_t r = 0; r < ng; ++r) {
        const uint32_t l = tb_len[ft_slot(r, t)];
        if (l > sub) {
          const uint64_t st = tb_st[ft_slot(r, t)];
          const unsigned long long add = (unsigned long long)tb_fx[ft_slot(r, t)];
          for (uint32_t i = sub; i < l; i += kFtLanes) {
            const uint32_t dl = ivf_ids[st + i] & (kSeg - 1);
            const uint32_t old = atomicOr(&mask[dl], bit);
            if (!(old & bit)) atomicAdd(&acc[dl], add);
          }
        }
        rank_order_fence();
      }
      wave_lds_sync();  // table reads done before the next group rewrites them
    }
    __syncthreads();
    filter_emit(mask, acc, wsum, reinterpret_cast<uint16_t *>(tb_st), q, sg, n_doc, q_fill, keys_per_q, cand_keys);
  }
}

// ---------------------------------------------------------------------------
// Bitonic sort (descending) of n (power of two) 64-bit keys in LDS.
// ---------------------------------------------------------------------------
__device__ void bitonic_desc(unsigned long long *s, uint32_t n) {
  for (uint32_t k = 2; k <= n; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const unsigned long long a = s[i], b = s[ixj];
          const bool up = (i & k) == 0;
          if (up ? (a < b) : (a > b)) { s[i] = b; s[ixj] = a; }
        }
      }
      __syncthreads();
    }
  }
}

__device__ __forceinline__ uint32_t pow2_ceil(uint32_t n) {
  uint32_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// ---------------------------------------------------------------------------
// Top-R selection per query over its candidate keys: MSB-first radix select
// (11-bit digits, LDS histogram), keys above the pivot digit go straight to
// the output; the pivot bucket is refined until it fits in LDS, then sorted.
// ---------------------------------------------------------------------------
constexpr uint32_t kSelThreads = 512;
constexpr uint32_t kSelSmall = 4096;  // key capacity of the common variant (R <= 4096)

// CAP: keys held in LDS (R <= CAP).  The 4,096-key variant keeps a 2,048-key
// pivot scratch and indexes up to 4,096 segments (16.7M docs); the
// 16,384-key one (128 KB of keys) a 1,024-key scratch and 2,048 segments,
// walking larger lists segment by segment.
// SEG: the candidate keys of query q lie in per-segment regions (seg_cnt[q][sg]
// keys at keys + sg * kSeg, the PLAID filter's layout); otherwise they are
// contiguous (q_fill[q] keys from q * keys_per_q, the IGP filter's layout).
template <uint32_t CAP, bool SEG>
__global__ __launch_bounds__(kSelThreads) void k_select_topr(
    const uint64_t *__restrict__ cand_keys, const uint32_t *__restrict__ seg_cnt, uint32_t n_seg,
    const uint32_t *__restrict__ q_fill, uint64_t keys_per_q,
    const uint32_t *__restrict__ q_r, uint32_t R_out, uint32_t *__restrict__ sel_ids,
    float *__restrict__ sel_F, uint32_t *__restrict__ sel_n, uint32_t *__restrict__ cand_tot) {
  constexpr uint32_t kSelScratch = CAP > kSelSmall ? 1024 : 2048;
  constexpr uint32_t kSelMaxSeg = SEG ? (CAP > kSelSmall ? 2048 : 4096) : 1;
  __shared__ unsigned long long outk[CAP];
  __shared__ unsigned long long scr[kSelScratch];
  __shared__ uint32_t hist[2048];
  __shared__ uint32_t spfx[kSelMaxSeg + 1];  // exclusive prefix of the segment counts
  __shared__ uint32_t s_out_n, s_scr_n, s_digit, s_gt, s_eq;
  __shared__ uint32_t s_n, s_wsum[kSelThreads / kWave];

  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  const uint32_t sel_lane = tid & (kWave - 1);
  const uint64_t sel_lt_mask = sel_lane == 0 ? 0ull : (~0ull >> (kWave - sel_lane));
  const uint32_t *cnt = SEG ? seg_cnt + (size_t)q * n_seg : nullptr;
  const uint64_t *keys = SEG ? cand_keys + (size_t)q * n_seg * kSeg : cand_keys + (size_t)q * keys_per_q;
  const bool flat = !SEG || n_seg <= kSelMaxSeg;
  if constexpr (SEG) {
    // the query's candidates: cnt[sg] keys at keys + sg * kSeg for every segment
    if (flat) {
      // prefix sum of the counts: each thread sums a run of segments, then a
      // scan over the threads' run totals
      const uint32_t per = (n_seg + kSelThreads - 1) / kSelThreads;
      const uint32_t s0 = min(n_seg, tid * per), s1 = min(n_seg, s0 + per);
      uint32_t part = 0;
      for (uint32_t sg = s0; sg < s1; ++sg) part += cnt[sg];
      const uint32_t lane = tid & (kWave - 1), wv = tid / kWave;
      uint32_t incl = part;
#pragma unroll
      for (uint32_t o = 1; o < kWave; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
      }
      if (lane == kWave - 1) s_wsum[wv] = incl;
      __syncthreads();
      uint32_t run = incl - part;
      for (uint32_t w = 0; w < wv; ++w) run += s_wsum[w];
      for (uint32_t sg = s0; sg < s1; ++sg) {
        spfx[sg] = run;
        run += cnt[sg];
      }
      if (tid == kSelThreads - 1) {
        spfx[n_seg] = run;
        s_n = run;
      }
    } else {
      if (tid == 0) s_n = 0;
      __syncthreads();
      uint32_t part = 0;
      for (uint32_t sg = tid; sg < n_seg; sg += kSelThreads) part += cnt[sg];
      if (part) atomicAdd(&s_n, part);
    }
  } else {
    if (tid == 0) s_n = (uint32_t)min((uint64_t)q_fill[q], keys_per_q);
  }
  __syncthreads();
  const uint32_t n = s_n;
  if (tid == 0 && cand_tot) cand_tot[q] = n;
  // key i of the query's list: contiguous, or (segment layout) in the
  // segment sg with spfx[sg] <= i < spfx[sg + 1], found with a cursor that
  // only moves forward
  auto key_at = [&](uint32_t i, uint32_t &sg) -> uint64_t {
    if constexpr (!SEG) {
      return keys[i];
    } else {
      while (spfx[sg + 1] <= i) ++sg;
      return keys[(size_t)sg * kSeg + (i - spfx[sg])];
    }
  };
  // visit every candidate key: thread tid takes i = tid, tid + kSelThreads,
  // ..., eight loads in flight; a segment layout with too many segments for
  // the prefix table goes segment by segment
  auto for_each_key = [&](auto &&body) {
    if (flat) {
      constexpr uint32_t U = 8;
      uint32_t sg = 0;
      for (uint32_t i0 = tid; i0 < n; i0 += kSelThreads * U) {
        uint64_t kv[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          const uint32_t i = i0 + u * kSelThreads;
          if (i < n) kv[u] = key_at(i, sg);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
          if (i0 + u * kSelThreads < n) body(kv[u]);
      }
    } else {
      for (uint32_t sg = 0; sg < n_seg; ++sg) {
        const uint32_t c = cnt[sg];
        const uint64_t *kk = keys + (size_t)sg * kSeg;
        for (uint32_t i = tid; i < c; i += kSelThreads) body(kk[i]);
      }
    }
  };
  // R_out is the row stride of the outputs; q_r (nullable) lowers R per query
  const uint32_t R = q_r ? max(1u, min(R_out, q_r[q])) : R_out;
  const uint32_t n_sel = min(n, R);

  if (tid == 0) { s_out_n = 0; s_scr_n = 0; }
  __syncthreads();

  if (n <= CAP) {
    const uint32_t np = pow2_ceil(max(n, 2u));
    if (flat) {
      uint32_t sg = 0;
      for (uint32_t i = tid; i < n; i += kSelThreads) outk[i] = key_at(i, sg);
    } else {
      uint32_t base = 0;
      for (uint32_t sg = 0; sg < n_seg; ++sg) {
        const uint32_t c = cnt[sg];
        const uint64_t *kk = keys + (size_t)sg * kSeg;
        for (uint32_t i = tid; i < c; i += kSelThreads) outk[base + i] = kk[i];
        base += c;
      }
    }
    for (uint32_t i = n + tid; i < np; i += kSelThreads) outk[i] = 0ull;
    __syncthreads();
    bitonic_desc(outk, np);
  } else {
    // n > CAP >= R: the pivot search below always finds a digit.
    uint64_t prefix = 0;
    uint32_t shift = 64;
    uint32_t need = R;
    bool done = false;
    while (!done) {
      const uint32_t bits = shift >= 11 ? 11u : shift;
      const uint32_t hi_shift = shift;  // bits above the current digit
      shift -= bits;
      const uint32_t nb = 1u << bits;
      for (uint32_t b = tid; b < nb; b += kSelThreads) hist[b] = 0;
      __syncthreads();
      if (hi_shift == 64) {
        // the first digit (sign, exponent and two mantissa bits of the
        // score) crowds into a few bins: the lanes of a wave holding the same
        // digit add their count once
        for_each_key([&](uint64_t k) {
          const uint32_t dg = (uint32_t)(k >> shift) & (nb - 1);
          uint64_t m = __ballot(true);
          while (m) {
            const uint32_t d0 = (uint32_t)__shfl((int)dg, (int)__builtin_ctzll(m));
            const uint64_t same = __ballot(dg == d0);
            if (dg == d0 && (same & sel_lt_mask) == 0) atomicAdd(&hist[d0], (uint32_t)__popcll(same));
            m &= ~same;
          }
        });
      } else {
        for_each_key([&](uint64_t k) {
          if ((k >> hi_shift) == prefix) atomicAdd(&hist[(uint32_t)(k >> shift) & (nb - 1)], 1u);
        });
      }
      __syncthreads();
      if (tid == 0) {
        uint32_t cum = 0, b = nb;
        while (b > 0) {
          --b;
          if (cum + hist[b] >= need) break;
          cum += hist[b];
        }
        s_digit = b; s_gt = cum; s_eq = hist[b];
      }
      __syncthreads();
      const uint32_t dsel = s_digit, gt = s_gt, eq = s_eq;
      const bool take_eq = (gt + eq == need);
      const bool eq_fits = eq <= kSelScratch;
      for_each_key([&](uint64_t k) {
        if (hi_shift == 64 || (k >> hi_shift) == prefix) {
          const uint32_t dg = (uint32_t)(k >> shift) & (nb - 1);
          if (dg > dsel || (dg == dsel && take_eq)) outk[atomicAdd(&s_out_n, 1u)] = k;
          else if (dg == dsel && eq_fits) scr[atomicAdd(&s_scr_n, 1u)] = k;
        }
      });
      __syncthreads();
      need -= gt;
      if (take_eq) {
        done = true;
      } else if (shift == 0 && !eq_fits) {
        // only reachable with duplicate keys: take the first `need` of them
        for_each_key([&](uint64_t k) {
          if ((k >> hi_shift) == prefix && (uint32_t)(k & (nb - 1)) == dsel) {
            const uint32_t slot = atomicAdd(&s_scr_n, 1u);
            if (slot < need) outk[s_out_n + slot] = k;
          }
        });
        __syncthreads();
        done = true;
      } else if (eq_fits) {
        const uint32_t ns = s_scr_n;
        const uint32_t np = pow2_ceil(max(ns, 2u));
        for (uint32_t i = ns + tid; i < np; i += kSelThreads) scr[i] = 0ull;
        __syncthreads();
        bitonic_desc(scr, np);
        const uint32_t base = s_out_n;
        for (uint32_t i = tid; i < need; i += kSelThreads) outk[base + i] = scr[i];
        __syncthreads();
        done = true;
      } else {
        prefix = (prefix << bits) | dsel;
      }
    }
    const uint32_t np = pow2_ceil(max(n_sel, 2u));
    for (uint32_t i = n_sel + tid; i < np; i += kSelThreads) outk[i] = 0ull;
    __syncthreads();
    bitonic_desc(outk, np);
  }
  for (uint32_t i = tid; i < R_out; i += kSelThreads) {
    const size_t o = (size_t)q * R_out + i;
    if (i < n_sel) {
      const unsigned long long k = outk[i];
      sel_ids[o] = ~(uint32_t)(k & 0xFFFFFFFFull);
      sel_F[o] = ord2f((uint32_t)(k >> 32));
    } else {
      sel_ids[o] = 0xFFFFFFFFu;
      sel_F[o] = -INFINITY;
    }
  }
  if (tid == 0) sel_n[q] = n_sel;
}

// ---------------------------------------------------------------------------
// Fused decompress + MaxSim, FP32.  Workgroup (slice, query), 4 waves; each
// wave takes candidate docs r = slice*4 + wave, stepping by 4*n_slice.  Lane
// l owns query token t = l & 31 (its 128 values in VGPRs) and parity
// h = l >> 5.  A step covers eight doc tokens: lane l decodes dims 4t..4t+3 of
// tokens h, h+2, h+4, h+6 straight from the packed codes into LDS (code word,
// one residual byte, one float4 of the centroid row each), then runs four
// independent FMA chains -- its query token against those four tokens --
// reading the decoded tokens as LDS broadcasts.  Each dot product is still
// one sequential chain over d, so results are bit-identical to the oracle.
// No decompressed embedding touches HBM.  NB = residual bits per dim: 2 keeps
// the decode weights of the lane's four dims in VGPRs; 1, 4 and 8 read the
// lane's NB/2 bytes and look each value up in the [128][2^NB] table in global
// memory (the reference's decode for 8/NB values per byte,
// DataStructPP.hpp:184-200, RerankPP.hpp:100-108).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pick4(uint32_t b, float w0, float w1, float w2, float w3) {
  const float lo = (b & 1u) ? w1 : w0;
  const float hi = (b & 1u) ? w3 : w2;
  return (b & 2u) ? hi : lo;
}

constexpr uint32_t kRrTok = 8;   // doc tokens per step
constexpr uint32_t kRrChains = kRrTok / 2;

template <uint32_t NB>
__global__ __launch_bounds__(256) void k_rerank_f32(
    const float *__restrict__ Q, uint32_t t_max, const uint32_t *__restrict__ q_ntok,
    const uint32_t *__restrict__ sel_ids, const uint32_t *__restrict__ sel_n, uint32_t R,
    uint32_t n_slice, const uint32_t *__restrict__ doclens, const uint64_t *__restrict__ docoff,
    const uint32_t *__restrict__ codes, const uint8_t *__restrict__ residual,
    const float *__restrict__ C, const float *__restrict__ W, int clip_zero,
    float *__restrict__ sel_score) {
  __shared__ __align__(16) float4 xs[4][kRrTok][kDim / 4];
  __shared__ float Ms[4][kMaxQTok];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, t = lane & 31;
  const uint32_t slice = blockIdx.x, q = blockIdx.y;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  const uint32_t n = sel_n[q];

  float qv[kDim];
  {
    const float4 *q4 = reinterpret_cast<const float4 *>(Q + ((size_t)q * t_max + t) * kDim);
#pragma unroll
    for (uint32_t i = 0; i < kDim / 4; ++i) {
      const float4 v = (t < T) ? q4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      qv[4 * i] = v.x; qv[4 * i + 1] = v.y; qv[4 * i + 2] = v.z; qv[4 * i + 3] = v.w;
    }
  }
  // decode weights of dims 4t..4t+3 (2-bit codes)
  constexpr uint32_t kRb = kDim * NB / 8, kNq = 1u << NB, kMask = kNq - 1;
  float w[4][4];
  if constexpr (NB == 2) {
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i)
#pragma unroll
      for (uint32_t b = 0; b < 4; ++b) w[i][b] = W[(4 * t + i) * 4 + b];
  }

  for (uint32_t r = slice * 4 + wave; r < n; r += n_slice * 4) {
    const uint32_t doc = sel_ids[(size_t)q * R + r];
    const uint32_t L = doclens[doc];
    const uint64_t off = docoff[doc];
    float M = clip_zero ? 0.0f : -FLT_MAX;
    for (uint32_t i = 0; i < L; i += kRrTok) {
#pragma unroll
      for (uint32_t c = 0; c < kRrChains; ++c) {
        const uint32_t tk = 2 * c + h;
        if (i + tk < L) {
          const uint64_t g = off + i + tk;
          const uint32_t code = codes[g];
          const float4 cv = reinterpret_cast<const float4 *>(C + (size_t)code * kDim)[t];
          float4 x;
          if constexpr (NB == 2) {
            const uint32_t byte = residual[g * kRb + t];
            x.x = cv.x + pick4((byte >> 6) & 3u, w[0][0], w[0][1], w[0][2], w[0][3]);
            x.y = cv.y + pick4((byte >> 4) & 3u, w[1][0], w[1][1], w[1][2], w[1][3]);
            x.z = cv.z + pick4((byte >> 2) & 3u, w[2][0], w[2][1], w[2][2], w[2][3]);
            x.w = cv.w + pick4(byte & 3u, w[3][0], w[3][1], w[3][2], w[3][3]);
          } else {
            // the 4*NB bits of dims 4t..4t+3, MSB-first from bit 4t*NB of the row
            uint32_t bits;
            const uint8_t *row = residual + g * kRb;
            if constexpr (NB == 1) bits = (row[t >> 1] >> ((t & 1u) ? 0 : 4)) & 0xFu;
            else if constexpr (NB == 4) bits = ((uint32_t)row[2 * t] << 8) | row[2 * t + 1];
            else bits = __builtin_bswap32(reinterpret_cast<const uint32_t *>(row)[t]);
            const float *w0 = W + (size_t)(4 * t) * kNq;
            x.x = cv.x + w0[(bits >> (3 * NB)) & kMask];
            x.y = cv.y + w0[kNq + ((bits >> (2 * NB)) & kMask)];
            x.z = cv.z + w0[2 * kNq + ((bits >> NB) & kMask)];
            x.w = cv.w + w0[3 * kNq + (bits & kMask)];
          }
          xs[wave][tk][t] = x;
        }
      }
      wave_lds_sync();
      float a[kRrChains];
#pragma unroll
      for (uint32_t c = 0; c < kRrChains; ++c) a[c] = 0.f;
#pragma unroll
      for (uint32_t d4 = 0; d4 < kDim / 4; ++d4) {
#pragma unroll
        for (uint32_t c = 0; c < kRrChains; ++c) {
          const float4 xv = xs[wave][2 * c + h][d4];
          a[c] = __builtin_fmaf(qv[4 * d4 + 0], xv.x, a[c]);
          a[c] = __builtin_fmaf(qv[4 * d4 + 1], xv.y, a[c]);
          a[c] = __builtin_fmaf(qv[4 * d4 + 2], xv.z, a[c]);
          a[c] = __builtin_fmaf(qv[4 * d4 + 3], xv.w, a[c]);
        }
      }
#pragma unroll
      for (uint32_t c = 0; c < kRrChains; ++c)
        if (i + 2 * c + h < L) M = fmaxf(M, a[c]);
      wave_lds_sync();
    }
    M = fmaxf(M, __shfl_xor(M, 32));
    if (h == 0) Ms[wave][t] = M;
    wave_lds_sync();
    if (lane == 0) {
      float sc = 0.0f;
      for (uint32_t tt = 0; tt < T; ++tt) sc += Ms[wave][tt];
      sel_score[(size_t)q * R + r] = sc;
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Fused decompress + MaxSim on MFMA (BF16 inputs, FP32 accumulation).
// x_i = C[code_i] + r_i is never formed: S = C_rows.Q^T + R.Q^T with both
// products on v_mfma_f32_32x32x16_bf16 into one accumulator.  A operand =
// 32 doc tokens (rows), B operand = the 32 query tokens (columns, fixed per
// wave).  Lane l = (row/col c = l & 31, half h = l >> 5): it loads its token's
// code word, 16 residual bytes and 128 bytes of the bf16 centroid row (dims
// [64h, 64h+64)), so the k index is permuted identically in A and B: step s,
// half h covers dims 64h + 8s .. +7.  The residual operand comes from a
// per-byte-position table in LDS (32 positions x 256 byte values x 4 bf16 =
// 64 KB): one 8-byte LDS read decodes four dims.  The accumulator leaves each
// lane 16 token rows of one query-token column, so the max over tokens is
// local plus one cross-half shuffle.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8_t as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8_t, v); }

// ---------------------------------------------------------------------------
// BF16 MaxSim core on v_mfma_f32_16x16x32_bf16, shared by the per-doc and the
// persistent kernels (so their scores are bit-identical).  A tile is 16 doc
// tokens (rows) x 128 dims against the 32 query tokens in two 16-column
// halves qh.  Lane l = (row r = l & 15, group g = l >> 4): k-step s (s = 0..3)
// of row r covers dims 32g + 8s .. +7.  The centroid rows are read from the
// MaxSim-order copy C_bf16m, whose piece 4s + g holds exactly those dims, so
// the four lanes of a row read 64 contiguous bytes per load, and group g's
// residual bits are bytes 8g .. 8g + 7 of the token's row: one 8-byte load per
// lane, no byte loaded twice.  Each k-step is two MFMAs per query half, the
// centroid part and then the residual part, into one accumulator:
// acc[qh][j] = row 4g + j, query token 16 qh + (l & 15).
//
// Residual operand: the nibble-pair table tab[pair_slot(p) + v] holds, as two
// bf16, W[2p][v >> 2] and W[2p + 1][v & 3] (pair p = dims 2p, 2p + 1; the
// nibble's high bits select dim 2p).  The four pairs 4k .. 4k+3 form a block
// of 64 words at 64 k + 16 ((k >> 2) & 3).  A lookup instruction reads pair
// 16g + 4s + i in group g, i.e. block 4g + s, and the blocks of the four
// groups then start 272 words apart (16 g mod 64), in disjoint 16-bank sets of
// a 64-bank LDS (and of each 32-lane half of a 32-bank one): no conflicts.
// ---------------------------------------------------------------------------
constexpr uint32_t kPairTab = 64 * 16 + 48;  // words of the padded nibble-pair table

__device__ __forceinline__ uint32_t pair_slot(uint32_t p) {
  return 64 * (p >> 2) + 16 * ((p >> 4) & 3u) + 16 * (p & 3u);
}

// fills tab (kPairTab words of LDS) from the [128][4] weights W; nt threads
__device__ __forceinline__ void fill_pair_tab(uint32_t *tab, const float *__restrict__ W, uint32_t tid,
                                              uint32_t nt) {
  for (uint32_t i = tid; i < kDim / 2 * 16; i += nt) {
    const uint32_t pp = i >> 4, v = i & 15u;
    const uint32_t lo = __builtin_bit_cast(uint16_t, (__bf16)W[(2 * pp) * 4 + (v >> 2)]);
    const uint32_t hi = __builtin_bit_cast(uint16_t, (__bf16)W[(2 * pp + 1) * 4 + (v & 3u)]);
    tab[pair_slot(pp) + v] = lo | (hi << 16);
  }
}

// The row's operands: pieces cv[s] = piece 4s + g of its MaxSim-order
// centroid row Cm and residual bytes 8g .. 8g + 7 in rb.
__device__ __forceinline__ void load_row16_bf16(const uint16_t *__restrict__ Cm, const uint8_t *__restrict__ residual,
                                                uint64_t tok, uint32_t code, uint32_t g, uint4 (&cv)[4], uint2 &rb) {
  const uint4 *cp = reinterpret_cast<const uint4 *>(Cm + (size_t)code * kDim) + g;
#pragma unroll
  for (uint32_t s = 0; s < 4; ++s) cv[s] = cp[4 * s];
  rb = *reinterpret_cast<const uint2 *>(residual + tok * kResBytes + 8 * g);
}

// Residual operand of k-step s: dims 32g + 8s .. +7 from bytes 8g + 2s, +1.
__device__ __forceinline__ uint4 res16_bf16(const uint32_t *tab, const uint2 &rb, uint32_t s, uint32_t g) {
  const uint32_t w = ((s >> 1) ? rb.y : rb.x) >> (16 * (s & 1));
  const uint32_t b0 = w & 0xFFu, b1 = (w >> 8) & 0xFFu;
  const uint32_t *t = tab + 272 * g + 64 * s;  // pair_slot(16g + 4s)
  return make_uint4(t[b0 >> 4], t[16 + (b0 & 15u)], t[32 + (b1 >> 4)], t[48 + (b1 & 15u)]);
}

// The residual operands of one tile, all four k-steps.
__device__ __forceinline__ void dec_res16_bf16(const uint32_t *tab, const uint2 &rb, uint32_t g, uint4 (&rv)[4]) {
#pragma unroll
  for (uint32_t s = 0; s < 4; ++s) rv[s] = res16_bf16(tab, rb, s, g);
}

// The 16 x 32 scores of one tile from its decoded operands (see above).
__device__ __forceinline__ void mma16_bf16(const uint4 (&cv)[4], const uint4 (&rv)[4], const bf16x8_t (&qv)[2][4],
                                           floatx4_t (&acc)[2]) {
  acc[0] = floatx4_t{};
  acc[1] = floatx4_t{};
#pragma unroll
  for (uint32_t s = 0; s < 4; ++s) {
#pragma unroll
    for (uint32_t qh = 0; qh < 2; ++qh) {
      acc[qh] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(cv[s]), qv[qh][s], acc[qh], 0, 0, 0);
      acc[qh] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(rv[s]), qv[qh][s], acc[qh], 0, 0, 0);
    }
  }
}

// The 16 x 32 scores of one tile (see above).
__device__ __forceinline__ void mma_tile16_bf16(const uint32_t *tab, const uint4 (&cv)[4], const uint2 &rb,
                                                const bf16x8_t (&qv)[2][4], uint32_t g, floatx4_t (&acc)[2]) {
  uint4 rv[4];
  dec_res16_bf16(tab, rb, g, rv);
  mma16_bf16(cv, rv, qv, acc);
}

// Query tokens as the 16x16x32 B operand: lane (col = l & 15, group g) holds
// qv[qh][s] = Q[16 qh + col][32g + 8s .. +7] in bf16 (zero past T tokens).
__device__ __forceinline__ void load_q16_bf16(const float *__restrict__ Q, uint32_t q, uint32_t t_max, uint32_t T,
                                              uint32_t lane, bf16x8_t (&qv)[2][4]) {
  const uint32_t col = lane & 15u, g = lane >> 4;
#pragma unroll
  for (uint32_t qh = 0; qh < 2; ++qh) {
    const uint32_t t = 16 * qh + col;
    const float4 *qr = reinterpret_cast<const float4 *>(Q + ((size_t)q * t_max + (t < T ? t : 0)) * kDim) + 8 * g;
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s) {
      const float4 a = qr[2 * s], b = qr[2 * s + 1];
      const float x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
      bf16x8_t v;
#pragma unroll
      for (uint32_t e = 0; e < 8; ++e) v[e] = (__bf16)(t < T ? x[e] : 0.0f);
      qv[qh][s] = v;
    }
  }
}

// Running max m of the values x, as a compare and a select (fmaxf quiets
// signalling-NaN inputs first: two more VALU instructions per max); a NaN x
// leaves m as it is.
__device__ __forceinline__ float vmax(float m, float x) { return x > m ? x : m; }

// Running per-query-token maxima over one 16-row tile's scores (lane: rows
// 4g .. 4g+3 of query tokens col, 16 + col); only rows below `valid` (rows
// of the tile that hold doc tokens) count, and a full tile skips the checks.
__device__ __forceinline__ void tile_max16(const floatx4_t (&acc)[2], uint32_t g, int valid, float &M0,
                                           float &M1) {
  if (valid >= 16) {
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
      M0 = vmax(M0, acc[0][e]);
      M1 = vmax(M1, acc[1][e]);
    }
  } else {
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
      if ((int)(4 * g + e) < valid) {
        M0 = vmax(M0, acc[0][e]);
        M1 = vmax(M1, acc[1][e]);
      }
    }
  }
}

// lane l < 32 of the result: the max of query token l over the 16 row groups
__device__ __forceinline__ float tile_max_reduce(float M0, float M1, uint32_t lane) {
  M0 = vmax(M0, __shfl_xor(M0, 16));
  M1 = vmax(M1, __shfl_xor(M1, 16));
  M0 = vmax(M0, __shfl_xor(M0, 32));
  M1 = vmax(M1, __shfl_xor(M1, 32));
  return (lane & 16u) ? M1 : M0;
}

// Per-query-token max over one doc's tokens; returns in lane l < 32 the max
// for query token l.  TPS tiles of 16 tokens in flight per step; rows past
// the doc's end load its last token and are left out of the max.
template <int TPS>
__device__ __forceinline__ float maxsim16_doc_bf16(uint32_t L, uint64_t off, uint32_t lane,
                                                   const uint32_t *__restrict__ codes,
                                                   const uint8_t *__restrict__ residual,
                                                   const uint16_t *__restrict__ Cm, const uint32_t *tab,
                                                   const bf16x8_t (&qv)[2][4], int clip_zero) {
  const uint32_t r = lane & 15u, g = lane >> 4;
  float M0 = clip_zero ? 0.0f : -FLT_MAX, M1 = M0;
  for (uint32_t i0 = 0; i0 < L; i0 += 16 * TPS) {
    uint32_t cd[TPS];
#pragma unroll
    for (int u = 0; u < TPS; ++u) cd[u] = codes[off + min(i0 + 16 * u + r, L - 1)];
    uint4 cv[TPS][4];
    uint2 rb[TPS];
#pragma unroll
    for (int u = 0; u < TPS; ++u) {
      const uint64_t tok = off + min(i0 + 16 * u + r, L - 1);
      load_row16_bf16(Cm, residual, tok, cd[u], g, cv[u], rb[u]);
    }
#pragma unroll
    for (int u = 0; u < TPS; ++u) {
      floatx4_t acc[2];
      mma_tile16_bf16(tab, cv[u], rb[u], qv, g, acc);
      tile_max16(acc, g, (int)L - (int)(i0 + 16 * u), M0, M1);
    }
  }
  return tile_max_reduce(M0, M1, lane);
}

__global__ __launch_bounds__(256) void k_rerank_bf16(
    const float *__restrict__ Q, uint32_t t_max, const uint32_t *__restrict__ q_ntok,
    const uint32_t *__restrict__ sel_ids, const uint32_t *__restrict__ sel_n, uint32_t R,
    uint32_t n_slice, const uint32_t *__restrict__ doclens, const uint64_t *__restrict__ docoff,
    const uint32_t *__restrict__ codes, const uint8_t *__restrict__ residual,
    const uint16_t *__restrict__ Cm, const float *__restrict__ W, int clip_zero,
    float *__restrict__ sel_score) {
  __shared__ uint32_t tab[kPairTab];
  __shared__ float Ms[4][kMaxQTok];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, col = lane & 31;
  const uint32_t slice = blockIdx.x, q = blockIdx.y;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  const uint32_t n = sel_n[q];

  fill_pair_tab(tab, W, tid, 256);
  bf16x8_t qv[2][4];
  load_q16_bf16(Q, q, t_max, T, lane, qv);
  __syncthreads();

  for (uint32_t r = slice * 4 + wave; r < n; r += n_slice * 4) {
    const uint32_t doc = sel_ids[(size_t)q * R + r];
    const float M = maxsim16_doc_bf16<1>(doclens[doc], docoff[doc], lane, codes, residual, Cm, tab, qv,
                                         clip_zero);
    if (h == 0) Ms[wave][col] = M;
    wave_lds_sync();
    if (lane == 0) {
      float sc = 0.0f;
      for (uint32_t tt = 0; tt < T; ++tt) sc += Ms[wave][tt];
      sel_score[(size_t)q * R + r] = sc;
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Doc-major BF16 MaxSim.  A pass of nb queries x R candidates holds each doc
// about nb*R/n_doc times (about four at 4,096 queries x 1,000 over 1M docs),
// so the query-major kernel above gathers every candidate doc's centroid rows
// once per query that chose it.  Here the (query, slot) pairs are grouped by
// doc (counting sort: k_dm_count, k_dm_bsum, k_dm_bscan, k_dm_offsets,
// k_dm_scatter), and a wave loads a doc's tiles once -- up to kDmTiles tiles
// of 16 tokens in registers -- then scores every query of the doc against
// them, reading each query's B operand as eight coalesced 1 KB loads from a
// per-pass bf16 copy in MFMA lane order (k_q_bf16m).  The MFMA operands, their
// order, the row mask and the left-to-right sum over query tokens are those
// of maxsim16_doc_bf16, so every score is bit-identical to k_rerank_bf16.
// Docs longer than kDmTiles tiles go chunk by chunk over groups of up to
// kDmGroup of their pairs, the per-token maxima of a group kept in LDS
// between chunks (max is exact, so the chunking changes no bit).
// ---------------------------------------------------------------------------
constexpr uint32_t kDmTiles = 8;     // 16-token doc tiles held in registers (128 tokens)
constexpr uint32_t kDmGroup = 16;    // pairs per group of a multi-chunk doc
constexpr uint32_t kDmGrab = 16;     // work items a wave takes per queue step
constexpr uint32_t kDmHoist = 5;     // chunks of up to this many tiles decode residuals once per chunk
constexpr uint32_t kDmBlockDocs = kDmDocsPerBlock;  // docs per block of the offset scan (256 threads x 16)

// Qm[q][qh][s][lane] = the B operand qv[qh][s] of load_q16_bf16 for lane.
__global__ __launch_bounds__(256) void k_q_bf16m(const float *__restrict__ Q, uint32_t t_max,
                                                 const uint32_t *__restrict__ q_ntok, uint32_t nb,
                                                 uint4 *__restrict__ Qm) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;  // (q, qh, s, lane)
  if (i >= nb * 512u) return;
  const uint32_t q = i >> 9, qh = (i >> 8) & 1u, s = (i >> 6) & 3u, lane = i & 63u;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  const uint32_t t = 16 * qh + (lane & 15u), g = lane >> 4;
  const float4 *qr = reinterpret_cast<const float4 *>(Q + ((size_t)q * t_max + (t < T ? t : 0)) * kDim) + 8 * g;
  const float4 a = qr[2 * s], b = qr[2 * s + 1];
  const float x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  bf16x8_t v;
#pragma unroll
  for (uint32_t e = 0; e < 8; ++e) v[e] = (__bf16)(t < T ? x[e] : 0.0f);
  Qm[i] = __builtin_bit_cast(uint4, v);
}

// cnt[doc] += 1 for every selected (query, slot)
__global__ __launch_bounds__(256) void k_dm_count(const uint32_t *__restrict__ sel_ids,
                                                  const uint32_t *__restrict__ sel_n, uint32_t R,
                                                  uint32_t *__restrict__ cnt) {
  const uint32_t q = blockIdx.y, n = min(sel_n[q], R);
  for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256)
    atomicAdd(&cnt[sel_ids[(size_t)q * R + j]], 1u);
}

// exclusive scan over a 256-thread block of (a, b); returns the block totals
__device__ __forceinline__ void dm_scan256(uint32_t a, uint32_t b, uint32_t &ex_a, uint32_t &ex_b,
                                           uint32_t &tot_a, uint32_t &tot_b) {
  __shared__ uint32_t sa[4], sb[4];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t ia = a, ib = b;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t ua = __shfl_up(ia, o), ub = __shfl_up(ib, o);
    if (lane >= o) {
      ia += ua;
      ib += ub;
    }
  }
  if (lane == 63) {
    sa[wv] = ia;
    sb[wv] = ib;
  }
  __syncthreads();
  uint32_t ba = 0, bb = 0;
  tot_a = tot_b = 0;
  for (uint32_t i = 0; i < 4; ++i) {
    if (i < wv) {
      ba += sa[i];
      bb += sb[i];
    }
    tot_a += sa[i];
    tot_b += sb[i];
  }
  ex_a = ba + ia - a;
  ex_b = bb + ib - b;
}

// per block of kDmBlockDocs docs: pairs and docs with pairs
__global__ __launch_bounds__(256) void k_dm_bsum(const uint32_t *__restrict__ cnt, uint32_t n_doc,
                                                 uint32_t *__restrict__ b_p, uint32_t *__restrict__ b_d) {
  const size_t d0 = (size_t)blockIdx.x * kDmBlockDocs + threadIdx.x * 16;
  uint32_t np = 0, nd = 0;
  for (uint32_t e = 0; e < 16; ++e) {
    const uint32_t c = d0 + e < n_doc ? cnt[d0 + e] : 0u;
    np += c;
    nd += c != 0;
  }
  uint32_t ea, eb, ta, tb;
  dm_scan256(np, nd, ea, eb, ta, tb);
  if (threadIdx.x == 0) {
    b_p[blockIdx.x] = ta;
    b_d[blockIdx.x] = tb;
  }
}

// in-place exclusive scan of the block sums; tot = (pairs, work items);
// resets the work-queue counter
__global__ __launch_bounds__(1024) void k_dm_bscan(uint32_t *__restrict__ b_p, uint32_t *__restrict__ b_d,
                                                   uint32_t nblk, uint32_t *__restrict__ tot) {
  __shared__ uint32_t sp[16], sd[16];
  const uint32_t per = (nblk + 1023) / 1024;
  const uint32_t i0 = min(nblk, threadIdx.x * per), i1 = min(nblk, i0 + per);
  uint32_t np = 0, nd = 0;
  for (uint32_t i = i0; i < i1; ++i) {
    np += b_p[i];
    nd += b_d[i];
  }
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t ip = np, id = nd;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t up = __shfl_up(ip, o), ud = __shfl_up(id, o);
    if (lane >= o) {
      ip += up;
      id += ud;
    }
  }
  if (lane == 63) {
    sp[wv] = ip;
    sd[wv] = id;
  }
  __syncthreads();
  uint32_t bp = 0, bd = 0, tp = 0, td = 0;
  for (uint32_t i = 0; i < 16; ++i) {
    if (i < wv) {
      bp += sp[i];
      bd += sd[i];
    }
    tp += sp[i];
    td += sd[i];
  }
  uint32_t base_p = bp + ip - np, base_d = bd + id - nd;
  for (uint32_t i = i0; i < i1; ++i) {
    const uint32_t a = b_p[i], b = b_d[i];
    b_p[i] = base_p;
    b_d[i] = base_d;
    base_p += a;
    base_d += b;
  }
  if (threadIdx.x == 0) {
    tot[0] = tp;
    tot[1] = td;
    tot[2] = 0;  // work-queue counter
  }
}

// pair offsets: work[i] = (doc, first pair) for the i-th doc with pairs, and
// cnt[doc] = its first pair (the scatter's cursor)
__global__ __launch_bounds__(256) void k_dm_offsets(uint32_t *__restrict__ cnt, uint32_t n_doc,
                                                    const uint32_t *__restrict__ b_p,
                                                    const uint32_t *__restrict__ b_d, uint2 *__restrict__ work) {
  const size_t d0 = (size_t)blockIdx.x * kDmBlockDocs + threadIdx.x * 16;
  uint32_t c[16], np = 0, nd = 0;
#pragma unroll
  for (uint32_t e = 0; e < 16; ++e) {
    c[e] = d0 + e < n_doc ? cnt[d0 + e] : 0u;
    np += c[e];
    nd += c[e] != 0;
  }
  uint32_t ea, eb, ta, tb;
  dm_scan256(np, nd, ea, eb, ta, tb);
  uint32_t p = b_p[blockIdx.x] + ea, w = b_d[blockIdx.x] + eb;
#pragma unroll
  for (uint32_t e = 0; e < 16; ++e) {
    if (c[e]) {
      work[w++] = make_uint2((uint32_t)(d0 + e), p);
      cnt[d0 + e] = p;
      p += c[e];
    }
  }
}

// plist[cursor(doc)++] = q * R + slot; afterwards cnt[doc] = the doc's end
__global__ __launch_bounds__(256) void k_dm_scatter(const uint32_t *__restrict__ sel_ids,
                                                    const uint32_t *__restrict__ sel_n, uint32_t R,
                                                    uint32_t *__restrict__ cnt, uint32_t *__restrict__ plist) {
  const uint32_t q = blockIdx.y, n = min(sel_n[q], R);
  for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
    const uint32_t pos = atomicAdd(&cnt[sel_ids[(size_t)q * R + j]], 1u);
    plist[pos] = q * R + j;
  }
}

__device__ __forceinline__ void load_qm(const uint4 *__restrict__ Qm, uint32_t q, uint32_t lane,
                                        bf16x8_t (&qv)[2][4]) {
  const uint4 *p = Qm + (size_t)q * 512 + lane;
#pragma unroll
  for (uint32_t qh = 0; qh < 2; ++qh)
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s) qv[qh][s] = as_bf16x8(p[64 * (4 * qh + s)]);
}

// One chunk of NT doc tiles (rows t0*16 ..) against the ng pairs pg .. of the
// doc: per-token maxima into mg[j] (combined with earlier chunks unless t0 = 0).
template <uint32_t NT>
__device__ __forceinline__ void dm_chunk(const uint4 *__restrict__ Qm, uint32_t R, const uint32_t *__restrict__ plist,
                                         uint32_t pg, uint32_t ng, uint32_t L, uint64_t off, uint32_t t0,
                                         const uint32_t *__restrict__ codes, const uint8_t *__restrict__ residual,
                                         const uint16_t *__restrict__ Cm, const uint32_t *tab, float init,
                                         uint32_t lane, float (*mg)[kMaxQTok + 1]) {
  const uint32_t r = lane & 15u, g = lane >> 4;
  // every tile's code first, then the rows: each row load depends on its
  // code, and waiting for one load waits for all issued before it
  uint32_t cd[NT];
#pragma unroll
  for (uint32_t u = 0; u < NT; ++u) cd[u] = codes[off + min(16 * (t0 + u) + r, L - 1)];
  uint4 cv[NT][4];
  uint2 rb[NT];
#pragma unroll
  for (uint32_t u = 0; u < NT; ++u)
    load_row16_bf16(Cm, residual, off + min(16 * (t0 + u) + r, L - 1), cd[u], g, cv[u], rb[u]);
  // chunks of up to kDmHoist tiles decode their residual operands once for
  // all pairs; longer ones decode per pair (registers)
  constexpr bool kHoist = NT <= kDmHoist;
  uint4 rh[kHoist ? NT : 1][4];
  if (kHoist) {
#pragma unroll
    for (uint32_t u = 0; u < NT; ++u) dec_res16_bf16(tab, rb[u], g, rh[kHoist ? u : 0]);
  }
  // rows of the chunk's last tile that hold doc tokens (the others are full)
  const int last_valid = (int)L - (int)(16 * (t0 + NT - 1));
  for (uint32_t j = 0; j < ng; ++j) {
    bf16x8_t qv[2][4];
    load_qm(Qm, plist[pg + j] / R, lane, qv);
    // all eight loads issued before the first MFMA waits on one of them
    __builtin_amdgcn_sched_barrier(0);
    // per-pair decode one tile ahead; the empty asm keeps the compiler from
    // hoisting it out of the pair loop (16 more live registers per tile)
    uint4 rv[2][4];
    if (!kHoist) {
      uint2 b = rb[0];
      asm volatile("" : "+v"(b.x), "+v"(b.y));
      dec_res16_bf16(tab, b, g, rv[0]);
    }
    float M0 = init, M1 = init;
#pragma unroll
    for (uint32_t u = 0; u < NT; ++u) {
      if (!kHoist && u + 1 < NT) {
        uint2 b = rb[u + 1];
        asm volatile("" : "+v"(b.x), "+v"(b.y));
        dec_res16_bf16(tab, b, g, rv[(u + 1) & 1]);
      }
      floatx4_t acc[2];
      mma16_bf16(cv[u], kHoist ? rh[kHoist ? u : 0] : rv[u & 1], qv, acc);
      tile_max16(acc, g, u + 1 < NT ? 16 : last_valid, M0, M1);
    }
    const float M = tile_max_reduce(M0, M1, lane);
    if (lane < 32) {  // lane l: query token l; each lane rereads only its own entry
      mg[j][lane] = t0 != 0 ? vmax(M, mg[j][lane]) : M;
    }
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_rerank_bf16_dm(
    const uint4 *__restrict__ Qm, uint32_t t_max, const uint32_t *__restrict__ q_ntok, uint32_t R,
    const uint2 *__restrict__ work, uint32_t *__restrict__ tot, const uint32_t *__restrict__ pend,
    const uint32_t *__restrict__ plist, const uint32_t *__restrict__ doclens,
    const uint64_t *__restrict__ docoff, const uint32_t *__restrict__ codes,
    const uint8_t *__restrict__ residual, const uint16_t *__restrict__ Cm, const float *__restrict__ W,
    int clip_zero, float *__restrict__ sel_score) {
  static_assert(kDmTiles == 8, "dm_chunk is dispatched for 1 .. 8 tiles");
  __shared__ uint32_t tab[kPairTab];
  // per-token maxima of a wave's pair group; rows padded so that the lanes
  // summing one row each read distinct banks
  __shared__ float Mg[4][kDmGroup][kMaxQTok + 1];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  fill_pair_tab(tab, W, tid, 256);
  __syncthreads();
  float(*mg)[kMaxQTok + 1] = Mg[wave];
  const uint32_t n_work = tot[1];
  const float init = clip_zero ? 0.0f : -FLT_MAX;
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&tot[2], kDmGrab);
    base = __builtin_amdgcn_readfirstlane(base);
    if (base >= n_work) break;
    const uint32_t w_end = min(base + kDmGrab, n_work);
    for (uint32_t wi = base; wi < w_end; ++wi) {
      const uint2 it = work[wi];
      const uint32_t doc = it.x, p0 = it.y, p1 = pend[doc];
      const uint32_t L = doclens[doc];
      const uint64_t off = docoff[doc];
      const uint32_t n_t = (L + 15) >> 4;
      for (uint32_t pg = p0; pg < p1; pg += kDmGroup) {
        const uint32_t ng = min(kDmGroup, p1 - pg);
        if (n_t == 0 && lane < 32)  // an empty doc: the identity for every token
          for (uint32_t j = 0; j < ng; ++j) mg[j][lane] = init;
        for (uint32_t t0 = 0; t0 < n_t; t0 += kDmTiles) {
#define MVR_DM_CHUNK(N) dm_chunk<N>(Qm, R, plist, pg, ng, L, off, t0, codes, residual, Cm, tab, init, lane, mg)
          switch (min(kDmTiles, n_t - t0)) {
            case 1: MVR_DM_CHUNK(1); break;
            case 2: MVR_DM_CHUNK(2); break;
            case 3: MVR_DM_CHUNK(3); break;
            case 4: MVR_DM_CHUNK(4); break;
            case 5: MVR_DM_CHUNK(5); break;
            case 6: MVR_DM_CHUNK(6); break;
            case 7: MVR_DM_CHUNK(7); break;
            default: MVR_DM_CHUNK(8); break;
          }
#undef MVR_DM_CHUNK
        }
        // scores of the group: lane j sums pair j's maxima left to right
        wave_lds_sync();
        if (lane < ng) {
          const uint32_t pq = plist[pg + lane], q = pq / R;
          const uint32_t T = q_ntok ? q_ntok[q] : t_max;
          float sc = 0.0f;
          for (uint32_t tt = 0; tt < T; ++tt) sc += mg[lane][tt];
          sel_score[pq] = sc;
        }
        wave_lds_sync();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Fused decompress + MaxSim on FP32 MFMA (v_mfma_f32_32x32x2_f32).  Same tile
// scheme as the BF16 kernel: A = 32 doc tokens (rows), B = the 32 query tokens
// (columns).  Lane l = (row/col c = l & 31, half h = l >> 5) decodes its token's dims
// [64h, 64h + 64) in registers -- x = C[code] + w[d][bucket], the exact FP32
// values of the decompressed embedding -- from one 16-byte residual load and
// 16 float4 loads of the centroid row; instruction s, half h covers dim
// 64h + s in both operands.  Dot products are summed in the MFMA's order, so
// scores agree with the sequential-FMA kernel to about 1e-7 relative (not
// bit-identical); the max over tokens is local plus one cross-half shuffle
// and the sum over query tokens runs left to right.
// ---------------------------------------------------------------------------
// FP32 MFMA MaxSim kernel.  The query tile sits in LDS instead of 64
// registers per lane: ql[col * kQlStride + d] = q[col][d] (zero for col >= T),
// read as float4 per four MFMAs (rows 528 B apart, so the 16 lanes of a
// 128-bit read phase hit 64 distinct banks), which leaves room for 4 waves per
// SIMD.  Each wave streams its docs (slots slice*4 + wave, + 4*n_slice, ...)
// as chunks of 8 tokens, four chunks to a 32-row tile, so a tile can span doc
// boundaries and only the last chunk of a doc is padded (95% of the MFMA rows
// carry tokens on a lognormal length mix around 67 tokens, against 81% with a
// tile per 32 tokens of one doc).  Row m of a tile is token 8*(m>>3) + (m&7)'s
// slot: lane (col, h) holds rows 8g + 4h + e (g = j>>2, e = j&3) of acc, so
// the max over a chunk is over j in [4g, 4g+4) plus one cross-half shuffle.
// The persistent kernel's FP32 phase runs the same loop (f32mfma_wave_docs,
// kept as a copy: written inline here the launch kernel stays within 128
// registers, 4 waves per SIMD), so their scores are bit-identical.
constexpr uint32_t kQlStride = kDim + 4;

// Decode-pair table of the FP32 MFMA MaxSim: tab2[p * 16 + v] = (W[2p][v >> 2],
// W[2p + 1][v & 3]), the two weights that a 4-bit nibble of residual bits
// selects for dims 2p, 2p + 1 (256 threads).
__device__ __forceinline__ void f32mfma_fill_tab(float2 *tab2, const float *__restrict__ W, uint32_t tid) {
  for (uint32_t i = tid; i < kDim / 2 * 16; i += 256) {
    const uint32_t pp = i >> 4, v = i & 15u;
    tab2[i] = make_float2(W[(2 * pp) * 4 + (v >> 2)], W[(2 * pp + 1) * 4 + (v & 3u)]);
  }
}

// The query tile: ql[c * kQlStride + d] = Q[q][c][d], zero for c >= T (256 threads).
__device__ __forceinline__ void f32mfma_load_query(float *ql, const float *__restrict__ Q, uint32_t q,
                                                   uint32_t t_max, uint32_t T, uint32_t tid) {
  for (uint32_t i = tid; i < kMaxQTok * kDim / 4; i += 256) {
    const uint32_t c = i / (kDim / 4), d4 = i % (kDim / 4);
    const float4 v = c < T ? reinterpret_cast<const float4 *>(Q + ((size_t)q * t_max + c) * kDim)[d4]
                           : make_float4(0.f, 0.f, 0.f, 0.f);
    *reinterpret_cast<float4 *>(ql + c * kQlStride + 4 * d4) = v;
  }
}

// One wave's share of a query's candidates: slots r = r_first, r_first +
// stride, ... < n of ids, scores to out[r].  tab2 and ql as filled above; Ms
// is this wave's [4][kMaxQTok + 1] LDS scratch.
__device__ __forceinline__ void f32mfma_wave_docs(const float2 *tab2, const float *ql, float (*Ms)[kMaxQTok + 1],
                                                  uint32_t lane, uint32_t T, const uint32_t *__restrict__ ids,
                                                  uint32_t n, uint32_t r_first, uint32_t stride,
                                                  const uint32_t *__restrict__ doclens,
                                                  const uint64_t *__restrict__ docoff,
                                                  const uint32_t *__restrict__ codes,
                                                  const uint8_t *__restrict__ residual,
                                                  const float *__restrict__ C, int clip_zero,
                                                  float *__restrict__ out) {
  const uint32_t h = lane >> 5, col = lane & 31;
  const float init = clip_zero ? 0.0f : -FLT_MAX;

  // the open doc: slot r, length L, first token off, next chunk at pos (an
  // empty doc is one chunk without tokens, so its score is the sum of the
  // initial maxima)
  uint32_t r = r_first, L = 0, pos = 0;
  uint64_t off = 0;
  auto open = [&]() {
    pos = 0;
    if (r < n) {
      const uint32_t doc = ids[r];
      L = doclens[doc];
      off = docoff[doc];
    }
  };
  open();
  float Mrun = init;
  const uint32_t qoff0 = col * kQlStride + 64 * h;
  while (r < n) {
    // the tile's four chunks (wave-uniform): slot, first token, length, last chunk of its doc
    uint32_t cr[4], cp[4], cl[4];
    uint64_t co[4];
    bool cv_[4], last[4];
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) {
      cv_[g] = r < n;
      cr[g] = r;
      cp[g] = pos;
      cl[g] = L;
      co[g] = off;
      last[g] = cv_[g] && pos + 8 >= L;
      if (cv_[g]) {
        pos += 8;
        if (last[g]) {
          r += stride;
          open();
        }
      }
    }
    // this lane's row: chunk col >> 3, token (col & 7) of it
    const uint32_t g = col >> 3;
    const uint32_t tok = (g == 0 ? cp[0] : g == 1 ? cp[1] : g == 2 ? cp[2] : cp[3]) + (col & 7);
    const uint32_t len = g == 0 ? cl[0] : g == 1 ? cl[1] : g == 2 ? cl[2] : cl[3];
    const uint64_t base = g == 0 ? co[0] : g == 1 ? co[1] : g == 2 ? co[2] : co[3];
    const bool ok = (g == 0 ? cv_[0] : g == 1 ? cv_[1] : g == 2 ? cv_[2] : cv_[3]) && tok < len;
    // rows without a token decode centroid 0 with zero residual bits (any
    // finite values: the max skips those rows), so no register is zeroed
    uint32_t code = 0;
    uint4 rb = make_uint4(0u, 0u, 0u, 0u);
    if (ok) {
      const uint64_t gt = base + tok;
      code = codes[gt];
      rb = *reinterpret_cast<const uint4 *>(residual + gt * kResBytes + 16 * h);
    }
    float4 cv[16];
    const float4 *cptr = reinterpret_cast<const float4 *>(C + (size_t)code * kDim + 64 * h);
#pragma unroll
    for (uint32_t s2 = 0; s2 < 16; ++s2) cv[s2] = cptr[s2];
    // keep the query reads inside the loop (not hoisted into 64 registers)
    uint32_t qoff = qoff0;
    asm volatile("" : "+v"(qoff));
    const float4 *qr = reinterpret_cast<const float4 *>(ql + qoff);
    // decode the whole tile first (the 32 data-dependent weight-pair reads
    // in flight together, added into the centroid values in place), then the
    // MFMA chain with no LDS wait between instructions but the query reads.
    // Residual byte 16h + s2 holds dims 64h + 4 s2 .. +3, most significant
    // bits first: its high nibble selects pair 32h + 2 s2, the low one the next.
    const uint32_t rw[4] = {rb.x, rb.y, rb.z, rb.w};
    const float2 *tp = tab2 + 32 * 16 * h;
#pragma unroll
    for (uint32_t s2 = 0; s2 < 16; ++s2) {
      const uint32_t sh = 8 * (s2 & 3);
      const float2 a2 = tp[(2 * s2) * 16 + ((rw[s2 >> 2] >> (sh + 4)) & 15u)];
      const float2 b2 = tp[(2 * s2 + 1) * 16 + ((rw[s2 >> 2] >> sh) & 15u)];
      cv[s2].x += a2.x;
      cv[s2].y += a2.y;
      cv[s2].z += b2.x;
      cv[s2].w += b2.y;
    }
    __builtin_amdgcn_sched_barrier(0);
    floatx16_t acc = floatx16_t{};
#pragma unroll
    for (uint32_t s2 = 0; s2 < 16; ++s2) {
      const float4 qv = qr[s2];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].x, qv.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].y, qv.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].z, qv.z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].w, qv.w, acc, 0, 0, 0);
    }
    // per chunk: max over its valid rows, folded into the open doc's max;
    // a doc's last chunk parks its per-token maxima (halves combined) in LDS
    uint32_t nfin = 0, fin_r[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (uint32_t gg = 0; gg < 4; ++gg) {
      if (cv_[gg]) {
        if (cp[gg] == 0) Mrun = init;
        if (cp[gg] + 8 <= cl[gg]) {
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e) Mrun = fmaxf(Mrun, acc[4 * gg + e]);
        } else {
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e)
            if (cp[gg] + 4 * h + e < cl[gg]) Mrun = fmaxf(Mrun, acc[4 * gg + e]);
        }
        if (last[gg]) {
          const float M2 = fmaxf(Mrun, __shfl_xor(Mrun, 32));
          if (h == 0) Ms[nfin][col] = M2;
          fin_r[nfin] = cr[gg];
          ++nfin;
        }
      }
    }
    // scores of the docs finished in this tile, one lane each: the
    // left-to-right sum over query tokens
    if (nfin) {
      wave_lds_sync();
      if (lane < nfin) {
        const uint32_t rr = lane == 0 ? fin_r[0] : lane == 1 ? fin_r[1] : lane == 2 ? fin_r[2] : fin_r[3];
        float sc = 0.0f;
        for (uint32_t tt = 0; tt < T; ++tt) sc += Ms[lane][tt];
        out[rr] = sc;
      }
      wave_lds_sync();
    }
  }
}


__global__ __launch_bounds__(256) void k_rerank_f32mfma(
    const float *__restrict__ Q, uint32_t t_max, const uint32_t *__restrict__ q_ntok,
    const uint32_t *__restrict__ sel_ids, const uint32_t *__restrict__ sel_n, uint32_t R,
    uint32_t n_slice, const uint32_t *__restrict__ doclens, const uint64_t *__restrict__ docoff,
    const uint32_t *__restrict__ codes, const uint8_t *__restrict__ residual,
    const float *__restrict__ C, const float *__restrict__ W, int clip_zero,
    float *__restrict__ sel_score) {
  // decode pairs: tab2[p * 16 + v] = (W[2p][v >> 2], W[2p + 1][v & 3]), the
  // two weights that a 4-bit nibble of residual bits selects for dims 2p, 2p+1
  __shared__ float2 tab2[kDim / 2 * 16];
  __shared__ float Ms[4][4][kMaxQTok + 1];  // [wave][doc finished in the tile][query token]
  __shared__ __align__(16) float ql[kMaxQTok * kQlStride];
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t h = lane >> 5, col = lane & 31;
  const uint32_t slice = blockIdx.x, q = blockIdx.y;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  const uint32_t n = sel_n[q];

  for (uint32_t i = tid; i < kDim / 2 * 16; i += 256) {
    const uint32_t pp = i >> 4, v = i & 15u;
    tab2[i] = make_float2(W[(2 * pp) * 4 + (v >> 2)], W[(2 * pp + 1) * 4 + (v & 3u)]);
  }
  for (uint32_t i = tid; i < kMaxQTok * kDim / 4; i += 256) {
    const uint32_t c = i / (kDim / 4), d4 = i % (kDim / 4);
    const float4 v = c < T ? reinterpret_cast<const float4 *>(Q + ((size_t)q * t_max + c) * kDim)[d4]
                           : make_float4(0.f, 0.f, 0.f, 0.f);
    *reinterpret_cast<float4 *>(ql + c * kQlStride + 4 * d4) = v;
  }
  __syncthreads();

  const float init = clip_zero ? 0.0f : -FLT_MAX;
  const uint32_t stride = n_slice * 4;
  const uint32_t *ids = sel_ids + (size_t)q * R;

  // the open doc: slot r, length L, first token off, next chunk at pos (an
  // empty doc is one chunk without tokens, so its score is the sum of the
  // initial maxima)
  uint32_t r = slice * 4 + wave, L = 0, pos = 0;
  uint64_t off = 0;
  auto open = [&]() {
    pos = 0;
    if (r < n) {
      const uint32_t doc = ids[r];
      L = doclens[doc];
      off = docoff[doc];
    }
  };
  open();
  float Mrun = init;
  const uint32_t qoff0 = col * kQlStride + 64 * h;
  while (r < n) {
    // the tile's four chunks (wave-uniform): slot, first token, length, last chunk of its doc
    uint32_t cr[4], cp[4], cl[4];
    uint64_t co[4];
    bool cv_[4], last[4];
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) {
      cv_[g] = r < n;
      cr[g] = r;
      cp[g] = pos;
      cl[g] = L;
      co[g] = off;
      last[g] = cv_[g] && pos + 8 >= L;
      if (cv_[g]) {
        pos += 8;
        if (last[g]) {
          r += stride;
          open();
        }
      }
    }
    // this lane's row: chunk col >> 3, token (col & 7) of it
    const uint32_t g = col >> 3;
    const uint32_t tok = (g == 0 ? cp[0] : g == 1 ? cp[1] : g == 2 ? cp[2] : cp[3]) + (col & 7);
    const uint32_t len = g == 0 ? cl[0] : g == 1 ? cl[1] : g == 2 ? cl[2] : cl[3];
    const uint64_t base = g == 0 ? co[0] : g == 1 ? co[1] : g == 2 ? co[2] : co[3];
    const bool ok = (g == 0 ? cv_[0] : g == 1 ? cv_[1] : g == 2 ? cv_[2] : cv_[3]) && tok < len;
    // rows without a token decode centroid 0 with zero residual bits (any
    // finite values: the max skips those rows), so no register is zeroed
    uint32_t code = 0;
    uint4 rb = make_uint4(0u, 0u, 0u, 0u);
    if (ok) {
      const uint64_t gt = base + tok;
      code = codes[gt];
      rb = *reinterpret_cast<const uint4 *>(residual + gt * kResBytes + 16 * h);
    }
    float4 cv[16];
    const float4 *cptr = reinterpret_cast<const float4 *>(C + (size_t)code * kDim + 64 * h);
#pragma unroll
    for (uint32_t s2 = 0; s2 < 16; ++s2) cv[s2] = cptr[s2];
    // keep the query reads inside the loop (not hoisted into 64 registers)
    uint32_t qoff = qoff0;
    asm volatile("" : "+v"(qoff));
    const float4 *qr = reinterpret_cast<const float4 *>(ql + qoff);
    // decode the whole tile first (the 32 data-dependent weight-pair reads
    // in flight together, added into the centroid values in place), then the
    // MFMA chain with no LDS wait between instructions but the query reads.
    // Residual byte 16h + s2 holds dims 64h + 4 s2 .. +3, most significant
    // bits first: its high nibble selects pair 32h + 2 s2, the low one the next.
    const uint32_t rw[4] = {rb.x, rb.y, rb.z, rb.w};
    const float2 *tp = tab2 + 32 * 16 * h;
#pragma unroll
    for (uint32_t s2 = 0; s2 < 16; ++s2) {
      const uint32_t sh = 8 * (s2 & 3);
      const float2 a2 = tp[(2 * s2) * 16 + ((rw[s2 >> 2] >> (sh + 4)) & 15u)];
      const float2 b2 = tp[(2 * s2 + 1) * 16 + ((rw[s2 >> 2] >> sh) & 15u)];
      cv[s2].x += a2.x;
      cv[s2].y += a2.y;
      cv[s2].z += b2.x;
      cv[s2].w += b2.y;
    }
    __builtin_amdgcn_sched_barrier(0);
    floatx16_t acc = floatx16_t{};
#pragma unroll
    for (uint32_t s2 = 0; s2 < 16; ++s2) {
      const float4 qv = qr[s2];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].x, qv.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].y, qv.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].z, qv.z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[s2].w, qv.w, acc, 0, 0, 0);
    }
    // per chunk: max over its valid rows, folded into the open doc's max;
    // a doc's last chunk parks its per-token maxima (halves combined) in LDS
    uint32_t nfin = 0, fin_r[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (uint32_t gg = 0; gg < 4; ++gg) {
      if (cv_[gg]) {
        if (cp[gg] == 0) Mrun = init;
        if (cp[gg] + 8 <= cl[gg]) {
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e) Mrun = fmaxf(Mrun, acc[4 * gg + e]);
        } else {
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e)
            if (cp[gg] + 4 * h + e < cl[gg]) Mrun = fmaxf(Mrun, acc[4 * gg + e]);
        }
        if (last[gg]) {
          const float M2 = fmaxf(Mrun, __shfl_xor(Mrun, 32));
          if (h == 0) Ms[wave][nfin][col] = M2;
          fin_r[nfin] = cr[gg];
          ++nfin;
        }
      }
    }
    // scores of the docs finished in this tile, one lane each: the
    // left-to-right sum over query tokens
    if (nfin) {
      wave_lds_sync();
      if (lane < nfin) {
        const uint32_t rr = lane == 0 ? fin_r[0] : lane == 1 ? fin_r[1] : lane == 2 ? fin_r[2] : fin_r[3];
        float sc = 0.0f;
        for (uint32_t tt = 0; tt < T; ++tt) sc += Ms[wave][lane][tt];
        sel_score[(size_t)q * R + rr] = sc;
      }
      wave_lds_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// Fused decompress + MaxSim on FP8 MFMA (OCP e4m3 inputs, FP32 accumulation),
// v_mfma_f32_32x32x16_fp8_fp8 in the BF16 kernel's tile scheme.  Scales:
// per centroid row (s_c = max|C_c| / 448, stored with the fp8 rows), per query
// token (s_q = max|q_t| / 448, computed when the query is loaded) and one for
// the residual weights (s_r).  x = C[code] + r is never formed: two
// accumulators hold C8.q8 and R8.q8, combined per output as
// s_q (s_c acc_C + s_r acc_R), which applies each scale to whole rows or
// columns of the products.  A centroid row costs 64 bytes per lane pair
// (128 B per token) instead of 256 B in BF16.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t fp8x2(float a, float b) {
  return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xFFFFu;
}

// Query tokens as the fp8 B operand (lane (col, h): dims 64h + 8s .. +7 in
// qf8[s]) and the token's scale.
__device__ __forceinline__ void load_q_fp8(const float *__restrict__ Q, uint32_t q, uint32_t t_max,
                                           uint32_t T, uint32_t h, uint32_t col, long (&qf8)[8],
                                           float &sq) {
  const float4 *qr = reinterpret_cast<const float4 *>(Q + ((size_t)q * t_max + col) * kDim + 64 * h);
  float v[64];
  float amax = 0.0f;
#pragma unroll
  for (uint32_t i = 0; i < 16; ++i) {
    const float4 x = (col < T) ? qr[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    v[4 * i] = x.x; v[4 * i + 1] = x.y; v[4 * i + 2] = x.z; v[4 * i + 3] = x.w;
    amax = fmaxf(amax, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))));
  }
  amax = fmaxf(amax, __shfl_xor(amax, 32));
  sq = amax / 448.0f;
  const float inv = amax > 0.0f ? 448.0f / amax : 0.0f;
#pragma unroll
  for (uint32_t s = 0; s < 8; ++s) {
    const uint32_t lo = fp8x2(v[8 * s] * inv, v[8 * s + 1] * inv) | (fp8x2(v[8 * s + 2] * inv, v[8 * s + 3] * inv) << 16);
    const uint32_t hi = fp8x2(v[8 * s + 4] * inv, v[8 * s + 5] * inv) | (fp8x2(v[8 * s + 6] * inv, v[8 * s + 7] * inv) << 16);
    qf8[s] = (long)(((uint64_t)hi << 32) | lo);
  }
}

// Per-query-token max over one doc on FP8 tiles; rsc = this wave's LDS row
// scales [TPS][32].
template <int TPS>
__device__ __forceinline__ float maxsim_doc_fp8(uint32_t L, uint64_t off, uint32_t h, uint32_t col,
                                                const uint32_t *__restrict__ codes,
                                                const uint8_t *__restrict__ residual,
                                                const uint8_t *__restrict__ C8,
                                                const float *__restrict__ csc, const uint32_t *lut8,
                                                float sr, const long (&qf8)[8], float sq, float *rsc,
                                                int clip_zero) {
  float M = clip_zero ? 0.0f : -FLT_MAX;
  for (uint32_t i0 = 0; i0 < L; i0 += 32 * TPS) {
    uint4 cv[TPS][4];
    uint4 rb[TPS];
#pragma unroll
    for (int u = 0; u < TPS; ++u) {
      const uint32_t i = i0 + 32 * u + col;
      rb[u] = make_uint4(0u, 0u, 0u, 0u);
      float sc = 0.0f;
      if (i < L) {
        const uint64_t g = off + i;
        const uint32_t code = codes[g];
        rb[u] = *reinterpret_cast<const uint4 *>(residual + g * kResBytes + 16 * h);
        const uint4 *cp = reinterpret_cast<const uint4 *>(C8 + (size_t)code * kDim + 64 * h);
#pragma unroll
        for (uint32_t s = 0; s < 4; ++s) cv[u][s] = cp[s];
        sc = csc[code];
      } else {
#pragma unroll
        for (uint32_t s = 0; s < 4; ++s) cv[u][s] = make_uint4(0u, 0u, 0u, 0u);
      }
      if (h == 0) rsc[32 * u + col] = sc;
    }
    floatx16_t accC[TPS], accR[TPS];
#pragma unroll
    for (int u = 0; u < TPS; ++u) { accC[u] = floatx16_t{}; accR[u] = floatx16_t{}; }
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
#pragma unroll
      for (int u = 0; u < TPS; ++u) {
        const uint4 c4 = cv[u][s >> 1];
        const long a = (long)((s & 1) ? (((uint64_t)c4.w << 32) | c4.z) : (((uint64_t)c4.y << 32) | c4.x));
        accC[u] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a, qf8[s], accC[u], 0, 0, 0);
        const uint32_t rw[4] = {rb[u].x, rb[u].y, rb[u].z, rb[u].w};
        const uint32_t word = rw[s >> 1];
        const uint32_t b0 = (word >> (16 * (s & 1))) & 0xFFu;
        const uint32_t b1 = (word >> (16 * (s & 1) + 8)) & 0xFFu;
        const uint32_t pos = 16 * h + 2 * s;
        const long r = (long)(((uint64_t)lut8[(pos + 1) * 256 + b1] << 32) | lut8[pos * 256 + b0]);
        accR[u] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(r, qf8[s], accR[u], 0, 0, 0);
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int u = 0; u < TPS; ++u)
#pragma unroll
      for (uint32_t j = 0; j < 16; ++j) {
        const uint32_t m = 8 * (j >> 2) + 4 * h + (j & 3);
        if (i0 + 32 * u + m < L) M = fmaxf(M, sq * (rsc[32 * u + m] * accC[u][j] + sr * accR[u][j]));
      }
    wave_lds_sync();
  }
  return fmaxf(M, __shfl_xor(M, 32));
}

__global__ __launch_bounds__(256) void k_rerank_fp8(
    const float *__restrict__ Q, uint32_t t_max, const uint32_t *__restrict__ q_ntok,
    const uint32_t *__restrict__ sel_ids, const uint32_t *__restrict__ sel_n, uint32_t R,
    uint32_t n_slice, const uint32_t *__restrict__ doclens, const uint64_t *__restrict__ docoff,
    const uint32_t *__restrict__ codes, const uint8_t *__restrict__ residual,
    const uint8_t *__restrict__ C8, const float *__restrict__ csc, const uint32_t *__restrict__ lut_g,
    float sr, int clip_zero, float *__restrict__ sel_score) {
  __shared__ uint32_t lut[kResBytes * 256];
  __shared__ float Ms[4][kMaxQTok];
  __shared__ float rsc[4][32];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, col = lane & 31;
  const uint32_t slice = blockIdx.x, q = blockIdx.y;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  const uint32_t n = sel_n[q];
  for (uint32_t i = tid; i < kResBytes * 256; i += 256) lut[i] = lut_g[i];
  long qf8[8];
  float sq;
  load_q_fp8(Q, q, t_max, T, h, col, qf8, sq);
  __syncthreads();
  for (uint32_t r = slice * 4 + wave; r < n; r += n_slice * 4) {
    const uint32_t doc = sel_ids[(size_t)q * R + r];
    const float M = maxsim_doc_fp8<1>(doclens[doc], docoff[doc], h, col, codes, residual, C8, csc, lut, sr,
                                      qf8, sq, rsc[wave], clip_zero);
    if (h == 0) Ms[wave][col] = M;
    wave_lds_sync();
    if (lane == 0) {
      float sc = 0.0f;
      for (uint32_t tt = 0; tt < T; ++tt) sc += Ms[wave][tt];
      sel_score[(size_t)q * R + r] = sc;
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Doc-major FP8 MaxSim: the pass's pairs grouped by doc as for the doc-major
// BF16 kernel, each doc's 32-token FP8 tiles loaded once per pass for every
// query that selected it.  Per pass, k_q_fp8m encodes the queries once: the
// B operand of lane (col, h) -- qf8[s] of load_q_fp8 -- in MFMA lane order
// (four coalesced 1 KB loads per pair, half the BF16 operand) and each
// token's scale.  A wave holds up to kDm8Tiles tiles of a doc in registers
// (centroid bytes, residual operands decoded once per chunk through the LDS
// table, row scales in LDS) and runs, for each pair of the doc, the
// per-query kernel's MFMA sequence and combine s_q (s_c acc_C + s_r acc_R)
// on the same operands: scores are bit-identical to k_rerank_fp8.  Longer
// docs go chunk by chunk over groups of up to kDmGroup pairs, the maxima
// combined in LDS.
// ---------------------------------------------------------------------------
constexpr uint32_t kDm8Tiles = 4;  // 32-token FP8 tiles held in registers (128 tokens)

// Qm8[(q * 4 + k) * 64 + lane] = (qf8[2k], qf8[2k + 1]) of lane (col, h) and
// sqm[q * 32 + col] = s_q of token col; one wave per query.
__global__ __launch_bounds__(256) void k_q_fp8m(const float *__restrict__ Q, uint32_t t_max,
                                                const uint32_t *__restrict__ q_ntok, uint32_t nb,
                                                uint4 *__restrict__ Qm8, float *__restrict__ sqm) {
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nb) return;  // whole waves
  const uint32_t h = lane >> 5, col = lane & 31;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  long qf8[8];
  float sq;
  load_q_fp8(Q, q, t_max, T, h, col, qf8, sq);
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint64_t a = (uint64_t)qf8[2 * k], b = (uint64_t)qf8[2 * k + 1];
    Qm8[((size_t)q * 4 + k) * 64 + lane] = make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b,
                                                      (uint32_t)(b >> 32));
  }
  if (h == 0) sqm[(size_t)q * 32 + col] = sq;
}

// One chunk of NT tiles (rows 32 t0 ..) of a doc against the ng pairs pg ..;
// per-token maxima into mg[j] (combined with earlier chunks unless t0 = 0).
template <uint32_t NT>
__device__ __forceinline__ void dm8_chunk(const uint4 *__restrict__ Qm8, const float *__restrict__ sqm, uint32_t R,
                                          const uint32_t *__restrict__ plist, uint32_t pg, uint32_t ng, uint32_t L,
                                          uint64_t off, uint32_t t0, const uint32_t *__restrict__ codes,
                                          const uint8_t *__restrict__ residual, const uint8_t *__restrict__ C8,
                                          const float *__restrict__ csc, const uint32_t *lut8, float sr, float init,
                                          uint32_t lane, float *rsc, float (*mg)[kMaxQTok + 1]) {
  const uint32_t h = lane >> 5, col = lane & 31;
  uint32_t cd[NT];
#pragma unroll
  for (uint32_t u = 0; u < NT; ++u) cd[u] = codes[off + min(32 * (t0 + u) + col, L - 1)];
  uint4 cv[NT][4];
  uint4 rb[NT];
  float scl[NT];
#pragma unroll
  for (uint32_t u = 0; u < NT; ++u) {
    const uint64_t g = off + min(32 * (t0 + u) + col, L - 1);
    rb[u] = *reinterpret_cast<const uint4 *>(residual + g * kResBytes + 16 * h);
    const uint4 *cp = reinterpret_cast<const uint4 *>(C8 + (size_t)cd[u] * kDim + 64 * h);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) cv[u][k] = cp[k];
    scl[u] = csc[cd[u]];
  }
  // row scales of the chunk for the combine (rows are lane-local in the
  // MFMA input, spread over lanes in its output)
  if (h == 0) {
#pragma unroll
    for (uint32_t u = 0; u < NT; ++u) rsc[32 * u + col] = scl[u];
  }
  // residual operands, decoded once for all pairs
  long rv[NT][8];
#pragma unroll
  for (uint32_t u = 0; u < NT; ++u) {
    const uint32_t rw[4] = {rb[u].x, rb[u].y, rb[u].z, rb[u].w};
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
      const uint32_t word = rw[s >> 1];
      const uint32_t b0 = (word >> (16 * (s & 1))) & 0xFFu;
      const uint32_t b1 = (word >> (16 * (s & 1) + 8)) & 0xFFu;
      const uint32_t pos = 16 * h + 2 * s;
      rv[u][s] = (long)(((uint64_t)lut8[(pos + 1) * 256 + b1] << 32) | lut8[pos * 256 + b0]);
    }
  }
  wave_lds_sync();
  // the next pair's query operand and scale are loaded while the current
  // pair's MFMA chains run (17 registers; the load latency is not exposed)
  uint4 qn[4];
  float sqn;
  auto load_pair = [&](uint32_t j) {
    const uint32_t q = plist[pg + j] / R;
    const uint4 *qp = Qm8 + (size_t)q * 4 * 64 + lane;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) qn[k] = qp[64 * k];
    sqn = sqm[(size_t)q * 32 + col];
  };
  load_pair(0);
  for (uint32_t j = 0; j < ng; ++j) {
    long qf8[8];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      qf8[2 * k] = (long)(((uint64_t)qn[k].y << 32) | qn[k].x);
      qf8[2 * k + 1] = (long)(((uint64_t)qn[k].w << 32) | qn[k].z);
    }
    const float sq = sqn;
    if (j + 1 < ng) load_pair(j + 1);
    float M = init;
#pragma unroll
    for (uint32_t u = 0; u < NT; ++u) {
      floatx16_t accC = floatx16_t{}, accR = floatx16_t{};
#pragma unroll
      for (uint32_t s = 0; s < 8; ++s) {
        const uint4 c4 = cv[u][s >> 1];
        const long a = (long)((s & 1) ? (((uint64_t)c4.w << 32) | c4.z) : (((uint64_t)c4.y << 32) | c4.x));
        accC = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a, qf8[s], accC, 0, 0, 0);
        accR = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(rv[u][s], qf8[s], accR, 0, 0, 0);
      }
      const int valid = (int)L - (int)(32 * (t0 + u));  // rows of this tile that hold tokens
#pragma unroll
      for (uint32_t e = 0; e < 16; ++e) {
        const uint32_t m = 8 * (e >> 2) + 4 * h + (e & 3);
        if ((int)m < valid) M = fmaxf(M, sq * (rsc[32 * u + m] * accC[e] + sr * accR[e]));
      }
    }
    M = fmaxf(M, __shfl_xor(M, 32));
    if (h == 0) mg[j][col] = t0 != 0 ? fmaxf(M, mg[j][col]) : M;
  }
  wave_lds_sync();  // rsc is rewritten by the next chunk
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_rerank_fp8_dm(
    const uint4 *__restrict__ Qm8, const float *__restrict__ sqm, uint32_t t_max,
    const uint32_t *__restrict__ q_ntok, uint32_t R, const uint2 *__restrict__ work, uint32_t *__restrict__ tot,
    const uint32_t *__restrict__ pend, const uint32_t *__restrict__ plist, const uint32_t *__restrict__ doclens,
    const uint64_t *__restrict__ docoff, const uint32_t *__restrict__ codes,
    const uint8_t *__restrict__ residual, const uint8_t *__restrict__ C8, const float *__restrict__ csc,
    const uint32_t *__restrict__ lut_g, float sr, int clip_zero, float *__restrict__ sel_score) {
  static_assert(kDm8Tiles == 4, "dm8_chunk is dispatched for 1 .. 4 tiles");
  __shared__ uint32_t lut[kResBytes * 256];
  __shared__ float rs[4][32 * kDm8Tiles];
  __shared__ float Mg[4][kDmGroup][kMaxQTok + 1];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (uint32_t i = tid; i < kResBytes * 256; i += 256) lut[i] = lut_g[i];
  __syncthreads();
  float(*mg)[kMaxQTok + 1] = Mg[wave];
  const uint32_t n_work = tot[1];
  const float init = clip_zero ? 0.0f : -FLT_MAX;
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&tot[2], kDmGrab);
    base = __builtin_amdgcn_readfirstlane(base);
    if (base >= n_work) break;
    const uint32_t w_end = min(base + kDmGrab, n_work);
    for (uint32_t wi = base; wi < w_end; ++wi) {
      const uint2 it = work[wi];
      const uint32_t doc = it.x, p0 = it.y, p1 = pend[doc];
      const uint32_t L = doclens[doc];
      const uint64_t off = docoff[doc];
      const uint32_t n_t = (L + 31) >> 5;
      for (uint32_t pg = p0; pg < p1; pg += kDmGroup) {
        const uint32_t ng = min(kDmGroup, p1 - pg);
        if (n_t == 0 && lane < 32)  // an empty doc: the identity for every token
          for (uint32_t j = 0; j < ng; ++j) mg[j][lane] = init;
        for (uint32_t t0 = 0; t0 < n_t; t0 += kDm8Tiles) {
#define MVR_DM8_CHUNK(N)                                                                                   \
  dm8_chunk<N>(Qm8, sqm, R, plist, pg, ng, L, off, t0, codes, residual, C8, csc, lut, sr, init, lane, rs[wave], \
               mg)
          switch (min(kDm8Tiles, n_t - t0)) {
            case 1: MVR_DM8_CHUNK(1); break;
            case 2: MVR_DM8_CHUNK(2); break;
            case 3: MVR_DM8_CHUNK(3); break;
            default: MVR_DM8_CHUNK(4); break;
          }
#undef MVR_DM8_CHUNK
        }
        // scores of the group: lane j sums pair j's maxima left to right
        wave_lds_sync();
        if (lane < ng) {
          const uint32_t pq = plist[pg + lane], q = pq / R;
          const uint32_t T = q_ntok ? q_ntok[q] : t_max;
          float sc = 0.0f;
          for (uint32_t tt = 0; tt < T; ++tt) sc += mg[lane][tt];
          sel_score[pq] = sc;
        }
        wave_lds_sync();
      }
    }
  }
}

// FP8 copies of the centroids (one wave per centroid) and of the residual
// weight table.
__global__ void k_to_fp8_rows(const float *__restrict__ C, uint32_t n_c, uint8_t *__restrict__ C8,
                              float *__restrict__ csc) {
  const uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= n_c) return;
  const float a = C[(size_t)c * kDim + 2 * lane], b = C[(size_t)c * kDim + 2 * lane + 1];
  float m = fmaxf(fabsf(a), fabsf(b));
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  const float inv = m > 0.0f ? 448.0f / m : 0.0f;
  reinterpret_cast<uint16_t *>(C8 + (size_t)c * kDim)[lane] = (uint16_t)fp8x2(a * inv, b * inv);
  if (lane == 0) csc[c] = m / 448.0f;
}

__global__ void k_lut_fp8(const float *__restrict__ W, float inv_sr, uint32_t *__restrict__ lut) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kResBytes * 256) return;
  const uint32_t pos = i >> 8, byte = i & 255u;
  float v[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) v[k] = W[(4 * pos + k) * 4 + ((byte >> (6 - 2 * k)) & 3u)] * inv_sr;
  lut[i] = fp8x2(v[0], v[1]) | (fp8x2(v[2], v[3]) << 16);
}

// ---------------------------------------------------------------------------
// Block-wide top-kk (kk <= 256) of n unique keys make_key(sc[i], ids[i]),
// written best first.  n <= 512: one bitonic sort in LDS.  Otherwise MSB-first
// radix select with 8-bit digits: keys above the pivot digit go to the output,
// the pivot bucket is refined until it fits the 512-key scratch, then sorted.
// ---------------------------------------------------------------------------
constexpr uint32_t kTopkMax = 256;
constexpr uint32_t kTopkScr = 512;

struct TopkLds {
  unsigned long long out[kTopkMax];
  unsigned long long scr[kTopkScr];
  uint32_t hist[256];
  uint32_t misc[8];
};

__device__ void block_topk_small(const float *__restrict__ sc, const uint32_t *__restrict__ ids,
                                 uint32_t n, uint32_t kk, TopkLds &L, uint32_t *__restrict__ o_ids,
                                 float *__restrict__ o_s, uint32_t *__restrict__ o_n) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const uint32_t n_sel = min(n, kk);
  unsigned long long *res;
  if (n <= kTopkScr) {
    const uint32_t np = pow2_ceil(max(n, 2u));
    for (uint32_t i = tid; i < np; i += nt) L.scr[i] = (i < n) ? make_key(sc[i], ids[i]) : 0ull;
    __syncthreads();
    bitonic_desc(L.scr, np);
    res = L.scr;
  } else {
    if (tid == 0) L.misc[0] = 0;
    uint64_t prefix = 0;
    uint32_t shift = 64, need = kk;
    for (;;) {
      const uint32_t hi = shift;
      shift -= 8;
      for (uint32_t b = tid; b < 256; b += nt) L.hist[b] = 0;
      __syncthreads();
      for (uint32_t i = tid; i < n; i += nt) {
        const uint64_t k = make_key(sc[i], ids[i]);
        if (hi == 64 || (k >> hi) == prefix) atomicAdd(&L.hist[(uint32_t)(k >> shift) & 255u], 1u);
      }
      __syncthreads();
      if (tid == 0) {
        uint32_t cum = 0, b = 256;
        while (b > 0) {
          --b;
          if (cum + L.hist[b] >= need) break;
          cum += L.hist[b];
        }
        L.misc[2] = b; L.misc[3] = cum; L.misc[4] = L.hist[b]; L.misc[1] = 0;
      }
      __syncthreads();
      const uint32_t dsel = L.misc[2], gt = L.misc[3], eq = L.misc[4];
      const bool take_eq = gt + eq == need;
      const bool eq_fits = eq <= kTopkScr;
      for (uint32_t i = tid; i < n; i += nt) {
        const uint64_t k = make_key(sc[i], ids[i]);
        if (hi == 64 || (k >> hi) == prefix) {
          const uint32_t dg = (uint32_t)(k >> shift) & 255u;
          if (dg > dsel || (dg == dsel && take_eq)) L.out[atomicAdd(&L.misc[0], 1u)] = k;
          else if (dg == dsel && eq_fits) L.scr[atomicAdd(&L.misc[1], 1u)] = k;
        }
      }
      __syncthreads();
      need -= gt;
      if (take_eq) break;
      if (eq_fits) {
        const uint32_t ns = L.misc[1];
        const uint32_t np = pow2_ceil(max(ns, 2u));
        for (uint32_t i = ns + tid; i < np; i += nt) L.scr[i] = 0ull;
        __syncthreads();
        bitonic_desc(L.scr, np);
        const uint32_t base = L.misc[0];
        for (uint32_t i = tid; i < need; i += nt) L.out[base + i] = L.scr[i];
        __syncthreads();
        break;
      }
      prefix = (prefix << 8) | dsel;  // keys are unique, so the bucket shrinks to fit
    }
    const uint32_t np = pow2_ceil(max(n_sel, 2u));
    for (uint32_t i = n_sel + tid; i < np; i += nt) L.out[i] = 0ull;
    __syncthreads();
    bitonic_desc(L.out, np);
    res = L.out;
  }
  for (uint32_t i = tid; i < kk; i += nt) {
    if (i < n_sel) {
      o_ids[i] = ~(uint32_t)(res[i] & 0xFFFFFFFFull);
      o_s[i] = ord2f((uint32_t)(res[i] >> 32));
    } else {
      o_ids[i] = 0xFFFFFFFFu;
      o_s[i] = -INFINITY;
    }
  }
  if (o_n && tid == 0) *o_n = n_sel;
}

// ---------------------------------------------------------------------------
// Persistent fused decompress + MaxSim + top-k.  A fixed grid of workgroups
// (sized to the resident capacity) pulls queries from a device queue; the 4
// waves of a workgroup score the query's candidates (one doc per wave at a
// time), then the same workgroup selects the query's top-kk from those scores
// -- no cross-workgroup hand-off, so candidate lists of very different lengths
// spread over the chip by the queue alone.  The decode tables are loaded into
// LDS once per workgroup, the query operand once per query.
// PREC 1: BF16 MFMA; PREC 3: FP32 MFMA.  TPS: tiles in flight per step
// (16 doc tokens each in BF16, 32 in FP32).
// ---------------------------------------------------------------------------
template <int PREC, int TPS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PREC == 3 ? 4 : 1))) void k_rerank_persist(
    const float *__restrict__ Q, uint32_t t_max, const uint32_t *__restrict__ q_ntok,
    const uint32_t *__restrict__ in_ids, const uint32_t *__restrict__ in_n, uint32_t R_in,
    uint32_t nb, uint32_t *__restrict__ queue, const uint32_t *__restrict__ doclens,
    const uint64_t *__restrict__ docoff, const uint32_t *__restrict__ codes,
    const uint8_t *__restrict__ residual, const float *__restrict__ C,
    const float *__restrict__ W, const uint16_t *__restrict__ Cm,
    const uint8_t *__restrict__ C8, const float *__restrict__ csc, const uint32_t *__restrict__ lut8_g,
    float sr, int clip_zero, float *__restrict__ score, uint32_t kk, uint32_t *__restrict__ out_ids,
    float *__restrict__ out_s, uint32_t *__restrict__ out_n, uint32_t out_stride) {
  __shared__ uint32_t tabb[PREC == 1 ? kPairTab : 1];
  __shared__ uint32_t lut8[PREC == 2 ? kResBytes * 256 : 1];
  __shared__ float2 tab2[PREC == 3 ? kDim / 2 * 16 : 1];
  __shared__ __align__(16) float ql[PREC == 3 ? kMaxQTok * kQlStride : 1];
  __shared__ float Ms3[PREC == 3 ? 4 : 1][4][kMaxQTok + 1];
  __shared__ float Ms[4][kMaxQTok];
  __shared__ float rsc[4][PREC == 2 ? 32 * TPS : 1];
  __shared__ TopkLds tk;
  __shared__ uint32_t s_q;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, col = lane & 31;
  if constexpr (PREC == 1)
    fill_pair_tab(tabb, W, tid, 256);
  else if constexpr (PREC == 2)
    for (uint32_t i = tid; i < kResBytes * 256; i += 256) lut8[i] = lut8_g[i];
  else
    f32mfma_fill_tab(tab2, W, tid);
  bf16x8_t qb[PREC == 1 ? 2 : 1][4];
  long q8[PREC == 2 ? 8 : 1];
  float sq = 0.0f;
  for (;;) {
    __syncthreads();  // the previous query's top-k is done with the LDS
    if (tid == 0) s_q = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t q = s_q;
    if (q >= nb) break;
    const uint32_t n = min(in_n[q], R_in);
    const uint32_t T = q_ntok ? q_ntok[q] : t_max;
    float *qs = score + (size_t)q * R_in;
    if constexpr (PREC == 3) {
      // the launch kernel's FP32 MFMA loop: query tile in LDS, docs streamed
      // as 8-token chunks packed four to a tile
      f32mfma_load_query(ql, Q, q, t_max, T, tid);
      __syncthreads();
      f32mfma_wave_docs(tab2, ql, Ms3[wave], lane, T, in_ids + (size_t)q * R_in, n, wave, 4, doclens, docoff,
                        codes, residual, C, clip_zero, qs);
    }
    if constexpr (PREC == 1) load_q16_bf16(Q, q, t_max, T, lane, qb);
    else if constexpr (PREC == 2) load_q_fp8(Q, q, t_max, T, h, col, q8, sq);
    for (uint32_t r = wave; r < (PREC == 3 ? 0u : n); r += 4) {
      const uint32_t doc = in_ids[(size_t)q * R_in + r];
      float M = 0.0f;
      if constexpr (PREC == 1)
        M = maxsim16_doc_bf16<TPS>(doclens[doc], docoff[doc], lane, codes, residual, Cm, tabb, qb,
                                   clip_zero);
      else if constexpr (PREC == 2)
        M = maxsim_doc_fp8<TPS>(doclens[doc], docoff[doc], h, col, codes, residual, C8, csc, lut8, sr, q8,
                                sq, rsc[wave], clip_zero);
      if (h == 0) Ms[wave][col] = M;
      wave_lds_sync();
      if (lane == 0) {
        float sc = 0.0f;
        for (uint32_t tt = 0; tt < T; ++tt) sc += Ms[wave][tt];
        qs[r] = sc;
      }
      wave_lds_sync();
    }
    __syncthreads();  // the query's scores are stored (same workgroup reads them)
    block_topk_small(qs, in_ids + (size_t)q * R_in, n, kk, tk, out_ids + (size_t)q * out_stride,
                     out_s + (size_t)q * out_stride, out_n ? out_n + q : nullptr);
  }
}

// ---------------------------------------------------------------------------
// Centroid scoring + top-P in BF16, pass 1: block maxima on MFMA.
// Workgroup = 8 waves = 8 queries x one slice of centroid tiles (32 centroids
// each).  A tile (32 rows x 128 bf16) is staged once in LDS and read by all 8
// waves as the A operand; each wave's B operand is its query's 32 tokens.
// For each (query token, tile) the kernel stores only the tile maximum.  The
// top-P centroids of a token always lie in its top-P tiles by maximum (a
// tile holding one of the P best scores has a maximum at least that large),
// so pass 2 only rescans those tiles.
// ---------------------------------------------------------------------------
constexpr uint32_t kBmWaves = 8;
// |bf16-input dot - exact FP32 dot| <= kBf16Margin ||q|| ||C|| for FP32 inputs:
// input rounding 2u + u^2 (u = 2^-8) plus two 128-term FP32 accumulations
constexpr float kBf16Margin = 0.008f;
constexpr uint32_t kBmStage = 8;  // tiles per LDS stage

__global__ __launch_bounds__(512) void k_ct_bf16_blockmax(
    const float *__restrict__ Q, uint32_t t_max, uint32_t n_q, const uint16_t *__restrict__ Cb,
    uint32_t n_c, uint32_t n_tiles, uint32_t tiles_per_split, float *__restrict__ bmax) {
  __shared__ uint4 As[kBmStage][32][17];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, col = lane & 31;
  const uint32_t q = blockIdx.x * kBmWaves + wave;
  const bool vq = q < n_q;
  bf16x8_t qf[8];
  {
    // the row pointer is clamped into the query array, so all 16 loads are
    // issued unconditionally (in flight together) and rows past the query's
    // tokens are zeroed afterwards
    const bool vr = vq && col < t_max;
    const float4 *qr =
        reinterpret_cast<const float4 *>(Q + ((size_t)(vq ? q : 0) * t_max + (vr ? col : 0)) * kDim + 64 * h);
    float4 x[16];
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) x[i] = qr[i];
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
      const float4 a = x[2 * s], b = x[2 * s + 1];
      bf16x8_t v;
      v[0] = (__bf16)(vr ? a.x : 0.0f); v[1] = (__bf16)(vr ? a.y : 0.0f);
      v[2] = (__bf16)(vr ? a.z : 0.0f); v[3] = (__bf16)(vr ? a.w : 0.0f);
      v[4] = (__bf16)(vr ? b.x : 0.0f); v[5] = (__bf16)(vr ? b.y : 0.0f);
      v[6] = (__bf16)(vr ? b.z : 0.0f); v[7] = (__bf16)(vr ? b.w : 0.0f);
      qf[s] = v;
    }
  }
  const uint32_t tb = blockIdx.y * tiles_per_split;
  const uint32_t te = min(n_tiles, tb + tiles_per_split);
  const uint4 *C4 = reinterpret_cast<const uint4 *>(Cb);
  // kBmStage tiles share one LDS stage (two barriers per stage, not per
  // tile); the next stage's rows are loaded into registers while the waves
  // run the current stage's MFMAs.  Thread tid stages 16-byte piece tid & 15
  // of row tid >> 4 of every tile of the stage.
  const uint32_t sr = tid >> 4, sk = tid & 15;
  uint4 pf[kBmStage];
  auto load_stage = [&](uint32_t t0) {
#pragma unroll
    for (uint32_t u = 0; u < kBmStage; ++u) {
      const uint32_t tile = t0 + u, c = tile * 32 + sr;
      pf[u] = (tile < te && c < n_c) ? C4[(size_t)c * 16 + sk] : make_uint4(0u, 0u, 0u, 0u);
    }
  };
  load_stage(tb);
  // maxima of the stage's tiles are kept in registers and stored together,
  // so each row's run of tile maxima is written as whole 32-byte pieces
  for (uint32_t t0 = tb; t0 < te; t0 += kBmStage) {
    __syncthreads();  // the previous stage's readers are done
#pragma unroll
    for (uint32_t u = 0; u < kBmStage; ++u) As[u][sr][sk] = pf[u];
    __syncthreads();
    if (t0 + kBmStage < te) load_stage(t0 + kBmStage);
    float mxb[kBmStage];
#pragma unroll
    for (uint32_t u = 0; u < kBmStage; ++u) {
      const uint32_t tile = t0 + u;
      mxb[u] = -INFINITY;
      if (tile < te) {  // uniform across the workgroup
        floatx16_t acc = {};
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(As[u][col][8 * h + s]), qf[s], acc, 0, 0, 0);
        float mx = -INFINITY;
        if (tile * 32 + 32 <= n_c) {  // a full tile (uniform): no row checks
          float t4[4];
#pragma unroll
          for (uint32_t g = 0; g < 4; ++g)
            t4[g] = fmaxf(fmaxf(acc[4 * g], acc[4 * g + 1]), fmaxf(acc[4 * g + 2], acc[4 * g + 3]));
          mx = fmaxf(fmaxf(t4[0], t4[1]), fmaxf(t4[2], t4[3]));
        } else {
#pragma unroll
          for (uint32_t j = 0; j < 16; ++j) {
            const uint32_t m = 8 * (j >> 2) + 4 * h + (j & 3);
            if (tile * 32 + m < n_c) mx = fmaxf(mx, acc[j]);
          }
        }
        mxb[u] = fmaxf(mx, __shfl_xor(mx, 32));
      }
    }
    if (h == 0 && vq && col < t_max) {
      float *row = bmax + ((size_t)q * t_max + col) * n_tiles;
#pragma unroll
      for (uint32_t u = 0; u < kBmStage; ++u)
        if (t0 + u < te) row[t0 + u] = mxb[u];
    }
  }
}

// Centroid scoring in FP8 (the FP8 mode), pass 1: tile maxima on
// v_mfma_f32_32x32x64_f8f6f4.  Centroid rows are the e4m3 codes of C_c / s_c
// (s_c = max_d |C_c[d]| / 448, the FP8 MaxSim's copy) and query tokens the
// e4m3 codes of q_t / s_t (s_t = max_d |q_t[d]| / 448, the FP8 MaxSim's
// per-token scale); score(t, c) = s_t s_c sum_d q8[d] c8[d], the products
// exact in FP32.  Lane (row / column l & 31, h = l >> 5) holds dims
// 64m + 32h .. +31 of its row in MFMA m = 0, 1 -- the same assignment for both
// operands, so the two MFMAs sum over all 128 dims.  Workgroup, staging and
// output as in k_ct_bf16_blockmax; a tile's 32 row scales are staged with its
// rows and applied to the accumulators before the maximum (s_t after it).
typedef int intx8_t __attribute__((ext_vector_type(8)));
constexpr uint32_t kF8Row = 9;  // uint4 per LDS row: 128 bytes + 16 of padding (distinct banks per 16 lanes)

__global__ __launch_bounds__(512) void k_ct_fp8_blockmax(
    const float *__restrict__ Q, uint32_t t_max, uint32_t n_q, const uint8_t *__restrict__ C8,
    const float *__restrict__ csc, uint32_t n_c, uint32_t n_tiles, uint32_t tiles_per_split,
    float *__restrict__ bmax) {
  __shared__ uint4 As[kBmStage][32][kF8Row];
  __shared__ __align__(16) float Ss[kBmStage][32];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, col = lane & 31;
  const uint32_t q = blockIdx.x * kBmWaves + wave;
  const bool vq = q < n_q;
  intx8_t qf[2];
  float st;
  {
    // the row pointer is clamped into the query array so that every load is
    // issued unconditionally; rows past the query's tokens are zeroed
    const bool vr = vq && col < t_max;
    const float *qr = Q + ((size_t)(vq ? q : 0) * t_max + (vr ? col : 0)) * kDim + 32 * h;
    float4 x[16];  // dims 64m + 32h + 4i .. +3 in x[8m + i]
#pragma unroll
    for (uint32_t m = 0; m < 2; ++m)
#pragma unroll
      for (uint32_t i = 0; i < 8; ++i) x[8 * m + i] = reinterpret_cast<const float4 *>(qr + 64 * m)[i];
    float amax = 0.0f;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) {
      if (!vr) x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      amax = fmaxf(amax, fmaxf(fmaxf(fabsf(x[i].x), fabsf(x[i].y)), fmaxf(fabsf(x[i].z), fabsf(x[i].w))));
    }
    amax = fmaxf(amax, __shfl_xor(amax, 32));  // the other half holds the other 64 dims
    st = amax / 448.0f;
    const float inv = amax > 0.0f ? 448.0f / amax : 0.0f;
#pragma unroll
    for (uint32_t m = 0; m < 2; ++m)
#pragma unroll
      for (uint32_t i = 0; i < 8; ++i) {
        const float4 v = x[8 * m + i];
        qf[m][i] = (int)(fp8x2(v.x * inv, v.y * inv) | (fp8x2(v.z * inv, v.w * inv) << 16));
      }
  }
  const uint32_t tb = blockIdx.y * tiles_per_split;
  const uint32_t te = min(n_tiles, tb + tiles_per_split);
  const uint4 *C4 = reinterpret_cast<const uint4 *>(C8);
  // a stage is kBmStage tiles of 32 rows x 8 pieces of 16 bytes: thread tid
  // loads pieces tid + 512 i (i = 0..3); threads below 256 load one row scale
  uint4 pf[4];
  float ps = 0.0f;
  auto load_stage = [&](uint32_t t0) {
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t pc = tid + 512 * i, u = pc >> 8, r = (pc >> 3) & 31u, k = pc & 7u;
      const uint32_t tile = t0 + u, c = tile * 32 + r;
      pf[i] = (tile < te && c < n_c) ? C4[(size_t)c * 8 + k] : make_uint4(0u, 0u, 0u, 0u);
    }
    const uint32_t tile = t0 + (tid >> 5), c = tile * 32 + (tid & 31u);
    ps = (tid < 256 && tile < te && c < n_c) ? csc[c] : 0.0f;
  };
  load_stage(tb);
  for (uint32_t t0 = tb; t0 < te; t0 += kBmStage) {
    __syncthreads();  // the previous stage's readers are done
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
      const uint32_t pc = tid + 512 * i;
      As[pc >> 8][(pc >> 3) & 31u][pc & 7u] = pf[i];
    }
    if (tid < 256) Ss[tid >> 5][tid & 31u] = ps;
    __syncthreads();
    if (t0 + kBmStage < te) load_stage(t0 + kBmStage);
    float mxb[kBmStage];
#pragma unroll
    for (uint32_t u = 0; u < kBmStage; ++u) {
      const uint32_t tile = t0 + u;
      mxb[u] = -INFINITY;
      if (tile < te) {  // uniform across the workgroup
        floatx16_t acc = {};
#pragma unroll
        for (uint32_t m = 0; m < 2; ++m) {
          const uint4 a0 = As[u][col][4 * m + 2 * h], a1 = As[u][col][4 * m + 2 * h + 1];
          const intx8_t av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                              (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
          acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, qf[m], acc, 0, 0, 0, 0, 0, 0);
        }
        // element 4g + e is row 8g + 4h + e: scaled by its centroid's s_c
        const bool full = tile * 32 + 32 <= n_c;  // uniform
        float mx = -INFINITY;
#pragma unroll
        for (uint32_t g = 0; g < 4; ++g) {
          const float4 sc = *reinterpret_cast<const float4 *>(&Ss[u][8 * g + 4 * h]);
          const float sv[4] = {sc.x, sc.y, sc.z, sc.w};
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e)
            if (full || tile * 32 + 8 * g + 4 * h + e < n_c) mx = vmax(mx, acc[4 * g + e] * sv[e]);
        }
        mxb[u] = vmax(mx, __shfl_xor(mx, 32)) * st;
      }
    }
    if (h == 0 && vq && col < t_max) {
      float *row = bmax + ((size_t)q * t_max + col) * n_tiles;
#pragma unroll
      for (uint32_t u = 0; u < kBmStage; ++u)
        if (t0 + u < te) row[t0 + u] = mxb[u];
    }
  }
}

// Pass 2: one wave per query-token row.  The tile maxima of pass 1 are
// approximate (bf16 inputs, FP32 accumulation in the matrix unit's order);
// |approx - exact| <= eps for every centroid, with
//   eps = margin * ||q|| * max_c ||C_c||
// (Cauchy-Schwarz on the rounding-error bound of a 128-term dot product).
// If A_P is the P-th largest approximate tile maximum, every centroid of the
// exact top-P (ties broken by id) lies in a tile whose approximate maximum is
// >= A_P - 2 eps: P tiles reach A_P, so P distinct centroids score >= A_P - eps
// exactly, hence the exact P-th score tau >= A_P - eps, and a centroid scoring
// >= tau has a tile maximum >= tau - eps.  Every centroid of those tiles is
// scored exactly -- one sequential FP32 FMA chain over d, on FP32 inputs
// (exact_f32) or bf16-rounded inputs (the BF16 mode's definition) -- and the
// top-P by (score desc, id asc) is kept.  Output S' = score + 1.
// The FP8 probe calls the same kernel with margin 0 and P_thr = the fixed
// number of tiles it rescans (every tile reaching the P_thr-th largest FP8
// maximum), without the bf16 pre-check (precheck = 0).
// A wave's sorted list of up to 64 J (score, id) pairs, best first by (score
// desc, id asc): element e = 64 j + lane sits in register j of lane `lane`.
// Only elements e < P are meaningful.
template <int J>
__device__ __forceinline__ void wave_list_tail(const float (&ls)[J], const uint32_t (&li)[J], uint32_t P,
                                               float &ts, uint32_t &ti) {
  const uint32_t jp = (P - 1) >> 6, lp = (P - 1) & 63u;
  ts = -INFINITY;
  ti = 0xFFFFFFFFu;
#pragma unroll
  for (int j = 0; j < J; ++j)
    if ((uint32_t)j == jp) {
      ts = readlane_f(ls[j], lp);
      ti = readlane_u(li[j], lp);
    }
}

// Inserts every lane's valid candidate (s, id) that beats element P-1, one
// at a time (lowest lane first): its position is the number of better
// elements, and the elements from there on move up by one (lane 0 of
// register j takes lane 63 of register j-1).
template <int J>
__device__ __forceinline__ void wave_list_insert(float (&ls)[J], uint32_t (&li)[J], uint32_t lane, uint32_t P,
                                                 float s, uint32_t id, bool valid) {
  float ts;
  uint32_t ti;
  wave_list_tail<J>(ls, li, P, ts, ti);
  uint64_t bal = __ballot(valid && better(s, id, ts, ti));
  while (bal) {
    const uint32_t l = (uint32_t)__builtin_ctzll(bal);
    const float cs = readlane_f(s, l);
    const uint32_t ci = readlane_u(id, l);
    uint32_t pos = 0;
#pragma unroll
    for (int j = 0; j < J; ++j)
      pos += (uint32_t)__popcll(__ballot(64u * j + lane < P && better(ls[j], li[j], cs, ci)));
    float carry_s = -INFINITY;
    uint32_t carry_i = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const float us = __shfl_up(ls[j], 1);
      const uint32_t ui = __shfl_up(li[j], 1);
      const float top_s = readlane_f(ls[j], 63);
      const uint32_t top_i = readlane_u(li[j], 63);
      const uint32_t e = 64u * j + lane;
      if (e > pos) {
        ls[j] = lane == 0 ? carry_s : us;
        li[j] = lane == 0 ? carry_i : ui;
      } else if (e == pos) {
        ls[j] = cs;
        li[j] = ci;
      }
      carry_s = top_s;
      carry_i = top_i;
    }
    bal &= bal - 1;
    wave_list_tail<J>(ls, li, P, ts, ti);
    bal &= __ballot(valid && better(s, id, ts, ti));
  }
}

template <bool kF32, int J>
__global__ __launch_bounds__(256) void k_ct_select(
    const float *__restrict__ Q, const uint32_t *__restrict__ row_map, uint32_t n_rows,
    const float *__restrict__ C, const uint16_t *__restrict__ Cb, uint32_t n_c, uint32_t n_tiles,
    const float *__restrict__ bmax, uint32_t P, uint32_t P_thr, float margin, float c_norm_max,
    int precheck, uint32_t *__restrict__ top_ids, float *__restrict__ top_s) {
  __shared__ float qs[4][kDim];
  __shared__ float qsb[4][kDim];  // bf16-rounded query (FP32 mode's pre-check)
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t row = blockIdx.x * 4 + wave;
  if (row >= n_rows) return;  // whole wave leaves; no block barrier below
  const uint64_t qrow = row_map ? row_map[row] : row;
  float nn = 0.0f;
  for (uint32_t d = lane; d < kDim; d += 64) {
    const float v = kF32 ? Q[qrow * kDim + d] : (float)(__bf16)Q[qrow * kDim + d];
    qs[wave][d] = v;
    qsb[wave][d] = (float)(__bf16)v;
    nn += v * v;
  }
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) nn += __shfl_xor(nn, off);
  wave_lds_sync();
  // A_P: the P_thr-th largest approximate tile maximum (P_thr = P with an
  // error margin; a fixed rank P_thr > P without one)
  const uint32_t P1 = min(P_thr, n_tiles);
  float ls[J];
  uint32_t li[J];
#pragma unroll
  for (int j = 0; j < J; ++j) { ls[j] = -INFINITY; li[j] = 0xFFFFFFFFu; }
  const float *bm = bmax + qrow * n_tiles;
  for (uint32_t base = 0; base < n_tiles; base += 64) {
    const uint32_t tl = base + lane;
    const bool v = tl < n_tiles;
    wave_list_insert<J>(ls, li, lane, P1, v ? bm[tl] : -INFINITY, tl, v);
  }
  float a_p;
  uint32_t a_pi;
  wave_list_tail<J>(ls, li, P1, a_p, a_pi);
  // 1.01: slack for rounding in the bound itself
  const float thr = a_p - 2.02f * margin * sqrtf(nn) * c_norm_max;
  // exact scores of every centroid in a tile reaching thr, two tiles per round
  float fs[J];
  uint32_t fi[J];
#pragma unroll
  for (int j = 0; j < J; ++j) { fs[j] = -INFINITY; fi[j] = 0xFFFFFFFFu; }
  for (uint32_t base = 0; base < n_tiles; base += 64) {
    const uint32_t tl = base + lane;
    uint64_t bal = __ballot(tl < n_tiles && bm[tl] >= thr);
    while (bal) {
      const uint32_t t0 = base + (uint32_t)__builtin_ctzll(bal);
      bal &= bal - 1;
      uint32_t t1 = 0xFFFFFFFFu;
      if (bal) { t1 = base + (uint32_t)__builtin_ctzll(bal); bal &= bal - 1; }
      const uint32_t tile = (lane < 32) ? t0 : t1;
      const uint32_t c = tile * 32 + (lane & 31);
      const bool v = tile != 0xFFFFFFFFu && c < n_c;
      float a = 0.0f;
      bool keep = v;
      if (kF32 && v && precheck && P1 == P) {
        // bf16 pre-check from the 256-byte bf16 row: this approximation is
        // within the same eps of the exact score as the tile maxima, so a
        // centroid of the exact top-P has it >= thr (exact >= tau >= A_P -
        // eps, which needs P tiles, i.e. P <= n_tiles); only those read the
        // 512-byte FP32 row
        const uint4 *cb = reinterpret_cast<const uint4 *>(Cb + (size_t)c * kDim);
        float ap = 0.0f;
#pragma unroll 4
        for (uint32_t k2 = 0; k2 < 16; ++k2) {
          const uint4 w = cb[k2];
          const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
          for (uint32_t e = 0; e < 4; ++e) {
            ap = __builtin_fmaf(qsb[wave][8 * k2 + 2 * e], __uint_as_float(ww[e] << 16), ap);
            ap = __builtin_fmaf(qsb[wave][8 * k2 + 2 * e + 1], __uint_as_float(ww[e] & 0xFFFF0000u), ap);
          }
        }
        keep = ap >= thr;
      }
      if (keep) {
        if (kF32) {
          const float4 *cr = reinterpret_cast<const float4 *>(C + (size_t)c * kDim);
#pragma unroll 8
          for (uint32_t k = 0; k < kDim / 4; ++k) {
            const float4 w = cr[k];
            a = __builtin_fmaf(qs[wave][4 * k], w.x, a);
            a = __builtin_fmaf(qs[wave][4 * k + 1], w.y, a);
            a = __builtin_fmaf(qs[wave][4 * k + 2], w.z, a);
            a = __builtin_fmaf(qs[wave][4 * k + 3], w.w, a);
          }
        } else {
          const uint4 *cr = reinterpret_cast<const uint4 *>(Cb + (size_t)c * kDim);
#pragma unroll 4
          for (uint32_t k = 0; k < 16; ++k) {
            const uint4 w = cr[k];
            const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
              a = __builtin_fmaf(qs[wave][8 * k + 2 * e], __uint_as_float(ww[e] << 16), a);
              a = __builtin_fmaf(qs[wave][8 * k + 2 * e + 1], __uint_as_float(ww[e] & 0xFFFF0000u), a);
            }
          }
        }
      }
      wave_list_insert<J>(fs, fi, lane, P, a + 1.0f, c, keep);
    }
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const uint32_t e = 64u * j + lane;
    if (e < P) {
      top_ids[qrow * P + e] = fi[j];
      top_s[qrow * P + e] = fs[j];
    }
  }
}

__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __builtin_bit_cast(uint16_t, (__bf16)x[i]);
}

// The MaxSim-order bf16 copy: y[c][8 (4s + g) + e] = bf16(x[c][32g + 8s + e]).
__global__ void k_to_bf16_maxsim(const float *__restrict__ x, uint16_t *__restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = (uint32_t)(i % kDim), piece = d >> 3, e = d & 7u;
  const uint32_t s = piece >> 2, g = piece & 3u;
  y[i] = __builtin_bit_cast(uint16_t, (__bf16)x[i - d + 32 * g + 8 * s + e]);
}

// ---------------------------------------------------------------------------
// Exhaustive MaxSim in FP64 over raw FP32 embeddings (the accuracy oracle that
// recall is measured against): out[q][doc] = sum_t max_i <Q_t, X_i> with every
// product and sum in double (products of two floats are exact in double).
// Same lane layout as k_rerank_f32: lane (t, h) keeps Q_t as 128 floats in
// VGPRs, doc tokens go through LDS as doubles two at a time.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_exhaustive_f64(
    const float *__restrict__ Q, uint32_t t_max, const uint32_t *__restrict__ q_ntok,
    const float *__restrict__ X, const uint32_t *__restrict__ doclens,
    const uint64_t *__restrict__ docoff, uint32_t n_doc, uint32_t n_slice,
    double *__restrict__ out) {
  __shared__ __align__(16) double xs[4][2][kDim];
  __shared__ double Ms[4][kMaxQTok];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t h = lane >> 5, t = lane & 31;
  const uint32_t slice = blockIdx.x, q = blockIdx.y;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  float qv[kDim];
  {
    const float4 *q4 = reinterpret_cast<const float4 *>(Q + ((size_t)q * t_max + t) * kDim);
#pragma unroll
    for (uint32_t i = 0; i < kDim / 4; ++i) {
      const float4 v = (t < T) ? q4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      qv[4 * i] = v.x; qv[4 * i + 1] = v.y; qv[4 * i + 2] = v.z; qv[4 * i + 3] = v.w;
    }
  }
  for (uint32_t doc = slice * 4 + wave; doc < n_doc; doc += n_slice * 4) {
    const uint32_t L = doclens[doc];
    const uint64_t off = docoff[doc];
    double M = -INFINITY;
    for (uint32_t i = 0; i < L; i += 2) {
      const uint32_t tok = i + h;
      const bool vt = tok < L;
      if (vt) {
        const float4 x = reinterpret_cast<const float4 *>(X + (off + tok) * kDim)[t];
        xs[wave][h][4 * t + 0] = (double)x.x;
        xs[wave][h][4 * t + 1] = (double)x.y;
        xs[wave][h][4 * t + 2] = (double)x.z;
        xs[wave][h][4 * t + 3] = (double)x.w;
      }
      wave_lds_sync();
      if (vt) {
        double a = 0.0;
#pragma unroll
        for (uint32_t d = 0; d < kDim; d += 2) {
          const double2 xv = *reinterpret_cast<const double2 *>(&xs[wave][h][d]);
          a = __builtin_fma((double)qv[d], xv.x, a);
          a = __builtin_fma((double)qv[d + 1], xv.y, a);
        }
        M = fmax(M, a);
      }
      wave_lds_sync();
    }
    M = fmax(M, __shfl_xor(M, 32));
    if (h == 0) Ms[wave][t] = M;
    wave_lds_sync();
    if (lane == 0) {
      double sc = 0.0;
      for (uint32_t tt = 0; tt < T; ++tt) sc += Ms[wave][tt];
      out[(size_t)q * n_doc + doc] = sc;
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Final top-k per query: sort the n reranked (score, doc) keys in LDS.
// ---------------------------------------------------------------------------
template <uint32_t CAP>
__global__ __launch_bounds__(kSelThreads) void k_topk_final(
    const uint32_t *__restrict__ ids, const float *__restrict__ score,
    const uint32_t *__restrict__ n_in, uint32_t R, uint32_t k, uint32_t *__restrict__ out_ids,
    float *__restrict__ out_s, uint32_t *__restrict__ out_n) {
  __shared__ unsigned long long s[CAP];
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  const uint32_t n = n_in[q];
  const uint32_t np = pow2_ceil(max(n, 2u));
  for (uint32_t i = tid; i < np; i += kSelThreads)
    s[i] = (i < n) ? make_key(score[(size_t)q * R + i], ids[(size_t)q * R + i]) : 0ull;
  __syncthreads();
  bitonic_desc(s, np);
  for (uint32_t i = tid; i < k; i += kSelThreads) {
    const size_t o = (size_t)q * k + i;
    if (i < n) {
      out_ids[o] = ~(uint32_t)(s[i] & 0xFFFFFFFFull);
      out_s[o] = ord2f((uint32_t)(s[i] >> 32));
    } else {
      out_ids[o] = 0xFFFFFFFFu;
      out_s[o] = -INFINITY;
    }
  }
  if (out_n && tid == 0) out_n[q] = min(n, k);
}

// ---------------------------------------------------------------------------
// Index preparation.
// ---------------------------------------------------------------------------
// seg_off[c][s] = first position in list c whose doc id is >= s*4096.
__global__ void k_seg_offsets(const uint64_t *__restrict__ ivf_off,
                              const uint32_t *__restrict__ ivf_size,
                              const uint32_t *__restrict__ ivf_ids, uint32_t n_c, uint32_t n_seg,
                              uint32_t *__restrict__ seg_off) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)n_c * (n_seg + 1);
  if (g >= total) return;
  const uint32_t c = (uint32_t)(g / (n_seg + 1)), s = (uint32_t)(g % (n_seg + 1));
  const uint32_t *lst = ivf_ids + ivf_off[c];
  const uint32_t sz = ivf_size[c];
  const uint64_t key = (uint64_t)s * kSeg;
  uint32_t lo = 0, hi = sz;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)lst[mid] < key) lo = mid + 1; else hi = mid;
  }
  seg_off[g] = lo;
}

// Index checks every kernel relies on: bad[0] counts codes >= n_c (a code is
// a centroid row index), bad[1] the IVF entries that are >= n_doc or not above
// their list predecessor (the filter's segment runs and the seg_off binary
// search need ascending lists).  One wave per list; codes grid-strided.
__global__ __launch_bounds__(256) void k_validate_index(const uint32_t *__restrict__ codes, uint64_t n_tok,
                                                        uint32_t n_c, const uint64_t *__restrict__ ivf_off,
                                                        const uint32_t *__restrict__ ivf_size,
                                                        const uint32_t *__restrict__ ivf_ids, uint32_t n_doc,
                                                        unsigned long long *__restrict__ bad) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
  uint32_t nb_code = 0, nb_ivf = 0;
  if (codes)
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_tok; i += nthr)
      nb_code += codes[i] >= n_c;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, n_wave = (uint32_t)(nthr >> 6);
  for (uint32_t c = wave; c < n_c; c += n_wave) {
    const uint32_t *lst = ivf_ids + ivf_off[c];
    const uint32_t sz = ivf_size[c];
    for (uint32_t i = lane; i < sz; i += 64) {
      const uint32_t d = lst[i];
      nb_ivf += (d >= n_doc) || (i > 0 && lst[i - 1] >= d);
    }
  }
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) {
    nb_code += __shfl_xor(nb_code, off);
    nb_ivf += __shfl_xor(nb_ivf, off);
  }
  if (lane == 0 && (nb_code | nb_ivf)) {
    atomicAdd(&bad[0], (unsigned long long)nb_code);
    atomicAdd(&bad[1], (unsigned long long)nb_ivf);
  }
}

// counters[0] += sum of the selected docs' lengths, counters[1] += candidates.
__global__ __launch_bounds__(256) void k_count_work(const uint32_t *__restrict__ sel_ids,
                                                    const uint32_t *__restrict__ sel_n, uint32_t R,
                                                    const uint32_t *__restrict__ doclens,
                                                    const uint32_t *__restrict__ cand_cnt,
                                                    const uint32_t *__restrict__ q_ntok,
                                                    unsigned long long *__restrict__ counters) {
  __shared__ unsigned long long wsum[256 / kWave];
  const uint32_t q = blockIdx.x;
  const uint32_t n = sel_n[q];
  unsigned long long t = 0;
  for (uint32_t r = threadIdx.x; r < n; r += 256) t += doclens[sel_ids[(size_t)q * R + r]];
  // one atomic per counter and query: contended atomics from every thread
  // would serialise on the three counters
#pragma unroll
  for (uint32_t off = 1; off < kWave; off <<= 1) t += __shfl_xor(t, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (uint32_t w = 0; w < 256 / kWave; ++w) tot += wsum[w];
    if (tot) {
      atomicAdd(&counters[0], tot);
      atomicAdd(&counters[2], tot * q_ntok[q]);
    }
    atomicAdd(&counters[1], (unsigned long long)cand_cnt[q]);
  }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// One thread per (token, 4 dims): x = C[code] + W[d][value], the value of
// dim d read MSB-first from bit d*n_bit of the token's packed row.
__global__ void k_decompress(const uint64_t *__restrict__ tok_ids, uint64_t n,
                             const uint32_t *__restrict__ codes,
                             const uint8_t *__restrict__ residual, const float *__restrict__ C,
                             const float *__restrict__ W, uint32_t n_bit, float *__restrict__ out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * (kDim / 4)) return;
  const uint64_t k = g / (kDim / 4);
  const uint32_t j = (uint32_t)(g % (kDim / 4));
  const uint64_t tok = tok_ids[k];
  const uint8_t *row = residual + tok * (kDim * n_bit / 8);
  const uint32_t nq = 1u << n_bit, mask = nq - 1;
  const float4 cv = reinterpret_cast<const float4 *>(C + (size_t)codes[tok] * kDim)[j];
  const float c[4] = {cv.x, cv.y, cv.z, cv.w};
  float x[4];
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i) {
    const uint32_t d = 4 * j + i, bit = d * n_bit;
    const uint32_t v = (row[bit >> 3] >> (8 - n_bit - (bit & 7u))) & mask;
    x[i] = c[i] + W[d * nq + v];
  }
  reinterpret_cast<float4 *>(out)[g] = make_float4(x[0], x[1], x[2], x[3]);
}

// ---------------------------------------------------------------------------
// Synthetic corpus and index build.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t key, uint64_t i) {
  return (double)(mix64(key + i) >> 11) * (1.0 / 9007199254740992.0);
}

// One wave per token; lane l produces normals 2l and 2l+1 of the token with
// one Box-Muller pair (same pairing as the CPU generator).
__global__ __launch_bounds__(256) void k_synth_tokens(uint64_t seed, uint64_t sc, uint64_t sn,
                                                      const float *__restrict__ centers,
                                                      uint32_t n_comp, double noise,
                                                      uint64_t tok_begin, uint64_t n_tok,
                                                      float *__restrict__ out) {
  __shared__ double red[4][kWave];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t k = (uint64_t)blockIdx.x * 4 + wave;
  const bool valid = k < n_tok;
  const uint64_t i = tok_begin + (valid ? k : 0);
  const uint64_t kc = mix64(seed ^ (sc * 0xD1B54A32D192ED03ull));
  const uint64_t kn = mix64(seed ^ (sn * 0xD1B54A32D192ED03ull));
  uint32_t comp = (uint32_t)(uniform01(kc, i) * (double)n_comp);
  if (comp >= n_comp) comp = n_comp - 1;
  const uint64_t pair = (i * kDim) / 2 + lane;
  const double u1 = uniform01(kn, 2 * pair), u2 = uniform01(kn, 2 * pair + 1);
  const double r = sqrt(-2.0 * log(1.0 - u1));
  const double a = 6.283185307179586 * u2;
  const double v0 = (double)centers[(size_t)comp * kDim + 2 * lane] + noise * (r * cos(a));
  const double v1 = (double)centers[(size_t)comp * kDim + 2 * lane + 1] + noise * (r * sin(a));
  red[wave][lane] = v0 * v0 + v1 * v1;
  wave_lds_sync();
  double nrm = 0.0;
  for (uint32_t l = 0; l < kWave; ++l) nrm += red[wave][l];
  nrm = sqrt(nrm);
  if (valid) {
    out[k * kDim + 2 * lane] = (float)(v0 / nrm);
    out[k * kDim + 2 * lane + 1] = (float)(v1 / nrm);
  }
}

// One thread per (token, packed byte): bucket = number of cutoffs <= residual
// (first cutoff strictly greater), packed MSB-first.
__global__ void k_compress(const float *__restrict__ x, const uint32_t *__restrict__ codes,
                           uint64_t n, const float *__restrict__ C, const float *__restrict__ cut,
                           int per_dim, uint8_t *__restrict__ out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * kResBytes) return;
  const uint64_t i = g / kResBytes;
  const uint32_t p = (uint32_t)(g % kResBytes);
  const float *xi = x + i * kDim;
  const float *ci = C + (size_t)codes[i] * kDim;
  uint32_t byte = 0;
  for (uint32_t v = 0; v < 4; ++v) {
    const uint32_t d = 4 * p + v;
    const float e = xi[d] - ci[d];
    const float *ct = cut + (per_dim ? d * 3 : 0);
    uint32_t b = 3;
    if (ct[0] > e) b = 0;
    else if (ct[1] > e) b = 1;
    else if (ct[2] > e) b = 2;
    byte = (byte << 2) | b;
  }
  out[g] = (uint8_t)byte;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_count_work(const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R, uint32_t n_q,
                             const DevIndex &ix, const uint32_t *cand_cnt, const uint32_t *q_ntok,
                             unsigned long long *counters, hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  k_count_work<<<n_q, 256, 0, s>>>(sel_ids, sel_n, R, ix.doclens, cand_cnt, q_ntok, counters);
  return hipGetLastError();
}

hipError_t launch_validate_index(const DevIndex &ix, unsigned long long *bad, hipStream_t s) {
  k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t *>(bad), 0u, 4);
  k_validate_index<<<2048, 256, 0, s>>>(ix.codes, ix.n_tok, ix.n_c, ix.ivf_off, ix.ivf_size, ix.ivf_ids, ix.n_doc,
                                        bad);
  return hipGetLastError();
}

hipError_t launch_seg_offsets(const DevIndex &ix, hipStream_t s) {
  const uint64_t total = (uint64_t)ix.n_c * (ix.n_seg + 1);
  k_seg_offsets<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(ix.ivf_off, ix.ivf_size,
                                                                ix.ivf_ids, ix.n_c, ix.n_seg,
                                                                ix.seg_off);
  return hipGetLastError();
}

hipError_t launch_centroid_topk(const float *Q, const uint32_t *row_map, uint32_t n_rows,
                                uint32_t n_out_rows, const DevIndex &ix, uint32_t P, int precision,
                                uint32_t n_split, uint32_t *part_ids, float *part_s,
                                uint32_t *top_ids, float *top_s, const uint16_t *Qbf,
                                hipStream_t s) {
  (void)n_out_rows;
  (void)Qbf;
  if (precision != 0) return hipErrorNotSupported;
  if (n_rows == 0) return hipSuccess;
  if (P > 64) return hipErrorInvalidValue;
  const uint32_t cps = ((ix.n_c + n_split - 1) / n_split + kCtChunk - 1) / kCtChunk * kCtChunk;
  const uint32_t ns = (ix.n_c + cps - 1) / cps;
  dim3 grid((n_rows + kRt - 1) / kRt, ns);
  k_centroid_topk_f32<0><<<grid, 256, 0, s>>>(Q, row_map, n_rows, ix.C, nullptr, ix.n_c, cps, P,
                                               part_ids, part_s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_topk_merge<<<(n_rows + 255) / 256, 256, 0, s>>>(part_ids, part_s, ns, n_rows, P, row_map,
                                                    top_ids, top_s);
  return hipGetLastError();
}

hipError_t launch_centroid_topk_bf16(const float *Q, uint32_t n_q, uint32_t t_max,
                                     const uint32_t *row_map, uint32_t n_rows, const DevIndex &ix,
                                     uint32_t P, float *bmax, uint32_t *top_ids, float *top_s,
                                     bool exact_f32, hipStream_t s) {
  if (n_q == 0 || n_rows == 0) return hipSuccess;
  if (!ix.C_bf16 || P == 0 || P > kMaxProbe) return hipErrorInvalidValue;
  const uint32_t n_tiles = (ix.n_c + 31) / 32;
  const uint32_t qblocks = (n_q + kBmWaves - 1) / kBmWaves;
  const uint32_t n_split = std::max(1u, std::min(n_tiles, (1024u + qblocks - 1) / qblocks));
  const uint32_t tps = (n_tiles + n_split - 1) / n_split;
  dim3 grid(qblocks, (n_tiles + tps - 1) / tps);
  k_ct_bf16_blockmax<<<grid, 512, 0, s>>>(Q, t_max, n_q, ix.C_bf16, ix.n_c, n_tiles, tps, bmax);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // error bounds per unit ||q|| ||C||: FP32 inputs vs bf16 MFMA = input rounding
  // (2u + u^2, u = 2^-8) plus two 128-term FP32 accumulations; bf16 inputs =
  // accumulation order only (products of bf16 values are exact in FP32), with
  // a wide allowance for the matrix unit's internal summation
  const float margin = exact_f32 ? kBf16Margin : 1e-4f;
  const float cmax = exact_f32 ? ix.c_norm_max : ix.c_norm_max * 1.004f;
  const dim3 g((n_rows + 3) / 4);
  // list registers per lane: 64 elements each
#define MVR_CT_SELECT(F32, J)                                                                            \
  k_ct_select<F32, J><<<g, 256, 0, s>>>(Q, row_map, n_rows, ix.C, ix.C_bf16, ix.n_c, n_tiles, bmax, P,  \
                                        P, margin, cmax, 1, top_ids, top_s)
  if (exact_f32) {
    if (P <= 64) MVR_CT_SELECT(true, 1);
    else if (P <= 128) MVR_CT_SELECT(true, 2);
    else MVR_CT_SELECT(true, 4);
  } else {
    if (P <= 64) MVR_CT_SELECT(false, 1);
    else if (P <= 128) MVR_CT_SELECT(false, 2);
    else MVR_CT_SELECT(false, 4);
  }
#undef MVR_CT_SELECT
  return hipGetLastError();
}

hipError_t launch_centroid_topk_fp8(const float *Q, uint32_t n_q, uint32_t t_max, const uint32_t *row_map,
                                    uint32_t n_rows, const DevIndex &ix, uint32_t P, float *bmax,
                                    uint32_t *top_ids, float *top_s, hipStream_t s) {
  if (n_q == 0 || n_rows == 0) return hipSuccess;
  if (!ix.C_fp8 || !ix.C_fp8_scale || P == 0 || P > kMaxProbe) return hipErrorInvalidValue;
  const uint32_t n_tiles = (ix.n_c + 31) / 32;
  const uint32_t qblocks = (n_q + kBmWaves - 1) / kBmWaves;
  const uint32_t n_split = std::max(1u, std::min(n_tiles, (1024u + qblocks - 1) / qblocks));
  const uint32_t tps = (n_tiles + n_split - 1) / n_split;
  dim3 grid(qblocks, (n_tiles + tps - 1) / tps);
  k_ct_fp8_blockmax<<<grid, 512, 0, s>>>(Q, t_max, n_q, ix.C_fp8, ix.C_fp8_scale, ix.n_c, n_tiles, tps, bmax);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // exact FP32 scores of every centroid in the kFp8ProbeTiles * P tiles with
  // the highest FP8 maxima (no error margin: the FP8 probe is approximate)
  // (the wave lists hold at most 256 tiles)
  const uint32_t P_thr = std::min(std::min(n_tiles, kFp8ProbeTiles * P), std::max(P, 256u));
  const dim3 g((n_rows + 3) / 4);
#define MVR_CT_SELECT(J)                                                                                   \
  k_ct_select<true, J><<<g, 256, 0, s>>>(Q, row_map, n_rows, ix.C, ix.C_bf16, ix.n_c, n_tiles, bmax, P, \
                                         P_thr, 0.0f, ix.c_norm_max, 0, top_ids, top_s)
  if (P_thr <= 64 && P <= 64) MVR_CT_SELECT(1);
  else if (P_thr <= 128 && P <= 128) MVR_CT_SELECT(2);
  else MVR_CT_SELECT(4);
#undef MVR_CT_SELECT
  return hipGetLastError();
}

hipError_t launch_filter(const uint32_t *top_ids, const float *top_s, uint32_t t_max, uint32_t P,
                         const uint32_t *q_ntok, const uint32_t *q_np, uint32_t n_q,
                         const DevIndex &ix, int clamp0, uint32_t *q_fill, uint64_t keys_per_q,
                         uint64_t *cand_keys, hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  const uint64_t n_wg = (uint64_t)ix.n_seg * n_q;
  // the fill counters start at zero (a fill kernel: graph-safe)
  k_fill_u32<<<(n_q + 255) / 256, 256, 0, s>>>(q_fill, 0u, n_q);
  if (n_wg > 0x7FFFFFF8ull || P == 0 || P > kMaxProbe) return hipErrorInvalidValue;
  const uint64_t n_grp = ((uint64_t)ix.n_seg + kFiltGroup - 1) / kFiltGroup * n_q;
  const uint32_t grid = (uint32_t)((n_grp + 7) / 8 * 8);  // the kernel's XCD remap needs a multiple of 8
#define MVR_FILTER(KERNEL)                                                                                  \
  KERNEL<<<grid, kFtThreads, 0, s>>>(top_ids, top_s, t_max, P, q_ntok, q_np, ix.ivf_off, ix.ivf_ids, ix.seg_off, \
                                     ix.n_seg, ix.n_doc, n_q, clamp0, q_fill, keys_per_q, cand_keys)
  // probe ranks per lane: the fewest that hold P
  if (P <= kFtLanes) MVR_FILTER(k_filter_tok<1>);
  else if (P <= 2 * kFtLanes) MVR_FILTER(k_filter_tok<2>);
  else if (P <= 4 * kFtLanes) MVR_FILTER(k_filter_tok<4>);
  else MVR_FILTER(k_filter_wide);
#undef MVR_FILTER
  return hipGetLastError();
}

hipError_t launch_select(const uint64_t *cand_keys, const uint32_t *q_fill, uint64_t keys_per_q,
                         const uint32_t *q_r, uint32_t n_q, uint32_t R, uint32_t *sel_ids,
                         float *sel_F, uint32_t *sel_n, uint32_t *cand_tot, hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  if (R <= kSelSmall)
    k_select_topr<kSelSmall, false><<<n_q, kSelThreads, 0, s>>>(cand_keys, nullptr, 0, q_fill, keys_per_q, q_r,
                                                                R, sel_ids, sel_F, sel_n, cand_tot);
  else
    k_select_topr<kMaxSelect, false><<<n_q, kSelThreads, 0, s>>>(cand_keys, nullptr, 0, q_fill, keys_per_q, q_r,
                                                                 R, sel_ids, sel_F, sel_n, cand_tot);
  return hipGetLastError();
}

hipError_t launch_select_segments(const uint64_t *cand_keys, const uint32_t *seg_cnt, uint32_t n_seg,
                                  const uint32_t *q_r, uint32_t n_q, uint32_t R, uint32_t *sel_ids,
                                  float *sel_F, uint32_t *sel_n, uint32_t *cand_tot, hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  if (R <= kSelSmall)
    k_select_topr<kSelSmall, true><<<n_q, kSelThreads, 0, s>>>(cand_keys, seg_cnt, n_seg, nullptr, 0, q_r, R,
                                                               sel_ids, sel_F, sel_n, cand_tot);
  else
    k_select_topr<kMaxSelect, true><<<n_q, kSelThreads, 0, s>>>(cand_keys, seg_cnt, n_seg, nullptr, 0, q_r, R,
                                                                sel_ids, sel_F, sel_n, cand_tot);
  return hipGetLastError();
}

hipError_t launch_rerank(const float *Q, uint32_t t_max, const uint32_t *q_ntok, uint32_t n_q,
                         const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R,
                         const DevIndex &ix, int precision, int clip_zero, float *sel_score,
                         hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  // one workgroup per 64 candidates of a query, capped at about 8k
  // workgroups in all (at least 4 per query)
  const uint32_t n_slice = std::min((R + 63) / 64, std::max(4u, 8192u / n_q));
  dim3 grid(n_slice, n_q);
  if (precision == 0) {
#define MVR_RR_F32(NB)                                                                                      \
  k_rerank_f32<NB><<<grid, 256, 0, s>>>(Q, t_max, q_ntok, sel_ids, sel_n, R, n_slice, ix.doclens, ix.docoff, \
                                        ix.codes, ix.residual, ix.C, ix.W, clip_zero, sel_score)
    switch (ix.n_bit) {
      case 1: MVR_RR_F32(1); break;
      case 2: MVR_RR_F32(2); break;
      case 4: MVR_RR_F32(4); break;
      case 8: MVR_RR_F32(8); break;
      default: return hipErrorNotSupported;
    }
#undef MVR_RR_F32
  } else if (ix.n_bit != 2) {
    return hipErrorNotSupported;  // the MFMA kernels decode 2-bit codes only
  } else if (precision == 3) {
    k_rerank_f32mfma<<<grid, 256, 0, s>>>(Q, t_max, q_ntok, sel_ids, sel_n, R, n_slice, ix.doclens,
                                          ix.docoff, ix.codes, ix.residual, ix.C, ix.W, clip_zero,
                                          sel_score);
  } else if (precision == 2) {
    if (!ix.C_fp8 || !ix.lut_fp8) return hipErrorInvalidValue;
    k_rerank_fp8<<<grid, 256, 0, s>>>(Q, t_max, q_ntok, sel_ids, sel_n, R, n_slice, ix.doclens, ix.docoff,
                                      ix.codes, ix.residual, ix.C_fp8, ix.C_fp8_scale, ix.lut_fp8,
                                      ix.fp8_res_scale, clip_zero, sel_score);
  } else if (precision == 1) {
    if (!ix.C_bf16m) return hipErrorInvalidValue;
    k_rerank_bf16<<<grid, 256, 0, s>>>(Q, t_max, q_ntok, sel_ids, sel_n, R, n_slice, ix.doclens,
                                       ix.docoff, ix.codes, ix.residual, ix.C_bf16m, ix.W, clip_zero,
                                       sel_score);
  } else {
    return hipErrorNotSupported;
  }
  return hipGetLastError();
}

hipError_t launch_rerank_bf16_dm(const float *Q, uint32_t t_max, const uint32_t *q_ntok, uint32_t n_q,
                                 const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R,
                                 const DevIndex &ix, int clip_zero, const DocMajorBufs &b,
                                 float *sel_score, hipStream_t s) {
  if (n_q == 0 || ix.n_doc == 0) return hipSuccess;
  if (ix.n_bit != 2) return hipErrorNotSupported;
  if (!ix.C_bf16m || (uint64_t)n_q * R > UINT32_MAX) return hipErrorInvalidValue;
  const uint32_t nblk = (ix.n_doc + kDmBlockDocs - 1) / kDmBlockDocs;
  static int occ = [] {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_rerank_bf16_dm, 256, 0) != hipSuccess || o < 1) o = 2;
    return o;
  }();
  k_fill_u32<<<(ix.n_doc + 255) / 256, 256, 0, s>>>(b.cnt, 0u, ix.n_doc);
  k_q_bf16m<<<(n_q * 512u + 255) / 256, 256, 0, s>>>(Q, t_max, q_ntok, n_q, b.Qm);
  const dim3 gq(std::min((R + 255) / 256, 16u), n_q);
  k_dm_count<<<gq, 256, 0, s>>>(sel_ids, sel_n, R, b.cnt);
  k_dm_bsum<<<nblk, 256, 0, s>>>(b.cnt, ix.n_doc, b.b_p, b.b_d);
  k_dm_bscan<<<1, 1024, 0, s>>>(b.b_p, b.b_d, nblk, b.tot);
  k_dm_offsets<<<nblk, 256, 0, s>>>(b.cnt, ix.n_doc, b.b_p, b.b_d, b.work);
  k_dm_scatter<<<gq, 256, 0, s>>>(sel_ids, sel_n, R, b.cnt, b.plist);
  const uint32_t grid = std::max(1u, ix.n_cu) * (uint32_t)occ;
  k_rerank_bf16_dm<<<grid, 256, 0, s>>>(b.Qm, t_max, q_ntok, R, b.work, b.tot, b.cnt, b.plist, ix.doclens,
                                        ix.docoff, ix.codes, ix.residual, ix.C_bf16m, ix.W, clip_zero,
                                        sel_score);
  return hipGetLastError();
}

hipError_t launch_rerank_fp8_dm(const float *Q, uint32_t t_max, const uint32_t *q_ntok, uint32_t n_q,
                                const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R,
                                const DevIndex &ix, int clip_zero, const DocMajorBufs &b,
                                float *sel_score, hipStream_t s) {
  if (n_q == 0 || ix.n_doc == 0) return hipSuccess;
  if (ix.n_bit != 2) return hipErrorNotSupported;
  if (!ix.C_fp8 || !ix.lut_fp8 || (uint64_t)n_q * R > UINT32_MAX) return hipErrorInvalidValue;
  const uint32_t nblk = (ix.n_doc + kDmBlockDocs - 1) / kDmBlockDocs;
  static const int occ = [] {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_rerank_fp8_dm, 256, 0) != hipSuccess || o < 1) o = 1;
    return o;
  }();
  // b.Qm: the FP8 operands (n_q x 4 KB), then the token scales (n_q x 128 B)
  uint4 *Qm8 = b.Qm;
  float *sqm = reinterpret_cast<float *>(b.Qm + (size_t)n_q * 256);
  k_fill_u32<<<(ix.n_doc + 255) / 256, 256, 0, s>>>(b.cnt, 0u, ix.n_doc);
  k_q_fp8m<<<(n_q + 3) / 4, 256, 0, s>>>(Q, t_max, q_ntok, n_q, Qm8, sqm);
  const dim3 gq(std::min((R + 255) / 256, 16u), n_q);
  k_dm_count<<<gq, 256, 0, s>>>(sel_ids, sel_n, R, b.cnt);
  k_dm_bsum<<<nblk, 256, 0, s>>>(b.cnt, ix.n_doc, b.b_p, b.b_d);
  k_dm_bscan<<<1, 1024, 0, s>>>(b.b_p, b.b_d, nblk, b.tot);
  k_dm_offsets<<<nblk, 256, 0, s>>>(b.cnt, ix.n_doc, b.b_p, b.b_d, b.work);
  k_dm_scatter<<<gq, 256, 0, s>>>(sel_ids, sel_n, R, b.cnt, b.plist);
  const uint32_t grid = std::max(1u, ix.n_cu) * (uint32_t)occ;
  k_rerank_fp8_dm<<<grid, 256, 0, s>>>(Qm8, sqm, t_max, q_ntok, R, b.work, b.tot, b.cnt, b.plist, ix.doclens,
                                       ix.docoff, ix.codes, ix.residual, ix.C_fp8, ix.C_fp8_scale, ix.lut_fp8,
                                       ix.fp8_res_scale, clip_zero, sel_score);
  return hipGetLastError();
}

hipError_t launch_persist(const float *Q, uint32_t t_max, const uint32_t *q_ntok,
                          const uint32_t *in_ids, const uint32_t *in_n, uint32_t R_in, uint32_t nb,
                          const DevIndex &ix, int precision, int clip_zero, const PersistCfg &cfg,
                          uint32_t *queue, float *score, uint32_t kk, uint32_t *out_ids, float *out_s,
                          uint32_t *out_n, uint32_t out_stride, hipStream_t s) {
  if (nb == 0) return hipSuccess;
  if (ix.n_bit != 2) return hipErrorNotSupported;
  if (kk == 0 || kk > kTopkMax || cfg.n_wg == 0) return hipErrorInvalidValue;
  if (precision == 1 && !ix.C_bf16m) return hipErrorInvalidValue;
  if (precision == 2 && (!ix.C_fp8 || !ix.lut_fp8)) return hipErrorInvalidValue;
  k_fill_u32<<<1, 64, 0, s>>>(queue, 0u, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint32_t grid = std::min(cfg.n_wg, nb);
#define MVR_PERSIST_ARGS                                                                           \
  Q, t_max, q_ntok, in_ids, in_n, R_in, nb, queue, ix.doclens, ix.docoff, ix.codes, ix.residual, \
      ix.C, ix.W, ix.C_bf16m, ix.C_fp8, ix.C_fp8_scale, ix.lut_fp8, ix.fp8_res_scale,   \
      clip_zero, score, kk, out_ids, out_s, out_n, out_stride
  if (precision == 1 && cfg.tps == 1) k_rerank_persist<1, 1><<<grid, 256, 0, s>>>(MVR_PERSIST_ARGS);
  else if (precision == 1) k_rerank_persist<1, 2><<<grid, 256, 0, s>>>(MVR_PERSIST_ARGS);
  else if (precision == 2 && cfg.tps == 1) k_rerank_persist<2, 1><<<grid, 256, 0, s>>>(MVR_PERSIST_ARGS);
  else if (precision == 2) k_rerank_persist<2, 2><<<grid, 256, 0, s>>>(MVR_PERSIST_ARGS);
  else if (precision == 3) k_rerank_persist<3, 1><<<grid, 256, 0, s>>>(MVR_PERSIST_ARGS);
  else return hipErrorNotSupported;
#undef MVR_PERSIST_ARGS
  return hipGetLastError();
}

uint32_t persist_max_wgs(int precision, uint32_t tps) {
  // resident workgroups per CU of each persistent kernel (registers and LDS)
  auto occ = [](const void *f) -> uint32_t {
    int o = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, f, 256, 0) == hipSuccess && o > 0 ? (uint32_t)o : 1u;
  };
  // (the FP32 kernel streams one packed 32-row tile per wave step: no
  // second tile setting)
  static const uint32_t t[3][2] = {
      {occ(reinterpret_cast<const void *>(k_rerank_persist<1, 1>)),
       occ(reinterpret_cast<const void *>(k_rerank_persist<1, 2>))},
      {occ(reinterpret_cast<const void *>(k_rerank_persist<2, 1>)),
       occ(reinterpret_cast<const void *>(k_rerank_persist<2, 2>))},
      {occ(reinterpret_cast<const void *>(k_rerank_persist<3, 1>)), 0u}};
  if (precision < 1 || precision > 3 || tps < 1 || tps > 2) return 0;
  return t[precision - 1][tps - 1];
}

hipError_t launch_exhaustive_f64(const float *Q, uint32_t n_q, uint32_t t_max,
                                 const uint32_t *q_ntok, const float *X, const uint32_t *doclens,
                                 const uint64_t *docoff, uint32_t n_doc, double *out,
                                 hipStream_t s) {
  if (n_q == 0 || n_doc == 0) return hipSuccess;
  const uint32_t n_slice = std::max(1u, std::min((n_doc + 3) / 4, 8192u / n_q));
  dim3 grid(n_slice, n_q);
  k_exhaustive_f64<<<grid, 256, 0, s>>>(Q, t_max, q_ntok, X, doclens, docoff, n_doc, n_slice, out);
  return hipGetLastError();
}

// ---- host-resident residuals: per-pass compaction of the selected docs ----
// The docs selected by a pass (over all its queries) are marked in a bitmap;
// compact id = rank of the doc among the marked ones, so compact ids follow
// doc ids.  Words of the bitmap go 256 to a block; per block the marked docs
// and their tokens are counted, one block scans those counts, and the list
// kernel writes (doc, length, token offset) per compact id.

__global__ void k_hs_mark(const uint32_t *__restrict__ sel_ids, const uint32_t *__restrict__ sel_n, uint32_t R,
                          uint32_t *__restrict__ bm) {
  const uint32_t q = blockIdx.x, n = min(sel_n[q], R);
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
    const uint32_t doc = sel_ids[(size_t)q * R + j];
    atomicOr(&bm[doc >> 5], 1u << (doc & 31u));
  }
}

// marked docs and their token count in word w
__device__ inline void hs_word(const uint32_t *__restrict__ bm, const uint32_t *__restrict__ doclens,
                               uint32_t n_words, uint32_t w, uint32_t &word, uint32_t &n,
                               unsigned long long &t) {
  word = w < n_words ? bm[w] : 0u;
  n = __popc(word);
  t = 0;
  for (uint32_t m = word; m; m &= m - 1) t += doclens[(size_t)w * 32 + (__ffs(m) - 1)];
}

// exclusive scan of (n, t) over a 256-thread block, plus the block totals
__device__ inline void hs_scan256(uint32_t n, unsigned long long t, uint32_t &ex_n, unsigned long long &ex_t,
                                  uint32_t &tot_n, unsigned long long &tot_t) {
  __shared__ uint32_t sn[4];
  __shared__ unsigned long long st[4];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t in = n;
  unsigned long long it = t;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t a = __shfl_up(in, o);
    const unsigned long long b = __shfl_up(it, o);
    if (lane >= o) {
      in += a;
      it += b;
    }
  }
  if (lane == 63) {
    sn[wv] = in;
    st[wv] = it;
  }
  __syncthreads();
  uint32_t bn = 0;
  unsigned long long bt = 0;
  tot_n = 0;
  tot_t = 0;
  for (uint32_t i = 0; i < 4; ++i) {
    if (i < wv) {
      bn += sn[i];
      bt += st[i];
    }
    tot_n += sn[i];
    tot_t += st[i];
  }
  ex_n = bn + in - n;
  ex_t = bt + it - t;
}

__global__ __launch_bounds__(256) void k_hs_wsum(const uint32_t *__restrict__ bm, const uint32_t *__restrict__ doclens,
                                                 uint32_t n_words, uint32_t *__restrict__ b_n,
                                                 unsigned long long *__restrict__ b_t) {
  uint32_t word, n;
  unsigned long long t;
  hs_word(bm, doclens, n_words, blockIdx.x * 256 + threadIdx.x, word, n, t);
  uint32_t ex_n, tot_n;
  unsigned long long ex_t, tot_t;
  hs_scan256(n, t, ex_n, ex_t, tot_n, tot_t);
  if (threadIdx.x == 0) {
    b_n[blockIdx.x] = tot_n;
    b_t[blockIdx.x] = tot_t;
  }
}

// in-place exclusive scan of the block counts; cnt = (docs, tokens) in total
__global__ __launch_bounds__(1024) void k_hs_bscan(uint32_t *__restrict__ b_n, unsigned long long *__restrict__ b_t,
                                                   uint32_t nblk, unsigned long long *__restrict__ cnt) {
  __shared__ uint32_t sn[16];
  __shared__ unsigned long long st[16];
  const uint32_t per = (nblk + 1023) / 1024;
  const uint32_t i0 = min(nblk, threadIdx.x * per), i1 = min(nblk, i0 + per);
  uint32_t n = 0;
  unsigned long long t = 0;
  for (uint32_t i = i0; i < i1; ++i) {
    n += b_n[i];
    t += b_t[i];
  }
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t in = n;
  unsigned long long it = t;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t a = __shfl_up(in, o);
    const unsigned long long b = __shfl_up(it, o);
    if (lane >= o) {
      in += a;
      it += b;
    }
  }
  if (lane == 63) {
    sn[wv] = in;
    st[wv] = it;
  }
  __syncthreads();
  uint32_t bn = 0, tn = 0;
  unsigned long long bt = 0, tt = 0;
  for (uint32_t i = 0; i < 16; ++i) {
    if (i < wv) {
      bn += sn[i];
      bt += st[i];
    }
    tn += sn[i];
    tt += st[i];
  }
  uint32_t base_n = bn + in - n;
  unsigned long long base_t = bt + it - t;
  for (uint32_t i = i0; i < i1; ++i) {
    const uint32_t a = b_n[i];
    const unsigned long long b = b_t[i];
    b_n[i] = base_n;
    b_t[i] = base_t;
    base_n += a;
    base_t += b;
  }
  if (threadIdx.x == 0) {
    cnt[0] = tn;
    cnt[1] = tt;
  }
}

__global__ __launch_bounds__(256) void k_hs_list(const uint32_t *__restrict__ bm, const uint32_t *__restrict__ doclens,
                                                 uint32_t n_words, const uint32_t *__restrict__ b_n,
                                                 const unsigned long long *__restrict__ b_t,
                                                 uint32_t *__restrict__ wbase, uint32_t *__restrict__ docs,
                                                 uint32_t *__restrict__ lens, uint64_t *__restrict__ offs) {
  const uint32_t w = blockIdx.x * 256 + threadIdx.x;
  uint32_t word, n;
  unsigned long long t;
  hs_word(bm, doclens, n_words, w, word, n, t);
  uint32_t ex_n, tot_n;
  unsigned long long ex_t, tot_t;
  hs_scan256(n, t, ex_n, ex_t, tot_n, tot_t);
  uint32_t c = b_n[blockIdx.x] + ex_n;
  unsigned long long tb = b_t[blockIdx.x] + ex_t;
  if (w < n_words) wbase[w] = c;
  for (uint32_t m = word; m; m &= m - 1) {
    const uint32_t doc = w * 32 + (__ffs(m) - 1);
    const uint32_t len = doclens[doc];
    docs[c] = doc;
    lens[c] = len;
    offs[c] = tb;
    ++c;
    tb += len;
  }
}

__global__ void k_hs_remap(const uint32_t *__restrict__ sel_ids, const uint32_t *__restrict__ sel_n, uint32_t R,
                           const uint32_t *__restrict__ bm, const uint32_t *__restrict__ wbase,
                           uint32_t *__restrict__ cids) {
  const uint32_t q = blockIdx.x, n = min(sel_n[q], R);
  for (uint32_t j = threadIdx.x; j < R; j += blockDim.x) {
    uint32_t c = 0;
    if (j < n) {
      const uint32_t doc = sel_ids[(size_t)q * R + j], w = doc >> 5;
      c = wbase[w] + __popc(bm[w] & ((1u << (doc & 31u)) - 1u));
    }
    cids[(size_t)q * R + j] = c;
  }
}

// One wave per compact doc: its codes and residual rows read from host
// memory mapped into the device's address space, written to the pass buffers.
__global__ __launch_bounds__(256) void k_hs_gather(const uint32_t *__restrict__ docs, const uint32_t *__restrict__ lens,
                                                   const uint64_t *__restrict__ offs, const uint64_t *__restrict__ docoff,
                                                   uint32_t U, const uint32_t *__restrict__ h_codes,
                                                   const uint4 *__restrict__ h_res, uint32_t *__restrict__ codes,
                                                   uint4 *__restrict__ res) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n_waves = gridDim.x * 4;
  for (uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6); c < U; c += n_waves) {
    const uint64_t src = docoff[docs[c]], dst = offs[c];
    const uint32_t len = lens[c];
    for (uint32_t i = lane; i < len; i += 64) codes[dst + i] = h_codes[src + i];
    // 32 bytes per token = two uint4; four loads in flight per lane
    const uint32_t n4 = len * 2;
    const uint4 *sp = h_res + src * 2;
    uint4 *dp = res + dst * 2;
    for (uint32_t i = lane; i < n4; i += 256) {
      uint4 v[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u)
        if (i + u * 64 < n4) v[u] = sp[i + u * 64];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u)
        if (i + u * 64 < n4) dp[i + u * 64] = v[u];
    }
  }
}

hipError_t launch_hs_compact(const uint32_t *sel_ids, const uint32_t *sel_n, uint32_t R, uint32_t nb,
                             const uint32_t *doclens, uint32_t n_doc, const HostStage &hs, hipStream_t s) {
  const uint32_t n_words = (n_doc + 31) / 32, nblk = (n_words + 255) / 256;
  hipError_t e = launch_fill_u32(hs.bm, 0u, n_words, s);
  if (e != hipSuccess || nb == 0) return e;
  k_hs_mark<<<nb, 256, 0, s>>>(sel_ids, sel_n, R, hs.bm);
  k_hs_wsum<<<nblk, 256, 0, s>>>(hs.bm, doclens, n_words, hs.b_n, hs.b_t);
  k_hs_bscan<<<1, 1024, 0, s>>>(hs.b_n, hs.b_t, nblk, hs.cnt);
  k_hs_list<<<nblk, 256, 0, s>>>(hs.bm, doclens, n_words, hs.b_n, hs.b_t, hs.wbase, hs.docs, hs.lens, hs.offs);
  k_hs_remap<<<nb, 256, 0, s>>>(sel_ids, sel_n, R, hs.bm, hs.wbase, hs.cids);
  return hipGetLastError();
}

// The gather is bound by the host link, not by the CUs: a few blocks keep
// enough reads in flight, and the rest of the chip stays free for the MaxSim
// of the other pass in flight (MVR_HS_GATHER_BLOCKS overrides the count).
hipError_t launch_hs_gather(const HostStage &hs, uint32_t U, const uint64_t *docoff, const uint32_t *h_codes,
                            const uint8_t *h_res, uint32_t *codes, uint8_t *res, hipStream_t s) {
  if (U == 0) return hipSuccess;
  static const uint32_t max_blocks = [] {
    const char *e = std::getenv("MVR_HS_GATHER_BLOCKS");
    const long v = e ? std::strtol(e, nullptr, 10) : 0;
    return v > 0 && v <= 65536 ? (uint32_t)v : 128u;
  }();
  const uint32_t grid = std::min<uint32_t>((U + 3) / 4, max_blocks);
  k_hs_gather<<<grid, 256, 0, s>>>(hs.docs, hs.lens, hs.offs, docoff, U, h_codes,
                                   reinterpret_cast<const uint4 *>(h_res), codes, reinterpret_cast<uint4 *>(res));
  return hipGetLastError();
}

// ids[i] = map[ids[i]] where ids[i] < n_map (compact ids of a host-staged pass
// back to doc ids; the 0xFFFFFFFF fill of short result lists stays)
__global__ void k_map_ids(uint32_t *ids, uint32_t n, const uint32_t *map, uint32_t n_map) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t v = ids[i];
    if (v < n_map) ids[i] = map[v];
  }
}

hipError_t launch_map_ids(uint32_t *ids, uint32_t n, const uint32_t *map, uint32_t n_map, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_map_ids<<<(n + 255) / 256, 256, 0, s>>>(ids, n, map, n_map);
  return hipGetLastError();
}

hipError_t launch_topk(const uint32_t *ids, const float *score, const uint32_t *n_in, uint32_t R,
                       uint32_t n_q, uint32_t k, uint32_t *out_ids, float *out_s, uint32_t *out_n,
                       hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  if (R <= kSelSmall)
    k_topk_final<kSelSmall><<<n_q, kSelThreads, 0, s>>>(ids, score, n_in, R, k, out_ids, out_s, out_n);
  else
    k_topk_final<kMaxSelect><<<n_q, kSelThreads, 0, s>>>(ids, score, n_in, R, k, out_ids, out_s, out_n);
  return hipGetLastError();
}

hipError_t launch_decompress(const uint64_t *tok_ids, uint64_t n, const DevIndex &ix, float *out,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t total = n * (kDim / 4);
  if (ix.n_bit != 1 && ix.n_bit != 2 && ix.n_bit != 4 && ix.n_bit != 8) return hipErrorNotSupported;
  k_decompress<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(tok_ids, n, ix.codes, ix.residual,
                                                               ix.C, ix.W, ix.n_bit, out);
  return hipGetLastError();
}

hipError_t launch_synth_tokens(uint64_t seed, uint64_t sc, uint64_t sn, const float *centers,
                               uint32_t n_comp, double noise, uint64_t tok_begin, uint64_t n_tok,
                               float *out, hipStream_t s) {
  if (n_tok == 0) return hipSuccess;
  k_synth_tokens<<<(unsigned)((n_tok + 3) / 4), 256, 0, s>>>(seed, sc, sn, centers, n_comp, noise,
                                                             tok_begin, n_tok, out);
  return hipGetLastError();
}

hipError_t launch_assign(const float *x, uint64_t n, const float *C, const float *bias,
                         uint32_t n_c, uint32_t *codes, hipStream_t s) {
  // Rows go in launches of at most 2^20 so the per-row score scratch stays small.
  const uint32_t kRowsPerLaunch = 1u << 20;
  float *ps = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&ps), sizeof(float) * kRowsPerLaunch);
  if (e != hipSuccess) return e;
  const uint32_t cps = (n_c + kCtChunk - 1) / kCtChunk * kCtChunk;
  for (uint64_t r0 = 0; r0 < n; r0 += kRowsPerLaunch) {
    const uint32_t nr = (uint32_t)((n - r0) < kRowsPerLaunch ? (n - r0) : kRowsPerLaunch);
    dim3 grid((nr + kRt - 1) / kRt, 1);
    k_centroid_topk_f32<1><<<grid, 256, 0, s>>>(x + r0 * kDim, nullptr, nr, C, bias, n_c, cps, 1,
                                                codes + r0, ps);
    e = hipGetLastError();
    if (e != hipSuccess) break;
  }
  hipError_t e2 = hipStreamSynchronize(s);
  hipError_t e3 = hipFree(ps);
  return e != hipSuccess ? e : (e2 != hipSuccess ? e2 : e3);
}

hipError_t launch_compress(const float *x, const uint32_t *codes, uint64_t n, const float *C,
                           const float *cut, int per_dim, uint8_t *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t total = n * kResBytes;
  k_compress<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(x, codes, n, C, cut, per_dim, out);
  return hipGetLastError();
}

hipError_t launch_convert_centroids(const DevIndex &ix, hipStream_t s) {
  const size_t n = (size_t)ix.n_c * kDim;
  if (ix.C_bf16) {
    k_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ix.C, ix.C_bf16, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (ix.C_bf16m) {
    k_to_bf16_maxsim<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ix.C, ix.C_bf16m, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (ix.C_fp8 && ix.C_fp8_scale) {
    k_to_fp8_rows<<<(ix.n_c + 3) / 4, 256, 0, s>>>(ix.C, ix.n_c, ix.C_fp8, ix.C_fp8_scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (ix.lut_fp8) {
    k_lut_fp8<<<(kResBytes * 256 + 255) / 256, 256, 0, s>>>(ix.W, ix.fp8_res_scale > 0 ? 1.0f / ix.fp8_res_scale : 0.0f,
                                                             ix.lut_fp8);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace mvr
