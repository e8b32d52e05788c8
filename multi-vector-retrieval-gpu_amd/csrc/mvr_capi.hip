// C ABI of libmvr: index upload/load, batched search orchestration,
// decompression and the device-side synthetic index helpers.  Host-only code
// apart from the launch calls into mvr_kernels.hip.
//
// Replaces (reference paths): IGPGPUPP ctor/buildIndex/search
// (src/impl/IGPGPUPP.hpp:67-332), the per-thread GPGPURetrievalPP pipeline
// (src/include/alg/igp_gpu_pp/IGPGPURetrievalPP.hpp:73-176) and the device
// structs of DataStructPP.hpp:16-205.  Differences by design: the whole index
// (including residual codes) is device resident unless the index is created
// with MVR_INDEX_RESIDUAL_ON_HOST, queries are processed in batches of up to
// `batch` per device pass, and errors are returned as status codes instead of
// exit().
#include "../../include/mvr.h"
#include "mvr_internal.h"
#include "npy_io.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

using namespace mvr;

// The no-alloc entry point's scratch (see kNoAllocMaxQ below): one copy per
// device, held until the process ends.
static __device__ __align__(256) uint32_t scratch[2 * 16384];

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail(MVR_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                   \
  } while (0)

// Scratch of mvr_rerank_docs_noalloc: the per-query token counts [0, n_q)
// and doc counts [n_q, 2 n_q) of one call.  Host code never uses the array's
// name as an address: each device's copy is reached through the address
// hipGetSymbolAddress returns for that device (scratch_on).
constexpr uint32_t kNoAllocMaxQ = 16384;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_dev[kMaxDevices] = {};

// Writes defaults for absent host arrays and clamps doc counts to R.
__global__ void k_noalloc_meta(uint32_t *meta, uint32_t n_q, uint32_t t_max, uint32_t R, int have_nt,
                               int have_nd) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_q) return;
  if (!have_nt) meta[q] = t_max;
  meta[n_q + q] = have_nd ? min(meta[n_q + q], R) : R;
}

// Growable device buffer.
struct DBuf {
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

// Growable pinned host buffer (metadata staging).
struct HBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t b = std::max<size_t>(bytes, 256);
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e == hipSuccess) cap = b;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// A captured device pass, replayed while the pass shape and the workspace
// buffers stay the same.
struct PassGraph {
  std::vector<uint64_t> key;
  hipGraphExec_t exec = nullptr;
};

struct Workspace {
  DBuf Q, meta, part_ids, part_s, top_ids, top_s, cand_cnt, cand_tot, cand_keys, sel_ids, sel_F, sel_n,
      sel_score, out_ids, out_s, counters, res_ids, res_s, res_n, bmax, feat, mlp_out, dyn_np, dyn_r,
      rr_meta, queue, pass_nt, pass_np, pass_rm, hs_bm, hs_bn, hs_bt, hs_wbase;
  // doc-major BF16 MaxSim (DocMajorBufs)
  DBuf dm_cnt, dm_work, dm_plist, dm_bp, dm_bd, dm_tot, dm_Qm;
  // host-resident residuals, per pass slot: compact lists and the staged docs
  DBuf hs_docs[2], hs_lens[2], hs_offs[2], hs_cids[2], hs_cnt[2], hs_codes[2], hs_res[2];
  HBuf staging;
  HBuf hs_cnt_h;  // [2][2] (docs, tokens) of each slot, read back
  hipStream_t hs_stream = nullptr;              // the staging copies of a pipelined search
  hipEvent_t hs_compacted[2] = {nullptr, nullptr}, hs_staged[2] = {nullptr, nullptr};
  hipEvent_t staging_done = nullptr;
  std::vector<PassGraph> graphs;  // most recently used last
  std::vector<uint64_t> graph_sig;  // buffer addresses the graphs were captured with
  hipStream_t cap_stream = nullptr;
  std::vector<DBuf *> all() {
    return {&Q, &meta, &part_ids, &part_s, &top_ids, &top_s, &cand_cnt, &cand_tot, &cand_keys,
            &sel_ids, &sel_F, &sel_n, &sel_score, &out_ids, &out_s, &counters, &res_ids, &res_s,
            &res_n, &bmax, &feat, &mlp_out, &dyn_np, &dyn_r, &rr_meta, &queue, &pass_nt, &pass_np,
            &pass_rm, &hs_bm, &hs_bn, &hs_bt, &hs_wbase, &hs_docs[0], &hs_docs[1],
            &hs_lens[0], &hs_lens[1], &hs_offs[0], &hs_offs[1], &hs_cids[0], &hs_cids[1], &hs_cnt[0],
            &hs_cnt[1], &hs_codes[0], &hs_codes[1], &hs_res[0], &hs_res[1], &dm_cnt, &dm_work,
            &dm_plist, &dm_bp, &dm_bd, &dm_tot, &dm_Qm};
  }
  void drop_graphs() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  void release() {
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    cap_stream = nullptr;
    for (DBuf *b : all()) b->release();
    staging.release();
    hs_cnt_h.release();
    if (hs_stream) (void)hipStreamDestroy(hs_stream);
    hs_stream = nullptr;
    for (int i = 0; i < 2; ++i) {
      if (hs_compacted[i]) (void)hipEventDestroy(hs_compacted[i]);
      if (hs_staged[i]) (void)hipEventDestroy(hs_staged[i]);
      hs_compacted[i] = hs_staged[i] = nullptr;
    }
    if (staging_done) (void)hipEventDestroy(staging_done);
    staging_done = nullptr;
  }
};

}  // namespace

struct mvr_index {
  int device = 0;
  mvr_mlp *probe = nullptr;    // learned probe model (not owned)
  float probe_conf = 0.6f;
  uint32_t probe_fallback = 16, probe_rmin = 0;
  DevIndex d;
  std::vector<uint32_t> h_doclens;
  Workspace ws;
  PersistCfg tune[4];          // work-queue MaxSim settings per precision
  uint32_t maxsim_order = MVR_MAXSIM_AUTO;  // BF16 MaxSim loop order (mvr_index_set_maxsim_order)
  bool tuned[4] = {false, false, false, false};
  // MVR_INDEX_RESIDUAL_ON_HOST: residual bytes and a copy of the codes stay in
  // host memory (residual pinned); each pass stages its selected docs
  uint64_t pool_budget = kPoolBudget;  // candidate pool bytes per pass (from free HBM at creation)
  bool res_host = false;
  uint8_t *h_res = nullptr;          // [n_tok][32], pinned and mapped
  uint32_t *h_codes = nullptr;       // [n_tok], pinned and mapped
  const uint8_t *h_res_dev = nullptr;  // their device-side addresses
  const uint32_t *h_codes_dev = nullptr;
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// Device address of `scratch` on `device` (that device's copy), looked up once
// per device with the device current and kept.
int scratch_on(int device, uint32_t **out) {
  static_assert(sizeof(scratch) == 2 * kNoAllocMaxQ * sizeof(uint32_t), "scratch size");
  if (device < 0 || device >= kMaxDevices) return fail(MVR_ERR_UNSUPPORTED, "device %d: no scratch slot", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    DeviceGuard g(device);
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MVR_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<uint32_t *>(p);
  }
  *out = g_scratch_dev[device];
  return MVR_OK;
}

void free_devindex(DevIndex &d) {
  for (void *p : {(void *)d.C, (void *)d.C_bf16, (void *)d.C_bf16m, (void *)d.C_fp8, (void *)d.C_fp8_scale,
                  (void *)d.codes, (void *)d.residual, (void *)d.W, (void *)d.doclens,
                  (void *)d.docoff, (void *)d.ivf_off, (void *)d.ivf_size, (void *)d.ivf_ids,
                  (void *)d.seg_off, (void *)d.lut_fp8})
    if (p) (void)hipFree(p);
  d = DevIndex();
}

template <class T>
hipError_t upload(T **dst, const T *src, size_t n, bool src_on_device) {
  hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess || n == 0) return e;
  return hipMemcpy(*dst, src, n * sizeof(T),
                   src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
}

template <class T> std::vector<T> fetch(const T *src, size_t n, bool on_device) {
  std::vector<T> v(n);
  if (n == 0) return v;
  if (on_device) (void)hipMemcpy(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost);
  else std::memcpy(v.data(), src, n * sizeof(T));
  return v;
}

}  // namespace

extern "C" {

const char *mvr_last_error(void) { return g_err.c_str(); }

int mvr_device_count(int *count) {
  if (!count) return fail(MVR_ERR_INVALID, "count is NULL");
  HIP_TRY(hipGetDeviceCount(count));
  return MVR_OK;
}

namespace {

// Codes and residual bytes read from the data sections of vq_code_l.npy and
// residual_code_l.npy (files open at their first data byte) instead of from
// a descriptor's arrays.
struct CodeFiles {
  FILE *codes = nullptr, *res = nullptr;
  ~CodeFiles() {
    if (codes) std::fclose(codes);
    if (res) std::fclose(res);
  }
};

// Bytes per read of the streaming loader (MVR_LOAD_CHUNK_BYTES overrides:
// at least 4 KiB, rounded down to a multiple of 4 so code chunks hold whole
// codes).
size_t load_chunk_bytes() {
  const char *e = std::getenv("MVR_LOAD_CHUNK_BYTES");
  const long long v = e ? std::atoll(e) : 0;
  if (v >= 4096) return (size_t)v & ~(size_t)3;
  return (size_t)64 << 20;
}

// codes in [0, n_c) (a code is a centroid row index in every kernel)
int check_codes(const uint32_t *c, size_t n, uint32_t n_c, uint64_t first) {
  for (size_t i = 0; i < n; ++i)
    if (c[i] >= n_c)
      return fail(MVR_ERR_IO, "vq code %u at %llu >= n_centroid %u", c[i], (unsigned long long)(first + i), n_c);
  return MVR_OK;
}

// Streams `bytes` from f into a new device buffer *dst: chunks are read
// into two pinned buffers in turn and copied with hipMemcpyAsync, so the
// read of one chunk overlaps the copy of the previous.  n_c > 0: the data are
// u32 codes, checked per chunk before they are copied.
int stream_to_device(FILE *f, uint8_t **dst, size_t bytes, uint32_t n_c) {
  struct Res {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    ~Res() {
      if (st) (void)hipStreamSynchronize(st);
      for (int i = 0; i < 2; ++i) {
        if (buf[i]) (void)hipHostFree(buf[i]);
        if (ev[i]) (void)hipEventDestroy(ev[i]);
      }
      if (st) (void)hipStreamDestroy(st);
    }
  } r;
  HIP_TRY(hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(bytes, 256)));
  const size_t chunk = std::min(load_chunk_bytes(), std::max<size_t>(bytes, 4));
  for (int i = 0; i < 2; ++i) {
    HIP_TRY(hipHostMalloc(&r.buf[i], chunk, hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&r.ev[i], hipEventDisableTiming));
  }
  HIP_TRY(hipStreamCreateWithFlags(&r.st, hipStreamNonBlocking));
  uint64_t i = 0;
  for (size_t off = 0; off < bytes; off += chunk, ++i) {
    const int b = (int)(i & 1);
    if (i >= 2) HIP_TRY(hipEventSynchronize(r.ev[b]));  // its previous copy is done
    const size_t n = std::min(chunk, bytes - off);
    if (std::fread(r.buf[b], 1, n, f) != n) return fail(MVR_ERR_IO, "npy data truncated at byte %zu", off);
    if (n_c) {
      const int rc = check_codes(static_cast<const uint32_t *>(r.buf[b]), n / 4, n_c, off / 4);
      if (rc != MVR_OK) return rc;
    }
    HIP_TRY(hipMemcpyAsync(*dst + off, r.buf[b], n, hipMemcpyHostToDevice, r.st));
    HIP_TRY(hipEventRecord(r.ev[b], r.st));
  }
  HIP_TRY(hipStreamSynchronize(r.st));
  return MVR_OK;
}

int index_create(const mvr_index_desc *desc, int device, uint32_t flags, const CodeFiles *cf, mvr_index **out);

}  // namespace

int mvr_index_create(const mvr_index_desc *desc, int device, mvr_index **out) {
  return mvr_index_create_ex(desc, device, 0u, out);
}

int mvr_index_create_ex(const mvr_index_desc *desc, int device, uint32_t flags, mvr_index **out) {
  return index_create(desc, device, flags, nullptr, out);
}

namespace {

int index_create(const mvr_index_desc *desc, int device, uint32_t flags, const CodeFiles *cf, mvr_index **out) {
  if (!desc || !out) return fail(MVR_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (flags & ~MVR_INDEX_RESIDUAL_ON_HOST) return fail(MVR_ERR_INVALID, "unknown index flags 0x%x", flags);
  if (desc->dim != kDim) return fail(MVR_ERR_UNSUPPORTED, "dim %u unsupported (kernels need 128)", desc->dim);
  if (desc->n_bit != 1 && desc->n_bit != 2 && desc->n_bit != 4 && desc->n_bit != 8)
    return fail(MVR_ERR_INVALID, "n_bit %u must be 1, 2, 4 or 8", desc->n_bit);
  if (desc->n_bit != 2 && (flags & MVR_INDEX_RESIDUAL_ON_HOST))
    return fail(MVR_ERR_UNSUPPORTED, "host-resident residuals need 2-bit codes (n_bit %u)", desc->n_bit);
  if (desc->n_doc == 0 || desc->n_centroid == 0)
    return fail(MVR_ERR_INVALID, "empty index (n_doc %u, n_centroid %u)", desc->n_doc, desc->n_centroid);
  if (!desc->doclens || !desc->centroids || (!cf && (!desc->codes || !desc->residual)) || !desc->weights ||
      !desc->ivf_offsets || !desc->ivf_sizes || (!desc->ivf_ids && desc->ivf_n_ele))
    return fail(MVR_ERR_INVALID, "NULL index array");
  int n_dev = 0;
  HIP_TRY(hipGetDeviceCount(&n_dev));
  if (device < 0 || device >= n_dev) return fail(MVR_ERR_INVALID, "device %d out of range (%d)", device, n_dev);
  DeviceGuard g(device);
  const bool od = desc->on_device != 0;

  mvr_index *ix = new mvr_index();
  ix->device = device;
  DevIndex &d = ix->d;
  d.n_doc = desc->n_doc;
  d.n_tok = desc->n_tok;
  d.n_c = desc->n_centroid;
  d.n_ele = desc->ivf_n_ele;
  d.n_bit = desc->n_bit;
  const size_t res_bytes = (size_t)kDim * d.n_bit / 8;  // packed residual bytes per token
  const uint32_t nq = 1u << d.n_bit;
  d.n_seg = (d.n_doc + kSeg - 1) / kSeg;
  {
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n_cu <= 0)
      n_cu = 256;
    d.n_cu = (uint32_t)n_cu;
  }

  auto bail = [&](int code) {
    free_devindex(ix->d);
    if (ix->h_res) (void)hipHostFree(ix->h_res);
    if (ix->h_codes) (void)hipHostFree(ix->h_codes);
    delete ix;
    return code;
  };
#define UP(dst, src, n)                                                                      \
  do {                                                                                       \
    hipError_t _e = upload(&(dst), (src), (n), od);                                          \
    if (_e != hipSuccess)                                                                    \
      return bail(fail(MVR_ERR_DEVICE, "upload of %s failed: %s", #src, hipGetErrorString(_e))); \
  } while (0)

  // host-side metadata: doc lengths / offsets, list sizes, decode table
  ix->h_doclens = fetch(desc->doclens, d.n_doc, od);
  std::vector<uint64_t> docoff(d.n_doc);
  uint64_t acc = 0;
  for (uint32_t i = 0; i < d.n_doc; ++i) {
    docoff[i] = acc;
    acc += ix->h_doclens[i];
    d.max_doclen = std::max(d.max_doclen, ix->h_doclens[i]);
  }
  if (acc != d.n_tok)
    return bail(fail(MVR_ERR_INVALID, "sum of doc lengths %llu != n_tok %llu",
                     (unsigned long long)acc, (unsigned long long)d.n_tok));
  std::vector<uint32_t> sizes = fetch(desc->ivf_sizes, d.n_c, od);
  std::vector<uint64_t> offs = fetch(desc->ivf_offsets, d.n_c, od);
  for (uint32_t c = 0; c < d.n_c; ++c) {
    d.max_list = std::max(d.max_list, sizes[c]);
    if (offs[c] + sizes[c] > d.n_ele)
      return bail(fail(MVR_ERR_INVALID, "IVF list %u exceeds n_ele", c));
  }
  std::vector<float> W((size_t)kDim * nq);
  {
    const size_t nw = desc->weights_per_dim ? (size_t)kDim * nq : nq;
    std::vector<float> w = fetch(desc->weights, nw, od);
    for (uint32_t dd = 0; dd < kDim; ++dd)
      for (uint32_t b = 0; b < nq; ++b) W[dd * nq + b] = w[(desc->weights_per_dim ? dd * nq : 0) + b];
  }

  UP(d.C, desc->centroids, (size_t)d.n_c * kDim);
  {
    // largest centroid norm (host copy: the descriptor may hold device pointers)
    const std::vector<float> Ch = fetch(desc->centroids, (size_t)d.n_c * kDim, od);
    for (uint32_t c = 0; c < d.n_c; ++c) {
      double nn = 0.0;
      for (uint32_t k = 0; k < kDim; ++k) nn += (double)Ch[(size_t)c * kDim + k] * Ch[(size_t)c * kDim + k];
      d.c_norm_max = std::max(d.c_norm_max, (float)std::sqrt(nn));
    }
  }
  if (flags & MVR_INDEX_RESIDUAL_ON_HOST) {
    // pinned, mapped host copies that the per-pass gather kernel reads
    ix->res_host = true;
    const hipMemcpyKind kind = od ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
    hipError_t e0 = hipHostMalloc(reinterpret_cast<void **>(&ix->h_res), std::max<size_t>(d.n_tok * kResBytes, 256),
                                  hipHostMallocMapped);
    if (e0 == hipSuccess)
      e0 = hipHostMalloc(reinterpret_cast<void **>(&ix->h_codes), std::max<size_t>(d.n_tok * 4, 256),
                         hipHostMallocMapped);
    if (cf) {
      // straight from the files into the pinned arrays
      if (e0 == hipSuccess && d.n_tok &&
          (std::fread(ix->h_res, 1, d.n_tok * kResBytes, cf->res) != d.n_tok * kResBytes ||
           std::fread(ix->h_codes, 4, d.n_tok, cf->codes) != d.n_tok))
        return bail(fail(MVR_ERR_IO, "npy data truncated"));
    } else {
      if (e0 == hipSuccess && d.n_tok) e0 = hipMemcpy(ix->h_res, desc->residual, d.n_tok * kResBytes, kind);
      if (e0 == hipSuccess && d.n_tok) e0 = hipMemcpy(ix->h_codes, desc->codes, d.n_tok * 4, kind);
    }
    void *pr = nullptr, *pc = nullptr;
    if (e0 == hipSuccess) e0 = hipHostGetDevicePointer(&pr, ix->h_res, 0);
    if (e0 == hipSuccess) e0 = hipHostGetDevicePointer(&pc, ix->h_codes, 0);
    if (e0 != hipSuccess)
      return bail(fail(MVR_ERR_DEVICE, "host residual copy failed: %s", hipGetErrorString(e0)));
    ix->h_res_dev = static_cast<const uint8_t *>(pr);
    ix->h_codes_dev = static_cast<const uint32_t *>(pc);
    if (!od) {
      // codes in range (a device descriptor's codes are trusted like a device-resident index's)
      for (uint64_t i = 0; i < d.n_tok; ++i)
        if (ix->h_codes[i] >= d.n_c) return bail(fail(MVR_ERR_INVALID, "code %u out of range", ix->h_codes[i]));
    }
  } else if (cf) {
    int rc = stream_to_device(cf->codes, reinterpret_cast<uint8_t **>(&d.codes), d.n_tok * 4, d.n_c);
    if (rc == MVR_OK) rc = stream_to_device(cf->res, &d.residual, d.n_tok * res_bytes, 0);
    if (rc != MVR_OK) return bail(rc);
  } else {
    UP(d.codes, desc->codes, d.n_tok);
    UP(d.residual, desc->residual, d.n_tok * res_bytes);
  }
  UP(d.doclens, desc->doclens, d.n_doc);
  UP(d.ivf_off, desc->ivf_offsets, d.n_c);
  UP(d.ivf_size, desc->ivf_sizes, d.n_c);
  UP(d.ivf_ids, desc->ivf_ids, d.n_ele);
#undef UP
  hipError_t e = upload(&d.docoff, docoff.data(), docoff.size(), false);
  if (e == hipSuccess) e = upload(&d.W, W.data(), W.size(), false);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void **>(&d.C_bf16), (size_t)d.n_c * kDim * sizeof(uint16_t));
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void **>(&d.C_bf16m), (size_t)d.n_c * kDim * sizeof(uint16_t));
  {
    float wmax = 0.0f;
    for (float v : W) wmax = std::max(wmax, std::fabs(v));
    d.fp8_res_scale = wmax / 448.0f;
  }
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d.C_fp8), (size_t)d.n_c * kDim);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d.C_fp8_scale), (size_t)d.n_c * sizeof(float));
  // the FP8 residual table decodes 2-bit codes (other widths: exact FP32 path only)
  if (e == hipSuccess && d.n_bit == 2)
    e = hipMalloc(reinterpret_cast<void **>(&d.lut_fp8), (size_t)kResBytes * 256 * 4);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void **>(&d.seg_off), (size_t)d.n_c * (d.n_seg + 1) * sizeof(uint32_t));
  {
    // codes in range and IVF lists ascending within [0, n_doc): checked on the
    // device for every descriptor (host or device arrays); the codes of a
    // host-resident index were checked on the host above
    unsigned long long *bad = nullptr, bad_h[2] = {0, 0};
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&bad), 16);
    if (e == hipSuccess) {
      DevIndex dv = d;
      if (ix->res_host) dv.codes = nullptr;
      e = launch_validate_index(dv, bad, nullptr);
    }
    if (e == hipSuccess) e = hipMemcpy(bad_h, bad, 16, hipMemcpyDeviceToHost);
    if (bad) (void)hipFree(bad);
    if (e == hipSuccess && bad_h[0])
      return bail(fail(MVR_ERR_INVALID, "%llu codes are >= n_centroid %u", bad_h[0], d.n_c));
    if (e == hipSuccess && bad_h[1])
      return bail(fail(MVR_ERR_INVALID, "%llu IVF entries are >= n_doc %u or not strictly ascending in their list",
                       bad_h[1], d.n_doc));
  }
  if (e == hipSuccess) e = launch_seg_offsets(d, nullptr);
  if (e == hipSuccess) e = launch_convert_centroids(d, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return bail(fail(MVR_ERR_DEVICE, "index setup failed: %s", hipGetErrorString(e)));
  {
    // a quarter of the HBM left free by the index, between 2 and 64 GiB
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      ix->pool_budget = std::min<uint64_t>(kPoolBudgetMax, std::max<uint64_t>(kPoolBudgetMin, free_b / 4));
  }
  {
    // the no-alloc entry point's scratch address on this device (set-up, once per device)
    uint32_t *scr = nullptr;
    const int rc = scratch_on(device, &scr);
    if (rc != MVR_OK) return bail(rc);
  }
  *out = ix;
  return MVR_OK;
}

}  // namespace

int mvr_index_load(const char *index_dir, int device, mvr_index **out) {
  return mvr_index_load_ex(index_dir, device, 0u, out);
}

int mvr_index_load_ex(const char *index_dir, int device, uint32_t flags, mvr_index **out) {
  if (!index_dir || !out) return fail(MVR_ERR_INVALID, "NULL argument");
  const std::string dir(index_dir);
  NpyArray cent, codes, weight, res, lens;
  std::string err;
  // vq_code_l.npy and residual_code_l.npy: headers only here; their data are
  // streamed by the index creation (never held whole in host memory, unless
  // the residuals are to stay on the host)
  CodeFiles cf;
  if (!(err = npy_load(dir + "/centroid_l.npy", cent)).empty()) return fail(MVR_ERR_IO, "%s", err.c_str());
  if (!(err = npy_open(dir + "/vq_code_l.npy", codes, &cf.codes)).empty()) return fail(MVR_ERR_IO, "%s", err.c_str());
  if (!(err = npy_load(dir + "/weight_l.npy", weight)).empty()) return fail(MVR_ERR_IO, "%s", err.c_str());
  if (!(err = npy_open(dir + "/residual_code_l.npy", res, &cf.res)).empty()) return fail(MVR_ERR_IO, "%s", err.c_str());
  if (!(err = npy_load(dir + "/item_n_vec_l.npy", lens)).empty()) return fail(MVR_ERR_IO, "%s", err.c_str());
  if (cent.word_size != 4 || cent.descr[1] != 'f' || cent.shape.size() != 2)
    return fail(MVR_ERR_IO, "centroid_l.npy must be float32 [n_centroid, dim]");
  if (codes.word_size != 4 || codes.shape.size() != 1)
    return fail(MVR_ERR_IO, "vq_code_l.npy must be (u)int32 [n_vec]");
  if (weight.word_size != 4 || weight.descr[1] != 'f' || weight.shape.empty() || weight.shape.size() > 2)
    return fail(MVR_ERR_IO, "weight_l.npy must be float32 [2^b] or [dim, 2^b]");
  if (res.word_size != 1 || res.shape.size() != 1)
    return fail(MVR_ERR_IO, "residual_code_l.npy must be uint8 [n_vec * packed]");
  if (lens.word_size != 4 || lens.shape.size() != 1)
    return fail(MVR_ERR_IO, "item_n_vec_l.npy must be uint32 [n_item]");
  const uint32_t n_c = (uint32_t)cent.shape[0], dim = (uint32_t)cent.shape[1];
  const uint64_t n_tok = codes.shape[0];
  // bits per dim from the decode table's width: weight_l.npy is [2^b] or [dim, 2^b]
  const size_t n_w = weight.shape.back();
  uint32_t n_bit = 0;
  for (uint32_t b : {1u, 2u, 4u, 8u})
    if (n_w == (1u << b)) n_bit = b;
  if (n_bit == 0) return fail(MVR_ERR_IO, "weight_l.npy holds %zu values per dim (expected 2, 4, 16 or 256)", n_w);
  if (weight.shape.size() == 2 && weight.shape[0] != dim)
    return fail(MVR_ERR_IO, "weight_l.npy has %zu rows, expected dim %u", weight.shape[0], dim);
  if (res.shape[0] != n_tok * dim * n_bit / 8)
    return fail(MVR_ERR_IO, "residual_code_l.npy has %zu bytes, expected %llu (%u-bit codes)", res.shape[0],
                (unsigned long long)(n_tok * dim * n_bit / 8), n_bit);

  // ivf.index: u32 n_centroid, u64 n_ele, u64[n_c] offsets, u32[n_c] sizes,
  // u32[n_ele] ids (src/include/struct/IVFIndex.hpp:168-183)
  std::ifstream in(dir + "/ivf.index", std::ios::binary);
  if (!in) return fail(MVR_ERR_IO, "cannot open %s/ivf.index", index_dir);
  uint32_t ivf_nc = 0;
  uint64_t n_ele = 0;
  in.read(reinterpret_cast<char *>(&ivf_nc), 4);
  in.read(reinterpret_cast<char *>(&n_ele), 8);
  if (!in || ivf_nc != n_c) return fail(MVR_ERR_IO, "ivf.index header mismatch (n_centroid %u vs %u)", ivf_nc, n_c);
  std::vector<uint64_t> off(n_c);
  std::vector<uint32_t> size(n_c), ids(n_ele);
  in.read(reinterpret_cast<char *>(off.data()), n_c * 8);
  in.read(reinterpret_cast<char *>(size.data()), n_c * 4);
  in.read(reinterpret_cast<char *>(ids.data()), n_ele * 4);
  if (!in) return fail(MVR_ERR_IO, "ivf.index truncated");

  mvr_index_desc desc{};
  desc.n_doc = (uint32_t)lens.shape[0];
  desc.n_tok = n_tok;
  desc.n_centroid = n_c;
  desc.dim = dim;
  desc.n_bit = n_bit;
  desc.doclens = lens.as<uint32_t>();
  desc.centroids = cent.as<float>();
  desc.codes = nullptr;  // from cf
  desc.residual = nullptr;
  desc.weights = weight.as<float>();
  desc.weights_per_dim = weight.shape.size() == 2 ? 1u : 0u;
  desc.ivf_offsets = off.data();
  desc.ivf_sizes = size.data();
  desc.ivf_ids = ids.data();
  desc.ivf_n_ele = n_ele;
  desc.on_device = 0;
  if (flags & ~MVR_INDEX_RESIDUAL_ON_HOST) return fail(MVR_ERR_INVALID, "unknown index flags 0x%x", flags);
  return index_create(&desc, device, flags, &cf, out);
}

void mvr_index_free(mvr_index *ix) {
  if (!ix) return;
  DeviceGuard g(ix->device);
  (void)hipDeviceSynchronize();
  ix->ws.release();
  free_devindex(ix->d);
  if (ix->h_res) (void)hipHostFree(ix->h_res);
  if (ix->h_codes) (void)hipHostFree(ix->h_codes);
  delete ix;
}

namespace {

// The compact index of one pass over a host-resident-residual index.
struct HostPass {
  DevIndex d;                     // the index with codes/residual/doclens/docoff of the pass's docs
  const uint32_t *ids = nullptr;  // [nb][R] compact ids of the selected docs
  const uint32_t *docs = nullptr; // [n] compact id -> doc id (ascending)
  uint32_t n = 0;
};

// Host-resident residuals (the reference's RerankPP::rerank / transmitSQCode,
// src/include/alg/igp_gpu_pp/RerankPP.hpp:289-361, gathers the candidates'
// residual codes on the host for every query and copies them over PCIe).  Here
// the device marks the docs selected by the pass's queries in a bitmap and
// ranks them (compact ids follow doc ids, so a ranking by (score, id) is the
// same on compact and original ids); the host reads back only the pass's doc
// and token counts to size the buffers; then a kernel reads those docs' codes
// and residual rows straight from the pinned, mapped host arrays, each doc
// once however many queries selected it.
//
// hs_compact queues the compaction of a pass into `slot` and the read-back of
// its counts on st; hs_gather (once those counts are on the host) queues the
// copy of the docs on gs.
int hs_compact(mvr_index *ix, Workspace &ws, int slot, const uint32_t *sel_ids, const uint32_t *sel_n,
               uint32_t R, uint32_t nb, hipStream_t st) {
  const DevIndex &d = ix->d;
  const uint32_t n_words = (d.n_doc + 31) / 32, nblk = (n_words + 255) / 256;
  const size_t umax = std::min<size_t>(d.n_doc, (size_t)nb * R);
  HIP_TRY(ws.hs_bm.ensure((size_t)n_words * 4));
  HIP_TRY(ws.hs_bn.ensure((size_t)nblk * 4));
  HIP_TRY(ws.hs_bt.ensure((size_t)nblk * 8));
  HIP_TRY(ws.hs_wbase.ensure((size_t)n_words * 4));
  HIP_TRY(ws.hs_docs[slot].ensure(umax * 4));
  HIP_TRY(ws.hs_lens[slot].ensure(umax * 4));
  HIP_TRY(ws.hs_offs[slot].ensure(umax * 8));
  HIP_TRY(ws.hs_cids[slot].ensure((size_t)nb * R * 4));
  HIP_TRY(ws.hs_cnt[slot].ensure(16));
  HostStage hs;
  hs.bm = ws.hs_bm.as<uint32_t>();
  hs.b_n = ws.hs_bn.as<uint32_t>();
  hs.b_t = ws.hs_bt.as<unsigned long long>();
  hs.wbase = ws.hs_wbase.as<uint32_t>();
  hs.docs = ws.hs_docs[slot].as<uint32_t>();
  hs.lens = ws.hs_lens[slot].as<uint32_t>();
  hs.offs = ws.hs_offs[slot].as<uint64_t>();
  hs.cids = ws.hs_cids[slot].as<uint32_t>();
  hs.cnt = ws.hs_cnt[slot].as<unsigned long long>();
  HIP_TRY(launch_hs_compact(sel_ids, sel_n, R, nb, d.doclens, d.n_doc, hs, st));
  HIP_TRY(hipMemcpyAsync(static_cast<char *>(ws.hs_cnt_h.p) + slot * 16, hs.cnt, 16, hipMemcpyDeviceToHost, st));
  return MVR_OK;
}

int hs_gather(mvr_index *ix, Workspace &ws, int slot, uint32_t nb, uint32_t R, hipStream_t gs, HostPass &hp) {
  const DevIndex &d = ix->d;
  const unsigned long long *cnt =
      reinterpret_cast<const unsigned long long *>(static_cast<const char *>(ws.hs_cnt_h.p) + slot * 16);
  const uint32_t U = (uint32_t)cnt[0];
  const uint64_t T = cnt[1];
  if (U > std::min<size_t>(d.n_doc, (size_t)nb * R) || T > d.n_tok)
    return fail(MVR_ERR_DEVICE, "pass compaction counted %u docs / %llu tokens", U, (unsigned long long)T);
  HIP_TRY(ws.hs_codes[slot].ensure(T * 4));
  HIP_TRY(ws.hs_res[slot].ensure(T * kResBytes));
  HostStage hs;
  hs.docs = ws.hs_docs[slot].as<uint32_t>();
  hs.lens = ws.hs_lens[slot].as<uint32_t>();
  hs.offs = ws.hs_offs[slot].as<uint64_t>();
  HIP_TRY(launch_hs_gather(hs, U, d.docoff, ix->h_codes_dev, ix->h_res_dev, ws.hs_codes[slot].as<uint32_t>(),
                           ws.hs_res[slot].as<uint8_t>(), gs));
  hp.d = d;
  hp.d.n_doc = U;
  hp.d.n_tok = T;
  hp.d.doclens = hs.lens;
  hp.d.docoff = hs.offs;
  hp.d.codes = ws.hs_codes[slot].as<uint32_t>();
  hp.d.residual = ws.hs_res[slot].as<uint8_t>();
  hp.ids = ws.hs_cids[slot].as<uint32_t>();
  hp.docs = hs.docs;
  hp.n = U;
  return MVR_OK;
}

}  // namespace

int mvr_index_set_maxsim_order(mvr_index *ix, uint32_t order) {
  if (!ix) return fail(MVR_ERR_INVALID, "NULL index");
  if (order > MVR_MAXSIM_DOC) return fail(MVR_ERR_INVALID, "unknown MaxSim order %u", order);
  ix->maxsim_order = order;
  return MVR_OK;
}

int mvr_index_info(const mvr_index *ix, uint32_t *n_doc, uint64_t *n_tok, uint32_t *n_centroid,
                   uint32_t *max_list, uint32_t *max_doclen) {
  if (!ix) return fail(MVR_ERR_INVALID, "NULL index");
  if (n_doc) *n_doc = ix->d.n_doc;
  if (n_tok) *n_tok = ix->d.n_tok;
  if (n_centroid) *n_centroid = ix->d.n_c;
  if (max_list) *max_list = ix->d.max_list;
  if (max_doclen) *max_doclen = ix->d.max_doclen;
  return MVR_OK;
}

int mvr_search(mvr_index *ix, const float *Q, uint32_t n_q, uint32_t t_max, const uint32_t *q_ntok,
               const uint32_t *q_nprobe, const mvr_search_params *prm, uint32_t *out_ids,
               float *out_scores, int on_device, void *stream, mvr_stage_out *stages,
               mvr_timings *timings) {
  if (!ix || !prm) return fail(MVR_ERR_INVALID, "NULL argument");
  if (n_q == 0) return MVR_OK;
  if (!Q || !out_ids || !out_scores) return fail(MVR_ERR_INVALID, "NULL query or output buffer");
  const DevIndex &d = ix->d;
  const uint32_t R = prm->probe_topk, k = prm->topk;
  if (t_max == 0 || t_max > kMaxQTok) return fail(MVR_ERR_INVALID, "query tokens %u not in [1, 32]", t_max);
  if (R < k) return fail(MVR_ERR_INVALID, "probe_topk %u is smaller than topk %u", R, k);
  if (k == 0) return fail(MVR_ERR_INVALID, "topk must be positive");
  if (R > kMaxSelect) return fail(MVR_ERR_UNSUPPORTED, "probe_topk %u > %u", R, kMaxSelect);
  if (prm->precision > MVR_FP32_MFMA) return fail(MVR_ERR_INVALID, "unknown precision %u", prm->precision);
  if (d.n_bit != 2 && (prm->precision != MVR_FP32 || prm->persistent))
    return fail(MVR_ERR_UNSUPPORTED, "%u-bit residual codes are searched with MVR_FP32 only (the MFMA and "
                "persistent kernels decode 2-bit codes)", d.n_bit);
  const bool fp32_mode = prm->precision == MVR_FP32 || prm->precision == MVR_FP32_MFMA;
  // FP32 centroid top-P: BF16 MFMA tile maxima + exact FP32 rescan (same
  // result as the all-FP32 scan); MVR_CT_SCAN=1 forces the all-FP32 scan
  static const bool force_scan = [] {
    const char *e = std::getenv("MVR_CT_SCAN");
    return e && e[0] == '1';
  }();
  // BF16/FP8: the best `rescore` docs by the low-precision MaxSim are scored
  // again in FP32 (on MFMA) and the final top-k comes from those FP32 scores
  const uint32_t R2 = (!fp32_mode && prm->rescore) ? std::min(prm->rescore, R) : 0;
  if (R2 && R2 < k) return fail(MVR_ERR_INVALID, "rescore %u is smaller than topk %u", R2, k);
  if (stages && on_device) return fail(MVR_ERR_INVALID, "stage outputs need host buffers");
  uint32_t P = prm->nprobe;
  std::vector<uint32_t> np(n_q), nt(n_q);
  for (uint32_t q = 0; q < n_q; ++q) {
    np[q] = q_nprobe ? q_nprobe[q] : prm->nprobe;
    nt[q] = q_ntok ? q_ntok[q] : t_max;
    if (np[q] == 0) return fail(MVR_ERR_INVALID, "nprobe must be positive");
    if (np[q] > kMaxProbe) return fail(MVR_ERR_UNSUPPORTED, "nprobe %u > %u", np[q], kMaxProbe);
    if (np[q] > d.n_c) return fail(MVR_ERR_INVALID, "nprobe %u is larger than n_centroid %u", np[q], d.n_c);
    if (nt[q] == 0 || nt[q] > t_max) return fail(MVR_ERR_INVALID, "query %u has %u tokens (t_max %u)", q, nt[q], t_max);
    P = std::max(P, np[q]);
  }
  const bool learned = prm->learned != 0;
  if (learned) {
    if (!ix->probe) return fail(MVR_ERR_INVALID, "learned probing needs a model (mvr_index_set_probe)");
    if (q_nprobe) return fail(MVR_ERR_INVALID, "learned probing chooses nprobe; pass q_nprobe = NULL");
    // the model reads the top-32 scores of every token and may pick up to 32 probes
    P = kProbeFeat;
    for (uint32_t q = 0; q < n_q; ++q) np[q] = kProbeFeat;
  }
  P = std::min(P, d.n_c);
  if (learned && P < kProbeFeat) return fail(MVR_ERR_INVALID, "learned probing needs >= 32 centroids");
  const bool fast_ct = !force_scan || P > 64;  // the all-FP32 scan keeps 64 per row
  // queries per device pass; the candidate pool holds each query's keys (at
  // most the docs its t_max x P probed lists hold), so a pass is also capped
  // by the pool budget
  const uint64_t keys_per_q = candidate_keys_per_query(d, t_max, P);
  const uint64_t pool_per_q = keys_per_q * 8;
  const uint32_t B = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(prm->batch ? prm->batch : 1024, ix->pool_budget / pool_per_q));
  DeviceGuard g(ix->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Workspace &ws = ix->ws;

  // ---- per-call metadata (host) : q_ntok, q_np, row_map ----
  bool uniform = true;
  for (uint32_t q = 0; q < n_q; ++q) uniform &= (nt[q] == t_max);
  std::vector<uint32_t> rmap;
  std::vector<uint32_t> rows_begin;  // per batch: first row-map entry
  size_t max_rows = 0;
  for (uint32_t b0 = 0; b0 < n_q; b0 += B) {
    const uint32_t nb = std::min(B, n_q - b0);
    rows_begin.push_back((uint32_t)rmap.size());
    size_t rows = 0;
    for (uint32_t i = 0; i < nb; ++i) {
      const uint32_t q = b0 + i;
      if (!uniform)
        for (uint32_t t = 0; t < nt[q]; ++t) rmap.push_back(i * t_max + t);
      rows += uniform ? t_max : nt[q];
    }
    max_rows = std::max(max_rows, rows);
  }
  const size_t meta_u32 = 2 * (size_t)n_q + rmap.size();
  const size_t meta_bytes = meta_u32 * 4;
  // wait until the previous call's staging copy finished before reuse
  if (ws.staging_done) HIP_TRY(hipEventSynchronize(ws.staging_done));
  else HIP_TRY(hipEventCreateWithFlags(&ws.staging_done, hipEventDisableTiming));
  HIP_TRY(ws.staging.ensure(meta_bytes));
  uint8_t *hs = reinterpret_cast<uint8_t *>(ws.staging.p);
  std::memcpy(hs, nt.data(), (size_t)n_q * 4);
  std::memcpy(hs + (size_t)n_q * 4, np.data(), (size_t)n_q * 4);
  if (!rmap.empty()) std::memcpy(hs + (size_t)n_q * 8, rmap.data(), rmap.size() * 4);
  HIP_TRY(ws.meta.ensure(meta_bytes));
  HIP_TRY(hipMemcpyAsync(ws.meta.p, hs, meta_bytes, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(ws.staging_done, s));
  const uint32_t *d_nt = ws.meta.as<uint32_t>();
  const uint32_t *d_np = d_nt + n_q;
  const uint32_t *d_rmap = d_np + n_q;

  // ---- workspace ----
  const uint32_t n_rows_max = (uint32_t)max_rows;
  const uint32_t rows_blocks = (n_rows_max + 255) / 256;
  uint32_t n_split = std::max(1u, std::min(16u, (1024u + rows_blocks - 1) / std::max(1u, rows_blocks)));
  n_split = std::min(n_split, std::max(1u, d.n_c / 256));
  const size_t Bc = std::min<size_t>(B, n_q);
  HIP_TRY(ws.part_ids.ensure((size_t)n_split * n_rows_max * P * 4));
  HIP_TRY(ws.part_s.ensure((size_t)n_split * n_rows_max * P * 4));
  HIP_TRY(ws.top_ids.ensure(Bc * t_max * P * 4));
  HIP_TRY(ws.top_s.ensure(Bc * t_max * P * 4));
  HIP_TRY(ws.cand_cnt.ensure(Bc * 4));  // per-query fill counters of the pool
  HIP_TRY(ws.cand_keys.ensure(Bc * pool_per_q));
  if (!fp32_mode || fast_ct) HIP_TRY(ws.bmax.ensure(Bc * t_max * ((d.n_c + 31) / 32) * 4));
  if (learned) {
    HIP_TRY(ws.feat.ensure(Bc * 1024 * 4));
    HIP_TRY(ws.mlp_out.ensure(Bc * 6 * 4));
    HIP_TRY(ws.dyn_np.ensure(Bc * 4));
    HIP_TRY(ws.dyn_r.ensure(Bc * 4));
  }
  if (R2) {
    HIP_TRY(ws.res_ids.ensure(Bc * R2 * 4));
    HIP_TRY(ws.res_s.ensure(Bc * R2 * 4));
    HIP_TRY(ws.res_n.ensure(Bc * 4));
  }
  // persistent work-queue MaxSim: BF16 / FP8 / FP32-MFMA, top-k (and rescore) <= 256
  if (prm->persistent && prm->precision == MVR_FP32)
    return fail(MVR_ERR_UNSUPPORTED, "the persistent kernel has no exact-FP32 variant (use MVR_FP32_MFMA)");
  if (prm->persistent && (k > kPersistTopkMax || R2 > kPersistTopkMax))
    return fail(MVR_ERR_UNSUPPORTED, "the persistent kernel selects at most %u (topk %u, rescore %u)",
                kPersistTopkMax, k, R2);
  const bool persistent = prm->persistent != 0;
  // BF16 and FP8 MaxSim doc-major (each selected doc's tiles loaded once per pass for
  // all of its queries) when a pass holds at least a quarter as many (query,
  // doc) pairs as the index has docs; query-major below that, where the
  // grouping's passes over every doc cost more than the sharing saves.
  // (Measured at 1M docs: doc-major is faster from 512 queries x 1,000
  // candidates, 12.7 vs 13.4 ms, to 4,096 x 1,000, 8.4 vs 12.8 ms.)
  // In a persistent search the doc-major kernel is the first phase's work
  // queue (of docs), followed by the per-query top-R2 (or top-k); with the
  // automatic order the autotuner times it against the per-query persistent
  // kernel and keeps the faster (PersistCfg::dm).
  const bool dm_ok = (prm->precision == MVR_BF16 || prm->precision == MVR_FP8) && (uint64_t)Bc * R <= UINT32_MAX;
  const bool doc_major = dm_ok && !persistent &&
                         (ix->maxsim_order == MVR_MAXSIM_DOC ||
                          (ix->maxsim_order == MVR_MAXSIM_AUTO && (uint64_t)Bc * R * 4 >= d.n_doc));
  const bool p_dm_fixed = persistent && dm_ok && ix->maxsim_order == MVR_MAXSIM_DOC;
  const bool p_dm_tuned = persistent && dm_ok && ix->maxsim_order == MVR_MAXSIM_AUTO;
  // the persistent first phase runs doc-major (read when the pass runs: the
  // autotuner changes the setting between its timed passes)
  auto persist_dm = [&]() { return p_dm_fixed || (p_dm_tuned && ix->tune[prm->precision].dm); };
  DocMajorBufs dmb;
  if (doc_major || p_dm_fixed || p_dm_tuned) {
    const size_t nblk = (d.n_doc + kDmDocsPerBlock - 1) / kDmDocsPerBlock;
    HIP_TRY(ws.dm_cnt.ensure((size_t)d.n_doc * 4));
    HIP_TRY(ws.dm_work.ensure((size_t)d.n_doc * 8));
    HIP_TRY(ws.dm_plist.ensure(Bc * R * 4));
    HIP_TRY(ws.dm_bp.ensure(nblk * 4));
    HIP_TRY(ws.dm_bd.ensure(nblk * 4));
    HIP_TRY(ws.dm_tot.ensure(16));
    HIP_TRY(ws.dm_Qm.ensure(Bc * 512 * 16));
    dmb.cnt = ws.dm_cnt.as<uint32_t>();
    dmb.work = ws.dm_work.as<uint2>();
    dmb.plist = ws.dm_plist.as<uint32_t>();
    dmb.b_p = ws.dm_bp.as<uint32_t>();
    dmb.b_d = ws.dm_bd.as<uint32_t>();
    dmb.tot = ws.dm_tot.as<uint32_t>();
    dmb.Qm = ws.dm_Qm.as<uint4>();
  }
  // graph replay: not with learned probing (the model's launches) or stage outputs
  // nor with host-resident residuals (each pass synchronises to stage its docs)
  const bool use_graph = prm->graph && !learned && !stages && !ix->res_host;
  const uint32_t n_pass = (n_q + B - 1) / B;
  // host-resident residuals: two passes in flight, the staging copy of one (on
  // a second stream) overlapping the device work of the other; the selection
  // outputs and the query buffer then have one slot per pass in flight
  // persistent settings still to tune: the first phase's (unless it is fixed
  // doc-major) and, with a rescore, the FP32 phase's
  const bool tune_p1 = persistent && !p_dm_fixed && !ix->tuned[prm->precision];
  const bool tune_p2 = persistent && R2 && prm->precision != MVR_FP32_MFMA && !ix->tuned[MVR_FP32_MFMA];
  const bool pipelined = ix->res_host && !stages && n_pass > 1 && !tune_p1 && !tune_p2;
  const size_t n_slot = pipelined ? 2 : 1;
  HIP_TRY(ws.cand_tot.ensure(n_slot * Bc * 4));
  HIP_TRY(ws.sel_ids.ensure(n_slot * Bc * R * 4));
  HIP_TRY(ws.sel_F.ensure(n_slot * Bc * R * 4));
  HIP_TRY(ws.sel_n.ensure(n_slot * Bc * 4));
  HIP_TRY(ws.sel_score.ensure(n_slot * Bc * R * 4));
  if (ix->res_host) HIP_TRY(ws.hs_cnt_h.ensure(32));
  if (pipelined) {
    if (!ws.hs_stream) HIP_TRY(hipStreamCreateWithFlags(&ws.hs_stream, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
      if (!ws.hs_compacted[i]) HIP_TRY(hipEventCreateWithFlags(&ws.hs_compacted[i], hipEventDisableTiming));
      if (!ws.hs_staged[i]) HIP_TRY(hipEventCreateWithFlags(&ws.hs_staged[i], hipEventDisableTiming));
    }
  }
  // per-slot views of the selection outputs
  struct Sel {
    uint32_t *ids, *n, *tot;
    float *F, *score;
  };
  auto sel = [&](size_t slot) {
    return Sel{ws.sel_ids.as<uint32_t>() + slot * Bc * R, ws.sel_n.as<uint32_t>() + slot * Bc,
               ws.cand_tot.as<uint32_t>() + slot * Bc, ws.sel_F.as<float>() + slot * Bc * R,
               ws.sel_score.as<float>() + slot * Bc * R};
  };
  if (!on_device || use_graph) {
    HIP_TRY(ws.Q.ensure(n_slot * Bc * t_max * kDim * 4));
    HIP_TRY(ws.out_ids.ensure(Bc * k * 4));
    HIP_TRY(ws.out_s.ensure(Bc * k * 4));
  }
  if (persistent) HIP_TRY(ws.queue.ensure(4));
  if (use_graph) {
    HIP_TRY(ws.pass_nt.ensure(Bc * 4));
    HIP_TRY(ws.pass_np.ensure(Bc * 4));
    HIP_TRY(ws.pass_rm.ensure(Bc * t_max * 4));
  }

  // Stage timings: events per device pass, read once at the end of the call
  // so that timing adds no host synchronisation between passes.  0..6: stage
  // boundaries; host-resident residuals: 7 the staged docs on the device, 8
  // the MaxSim stage starting.
  constexpr int kNev = 9;
  std::vector<hipEvent_t> ev;
  if (timings) {
    ev.resize((size_t)n_pass * kNev);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    *timings = mvr_timings{};
    HIP_TRY(ws.counters.ensure(24));
    HIP_TRY(launch_fill_u32(ws.counters.as<uint32_t>(), 0u, 6, s));
  }
  auto rec = [&](uint32_t pass, int i, hipStream_t on = nullptr) -> hipError_t {
    return timings ? hipEventRecord(ev[(size_t)pass * kNev + i], on ? on : s) : hipSuccess;
  };

  // A device pass over nb queries whose inputs are in place, in two halves.
  // `timed` records the stage events (not while a graph is being captured).
  // Front: centroid scoring, probing, filter and selection into `slot`.
  auto run_front = [&](hipStream_t st, const float *Qb, const uint32_t *rm, uint32_t n_rows,
                       const uint32_t *bnt, const uint32_t *bnp, uint32_t nb, int bi, bool timed,
                       size_t slot) -> int {
    auto mark = [&](int i) -> hipError_t { return timed ? rec(bi, i) : hipSuccess; };
    const Sel sl = sel(slot);
    HIP_TRY(mark(0));
    if (fp32_mode && !fast_ct)
      HIP_TRY(launch_centroid_topk(Qb, rm, n_rows, nb * t_max, d, P, MVR_FP32, n_split,
                                   ws.part_ids.as<uint32_t>(), ws.part_s.as<float>(),
                                   ws.top_ids.as<uint32_t>(), ws.top_s.as<float>(), nullptr, st));
    else if (prm->precision == MVR_FP8)
      HIP_TRY(launch_centroid_topk_fp8(Qb, nb, t_max, rm, n_rows, d, P, ws.bmax.as<float>(),
                                       ws.top_ids.as<uint32_t>(), ws.top_s.as<float>(), st));
    else
      HIP_TRY(launch_centroid_topk_bf16(Qb, nb, t_max, rm, n_rows, d, P, ws.bmax.as<float>(),
                                        ws.top_ids.as<uint32_t>(), ws.top_s.as<float>(), fp32_mode, st));
    HIP_TRY(mark(1));
    const uint32_t *q_r = nullptr;
    if (learned) {
      HIP_TRY(launch_probe_features(ws.top_s.as<float>(), t_max, P, bnt, nb, ws.feat.as<float>(), st));
      if (mvr_mlp_forward(ix->probe, ws.feat.as<float>(), nb, ws.mlp_out.as<float>(), st) != MVR_OK)
        return fail(MVR_ERR_DEVICE, "probe model: %s", mvr_mlp_last_error());
      HIP_TRY(launch_probe_decide(ws.mlp_out.as<float>(), nb, ix->probe_conf, ix->probe_fallback,
                                  std::max(ix->probe_rmin, k), R, ws.dyn_np.as<uint32_t>(),
                                  ws.dyn_r.as<uint32_t>(), st));
      bnp = ws.dyn_np.as<uint32_t>();
      q_r = ws.dyn_r.as<uint32_t>();
    }
    HIP_TRY(launch_filter(ws.top_ids.as<uint32_t>(), ws.top_s.as<float>(), t_max, P, bnt, bnp, nb,
                          d, prm->clip_zero ? 0 : 1, ws.cand_cnt.as<uint32_t>(), keys_per_q,
                          ws.cand_keys.as<uint64_t>(), st));
    HIP_TRY(mark(2));
    HIP_TRY(launch_select(ws.cand_keys.as<uint64_t>(), ws.cand_cnt.as<uint32_t>(), keys_per_q, q_r, nb, R,
                          sl.ids, sl.F, sl.n, sl.tot, st));
    HIP_TRY(mark(3));
    return MVR_OK;
  };

  // Back: MaxSim of the selected docs (with host-resident residuals, of the
  // pass's staged docs under their compact ids, mapped back at the end),
  // the FP32 rescore and the final top-k.
  auto run_back = [&](hipStream_t st, const float *Qb, const uint32_t *bnt, uint32_t nb, uint32_t *oid, float *os,
                      int bi, bool timed, size_t slot, const HostPass *hp) -> int {
    auto mark = [&](int i) -> hipError_t { return timed ? rec(bi, i) : hipSuccess; };
    const Sel sl = sel(slot);
    const uint32_t *sid = hp ? hp->ids : sl.ids;
    const DevIndex &dx = hp ? hp->d : d;
    const int clip = (int)prm->clip_zero;
    auto rerank_dm = [&]() -> hipError_t {
      return prm->precision == MVR_FP8
                 ? launch_rerank_fp8_dm(Qb, t_max, bnt, nb, sid, sl.n, R, dx, clip, dmb, sl.score, st)
                 : launch_rerank_bf16_dm(Qb, t_max, bnt, nb, sid, sl.n, R, dx, clip, dmb, sl.score, st);
    };
    if (persistent && persist_dm()) {
      // phase 1: doc-major BF16/FP8 MaxSim of all candidates (a work queue of
      // docs), then each query's top-R2 -- or its top-k without a rescore
      HIP_TRY(rerank_dm());
      if (R2) {
        HIP_TRY(launch_topk(sid, sl.score, sl.n, R, nb, R2, ws.res_ids.as<uint32_t>(), ws.res_s.as<float>(),
                            ws.res_n.as<uint32_t>(), st));
        HIP_TRY(mark(4));
        // phase 2: FP32 MaxSim of those, fused top-k
        const PersistCfg &c2 = ix->tuned[MVR_FP32_MFMA] ? ix->tune[MVR_FP32_MFMA] : ix->tune[prm->precision];
        HIP_TRY(launch_persist(Qb, t_max, bnt, ws.res_ids.as<uint32_t>(), ws.res_n.as<uint32_t>(), R2, nb, dx,
                               MVR_FP32_MFMA, clip, c2, ws.queue.as<uint32_t>(), ws.res_s.as<float>(), k, oid,
                               os, nullptr, k, st));
        HIP_TRY(mark(5));
      } else {
        HIP_TRY(launch_topk(sid, sl.score, sl.n, R, nb, k, oid, os, nullptr, st));
        HIP_TRY(mark(4));
        HIP_TRY(mark(5));
      }
      if (hp) HIP_TRY(launch_map_ids(oid, nb * k, hp->docs, hp->n, st));
      HIP_TRY(mark(6));
      return MVR_OK;
    }
    if (persistent) {
      const PersistCfg &c1 = ix->tune[prm->precision];
      uint32_t *qu = ws.queue.as<uint32_t>();
      if (R2) {
        // phase 1: low-precision MaxSim of all candidates, fused top-R2;
        // phase 2: FP32 MaxSim of those, fused top-k
        HIP_TRY(launch_persist(Qb, t_max, bnt, sid, sl.n, R, nb, dx,
                               (int)prm->precision, clip, c1, qu, sl.score, R2,
                               ws.res_ids.as<uint32_t>(), ws.res_s.as<float>(), ws.res_n.as<uint32_t>(), R2,
                               st));
        HIP_TRY(mark(4));
        const PersistCfg &c2 = ix->tuned[MVR_FP32_MFMA] ? ix->tune[MVR_FP32_MFMA] : c1;
        HIP_TRY(launch_persist(Qb, t_max, bnt, ws.res_ids.as<uint32_t>(), ws.res_n.as<uint32_t>(), R2, nb, dx,
                               MVR_FP32_MFMA, clip, c2, qu, ws.res_s.as<float>(), k, oid, os, nullptr,
                               k, st));
        HIP_TRY(mark(5));
      } else {
        HIP_TRY(launch_persist(Qb, t_max, bnt, sid, sl.n, R, nb, dx,
                               (int)prm->precision, clip, c1, qu, sl.score, k, oid,
                               os, nullptr, k, st));
        HIP_TRY(mark(4));
        HIP_TRY(mark(5));
      }
      if (hp) HIP_TRY(launch_map_ids(oid, nb * k, hp->docs, hp->n, st));
      HIP_TRY(mark(6));
      return MVR_OK;
    }
    if (doc_major)
      HIP_TRY(rerank_dm());
    else
      HIP_TRY(launch_rerank(Qb, t_max, bnt, nb, sid, sl.n, R, dx, prm->precision, clip, sl.score, st));
    HIP_TRY(mark(4));
    if (R2) {
      // best R2 by the low-precision score, FP32 MaxSim of those, top-k of the FP32 scores
      HIP_TRY(launch_topk(sid, sl.score, sl.n,
                          R, nb, R2, ws.res_ids.as<uint32_t>(), ws.res_s.as<float>(),
                          ws.res_n.as<uint32_t>(), st));
      HIP_TRY(launch_rerank(Qb, t_max, bnt, nb, ws.res_ids.as<uint32_t>(), ws.res_n.as<uint32_t>(), R2,
                            dx, MVR_FP32_MFMA, clip, ws.res_s.as<float>(), st));
    }
    HIP_TRY(mark(5));
    if (R2)
      HIP_TRY(launch_topk(ws.res_ids.as<uint32_t>(), ws.res_s.as<float>(), ws.res_n.as<uint32_t>(), R2,
                          nb, k, oid, os, nullptr, st));
    else
      HIP_TRY(launch_topk(sid, sl.score, sl.n,
                          R, nb, k, oid, os, nullptr, st));
    if (hp) HIP_TRY(launch_map_ids(oid, nb * k, hp->docs, hp->n, st));
    HIP_TRY(mark(6));
    return MVR_OK;
  };

  auto note_staged = [&](const HostPass &hp) {
    if (timings) {
      timings->staged_docs += hp.n;
      timings->staged_tokens += hp.d.n_tok;
    }
  };

  // One whole pass on st (slot 0).
  auto run_pass = [&](hipStream_t st, const float *Qb, const uint32_t *rm, uint32_t n_rows,
                      const uint32_t *bnt, const uint32_t *bnp, uint32_t nb, uint32_t *oid, float *os,
                      int bi, bool timed) -> int {
    int rc = run_front(st, Qb, rm, n_rows, bnt, bnp, nb, bi, timed, 0);
    if (rc != MVR_OK) return rc;
    if (!ix->res_host) return run_back(st, Qb, bnt, nb, oid, os, bi, timed, 0, nullptr);
    const Sel sl = sel(0);
    HostPass hp;
    if ((rc = hs_compact(ix, ws, 0, sl.ids, sl.n, R, nb, st)) != MVR_OK) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    if ((rc = hs_gather(ix, ws, 0, nb, R, st, hp)) != MVR_OK) return rc;
    if (timed) {
      HIP_TRY(rec(bi, 7));
      HIP_TRY(rec(bi, 8));
      note_staged(hp);
    }
    return run_back(st, Qb, bnt, nb, oid, os, bi, timed, 0, &hp);
  };

  // Graphs are only valid for the buffers they were captured with.
  if (use_graph) {
    std::vector<uint64_t> sig;
    for (DBuf *b : ws.all()) sig.push_back(reinterpret_cast<uint64_t>(b->p));
    if (sig != ws.graph_sig) {
      ws.drop_graphs();
      ws.graph_sig = sig;
    }
    if (!ws.cap_stream) HIP_TRY(hipStreamCreateWithFlags(&ws.cap_stream, hipStreamNonBlocking));
  }
  std::vector<char> pass_graphed(n_pass, 0);

  // Pipelined passes (host-resident residuals): the front of pass b+1 and
  // its compaction are queued on s before the back of pass b, and pass b+1's
  // staging copy runs on hs_stream while s works on pass b's MaxSim.
  struct PassIn {
    uint32_t b0 = 0, nb = 0, n_rows = 0;
    const float *Qb = nullptr;
    const uint32_t *rm = nullptr, *bnt = nullptr, *bnp = nullptr;
  };
  if (pipelined) {
    hipStream_t gs = ws.hs_stream;
    HostPass hps[2];
    PassIn pin[2];
    auto front_and_compact = [&](uint32_t bi) -> int {
      const size_t slot = bi & 1u;
      PassIn &pi = pin[slot];
      pi.b0 = bi * B;
      pi.nb = std::min(B, n_q - pi.b0);
      if (on_device) {
        pi.Qb = Q + (size_t)pi.b0 * t_max * kDim;
      } else {
        float *qs = ws.Q.as<float>() + slot * Bc * t_max * kDim;
        HIP_TRY(hipMemcpyAsync(qs, Q + (size_t)pi.b0 * t_max * kDim, (size_t)pi.nb * t_max * kDim * 4,
                               hipMemcpyHostToDevice, s));
        pi.Qb = qs;
      }
      pi.rm = uniform ? nullptr : d_rmap + rows_begin[bi];
      pi.n_rows = 0;
      if (uniform) pi.n_rows = pi.nb * t_max;
      else for (uint32_t i = 0; i < pi.nb; ++i) pi.n_rows += nt[pi.b0 + i];
      pi.bnt = d_nt + pi.b0;
      pi.bnp = d_np + pi.b0;
      int rc = run_front(s, pi.Qb, pi.rm, pi.n_rows, pi.bnt, pi.bnp, pi.nb, (int)bi, true, slot);
      if (rc != MVR_OK) return rc;
      const Sel sl = sel(slot);
      if ((rc = hs_compact(ix, ws, (int)slot, sl.ids, sl.n, R, pi.nb, s)) != MVR_OK) return rc;
      HIP_TRY(hipEventRecord(ws.hs_compacted[slot], s));
      return MVR_OK;
    };
    auto stage = [&](uint32_t bi) -> int {
      const size_t slot = bi & 1u;
      HIP_TRY(hipEventSynchronize(ws.hs_compacted[slot]));
      const int rc = hs_gather(ix, ws, (int)slot, pin[slot].nb, R, gs, hps[slot]);
      if (rc != MVR_OK) return rc;
      HIP_TRY(rec(bi, 7, gs));
      HIP_TRY(hipEventRecord(ws.hs_staged[slot], gs));
      note_staged(hps[slot]);
      return MVR_OK;
    };
    int rc = front_and_compact(0);
    if (rc == MVR_OK) rc = stage(0);
    for (uint32_t bi = 0; rc == MVR_OK && bi < n_pass; ++bi) {
      const size_t slot = bi & 1u;
      if (bi + 1 < n_pass && (rc = front_and_compact(bi + 1)) != MVR_OK) break;
      HIP_TRY(hipStreamWaitEvent(s, ws.hs_staged[slot], 0));
      HIP_TRY(rec(bi, 8));
      const PassIn &pi = pin[slot];
      uint32_t *oid = on_device ? out_ids + (size_t)pi.b0 * k : ws.out_ids.as<uint32_t>();
      float *os = on_device ? out_scores + (size_t)pi.b0 * k : ws.out_s.as<float>();
      if ((rc = run_back(s, pi.Qb, pi.bnt, pi.nb, oid, os, (int)bi, true, slot, &hps[slot])) != MVR_OK) break;
      if (!on_device) {
        HIP_TRY(hipMemcpyAsync(out_ids + (size_t)pi.b0 * k, oid, (size_t)pi.nb * k * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(out_scores + (size_t)pi.b0 * k, os, (size_t)pi.nb * k * 4, hipMemcpyDeviceToHost, s));
      }
      if (timings) {
        const Sel sl = sel(slot);
        HIP_TRY(launch_count_work(sl.ids, sl.n, R, pi.nb, d, sl.tot, pi.bnt, ws.counters.as<unsigned long long>(), s));
      }
      if (bi + 1 < n_pass) rc = stage(bi + 1);
    }
    if (rc != MVR_OK) return rc;
  }

  if (!pipelined) {
    int bi = 0;
    for (uint32_t b0 = 0; b0 < n_q; b0 += B, ++bi) {
      const uint32_t nb = std::min(B, n_q - b0);
      const float *Qb;
      uint32_t *oid;
      float *os;
      if (on_device && !use_graph) {
        Qb = Q + (size_t)b0 * t_max * kDim;
        oid = out_ids + (size_t)b0 * k;
        os = out_scores + (size_t)b0 * k;
      } else {
        HIP_TRY(hipMemcpyAsync(ws.Q.p, Q + (size_t)b0 * t_max * kDim, (size_t)nb * t_max * kDim * 4,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        Qb = ws.Q.as<float>();
        oid = ws.out_ids.as<uint32_t>();
        os = ws.out_s.as<float>();
      }
      const uint32_t *rm = uniform ? nullptr : d_rmap + rows_begin[bi];
      uint32_t n_rows = 0;
      if (uniform) n_rows = nb * t_max;
      else for (uint32_t i = 0; i < nb; ++i) n_rows += nt[b0 + i];
      const uint32_t *bnt = d_nt + b0;
      const uint32_t *bnp = d_np + b0;

      const bool need_p1 = tune_p1 && !ix->tuned[prm->precision];
      const bool need_p2 = tune_p2 && !ix->tuned[MVR_FP32_MFMA];
      if (need_p1 || need_p2) {
        // autotune on this pass: each candidate setting runs the whole pass
        // once to warm up and once timed (results are identical for every
        // setting, so the pass stays valid).  Candidates: tiles in flight per
        // wave x workgroups per CU up to the kernel's residency.  With a
        // rescore, the FP32 phase's setting is tuned next (unless FP32 MFMA
        // already has one), the first phase's best held fixed.
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        auto time_pass = [&](const PersistCfg &c, uint32_t prec, float &ms) -> int {
          ix->tune[prec] = c;
          for (int rep = 0; rep < 2; ++rep) {
            HIP_TRY(hipEventRecord(e0, s));
            const int rc = run_pass(s, Qb, rm, n_rows, bnt, bnp, nb, oid, os, bi, false);
            if (rc != MVR_OK) return rc;
            HIP_TRY(hipEventRecord(e1, s));
            HIP_TRY(hipEventSynchronize(e1));
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
          }
          return MVR_OK;
        };
        auto tune_one = [&](uint32_t prec) -> int {
          float best = 1e30f, ms = 0.0f;
          PersistCfg best_c;
          for (uint32_t tps = 1; tps <= 2; ++tps)
            for (uint32_t wpc = 1; wpc <= persist_max_wgs((int)prec, tps); ++wpc) {
              PersistCfg c;
              c.n_wg = wpc * d.n_cu;
              c.tps = tps;
              const int rc = time_pass(c, prec, ms);
              if (rc != MVR_OK) return rc;
              if (ms < best) {
                best = ms;
                best_c = c;
              }
            }
          // the doc-major first phase against the best per-query setting
          // (which stays recorded for searches in the query order)
          if (prec == prm->precision && p_dm_tuned) {
            PersistCfg c = best_c;
            c.dm = true;
            const int rc = time_pass(c, prec, ms);
            if (rc != MVR_OK) return rc;
            if (ms < best) best_c = c;
          }
          ix->tune[prec] = best_c;
          ix->tuned[prec] = true;
          return MVR_OK;
        };
        int rc = need_p1 ? tune_one(prm->precision) : MVR_OK;
        if (rc == MVR_OK && need_p2) {
          ix->tuned[MVR_FP32_MFMA] = true;  // run_pass's rescore phase reads tune[MVR_FP32_MFMA]
          rc = tune_one(MVR_FP32_MFMA);
          if (rc != MVR_OK) ix->tuned[MVR_FP32_MFMA] = false;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != MVR_OK) {
          if (need_p1) ix->tuned[prm->precision] = false;
          return rc;
        }
      }

      if (use_graph) {
        // inputs of the pass into fixed buffers, then replay (or capture) the pass
        HIP_TRY(hipMemcpyAsync(ws.pass_nt.p, bnt, (size_t)nb * 4, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(ws.pass_np.p, bnp, (size_t)nb * 4, hipMemcpyDeviceToDevice, s));
        if (rm) HIP_TRY(hipMemcpyAsync(ws.pass_rm.p, rm, (size_t)n_rows * 4, hipMemcpyDeviceToDevice, s));
        const PersistCfg &tc = ix->tune[prm->precision];
        const std::vector<uint64_t> key = {nb, n_rows, t_max, P, R, R2, k, prm->precision, prm->clip_zero,
                                           uniform ? 1u : 0u, persistent ? 1u : 0u, fast_ct ? 1u : 0u,
                                           tc.n_wg, tc.tps, tc.dm ? 1u : 0u};
        hipGraphExec_t exec = nullptr;
        for (size_t gi = 0; gi < ws.graphs.size(); ++gi)
          if (ws.graphs[gi].key == key) {
            exec = ws.graphs[gi].exec;
            std::rotate(ws.graphs.begin() + gi, ws.graphs.begin() + gi + 1, ws.graphs.end());
            break;
          }
        if (!exec) {
          hipGraph_t graph = nullptr;
          HIP_TRY(hipStreamBeginCapture(ws.cap_stream, hipStreamCaptureModeRelaxed));
          const int rc = run_pass(ws.cap_stream, ws.Q.as<float>(), rm ? ws.pass_rm.as<uint32_t>() : nullptr,
                                  n_rows, ws.pass_nt.as<uint32_t>(), ws.pass_np.as<uint32_t>(), nb,
                                  ws.out_ids.as<uint32_t>(), ws.out_s.as<float>(), bi, false);
          const hipError_t ce = hipStreamEndCapture(ws.cap_stream, &graph);
          if (rc != MVR_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
          }
          HIP_TRY(ce);
          const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          HIP_TRY(ie);
          if (ws.graphs.size() >= 64) {
            (void)hipGraphExecDestroy(ws.graphs.front().exec);
            ws.graphs.erase(ws.graphs.begin());
          }
          ws.graphs.push_back(PassGraph{key, exec});
        }
        HIP_TRY(rec(bi, 0));
        HIP_TRY(hipGraphLaunch(exec, s));
        for (int i = 1; i < kNev; ++i) HIP_TRY(rec(bi, i));
        pass_graphed[bi] = 1;
        if (on_device) {
          HIP_TRY(hipMemcpyAsync(out_ids + (size_t)b0 * k, oid, (size_t)nb * k * 4, hipMemcpyDeviceToDevice, s));
          HIP_TRY(hipMemcpyAsync(out_scores + (size_t)b0 * k, os, (size_t)nb * k * 4, hipMemcpyDeviceToDevice, s));
        }
      } else {
        const int rc = run_pass(s, Qb, rm, n_rows, bnt, bnp, nb, oid, os, bi, true);
        if (rc != MVR_OK) return rc;
      }
      if (!on_device) {
        HIP_TRY(hipMemcpyAsync(out_ids + (size_t)b0 * k, oid, (size_t)nb * k * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(out_scores + (size_t)b0 * k, os, (size_t)nb * k * 4, hipMemcpyDeviceToHost, s));
      }
      if (stages) {
        HIP_TRY(hipStreamSynchronize(s));
        if (stages->top_ids)
          HIP_TRY(hipMemcpy(stages->top_ids + (size_t)b0 * t_max * P, ws.top_ids.p, (size_t)nb * t_max * P * 4, hipMemcpyDeviceToHost));
        if (stages->top_s)
          HIP_TRY(hipMemcpy(stages->top_s + (size_t)b0 * t_max * P, ws.top_s.p, (size_t)nb * t_max * P * 4, hipMemcpyDeviceToHost));
        if (stages->cand_n)
          HIP_TRY(hipMemcpy(stages->cand_n + b0, ws.cand_tot.p, (size_t)nb * 4, hipMemcpyDeviceToHost));
        if (stages->sel_ids)
          HIP_TRY(hipMemcpy(stages->sel_ids + (size_t)b0 * R, ws.sel_ids.p, (size_t)nb * R * 4, hipMemcpyDeviceToHost));
        if (stages->sel_F)
          HIP_TRY(hipMemcpy(stages->sel_F + (size_t)b0 * R, ws.sel_F.p, (size_t)nb * R * 4, hipMemcpyDeviceToHost));
        if (stages->sel_score)
          HIP_TRY(hipMemcpy(stages->sel_score + (size_t)b0 * R, ws.sel_score.p, (size_t)nb * R * 4, hipMemcpyDeviceToHost));
        if (stages->nprobe)
          HIP_TRY(hipMemcpy(stages->nprobe + b0, learned ? ws.dyn_np.p : (const void *)(d_np + b0), (size_t)nb * 4,
                            hipMemcpyDeviceToHost));
        if (stages->rerank_n)
          HIP_TRY(hipMemcpy(stages->rerank_n + b0, ws.sel_n.p, (size_t)nb * 4, hipMemcpyDeviceToHost));
      }
      if (timings)
        HIP_TRY(launch_count_work(ws.sel_ids.as<uint32_t>(), ws.sel_n.as<uint32_t>(), R, nb, d,
                                  ws.cand_tot.as<uint32_t>(), d_nt + b0, ws.counters.as<unsigned long long>(), s));
    }
  }
  if (timings) {
    HIP_TRY(hipEventSynchronize(ev[(size_t)(n_pass - 1) * kNev + 6]));
    for (uint32_t b = 0; b < n_pass; ++b) {
      float ms[6];
      for (int i = 0; i < 6; ++i)
        HIP_TRY(hipEventElapsedTime(&ms[i], ev[b * kNev + i], ev[b * kNev + i + 1]));
      // pipelined passes overlap: the call's span instead of the sum
      if (!pipelined) timings->total_ms += ms[0] + ms[1] + ms[2] + ms[3] + ms[4] + ms[5];
      if (pass_graphed[b]) continue;  // a replayed graph has no stage boundaries
      timings->centroid_ms += ms[0];
      timings->filter_ms += ms[1];
      timings->select_ms += ms[2];
      timings->rerank_ms += ms[3];
      timings->rescore_ms += ms[4];
      timings->topk_ms += ms[5];
      if (ix->res_host) {
        // the MaxSim alone (from event 8) and the staging (selection end to staged docs)
        float rm = 0.0f, sm = 0.0f;
        HIP_TRY(hipEventElapsedTime(&rm, ev[b * kNev + 8], ev[b * kNev + 4]));
        HIP_TRY(hipEventElapsedTime(&sm, ev[b * kNev + 3], ev[b * kNev + 7]));
        timings->rerank_ms += rm - ms[3];
        timings->staging_ms += sm;
      }
    }
    if (pipelined) {
      float span = 0.0f;
      HIP_TRY(hipEventElapsedTime(&span, ev[0], ev[(size_t)(n_pass - 1) * kNev + 6]));
      timings->total_ms = span;
    }
    timings->n_pass = n_pass;
    unsigned long long cnt[3] = {0, 0, 0};
    HIP_TRY(hipMemcpy(cnt, ws.counters.p, 24, hipMemcpyDeviceToHost));
    timings->rerank_tokens = cnt[0];
    timings->candidates = cnt[1];
    timings->rerank_pairs = cnt[2];
    for (auto &e : ev) (void)hipEventDestroy(e);
  }
  if (!on_device) HIP_TRY(hipStreamSynchronize(s));
  return MVR_OK;
}

int mvr_index_set_tuning(mvr_index *ix, uint32_t precision, uint32_t wgs_per_cu, uint32_t tiles) {
  if (!ix) return fail(MVR_ERR_INVALID, "NULL index");
  if (precision != MVR_BF16 && precision != MVR_FP8 && precision != MVR_FP32_MFMA)
    return fail(MVR_ERR_UNSUPPORTED, "no work-queue kernel for precision %u", precision);
  if (tiles < 1 || tiles > 2) return fail(MVR_ERR_INVALID, "tiles %u not in {1,2}", tiles);
  const uint32_t wmax = persist_max_wgs((int)precision, tiles);
  if (wmax == 0) return fail(MVR_ERR_INVALID, "no work-queue kernel with %u tiles for precision %u", tiles, precision);
  if (wgs_per_cu < 1 || wgs_per_cu > wmax)
    return fail(MVR_ERR_INVALID, "wgs_per_cu %u not in 1..%u (residency of that kernel)", wgs_per_cu, wmax);
  PersistCfg c;
  c.n_wg = wgs_per_cu * ix->d.n_cu;
  c.tps = tiles;
  ix->tune[precision] = c;  // the per-query persistent kernel (dm = false)
  ix->tuned[precision] = true;
  return MVR_OK;
}

int mvr_index_get_tuning_ex(const mvr_index *ix, uint32_t precision, uint32_t out[3]) {
  if (!ix || !out) return fail(MVR_ERR_INVALID, "NULL argument");
  if (precision > 3 || !ix->tuned[precision]) return fail(MVR_ERR_INVALID, "precision %u not tuned", precision);
  out[0] = ix->tune[precision].dm ? MVR_MAXSIM_DOC : MVR_MAXSIM_QUERY;
  out[1] = ix->tune[precision].n_wg / std::max(1u, ix->d.n_cu);
  out[2] = ix->tune[precision].tps;
  return MVR_OK;
}

int mvr_index_get_tuning(const mvr_index *ix, uint32_t precision, uint32_t out[2]) {
  if (!ix || !out) return fail(MVR_ERR_INVALID, "NULL argument");
  if (precision > 3 || !ix->tuned[precision]) return fail(MVR_ERR_INVALID, "precision %u not tuned", precision);
  out[0] = ix->tune[precision].n_wg / std::max(1u, ix->d.n_cu);
  out[1] = ix->tune[precision].tps;
  return MVR_OK;
}

int mvr_rerank_docs(mvr_index *ix, const float *Q_dev, uint32_t n_q, uint32_t t_max,
                    const uint32_t *q_ntok, const uint32_t *doc_ids_dev, const uint32_t *n_docs,
                    uint32_t R, uint32_t precision, float *scores_dev, void *stream) {
  if (!ix || !Q_dev || !doc_ids_dev || !scores_dev) return fail(MVR_ERR_INVALID, "NULL argument");
  if (ix->res_host)
    return fail(MVR_ERR_UNSUPPORTED, "mvr_rerank_docs needs a device-resident index (residuals are on the host)");
  if (n_q == 0 || R == 0) return MVR_OK;
  if (t_max == 0 || t_max > kMaxQTok) return fail(MVR_ERR_INVALID, "t_max %u not in [1, 32]", t_max);
  if (precision > MVR_FP32_MFMA) return fail(MVR_ERR_INVALID, "unknown precision %u", precision);
  if (ix->d.n_bit != 2 && precision != MVR_FP32)
    return fail(MVR_ERR_UNSUPPORTED, "%u-bit residual codes are reranked with MVR_FP32 only", ix->d.n_bit);
  DeviceGuard g(ix->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Workspace &ws = ix->ws;
  // metadata: q_ntok and the per-query doc counts (pinned staging, one copy)
  if (ws.staging_done) HIP_TRY(hipEventSynchronize(ws.staging_done));
  else HIP_TRY(hipEventCreateWithFlags(&ws.staging_done, hipEventDisableTiming));
  HIP_TRY(ws.staging.ensure((size_t)n_q * 8));
  uint32_t *hs = reinterpret_cast<uint32_t *>(ws.staging.p);
  for (uint32_t q = 0; q < n_q; ++q) {
    hs[q] = q_ntok ? q_ntok[q] : t_max;
    hs[n_q + q] = n_docs ? std::min(n_docs[q], R) : R;
    if (hs[q] == 0 || hs[q] > t_max) return fail(MVR_ERR_INVALID, "query %u has %u tokens", q, hs[q]);
  }
  HIP_TRY(ws.rr_meta.ensure((size_t)n_q * 8));
  HIP_TRY(hipMemcpyAsync(ws.rr_meta.p, hs, (size_t)n_q * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipEventRecord(ws.staging_done, s));
  const uint32_t *d_nt = ws.rr_meta.as<uint32_t>(), *d_n = d_nt + n_q;
  HIP_TRY(launch_rerank(Q_dev, t_max, d_nt, n_q, doc_ids_dev, d_n, R, ix->d, (int)precision, 0,
                        scores_dev, s));
  return MVR_OK;
}

int mvr_rerank_docs_noalloc(mvr_index *ix, const float *Q_dev, uint32_t n_q, uint32_t t_max,
                            const uint32_t *q_ntok, const uint32_t *doc_ids_dev, const uint32_t *n_docs,
                            uint32_t R, uint32_t precision, float *scores_dev, void *stream) {
  if (!ix || !Q_dev || !doc_ids_dev || !scores_dev) return fail(MVR_ERR_INVALID, "NULL argument");
  if (ix->res_host)
    return fail(MVR_ERR_UNSUPPORTED, "mvr_rerank_docs_noalloc needs a device-resident index (residuals are on the host)");
  if (n_q == 0 || R == 0) return MVR_OK;
  if (n_q > kNoAllocMaxQ) return fail(MVR_ERR_INVALID, "n_q %u > %u (the no-alloc scratch holds %u queries)", n_q,
                                      kNoAllocMaxQ, kNoAllocMaxQ);
  if (t_max == 0 || t_max > kMaxQTok) return fail(MVR_ERR_INVALID, "t_max %u not in [1, 32]", t_max);
  if (precision > MVR_FP32_MFMA) return fail(MVR_ERR_INVALID, "unknown precision %u", precision);
  if (ix->d.n_bit != 2 && precision != MVR_FP32)
    return fail(MVR_ERR_UNSUPPORTED, "%u-bit residual codes are reranked with MVR_FP32 only", ix->d.n_bit);
  if (q_ntok)
    for (uint32_t q = 0; q < n_q; ++q)
      if (q_ntok[q] == 0 || q_ntok[q] > t_max) return fail(MVR_ERR_INVALID, "query %u has %u tokens", q, q_ntok[q]);
  uint32_t *meta = nullptr;
  const int rc = scratch_on(ix->device, &meta);
  if (rc != MVR_OK) return rc;
  DeviceGuard g(ix->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the caller's host arrays straight into the scratch, then defaults and the clamp to R
  if (q_ntok) HIP_TRY(hipMemcpyAsync(meta, q_ntok, (size_t)n_q * 4, hipMemcpyHostToDevice, s));
  if (n_docs) HIP_TRY(hipMemcpyAsync(meta + n_q, n_docs, (size_t)n_q * 4, hipMemcpyHostToDevice, s));
  k_noalloc_meta<<<(n_q + 255) / 256, 256, 0, s>>>(meta, n_q, t_max, R, q_ntok ? 1 : 0, n_docs ? 1 : 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(launch_rerank(Q_dev, t_max, meta, n_q, doc_ids_dev, meta + n_q, R, ix->d, (int)precision, 0,
                        scores_dev, s));
  return MVR_OK;
}

int mvr_index_set_probe(mvr_index *ix, mvr_mlp *mlp, float conf, uint32_t fallback_nprobe,
                        uint32_t r_min) {
  if (!ix) return fail(MVR_ERR_INVALID, "NULL index");
  if (mlp && (fallback_nprobe == 0 || fallback_nprobe > kProbeFeat))
    return fail(MVR_ERR_INVALID, "fallback nprobe %u not in [1, 32]", fallback_nprobe);
  ix->probe = mlp;
  ix->probe_conf = conf;
  ix->probe_fallback = fallback_nprobe;
  ix->probe_rmin = r_min;
  return MVR_OK;
}

int mvr_decompress(mvr_index *ix, const uint32_t *doc_ids, uint32_t n, float *out) {
  if (!ix || (!doc_ids && n) || (!out && n)) return fail(MVR_ERR_INVALID, "NULL argument");
  if (ix->res_host)
    return fail(MVR_ERR_UNSUPPORTED, "mvr_decompress needs a device-resident index (residuals are on the host)");
  if (n == 0) return MVR_OK;
  std::vector<uint64_t> toks;
  uint64_t acc = 0;
  std::vector<uint64_t> docoff(ix->d.n_doc);
  for (uint32_t i = 0; i < ix->d.n_doc; ++i) { docoff[i] = acc; acc += ix->h_doclens[i]; }
  for (uint32_t i = 0; i < n; ++i) {
    if (doc_ids[i] >= ix->d.n_doc) return fail(MVR_ERR_INVALID, "doc id %u >= n_doc %u", doc_ids[i], ix->d.n_doc);
    for (uint32_t t = 0; t < ix->h_doclens[doc_ids[i]]; ++t) toks.push_back(docoff[doc_ids[i]] + t);
  }
  DeviceGuard g(ix->device);
  uint64_t *dt = nullptr;
  float *dout = nullptr;
  HIP_TRY(hipMalloc(reinterpret_cast<void **>(&dt), std::max<size_t>(1, toks.size()) * 8));
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&dout), std::max<size_t>(1, toks.size()) * kDim * 4);
  if (e == hipSuccess) e = hipMemcpy(dt, toks.data(), toks.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_decompress(dt, toks.size(), ix->d, dout, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out, dout, toks.size() * kDim * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dt);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) return fail(MVR_ERR_DEVICE, "decompress failed: %s", hipGetErrorString(e));
  return MVR_OK;
}

int mvr_exhaustive_f64(const float *Q_dev, uint32_t n_q, uint32_t t_max, const uint32_t *q_ntok_dev,
                       const float *X_dev, const uint32_t *doclens_dev, uint32_t n_doc,
                       double *out_dev, void *stream) {
  if (!Q_dev || !X_dev || !doclens_dev || !out_dev) return fail(MVR_ERR_INVALID, "NULL argument");
  if (t_max == 0 || t_max > kMaxQTok) return fail(MVR_ERR_INVALID, "t_max %u not in [1, 32]", t_max);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::vector<uint32_t> lens(n_doc);
  HIP_TRY(hipMemcpy(lens.data(), doclens_dev, (size_t)n_doc * 4, hipMemcpyDeviceToHost));
  std::vector<uint64_t> off(n_doc);
  uint64_t acc = 0;
  for (uint32_t i = 0; i < n_doc; ++i) { off[i] = acc; acc += lens[i]; }
  uint64_t *doff = nullptr;
  HIP_TRY(hipMalloc(reinterpret_cast<void **>(&doff), std::max<size_t>(1, n_doc) * 8));
  hipError_t e = hipMemcpy(doff, off.data(), (size_t)n_doc * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_exhaustive_f64(Q_dev, n_q, t_max, q_ntok_dev, X_dev, doclens_dev, doff, n_doc, out_dev, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(doff);
  if (e != hipSuccess) return fail(MVR_ERR_DEVICE, "exhaustive MaxSim failed: %s", hipGetErrorString(e));
  return MVR_OK;
}

int mvr_synth_tokens(uint64_t seed, uint64_t stream_comp, uint64_t stream_noise,
                     const float *centers_dev, uint32_t n_comp, uint32_t dim, double noise_std,
                     uint64_t tok_begin, uint64_t n_tok, float *out_dev, void *stream) {
  if (dim != kDim) return fail(MVR_ERR_UNSUPPORTED, "dim must be 128");
  if (!centers_dev || !out_dev || n_comp == 0) return fail(MVR_ERR_INVALID, "bad arguments");
  HIP_TRY(launch_synth_tokens(seed, stream_comp, stream_noise, centers_dev, n_comp, noise_std,
                              tok_begin, n_tok, out_dev, reinterpret_cast<hipStream_t>(stream)));
  return MVR_OK;
}

int mvr_assign_bias(const float *x_dev, uint64_t n, const float *C_dev, const float *half_norm_dev,
                    uint32_t n_c, uint32_t dim, uint32_t *codes_dev, void *stream) {
  if (dim != kDim) return fail(MVR_ERR_UNSUPPORTED, "dim must be 128");
  if (!x_dev || !C_dev || !codes_dev || n_c == 0) return fail(MVR_ERR_INVALID, "bad arguments");
  HIP_TRY(launch_assign(x_dev, n, C_dev, half_norm_dev, n_c, codes_dev,
                        reinterpret_cast<hipStream_t>(stream)));
  return MVR_OK;
}

int mvr_assign(const float *x_dev, uint64_t n, const float *C_dev, uint32_t n_c, uint32_t dim,
               uint32_t *codes_dev, void *stream) {
  return mvr_assign_bias(x_dev, n, C_dev, nullptr, n_c, dim, codes_dev, stream);
}

int mvr_compress(const float *x_dev, const uint32_t *codes_dev, uint64_t n, const float *C_dev,
                 uint32_t dim, const float *cutoffs_dev, int per_dim, uint8_t *out_dev, void *stream) {
  if (dim != kDim) return fail(MVR_ERR_UNSUPPORTED, "dim must be 128");
  if (!x_dev || !codes_dev || !C_dev || !cutoffs_dev || !out_dev) return fail(MVR_ERR_INVALID, "bad arguments");
  HIP_TRY(launch_compress(x_dev, codes_dev, n, C_dev, cutoffs_dev, per_dim, out_dev,
                          reinterpret_cast<hipStream_t>(stream)));
  return MVR_OK;
}

int mvr_compress_host(const float *x, const uint32_t *codes, uint64_t n, const float *C, uint32_t n_c,
                      uint32_t dim, const float *cutoffs, int per_dim, uint8_t *out, int device) {
  if (dim != kDim) return fail(MVR_ERR_UNSUPPORTED, "dim must be 128");
  if (n == 0) return MVR_OK;
  if (!x || !codes || !C || !cutoffs || !out || n_c == 0) return fail(MVR_ERR_INVALID, "bad arguments");
  for (uint64_t i = 0; i < n; ++i)
    if (codes[i] >= n_c)
      return fail(MVR_ERR_INVALID, "code %u at %llu >= n_centroid %u", codes[i], (unsigned long long)i, n_c);
  int n_dev = 0;
  HIP_TRY(hipGetDeviceCount(&n_dev));
  if (device < 0 || device >= n_dev) return fail(MVR_ERR_INVALID, "device %d out of range (%d)", device, n_dev);
  HIP_TRY(hipSetDevice(device));
  const uint64_t chunk = std::min<uint64_t>(n, 1ull << 22);  // 2 GB of embeddings per step
  float *dx = nullptr, *dC = nullptr, *dcut = nullptr;
  uint32_t *dcodes = nullptr;
  uint8_t *dout = nullptr;
  hipStream_t s = nullptr;
  auto cleanup = [&]() {
    if (s) (void)hipStreamSynchronize(s);
    for (void *p : {(void *)dx, (void *)dC, (void *)dcut, (void *)dcodes, (void *)dout})
      if (p) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  };
  const size_t cut_n = per_dim ? (size_t)kDim * 3 : 3;
  hipError_t e = hipSuccess;
  if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&dx, sizeof(float) * chunk * kDim)) != hipSuccess ||
      (e = hipMalloc(&dcodes, sizeof(uint32_t) * chunk)) != hipSuccess ||
      (e = hipMalloc(&dout, (size_t)kResBytes * chunk)) != hipSuccess ||
      (e = hipMalloc(&dC, sizeof(float) * (size_t)n_c * kDim)) != hipSuccess ||
      (e = hipMalloc(&dcut, sizeof(float) * cut_n)) != hipSuccess ||
      (e = hipMemcpyAsync(dC, C, sizeof(float) * (size_t)n_c * kDim, hipMemcpyHostToDevice, s)) != hipSuccess ||
      (e = hipMemcpyAsync(dcut, cutoffs, sizeof(float) * cut_n, hipMemcpyHostToDevice, s)) != hipSuccess) {
    cleanup();
    return fail(MVR_ERR_DEVICE, "compress setup failed: %s", hipGetErrorString(e));
  }
  for (uint64_t i0 = 0; i0 < n; i0 += chunk) {
    const uint64_t m = std::min(chunk, n - i0);
    if ((e = hipMemcpyAsync(dx, x + i0 * kDim, sizeof(float) * m * kDim, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(dcodes, codes + i0, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = launch_compress(dx, dcodes, m, dC, dcut, per_dim, dout, s)) != hipSuccess ||
        (e = hipMemcpyAsync(out + i0 * kResBytes, dout, (size_t)kResBytes * m, hipMemcpyDeviceToHost, s)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
      cleanup();
      return fail(MVR_ERR_DEVICE, "compress failed: %s", hipGetErrorString(e));
    }
  }
  cleanup();
  return MVR_OK;
}

}  // extern "C"
