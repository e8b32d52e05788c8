// Learned-probe MLP: training and inference on hand-written HIP kernels.
//
// Network (dims d0 -> d1 -> d2 -> 6): h1 = GELU(x W1^T + b1),
// h2 = GELU(h1 W2^T + b2), o = h2 W3^T + b3; o[:5] are logits over nprobe
// classes {2, 4, 8, 16, 32}, o[5] regresses log(probe_topk needed).  Loss =
// mean cross-entropy + mean squared error; Adam (0.9, 0.999, 1e-8, bias
// corrected).  Every reduction has a fixed order, so a training step is
// deterministic.  The reference has no learned component (it probes a fixed
// nprobe, src/include/alg/igp_gpu_pp/GenerateVQScoreCandidatePP.hpp:166-174);
// this model chooses nprobe per query before the filter runs.
#include "../../include/mvr.h"
#include "mvr_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace mvr {
namespace {

// C[M][N] = A(M x K) * B(K x N) with strided operands:
//   A(i, k) = A[i * a_i + k * a_k],  B(k, j) = B[k * b_k + j * b_j].
// Epilogue: + bias[j] (nullable); Z (nullable) receives the pre-activation;
// C receives GELU(.) when gelu != 0.  FP32 on the matrix cores
// (v_mfma_f32_32x32x2_f32): a 64x64 tile per workgroup, each of the four
// waves owns a 32x32 quarter; K in chunks of 16 staged through LDS with zeros
// past M, N and K, eight MFMAs per chunk (MFMA s reads k = 2s from the lower
// lane half and k = 2s + 1 from the upper).  The products are summed in the
// matrix unit's fixed order, so a step is deterministic.
constexpr int kTm = 64, kTn = 64, kTk = 16;
typedef float mlp_floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = expf(-0.5f * z * z) * 0.39894228040143268f;
  return cdf + z * pdf;
}

__global__ __launch_bounds__(256) void k_gemm(const float *__restrict__ A, size_t a_i, size_t a_k,
                                              const float *__restrict__ B, size_t b_k, size_t b_j,
                                              float *__restrict__ Cm, const float *__restrict__ bias,
                                              float *__restrict__ Z, int gelu, uint32_t M,
                                              uint32_t N, uint32_t K) {
  __shared__ float As[kTk][kTm + 4];
  __shared__ float Bs[kTk][kTn + 4];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t t = lane & 31, h = lane >> 5;
  const uint32_t wi = (wave >> 1) * 32, wj = (wave & 1) * 32;  // the wave's quarter of the tile
  const uint32_t i0 = blockIdx.y * kTm, j0 = blockIdx.x * kTn;
  mlp_floatx16 acc = {};
  for (uint32_t k0 = 0; k0 < K; k0 += kTk) {
    for (uint32_t e = tid; e < kTm * kTk; e += 256) {
      const uint32_t r = e / kTk, kk = e % kTk;
      const uint32_t i = i0 + r, k = k0 + kk;
      As[kk][r] = (i < M && k < K) ? A[i * a_i + k * a_k] : 0.0f;
    }
    for (uint32_t e = tid; e < kTn * kTk; e += 256) {
      const uint32_t c = e % kTn, kk = e / kTn;
      const uint32_t j = j0 + c, k = k0 + kk;
      Bs[kk][c] = (j < N && k < K) ? B[k * b_k + j * b_j] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t st = 0; st < kTk / 2; ++st)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(As[2 * st + h][wi + t], Bs[2 * st + h][wj + t], acc, 0, 0, 0);
    __syncthreads();
  }
  // acc[4g + e] holds row 8g + 4h + e, column t of the wave's quarter
  const uint32_t j = j0 + wj + t;
  if (j >= N) return;
  const float bj = bias ? bias[j] : 0.0f;
#pragma unroll
  for (uint32_t r = 0; r < 16; ++r) {
    const uint32_t i = i0 + wi + 8 * (r >> 2) + 4 * h + (r & 3);
    if (i >= M) continue;
    const float z = acc[r] + bj;
    if (Z) Z[(size_t)i * N + j] = z;
    Cm[(size_t)i * N + j] = gelu ? gelu_f(z) : z;
  }
}

// dZ = dH * GELU'(Z), in place on dH.
__global__ void k_gelu_bwd(float *__restrict__ dH, const float *__restrict__ Z, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dH[i] *= gelu_grad_f(Z[i]);
}

// db[j] = sum_i dZ[i][j], i ascending.
__global__ void k_colsum(const float *__restrict__ dZ, uint32_t M, uint32_t N, float *__restrict__ db) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.0f;
  for (uint32_t i = 0; i < M; ++i) s += dZ[(size_t)i * N + j];
  db[j] = s;
}

// Loss head: softmax-CE over o[:5] and squared error on o[5]; dO scaled by 1/B.
__global__ void k_loss_head(const float *__restrict__ O, const int32_t *__restrict__ y,
                            const float *__restrict__ r, uint32_t B, float *__restrict__ dO,
                            float *__restrict__ row_loss) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float *o = O + (size_t)b * 6;
  float m = o[0];
  for (int c = 1; c < 5; ++c) m = fmaxf(m, o[c]);
  float e[5], se = 0.0f;
  for (int c = 0; c < 5; ++c) { e[c] = expf(o[c] - m); se += e[c]; }
  const int yc = y[b];
  const float inv_b = 1.0f / (float)B;
  for (int c = 0; c < 5; ++c) dO[(size_t)b * 6 + c] = (e[c] / se - (c == yc ? 1.0f : 0.0f)) * inv_b;
  const float diff = o[5] - r[b];
  dO[(size_t)b * 6 + 5] = 2.0f * diff * inv_b;
  row_loss[b] = (logf(se) + m - o[yc]) + diff * diff;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, size_t n, float lr, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// Features of a query: its tokens' top-32 centroid scores (S' - 1), sorted,
// 32 x 32 = 1024 values; rows beyond the query's token count are zero.
__global__ void k_probe_features(const float *__restrict__ top_s, uint32_t t_max, uint32_t P,
                                 const uint32_t *__restrict__ q_ntok, uint32_t n_q,
                                 float *__restrict__ X) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_q * 1024u) return;
  const uint32_t q = g >> 10, t = (g >> 5) & 31, j = g & 31;
  const uint32_t T = q_ntok ? q_ntok[q] : t_max;
  float v = 0.0f;
  if (t < T && j < P) v = top_s[((size_t)q * t_max + t) * P + j] - 1.0f;
  X[g] = v;
}

// nprobe = class of the largest softmax when that probability reaches conf,
// else fallback; probe_topk = clamp(round(exp(o[5])), r_min, r_max).
__global__ void k_probe_decide(const float *__restrict__ O, uint32_t n_q, float conf,
                               uint32_t fallback, uint32_t r_min, uint32_t r_max,
                               uint32_t *__restrict__ q_np, uint32_t *__restrict__ q_r) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_q) return;
  const float *o = O + (size_t)q * 6;
  float m = o[0];
  int best = 0;
  for (int c = 1; c < 5; ++c)
    if (o[c] > m) { m = o[c]; best = c; }
  float se = 0.0f;
  for (int c = 0; c < 5; ++c) se += expf(o[c] - m);
  const float pmax = 1.0f / se;
  q_np[q] = pmax >= conf ? (2u << best) : fallback;
  if (q_r) {
    const float rv = rintf(expf(fminf(o[5], 20.0f)));
    q_r[q] = (uint32_t)fminf(fmaxf(rv, (float)r_min), (float)r_max);
  }
}

}  // namespace

hipError_t launch_probe_features(const float *top_s, uint32_t t_max, uint32_t P,
                                 const uint32_t *q_ntok, uint32_t n_q, float *X, hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  k_probe_features<<<(n_q * 1024u + 255) / 256, 256, 0, s>>>(top_s, t_max, P, q_ntok, n_q, X);
  return hipGetLastError();
}

hipError_t launch_probe_decide(const float *O, uint32_t n_q, float conf, uint32_t fallback,
                               uint32_t r_min, uint32_t r_max, uint32_t *q_np, uint32_t *q_r,
                               hipStream_t s) {
  if (n_q == 0) return hipSuccess;
  k_probe_decide<<<(n_q + 255) / 256, 256, 0, s>>>(O, n_q, conf, fallback, r_min, r_max, q_np, q_r);
  return hipGetLastError();
}

}  // namespace mvr

using namespace mvr;

struct mvr_mlp {
  int device = 0;
  uint32_t dims[4] = {0, 0, 0, 0};
  size_t off[7] = {};            // parameter offsets: W1 b1 W2 b2 W3 b3, total
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint32_t step = 0;
  uint32_t cap_b = 0;            // activation buffers sized for this batch
  float *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *o = nullptr, *dO = nullptr,
        *d2 = nullptr, *d1 = nullptr, *rl = nullptr;
};

namespace {
thread_local std::string g_mlp_err;
int mfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_mlp_err = buf;
  return code;
}
#define MLP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess) return mfail(MVR_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

hipError_t gemm(const float *A, size_t a_i, size_t a_k, const float *B, size_t b_k, size_t b_j,
                float *Cm, const float *bias, float *Z, int gelu, uint32_t M, uint32_t N, uint32_t K,
                hipStream_t s) {
  dim3 grid((N + kTn - 1) / kTn, (M + kTm - 1) / kTm);
  k_gemm<<<grid, 256, 0, s>>>(A, a_i, a_k, B, b_k, b_j, Cm, bias, Z, gelu, M, N, K);
  return hipGetLastError();
}

hipError_t ensure_batch(mvr_mlp *n, uint32_t B) {
  if (B <= n->cap_b) return hipSuccess;
  for (float *q : {n->z1, n->h1, n->z2, n->h2, n->o, n->dO, n->d2, n->d1, n->rl})
    if (q) (void)hipFree(q);
  const uint32_t d1 = n->dims[1], d2 = n->dims[2], d3 = n->dims[3];
  hipError_t e = hipSuccess;
  auto al = [&](float **q, size_t cnt) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(q), cnt * sizeof(float));
  };
  al(&n->z1, (size_t)B * d1);
  al(&n->h1, (size_t)B * d1);
  al(&n->z2, (size_t)B * d2);
  al(&n->h2, (size_t)B * d2);
  al(&n->o, (size_t)B * d3);
  al(&n->dO, (size_t)B * d3);
  al(&n->d2, (size_t)B * d2);
  al(&n->d1, (size_t)B * d1);
  al(&n->rl, (size_t)B);
  if (e == hipSuccess) n->cap_b = B;
  return e;
}

// Forward pass into n->z*, n->h*, n->o.
hipError_t forward(mvr_mlp *n, const float *x, uint32_t B, hipStream_t s) {
  const uint32_t d0 = n->dims[0], d1 = n->dims[1], d2 = n->dims[2], d3 = n->dims[3];
  const float *W1 = n->p + n->off[0], *b1 = n->p + n->off[1], *W2 = n->p + n->off[2],
              *b2 = n->p + n->off[3], *W3 = n->p + n->off[4], *b3 = n->p + n->off[5];
  hipError_t e = gemm(x, d0, 1, W1, 1, d0, n->h1, b1, n->z1, 1, B, d1, d0, s);
  if (e == hipSuccess) e = gemm(n->h1, d1, 1, W2, 1, d1, n->h2, b2, n->z2, 1, B, d2, d1, s);
  if (e == hipSuccess) e = gemm(n->h2, d2, 1, W3, 1, d2, n->o, b3, nullptr, 0, B, d3, d2, s);
  return e;
}

// Loss + gradients into n->g (and per-row losses into n->rl).
hipError_t backward(mvr_mlp *n, const float *x, const int32_t *y, const float *r, uint32_t B,
                    hipStream_t s) {
  const uint32_t d0 = n->dims[0], d1 = n->dims[1], d2 = n->dims[2], d3 = n->dims[3];
  const float *W2 = n->p + n->off[2], *W3 = n->p + n->off[4];
  float *gW1 = n->g + n->off[0], *gb1 = n->g + n->off[1], *gW2 = n->g + n->off[2],
        *gb2 = n->g + n->off[3], *gW3 = n->g + n->off[4], *gb3 = n->g + n->off[5];
  k_loss_head<<<(B + 255) / 256, 256, 0, s>>>(n->o, y, r, B, n->dO, n->rl);
  hipError_t e = hipGetLastError();
  // layer 3: gW3 = dO^T h2, gb3 = colsum dO, d2 = (dO W3) * GELU'(z2)
  if (e == hipSuccess) e = gemm(n->dO, 1, d3, n->h2, d2, 1, gW3, nullptr, nullptr, 0, d3, d2, B, s);
  if (e == hipSuccess) { k_colsum<<<(d3 + 255) / 256, 256, 0, s>>>(n->dO, B, d3, gb3); e = hipGetLastError(); }
  if (e == hipSuccess) e = gemm(n->dO, d3, 1, W3, d2, 1, n->d2, nullptr, nullptr, 0, B, d2, d3, s);
  if (e == hipSuccess) {
    k_gelu_bwd<<<(unsigned)(((size_t)B * d2 + 255) / 256), 256, 0, s>>>(n->d2, n->z2, (size_t)B * d2);
    e = hipGetLastError();
  }
  // layer 2
  if (e == hipSuccess) e = gemm(n->d2, 1, d2, n->h1, d1, 1, gW2, nullptr, nullptr, 0, d2, d1, B, s);
  if (e == hipSuccess) { k_colsum<<<(d2 + 255) / 256, 256, 0, s>>>(n->d2, B, d2, gb2); e = hipGetLastError(); }
  if (e == hipSuccess) e = gemm(n->d2, d2, 1, W2, d1, 1, n->d1, nullptr, nullptr, 0, B, d1, d2, s);
  if (e == hipSuccess) {
    k_gelu_bwd<<<(unsigned)(((size_t)B * d1 + 255) / 256), 256, 0, s>>>(n->d1, n->z1, (size_t)B * d1);
    e = hipGetLastError();
  }
  // layer 1
  if (e == hipSuccess) e = gemm(n->d1, 1, d1, x, d0, 1, gW1, nullptr, nullptr, 0, d1, d0, B, s);
  if (e == hipSuccess) { k_colsum<<<(d1 + 255) / 256, 256, 0, s>>>(n->d1, B, d1, gb1); e = hipGetLastError(); }
  return e;
}

}  // namespace

extern "C" {

const char *mvr_mlp_last_error(void) { return g_mlp_err.c_str(); }

int mvr_mlp_create(const uint32_t *dims, int device, mvr_mlp **out) {
  if (!dims || !out) return mfail(MVR_ERR_INVALID, "NULL argument");
  if (dims[3] != 6) return mfail(MVR_ERR_INVALID, "output width must be 6 (5 classes + cutoff)");
  for (int i = 0; i < 4; ++i)
    if (dims[i] == 0) return mfail(MVR_ERR_INVALID, "zero layer width");
  MLP_TRY(hipSetDevice(device));
  mvr_mlp *n = new mvr_mlp();
  n->device = device;
  std::memcpy(n->dims, dims, sizeof(n->dims));
  size_t o = 0;
  for (int l = 0; l < 3; ++l) {
    n->off[2 * l] = o;
    o += (size_t)dims[l + 1] * dims[l];
    n->off[2 * l + 1] = o;
    o += dims[l + 1];
  }
  n->off[6] = o;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&n->p), o * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&n->g), o * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&n->m), o * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&n->v), o * 4);
  if (e == hipSuccess) e = hipMemset(n->p, 0, o * 4);
  if (e == hipSuccess) e = hipMemset(n->m, 0, o * 4);
  if (e == hipSuccess) e = hipMemset(n->v, 0, o * 4);
  if (e != hipSuccess) {
    mvr_mlp_free(n);
    return mfail(MVR_ERR_DEVICE, "allocation failed: %s", hipGetErrorString(e));
  }
  *out = n;
  return MVR_OK;
}

void mvr_mlp_free(mvr_mlp *n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  (void)hipDeviceSynchronize();
  for (float *q : {n->p, n->g, n->m, n->v, n->z1, n->h1, n->z2, n->h2, n->o, n->dO, n->d2, n->d1, n->rl})
    if (q) (void)hipFree(q);
  delete n;
}

uint64_t mvr_mlp_num_params(const mvr_mlp *n) { return n ? n->off[6] : 0; }

int mvr_mlp_set_params(mvr_mlp *n, const float *flat) {
  if (!n || !flat) return mfail(MVR_ERR_INVALID, "NULL argument");
  MLP_TRY(hipMemcpy(n->p, flat, n->off[6] * 4, hipMemcpyHostToDevice));
  MLP_TRY(hipMemset(n->m, 0, n->off[6] * 4));
  MLP_TRY(hipMemset(n->v, 0, n->off[6] * 4));
  n->step = 0;
  return MVR_OK;
}

int mvr_mlp_get_params(const mvr_mlp *n, float *flat) {
  if (!n || !flat) return mfail(MVR_ERR_INVALID, "NULL argument");
  MLP_TRY(hipMemcpy(flat, n->p, n->off[6] * 4, hipMemcpyDeviceToHost));
  return MVR_OK;
}

int mvr_mlp_forward(mvr_mlp *n, const float *x_dev, uint32_t B, float *out_dev, void *stream) {
  if (!n || !x_dev || !out_dev) return mfail(MVR_ERR_INVALID, "NULL argument");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  MLP_TRY(ensure_batch(n, B));
  MLP_TRY(forward(n, x_dev, B, s));
  MLP_TRY(hipMemcpyAsync(out_dev, n->o, (size_t)B * n->dims[3] * 4, hipMemcpyDeviceToDevice, s));
  return MVR_OK;
}

int mvr_mlp_grads(mvr_mlp *n, const float *x_dev, const int32_t *y_dev, const float *r_dev,
                  uint32_t B, float *grads_host, float *loss) {
  if (!n || !x_dev || !y_dev || !r_dev || B == 0) return mfail(MVR_ERR_INVALID, "bad arguments");
  MLP_TRY(ensure_batch(n, B));
  MLP_TRY(forward(n, x_dev, B, nullptr));
  MLP_TRY(backward(n, x_dev, y_dev, r_dev, B, nullptr));
  MLP_TRY(hipDeviceSynchronize());
  if (grads_host) MLP_TRY(hipMemcpy(grads_host, n->g, n->off[6] * 4, hipMemcpyDeviceToHost));
  if (loss) {
    std::vector<float> rl(B);
    MLP_TRY(hipMemcpy(rl.data(), n->rl, (size_t)B * 4, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (float v : rl) s += v;
    *loss = (float)(s / B);
  }
  return MVR_OK;
}

int mvr_mlp_train_step(mvr_mlp *n, const float *x_dev, const int32_t *y_dev, const float *r_dev,
                       uint32_t B, float lr, float *loss, void *stream) {
  if (!n || !x_dev || !y_dev || !r_dev || B == 0) return mfail(MVR_ERR_INVALID, "bad arguments");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  MLP_TRY(ensure_batch(n, B));
  MLP_TRY(forward(n, x_dev, B, s));
  MLP_TRY(backward(n, x_dev, y_dev, r_dev, B, s));
  n->step++;
  const float c1 = 1.0f - std::pow(0.9f, (float)n->step), c2 = 1.0f - std::pow(0.999f, (float)n->step);
  const size_t np = n->off[6];
  k_adam<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(n->p, n->g, n->m, n->v, np, lr, c1, c2);
  MLP_TRY(hipGetLastError());
  if (loss) {
    std::vector<float> rl(B);
    MLP_TRY(hipMemcpyAsync(rl.data(), n->rl, (size_t)B * 4, hipMemcpyDeviceToHost, s));
    MLP_TRY(hipStreamSynchronize(s));
    double sum = 0.0;
    for (float v : rl) sum += v;
    *loss = (float)(sum / B);
  }
  return MVR_OK;
}

}  // extern "C"
