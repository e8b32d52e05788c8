// Lane-layout probe for the MFMA builtins the bf16/fp8 kernels rely on
// (test infrastructure, built into tests/probes/libmvr_probe.so).  One wave;
// every index stays inside the 64-lane buffers the host passes.
#include <hip/hip_runtime.h>
#include <cstdint>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__global__ void k_probe_bf16(const bf16x8 *a, const bf16x8 *b, floatx16 *d) {
  const uint32_t l = threadIdx.x;  // 64 lanes
  floatx16 acc = {};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[l], b[l], acc, 0, 0, 0);
  d[l] = acc;
}

__global__ void k_probe_fp8(const long *a, const long *b, floatx16 *d) {
  const uint32_t l = threadIdx.x;
  floatx16 acc = {};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a[l], b[l], acc, 0, 0, 0);
  d[l] = acc;
}

// Packs pairs of floats to fp8 with the hardware conversion (e4m3).
__global__ void k_probe_cvt_fp8(const float *x, uint32_t *out, uint32_t n_pairs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pairs) return;
  out[i] = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(x[2 * i], x[2 * i + 1], 0, false) & 0xFFFFu;
}

// The first failing status of a sequence of HIP calls; later calls are
// skipped once one has failed.
struct FirstError {
  hipError_t e = hipSuccess;
  bool ok() const { return e == hipSuccess; }
  void operator()(hipError_t x) {
    if (e == hipSuccess) e = x;
  }
};

extern "C" int probe_mfma(int kind, const void *a, const void *b, float *d) {
  const size_t op = kind == 0 ? 16 : 8;
  void *da = nullptr, *db = nullptr, *dd = nullptr;
  FirstError err;
  err(hipMalloc(&da, 64 * op));
  if (err.ok()) err(hipMalloc(&db, 64 * op));
  if (err.ok()) err(hipMalloc(&dd, 64 * 64));
  if (err.ok()) err(hipMemcpy(da, a, 64 * op, hipMemcpyHostToDevice));
  if (err.ok()) err(hipMemcpy(db, b, 64 * op, hipMemcpyHostToDevice));
  if (err.ok()) {
    if (kind == 0)
      k_probe_bf16<<<1, 64>>>((const bf16x8 *)da, (const bf16x8 *)db, (floatx16 *)dd);
    else
      k_probe_fp8<<<1, 64>>>((const long *)da, (const long *)db, (floatx16 *)dd);
    err(hipGetLastError());
  }
  if (err.ok()) err(hipMemcpy(d, dd, 64 * 64, hipMemcpyDeviceToHost));
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dd);
  return !err.ok();
}

extern "C" int probe_cvt_fp8(const float *x, uint32_t *out, uint32_t n_pairs) {
  float *dx = nullptr;
  uint32_t *dout = nullptr;
  FirstError err;
  err(hipMalloc((void **)&dx, n_pairs * 8));
  if (err.ok()) err(hipMalloc((void **)&dout, n_pairs * 4));
  if (err.ok()) err(hipMemcpy(dx, x, n_pairs * 8, hipMemcpyHostToDevice));
  if (err.ok()) {
    k_probe_cvt_fp8<<<(n_pairs + 63) / 64, 64>>>(dx, dout, n_pairs);
    err(hipGetLastError());
  }
  if (err.ok()) err(hipMemcpy(out, dout, n_pairs * 4, hipMemcpyDeviceToHost));
  (void)hipFree(dx);
  (void)hipFree(dout);
  return !err.ok();
}

// Single-precision square root and division: the default operators (what the
// PLAID refine kernel uses) next to the __fsqrt_rn / __fdiv_rn intrinsics.
__global__ void k_probe_divsqrt(const float *a, const float *b, float *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[4 * i + 0] = __fsqrt_rn(a[i]);
  out[4 * i + 1] = __fdiv_rn(a[i], b[i]);
  out[4 * i + 2] = sqrtf(a[i]);
  out[4 * i + 3] = a[i] / b[i];
}

extern "C" int probe_divsqrt(const float *a, const float *b, float *out, uint32_t n) {
  float *da = nullptr, *db = nullptr, *dout = nullptr;
  FirstError err;
  err(hipMalloc((void **)&da, n * 4));
  if (err.ok()) err(hipMalloc((void **)&db, n * 4));
  if (err.ok()) err(hipMalloc((void **)&dout, n * 16));
  if (err.ok()) err(hipMemcpy(da, a, n * 4, hipMemcpyHostToDevice));
  if (err.ok()) err(hipMemcpy(db, b, n * 4, hipMemcpyHostToDevice));
  if (err.ok()) {
    k_probe_divsqrt<<<(n + 255) / 256, 256>>>(da, db, dout, n);
    err(hipGetLastError());
  }
  if (err.ok()) err(hipMemcpy(out, dout, (size_t)n * 16, hipMemcpyDeviceToHost));
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dout);
  return !err.ok();
}
