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

// v_mfma_f32_16x16x32_bf16: the 4 outputs of each lane in elements 0..3
typedef float floatx4 __attribute__((ext_vector_type(4)));
__global__ void k_probe_bf16_16(const bf16x8 *a, const bf16x8 *b, floatx16 *d) {
  const uint32_t l = threadIdx.x;
  floatx4 acc = {};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[l], b[l], acc, 0, 0, 0);
  floatx16 o = {};
  o[0] = acc[0];
  o[1] = acc[1];
  o[2] = acc[2];
  o[3] = acc[3];
  d[l] = o;
}

__global__ void k_probe_fp8(const long *a, const long *b, floatx16 *d) {
  const uint32_t l = threadIdx.x;
  floatx16 acc = {};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a[l], b[l], acc, 0, 0, 0);
  d[l] = acc;
}

// v_mfma_f32_32x32x64_f8f6f4 (e4m3 operands, 32 bytes per lane)
typedef int intx8 __attribute__((ext_vector_type(8)));
__global__ void k_probe_fp8_64(const intx8 *a, const intx8 *b, floatx16 *d) {
  const uint32_t l = threadIdx.x;
  floatx16 acc = {};
  acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[l], b[l], acc, 0, 0, 0, 0, 0, 0);
  d[l] = acc;
}

// Packs pairs of floats to fp8 with the hardware conversion (e4m3).
__global__ void k_probe_cvt_fp8(const float *x, uint32_t *out, uint32_t n_pairs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pairs) return;
  out[i] = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(x[2 * i], x[2 * i + 1], 0, false) & 0xFFFFu;
}

// Direct global -> LDS loads of 16 bytes per lane (global_load_lds_dwordx4):
// lane l's piece lands at the LDS address given as the destination plus the
// immediate offset plus 16 l; the immediate offset is added to the global
// address as well.  Two loads into a 192-entry buffer pre-filled with a
// marker; every address stays inside src (128 pieces) and buf.
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void glb_void;
__global__ void k_probe_gll(const uint4 *src, const uint32_t *perm, uint4 *out) {
  __shared__ uint4 buf[192];
  const uint32_t l = threadIdx.x;
  for (uint32_t i = l; i < 192; i += 64) buf[i] = make_uint4(0xDEADu, 0xDEADu, 0xDEADu, 0xDEADu);
  __syncthreads();
  const uint4 *g = src + (perm[l] & 63u);
  __builtin_amdgcn_global_load_lds((glb_void *)g, (lds_void *)(buf + 32), 16, 0, 0);
  __builtin_amdgcn_global_load_lds((glb_void *)g, (lds_void *)(buf + 32), 16, 1024, 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (uint32_t i = l; i < 192; i += 64) out[i] = buf[i];
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

// kind 0: bf16 32x32x16, 1: fp8 32x32x16, 2: bf16 16x16x32, 3: fp8 32x32x64
extern "C" int probe_mfma(int kind, const void *a, const void *b, float *d) {
  const size_t op = kind == 1 ? 8 : kind == 3 ? 32 : 16;
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
    else if (kind == 2)
      k_probe_bf16_16<<<1, 64>>>((const bf16x8 *)da, (const bf16x8 *)db, (floatx16 *)dd);
    else if (kind == 3)
      k_probe_fp8_64<<<1, 64>>>((const intx8 *)da, (const intx8 *)db, (floatx16 *)dd);
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

extern "C" int probe_gll(const uint32_t *src, const uint32_t *perm, uint32_t *out) {
  void *ds = nullptr, *dp = nullptr, *dout = nullptr;
  FirstError err;
  err(hipMalloc(&ds, 128 * 16));
  if (err.ok()) err(hipMalloc(&dp, 64 * 4));
  if (err.ok()) err(hipMalloc(&dout, 192 * 16));
  if (err.ok()) err(hipMemcpy(ds, src, 128 * 16, hipMemcpyHostToDevice));
  if (err.ok()) err(hipMemcpy(dp, perm, 64 * 4, hipMemcpyHostToDevice));
  if (err.ok()) {
    k_probe_gll<<<1, 64>>>((const uint4 *)ds, (const uint32_t *)dp, (uint4 *)dout);
    err(hipGetLastError());
  }
  if (err.ok()) err(hipMemcpy(out, dout, 192 * 16, hipMemcpyDeviceToHost));
  (void)hipFree(ds);
  (void)hipFree(dp);
  (void)hipFree(dout);
  return err.ok() ? 0 : (int)err.e;
}
