// GPGPUBuildIndexPackGPU: pybind11 index-build module with the reference's
// Python surface (src/app/build_index/GPGPUBuildIndexPackGPU.cpp:14-42):
//   compute_quantized_scalar(item_vec_l, centroid_l, code_l, n_bit) -> (cutoffs, weights)
//       ResidualScalarQuantizationPack.hpp:17-99
//   CompressResidualCode(centroid_l, cutoff_l, weight_l, n_bit)
//       .n_val_per_vec, .compute_residual_code(vec_l, code_l) -> (codes,)
//       ResidualScalarQuantizationPackGPU.hpp:30-80 (first cutoff > e, MSB-first)
//   IVFIndex(vq_code_l, item_n_vec_l, n_vec, n_item, n_centroid).build() / .save(location)
//       src/include/struct/IVFIndex.hpp:78-183
// plus compute_quantized_scalar_per_dim (per-dimension cutoffs, bucket-mean
// weights) for the synthetic configurations.  compute_residual_code runs on
// the GPU through libmvr (mvr_compress_host) when one is present, as the
// reference's does (MVR_BUILD_DEVICE=cpu|gpu|auto); `.device` tells where the
// last call ran.  Both give the same bytes.  Arrays are taken C-contiguous
// with forcecast (src/include/struct/TypeDef.hpp:13-21).  Invalid arguments
// raise ValueError instead of exit(-1).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace py = pybind11;
using arr_f32 = py::array_t<float, py::array::c_style | py::array::forcecast>;
using arr_u32 = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>;

namespace {

void check_nbit(uint32_t n_bit) {
  if (n_bit != 1 && n_bit != 2 && n_bit != 4 && n_bit != 8)
    throw py::value_error("n_bit must be 1, 2, 4 or 8");
}

// Value of rank floor(q * n) of the sorted values (what std::nth_element
// returns at that index).
float rank_value(std::vector<float> &v, double q) {
  const size_t idx = (size_t)std::floor(q * (double)v.size());
  std::nth_element(v.begin(), v.begin() + idx, v.end());
  return v[idx];
}

std::vector<float> residual_values(const arr_f32 &x, const arr_f32 &C, const arr_u32 &codes,
                                   size_t &n, uint32_t &dim) {
  if (x.ndim() != 2 || C.ndim() != 2 || codes.ndim() != 1)
    throw py::value_error("expected item_vec_l [n, d], centroid_l [c, d], code_l [n]");
  n = x.shape(0);
  dim = (uint32_t)x.shape(1);
  if ((uint32_t)C.shape(1) != dim || (size_t)codes.shape(0) != n)
    throw py::value_error("shape mismatch between item_vec_l, centroid_l and code_l");
  const float *xp = x.data(), *cp = C.data();
  const uint32_t *kp = codes.data();
  const size_t n_c = C.shape(0);
  std::vector<float> e(n * dim);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    if (kp[i] >= n_c) continue;
    for (uint32_t d = 0; d < dim; ++d) e[i * dim + d] = xp[i * dim + d] - cp[(size_t)kp[i] * dim + d];
  }
  for (size_t i = 0; i < n; ++i)
    if (kp[i] >= n_c) throw py::value_error("code_l holds an id >= number of centroids");
  return e;
}

py::tuple compute_quantized_scalar(const arr_f32 &x, const arr_f32 &C, const arr_u32 &codes,
                                   uint32_t n_bit) {
  check_nbit(n_bit);
  size_t n;
  uint32_t dim;
  std::vector<float> e = residual_values(x, C, codes, n, dim);
  if (e.empty()) throw py::value_error("no residuals");
  const uint32_t nq = 1u << n_bit;
  std::vector<float> cut(nq - 1), w(nq);
  for (uint32_t i = 1; i < nq; ++i) cut[i - 1] = rank_value(e, (double)i / nq);
  for (uint32_t i = 0; i < nq; ++i) w[i] = rank_value(e, (double)i / nq + 0.5 / nq);
  return py::make_tuple(cut, w);
}

uint32_t bucket(const float *cut, uint32_t ncut, float e) {
  for (uint32_t k = 0; k < ncut; ++k)
    if (cut[k] > e) return k;
  return ncut;
}

py::tuple compute_quantized_scalar_per_dim(const arr_f32 &x, const arr_f32 &C, const arr_u32 &codes,
                                           uint32_t n_bit) {
  check_nbit(n_bit);
  size_t n;
  uint32_t dim;
  std::vector<float> e = residual_values(x, C, codes, n, dim);
  if (n == 0) throw py::value_error("no residuals");
  const uint32_t nq = 1u << n_bit, ncut = nq - 1;
  py::array_t<float> cut_arr({(py::ssize_t)dim, (py::ssize_t)ncut});
  py::array_t<float> w_arr({(py::ssize_t)dim, (py::ssize_t)nq});
  float *cut = cut_arr.mutable_data(), *w = w_arr.mutable_data();
#pragma omp parallel for schedule(dynamic)
  for (long long d = 0; d < (long long)dim; ++d) {
    std::vector<float> col(n);
    for (size_t i = 0; i < n; ++i) col[i] = e[i * dim + d];
    std::vector<float> sorted(col);
    std::sort(sorted.begin(), sorted.end());
    for (uint32_t i = 1; i < nq; ++i)
      cut[d * ncut + i - 1] = sorted[(size_t)std::floor((double)i / nq * (double)n)];
    std::vector<double> sum(nq, 0.0);
    std::vector<size_t> cnt(nq, 0);
    for (size_t i = 0; i < n; ++i) {
      const uint32_t b = bucket(cut + d * ncut, ncut, col[i]);
      sum[b] += col[i];
      cnt[b]++;
    }
    for (uint32_t b = 0; b < nq; ++b) {
      float v;
      if (cnt[b]) v = (float)(sum[b] / (double)cnt[b]);
      else if (b == 0) v = cut[d * ncut];
      else if (b == ncut) v = cut[d * ncut + ncut - 1];
      else v = 0.5f * (cut[d * ncut + b - 1] + cut[d * ncut + b]);
      w[d * nq + b] = v;
    }
  }
  return py::make_tuple(cut_arr, w_arr);
}

// The GPU residual coder of libmvr (mvr_compress_host), loaded from the
// directory of this module on first use.  torch is imported first when it is
// installed, so that libmvr binds to the HIP runtime torch already carries.
struct GpuCoder {
  using compress_fn = int (*)(const float *, const uint32_t *, uint64_t, const float *, uint32_t, uint32_t,
                              const float *, int, uint8_t *, int);
  compress_fn compress = nullptr;
  int (*device_count)(int *) = nullptr;
  const char *(*last_error)() = nullptr;
  bool tried = false;
  int n_dev = 0;
};

GpuCoder &gpu_coder() {
  static GpuCoder g;
  if (g.tried) return g;
  g.tried = true;
  try {
    py::module_::import("torch");
  } catch (...) {
  }
  Dl_info info;
  if (!dladdr(reinterpret_cast<void *>(&gpu_coder), &info) || !info.dli_fname) return g;
  std::string dir(info.dli_fname);
  dir = dir.substr(0, dir.find_last_of('/') + 1);
  void *h = dlopen((dir + "libmvr.so").c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h) return g;
  g.compress = reinterpret_cast<GpuCoder::compress_fn>(dlsym(h, "mvr_compress_host"));
  g.device_count = reinterpret_cast<int (*)(int *)>(dlsym(h, "mvr_device_count"));
  g.last_error = reinterpret_cast<const char *(*)()>(dlsym(h, "mvr_last_error"));
  if (!g.compress || !g.device_count || !g.last_error || g.device_count(&g.n_dev) != 0) g.n_dev = 0;
  return g;
}

// MVR_BUILD_DEVICE: "cpu" (host threads), "gpu" (device 0, error when there
// is none), or unset / "auto" (the GPU when one is there and the shape is the
// device kernel's: 2-bit codes of 128-dim vectors).
std::string build_device_mode() {
  const char *e = std::getenv("MVR_BUILD_DEVICE");
  return e ? std::string(e) : std::string("auto");
}

class CompressResidualCode {
 public:
  uint32_t n_centroid_ = 0, vec_dim_ = 0, n_bit_ = 0, n_val_per_byte_ = 0, n_packed_val_per_vec_ = 0;
  uint32_t n_cutoff_ = 0, cut_stride_ = 0;
  std::vector<float> centroid_l_, cutoff_l_, weight_l_;

  CompressResidualCode(const arr_f32 &C, const arr_f32 &cut, const arr_f32 &w, uint32_t n_bit) {
    check_nbit(n_bit);
    if (C.ndim() != 2) throw py::value_error("centroid_l must be [n_centroid, dim]");
    n_centroid_ = (uint32_t)C.shape(0);
    vec_dim_ = (uint32_t)C.shape(1);
    centroid_l_.assign(C.data(), C.data() + (size_t)n_centroid_ * vec_dim_);
    n_bit_ = n_bit;
    const uint32_t nq = 1u << n_bit;
    n_cutoff_ = nq - 1;
    if (cut.ndim() == 1) {
      if ((uint32_t)cut.shape(0) != n_cutoff_) throw py::value_error("cutoff_l must hold 2^n_bit - 1 values");
      cut_stride_ = 0;
    } else if (cut.ndim() == 2) {
      if ((uint32_t)cut.shape(0) != vec_dim_ || (uint32_t)cut.shape(1) != n_cutoff_)
        throw py::value_error("per-dimension cutoff_l must be [dim, 2^n_bit - 1]");
      cut_stride_ = n_cutoff_;
    } else {
      throw py::value_error("cutoff_l must be 1-D or 2-D");
    }
    cutoff_l_.assign(cut.data(), cut.data() + cut.size());
    weight_l_.assign(w.data(), w.data() + w.size());
    n_val_per_byte_ = 8 / n_bit_;
    n_packed_val_per_vec_ = (vec_dim_ + n_val_per_byte_ - 1) / n_val_per_byte_;
  }

  std::string device_ = "cpu";  // where the last compute_residual_code ran

  py::tuple compute_residual_code(const arr_f32 &x, const arr_u32 &codes) {
    if (x.ndim() != 2 || (uint32_t)x.shape(1) != vec_dim_ || codes.ndim() != 1 || codes.shape(0) != x.shape(0))
      throw py::value_error("expected vec_l [n, dim] and code_l [n]");
    const size_t n = x.shape(0);
    const float *xp = x.data();
    const uint32_t *kp = codes.data();
    for (size_t i = 0; i < n; ++i)
      if (kp[i] >= n_centroid_) throw py::value_error("code_l holds an id >= number of centroids");
    std::vector<uint8_t> out(n * n_packed_val_per_vec_);
    const std::string mode = build_device_mode();
    if (mode != "cpu" && mode != "gpu" && mode != "auto")
      throw py::value_error("MVR_BUILD_DEVICE must be cpu, gpu or auto");
    const bool shape_ok = n_bit_ == 2 && vec_dim_ == 128;
    if (mode != "cpu" && n > 0) {
      GpuCoder &g = gpu_coder();
      const bool avail = g.compress && g.n_dev > 0;
      if (mode == "gpu" && (!avail || !shape_ok))
        throw std::runtime_error(!avail ? "MVR_BUILD_DEVICE=gpu but no GPU / libmvr.so is available"
                                        : "the GPU residual coder needs n_bit 2 and dim 128");
      if (avail && shape_ok) {
        int st;
        {
          py::gil_scoped_release nogil;
          st = g.compress(xp, kp, n, centroid_l_.data(), n_centroid_, vec_dim_, cutoff_l_.data(),
                          cut_stride_ ? 1 : 0, out.data(), 0);
        }
        if (st != 0) throw std::runtime_error(std::string("GPU residual coder: ") + g.last_error());
        device_ = "gpu";
        return py::make_tuple(out);
      }
    }
    device_ = "cpu";
    const uint8_t mask = (uint8_t)((1u << n_bit_) - 1);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
      const float *xi = xp + i * vec_dim_;
      const float *ci = centroid_l_.data() + (size_t)kp[i] * vec_dim_;
      for (uint32_t p = 0; p < n_packed_val_per_vec_; ++p) {
        uint8_t byte = 0;
        for (uint32_t v = 0; v < n_val_per_byte_; ++v) {
          const uint32_t d = p * n_val_per_byte_ + v;
          uint32_t b = 0;
          if (d < vec_dim_)
            b = bucket(cutoff_l_.data() + (cut_stride_ ? d * cut_stride_ : 0), n_cutoff_, xi[d] - ci[d]);
          byte = (uint8_t)((byte << n_bit_) | (b & mask));
        }
        out[i * n_packed_val_per_vec_ + p] = byte;
      }
    }
    return py::make_tuple(out);
  }
};

class IVFIndex {
 public:
  size_t n_vecs_ = 0;
  uint32_t n_item_ = 0, n_centroid_ = 0;
  std::vector<uint32_t> code_l_, item_n_vecs_l_, ivf_size_l_, ivf_;
  std::vector<size_t> ivf_offset_l_;

  IVFIndex(std::vector<uint32_t> codes, std::vector<uint32_t> lens, size_t n_vec, uint32_t n_item,
           uint32_t n_centroid)
      : n_vecs_(n_vec), n_item_(n_item), n_centroid_(n_centroid), code_l_(std::move(codes)),
        item_n_vecs_l_(std::move(lens)) {
    if (code_l_.size() != n_vecs_) throw py::value_error("len(vq_code_l) != n_vec");
    if (item_n_vecs_l_.size() != n_item_) throw py::value_error("len(item_n_vec_l) != n_item");
  }

  void build() {
    size_t total = 0;
    for (uint32_t v : item_n_vecs_l_) total += v;
    if (total != n_vecs_) throw py::value_error("sum(item_n_vec_l) != n_vec");
    for (uint32_t c : code_l_)
      if (c >= n_centroid_) throw py::value_error("vq_code_l holds an id >= n_centroid");
    // pass 1: per-document unique centroid ids (sorted), counted per list
    ivf_size_l_.assign(n_centroid_, 0);
    std::vector<uint32_t> uniq;
    size_t off = 0;
    std::vector<std::vector<uint32_t>> doc_sets(n_item_);
    for (uint32_t d = 0; d < n_item_; ++d) {
      uniq.assign(code_l_.begin() + off, code_l_.begin() + off + item_n_vecs_l_[d]);
      std::sort(uniq.begin(), uniq.end());
      uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
      for (uint32_t c : uniq) ivf_size_l_[c]++;
      doc_sets[d] = uniq;
      off += item_n_vecs_l_[d];
    }
    ivf_offset_l_.assign(n_centroid_, 0);
    size_t n_ele = 0;
    for (uint32_t c = 0; c < n_centroid_; ++c) { ivf_offset_l_[c] = n_ele; n_ele += ivf_size_l_[c]; }
    ivf_.assign(n_ele, 0);
    std::vector<size_t> fill(ivf_offset_l_);
    for (uint32_t d = 0; d < n_item_; ++d)
      for (uint32_t c : doc_sets[d]) ivf_[fill[c]++] = d;  // docs ascending per list
  }

  void save(const std::string &location) const {
    std::ofstream out(location, std::ios::binary);
    if (!out) throw py::value_error("cannot write " + location);
    const size_t n_ele = ivf_.size();
    out.write(reinterpret_cast<const char *>(&n_centroid_), 4);
    out.write(reinterpret_cast<const char *>(&n_ele), 8);
    out.write(reinterpret_cast<const char *>(ivf_offset_l_.data()), (std::streamsize)n_centroid_ * 8);
    out.write(reinterpret_cast<const char *>(ivf_size_l_.data()), (std::streamsize)n_centroid_ * 4);
    out.write(reinterpret_cast<const char *>(ivf_.data()), (std::streamsize)n_ele * 4);
  }
};

}  // namespace

PYBIND11_MODULE(GPGPUBuildIndexPackGPU, m) {
  m.def("compute_quantized_scalar", &compute_quantized_scalar, py::arg("item_vec_l"),
        py::arg("centroid_l"), py::arg("code_l"), py::arg("n_bit"));
  m.def("compute_quantized_scalar_per_dim", &compute_quantized_scalar_per_dim, py::arg("item_vec_l"),
        py::arg("centroid_l"), py::arg("code_l"), py::arg("n_bit"));
  py::class_<CompressResidualCode>(m, "CompressResidualCode")
      .def(py::init<const arr_f32 &, const arr_f32 &, const arr_f32 &, uint32_t>(),
           py::arg("centroid_l"), py::arg("cutoff_l"), py::arg("weight_l"), py::arg("n_bit"))
      .def_readonly("n_val_per_vec", &CompressResidualCode::n_packed_val_per_vec_)
      .def_readonly("device", &CompressResidualCode::device_)
      .def("compute_residual_code", &CompressResidualCode::compute_residual_code, py::arg("vec_l"),
           py::arg("code_l"));
  py::class_<IVFIndex>(m, "IVFIndex")
      .def(py::init<std::vector<uint32_t>, std::vector<uint32_t>, size_t, uint32_t, uint32_t>(),
           py::arg("vq_code_l"), py::arg("item_n_vec_l"), py::arg("n_vec"), py::arg("n_item"),
           py::arg("n_centroid"))
      .def("build", &IVFIndex::build)
      .def("save", &IVFIndex::save, py::arg("location"));
}
