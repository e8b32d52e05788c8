// BruteForceProgressive: pybind11 module with the reference's Python surface
// (src/app/BruteForceProgressive.cpp:16-33, src/impl/BruteForceProgressive.hpp):
//   DocRetrieval(query_l [n_q, T, dim], vec_dim, precision="fp32")  ("fp32_mfma": the
//     dots on FP32 MFMA, within 1e-5 relative; not in the reference's signature)
//   .computeScore(item_l: list of [L_i, dim], itemID_l)   all-pairs MaxSim on the GPU
//   .searchKNN(topk) -> (scores [n_q, topk], ids [n_q, topk])
//   .finish_compute()
//   DocRetrieval.merge_result(final_distance_l, final_id_l, distance_l, id_l)  (static)
// It is what script/data/groundtruth.py:40-109 drives to build groundtruth.
// The scores come from libmvr's mvr_maxsim_scores_host (loaded from this
// module's directory; torch is imported first when installed so both share
// one HIP runtime); there is no host fallback, as the reference has none.
// Result order follows the reference's min-heap pops: ascending score, the
// unfilled tail (topk > items) left at (-FLT_MAX, 0xFFFFFFFF).  Ties, which
// the reference's heap leaves unordered, keep the earlier item (lower
// position) in the top-k and list it after the later one.
#include <dlfcn.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace py = pybind11;
using arr_f32 = py::array_t<float, py::array::c_style | py::array::forcecast>;
using arr_u32 = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>;

namespace {

struct MaxSimLib {
  int (*scores)(const float *, uint32_t, uint32_t, uint32_t, const float *, const uint32_t *, uint32_t, float *,
                int, uint32_t) = nullptr;
  const char *(*last_error)() = nullptr;
  bool tried = false;
};

MaxSimLib &maxsim_lib() {
  static MaxSimLib g;
  if (g.tried) return g;
  g.tried = true;
  try {
    py::module_::import("torch");
  } catch (...) {
  }
  Dl_info info;
  if (!dladdr(reinterpret_cast<void *>(&maxsim_lib), &info) || !info.dli_fname) return g;
  std::string dir(info.dli_fname);
  dir = dir.substr(0, dir.find_last_of('/') + 1);
  void *h = dlopen((dir + "libmvr.so").c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h) return g;
  g.scores = reinterpret_cast<decltype(g.scores)>(dlsym(h, "mvr_maxsim_scores_host_ex"));
  g.last_error = reinterpret_cast<const char *(*)()>(dlsym(h, "mvr_maxsim_last_error"));
  return g;
}

// (score, position) with "better" = higher score, then lower position
struct Cand {
  float s;
  uint32_t pos;
};
inline bool better(const Cand &a, const Cand &b) { return a.s > b.s || (a.s == b.s && a.pos < b.pos); }

// The best k of cand (k <= cand.size()) written ascending (worst first).
void topk_ascending(std::vector<Cand> &cand, uint32_t k, Cand *out) {
  std::partial_sort(cand.begin(), cand.begin() + k, cand.end(), better);
  for (uint32_t r = 0; r < k; ++r) out[r] = cand[k - 1 - r];
}

class DocRetrieval {
 public:
  uint32_t n_query_ = 0, query_n_vecs_ = 0, vec_dim_ = 0, n_item_ = 0;
  uint32_t precision_ = 0;  // mvr_precision: MVR_FP32 (0) or MVR_FP32_MFMA (3)
  std::vector<float> query_;
  std::vector<float> score_;        // [n_query][n_item]
  std::vector<uint32_t> external_;  // [n_item]

  DocRetrieval(const arr_f32 &query_l, uint32_t vec_dim, const std::string &precision) : vec_dim_(vec_dim) {
    if (precision == "fp32") precision_ = 0;
    else if (precision == "fp32_mfma") precision_ = 3;
    else throw std::invalid_argument("precision must be \"fp32\" or \"fp32_mfma\"");
    if (query_l.ndim() != 3) throw std::runtime_error("the vector set list dimension should be 3");
    if ((uint32_t)query_l.shape(2) != vec_dim_)
      throw std::runtime_error("the embedding dimension should equal to the pre-assignment");
    n_query_ = (uint32_t)query_l.shape(0);
    query_n_vecs_ = (uint32_t)query_l.shape(1);
    query_.assign(query_l.data(), query_l.data() + query_l.size());
  }

  void computeScore(const std::vector<arr_f32> &item_l, const std::vector<uint32_t> &itemID_l) {
    if (item_l.size() != itemID_l.size())
      throw std::runtime_error("the item embedding size does not equal to the externalID size");
    std::vector<uint32_t> lens(item_l.size());
    size_t total = 0;
    for (size_t i = 0; i < item_l.size(); ++i) {
      const arr_f32 &a = item_l[i];
      if (a.ndim() != 2) throw std::runtime_error("the item embeddings' should be 2 dimension");
      if ((uint32_t)a.shape(1) != vec_dim_)
        throw std::runtime_error("the item embeddings dimenision should be equal to the preassign");
      lens[i] = (uint32_t)a.shape(0);
      total += lens[i];
    }
    std::vector<float> X(total * vec_dim_);
    size_t o = 0;
    for (const arr_f32 &a : item_l) {
      std::memcpy(X.data() + o, a.data(), sizeof(float) * a.size());
      o += a.size();
    }
    MaxSimLib &L = maxsim_lib();
    if (!L.scores) throw std::runtime_error("libmvr.so (mvr_maxsim_scores_host_ex) is not available next to this module");
    n_item_ = (uint32_t)item_l.size();
    score_.assign((size_t)n_query_ * n_item_, 0.0f);
    external_ = itemID_l;
    int st;
    {
      py::gil_scoped_release nogil;
      st = L.scores(query_.data(), n_query_, query_n_vecs_, vec_dim_, X.data(), lens.data(), n_item_, score_.data(), 0,
                    precision_);
    }
    if (st != 0) throw std::runtime_error(std::string("computeScore: ") + L.last_error());
  }

  py::tuple searchKNN(uint32_t topk) {
    py::array_t<float> dist({(py::ssize_t)n_query_, (py::ssize_t)topk});
    py::array_t<uint32_t> ids({(py::ssize_t)n_query_, (py::ssize_t)topk});
    float *dp = dist.mutable_data();
    uint32_t *ip = ids.mutable_data();
    const uint32_t k = std::min(topk, n_item_);
    {
      py::gil_scoped_release nogil;
#pragma omp parallel for schedule(dynamic)
      for (long long q = 0; q < (long long)n_query_; ++q) {
        std::vector<Cand> cand(n_item_);
        for (uint32_t i = 0; i < n_item_; ++i) cand[i] = Cand{score_[(size_t)q * n_item_ + i], i};
        std::vector<Cand> best(k);
        if (k) topk_ascending(cand, k, best.data());
        for (uint32_t r = 0; r < topk; ++r) {
          dp[(size_t)q * topk + r] = r < k ? best[r].s : -std::numeric_limits<float>::max();
          ip[(size_t)q * topk + r] = r < k ? external_[best[r].pos] : 0xFFFFFFFFu;
        }
      }
    }
    return py::make_tuple(dist, ids);
  }

  void finish_compute() {
    std::vector<float>().swap(score_);
    std::vector<uint32_t>().swap(external_);
    n_item_ = 0;
  }

  static py::tuple merge_result(const arr_f32 &final_d, const arr_u32 &final_id, const arr_f32 &d, const arr_u32 &id) {
    if (final_d.ndim() == 1) return py::make_tuple(d, id);
    if (final_d.ndim() != 2 || final_id.ndim() != 2 || d.ndim() != 2 || id.ndim() != 2 ||
        final_d.shape(0) != d.shape(0) || final_d.shape(1) != d.shape(1) || final_id.shape(0) != d.shape(0) ||
        final_id.shape(1) != d.shape(1) || id.shape(0) != d.shape(0) || id.shape(1) != d.shape(1))
      throw std::runtime_error("merge_result needs four [n_query, topk] arrays");
    const uint32_t n_query = (uint32_t)d.shape(0), topk = (uint32_t)d.shape(1);
    py::array_t<float> od({(py::ssize_t)n_query, (py::ssize_t)topk});
    py::array_t<uint32_t> oi({(py::ssize_t)n_query, (py::ssize_t)topk});
    const float *fd = final_d.data(), *nd = d.data();
    const uint32_t *fi = final_id.data(), *ni = id.data();
    float *odp = od.mutable_data();
    uint32_t *oip = oi.mutable_data();
    {
      py::gil_scoped_release nogil;
#pragma omp parallel for schedule(static)
      for (long long q = 0; q < (long long)n_query; ++q) {
        // the earlier list first: a new entry must be strictly better to enter
        std::vector<Cand> cand(2 * (size_t)topk);
        for (uint32_t r = 0; r < topk; ++r) cand[r] = Cand{fd[(size_t)q * topk + r], r};
        for (uint32_t r = 0; r < topk; ++r) cand[topk + r] = Cand{nd[(size_t)q * topk + r], topk + r};
        std::vector<Cand> best(topk);
        if (topk) topk_ascending(cand, topk, best.data());
        for (uint32_t r = 0; r < topk; ++r) {
          const uint32_t p = best[r].pos;
          odp[(size_t)q * topk + r] = best[r].s;
          oip[(size_t)q * topk + r] = p < topk ? fi[(size_t)q * topk + p] : ni[(size_t)q * topk + p - topk];
        }
      }
    }
    return py::make_tuple(od, oi);
  }
};

}  // namespace

PYBIND11_MODULE(BruteForceProgressive, m) {
  py::class_<DocRetrieval>(m, "DocRetrieval",
                           "Exhaustive MaxSim of a query set against batches of documents (groundtruth).")
      .def(py::init<const arr_f32 &, uint32_t, const std::string &>(), py::arg("query_l"), py::arg("vec_dim"),
           py::arg("precision") = "fp32")
      .def("computeScore", &DocRetrieval::computeScore, py::arg("item_l"), py::arg("itemID_l"))
      .def("searchKNN", &DocRetrieval::searchKNN, py::arg("topk"))
      .def("finish_compute", &DocRetrieval::finish_compute)
      .def_static("merge_result", &DocRetrieval::merge_result, py::arg("final_distance_l"), py::arg("final_id_l"),
                  py::arg("distance_l"), py::arg("id_l"), "Merge the retrieval result by multiple searching");
}
