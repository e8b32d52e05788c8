// IGPGPUPP / IGPGPUBasic command-line drop-ins (same flags, input files and
// output files as the reference executables), calling libmvr through its C
// ABI only.
//
//   flags          src/app/IGPGPUPP.cpp:15-30 (IGPGPUBasic.cpp:15-30 without -n_thread)
//   input paths    src/app/IGPGPUPP.cpp:33-73, src/include/util/CPPIOUtil.hpp:11-39
//   outputs        src/app/IGPGPUPP.cpp:144-215 (IGPGPUBasic.cpp:140-204)
//   exit codes     1 on a flag error (IGPGPUPP.cpp:86-89) or a missing queries
//                  file (CPPIOUtil.hpp:19-22); 255 (exit(-1)) when probe_topk <
//                  topk or nprobe > n_centroid (IGPGPUPP.hpp:256-259,
//                  IGPGPURetrievalPP.hpp:80-85); 1 on any library error.
//
// Flag grammar follows deps/cmd_line_parser/include/parser.hpp:20-60: each
// flag once, "-name value", unknown flags are errors, every required flag must
// appear.  Extra optional flags: -precision fp32|fp32_mfma|bf16|fp8, -rescore N,
// -batch N, -residual_on_host true|false.  The dataset root /home/{username}/Dataset/multi-vector-retrieval-GPU
// can be replaced with the MVR_DATASET_ROOT environment variable.
#include "mvr.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#if defined(MVR_VARIANT_BASIC)
static const bool kBasic = true;
#else
static const bool kBasic = false;
#endif

namespace {

struct Flag {
  std::string name, descr;
  bool required;
  std::string value;
  bool seen = false;
};

struct Parser {
  std::vector<Flag> flags;
  void add(const std::string &name, const std::string &descr, bool required) {
    flags.push_back({name, descr, required, "", false});
  }
  void help(const char *argv0) const {
    std::cerr << "Usage: " << argv0;
    for (const auto &f : flags) std::cerr << " [" << f.name << " value]";
    std::cerr << "\n";
    for (const auto &f : flags)
      std::cerr << "\t" << f.name << ": " << (f.required ? "REQUIRED: " : "") << f.descr << "\n";
  }
  bool parse(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
      const std::string a(argv[i]);
      if (a == "-h" || a == "--help") { help(argv[0]); return false; }
      auto it = std::find_if(flags.begin(), flags.end(), [&](const Flag &f) { return f.name == a; });
      if (it == flags.end()) {
        std::cerr << "== error: shorthand '" << a << "' not found" << std::endl;
        help(argv[0]);
        return false;
      }
      if (it->seen) {
        std::cerr << "== error: shorthand '" << a << "' already parsed" << std::endl;
        help(argv[0]);
        return false;
      }
      if (++i == argc) { help(argv[0]); return false; }
      it->value = argv[i];
      it->seen = true;
    }
    for (const auto &f : flags)
      if (f.required && !f.seen) { help(argv[0]); return false; }
    return true;
  }
  const std::string &get(const std::string &name) const {
    for (const auto &f : flags)
      if (f.name == name) return f.value;
    static const std::string empty;
    return empty;
  }
  uint32_t u32(const std::string &name) const { return (uint32_t)std::strtol(get(name).c_str(), nullptr, 10); }
  bool boolean(const std::string &name) const {
    const std::string &v = get(name);
    if (v == "true") return true;
    if (v == "false") return false;
    return std::strtol(v.c_str(), nullptr, 10) != 0;
  }
};

bool read_npy_f32(const std::string &path, std::vector<float> &out, std::vector<size_t> &shape) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  char magic[8];
  in.read(magic, 8);
  if (!in || std::memcmp(magic, "\x93NUMPY", 6) != 0) return false;
  uint32_t hlen = 0;
  if (magic[6] == 1) {
    unsigned char b[2];
    in.read(reinterpret_cast<char *>(b), 2);
    hlen = b[0] | (b[1] << 8);
  } else {
    unsigned char b[4];
    in.read(reinterpret_cast<char *>(b), 4);
    hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
  }
  std::string h(hlen, '\0');
  in.read(&h[0], hlen);
  if (h.find("'<f4'") == std::string::npos) return false;
  const size_t l = h.find('(', h.find("'shape'")), r = h.find(')', l);
  std::stringstream ss(h.substr(l + 1, r - l - 1));
  std::string tok;
  shape.clear();
  size_t n = 1;
  while (std::getline(ss, tok, ','))
    if (tok.find_first_of("0123456789") != std::string::npos) {
      shape.push_back(std::stoull(tok));
      n *= shape.back();
    }
  out.resize(n);
  in.read(reinterpret_cast<char *>(out.data()), n * 4);
  return (bool)in;
}

}  // namespace

int main(int argc, char **argv) {
  Parser p;
  p.add("-username", "Username", true);
  p.add("-dataset", "Dataset", true);
  p.add("-method-name", "Method name", true);
  p.add("-num-centroid", "Number of centroid of the index", true);
  p.add("-num-bit", "Number of bits in the SQ index", true);
  p.add("-topk", "Number of nearest neighbours.", true);
  p.add("-nprobe", "number of probe", true);
  p.add("-probe_topk", "topk of the probe", true);
  if (!kBasic) p.add("-n_thread", "number of cpu thread", true);
  p.add("-is_profile", "is profile the result", true);
  p.add("-precision", "fp32 (default), fp32_mfma, bf16 or fp8 centroid scoring and MaxSim", false);
  p.add("-rescore", "bf16/fp8: rescore this many docs in fp32 (default 200)", false);
  p.add("-batch", "queries per device pass (default 1024)", false);
  p.add("-residual_on_host", "keep codes and residuals in host memory (default false)", false);
  if (!p.parse(argc, argv)) {
    std::cerr << "Error parsing command line" << std::endl;
    return 1;
  }
  const std::string user = p.get("-username"), ds = p.get("-dataset"), method = p.get("-method-name");
  const uint32_t n_centroid = p.u32("-num-centroid"), n_bit = p.u32("-num-bit");
  const uint32_t topk = p.u32("-topk"), nprobe = p.u32("-nprobe"), probe_topk = p.u32("-probe_topk");
  const uint32_t n_thread = kBasic ? 1 : p.u32("-n_thread");
  const bool is_profile = p.boolean("-is_profile");
  std::string prec = p.get("-precision");
  if (prec.empty()) prec = "fp32";
  uint32_t precision = MVR_FP32;
  if (prec == "bf16") precision = MVR_BF16;
  else if (prec == "fp8") precision = MVR_FP8;
  else if (prec == "fp32_mfma") precision = MVR_FP32_MFMA;
  else if (prec != "fp32") { std::cerr << "Error parsing command line: bad -precision" << std::endl; return 1; }
  const uint32_t rescore = p.get("-rescore").empty() ? 200 : p.u32("-rescore");
  const uint32_t batch = p.get("-batch").empty() ? 1024 : std::max(1u, p.u32("-batch"));
  const uint32_t ix_flags =
      !p.get("-residual_on_host").empty() && p.boolean("-residual_on_host") ? MVR_INDEX_RESIDUAL_ON_HOST : 0u;

  const char *env_root = std::getenv("MVR_DATASET_ROOT");
  const std::string root = env_root ? std::string(env_root)
                                    : "/home/" + user + "/Dataset/multi-vector-retrieval-GPU";
  char suffix[256];
  std::snprintf(suffix, sizeof(suffix), "n_centroid_%u-n_bit_%u", n_centroid, n_bit);
  const std::string index_path = root + "/Index/" + ds + "/" + method + "-" + suffix;
  const std::string query_path = root + "/Embedding/" + ds + "/query_embedding.npy";
  const std::string answer_path = root + "/Result/answer";
  const std::string qid_path = root + "/RawData/" + ds + "/document/queries.dev.tsv";

  mvr_index *ix = nullptr;
  if (mvr_index_load_ex(index_path.c_str(), 0, ix_flags, &ix) != MVR_OK) {
    std::cerr << "error: " << mvr_last_error() << std::endl;
    return 1;
  }
  uint32_t ix_nc = 0;
  mvr_index_info(ix, nullptr, nullptr, &ix_nc, nullptr, nullptr);
  if (ix_nc != n_centroid) {
    std::cerr << "error: index has " << ix_nc << " centroids, -num-centroid says " << n_centroid << std::endl;
    mvr_index_free(ix);
    return 1;
  }
  std::vector<float> Q;
  std::vector<size_t> qshape;
  if (!read_npy_f32(query_path, Q, qshape) || qshape.size() != 3 || qshape[2] != 128) {
    std::cerr << "error: cannot read float32 [n_query, n_vec, 128] " << query_path << std::endl;
    mvr_index_free(ix);
    return 1;
  }
  const uint32_t n_query = (uint32_t)qshape[0], query_n_vec = (uint32_t)qshape[1];
  if (probe_topk < topk) {
    std::cerr << "the number of refined topk is smaller than the number of returned topk, program exit" << std::endl;
    mvr_index_free(ix);
    return 255;
  }
  if (nprobe > n_centroid) {
    std::cerr << "nprobe is larger than n_centroid, nprobe " << nprobe << ", n_centroid " << n_centroid << std::endl;
    mvr_index_free(ix);
    return 255;
  }
  mvr_search_params prm{};
  prm.nprobe = nprobe;
  prm.probe_topk = probe_topk;
  prm.topk = topk;
  prm.precision = precision;
  prm.rescore = (precision == MVR_FP32 || precision == MVR_FP32_MFMA) ? 0 : std::min(rescore, probe_topk);
  prm.clip_zero = kBasic ? 1 : 0;
  prm.batch = batch;

  std::vector<uint32_t> ids((size_t)n_query * topk);
  std::vector<float> scores((size_t)n_query * topk);
  std::vector<double> search_time(n_query);
  // warm-up on the first pass worth of queries (IGPGPUPP.hpp:276-279 warms up on 10)
  {
    const uint32_t nw = std::min(n_query, batch);
    if (mvr_search(ix, Q.data(), nw, query_n_vec, nullptr, nullptr, &prm, ids.data(), scores.data(), 0,
                   nullptr, nullptr, nullptr) != MVR_OK) {
      std::cerr << "error: " << mvr_last_error() << std::endl;
      mvr_index_free(ix);
      return 1;
    }
  }
  const auto t_all = std::chrono::steady_clock::now();
  for (uint32_t q0 = 0; q0 < n_query; q0 += batch) {
    const uint32_t nb = std::min(batch, n_query - q0);
    const auto t0 = std::chrono::steady_clock::now();
    if (mvr_search(ix, Q.data() + (size_t)q0 * query_n_vec * 128, nb, query_n_vec, nullptr, nullptr,
                   &prm, ids.data() + (size_t)q0 * topk, scores.data() + (size_t)q0 * topk, 0,
                   nullptr, nullptr, nullptr) != MVR_OK) {
      std::cerr << "error: " << mvr_last_error() << std::endl;
      mvr_index_free(ix);
      return 1;
    }
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (uint32_t i = 0; i < nb; ++i) search_time[q0 + i] = dt;  // every query of a pass waits for the pass
  }
  const double batch_query_time =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_all).count();
  mvr_index_free(ix);

  std::vector<uint32_t> qids;
  {
    std::ifstream f(qid_path);
    if (!f.is_open()) {
      std::cerr << "Error: Could not open the file!" << std::endl;
      return 1;
    }
    std::string line;
    while (std::getline(f, line)) {
      if (line.empty()) continue;
      const size_t tab = line.find('\t');
      if (tab == std::string::npos) {
        std::cerr << "Warning: Invalid format in line: " << line << std::endl;
        continue;
      }
      qids.push_back((uint32_t)std::stoul(line.substr(0, tab)));
    }
  }
  if (qids.size() < n_query) {
    std::cerr << "error: " << qid_path << " lists " << qids.size() << " queries, need " << n_query << std::endl;
    return 1;
  }
  char retr[256];
  if (kBasic) std::snprintf(retr, sizeof(retr), "nprobe_%u-probe_topk_%u", nprobe, probe_topk);
  else std::snprintf(retr, sizeof(retr), "nprobe_%u-probe_topk_%u-n_thread_%u", nprobe, probe_topk, n_thread);
  const std::string prof = is_profile ? "-profile" : "";
  const std::string stem = ds + "-" + method;
  const std::string tail = std::string(suffix) + "-" + retr + prof;
  const std::string ans = answer_path + "/" + stem + "-top" + std::to_string(topk) + "-" + tail + ".tsv";
  const std::string perf = answer_path + "/" + stem + "-performance-top" + std::to_string(topk) + "-" + tail;
  {
    std::ofstream out(ans);
    if (!out) { std::cerr << "error: cannot write " << ans << std::endl; return 1; }
    for (uint32_t q = 0; q < n_query; ++q)
      for (uint32_t r = 0; r < topk; ++r)
        out << qids[q] << "\t" << ids[(size_t)q * topk + r] << "\t" << r + 1 << "\t"
            << scores[(size_t)q * topk + r] << "\n";
  }
  {
    std::ofstream out(perf + ".tsv");
    if (kBasic) {
      // the basic executable writes ten per-stage columns; the batched device
      // pass has no per-query stage split, so all but search_time are 0
      out << "search_time\treset_time\ttransmit_query_time\tcand_gen_time\tfilter_time\tfilter_transmit_time\t"
          << "filter_compute_time\trerank_time\trerank_transmit_time\trerank_compute_time\n";
      for (uint32_t q = 0; q < n_query; ++q) out << search_time[q] << "\t0\t0\t0\t0\t0\t0\t0\t0\t0\n";
    } else {
      out << "search_time\n";
      for (uint32_t q = 0; q < n_query; ++q) out << search_time[q] << "\n";
    }
  }
  {
    std::ofstream out(perf + ".txt");
    out << "batch_query_time(s)\t" << batch_query_time << "\n";
    out << "n_query\t" << n_query << "\n";
    out << "topk\t" << topk << "\n";
  }
  return 0;
}
