// Minimal .npy reader/writer (format version 1.0/2.0, little-endian, C order)
// for the reference index files: centroid_l.npy (f4), vq_code_l.npy (i4/u4),
// weight_l.npy (f4), residual_code_l.npy (u1), item_n_vec_l.npy (u4),
// query_embedding.npy (f4).  Replaces the vendored cnpy (deps/cnpy/cnpy.h:24-73).
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace mvr {

struct NpyArray {
  std::string descr;            // e.g. "<f4"
  std::vector<size_t> shape;
  size_t word_size = 0;
  std::vector<uint8_t> data;
  size_t numel() const {
    size_t n = 1;
    for (size_t s : shape) n *= s;
    return n;
  }
  template <class T> const T *as() const { return reinterpret_cast<const T *>(data.data()); }
};

// Parses the header of path into out (no data read) and leaves *fp open at
// the first data byte.  Returns an empty string on success, else an error
// message (and *fp closed).
inline std::string npy_open(const std::string &path, NpyArray &out, FILE **fp) {
  *fp = nullptr;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return "cannot open " + path;
  unsigned char magic[8];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0) {
    std::fclose(f);
    return "not a .npy file: " + path;
  }
  uint32_t hlen = 0;
  if (magic[6] == 1) {
    unsigned char b[2];
    if (std::fread(b, 1, 2, f) != 2) { std::fclose(f); return "truncated header: " + path; }
    hlen = b[0] | (b[1] << 8);
  } else {
    unsigned char b[4];
    if (std::fread(b, 1, 4, f) != 4) { std::fclose(f); return "truncated header: " + path; }
    hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
  }
  std::string h(hlen, '\0');
  if (std::fread(&h[0], 1, hlen, f) != hlen) { std::fclose(f); return "truncated header: " + path; }
  auto field = [&](const char *key) -> std::string {
    size_t p = h.find(key);
    if (p == std::string::npos) return "";
    p = h.find(':', p);
    return h.substr(p + 1);
  };
  std::string d = field("'descr'");
  size_t q1 = d.find('\''), q2 = d.find('\'', q1 + 1);
  if (q1 == std::string::npos || q2 == std::string::npos) { std::fclose(f); return "bad descr: " + path; }
  out.descr = d.substr(q1 + 1, q2 - q1 - 1);
  const std::string fo = field("'fortran_order'");
  if (fo.substr(0, fo.find(',')).find("True") != std::string::npos) {
    std::fclose(f);
    return "fortran order not supported: " + path;
  }
  std::string s = field("'shape'");
  size_t l = s.find('('), r = s.find(')');
  out.shape.clear();
  std::string dims = s.substr(l + 1, r - l - 1);
  size_t pos = 0;
  while (pos < dims.size()) {
    while (pos < dims.size() && (dims[pos] == ' ' || dims[pos] == ',')) ++pos;
    if (pos >= dims.size()) break;
    size_t e = pos;
    while (e < dims.size() && dims[e] >= '0' && dims[e] <= '9') ++e;
    if (e == pos) break;
    out.shape.push_back(std::stoull(dims.substr(pos, e - pos)));
    pos = e;
  }
  if (out.descr.size() < 3 || (out.descr[0] == '>' && out.descr[2] != '1')) {
    std::fclose(f);
    return "unsupported dtype " + out.descr + ": " + path;
  }
  out.word_size = std::stoul(out.descr.substr(2));
  *fp = f;
  return "";
}

// Returns an empty string on success, else an error message.
inline std::string npy_load(const std::string &path, NpyArray &out) {
  FILE *f = nullptr;
  std::string err = npy_open(path, out, &f);
  if (!err.empty()) return err;
  const size_t bytes = out.numel() * out.word_size;
  out.data.resize(bytes);
  if (bytes && std::fread(out.data.data(), 1, bytes, f) != bytes) {
    std::fclose(f);
    return "truncated data: " + path;
  }
  std::fclose(f);
  return "";
}

}  // namespace mvr
