// Matrix Market reader (host, OpenMP): the reference's input format, parsed into a 0-based edge
// list for rlc_graph_from_edges.
//
// Rules of the reference reader (inc/mtx.hxx:30-55 header, :151-187 body): comment lines start
// with '%'; the banner "%%MatrixMarket matrix coordinate <field> <symmetry>" makes the file
// symmetric when <symmetry> is "symmetric" or "skew-symmetric"; the first other line holds
// "rows cols size"; vertices are 1..max(rows, cols) (vertex 0 does not exist, inc/mtx.hxx:213);
// each body line "u v [w]" gives the edge (u, v) and, in a symmetric file, (v, u) right after it;
// the weight is read only for weighted input, and a weight of 0 (or a missing one) is 1. Here
// ids are shifted to 0..n-1.
#include <omp.h>

#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"
#include "rlc.h"

namespace {

const char *skip_spaces(const char *p, const char *e) {
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  return p;
}

bool parse_u64(const char *&p, const char *e, uint64_t &x) {
  p = skip_spaces(p, e);
  if (p >= e || *p < '0' || *p > '9') return false;
  uint64_t v = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    v = v * 10 + (uint64_t)(*p - '0');
    ++p;
  }
  x = v;
  return true;
}

}  // namespace

extern "C" {

rlc_status rlc_mtx_read(const char *path, int32_t weighted, rlc_edges *out) {
  rlc::clear_error();
  if (!path || !out) return rlc::fail(RLC_ERR_INVALID_ARGUMENT, "rlc_mtx_read: null argument");
  std::memset(out, 0, sizeof(*out));
  std::ifstream f(path, std::ios::binary);
  if (!f) return rlc::fail(RLC_ERR_INVALID_ARGUMENT, std::string("rlc_mtx_read: cannot open ") + path);
  std::string text;
  try {
    f.seekg(0, std::ios::end);
    text.resize((size_t)f.tellg());
    f.seekg(0);
    f.read(&text[0], (std::streamsize)text.size());
  } catch (...) {
    return rlc::fail(RLC_ERR_OUT_OF_MEMORY, "rlc_mtx_read: cannot hold the file");
  }
  text.push_back('\n');
  const char *p = text.data(), *end = text.data() + text.size();
  // header: comments, banner, then "rows cols size"
  bool symmetric = false, banner = false;
  uint64_t rows = 0, cols = 0, size = 0;
  for (;;) {
    if (p >= end) return rlc::fail(RLC_ERR_INVALID_GRAPH, "rlc_mtx_read: no size line");
    const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
    std::string line(p, nl);
    p = nl + 1;
    if (!line.empty() && line[0] == '%') {
      if (line.rfind("%%", 0) == 0) {
        std::istringstream s(line);
        std::string h0, h1, h2, h3, h4;
        s >> h0 >> h1 >> h2 >> h3 >> h4;
        if (h1 != "matrix" || h2 != "coordinate")
          return rlc::fail(RLC_ERR_INVALID_GRAPH, "rlc_mtx_read: not a coordinate matrix");
        symmetric = h4 == "symmetric" || h4 == "skew-symmetric";
        banner = true;
      }
      continue;
    }
    std::istringstream s(line);
    if (!(s >> rows >> cols >> size)) return rlc::fail(RLC_ERR_INVALID_GRAPH, "rlc_mtx_read: bad size line");
    break;
  }
  if (!banner) return rlc::fail(RLC_ERR_INVALID_GRAPH, "rlc_mtx_read: no %%MatrixMarket banner");
  const uint64_t n = rows > cols ? rows : cols;
  if (n >= 0xFFFFFFFFull) return rlc::fail(RLC_ERR_INVALID_GRAPH, "rlc_mtx_read: more than 2^32-2 vertices");
  // body: cut at line boundaries into one piece per thread, count, then parse in place
  const int T = omp_get_max_threads();
  std::vector<const char *> cut(T + 1);
  cut[0] = p;
  for (int t = 1; t < T; ++t) {
    const char *q = p + (size_t)(end - p) * t / T;
    if (q < cut[t - 1]) q = cut[t - 1];
    const char *nl = (const char *)std::memchr(q, '\n', (size_t)(end - q));
    cut[t] = nl ? nl + 1 : end;
  }
  cut[T] = end;
  auto is_entry = [](const char *a, const char *b) {
    a = skip_spaces(a, b);
    return a < b && *a != '%';
  };
  std::vector<uint64_t> lines(T + 1, 0);
#pragma omp parallel for schedule(static, 1)
  for (int t = 0; t < T; ++t) {
    uint64_t c = 0;
    for (const char *a = cut[t]; a < cut[t + 1];) {
      const char *nl = (const char *)std::memchr(a, '\n', (size_t)(cut[t + 1] - a));
      if (!nl) nl = cut[t + 1];
      c += is_entry(a, nl) ? 1 : 0;
      a = nl + 1;
    }
    lines[t + 1] = c;
  }
  for (int t = 0; t < T; ++t) lines[t + 1] += lines[t];
  const uint64_t L = lines[T], per = symmetric ? 2 : 1, count = L * per;
  uint32_t *src = (uint32_t *)std::malloc(std::max<uint64_t>(count, 1) * 4);
  uint32_t *dst = (uint32_t *)std::malloc(std::max<uint64_t>(count, 1) * 4);
  float *wei = weighted ? (float *)std::malloc(std::max<uint64_t>(count, 1) * 4) : nullptr;
  if (!src || !dst || (weighted && !wei)) {
    std::free(src);
    std::free(dst);
    std::free(wei);
    return rlc::fail(RLC_ERR_OUT_OF_MEMORY, "rlc_mtx_read: edge arrays");
  }
  std::vector<uint64_t> bad(T, 0);  // 1-based line of the first bad entry of each piece (0: none)
#pragma omp parallel for schedule(static, 1)
  for (int t = 0; t < T; ++t) {
    uint64_t k = lines[t];
    for (const char *a = cut[t]; a < cut[t + 1];) {
      const char *nl = (const char *)std::memchr(a, '\n', (size_t)(cut[t + 1] - a));
      if (!nl) nl = cut[t + 1];
      if (is_entry(a, nl)) {
        const char *q = a;
        uint64_t u = 0, v = 0;
        double w = 0;
        bool ok = parse_u64(q, nl, u) && parse_u64(q, nl, v) && u >= 1 && v >= 1 && u <= n && v <= n;
        if (ok && weighted) {
          q = skip_spaces(q, nl);
          if (q < nl) {
            char *qe = nullptr;
            w = std::strtod(q, &qe);  // the line ends with '\n', so strtod stops inside the buffer
            ok = qe != q;
          }
        }
        if (!ok) {
          if (!bad[t]) bad[t] = k + 1;
        } else {
          const float x = (float)(w != 0 ? w : 1.0);
          src[per * k] = (uint32_t)(u - 1);
          dst[per * k] = (uint32_t)(v - 1);
          if (wei) wei[per * k] = x;
          if (symmetric) {
            src[per * k + 1] = (uint32_t)(v - 1);
            dst[per * k + 1] = (uint32_t)(u - 1);
            if (wei) wei[per * k + 1] = x;
          }
        }
        ++k;
      }
      a = nl + 1;
    }
  }
  for (int t = 0; t < T; ++t)
    if (bad[t]) {
      std::free(src);
      std::free(dst);
      std::free(wei);
      return rlc::fail(RLC_ERR_INVALID_GRAPH, "rlc_mtx_read: bad entry " + std::to_string(bad[t]) +
                                                  " (expected \"u v [w]\" with 1 <= u, v <= " + std::to_string(n) + ")");
    }
  (void)size;
  out->n = n;
  out->count = count;
  out->src = src;
  out->dst = dst;
  out->weights = wei;
  out->symmetric = symmetric ? 1 : 0;
  return RLC_OK;
}

void rlc_edges_release(rlc_edges *e) {
  if (!e) return;
  std::free(e->src);
  std::free(e->dst);
  std::free(e->weights);
  std::memset(e, 0, sizeof(*e));
}

}  // extern "C"
