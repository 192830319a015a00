// Host ingest of the reference text format.
//
// Parsing rules follow Signed_TC/Multicore Implementations/graph_io.h:72-97
// (skip_whitespace / parse_uint / parse_sign) and the line loop of
// load_graph at graph_io.h:137-203: a '\n' or '\r' at a line start is skipped,
// a line starting with '#' or '%' is a comment, otherwise "u ws v ws sign" is
// read and the rest of the line ignored; u == v lines are dropped; n = max id
// + 1 over kept lines.  The file is split into per-thread chunks at newline
// boundaries as graph_io.h:117-133 does, but each chunk is parsed into its own
// slice of the output and the slices are closed up in file order, so a
// dropped self-loop leaves no hole (graph_io.h:147,159-161 reserve one and
// never fill it).
//
// A buffer that starts with the "%%MatrixMarket" banner is read as a
// MatrixMarket coordinate file instead, the input of the reference's GPU
// drivers (GPU Implementations/GlobalMEM/tc_split/signed_tc.cu:21-33, loaded
// there by Gunrock's io::matrix_market_t, which is not part of the reference):
// banner "%%MatrixMarket matrix coordinate <real|double|integer|pattern>
// <general|symmetric>", '%' comment lines, a size line "M N NNZ", then NNZ
// entries "i j [w]" with 1-based i <= M, j <= N.  An entry's sign is + iff its
// weight, read as a float, is > 0 (signed_tc.hxx:317); pattern entries are +.
// n = max(M, N).  Both stored directions of an edge, or one triangle of a
// symmetric file, give the same undirected edge.  Complex, hermitian,
// skew-symmetric and array files, indices out of range and an entry count
// other than NNZ are errors (STC_E_ARG).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/stc_gen.h"

namespace {

// One chunk's entries, written straight into its slice of the output arrays
// (at most one entry per line, so the slice holds the chunk's newline count
// plus one).  Returns the entries kept; max_id gets the largest kept id.
uint64_t parse_range(const char *p, const char *end, uint32_t *ou, uint32_t *ov, int8_t *os,
                     uint32_t &max_id) {
  uint64_t k = 0;
  uint32_t mx = 0;
  while (p < end) {
    if (*p == '\n' || *p == '\r') {
      p++;
      continue;
    }
    if (*p == '#' || *p == '%') {
      while (p < end && *p != '\n') p++;
      continue;
    }
    uint32_t a = 0, b = 0;
    while (p < end && *p >= '0' && *p <= '9') a = a * 10u + (uint32_t)(*p++ - '0');
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    while (p < end && *p >= '0' && *p <= '9') b = b * 10u + (uint32_t)(*p++ - '0');
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    int8_t s = 1;
    if (p < end && *p == '-') {
      s = -1;
      p++;
    } else if (p < end && *p == '+') {
      p++;
    }
    while (p < end && *p >= '0' && *p <= '9') p++;
    while (p < end && *p != '\n') p++;
    if (a == b) continue;
    ou[k] = a;
    ov[k] = b;
    os[k] = s;
    k++;
    const uint32_t hi = a > b ? a : b;
    if (hi > mx) mx = hi;
  }
  max_id = mx;
  return k;
}

uint64_t count_newlines(const char *p, const char *end) {
  uint64_t c = 0;
  while (p < end) {
    const void *q = std::memchr(p, '\n', (size_t)(end - p));
    if (!q) break;
    c++;
    p = (const char *)q + 1;
  }
  return c;
}

// ---- MatrixMarket coordinate files -----------------------------------------
struct MtxChunk {
  std::vector<uint32_t> u, v;
  std::vector<int8_t> s;
  uint64_t entries = 0;
  bool bad = false;
};

const char *next_line(const char *p, const char *end) {
  while (p < end && *p != '\n') p++;
  return p < end ? p + 1 : end;
}

// Lowercase words of one line into out (at most `max` words).
int line_words(const char *p, const char *end, std::string *out, int max) {
  int k = 0;
  while (p < end && *p != '\n' && k < max) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
    if (p >= end || *p == '\n') break;
    std::string w;
    while (p < end && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n')
      w.push_back((char)std::tolower((unsigned char)*p++));
    out[k++] = w;
  }
  return k;
}

bool parse_u64(const char *&p, const char *end, uint64_t &x) {
  while (p < end && (*p == ' ' || *p == '\t')) p++;
  if (p >= end || *p < '0' || *p > '9') return false;
  x = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    if (x > (~0ull - 9) / 10) return false;
    x = x * 10 + (uint64_t)(*p++ - '0');
  }
  return true;
}

void parse_mtx_range(const char *p, const char *end, uint64_t rows, uint64_t cols, bool pattern,
                     MtxChunk &out) {
  char num[64];
  while (p < end) {
    if (*p == '\n' || *p == '\r' || *p == '%') {
      p = next_line(p, end);
      continue;
    }
    const char *line = p;
    uint64_t i = 0, j = 0;
    if (!parse_u64(p, end, i) || !parse_u64(p, end, j) || i < 1 || j < 1 || i > rows ||
        j > cols) {
      // a line of only blanks is not an entry
      const char *q = line;
      while (q < end && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
      if (q < end && *q != '\n') out.bad = true;
      p = next_line(line, end);
      continue;
    }
    int8_t sg = 1;
    if (!pattern) {
      while (p < end && (*p == ' ' || *p == '\t')) p++;
      size_t k = 0;
      while (p < end && k + 1 < sizeof num && *p != ' ' && *p != '\t' && *p != '\r' &&
             *p != '\n')
        num[k++] = *p++;
      num[k] = 0;
      char *e = nullptr;
      const float w = std::strtof(num, &e);
      if (k == 0 || e == num) {
        out.bad = true;
      } else {
        sg = w > 0.0f ? 1 : -1;
      }
    }
    p = next_line(p, end);
    out.entries++;
    if (i == j) continue;
    out.u.push_back((uint32_t)(i - 1));
    out.v.push_back((uint32_t)(j - 1));
    out.s.push_back(sg);
  }
}

int parse_mtx(const char *buf, uint64_t len, int threads, stc_host_graph *g) {
  const char *end = buf + len, *p = buf;
  std::string w[5];
  if (line_words(p, end, w, 5) != 5 || w[0] != "%%matrixmarket" || w[1] != "matrix" ||
      w[2] != "coordinate")
    return STC_E_ARG;
  const bool pattern = w[3] == "pattern";
  if (!pattern && w[3] != "real" && w[3] != "double" && w[3] != "integer") return STC_E_ARG;
  if (w[4] != "general" && w[4] != "symmetric") return STC_E_ARG;
  p = next_line(p, end);
  while (p < end && (*p == '%' || *p == '\n' || *p == '\r')) p = next_line(p, end);
  uint64_t rows = 0, cols = 0, nnz = 0;
  if (!parse_u64(p, end, rows) || !parse_u64(p, end, cols) || !parse_u64(p, end, nnz))
    return STC_E_ARG;
  if (std::max(rows, cols) > 0xFFFFFFFFull) return STC_E_TOO_LARGE;
  p = next_line(p, end);
  const uint64_t rest = (uint64_t)(end - p);
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  if ((uint64_t)nt > rest / 4096 + 1) nt = (int)(rest / 4096 + 1);
  std::vector<const char *> lo(nt), hi(nt);
  const uint64_t chunk = rest / nt;
  for (int t = 0; t < nt; t++) {
    const char *a = p + (uint64_t)t * chunk;
    const char *b = (t == nt - 1) ? end : p + (uint64_t)(t + 1) * chunk;
    if (t > 0)
      while (a < end && *(a - 1) != '\n') a++;
    if (t < nt - 1)
      while (b < end && *(b - 1) != '\n') b++;
    lo[t] = a;
    hi[t] = b;
  }
  std::vector<MtxChunk> parts(nt);
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        if (lo[t] < hi[t]) parse_mtx_range(lo[t], hi[t], rows, cols, pattern, parts[t]);
      });
    for (auto &th : pool) th.join();
  }
  uint64_t m = 0, entries = 0;
  for (auto &c : parts) {
    if (c.bad) return STC_E_ARG;
    m += c.u.size();
    entries += c.entries;
  }
  if (entries != nnz) return STC_E_ARG;
  const size_t mm = m ? m : 1;
  g->u = (uint32_t *)std::malloc(mm * 4);
  g->v = (uint32_t *)std::malloc(mm * 4);
  g->sign = (int8_t *)std::malloc(mm);
  if (!g->u || !g->v || !g->sign) return STC_E_OOM;
  uint64_t off = 0;
  for (auto &c : parts) {
    const size_t k = c.u.size();
    if (k) {
      std::memcpy(g->u + off, c.u.data(), k * 4);
      std::memcpy(g->v + off, c.v.data(), k * 4);
      std::memcpy(g->sign + off, c.s.data(), k);
    }
    off += k;
  }
  g->m = m;
  g->n = (uint32_t)std::max(rows, cols);
  return STC_OK;
}

}  // namespace

extern "C" {

int stc_parse_graph(const char *buf, uint64_t len, int threads,
                    stc_host_graph **out) {
  if (!out || (!buf && len)) return STC_E_ARG;
  static const char kBanner[] = "%%MatrixMarket";
  if (len >= sizeof kBanner - 1 && std::memcmp(buf, kBanner, sizeof kBanner - 1) == 0) {
    *out = nullptr;
    auto *g = (stc_host_graph *)std::calloc(1, sizeof(stc_host_graph));
    if (!g) return STC_E_OOM;
    const int rc = parse_mtx(buf, len, threads, g);
    if (rc != STC_OK) {
      stc_graph_free(g);
      return rc;
    }
    *out = g;
    return STC_OK;
  }
  *out = nullptr;
  auto *g = (stc_host_graph *)std::calloc(1, sizeof(stc_host_graph));
  if (!g) return STC_E_OOM;
  if (len == 0) {
    g->u = (uint32_t *)std::malloc(4);
    g->v = (uint32_t *)std::malloc(4);
    g->sign = (int8_t *)std::malloc(1);
    *out = g;
    return STC_OK;
  }
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  if ((uint64_t)nt > len / 4096 + 1) nt = (int)(len / 4096 + 1);
  std::vector<const char *> lo(nt), hi(nt);
  const uint64_t chunk = len / nt;
  for (int t = 0; t < nt; t++) {
    const char *s = buf + (uint64_t)t * chunk;
    const char *e = (t == nt - 1) ? buf + len : buf + (uint64_t)(t + 1) * chunk;
    if (t > 0)
      while (s < buf + len && *(s - 1) != '\n') s++;
    if (t < nt - 1)
      while (e < buf + len && *(e - 1) != '\n') e++;
    lo[t] = s;
    hi[t] = e;
  }
  // Two passes, both threaded: each chunk's newline count bounds its entries,
  // so one allocation takes every chunk's slice and the threads parse
  // straight into the output; then only the chunks after a dropped line
  // (self-loop, comment, blank) move down.  No per-thread buffers, no serial
  // copy of the whole edge list.
  std::vector<uint64_t> base(nt + 1, 0), kept(nt, 0);
  std::vector<uint32_t> mx(nt, 0);
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] { base[t + 1] = count_newlines(lo[t], hi[t]) + 1; });
    for (auto &th : pool) th.join();
  }
  for (int t = 0; t < nt; t++) base[t + 1] += base[t];
  const size_t cap = base[nt] ? base[nt] : 1;
  g->u = (uint32_t *)std::malloc(cap * 4);
  g->v = (uint32_t *)std::malloc(cap * 4);
  g->sign = (int8_t *)std::malloc(cap);
  if (!g->u || !g->v || !g->sign) {
    stc_graph_free(g);
    return STC_E_OOM;
  }
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        if (lo[t] < hi[t])
          kept[t] = parse_range(lo[t], hi[t], g->u + base[t], g->v + base[t], g->sign + base[t],
                                mx[t]);
      });
    for (auto &th : pool) th.join();
  }
  uint64_t m = 0;
  uint32_t max_id = 0;
  for (int t = 0; t < nt; t++) {
    if (kept[t] && m != base[t]) {  // move the chunk down over the gap
      std::memmove(g->u + m, g->u + base[t], kept[t] * 4);
      std::memmove(g->v + m, g->v + base[t], kept[t] * 4);
      std::memmove(g->sign + m, g->sign + base[t], kept[t]);
    }
    m += kept[t];
    if (mx[t] > max_id) max_id = mx[t];
  }
  g->m = m;
  // Reference: n = max_id + 1 with max_id starting at 0 (graph_io.h:162,210).
  g->n = max_id + 1;
  *out = g;
  return STC_OK;
}

int stc_load_graph(const char *path, int threads, stc_host_graph **out) {
  if (!path || !out) return STC_E_ARG;
  *out = nullptr;
  int fd = open(path, O_RDONLY);
  if (fd < 0) return STC_E_IO;
  struct stat st;
  if (fstat(fd, &st) < 0) {
    close(fd);
    return STC_E_IO;
  }
  uint64_t size = (uint64_t)st.st_size;
  if (size == 0) {
    close(fd);
    return stc_parse_graph(nullptr, 0, threads, out);
  }
  void *data = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  if (data == MAP_FAILED) {
    close(fd);
    return STC_E_IO;
  }
  madvise(data, size, MADV_SEQUENTIAL);
  int rc = stc_parse_graph((const char *)data, size, threads, out);
  munmap(data, size);
  close(fd);
  return rc;
}

int stc_write_graph(const char *path, const stc_host_graph *g) {
  if (!path || !g) return STC_E_ARG;
  FILE *f = std::fopen(path, "wb");
  if (!f) return STC_E_IO;
  std::vector<char> buf(1 << 20);
  size_t pos = 0;
  for (uint64_t i = 0; i < g->m; i++) {
    if (pos + 40 > buf.size()) {
      if (std::fwrite(buf.data(), 1, pos, f) != pos) {
        std::fclose(f);
        return STC_E_IO;
      }
      pos = 0;
    }
    pos += (size_t)std::snprintf(buf.data() + pos, 40, "%u %u %s\n", g->u[i],
                                 g->v[i], g->sign[i] < 0 ? "-1" : "1");
  }
  if (pos && std::fwrite(buf.data(), 1, pos, f) != pos) {
    std::fclose(f);
    return STC_E_IO;
  }
  return std::fclose(f) == 0 ? STC_OK : STC_E_IO;
}

}  // extern "C"
