// Host ingest of the reference text format.
//
// Parsing rules follow Signed_TC/Multicore Implementations/graph_io.h:72-97
// (skip_whitespace / parse_uint / parse_sign) and the line loop of
// load_graph at graph_io.h:137-203: a '\n' or '\r' at a line start is skipped,
// a line starting with '#' or '%' is a comment, otherwise "u ws v ws sign" is
// read and the rest of the line ignored; u == v lines are dropped; n = max id
// + 1 over kept lines.  The file is split into per-thread chunks at newline
// boundaries as graph_io.h:117-133 does, but each chunk is parsed into its own
// buffer and the buffers are joined in file order, so a dropped self-loop
// leaves no hole (graph_io.h:147,159-161 reserve one and never fill it).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/stc_gen.h"

namespace {

struct Chunk {
  std::vector<uint32_t> u, v;
  std::vector<int8_t> s;
  uint32_t max_id = 0;
  bool any = false;
};

void parse_range(const char *p, const char *end, Chunk &out) {
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
    out.u.push_back(a);
    out.v.push_back(b);
    out.s.push_back(s);
    uint32_t hi = a > b ? a : b;
    if (hi > out.max_id) out.max_id = hi;
    out.any = true;
  }
}

}  // namespace

extern "C" {

int stc_parse_graph(const char *buf, uint64_t len, int threads,
                    stc_host_graph **out) {
  if (!out || (!buf && len)) return STC_E_ARG;
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
  std::vector<Chunk> parts(nt);
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        if (lo[t] < hi[t]) parse_range(lo[t], hi[t], parts[t]);
      });
    for (auto &th : pool) th.join();
  }
  uint64_t m = 0;
  uint32_t max_id = 0;
  for (auto &c : parts) {
    m += c.u.size();
    if (c.max_id > max_id) max_id = c.max_id;
  }
  size_t mm = m ? m : 1;
  g->u = (uint32_t *)std::malloc(mm * 4);
  g->v = (uint32_t *)std::malloc(mm * 4);
  g->sign = (int8_t *)std::malloc(mm);
  if (!g->u || !g->v || !g->sign) {
    stc_graph_free(g);
    return STC_E_OOM;
  }
  uint64_t off = 0;
  for (auto &c : parts) {
    size_t k = c.u.size();
    if (k) {
      std::memcpy(g->u + off, c.u.data(), k * 4);
      std::memcpy(g->v + off, c.v.data(), k * 4);
      std::memcpy(g->sign + off, c.s.data(), k);
    }
    off += k;
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
