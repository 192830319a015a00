// Driver for the sanitizer builds of the host-only sources (graph_io.cpp:
// the threaded text and MatrixMarket parsers; gen.cpp: the threaded
// generators).  tests/test_sanitizers.py compiles it with g++ under
// -fsanitize=thread and under -fsanitize=address,undefined and runs it: every
// threaded result must equal the single-threaded one, and the sanitizer
// reports any race, out-of-bounds access or leak.  Exit status 0 = all equal.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../include/stc_gen.h"

static int failures = 0;

static bool same(const stc_host_graph *a, const stc_host_graph *b) {
  if (a->n != b->n || a->m != b->m) return false;
  return a->m == 0 || (!std::memcmp(a->u, b->u, a->m * 4) && !std::memcmp(a->v, b->v, a->m * 4) &&
                       !std::memcmp(a->sign, b->sign, a->m));
}

static void check(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "MISMATCH: %s\n", what);
    failures++;
  }
}

// Parse `text` with 1 thread and with each of `threads`; all must agree.
static void parse_all(const std::string &text, const char *what) {
  stc_host_graph *ref = nullptr;
  if (stc_parse_graph(text.data(), text.size(), 1, &ref) != 0) {
    check(false, what);
    return;
  }
  for (int t : {2, 3, 7, 8, 16}) {
    stc_host_graph *g = nullptr;
    check(stc_parse_graph(text.data(), text.size(), t, &g) == 0 && same(ref, g), what);
    stc_graph_free(g);
  }
  stc_graph_free(ref);
}

static std::string to_text(const stc_host_graph *g, bool noise) {
  std::string s;
  char buf[64];
  for (uint64_t i = 0; i < g->m; i++) {
    if (noise && i % 97 == 0) s += "# a comment line\n";
    if (noise && i % 101 == 0) s += "\r\n";
    if (noise && i % 89 == 0) {  // a self-loop line: dropped
      std::snprintf(buf, sizeof buf, "%u %u 1\n", g->u[i], g->u[i]);
      s += buf;
    }
    std::snprintf(buf, sizeof buf, noise && i % 3 == 0 ? "%u\t%u\t%s\r\n" : "%u %u %s\n", g->u[i],
                  g->v[i], g->sign[i] < 0 ? "-1" : (i % 2 ? "+1" : "1"));
    s += buf;
  }
  return s;
}

int main() {
  // threaded generators: every thread count gives the same graph
  stc_host_graph *r1 = nullptr, *r8 = nullptr;
  check(stc_gen_rmat(14, 16, 0.57, 0.19, 0.19, 0.3, 11, 1, &r1) == 0, "rmat gen 1");
  check(stc_gen_rmat(14, 16, 0.57, 0.19, 0.19, 0.3, 11, 8, &r8) == 0, "rmat gen 8");
  check(r1 && r8 && same(r1, r8), "rmat threads");
  stc_host_graph *h1 = nullptr, *h8 = nullptr;
  check(stc_gen_hub(200000, 300000, 8, 1.2, 0.6, 2000, 0.1, 12, 1, &h1) == 0, "hub gen 1");
  check(stc_gen_hub(200000, 300000, 8, 1.2, 0.6, 2000, 0.1, 12, 8, &h8) == 0, "hub gen 8");
  check(h1 && h8 && same(h1, h8), "hub threads");
  stc_host_batch *b1 = nullptr, *b8 = nullptr;
  check(stc_gen_batch(64, 64, 4096, 4.0, 32.0, 0.05, 0.5, 13, 1, &b1) == 0, "batch gen 1");
  check(stc_gen_batch(64, 64, 4096, 4.0, 32.0, 0.05, 0.5, 13, 8, &b8) == 0, "batch gen 8");
  check(b1 && b8 && b1->total_m == b8->total_m &&
            !std::memcmp(b1->u, b8->u, b1->total_m * 4) &&
            !std::memcmp(b1->sign, b8->sign, b1->total_m),
        "batch threads");
  // threaded text parser on clean and noisy text, on text with no final
  // newline and on text smaller than the thread count
  if (r1) {
    parse_all(to_text(r1, false), "rmat text");
    parse_all(to_text(r1, true), "rmat noisy text");
    std::string t = to_text(r1, false);
    t.pop_back();
    parse_all(t, "no final newline");
  }
  parse_all("", "empty");
  parse_all("0 1 -1", "one line");
  parse_all("# only a comment\n% and another\n", "comments only");
  parse_all("3 3 1\n4 4 -1\n", "self-loops only");
  // MatrixMarket through the same entry point
  std::string mtx = "%%MatrixMarket matrix coordinate real general\n% c\n6 6 5\n";
  mtx += "1 2 1.5\n2 3 -2\n3 1 0.5\n2 1 1.5\n5 6 -1\n";
  parse_all(mtx, "matrix market");
  stc_graph_free(r1);
  stc_graph_free(r8);
  stc_graph_free(h1);
  stc_graph_free(h8);
  stc_batch_free(b1);
  stc_batch_free(b8);
  if (failures) {
    std::fprintf(stderr, "%d mismatches\n", failures);
    return 1;
  }
  std::printf("sanitize_host: ok\n");
  return 0;
}
