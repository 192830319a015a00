// `stc <graph_file>`: drop-in for the reference's counting programs
// (Signed_TC/Multicore Implementations/tc_merge_sign_enu.cpp:46-154 and its
// siblings): same argument, same phase lines "[label] took %.4f seconds"
// (timer.h:24-26) and the same three result lines
//   Balanced   = %llu / Unbalanced = %llu / Total      = %llu
// (tc_merge_sign_enu.cpp:146-149).  Usage errors and unreadable files return
// 1.  Optional flags: --dispatch=learned|fixed|merge|hash|bitmap|pull, --repeat N,
// --classes (adds the T0..T3 block of tc_serial.cpp:18-30), --json,
// --device N, --model-dir DIR.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/stc.h"
#include "../../include/stc_gen.h"

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
void phase(const char *label, double s) { std::printf("[%s] took %.4f seconds\n", label, s); }
}  // namespace

int main(int argc, char **argv) {
  const char *path = nullptr;
  int dispatch = STC_LEARNED, repeat = 1, device = 0;
  bool classes = false, json = false;
  const char *model_dir = nullptr;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a.rfind("--dispatch=", 0) == 0) {
      std::string d = a.substr(11);
      if (d == "learned") dispatch = STC_LEARNED;
      else if (d == "fixed") dispatch = STC_FIXED;
      else if (d == "merge") dispatch = STC_FORCE_MERGE;
      else if (d == "hash") dispatch = STC_FORCE_HASH;
      else if (d == "bitmap") dispatch = STC_FORCE_BITMAP;
      else if (d == "pull") dispatch = STC_FORCE_PULL;
      else {
        std::fprintf(stderr, "unknown dispatch %s\n", d.c_str());
        return 1;
      }
    } else if (a == "--repeat" && i + 1 < argc) {
      repeat = std::atoi(argv[++i]);
    } else if (a == "--device" && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else if (a == "--model-dir" && i + 1 < argc) {
      model_dir = argv[++i];
    } else if (a == "--classes") {
      classes = true;
    } else if (a == "--json") {
      json = true;
    } else if (!path) {
      path = argv[i];
    }
  }
  if (!path) {
    std::printf("Usage: %s <graph_file_path>\n", argv[0]);
    return 1;
  }
  double t_total = now_s();
  double t0 = now_s();
  stc_host_graph *g = nullptr;
  int rc = stc_load_graph(path, 0, &g);
  if (rc != STC_OK) {
    std::perror(path);
    return 1;
  }
  phase("Graph Load (I/O + parse)", now_s() - t0);

  stc_ctx *ctx = nullptr;
  rc = stc_create(device, model_dir, &ctx);
  if (rc != STC_OK) {
    std::fprintf(stderr, "stc: no usable GPU (code %d)\n", rc);
    stc_graph_free(g);
    return 1;
  }
  stc_edges e{g->n, g->m, g->u, g->v, g->sign};
  t0 = now_s();
  rc = stc_prepare(ctx, &e, 0);
  if (rc != STC_OK) {
    std::fprintf(stderr, "stc: %s (code %d)\n", stc_last_error(ctx), rc);
    stc_destroy(ctx);
    stc_graph_free(g);
    return 1;
  }
  phase("Build Forward CSR", now_s() - t0);
  stc_opts o{dispatch, repeat > 0 ? repeat : 1, classes ? 1 : 0};
  stc_counts cnt{0, 0};
  t0 = now_s();
  rc = stc_count_prepared(ctx, &o, &cnt);
  double t_count = now_s() - t0;
  if (rc != STC_OK) {
    std::fprintf(stderr, "stc: %s (code %d)\n", stc_last_error(ctx), rc);
    stc_destroy(ctx);
    stc_graph_free(g);
    return 1;
  }
  phase("Triangle Counting Execution", t_count);
  std::printf("Balanced   = %llu\n", (unsigned long long)cnt.balanced);
  std::printf("Unbalanced = %llu\n", (unsigned long long)cnt.unbalanced);
  std::printf("Total      = %llu\n", (unsigned long long)(cnt.balanced + cnt.unbalanced));
  stc_stats st{};
  stc_get_stats(ctx, &st);
  std::printf("GPU time   = %.4f ms\n", st.ms_count);
  if (classes) {
    stc_classes c{};
    if (stc_last_classes(ctx, &c) == STC_OK) {
      std::printf("\nTriangle counts\n");
      std::printf("  T3 (+++ balanced)   : %llu\n", (unsigned long long)c.t3);
      std::printf("  T1 (+-- balanced)   : %llu\n", (unsigned long long)c.t1);
      std::printf("  T2 (++- unbalanced) : %llu\n", (unsigned long long)c.t2);
      std::printf("  T0 (--- unbalanced) : %llu\n", (unsigned long long)c.t0);
    }
  }
  if (json) {
    std::printf(
        "{\"n\": %u, \"m_in\": %llu, \"m_oriented\": %llu, \"balanced\": %llu, "
        "\"unbalanced\": %llu, \"ms_upload\": %.4f, \"ms_build\": %.4f, \"ms_count\": %.4f, "
        "\"b_alg\": %.0f, \"edges_merge\": %llu, \"edges_hash\": %llu, \"edges_bitmap\": %llu, "
        "\"edges_pull\": %llu, \"dispatch\": %d}\n",
        st.n, (unsigned long long)st.m_in, (unsigned long long)st.m_oriented,
        (unsigned long long)cnt.balanced, (unsigned long long)cnt.unbalanced, st.ms_upload,
        st.ms_build, st.ms_count, st.b_alg, (unsigned long long)st.edges_merge,
        (unsigned long long)st.edges_hash, (unsigned long long)st.edges_bitmap,
        (unsigned long long)st.edges_pull, st.dispatch_used);
  }
  stc_destroy(ctx);
  stc_graph_free(g);
  phase("Total execution time", now_s() - t_total);
  return 0;
}
