/*
 * stc_oracle.h -- CPU oracle for the signed triangle counter.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C + OpenMP restatement of the
 * reference's CPU path (Signed_TC/Multicore Implementations/).  Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg may load it, and
 * only as the checker or the timed CPU baseline -- never as the product path.
 *
 * Every function cites the reference lines it restates.  How each one is
 * pinned (beyond restating the reference) is listed in the header of
 * stc_oracle.c.
 */
#ifndef STC_ORACLE_H
#define STC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_edges {
  uint32_t n;
  uint64_t m;
  uint32_t *u;
  uint32_t *v;
  int8_t *sign;
} orc_edges;

/* Forward (degree-oriented) CSR, the reference's ForwardCSR
 * (graph_io.h:22-30) plus bookkeeping. Lists are sorted ascending. */
typedef struct orc_csr {
  uint32_t n;
  uint64_t m;          /* oriented edges = unique undirected edges */
  uint64_t m_in;       /* input lines */
  uint64_t *offsets;   /* n + 1 */
  uint32_t *edges;     /* m */
  int8_t *signs;       /* m */
  uint32_t *fdeg;      /* n */
  uint32_t *rank_;     /* n */
  uint64_t graph_hash; /* FNV-1a-64 over the deduplicated canonical triples */
  uint32_t max_fdeg;
} orc_csr;

int orc_parse(const char *buf, uint64_t len, orc_edges **out);
int orc_load(const char *path, orc_edges **out);
void orc_free_edges(orc_edges *e);

int orc_build_csr(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                  const int8_t *sign, int threads, orc_csr **out);
void orc_free_csr(orc_csr *c);

void orc_count_merge(const orc_csr *c, int threads, uint64_t *bal,
                     uint64_t *unbal);
void orc_count_hash(const orc_csr *c, int threads, uint64_t *bal,
                    uint64_t *unbal);
void orc_count_classes(const orc_csr *c, int threads, uint64_t t[4]);
/* Per-vertex 4-class counts credited to the source vertex of each triangle's
 * lowest-ranked vertex (tc_serial.cpp:141-158): cnt[4*v + k]. */
void orc_count_classes_per_vertex(const orc_csr *c, int threads, uint64_t *cnt);

/* Merge count over the oriented edges of the source vertices v with
 * v % stride == phase; returns wall seconds of the counting loop, writes the
 * number of oriented edges processed and their algorithmic bytes
 * 4*(d+(u)+d+(v)) per edge. */
double orc_count_merge_sample(const orc_csr *c, int threads, uint32_t stride,
                              uint32_t phase, uint64_t *edges_done,
                              double *bytes_done, uint64_t *bal,
                              uint64_t *unbal);
/* Wall seconds of the full merge counting loop (the reference's
 * "Triangle Counting Execution" phase, tc_merge_sign_enu.cpp:131-145). */
double orc_time_merge(const orc_csr *c, int threads, uint64_t *bal,
                      uint64_t *unbal);
/* Wall seconds of the hash path's counting loop (tc_hash_sign_enu.cpp:124-165);
 * *build_s gets the seconds of its per-vertex table build (:100-122). */
double orc_time_hash(const orc_csr *c, int threads, double *build_s, uint64_t *bal,
                     uint64_t *unbal);

#ifdef __cplusplus
}
#endif
#endif
