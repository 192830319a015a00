/*
 * stc_gen.h -- host-side graph sources for the signed triangle counter:
 * the reference text format (read and write) and the deterministic synthetic
 * workload generators.  Everything here runs on the CPU only; nothing calls
 * HIP, so these entry points work on a machine without a GPU.
 *
 * Text format (reference: Signed_TC/Multicore Implementations/graph_io.h:72-212):
 *   one edge per line "u<ws>v[<ws>sign]", <ws> = spaces/tabs; a line whose first
 *   character is '#' or '%' is a comment; '\n' / '\r' at a line start are skipped;
 *   ids are unsigned decimals (no validation, a non-digit reads as 0, uint32
 *   wrap-around on overflow); the sign is -1 iff the third field starts with '-',
 *   otherwise +1 (including "+1", "1", "0" and a missing field); the rest of the
 *   line is ignored; lines with u == v are dropped; n = max id + 1 over the kept
 *   lines (n = 1 for a non-empty file whose lines are all dropped, n = 0 for an
 *   empty file).
 *
 * Deviation from the reference, by design: a self-loop line is dropped cleanly.
 * The reference reserves a slot for it and leaves it uninitialised
 * (graph_io.h:147,159-161,191-192), which is undefined behaviour.
 */
#ifndef STC_GEN_H
#define STC_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A host edge list owned by the library (free with stc_graph_free).
 * u[i], v[i] are raw endpoint ids as written/generated (not canonicalised),
 * sign[i] is +1 or -1.  n bounds every id (ids < n). */
typedef struct stc_host_graph {
  uint32_t n;
  uint64_t m;
  uint32_t *u;
  uint32_t *v;
  int8_t *sign;
} stc_host_graph;

/* k graphs stored back to back in one pool: graph g owns lines
 * [edge_off[g], edge_off[g+1]) of u/v/sign and has vertex-id bound n[g].
 * kind[g] = 0 for Erdos-Renyi, 1 for RMAT; p_neg[g] is its negative-sign
 * probability and avg_deg[g] the (clamped) requested average degree. */
typedef struct stc_host_batch {
  uint32_t k;
  uint64_t total_m;
  uint64_t *edge_off; /* k + 1 */
  uint32_t *n;        /* k */
  uint8_t *kind;      /* k */
  double *p_neg;      /* k */
  double *avg_deg;    /* k */
  uint32_t *u;
  uint32_t *v;
  int8_t *sign;
} stc_host_batch;

/* Return codes shared with stc.h. */
#ifndef STC_OK
#define STC_OK 0
#define STC_E_ARG (-1)
#define STC_E_TOO_LARGE (-2)
#define STC_E_OOM (-3)
#define STC_E_HIP (-4)
#define STC_E_MODEL (-5)
#define STC_E_IO (-6)
#define STC_E_NODEV (-7)
#endif

/* Parse a graph file in the reference text format (semantics above).
 * threads <= 0 uses all hardware threads.  Returns STC_E_IO if the file cannot
 * be opened or read. */
int stc_load_graph(const char *path, int threads, stc_host_graph **out);
/* Same parser over an in-memory buffer. */
int stc_parse_graph(const char *buf, uint64_t len, int threads,
                    stc_host_graph **out);
/* Write "u v 1" / "u v -1" lines, one per edge, in list order. */
int stc_write_graph(const char *path, const stc_host_graph *g);
void stc_graph_free(stc_host_graph *g);

/* --- deterministic generators (SplitMix64-seeded xoshiro256**; every
 * probability is applied as an integer threshold, so a seed gives the same
 * graph on every x86-64 host and for every thread count) --------------------
 * The sign of an undirected edge {u,v} is a pure function of (seed, min, max),
 * so duplicate lines of one edge always carry the same sign. */

/* G(n, m): exactly m distinct non-loop pairs by rejection sampling, written in
 * draw order with the endpoints as drawn. */
int stc_gen_er(uint32_t n, uint64_t m, double p_neg, uint64_t seed,
               stc_host_graph **out);
/* rows x cols 4-neighbour grid; each cell gets, with probability p_diag, one
 * diagonal whose direction is chosen uniformly. Vertex id = r * cols + c. */
int stc_gen_grid(uint32_t rows, uint32_t cols, double p_diag, double p_neg,
                 uint64_t seed, stc_host_graph **out);
/* RMAT: edge_factor << scale pairs with quadrant probabilities (a,b,c,1-a-b-c);
 * a self-loop draw is redrawn; duplicate pairs are kept (the counter dedupes). */
int stc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                 double c, double p_neg, uint64_t seed, int threads,
                 stc_host_graph **out);
/* Hub-dominated graph: `hubs` hubs joined in a clique whose neighbour counts are
 * Zipf(s) shares of hub_share * target_m distinct non-hub neighbours, plus
 * uniform random non-hub pairs, plus `planted` random triangles on non-hub
 * vertices, for target_m lines in all. */
int stc_gen_hub(uint32_t n, uint64_t target_m, uint32_t hubs, double zipf_s,
                double hub_share, uint32_t planted, double p_neg, uint64_t seed,
                int threads, stc_host_graph **out);
/* The batched serving workload: k graphs, n log-uniform in [n_lo, n_hi],
 * average degree uniform in [d_lo, d_hi] (clamped to n-1), p_neg uniform in
 * [p_lo, p_hi]; even g -> Erdos-Renyi, odd g -> RMAT(0.57,0.19,0.19) at scale
 * ceil(log2 n) with out-of-range ids redrawn. */
int stc_gen_batch(uint32_t k, uint32_t n_lo, uint32_t n_hi, double d_lo,
                  double d_hi, double p_lo, double p_hi, uint64_t seed,
                  int threads, stc_host_batch **out);
void stc_batch_free(stc_host_batch *b);

/* The five benchmark workloads by number (1..5) with their fixed seeds:
 * 1 ER n=10,000 m=100,000 p-=0.3; 2 grid 960x960 + diagonals p=0.5 p-=0.2;
 * 3 RMAT scale 22 ef 16 (0.57,0.19,0.19,0.05) p-=0.25; 4 hub n=18M m=19M
 * 8 hubs Zipf(1.2) 60% + 200k planted triangles p-=0.1.  Workload 5 is the
 * batch (use stc_gen_workload_batch). */
int stc_gen_workload(int id, int threads, stc_host_graph **out);
int stc_gen_workload_batch(uint32_t k, int threads, stc_host_batch **out);

#ifdef __cplusplus
}
#endif
#endif /* STC_GEN_H */
