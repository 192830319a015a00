/*
 * stc_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see stc_oracle.h).
 *
 * A plain C11 + OpenMP restatement of the reference CPU programs in
 * Signed_TC/Multicore Implementations/ (graph_io.h, tc_merge_sign_enu.cpp,
 * tc_hash_sign_enu.cpp, tc_serial.cpp).  Reference sources are not copied:
 * each function restates the algorithm and cites the lines it follows.
 *
 * Defined behaviour where the reference is undefined or unspecified
 * (excluded from parity with the reference by construction):
 *   - a self-loop line is dropped cleanly (the reference leaves an
 *     uninitialised slot, graph_io.h:147,159-161,191-192);
 *   - among duplicate lines of one undirected edge, the FIRST line in file
 *     order gives the sign (the reference keeps the first element after an
 *     unstable std::sort, graph_io.h:241-259, so conflicting signs are
 *     unspecified there).
 *
 * How each function is pinned (tests/test_oracle.py, tests/test_goldens.py):
 *   orc_parse            hand-written format fixtures (comments, CR/LF, tabs,
 *                        sign spellings, leading space, overflow) with
 *                        expected edge lists written out by hand.
 *   orc_build_csr        invariants (each unique edge in exactly one list,
 *                        out-neighbours have larger (deg,id), sorted lists,
 *                        m = unique pairs computed independently with numpy),
 *                        rank_ against an independent numpy lexsort.
 *   orc_count_merge /    golden counts produced by the reference binaries
 *   orc_count_hash /     (built from its sources by oracle/Makefile) on the
 *   orc_count_classes    generated workloads; closed forms (K_n all-+/all-,
 *                        2-colouring, grid + diagonals); trace identities
 *                        tr(A^3), tr(S^3) by scipy; brute force O(n^3) on
 *                        tiny graphs; switching and global-flip invariance.
 *   orc_count_classes_per_vertex  sums to orc_count_classes; brute force.
 */
#include "stc_oracle.h"

#include <fcntl.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

/* ---- parse: graph_io.h:72-97 (helpers) and :137-203 (line loop) ---------- */
int orc_parse(const char *buf, uint64_t len, orc_edges **out) {
  orc_edges *e = (orc_edges *)calloc(1, sizeof(orc_edges));
  if (!e) return -3;
  uint64_t cap = 1024, m = 0;
  e->u = (uint32_t *)malloc(cap * 4);
  e->v = (uint32_t *)malloc(cap * 4);
  e->sign = (int8_t *)malloc(cap);
  uint32_t max_id = 0;
  const char *p = buf, *end = buf + len;
  while (p < end) {
    if (*p == '\n' || *p == '\r') { p++; continue; }
    if (*p == '#' || *p == '%') {
      while (p < end && *p != '\n') p++;
      continue;
    }
    uint32_t a = 0, b = 0;
    int8_t s = 1;
    while (p < end && *p >= '0' && *p <= '9') a = a * 10u + (uint32_t)(*p++ - '0');
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    while (p < end && *p >= '0' && *p <= '9') b = b * 10u + (uint32_t)(*p++ - '0');
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    if (p < end && *p == '-') { s = -1; p++; }
    else if (p < end && *p == '+') p++;
    while (p < end && *p >= '0' && *p <= '9') p++;
    while (p < end && *p != '\n') p++;
    if (a == b) continue; /* graph_io.h:191-192, without the slot hole */
    if (m == cap) {
      cap *= 2;
      e->u = (uint32_t *)realloc(e->u, cap * 4);
      e->v = (uint32_t *)realloc(e->v, cap * 4);
      e->sign = (int8_t *)realloc(e->sign, cap);
    }
    e->u[m] = a; e->v[m] = b; e->sign[m] = s; m++;
    if (a > max_id) max_id = a;
    if (b > max_id) max_id = b;
  }
  e->m = m;
  e->n = len ? max_id + 1 : 0; /* graph_io.h:104-105,210 */
  *out = e;
  return 0;
}

int orc_load(const char *path, orc_edges **out) {
  int fd = open(path, O_RDONLY);
  if (fd < 0) return -6;
  struct stat st;
  if (fstat(fd, &st) < 0) { close(fd); return -6; }
  uint64_t size = (uint64_t)st.st_size;
  if (size == 0) { close(fd); return orc_parse("", 0, out); }
  const char *d = (const char *)mmap(NULL, size, PROT_READ, MAP_PRIVATE, fd, 0);
  if (d == MAP_FAILED) { close(fd); return -6; }
  int rc = orc_parse(d, size, out);
  munmap((void *)d, size);
  close(fd);
  return rc;
}

void orc_free_edges(orc_edges *e) {
  if (!e) return;
  free(e->u); free(e->v); free(e->sign); free(e);
}

/* ---- stable LSD radix sort of (key, index) by key ------------------------ */
static void radix_sort_kv(uint64_t *key, uint64_t *idx, uint64_t n, int bits) {
  uint64_t *k2 = (uint64_t *)malloc((n ? n : 1) * 8);
  uint64_t *i2 = (uint64_t *)malloc((n ? n : 1) * 8);
  uint64_t *sk = key, *si = idx, *dk = k2, *di = i2;
  for (int shift = 0; shift < bits; shift += 11) {
    uint64_t cnt[2049];
    memset(cnt, 0, sizeof cnt);
    for (uint64_t i = 0; i < n; i++) cnt[((sk[i] >> shift) & 2047) + 1]++;
    for (int d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
    for (uint64_t i = 0; i < n; i++) {
      uint64_t pos = cnt[(sk[i] >> shift) & 2047]++;
      dk[pos] = sk[i];
      di[pos] = si[i];
    }
    uint64_t *t = sk; sk = dk; dk = t;
    t = si; si = di; di = t;
  }
  if (sk != key) {
    memcpy(key, sk, n * 8);
    memcpy(idx, si, n * 8);
  }
  free(k2);
  free(i2);
}

static int cmp_u32_s8(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

/* ---- forward CSR: graph_io.h:231-402 -------------------------------------- */
int orc_build_csr(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                  const int8_t *sign, int threads, orc_csr **out) {
  if (threads > 0) omp_set_num_threads(threads);
  orc_csr *c = (orc_csr *)calloc(1, sizeof(orc_csr));
  if (!c) return -3;
  c->n = n;
  c->m_in = m;
  /* canonical (min,max) keys, stable sort, keep-first dedupe
     (graph_io.h:193-195 canonicalise, :241-259 sort + dedupe) */
  uint64_t *key = (uint64_t *)malloc((m ? m : 1) * 8);
  uint64_t *idx = (uint64_t *)malloc((m ? m : 1) * 8);
  /* self-loop lines are dropped as load_graph does (graph_io.h:191-192) */
  uint64_t mk = 0;
  for (uint64_t i = 0; i < m; i++) {
    uint32_t a = u[i], b = v[i];
    if (a == b) continue;
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    key[mk] = ((uint64_t)lo << 32) | hi;
    idx[mk] = i;
    mk++;
  }
  uint32_t nb = 0;
  while (nb < 32 && ((uint64_t)1 << nb) < (uint64_t)n) nb++;
  /* only bits [0,nb) of hi and [32,32+nb) of lo are non-zero: compress */
  #pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < mk; i++)
    key[i] = ((key[i] >> 32) << nb) | (key[i] & 0xffffffffull);
  radix_sort_kv(key, idx, mk, 2 * (int)nb > 0 ? 2 * (int)nb : 1);
  uint64_t um = 0;
  for (uint64_t i = 0; i < mk; i++) {
    if (um > 0 && key[um - 1] == key[i]) continue;
    key[um] = key[i];
    idx[um] = idx[i];
    um++;
  }
  const uint64_t lomask = nb ? (((uint64_t)1 << nb) - 1) : 0;
  uint32_t *eu = (uint32_t *)malloc((um ? um : 1) * 4);
  uint32_t *ev = (uint32_t *)malloc((um ? um : 1) * 4);
  int8_t *es = (int8_t *)malloc(um ? um : 1);
  uint64_t h = 1469598103934665603ull; /* FNV-1a-64 offset basis */
  for (uint64_t i = 0; i < um; i++) {
    eu[i] = (uint32_t)(key[i] >> nb);
    ev[i] = (uint32_t)(key[i] & lomask);
    es[i] = sign[idx[i]];
    unsigned char rec[9];
    memcpy(rec, &eu[i], 4);
    memcpy(rec + 4, &ev[i], 4);
    rec[8] = (unsigned char)es[i];
    for (int k = 0; k < 9; k++) { h ^= rec[k]; h *= 1099511628211ull; }
  }
  c->graph_hash = h;
  free(key);
  free(idx);
  /* degree (graph_io.h:261-268) */
  uint32_t *deg = (uint32_t *)calloc(n ? n : 1, 4);
  for (uint64_t i = 0; i < um; i++) { deg[eu[i]]++; deg[ev[i]]++; }
  /* rank_ = position in a stable counting sort by degree (graph_io.h:270-317) */
  uint32_t max_deg = 0;
  for (uint32_t i = 0; i < n; i++) if (deg[i] > max_deg) max_deg = deg[i];
  uint64_t *bucket = (uint64_t *)calloc((uint64_t)max_deg + 2, 8);
  for (uint32_t i = 0; i < n; i++) bucket[deg[i] + 1]++;
  for (uint32_t d = 1; d < max_deg + 2; d++) bucket[d] += bucket[d - 1];
  c->rank_ = (uint32_t *)malloc((n ? n : 1) * 4);
  for (uint32_t i = 0; i < n; i++) c->rank_[i] = (uint32_t)(bucket[deg[i]]++);
  free(bucket);
  /* orientation low -> high rank, fdeg, offsets (graph_io.h:319-367) */
  c->fdeg = (uint32_t *)calloc(n ? n : 1, 4);
  for (uint64_t i = 0; i < um; i++) {
    if (c->rank_[eu[i]] < c->rank_[ev[i]]) c->fdeg[eu[i]]++;
    else c->fdeg[ev[i]]++;
  }
  c->offsets = (uint64_t *)malloc(((uint64_t)n + 1) * 8);
  c->offsets[0] = 0;
  for (uint32_t i = 0; i < n; i++) c->offsets[i + 1] = c->offsets[i] + c->fdeg[i];
  c->m = um;
  /* scatter (graph_io.h:369-394), then sort each list (tc_merge_sign_enu.cpp:
     70-126 sorts by neighbour id, carrying the sign) */
  uint64_t *tmp = (uint64_t *)malloc((um ? um : 1) * 8);
  uint64_t *cur = (uint64_t *)malloc(((uint64_t)n + 1) * 8);
  memcpy(cur, c->offsets, (uint64_t)n * 8);
  for (uint64_t i = 0; i < um; i++) {
    uint32_t a = eu[i], b = ev[i];
    uint32_t src = c->rank_[a] < c->rank_[b] ? a : b;
    uint32_t dst = src == a ? b : a;
    tmp[cur[src]++] = ((uint64_t)dst << 8) | (uint8_t)es[i];
  }
  free(cur);
  free(eu); free(ev); free(es);
  uint32_t mf = 0;
  #pragma omp parallel for schedule(dynamic, 64) reduction(max : mf)
  for (uint32_t x = 0; x < n; x++) {
    uint64_t b = c->offsets[x], d = c->offsets[x + 1] - b;
    if (d > 1) qsort(tmp + b, d, 8, cmp_u32_s8);
    if ((uint32_t)d > mf) mf = (uint32_t)d;
  }
  c->max_fdeg = mf;
  c->edges = (uint32_t *)malloc((um ? um : 1) * 4);
  c->signs = (int8_t *)malloc(um ? um : 1);
  #pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < um; i++) {
    c->edges[i] = (uint32_t)(tmp[i] >> 8);
    c->signs[i] = (int8_t)(uint8_t)(tmp[i] & 0xff);
  }
  free(tmp);
  free(deg);
  *out = c;
  return 0;
}

void orc_free_csr(orc_csr *c) {
  if (!c) return;
  free(c->offsets); free(c->edges); free(c->signs); free(c->fdeg);
  free(c->rank_); free(c);
}

/* ---- merge: intersect_signed_ranked, tc_merge_sign_enu.cpp:12-44 --------- */
static inline void merge_one(const orc_csr *c, uint32_t v, uint64_t e,
                             uint64_t *bal, uint64_t *unbal) {
  const uint32_t *E = c->edges;
  const int8_t *S = c->signs;
  const uint32_t *R = c->rank_;
  uint32_t w = E[e];
  int s_vw = S[e];
  uint64_t a = c->offsets[v], a_end = c->offsets[v + 1];
  uint64_t b = c->offsets[w], b_end = c->offsets[w + 1];
  uint32_t rv = R[v], rw = R[w];
  while (a < a_end && b < b_end) {
    uint32_t x = E[a], y = E[b];
    if (x == y) {
      if (R[x] > rv && R[x] > rw) {
        int prod = s_vw * S[a] * S[b];
        if (prod > 0) (*bal)++; else (*unbal)++;
      }
      a++; b++;
    } else if (x < y) a++;
    else b++;
  }
}

/* count loop, tc_merge_sign_enu.cpp:128-145 */
void orc_count_merge(const orc_csr *c, int threads, uint64_t *bal,
                     uint64_t *unbal) {
  if (threads > 0) omp_set_num_threads(threads);
  unsigned long long B = 0, U = 0;
  #pragma omp parallel for schedule(dynamic, 16) reduction(+ : B, U)
  for (uint32_t v = 0; v < c->n; v++) {
    uint64_t b = 0, u = 0;
    for (uint64_t e = c->offsets[v]; e < c->offsets[v + 1]; e++) merge_one(c, v, e, &b, &u);
    B += b; U += u;
  }
  *bal = B; *unbal = U;
}

double orc_time_merge(const orc_csr *c, int threads, uint64_t *bal,
                      uint64_t *unbal) {
  double t0 = omp_get_wtime();
  orc_count_merge(c, threads, bal, unbal);
  return omp_get_wtime() - t0;
}

double orc_count_merge_sample(const orc_csr *c, int threads, uint32_t stride,
                              uint32_t phase, uint64_t *edges_done,
                              double *bytes_done, uint64_t *bal,
                              uint64_t *unbal) {
  if (threads > 0) omp_set_num_threads(threads);
  if (stride == 0) stride = 1;
  unsigned long long B = 0, U = 0, ne = 0;
  double bytes = 0;
  double t0 = omp_get_wtime();
  #pragma omp parallel for schedule(dynamic, 16) reduction(+ : B, U, ne, bytes)
  for (uint32_t v = phase % stride; v < c->n; v += stride) {
    uint64_t b = 0, u = 0;
    uint64_t beg = c->offsets[v], end = c->offsets[v + 1];
    for (uint64_t e = beg; e < end; e++) {
      merge_one(c, v, e, &b, &u);
      uint32_t w = c->edges[e];
      bytes += 4.0 * (double)((end - beg) + (c->offsets[w + 1] - c->offsets[w]));
    }
    ne += end - beg;
    B += b; U += u;
  }
  double t = omp_get_wtime() - t0;
  *edges_done = ne; *bytes_done = bytes; *bal = B; *unbal = U;
  return t;
}

/* ---- hash: AdjHashTable, tc_hash_sign_enu.cpp:11-74; loop :100-165 ------- */
typedef struct { uint32_t key; int8_t sign; } orc_ht_entry;
static inline uint32_t orc_hash32(uint32_t x) {
  x = ((x >> 16) ^ x) * 0x45d9f3bu;
  x = ((x >> 16) ^ x) * 0x45d9f3bu;
  return (x >> 16) ^ x;
}
static inline uint32_t orc_next_pow2(uint32_t x) {
  if (x == 0) return 1;
  x--;
  x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16;
  return x + 1;
}

/* The hash path in its two timed phases: the per-vertex tables (the
 * reference's "Build Hash Tables" timer, tc_hash_sign_enu.cpp:100-122) and the
 * count (its "Triangle Counting Execution", :124-165); wall seconds of each. */
static void hash_phases(const orc_csr *c, int threads, uint64_t *bal, uint64_t *unbal,
                        double *build_s, double *count_s) {
  if (threads > 0) omp_set_num_threads(threads);
  const double t0 = omp_get_wtime();
  const uint32_t n = c->n;
  uint64_t *ho = (uint64_t *)malloc(((uint64_t)n + 1) * 8);
  ho[0] = 0;
  for (uint32_t v = 0; v < n; v++) ho[v + 1] = ho[v] + orc_next_pow2(c->fdeg[v] * 2 + 1);
  orc_ht_entry *arena = (orc_ht_entry *)malloc((ho[n] ? ho[n] : 1) * sizeof(orc_ht_entry));
  #pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t v = 0; v < n; v++) {
    uint32_t cap = (uint32_t)(ho[v + 1] - ho[v]), mask = cap - 1;
    orc_ht_entry *t = arena + ho[v];
    for (uint32_t i = 0; i < cap; i++) t[i].key = UINT32_MAX;
    for (uint64_t e = c->offsets[v]; e < c->offsets[v + 1]; e++) {
      uint32_t pos = orc_hash32(c->edges[e]) & mask;
      while (t[pos].key != UINT32_MAX) pos = (pos + 1) & mask;
      t[pos].key = c->edges[e];
      t[pos].sign = c->signs[e];
    }
  }
  const double t1 = omp_get_wtime();
  const uint32_t *R = c->rank_;
  unsigned long long B = 0, U = 0;
  #pragma omp parallel for schedule(dynamic, 16) reduction(+ : B, U)
  for (uint32_t v = 0; v < n; v++) {
    for (uint64_t wi = c->offsets[v]; wi < c->offsets[v + 1]; wi++) {
      uint32_t w = c->edges[wi];
      if (R[v] >= R[w]) continue;
      int s_vw = c->signs[wi];
      int v_small = c->fdeg[v] <= c->fdeg[w];
      uint32_t it = v_small ? v : w, pr = v_small ? w : v;
      uint32_t mask = (uint32_t)(ho[pr + 1] - ho[pr]) - 1;
      const orc_ht_entry *t = arena + ho[pr];
      for (uint64_t i = c->offsets[it]; i < c->offsets[it + 1]; i++) {
        uint32_t x = c->edges[i];
        if (R[x] <= R[v] || R[x] <= R[w]) continue;
        uint32_t pos = orc_hash32(x) & mask;
        int found = 0, s_px = 0;
        while (t[pos].key != UINT32_MAX) {
          if (t[pos].key == x) { found = 1; s_px = t[pos].sign; break; }
          pos = (pos + 1) & mask;
        }
        if (!found) continue;
        int prod = s_vw * c->signs[i] * s_px;
        if (prod > 0) B++; else U++;
      }
    }
  }
  const double t2 = omp_get_wtime();
  free(arena);
  free(ho);
  *bal = B; *unbal = U;
  if (build_s) *build_s = t1 - t0;
  if (count_s) *count_s = t2 - t1;
}

void orc_count_hash(const orc_csr *c, int threads, uint64_t *bal,
                    uint64_t *unbal) {
  hash_phases(c, threads, bal, unbal, NULL, NULL);
}

double orc_time_hash(const orc_csr *c, int threads, double *build_s, uint64_t *bal,
                     uint64_t *unbal) {
  double count_s = 0;
  hash_phases(c, threads, bal, unbal, build_s, &count_s);
  return count_s;
}

/* ---- four classes: tc_serial.cpp:95-170 ---------------------------------- */
static inline void classes_one(const orc_csr *c, uint32_t v, uint64_t e,
                               uint64_t t[4]) {
  const uint32_t *E = c->edges;
  const int8_t *S = c->signs;
  const uint32_t *R = c->rank_;
  uint32_t w = E[e];
  int s_vw = S[e];
  uint64_t a = c->offsets[v], a_end = c->offsets[v + 1];
  uint64_t b = c->offsets[w], b_end = c->offsets[w + 1];
  while (a < a_end && b < b_end) {
    uint32_t x = E[a], y = E[b];
    if (x == y) {
      if (R[x] > R[v] && R[x] > R[w]) {
        int neg = ((1 - s_vw) >> 1) + ((1 - S[a]) >> 1) + ((1 - S[b]) >> 1);
        /* neg negatives: 0 -> T3, 1 -> T2, 2 -> T1, 3 -> T0 */
        t[3 - neg]++;
      }
      a++; b++;
    } else if (x < y) a++;
    else b++;
  }
}

void orc_count_classes_per_vertex(const orc_csr *c, int threads, uint64_t *cnt) {
  if (threads > 0) omp_set_num_threads(threads);
  #pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t v = 0; v < c->n; v++) {
    uint64_t t[4] = {0, 0, 0, 0};
    for (uint64_t e = c->offsets[v]; e < c->offsets[v + 1]; e++) classes_one(c, v, e, t);
    for (int k = 0; k < 4; k++) cnt[4 * (uint64_t)v + k] = t[k];
  }
}

void orc_count_classes(const orc_csr *c, int threads, uint64_t t[4]) {
  uint64_t *cnt = (uint64_t *)calloc(4 * ((uint64_t)c->n + 1), 8);
  orc_count_classes_per_vertex(c, threads, cnt);
  t[0] = t[1] = t[2] = t[3] = 0;
  for (uint32_t v = 0; v < c->n; v++)
    for (int k = 0; k < 4; k++) t[k] += cnt[4 * (uint64_t)v + k];
  free(cnt);
}
