// Deterministic synthetic signed graphs for the five benchmark workloads.
// No reference counterpart: the reference ships no generator and no datasets
// (its README points at SuiteSparse downloads). Output follows the reference
// input contract of Signed_TC/Multicore Implementations/graph_io.h:100-212:
// (u, v, sign) lines with sign in {+1, -1} and no self-loops.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/stc_gen.h"
#include "rng.h"

namespace stc {
namespace {

int hw_threads(int threads) {
  if (threads > 0) return threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

template <class F>
void parallel_for_chunks(uint64_t nchunks, int threads, F &&f) {
  threads = (int)std::min<uint64_t>((uint64_t)hw_threads(threads), nchunks);
  if (threads <= 1) {
    for (uint64_t c = 0; c < nchunks; c++) f(c);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      for (uint64_t c = t; c < nchunks; c += threads) f(c);
    });
  for (auto &th : pool) th.join();
}

stc_host_graph *alloc_graph(uint32_t n, uint64_t m) {
  auto *g = (stc_host_graph *)std::calloc(1, sizeof(stc_host_graph));
  if (!g) return nullptr;
  g->n = n;
  g->m = m;
  size_t mm = m ? m : 1;
  g->u = (uint32_t *)std::malloc(mm * sizeof(uint32_t));
  g->v = (uint32_t *)std::malloc(mm * sizeof(uint32_t));
  g->sign = (int8_t *)std::malloc(mm);
  if (!g->u || !g->v || !g->sign) {
    stc_graph_free(g);
    return nullptr;
  }
  return g;
}

// Open-addressing set of 64-bit keys (key 0 is never inserted: pairs are
// stored as (lo << 32 | hi) + 1).
struct PairSet {
  std::vector<uint64_t> slot;
  uint64_t mask;
  explicit PairSet(uint64_t expect) {
    uint64_t cap = 16;
    while (cap < expect * 2 + 16) cap <<= 1;
    slot.assign(cap, 0);
    mask = cap - 1;
  }
  bool insert(uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    uint64_t key = (((uint64_t)lo << 32) | hi) + 1;
    uint64_t pos = mix64(key) & mask;
    while (slot[pos]) {
      if (slot[pos] == key) return false;
      pos = (pos + 1) & mask;
    }
    slot[pos] = key;
    return true;
  }
};

// Erdos-Renyi G(n, m) into caller arrays (m distinct non-loop pairs).
void fill_er(uint32_t n, uint64_t m, uint64_t seed, uint64_t sign_seed,
             uint64_t thr, uint32_t *u, uint32_t *v, int8_t *s) {
  Xoshiro256ss rng(seed);
  PairSet set(m);
  uint64_t i = 0;
  while (i < m) {
    uint32_t a = rng.below(n), b = rng.below(n);
    if (a == b || !set.insert(a, b)) continue;
    u[i] = a;
    v[i] = b;
    s[i] = edge_sign(sign_seed, a, b, thr);
    i++;
  }
}

// One RMAT endpoint pair at `scale` levels; quadrant thresholds are cumulative
// 32-bit integers (ta < tb < tc).
inline void rmat_pair(Xoshiro256ss &rng, uint32_t scale, uint32_t ta,
                      uint32_t tb, uint32_t tc, uint32_t &a, uint32_t &b) {
  uint32_t x = 0, y = 0;
  uint32_t lvl = 0;
  while (lvl < scale) {
    uint64_t r = rng.next();
    for (int half = 0; half < 2 && lvl < scale; half++, lvl++) {
      uint32_t q = (uint32_t)(half ? (r >> 32) : r);
      uint32_t bit = 1u << (scale - 1 - lvl);
      if (q < ta) {
      } else if (q < tb) {
        y |= bit;
      } else if (q < tc) {
        x |= bit;
      } else {
        x |= bit;
        y |= bit;
      }
    }
  }
  a = x;
  b = y;
}

inline uint32_t u32_threshold(double p) {
  double t = p * 4294967296.0;
  if (t >= 4294967295.0) return 4294967295u;
  if (t <= 0) return 0;
  return (uint32_t)t;
}

// RMAT pairs with ids restricted to [0, n) (redraw if out of range or a loop).
void fill_rmat(uint32_t scale, uint32_t n, uint64_t m, double pa, double pb,
               double pc, uint64_t seed, uint64_t sign_seed, uint64_t thr,
               uint32_t *u, uint32_t *v, int8_t *s, int threads) {
  const uint32_t ta = u32_threshold(pa), tb = u32_threshold(pa + pb),
                 tc = u32_threshold(pa + pb + pc);
  const uint64_t chunk = 1ull << 16;
  const uint64_t nchunks = (m + chunk - 1) / chunk;
  parallel_for_chunks(nchunks, threads, [&](uint64_t c) {
    Xoshiro256ss rng(mix64(seed ^ mix64(c + 0x5eedull)));
    uint64_t lo = c * chunk, hi = std::min(m, lo + chunk);
    for (uint64_t i = lo; i < hi; i++) {
      uint32_t a, b;
      do {
        rmat_pair(rng, scale, ta, tb, tc, a, b);
      } while (a == b || a >= n || b >= n);
      u[i] = a;
      v[i] = b;
      s[i] = edge_sign(sign_seed, a, b, thr);
    }
  });
}

}  // namespace
}  // namespace stc

using namespace stc;

extern "C" {

void stc_graph_free(stc_host_graph *g) {
  if (!g) return;
  std::free(g->u);
  std::free(g->v);
  std::free(g->sign);
  std::free(g);
}

void stc_batch_free(stc_host_batch *b) {
  if (!b) return;
  std::free(b->edge_off);
  std::free(b->n);
  std::free(b->kind);
  std::free(b->p_neg);
  std::free(b->avg_deg);
  std::free(b->u);
  std::free(b->v);
  std::free(b->sign);
  std::free(b);
}

int stc_gen_er(uint32_t n, uint64_t m, double p_neg, uint64_t seed,
               stc_host_graph **out) {
  if (!out) return STC_E_ARG;
  *out = nullptr;
  if (n < 2 && m > 0) return STC_E_ARG;
  if (m > (uint64_t)n * (n - 1) / 2) return STC_E_ARG;
  stc_host_graph *g = alloc_graph(n, m);
  if (!g) return STC_E_OOM;
  fill_er(n, m, seed, mix64(seed ^ 0x5167ull), prob_threshold53(p_neg), g->u,
          g->v, g->sign);
  *out = g;
  return STC_OK;
}

int stc_gen_grid(uint32_t rows, uint32_t cols, double p_diag, double p_neg,
                 uint64_t seed, stc_host_graph **out) {
  if (!out || rows == 0 || cols == 0) return STC_E_ARG;
  *out = nullptr;
  if ((uint64_t)rows * cols >= (1ull << 31)) return STC_E_TOO_LARGE;
  const uint64_t thr_diag = prob_threshold53(p_diag);
  const uint64_t thr = prob_threshold53(p_neg);
  const uint64_t sign_seed = mix64(seed ^ 0x5167ull);
  const uint64_t cell_seed = mix64(seed ^ 0xce11ull);
  // Count diagonals first so the arrays are exact.
  std::vector<uint8_t> diag((uint64_t)(rows > 0 ? rows - 1 : 0) *
                            (cols > 0 ? cols - 1 : 0));
  uint64_t ndiag = 0;
  for (uint32_t r = 0; r + 1 < rows; r++)
    for (uint32_t c = 0; c + 1 < cols; c++) {
      uint64_t cell = (uint64_t)r * (cols - 1) + c;
      uint64_t h = mix64(cell_seed ^ mix64(cell));
      uint8_t d = 0;
      if ((h >> 11) < thr_diag) d = (h & 1) ? 2 : 1;  // 1: down-right, 2: down-left
      diag[cell] = d;
      ndiag += d != 0;
    }
  uint64_t m = (uint64_t)rows * (cols - 1) + (uint64_t)(rows - 1) * cols + ndiag;
  stc_host_graph *g = alloc_graph(rows * cols, m);
  if (!g) return STC_E_OOM;
  uint64_t i = 0;
  auto put = [&](uint32_t a, uint32_t b) {
    g->u[i] = a;
    g->v[i] = b;
    g->sign[i] = edge_sign(sign_seed, a, b, thr);
    i++;
  };
  for (uint32_t r = 0; r < rows; r++)
    for (uint32_t c = 0; c < cols; c++) {
      uint32_t id = r * cols + c;
      if (c + 1 < cols) put(id, id + 1);
      if (r + 1 < rows) put(id, id + cols);
      if (r + 1 < rows && c + 1 < cols) {
        uint8_t d = diag[(uint64_t)r * (cols - 1) + c];
        if (d == 1) put(id, id + cols + 1);
        if (d == 2) put(id + 1, id + cols);
      }
    }
  *out = g;
  return STC_OK;
}

int stc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                 double c, double p_neg, uint64_t seed, int threads,
                 stc_host_graph **out) {
  if (!out || scale == 0 || scale > 31 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return STC_E_ARG;
  *out = nullptr;
  uint64_t m = (uint64_t)edge_factor << scale;
  uint32_t n = 1u << scale;
  stc_host_graph *g = alloc_graph(n, m);
  if (!g) return STC_E_OOM;
  fill_rmat(scale, n, m, a, b, c, seed, mix64(seed ^ 0x5167ull),
            prob_threshold53(p_neg), g->u, g->v, g->sign, threads);
  *out = g;
  return STC_OK;
}

int stc_gen_hub(uint32_t n, uint64_t target_m, uint32_t hubs, double zipf_s,
                double hub_share, uint32_t planted, double p_neg, uint64_t seed,
                int threads, stc_host_graph **out) {
  if (!out || hubs == 0 || n < hubs + 3 || n >= (1u << 31)) return STC_E_ARG;
  *out = nullptr;
  const uint64_t thr = prob_threshold53(p_neg);
  const uint64_t sign_seed = mix64(seed ^ 0x5167ull);
  Xoshiro256ss rng(mix64(seed ^ 0x4ab5ull));
  // Hub ids: distinct uniform draws.
  std::vector<uint32_t> hub_id;
  std::vector<uint8_t> is_hub(n, 0);
  while (hub_id.size() < hubs) {
    uint32_t h = rng.below(n);
    if (is_hub[h]) continue;
    is_hub[h] = 1;
    hub_id.push_back(h);
  }
  // Zipf shares of the hub edges.
  double wsum = 0;
  for (uint32_t k = 1; k <= hubs; k++) wsum += std::pow((double)k, -zipf_s);
  const uint64_t hub_total = (uint64_t)std::llround(hub_share * (double)target_m);
  std::vector<uint64_t> hub_deg(hubs);
  uint64_t hub_sum = 0;
  for (uint32_t k = 0; k < hubs; k++) {
    hub_deg[k] = (uint64_t)std::llround((double)hub_total *
                                         std::pow((double)(k + 1), -zipf_s) / wsum);
    hub_deg[k] = std::min<uint64_t>(hub_deg[k], n - hubs);
    hub_sum += hub_deg[k];
  }
  const uint64_t clique = (uint64_t)hubs * (hubs - 1) / 2;
  const uint64_t tri_edges = 3ull * planted;
  if (hub_sum + clique + tri_edges > target_m) return STC_E_ARG;
  const uint64_t rand_edges = target_m - hub_sum - clique - tri_edges;
  stc_host_graph *g = alloc_graph(n, target_m);
  if (!g) return STC_E_OOM;
  uint64_t i = 0;
  auto put = [&](uint32_t a, uint32_t b) {
    g->u[i] = a;
    g->v[i] = b;
    g->sign[i] = edge_sign(sign_seed, a, b, thr);
    i++;
  };
  for (uint32_t x = 0; x < hubs; x++)
    for (uint32_t y = x + 1; y < hubs; y++) put(hub_id[x], hub_id[y]);
  // Each hub: distinct uniform non-hub neighbours (per-hub stream, parallel).
  std::vector<uint64_t> base(hubs + 1, i);
  for (uint32_t k = 0; k < hubs; k++) base[k + 1] = base[k] + hub_deg[k];
  parallel_for_chunks(hubs, threads, [&](uint64_t k) {
    Xoshiro256ss r(mix64(seed ^ mix64(0x40b0ull + k)));
    std::vector<uint64_t> seen((n + 63) / 64, 0);
    uint64_t j = base[k];
    while (j < base[k + 1]) {
      uint32_t x = r.below(n);
      if (is_hub[x] || (seen[x >> 6] >> (x & 63) & 1)) continue;
      seen[x >> 6] |= 1ull << (x & 63);
      g->u[j] = hub_id[k];
      g->v[j] = x;
      g->sign[j] = edge_sign(sign_seed, hub_id[k], x, thr);
      j++;
    }
  });
  i = base[hubs];
  // Uniform random non-hub pairs (duplicates possible and kept).
  {
    const uint64_t chunk = 1ull << 16;
    const uint64_t nch = (rand_edges + chunk - 1) / chunk;
    const uint64_t start = i;
    parallel_for_chunks(nch, threads, [&](uint64_t c) {
      Xoshiro256ss r(mix64(seed ^ mix64(0xa11dull + c)));
      uint64_t lo = c * chunk, hi = std::min(rand_edges, lo + chunk);
      for (uint64_t e = lo; e < hi; e++) {
        uint32_t a, b;
        do {
          a = r.below(n);
          b = r.below(n);
        } while (a == b || is_hub[a] || is_hub[b]);
        g->u[start + e] = a;
        g->v[start + e] = b;
        g->sign[start + e] = edge_sign(sign_seed, a, b, thr);
      }
    });
    i = start + rand_edges;
  }
  // Planted triangles on three distinct non-hub vertices.
  for (uint32_t t = 0; t < planted; t++) {
    uint32_t a, b, c;
    do {
      a = rng.below(n);
      b = rng.below(n);
      c = rng.below(n);
    } while (a == b || b == c || a == c || is_hub[a] || is_hub[b] || is_hub[c]);
    put(a, b);
    put(b, c);
    put(a, c);
  }
  g->m = i;
  *out = g;
  return STC_OK;
}

int stc_gen_batch(uint32_t k, uint32_t n_lo, uint32_t n_hi, double d_lo,
                  double d_hi, double p_lo, double p_hi, uint64_t seed,
                  int threads, stc_host_batch **out) {
  if (!out || k == 0 || n_lo < 4 || n_hi < n_lo || n_hi >= (1u << 30))
    return STC_E_ARG;
  *out = nullptr;
  auto *b = (stc_host_batch *)std::calloc(1, sizeof(stc_host_batch));
  if (!b) return STC_E_OOM;
  b->k = k;
  b->edge_off = (uint64_t *)std::calloc(k + 1, sizeof(uint64_t));
  b->n = (uint32_t *)std::calloc(k, sizeof(uint32_t));
  b->kind = (uint8_t *)std::calloc(k, 1);
  b->p_neg = (double *)std::calloc(k, sizeof(double));
  b->avg_deg = (double *)std::calloc(k, sizeof(double));
  if (!b->edge_off || !b->n || !b->kind || !b->p_neg || !b->avg_deg) {
    stc_batch_free(b);
    return STC_E_OOM;
  }
  const double lg_lo = std::log2((double)n_lo), lg_hi = std::log2((double)n_hi);
  for (uint32_t g = 0; g < k; g++) {
    Xoshiro256ss r(mix64(seed ^ mix64(0xba7c0000ull + g)));
    double e = lg_lo + (lg_hi - lg_lo) * r.unit();
    uint32_t n = (uint32_t)std::llround(std::exp2(e));
    n = std::min(std::max(n, n_lo), n_hi);
    double d = d_lo + (d_hi - d_lo) * r.unit();
    if (d > (double)(n - 1)) d = (double)(n - 1);
    double p = p_lo + (p_hi - p_lo) * r.unit();
    uint64_t m = (uint64_t)std::llround((double)n * d / 2.0);
    uint64_t mmax = (uint64_t)n * (n - 1) / 2;
    if (m > mmax) m = mmax;
    b->n[g] = n;
    b->kind[g] = (uint8_t)(g & 1);
    b->p_neg[g] = p;
    b->avg_deg[g] = d;
    b->edge_off[g + 1] = b->edge_off[g] + m;
  }
  b->total_m = b->edge_off[k];
  size_t mm = b->total_m ? b->total_m : 1;
  b->u = (uint32_t *)std::malloc(mm * sizeof(uint32_t));
  b->v = (uint32_t *)std::malloc(mm * sizeof(uint32_t));
  b->sign = (int8_t *)std::malloc(mm);
  if (!b->u || !b->v || !b->sign) {
    stc_batch_free(b);
    return STC_E_OOM;
  }
  // Largest graphs first so the thread pool stays balanced.
  std::vector<uint32_t> order(k);
  for (uint32_t g = 0; g < k; g++) order[g] = g;
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    uint64_t mx = b->edge_off[x + 1] - b->edge_off[x];
    uint64_t my = b->edge_off[y + 1] - b->edge_off[y];
    return mx != my ? mx > my : x < y;
  });
  parallel_for_chunks(k, threads, [&](uint64_t idx) {
    uint32_t g = order[idx];
    uint64_t off = b->edge_off[g], m = b->edge_off[g + 1] - off;
    uint64_t gseed = mix64(seed ^ mix64(0x9a9a0000ull + g));
    uint64_t thr = prob_threshold53(b->p_neg[g]);
    uint64_t sseed = mix64(gseed ^ 0x5167ull);
    if (b->kind[g] == 0) {
      fill_er(b->n[g], m, gseed, sseed, thr, b->u + off, b->v + off,
              b->sign + off);
    } else {
      uint32_t scale = 1;
      while ((1u << scale) < b->n[g]) scale++;
      fill_rmat(scale, b->n[g], m, 0.57, 0.19, 0.19, gseed, sseed, thr,
                b->u + off, b->v + off, b->sign + off, 1);
    }
  });
  *out = b;
  return STC_OK;
}

// Fixed seeds of the benchmark workloads.
static const uint64_t kWorkloadSeed[6] = {0, 0x5EED0001ull, 0x5EED0002ull,
                                          0x5EED0003ull, 0x5EED0004ull,
                                          0x5EED0005ull};

int stc_gen_workload(int id, int threads, stc_host_graph **out) {
  switch (id) {
    case 1:
      return stc_gen_er(10000, 100000, 0.3, kWorkloadSeed[1], out);
    case 2:
      return stc_gen_grid(960, 960, 0.5, 0.2, kWorkloadSeed[2], out);
    case 3:
      return stc_gen_rmat(22, 16, 0.57, 0.19, 0.19, 0.25, kWorkloadSeed[3],
                          threads, out);
    case 4:
      return stc_gen_hub(18000000u, 19000000ull, 8, 1.2, 0.6, 200000, 0.1,
                         kWorkloadSeed[4], threads, out);
    default:
      if (out) *out = nullptr;
      return STC_E_ARG;
  }
}

int stc_gen_workload_batch(uint32_t k, int threads, stc_host_batch **out) {
  return stc_gen_batch(k, 64, 65536, 4.0, 64.0, 0.05, 0.5, kWorkloadSeed[5],
                       threads, out);
}

}  // extern "C"
