// Binning + adjacency intersection with sign parity, for gfx950.
//
// Reference counterpart: the counting loops of
//   Signed_TC/Multicore Implementations/tc_merge_sign_enu.cpp:12-44,131-145
//     (per forward edge (v,w), merge N+(v) and N+(w), prod of three signs)
//   Signed_TC/Multicore Implementations/tc_hash_sign_enu.cpp:124-165
//     (iterate one list, probe a per-vertex table of the other)
//   Signed_TC/GPU Implementations/GlobalMEM/signed_tc.hxx:33-228
//     (warp-cooperative intersection, 4x skew rule)
// Here each oriented edge (u->v) of the rank-space CSR is binned by
// (ceil_log2 of the part of N+(u) ranked above v, ceil_log2 d+(v)); a
// dispatcher maps each of the 256 bins to one strategy:
//   merge-lane  one lane per edge, two-pointer merge (short lists)
//   merge-wave  one wave64 per edge, 64x64 tiles, ds_bpermute binary search
//   hash        one workgroup per source u: N+(u) in an LDS hash table,
//               waves stream N+(v) for u's hash edges and probe
//   bitmap      one workgroup per source u: N+(u) as a 2-bit (present,
//               negative) LDS bitmap over its rank span, waves stream N+(v)
// A triangle {u,v,x} with x in N+(u) and N+(v) has parity
// neg(uv) ^ neg(ux) ^ neg(vx); parity 0 is balanced.  Every kernel reduces
// with wave64 ballot + popcount into 64-bit counters.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace stc {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// Device control block, cleared at the start of every counting run.
struct Ctrl {
  uint32_t bucket_cnt[kBuckets];
  uint32_t done;
  uint32_t len[4];   // merge-lane list, merge-wave list, hash vertices, bitmap vertices
  uint32_t work[8];  // dynamic work counters
  uint32_t pad[3];
  unsigned long long t[4];           // class counters
  unsigned long long strat_edges[8]; // oriented edges per strategy
  uint8_t table[kBuckets];
};
static_assert(sizeof(Ctrl) % 8 == 0, "ctrl alignment");

namespace {

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << __lane_id()) - 1ull;
}

__device__ __forceinline__ int clog2(uint32_t x) {
  return x <= 1 ? 0 : 32 - __clz(x - 1);
}

__device__ __forceinline__ int bucket_of(uint32_t da, uint32_t db) {
  int a = min(clog2(da), kBucketSide - 1), b = min(clog2(db), kBucketSide - 1);
  return a * kBucketSide + b;
}

__device__ __forceinline__ uint8_t merge_variant(int bk, const TuneParams &t) {
  int la = bk / kBucketSide, lb = bk % kBucketSide;
  int lmin = min(la, lb), lmax = max(la, lb);
  return lmax <= t.merge_wave_min_log2[lmin] ? S_MERGE_LANE : S_MERGE_WAVE;
}

__device__ uint8_t fixed_rule(int bk, const TuneParams &t) {
  int la = bk / kBucketSide, lb = bk % kBucketSide;
  int skew = la > lb ? la - lb : lb - la;
  if (skew >= t.fixed_skew_log2) return S_HASH;     // reference 4x skew rule
  if (la >= t.fixed_hub_log2) return S_BITMAP;      // long source list
  return merge_variant(bk, t);
}

__device__ uint8_t mlp_rule(int bk, uint32_t cnt, const DispatchParams &p) {
  const MlpModel &M = p.model;
  float la = (float)(bk / kBucketSide), lb = (float)(bk % kBucketSide);
  float x[kMlpIn] = {la, lb, la - lb, __log2f((float)cnt + 1.0f), p.graph_feat[0],
                     p.graph_feat[1]};
  float h[kMlpHidden];
#pragma unroll
  for (int j = 0; j < kMlpHidden; j++) {
    float acc = M.b1[j];
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) acc += M.w1[j][i] * ((x[i] - M.mean[i]) * M.inv_std[i]);
    h[j] = acc > 0.f ? acc : 0.f;
  }
  int best = 0;
  float bv = -3.0e38f;
#pragma unroll
  for (int k = 0; k < kMlpOut; k++) {
    float acc = M.b2[k];
#pragma unroll
    for (int j = 0; j < kMlpHidden; j++) acc += M.w2[k][j] * h[j];
    if (acc > bv) {
      bv = acc;
      best = k;
    }
  }
  if (best == 1) return S_HASH;
  if (best == 2) return S_BITMAP;
  return merge_variant(bk, p.tune);
}

__device__ uint8_t decide(int bk, uint32_t cnt, const DispatchParams &p) {
  switch (p.mode) {
    case D_FORCE_MERGE:
      return merge_variant(bk, p.tune);
    case D_FORCE_HASH:
      return S_HASH;
    case D_FORCE_BITMAP:
      return S_BITMAP;
    case D_TABLE:
      return p.host_table[bk];
    case D_LEARNED:
      if (p.have_model) return mlp_rule(bk, cnt, p);
      return fixed_rule(bk, p.tune);
    default:
      return fixed_rule(bk, p.tune);
  }
}

// Final strategy of one edge: the bin's choice, demoted when the source list
// does not fit the LDS structure.  Used identically by binning and by the
// vertex-grouped kernels so both agree edge by edge.
__device__ __forceinline__ uint8_t edge_strategy(const uint8_t *table,
                                                 const TuneParams &t,
                                                 uint32_t dfull_a, uint32_t span_a,
                                                 uint32_t da_eff, uint32_t db) {
  if (da_eff == 0 || db == 0) return S_SKIP;
  int bk = bucket_of(da_eff, db);
  uint8_t s = table[bk];
  if (s == S_BITMAP && span_a > (1u << t.bitmap_span_log2)) s = S_HASH;
  if (s == S_HASH && dfull_a > (1u << t.hash_max_log2)) s = merge_variant(bk, t);
  return s;
}

__device__ __forceinline__ uint32_t span_of(const uint32_t *__restrict__ col,
                                            uint32_t la, uint32_t ea) {
  return ea > la ? (col[ea - 1] >> 1) - (col[la] >> 1) + 1u : 0u;
}

// Rank of this lane among lanes of the wave with the same 8-bit value.
__device__ __forceinline__ void wave_match8(uint32_t d, bool valid, uint32_t &rank,
                                            uint32_t &cnt) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    bool bit = (d >> b) & 1u;
    uint64_t bb = __ballot(bit);
    peers &= bit ? bb : ~bb;
  }
  rank = (uint32_t)__popcll(peers & lanemask_lt());
  cnt = (uint32_t)__popcll(peers);
}

// Wave-aggregated append: returns this lane's slot in `list` (valid lanes).
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t *len) {
  uint64_t mask = __ballot(pred);
  uint32_t base = 0;
  if (mask) {
    int leader = __builtin_ctzll(mask);
    if (__lane_id() == leader) base = atomicAdd(len, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
  }
  return base + (uint32_t)__popcll(mask & lanemask_lt());
}

__device__ __forceinline__ void wave_add(unsigned long long *dst, uint64_t v) {
  // v is wave-uniform
  if (__lane_id() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// ---- binning ----------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bin_hist(DevCsr g, Ctrl *ctrl,
                                                  const DispatchParams *P) {
  __shared__ uint32_t h[kBuckets];
  __shared__ int last;
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t mround = (g.m + 63) & ~63ull;  // keep waves converged
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + tid; e < mround; e += stride) {
    bool valid = e < g.m;
    uint32_t bk = 0;
    if (valid) {
      uint32_t a = g.src[e], b = g.col[e] >> 1;
      uint32_t da = g.off[a + 1] - (uint32_t)e - 1u;
      uint32_t db = g.off[b + 1] - g.off[b];
      valid = da && db;
      if (valid) bk = (uint32_t)bucket_of(da, db);
    }
    uint32_t rank, cnt;
    wave_match8(bk, valid, rank, cnt);
    if (valid && rank == 0) atomicAdd(&h[bk], cnt);
  }
  __syncthreads();
  if (h[tid]) atomicAdd(&ctrl->bucket_cnt[tid], h[tid]);
  __threadfence();
  __syncthreads();
  if (tid == 0) last = atomicAdd(&ctrl->done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (last) {
    __threadfence();
    uint32_t cnt = atomicAdd(&ctrl->bucket_cnt[tid], 0u);
    ctrl->table[tid] = cnt ? decide(tid, cnt, *P) : (uint8_t)S_SKIP;
  }
}

__global__ __launch_bounds__(256) void k_bin_scatter(DevCsr g, Ctrl *ctrl,
                                                     const DispatchParams *P,
                                                     uint32_t *__restrict__ lane_list,
                                                     uint32_t *__restrict__ wave_list,
                                                     uint32_t *__restrict__ vflag) {
  __shared__ uint8_t table[kBuckets];
  __shared__ TuneParams tune;
  const int tid = threadIdx.x;
  table[tid] = ctrl->table[tid];
  if (tid == 0) tune = P->tune;
  __syncthreads();
  uint64_t cnt_s[S_COUNT] = {0, 0, 0, 0, 0};
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t mround = (g.m + 63) & ~63ull;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + tid; e < mround; e += stride) {
    uint8_t s = S_SKIP;
    uint32_t a = 0;
    if (e < g.m) {
      a = g.src[e];
      uint32_t b = g.col[e] >> 1;
      uint32_t la = g.off[a], ea = g.off[a + 1];
      uint32_t da = ea - (uint32_t)e - 1u;
      uint32_t db = g.off[b + 1] - g.off[b];
      if (da && db) {
        uint8_t s0 = table[bucket_of(da, db)];
        uint32_t span = s0 == S_BITMAP ? span_of(g.col, la, ea) : 0u;
        s = edge_strategy(table, tune, ea - la, span, da, db);
      }
    }
    uint32_t p = wave_append(s == S_MERGE_LANE, &ctrl->len[0]);
    if (s == S_MERGE_LANE) lane_list[p] = (uint32_t)e;
    p = wave_append(s == S_MERGE_WAVE, &ctrl->len[1]);
    if (s == S_MERGE_WAVE) wave_list[p] = (uint32_t)e;
    if (s == S_HASH || s == S_BITMAP) {
      uint32_t bit = s == S_HASH ? 1u : 2u;
      if (!(vflag[a] & bit)) atomicOr(&vflag[a], bit);
    }
#pragma unroll
    for (int k = 1; k < S_COUNT; k++) cnt_s[k] += __popcll(__ballot(s == k));
    cnt_s[0] += __popcll(__ballot(e < g.m && s == S_SKIP));
  }
#pragma unroll
  for (int k = 0; k < S_COUNT; k++) wave_add(&ctrl->strat_edges[k], cnt_s[k]);
}

// Vertices with hash / bitmap work, in ascending rank, then flags reset.
__global__ void k_vflag_split(uint32_t *__restrict__ vflag, uint32_t n,
                              uint32_t *__restrict__ fh, uint32_t *__restrict__ fb) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    uint32_t f = vflag[v];
    fh[v] = f & 1u;
    fb[v] = (f >> 1) & 1u;
    if (f) vflag[v] = 0;
  }
}

__global__ void k_vlist_scatter(const uint32_t *__restrict__ fh,
                                const uint32_t *__restrict__ ph,
                                const uint32_t *__restrict__ fb,
                                const uint32_t *__restrict__ pb, uint32_t n,
                                uint32_t *__restrict__ vh, uint32_t *__restrict__ vb,
                                Ctrl *ctrl) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += gridDim.x * blockDim.x) {
    if (fh[v]) vh[ph[v]] = v;
    if (fb[v]) vb[pb[v]] = v;
    if (v == n - 1) {
      ctrl->len[2] = ph[v] + fh[v];
      ctrl->len[3] = pb[v] + fb[v];
    }
  }
}

// ---- class accumulation -----------------------------------------------------
// neg = number of negative edges of a found triangle (0..3): class T(3-neg).
template <bool CLS>
struct Acc {
  uint64_t c[4] = {0, 0, 0, 0};
  // found/neg per lane; adds wave-uniform popcounts
  __device__ __forceinline__ void add(bool found, uint32_t neg) {
    if (CLS) {
      c[0] += __popcll(__ballot(found && neg == 3));
      c[1] += __popcll(__ballot(found && neg == 2));
      c[2] += __popcll(__ballot(found && neg == 1));
      c[3] += __popcll(__ballot(found && neg == 0));
    } else {
      c[0] += __popcll(__ballot(found && !(neg & 1)));  // balanced
      c[1] += __popcll(__ballot(found && (neg & 1)));   // unbalanced
    }
  }
  __device__ __forceinline__ void flush(Ctrl *ctrl) {
#pragma unroll
    for (int k = 0; k < 4; k++) wave_add(&ctrl->t[k], c[k]);
  }
};

// ---- merge: one lane per edge ----------------------------------------------
template <bool CLS>
__global__ __launch_bounds__(256) void k_merge_lane(DevCsr g, Ctrl *ctrl,
                                                    const uint32_t *__restrict__ list) {
  const uint32_t len = ctrl->len[0];
  // per-lane class counters; in two-class mode c0 = balanced, c1 = unbalanced
  uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  const uint32_t *__restrict__ col = g.col;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    uint32_t e = list[i];
    uint32_t w = col[e];
    uint32_t a = g.src[e], b = w >> 1, sab = w & 1u;
    uint32_t pa = e + 1, ea = g.off[a + 1];
    uint32_t pb = g.off[b], eb = g.off[b + 1];
    uint32_t x = col[pa], y = col[pb];
    while (true) {
      uint32_t xr = x >> 1, yr = y >> 1;
      if (xr == yr) {
        uint32_t neg = sab + (x & 1u) + (y & 1u);
        if (CLS) {
          c0 += neg == 3;
          c1 += neg == 2;
          c2 += neg == 1;
          c3 += neg == 0;
        } else {
          c0 += (neg & 1u) ^ 1u;
          c1 += neg & 1u;
        }
        if (++pa >= ea || ++pb >= eb) break;
        x = col[pa];
        y = col[pb];
      } else if (xr < yr) {
        if (++pa >= ea) break;
        x = col[pa];
      } else {
        if (++pb >= eb) break;
        y = col[pb];
      }
    }
  }
  uint64_t v[4] = {c0, c1, c2, c3};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    for (int d = 32; d > 0; d >>= 1) v[k] += __shfl_down(v[k], d, 64);
    wave_add(&ctrl->t[k], v[k]);
  }
}

// ---- merge: one wave64 per edge, 64x64 tiles -----------------------------
template <bool CLS>
__global__ __launch_bounds__(256) void k_merge_wave(DevCsr g, Ctrl *ctrl,
                                                    const uint32_t *__restrict__ list) {
  const uint32_t len = ctrl->len[1];
  const uint32_t *__restrict__ col = g.col;
  const int lane = __lane_id();
  const uint32_t nw = gridDim.x * (blockDim.x >> 6);
  Acc<CLS> acc;
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < len; i += nw) {
    const uint32_t e = list[i];
    const uint32_t w = col[e];
    const uint32_t a = g.src[e], b = w >> 1, sab = w & 1u;
    uint32_t ia = e + 1, ea = g.off[a + 1];
    uint32_t ib = g.off[b], eb = g.off[b + 1];
    uint32_t xa = kEmpty, yb = kEmpty;
    uint32_t nA = 0, nB = 0;
    bool needA = true, needB = true;
    while (ia < ea && ib < eb) {
      if (needA) {
        nA = min(64u, ea - ia);
        xa = (uint32_t)lane < nA ? col[ia + lane] : kEmpty;
      }
      if (needB) {
        nB = min(64u, eb - ib);
        yb = (uint32_t)lane < nB ? col[ib + lane] : kEmpty;
      }
      // lower bound of xa among the 64 B entries (padding sorts last)
      const uint32_t key = xa >> 1;
      uint32_t pos = 0;
#pragma unroll
      for (uint32_t step = 32; step > 0; step >>= 1) {
        uint32_t cand = (uint32_t)__shfl((int)yb, (int)(pos + step - 1), 64);
        if ((cand >> 1) < key) pos += step;
      }
      uint32_t hit = (uint32_t)__shfl((int)yb, (int)min(pos, 63u), 64);
      bool found = (uint32_t)lane < nA && pos < nB && (hit >> 1) == key;
      acc.add(found, sab + (xa & 1u) + (hit & 1u));
      uint32_t lastA = (uint32_t)__shfl((int)xa, (int)(nA - 1), 64) >> 1;
      uint32_t lastB = (uint32_t)__shfl((int)yb, (int)(nB - 1), 64) >> 1;
      needA = lastA <= lastB;
      needB = lastB <= lastA;
      if (needA) ia += nA;
      if (needB) ib += nB;
    }
  }
  acc.flush(ctrl);
}

// ---- vertex-grouped: LDS hash table of N+(u) ------------------------------
__device__ __forceinline__ uint32_t fib_hash(uint32_t r, int shift) {
  return (r * 0x9E3779B1u) >> shift;
}

template <bool CLS>
__global__ __launch_bounds__(256) void k_hash(DevCsr g, Ctrl *ctrl,
                                              const DispatchParams *P,
                                              const uint32_t *__restrict__ vlist) {
  extern __shared__ uint32_t T[];
  __shared__ uint8_t table[kBuckets];
  __shared__ uint32_t s_item;
  __shared__ TuneParams tune;
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const int nwaves = blockDim.x >> 6;
  table[tid] = ctrl->table[tid];
  if (tid == 0) tune = P->tune;
  const uint32_t len = ctrl->len[2];
  const uint32_t *__restrict__ col = g.col;
  Acc<CLS> acc;
  while (true) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(&ctrl->work[2], 1u);
    __syncthreads();
    const uint32_t item = s_item;
    if (item >= len) break;
    const uint32_t a = vlist[len - 1 - item];  // highest ranks first
    const uint32_t la = g.off[a], ea = g.off[a + 1], d = ea - la;
    const uint32_t span = span_of(col, la, ea);
    int lg = clog2(2 * d);
    if (lg < 6) lg = 6;
    const uint32_t cap = 1u << lg, mask = cap - 1u;
    const int shift = 32 - lg;
    for (uint32_t k = tid; k < cap; k += blockDim.x) T[k] = kEmpty;
    __syncthreads();
    for (uint32_t k = tid; k < d; k += blockDim.x) {
      uint32_t x = col[la + k];
      uint32_t h = fib_hash(x >> 1, shift);
      while (atomicCAS(&T[h], kEmpty, x) != kEmpty) h = (h + 1) & mask;
    }
    __syncthreads();
    for (uint32_t j = la + wave; j < ea; j += nwaves) {
      const uint32_t w = col[j];
      const uint32_t b = w >> 1, sab = w & 1u;
      const uint32_t pb = g.off[b], eb = g.off[b + 1];
      if (edge_strategy(table, tune, d, span, ea - j - 1, eb - pb) != S_HASH) continue;
      for (uint32_t base = pb; base < eb; base += 64) {
        const bool valid = base + lane < eb;
        const uint32_t y = valid ? col[base + lane] : kEmpty;
        const uint32_t yr = y >> 1;
        uint32_t h = fib_hash(yr, shift);
        uint32_t t = kEmpty;
        bool found = false;
        if (valid) {
          while (true) {
            t = T[h];
            if (t == kEmpty) break;
            if ((t >> 1) == yr) {
              found = true;
              break;
            }
            h = (h + 1) & mask;
          }
        }
        acc.add(found, sab + (y & 1u) + (t & 1u));
      }
    }
  }
  acc.flush(ctrl);
}

// ---- vertex-grouped: LDS 2-bit bitmap over N+(u)'s rank span --------------
template <bool CLS>
__global__ __launch_bounds__(256) void k_bitmap(DevCsr g, Ctrl *ctrl,
                                                const DispatchParams *P,
                                                const uint32_t *__restrict__ vlist) {
  extern __shared__ uint32_t B[];
  __shared__ uint8_t table[kBuckets];
  __shared__ uint32_t s_item;
  __shared__ TuneParams tune;
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const int nwaves = blockDim.x >> 6;
  table[tid] = ctrl->table[tid];
  if (tid == 0) tune = P->tune;
  const uint32_t len = ctrl->len[3];
  const uint32_t *__restrict__ col = g.col;
  Acc<CLS> acc;
  while (true) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(&ctrl->work[3], 1u);
    __syncthreads();
    const uint32_t item = s_item;
    if (item >= len) break;
    const uint32_t a = vlist[len - 1 - item];
    const uint32_t la = g.off[a], ea = g.off[a + 1], d = ea - la;
    const uint32_t lo = col[la] >> 1;
    const uint32_t span = span_of(col, la, ea);
    const uint32_t words = (span + 15) >> 4;
    for (uint32_t k = tid; k < words; k += blockDim.x) B[k] = 0u;
    __syncthreads();
    for (uint32_t k = tid; k < d; k += blockDim.x) {
      uint32_t x = col[la + k];
      uint32_t r = (x >> 1) - lo;
      atomicOr(&B[r >> 4], (1u | ((x & 1u) << 1)) << ((r & 15u) * 2u));
    }
    __syncthreads();
    for (uint32_t j = la + wave; j < ea; j += nwaves) {
      const uint32_t w = col[j];
      const uint32_t b = w >> 1, sab = w & 1u;
      const uint32_t pb = g.off[b], eb = g.off[b + 1];
      if (edge_strategy(table, tune, d, span, ea - j - 1, eb - pb) != S_BITMAP) continue;
      for (uint32_t base = pb; base < eb; base += 64) {
        const bool valid = base + lane < eb;
        const uint32_t y = valid ? col[base + lane] : kEmpty;
        const uint32_t r = (y >> 1) - lo;  // wraps for y below the span
        uint32_t bits = 0;
        if (valid && r < span) bits = B[r >> 4] >> ((r & 15u) * 2u);
        acc.add((bits & 1u) != 0u, sab + (y & 1u) + ((bits >> 1) & 1u));
      }
    }
  }
  acc.flush(ctrl);
}

__global__ void k_fill_words(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

inline int grid_for(uint64_t n, int threads, int cap) {
  uint64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > (uint64_t)cap) g = cap;
  return (int)g;
}

}  // namespace

size_t hash_lds_bytes(const TuneParams &t) {
  return (size_t)4u << std::max(t.hash_max_log2 + 1, 6);
}
size_t bitmap_lds_bytes(const TuneParams &t) {
  return (size_t)4u * ((1u << t.bitmap_span_log2) / 16u);
}

#define STC_TRY(expr)                \
  do {                               \
    hipError_t e_ = (expr);          \
    if (e_ != hipSuccess) return e_; \
  } while (0)

// Buffers: cb.gcount holds Ctrl; cb.lists the merge-lane and merge-wave edge
// lists (m words each); cb.vflag n words kept zero between runs; cb.vlist 6n
// words (hash flags, their scan, bitmap flags, their scan, the two vertex
// lists).
hipError_t enqueue_count(const DevCsr &g, const DispatchParams &hp,
                         const DispatchParams *dp, CountBuffers &cb, bool classes,
                         int num_cus, hipStream_t st, hipEvent_t ev_bin_done) {
  const uint64_t m = g.m;
  const uint32_t n = g.n;
  STC_TRY(cb.gcount.ensure(sizeof(Ctrl)));
  STC_TRY(cb.lists.ensure(2 * (m ? m : 1) * 4));
  if (cb.vflag.bytes < ((uint64_t)n + 1) * 4) {
    STC_TRY(cb.vflag.ensure(((uint64_t)n + 1) * 4));
    hipLaunchKernelGGL(k_fill_words, dim3(grid_for((uint64_t)n + 1, 256, 8192)), dim3(256),
                       0, st, cb.vflag.as<uint32_t>(), cb.vflag.bytes / 4, 0u);
    STC_TRY(hipGetLastError());
  }
  STC_TRY(cb.vlist.ensure(6 * ((uint64_t)n + 1) * 4));
  Ctrl *ctrl = cb.gcount.as<Ctrl>();
  uint32_t *lane_list = cb.lists.as<uint32_t>();
  uint32_t *wave_list = lane_list + (m ? m : 1);
  uint32_t *vf = cb.vflag.as<uint32_t>();
  uint32_t *fh = cb.vlist.as<uint32_t>();
  uint32_t *ph = fh + (n + 1), *fb = ph + (n + 1), *pb = fb + (n + 1);
  uint32_t *vh = pb + (n + 1), *vb = vh + (n + 1);

  hipLaunchKernelGGL(k_fill_words, dim3(4), dim3(256), 0, st, (uint32_t *)ctrl,
                     (uint64_t)(sizeof(Ctrl) / 4), 0u);
  STC_TRY(hipGetLastError());
  const int bin_grid = grid_for(m, 256, num_cus * 8);
  hipLaunchKernelGGL(k_bin_hist, dim3(bin_grid), dim3(256), 0, st, g, ctrl, dp);
  STC_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bin_scatter, dim3(bin_grid), dim3(256), 0, st, g, ctrl, dp,
                     lane_list, wave_list, vf);
  STC_TRY(hipGetLastError());
  if (n > 0) {
    const int vgrid = grid_for(n, 256, num_cus * 8);
    hipLaunchKernelGGL(k_vflag_split, dim3(vgrid), dim3(256), 0, st, vf, n, fh, fb);
    STC_TRY(hipGetLastError());
    STC_TRY(scan_exclusive_u32(fh, ph, n, cb.scan, st));
    STC_TRY(scan_exclusive_u32(fb, pb, n, cb.scan, st));
    hipLaunchKernelGGL(k_vlist_scatter, dim3(vgrid), dim3(256), 0, st, fh, ph, fb, pb, n,
                       vh, vb, ctrl);
    STC_TRY(hipGetLastError());
  }
  if (ev_bin_done) STC_TRY(hipEventRecord(ev_bin_done, st));
  const int grid = num_cus * 8;
  const size_t hash_lds = hash_lds_bytes(hp.tune);
  const size_t bm_lds = bitmap_lds_bytes(hp.tune);
  const int hash_grid = num_cus * (int)std::max<size_t>(1, std::min<size_t>(8, (96 * 1024) / hash_lds));
  const int bm_grid = num_cus * (int)std::max<size_t>(1, std::min<size_t>(8, (128 * 1024) / bm_lds));
  if (classes) {
    hipLaunchKernelGGL(k_merge_lane<true>, dim3(grid), dim3(256), 0, st, g, ctrl, lane_list);
    hipLaunchKernelGGL(k_merge_wave<true>, dim3(grid), dim3(256), 0, st, g, ctrl, wave_list);
    hipLaunchKernelGGL(k_hash<true>, dim3(hash_grid), dim3(256), hash_lds, st, g, ctrl, dp, vh);
    hipLaunchKernelGGL(k_bitmap<true>, dim3(bm_grid), dim3(256), bm_lds, st, g, ctrl, dp, vb);
  } else {
    hipLaunchKernelGGL(k_merge_lane<false>, dim3(grid), dim3(256), 0, st, g, ctrl, lane_list);
    hipLaunchKernelGGL(k_merge_wave<false>, dim3(grid), dim3(256), 0, st, g, ctrl, wave_list);
    hipLaunchKernelGGL(k_hash<false>, dim3(hash_grid), dim3(256), hash_lds, st, g, ctrl, dp, vh);
    hipLaunchKernelGGL(k_bitmap<false>, dim3(bm_grid), dim3(256), bm_lds, st, g, ctrl, dp, vb);
  }
  return hipGetLastError();
}

hipError_t read_count(CountBuffers &cb, bool classes, CountResult &r, hipStream_t st) {
  Ctrl h;
  STC_TRY(hipMemcpyAsync(&h, cb.gcount.p, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  STC_TRY(hipStreamSynchronize(st));
  for (int k = 0; k < 4; k++) r.t[k] = h.t[k];
  (void)classes;
  for (int k = 0; k < S_COUNT; k++) r.edges_by_strat[k] = h.strat_edges[k];
  for (int k = 0; k < kBuckets; k++) {
    r.table[k] = h.table[k];
    r.bucket_cnt[k] = h.bucket_cnt[k];
  }
  return hipSuccess;
}

}  // namespace stc
