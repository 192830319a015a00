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
//   bitmap      one workgroup (or wave) per source u: N+(u) as an LDS bitmap
//               over a rank window, two bits per rank one-hot by sign, waves
//               stream N+(v)
//   pull        one workgroup (or wave) per chunk of the in-edges of a target
//               v: N+(v) as such a bitmap, waves stream the part of N+(u)
//               after v for each in-edge (u->v)
// A triangle {u,v,x} with x in N+(u) and N+(v) has parity
// neg(uv) ^ neg(ux) ^ neg(vx); parity 0 is balanced.  Every kernel reduces
// with wave64 ballot + popcount into 64-bit counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <vector>

#include "internal.h"

namespace stc {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// Source lists of at most kWaveMaxDeg entries are hashed by one wave; the
// bitmap kernel takes source lists of kWaveMaxDeg+1 .. kBitmapMaxDeg entries.
constexpr uint32_t kWaveMaxDeg = 64;
constexpr uint32_t kBitmapMaxDeg = 2048;
// Entries of a workgroup-bitmap source list that fall below its rank window
// are kept as a sorted LDS array of at most kLowCap words.
constexpr uint32_t kLowCap = 1024;
// Zero words after every LDS bitmap: a probe past the window reads the first
// of them (four keep what follows 16-byte aligned).
constexpr uint32_t kGuard = 4;
constexpr int kVertexBlock = 256;   // threads per workgroup of the hash kernel
constexpr int kBitmapBlock = 512;   // threads per workgroup of the bitmap kernel

// Device control block, cleared at the start of every counting run.
// Copies of the bucket histogram that k_bin_hist's workgroups add into (by
// blockIdx mod kHistSlots), so that no counter takes more than 1/32 of the
// workgroups' atomics; k_bin_table sums them into bucket_cnt.
constexpr int kHistSlots = 32;
// Copies of the class counters, one 128-byte line each: a wave adds its
// totals into copy blockIdx mod kTSlots (thousands of waves end together, and
// one counter pair took all of their atomics); the host sums the copies.
constexpr int kTSlots = 64;
constexpr int kTStride = 16;

struct Ctrl {
  uint32_t bucket_cnt[kBuckets];
  uint32_t len[8];   // merge-lane list, merge-wave list, then vertex lists:
                     // workgroup hash, workgroup bitmap, wave hash, wave bitmap,
                     // workgroup pull, wave pull
  uint32_t work[8];  // dynamic work counters
  uint32_t pad[2];
  uint32_t queue[7][8][32];          // XCD work queues (workgroup hash,
                                     // workgroup bitmap, wave hash, batch
                                     // phase 1, wave bitmap or batch phase
                                     // 2, workgroup pull, wave pull; a single
                                     // graph's two-phase pull launch uses the
                                     // last two): one counter per XCD, 128
                                     // bytes apart
  unsigned long long t[kTSlots][kTStride];  // class counters: copies, entries 0..3 each
  unsigned long long kern_edges[10]; // oriented edges counted per kernel
                                     // (launch order); slot 8: left out by
                                     // a partial table, slot 9: skipped
  unsigned long long kern_words[10]; // sum of d+(u) + d+(v) over those edges
  unsigned long long kern_bytes[10]; // bytes each kernel's strategy reads (see k_edge_stats)
  uint8_t table[kBuckets];
  // last: the host reads the control block up to here
  uint32_t bucket_part[kHistSlots][kBuckets];
};
static_assert(sizeof(Ctrl) % 8 == 0, "ctrl alignment");

namespace {

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << __lane_id()) - 1ull;
}

// Wave-uniform read of p[i] (i uniform) through the constant address space:
// a scalar load, counted on lgkmcnt, so it never waits behind the wave's
// outstanding vector loads.  Only for arrays no kernel writes while counting.
__device__ __forceinline__ uint32_t sload(const uint32_t *p, uint32_t i) {
  typedef const __attribute__((address_space(4))) uint32_t cu32;
  return ((const cu32 *)p)[i];
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

// The fixed rule: the fallback when no model is loaded, and the judge of the
// learned dispatcher's speed.  Its thresholds are tuned (autotune.py) on the
// dispatcher's training graphs; 16 switches the pull, skew and hub branches
// off (la, lb <= 15), -1 the merge branch, and a negative pull threshold
// also pulls where the suffix is the longer list.
//   1. Both lists short: merge (one lane, or one wave, per edge).
//   2. The source's suffix shorter than the target's list by 2^fixed_pull:
//      pull, which streams the suffix against the staged target list -- the
//      reference's hash path iterates the shorter of the two lists against a
//      table of the longer (tc_hash_sign_enu.cpp:138-144).
//   3. Lengths apart by 2^fixed_skew (2^2: the reference GPU kernel's 4x rule,
//      signed_tc.hxx:47,65, whose lanes binary-search the longer list): the
//      wave merge, whose lanes binary-search a 64-entry tile of the other list.
//   4. A long source suffix: push bitmap (the source list staged once, the
//      target lists streamed).
//   5. Otherwise merge.
__device__ uint8_t fixed_rule(int bk, const TuneParams &t) {
  const int la = bk / kBucketSide, lb = bk % kBucketSide;
  const int skew = la > lb ? la - lb : lb - la;
  if (max(la, lb) <= t.fixed_merge_log2) return merge_variant(bk, t);
  if (lb - la >= t.fixed_pull_log2) return S_PULL;
  if (skew >= t.fixed_skew_log2) return S_MERGE_WAVE;
  if (la >= t.fixed_hub_log2) return S_BITMAP;
  return merge_variant(bk, t);
}

__device__ uint8_t mlp_rule(int bk, uint32_t cnt, const DispatchParams &p) {
  const MlpModel &M = p.model;
  float la = (float)(bk / kBucketSide), lb = (float)(bk % kBucketSide);
  float x[kMlpIn] = {la, lb, la - lb, __log2f((float)cnt + 1.0f), p.graph_feat[0],
                     p.graph_feat[1], p.graph_feat[2]};
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
  if (best == 3) return S_PULL;
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
    case D_FORCE_PULL:
      return S_PULL;
    case D_TABLE:  // host table: 0 skip, 1 merge, 2 hash, 3 bitmap, 4 pull
      switch (p.host_table[bk]) {
        case 1: return merge_variant(bk, p.tune);
        case 2: return S_HASH;
        case 3: return S_BITMAP;
        case 4: return S_PULL;
        default: return S_SKIP;
      }
    case D_LEARNED:
      if (p.have_model) return mlp_rule(bk, cnt, p);
      return fixed_rule(bk, p.tune);
    default:
      return fixed_rule(bk, p.tune);
  }
}

// A source list of d entries spanning span ranks fits the bitmap kernels when
// it is short enough for one wave (any span: per-wave window plus at most d
// low entries), or when its entries below the workgroup window fit kLowCap
// (at most d - 1 of them lie below the window, none when span <= window).
// span() is evaluated only when the answer depends on it.
template <class Span>
__device__ __forceinline__ bool bitmap_fits(const TuneParams &t, uint32_t d, const Span &span) {
  if (d <= kWaveMaxDeg) return true;
  if (d > kBitmapMaxDeg) return false;
  return d - 1u <= kLowCap || span() <= (1u << t.bitmap_span_log2);
}

// A target list fits the pull kernels under the rule of the bitmap kernels:
// at most kLowCap entries (all of its entries below any window then fit the
// low array), or at most kBitmapMaxDeg entries over a rank span within the
// workgroup window (no entry below it).  The bucket's target class clog2(db)
// decides the first case (db <= 1024 <=> class <= 10) without reading the
// list; span_b() (the target list's rank span) is read only for class 11.
// Every in-edge of one target gets the same answer.
constexpr int kPullMaxClass = 10;  // clog2(kLowCap)
static_assert((1u << kPullMaxClass) == kLowCap, "pull fit class");
static_assert((2u << kPullMaxClass) == kBitmapMaxDeg, "pull fit by span: one class above");

template <class SpanB>
__device__ __forceinline__ bool pull_fits(const TuneParams &t, int bk, const SpanB &span_b) {
  if (!t.allow_pull) return false;
  const int cls = bk % kBucketSide;
  if (cls <= kPullMaxClass) return true;
  return cls == kPullMaxClass + 1 && span_b() <= (1u << t.bitmap_span_log2);
}

// Final strategy of an edge in bucket bk: the bin's choice, demoted when the
// list it stages does not fit the LDS structure -- pull (target list, rank
// span span_b) to bitmap, bitmap (source list of dfull_a entries over span_a
// ranks) to hash, hash to merge.  Binning (from the stored bucket) and the
// vertex-grouped kernels (through edge_strategy) apply the same rule, so both
// agree edge by edge.  The spans are read only when the chain reaches their
// test (span_of below).
template <class SpanA, class SpanB>
__device__ __forceinline__ uint8_t bucket_strategy(const uint8_t *table, const TuneParams &t,
                                                   uint32_t dfull_a, const SpanA &span_a,
                                                   const SpanB &span_b, int bk) {
  uint8_t s = table[bk];
  if (s == S_PULL && !pull_fits(t, bk, span_b)) s = S_BITMAP;
  if (s == S_BITMAP && !bitmap_fits(t, dfull_a, span_a)) s = S_HASH;
  if (s == S_HASH && dfull_a > (1u << t.hash_max_log2)) s = merge_variant(bk, t);
  return s;
}

__device__ __forceinline__ uint32_t span_of(const uint32_t *__restrict__ col,
                                            uint32_t la, uint32_t ea) {
  return ea > la ? (col[ea - 1] >> 1) - (col[la] >> 1) + 1u : 0u;
}

// Strategy of edge (a -> b) seen from a vertex-grouped kernel: the source
// list has dfull_a entries over span_a ranks, da_eff of them after b; b's
// list is col[pb, eb).
__device__ __forceinline__ uint8_t edge_strategy(const uint8_t *table, const TuneParams &t,
                                                 uint32_t dfull_a, uint32_t span_a,
                                                 uint32_t da_eff, const uint32_t *col,
                                                 uint32_t pb, uint32_t eb) {
  const uint32_t db = eb - pb;
  if (da_eff == 0 || db == 0) return S_SKIP;
  return bucket_strategy(table, t, dfull_a, [&] { return span_a; },
                         [&] { return span_of(col, pb, eb); }, bucket_of(da_eff, db));
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

// ---- XCD-aware work queue ---------------------------------------------------
// `len` items are split over the 8 XCDs (workgroups are dispatched round-robin
// over the 8 XCDs of an MI355X, so blockIdx % 8 names the XCD).  Each XCD's
// share has its own counter on its own 128-byte line and is drained from the
// top (highest ranks, the heaviest vertices, first); a drained share sends the
// caller on to the next XCD's.  Returns false when every share is empty.
// (Kernels read their list length first and exit at once when it is zero.)
//   Contiguous shares (workgroup kernels: one vertex per pop) keep a range of
// vertices on one XCD's L2.  INTERLEAVE deals pops of `chunk` items out
// round-robin instead (pop c of the list to XCD c % 8; wave kernels), so the
// heavy top of a list is spread over all XCDs: it took the wave-pull kernel
// from 0.70 to 0.59 ms on RMAT-22.
struct QueueCursor {
  uint32_t q, tried;
};

template <bool INTERLEAVE = false>
__device__ __forceinline__ bool queue_pop(uint32_t (*ctr)[32], uint32_t len, uint32_t chunk,
                                          QueueCursor &cur, uint32_t &first,
                                          uint32_t &count) {
  while (cur.tried < 8) {
    if constexpr (INTERLEAVE) {
      const uint32_t npops = (len + chunk - 1) / chunk;
      const uint32_t mine = npops > cur.q ? (npops - cur.q + 7) / 8 : 0u;  // pops q, q+8, ...
      const uint32_t k = atomicAdd(&ctr[cur.q][0], 1u);
      if (k < mine) {
        first = (npops - 1u - (cur.q + 8u * k)) * chunk;  // from the top
        count = min(chunk, len - first);
        return true;
      }
    } else {
      const uint32_t lo = (uint32_t)(((uint64_t)len * cur.q) >> 3);
      const uint32_t hi = (uint32_t)(((uint64_t)len * (cur.q + 1)) >> 3);
      const uint32_t k = atomicAdd(&ctr[cur.q][0], chunk);
      if (k < hi - lo) {
        count = min(chunk, hi - lo - k);
        first = hi - k - count;  // items [first, first + count), taken from the top
        return true;
      }
    }
    cur.q = (cur.q + 1) & 7u;
    cur.tried++;
  }
  return false;
}

// ---- binning ----------------------------------------------------------------
// Bucket histogram from the edge buckets the build stored (two bytes per
// edge, in edge order, read as 16-byte groups of eight): no edge needs a
// gather.  Read from the in-edge records (eight bytes per edge, the target's
// class packed beside the suffix length) the pass moved four times the
// bytes; gathering the target's class per edge was bound by the L2's
// request rate (RMAT-22: 0.32 ms for 64 M edges).
//   The workgroup keeps kHistCopies copies of each bucket's counter and lane
// l adds to copy l % kHistCopies, so at most 8 lanes of a wave add to one
// LDS word: the lanes of a wave often share a bucket (the edges of one
// source), and one counter per bucket would serialize them.  Grouping the
// lanes by bucket first (an 8-round ballot match) cost about 55 vector
// instructions per round and left the pass bound by the vector unit.
constexpr uint32_t kHistCopies = 8;

__global__ __launch_bounds__(256, 8) void k_bin_hist(DevCsr g, Ctrl *ctrl) {
  __shared__ uint32_t h[kBuckets * kHistCopies];
  const int tid = threadIdx.x;
  for (uint32_t k = tid; k < kBuckets * kHistCopies; k += 256) h[k] = 0;
  __syncthreads();
  const uint32_t copy = __lane_id() % kHistCopies;
  const uint4 *q = (const uint4 *)g.ebk;  // 8 edge buckets per group
  const uint64_t ng = (g.m + 7) >> 3;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + tid; i < ng; i += stride) {
    const uint4 x = q[i];
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t bk = (w[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
      if (bk != kNoEdgeBucket) atomicAdd(&h[bk * kHistCopies + copy], 1u);
    }
  }
  __syncthreads();
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t c = 0; c < kHistCopies; c++) sum += h[tid * kHistCopies + c];
  if (sum) atomicAdd(&ctrl->bucket_part[blockIdx.x % kHistSlots][tid], sum);
}

// The strategy table from the bucket histogram (one workgroup, thread =
// bucket), after every workgroup of k_bin_hist has added its counts: a kernel
// of its own, because the last-workgroup-decides pattern cost every workgroup
// a device-scope fence.
__global__ __launch_bounds__(kBuckets) void k_bin_table(Ctrl *ctrl, const DispatchParams *P) {
  const int tid = threadIdx.x;
  uint32_t cnt = 0;
#pragma unroll
  for (int s = 0; s < kHistSlots; s++) cnt += ctrl->bucket_part[s][tid];
  ctrl->bucket_cnt[tid] = cnt;
  ctrl->table[tid] = cnt ? decide(tid, cnt, *P) : (uint8_t)S_SKIP;
}

// Edge lists are written per workgroup: workgroup b owns the segment
// [b * seg, (b + 1) * seg) of each list (seg = the most edges one workgroup
// visits) and appends through an LDS counter; k_list_compact then packs the
// segments.  No global atomic sits on a hot address.
__device__ __forceinline__ uint32_t lds_append(bool pred, uint32_t *s_len) {
  uint64_t mask = __ballot(pred);
  uint32_t base = 0;
  if (__lane_id() == 0 && mask) base = atomicAdd(s_len, (uint32_t)__popcll(mask));
  base = __shfl(base, 0, 64);
  return base + (uint32_t)__popcll(mask & lanemask_lt());
}

// Strategy of every edge from its bucket; merge edges go to the
// workgroup's list segments, hash and bitmap edges flag their source vertex
// (bit 0 hash, bit 1 bitmap).  Edges are sorted by source, so the lanes of one
// source form runs: each run's first lane sets the run's flag bits with one
// atomic, only when they are not set yet.  Pull edges need nothing here: the
// vertex-list pass finds their targets from each target's suffix-class mask
// (a per-edge flag on the target cost one device-scope atomic per pull edge:
// 23 ms of a 92 ms batch of 605 M edges).  A table whose buckets are all
// skipped or pull of target classes that always fit (class <= 10, or 11 when
// the prepared graph's class-11 lists all fit the window) leaves nothing to
// scatter, so the workgroups return at once without reading an edge; else an
// edge that such a bucket settles needs no list offsets.
__global__ __launch_bounds__(256) void k_bin_scatter(DevCsr g, Ctrl *ctrl,
                                                     const DispatchParams *P,
                                                     uint32_t *__restrict__ lane_seg,
                                                     uint32_t *__restrict__ wave_seg,
                                                     uint32_t seg,
                                                     uint32_t *__restrict__ blk_cnt,
                                                     uint32_t *__restrict__ vflag) {
  __shared__ uint8_t table[kBuckets];
  __shared__ TuneParams tune;
  __shared__ uint32_t s_len[2];
  const int tid = threadIdx.x;
  const uint32_t lane = __lane_id();
  const uint8_t mine = ctrl->table[tid];
  table[tid] = mine;
  if (tid == 0) {
    tune = P->tune;
    s_len[0] = s_len[1] = 0;
  }
  // pull of a target class whose every list fits the pull kernels
  const uint32_t fit_cls = kPullMaxClass + (P->pull_c11_fits ? 1u : 0u);
  // a bucket that lists merge edges or flags sources, or may demote pull
  const bool other = mine != S_SKIP && (mine != S_PULL || tid % kBucketSide > (int)fit_cls ||
                                        !P->tune.allow_pull);
  if (!__syncthreads_or(other)) {
    if (tid == 0) blk_cnt[blockIdx.x] = blk_cnt[gridDim.x + blockIdx.x] = 0u;
    return;
  }
  uint32_t *lane_out = lane_seg + (uint64_t)blockIdx.x * seg;
  uint32_t *wave_out = wave_seg + (uint64_t)blockIdx.x * seg;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t mround = (g.m + 63) & ~63ull;  // keep waves converged
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + tid; e < mround; e += stride) {
    uint8_t s = S_SKIP;
    if (e < g.m) {
      const uint32_t bk = g.ebk[e];
      if (bk != kNoEdgeBucket) {
        s = table[bk];
        // merge and skipped buckets keep their strategy, and so does pull of
        // a target class that always fits; the others may be demoted
        if (s == S_HASH || s == S_BITMAP ||
            (s == S_PULL && (bk % kBucketSide > fit_cls || !tune.allow_pull))) {
          const uint32_t a = g.src[e];
          const uint32_t la = g.off[a], ea = g.off[a + 1];
          s = bucket_strategy(
              table, tune, ea - la, [&] { return span_of(g.col, la, ea); },
              [&] {
                const uint32_t b = g.col[e] >> 1;
                return span_of(g.col, g.off[b], g.off[b + 1]);
              },
              (int)bk);
        }
      }
    }
    uint32_t p = lds_append(s == S_MERGE_LANE, &s_len[0]);
    if (s == S_MERGE_LANE) lane_out[p] = (uint32_t)e;
    p = lds_append(s == S_MERGE_WAVE, &s_len[1]);
    if (s == S_MERGE_WAVE) wave_out[p] = (uint32_t)e;
    const uint64_t hb = __ballot(s == S_HASH), bb = __ballot(s == S_BITMAP);
    if ((hb | bb) == 0ull) continue;  // wave-uniform: no source to flag
    const uint32_t a = e < g.m ? g.src[e] : 0xFFFFFFFFu;
    // runs of equal sources: heads, and each head's run as a lane mask
    const uint32_t prev = (uint32_t)__shfl_up((int)a, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || prev != a);
    if ((heads >> lane) & 1u) {
      const uint64_t above = heads & ~((2ull << lane) - 1ull);
      const uint64_t run = (above ? (1ull << __builtin_ctzll(above)) - 1ull : ~0ull) &
                           ~((1ull << lane) - 1ull);
      const uint32_t bits = ((hb & run) ? 1u : 0u) | ((bb & run) ? 2u : 0u);
      if (bits && (vflag[a] & bits) != bits) atomicOr(&vflag[a], bits);
    }
  }
  __syncthreads();
  if (tid == 0) {
    blk_cnt[blockIdx.x] = s_len[0];
    blk_cnt[gridDim.x + blockIdx.x] = s_len[1];
  }
}

// Pack the per-workgroup segments: blk_pos holds the exclusive scan of
// blk_cnt (2 * nblk entries: merge-lane counts, then merge-wave counts).
__global__ __launch_bounds__(256) void k_list_compact(const uint32_t *__restrict__ lane_seg,
                                                      const uint32_t *__restrict__ wave_seg,
                                                      uint32_t seg,
                                                      const uint32_t *__restrict__ blk_cnt,
                                                      const uint32_t *__restrict__ blk_pos,
                                                      uint32_t nblk,
                                                      uint32_t *__restrict__ lane_list,
                                                      uint32_t *__restrict__ wave_list,
                                                      Ctrl *ctrl) {
  const uint32_t b = blockIdx.x % nblk;
  const bool lanes = blockIdx.x < nblk;
  const uint32_t *src = (lanes ? lane_seg : wave_seg) + (uint64_t)b * seg;
  uint32_t *dst = lanes ? lane_list : wave_list;
  const uint32_t cnt = blk_cnt[blockIdx.x];
  const uint32_t base = blk_pos[blockIdx.x] - (lanes ? 0u : blk_pos[nblk]);
  for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) dst[base + i] = src[i];
  if (blockIdx.x == nblk - 1 && threadIdx.x == 0) ctrl->len[0] = blk_pos[nblk - 1] + cnt;
  if (blockIdx.x == 2 * nblk - 1 && threadIdx.x == 0)
    ctrl->len[1] = blk_pos[2 * nblk - 1] + cnt - blk_pos[nblk];
}

// Vertices with hash / bitmap / pull work go to six lists in ascending rank:
// workgroup hash, workgroup bitmap, wave hash, wave bitmap, workgroup pull,
// wave pull (vertices with at most kWaveMaxDeg out-edges use the wave
// kernels).  For pull the listed vertex is the target v whose list is
// staged, and its entries are chunks of its in-edge range [tin_off[v],
// tin_off[v + 1]): one entry per 2^pull_block_chunk_log2 (workgroup) or
// 2^pull_wave_chunk_log2 (wave) in-edges (tuned; 2048 and 256 by default),
// holding the chunk's first position, so a
// target with a huge in-degree is spread over many workgroups.  A workgroup
// owns a tile of kVTile consecutive vertices, visited in rounds of 256:
// k_vcount counts each list's entries per tile, an exclusive scan over the
// tile counts gives each tile's list offsets, and k_vscatter writes the
// entries in order and clears the flags.
constexpr uint32_t kVTile = 4096;
constexpr int kVLists = 6;
// The smallest pull chunk the tune table admits (it bounds the list sizes).
constexpr uint32_t kPullChunkMin = 32;

// The strategy table by target class: pull_src[cb] has bit c set when bucket
// (c, cb) is binned to pull (built in LDS by every workgroup of the pass;
// any_pull_in reads it back after the barrier that follows).
__device__ __forceinline__ void pull_sources_of(const Ctrl *ctrl, uint16_t *pull_src) {
  const uint32_t cb = threadIdx.x;
  if (cb < (uint32_t)kBucketSide) {
    uint32_t msk = 0;
    for (int c = 0; c < kBucketSide; c++)
      if (ctrl->table[c * kBucketSide + cb] == S_PULL) msk |= 1u << c;
    pull_src[cb] = (uint16_t)msk;
  }
}

__device__ __forceinline__ bool any_pull_in(const uint16_t *pull_src) {
  uint32_t any = 0;
#pragma unroll
  for (int c = 0; c < kBucketSide; c++) any |= pull_src[c];
  return any != 0;
}

// Entries of each list a vertex adds (0 or 1, pull lists one per chunk).  A
// target is a pull vertex when some in-edge lies in a bucket binned to pull
// (its suffix-class mask meets pull_src of its own class) and its list fits
// the pull kernels -- exactly when some in-edge's strategy is pull, since
// the fit test reads the target alone.  A pull target's chunks are wave
// items when its list has at most kWaveMaxDeg entries, or at most
// 2^wave_max_log2 entries over a rank span within the wave window (no entry
// then falls below the window); else workgroup items.
__device__ __forceinline__ void vertex_lists_of(uint32_t v, const DevCsr &g, const uint32_t *vflag,
                                                const uint16_t *pull_src, bool any_pull,
                                                const TuneParams &t, uint32_t wave_max_log2,
                                                uint32_t c[kVLists], uint32_t &f) {
  f = 0;
  bool small = false, small_pull = false, pull = false;
  uint32_t indeg = 0;
  if (v < g.n) {
    f = vflag[v];
    const uint32_t cb = any_pull ? (uint32_t)g.dcls[v] : 0xFFu;  // no pull bucket: no reads
    if (cb != 0xFFu && (g.smask[v] & pull_src[cb]))
      pull = pull_fits(t, (int)cb, [&] { return span_of(g.col, g.off[v], g.off[v + 1]); });
    if (f || pull) {
      const uint32_t d = g.off[v + 1] - g.off[v];
      small = d <= kWaveMaxDeg;
      small_pull = small;
      if (pull && !small && d <= (1u << wave_max_log2))
        small_pull = span_of(g.col, g.off[v], g.off[v + 1]) <= (1u << t.wave_window_log2);
    }
    if (pull) indeg = g.tin_off[v + 1] - g.tin_off[v];
  }
  c[0] = (f & 1u) && !small;
  c[1] = (f & 2u) && !small;
  c[2] = (f & 1u) && small;
  c[3] = (f & 2u) && small;
  const uint32_t cb_ = 1u << t.pull_block_chunk_log2, cw = 1u << t.pull_wave_chunk_log2;
  c[4] = pull && !small_pull ? (indeg + cb_ - 1) / cb_ : 0u;
  c[5] = pull && small_pull ? (indeg + cw - 1) / cw : 0u;
}

// Inclusive sum over the wave (every lane active) by DPP: row_shr 1, 2, 4, 8
// scan each row of 16 lanes, row_bcast 15 adds row 0's total to row 1 and
// row 2's to row 3, row_bcast 31 adds the first two rows' total to rows 2
// and 3.  No LDS round trip sits between the steps (a lane-shuffle scan takes
// six dependent ds_bpermute reads).
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// Sum over the wave (every lane active), wave-uniform.
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_incl(x), 63);
}

// Upper bound of the entries of one list: every vertex once plus the extra
// chunks of the pull lists (at most m / kPullChunkMin).
__host__ __device__ inline uint64_t vlist_stride(uint32_t n, uint64_t m) {
  return (uint64_t)n + 1 + m / kPullChunkMin + 1;
}

__global__ __launch_bounds__(256) void k_vcount(DevCsr g, const Ctrl *ctrl,
                                                const DispatchParams *P,
                                                const uint32_t *__restrict__ vflag,
                                                uint32_t *__restrict__ tile_cnt) {
  __shared__ uint32_t s_cnt[kVLists];
  __shared__ uint16_t s_pull[kBucketSide];
  __shared__ TuneParams tune;
  const uint32_t tid = threadIdx.x, ntiles = gridDim.x;
  if (tid < kVLists) s_cnt[tid] = 0;
  if (tid == 0) tune = P->tune;
  pull_sources_of(ctrl, s_pull);
  // a table of merge buckets only flags no vertex (demotion runs towards
  // merge): every list is empty and no flag needs reading
  const uint8_t st = ctrl->table[tid];
  const bool listed =
      __syncthreads_or(st == S_HASH || st == S_BITMAP || st == S_PULL) != 0;
  if (!listed) {
    if (tid < kVLists) tile_cnt[tid * ntiles + blockIdx.x] = 0u;
    return;
  }
  const bool any_pull = any_pull_in(s_pull);
  const uint32_t wmax = P->wave_pull_max_log2;
  uint32_t cnt[kVLists] = {0, 0, 0, 0, 0, 0};
  const uint32_t base = blockIdx.x * kVTile;
#pragma unroll 4  // rounds are independent: their loads overlap
  for (uint32_t r = 0; r < kVTile; r += 256) {
    uint32_t c[kVLists];
    uint32_t f;
    vertex_lists_of(base + r + tid, g, vflag, s_pull, any_pull, tune, wmax, c, f);
#pragma unroll
    for (int L = 0; L < kVLists; L++) cnt[L] += c[L];
  }
#pragma unroll
  for (int L = 0; L < kVLists; L++)
    if (cnt[L]) atomicAdd(&s_cnt[L], cnt[L]);
  __syncthreads();
  if (tid < kVLists) tile_cnt[tid * ntiles + blockIdx.x] = s_cnt[tid];
}

// tile_pos: exclusive scan of tile_cnt (list-major, kVLists * ntiles
// entries); list L goes to vlist + L * stride.  A tile that adds no entry to
// any list has no flagged vertex either (every flag adds an entry), so its
// workgroup returns at once: on a graph counted by merge alone every tile does.
__global__ __launch_bounds__(256) void k_vscatter(DevCsr g, const DispatchParams *P,
                                                  uint32_t *__restrict__ vflag,
                                                  const uint32_t *__restrict__ tile_cnt,
                                                  const uint32_t *__restrict__ tile_pos,
                                                  uint32_t *__restrict__ vlist, uint64_t stride,
                                                  Ctrl *ctrl) {
  __shared__ uint32_t s_wcnt[kVLists][4];  // [list][wave] entries in this round
  __shared__ uint32_t s_run[kVLists];      // entries of earlier rounds in this tile
  __shared__ uint16_t s_pull[kBucketSide];
  __shared__ TuneParams tune;
  const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const uint32_t ntiles = gridDim.x, b = blockIdx.x;
  uint32_t tile_off[kVLists];
  uint32_t any = 0;
#pragma unroll
  for (int L = 0; L < kVLists; L++) {
    tile_off[L] = tile_pos[L * ntiles + b] - tile_pos[L * ntiles];
    any |= tile_cnt[L * ntiles + b];
  }
  if (any == 0u) {  // workgroup-uniform: no vertex of the tile is flagged or listed
    if (b == ntiles - 1 && tid < kVLists) {
      const uint32_t last = tid * ntiles + ntiles - 1;
      ctrl->len[2 + tid] = tile_pos[last] + tile_cnt[last] - tile_pos[tid * ntiles];
    }
    return;
  }
  if (tid < kVLists) s_run[tid] = 0;
  if (tid == 0) tune = P->tune;
  pull_sources_of(ctrl, s_pull);
  __syncthreads();
  const bool any_pull = any_pull_in(s_pull);
  const uint32_t wmax = P->wave_pull_max_log2;
  const uint32_t base = b * kVTile;
  // the next round's lists are found while this round's are written (their
  // loads are in flight across this round's barriers)
  uint32_t cn[kVLists], fn;
  vertex_lists_of(base + tid, g, vflag, s_pull, any_pull, tune, wmax, cn, fn);
  for (uint32_t r = 0; r < kVTile; r += 256) {
    const uint32_t v = base + r + tid;
    uint32_t c[kVLists];
#pragma unroll
    for (int L = 0; L < kVLists; L++) c[L] = cn[L];
    const uint32_t f = fn;
    if (r + 256 < kVTile)
      vertex_lists_of(v + 256, g, vflag, s_pull, any_pull, tune, wmax, cn, fn);
    uint32_t incl[kVLists];
#pragma unroll
    for (int L = 0; L < kVLists; L++) {
      incl[L] = wave_scan_incl(c[L]);
      if (lane == 63) s_wcnt[L][wave] = incl[L];
    }
    __syncthreads();
#pragma unroll
    for (int L = 0; L < kVLists; L++) {
      if (c[L]) {
        uint32_t p = tile_off[L] + s_run[L] + incl[L] - c[L];
        for (uint32_t w = 0; w < wave; w++) p += s_wcnt[L][w];
        uint32_t *dst = vlist + (uint64_t)L * stride + p;
        if (L < 4) {
          dst[0] = v;
        } else {
          const uint32_t chunk =
              1u << (L == 4 ? tune.pull_block_chunk_log2 : tune.pull_wave_chunk_log2);
          const uint32_t t0 = g.tin_off[v];
          uint32_t *dst_t = dst + 2 * stride;  // the entries' targets
          for (uint32_t k = 0; k < c[L]; k++) {
            dst[k] = t0 + k * chunk;
            dst_t[k] = v;
          }
        }
      }
    }
    if (f) vflag[v] = 0u;
    __syncthreads();
    if (tid < kVLists)
      s_run[tid] += s_wcnt[tid][0] + s_wcnt[tid][1] + s_wcnt[tid][2] + s_wcnt[tid][3];
    __syncthreads();
  }
  if (b == ntiles - 1 && tid < kVLists) {
    const uint32_t last = tid * ntiles + ntiles - 1;
    ctrl->len[2 + tid] = tile_pos[last] + tile_cnt[last] - tile_pos[tid * ntiles];
  }
}

// Per-kernel statistics of the last counting run, outside the timed path:
// every edge's kernel by the same rule binning applies (bucket byte, source
// list fit, source length for wave or workgroup kernels), with its oriented
// edge count and d+(u) + d+(v) sum.  Slots follow the launch order
// (merge-lane, merge-wave, wave-hash, workgroup-hash, wave-bitmap,
// workgroup-bitmap, wave-pull, workgroup-pull; pull edges by the target
// list's length); slot 8 the edges a partial table leaves out, slot 9 the
// edges that cannot close a triangle.
//   kern_bytes is what each kernel's strategy must read, its algorithmic
// bytes (the roofline numerator), in two parts.  Per counted edge (u -> v),
// with da = |N+(u) after v| and db = d+(v), added here:
//   merge (lane or wave)  24 + 4 (da + db): the list entry, col[e], src[e],
//                         off[u + 1], off[v], off[v + 1], then both lists
//                         (the merge walks the suffix of N+(u) and N+(v));
//   hash, push bitmap     4 db: the streamed target list N+(v);
//   pull                  4 da: the streamed suffix of N+(u) after v;
// plus 4 for the graph_of read of a merge edge in a batch.  Per item of the
// vertex lists, added by k_list_stats: the staged list and the metadata of
// every edge the item visits.
__global__ __launch_bounds__(256) void k_edge_stats(DevCsr g, Ctrl *ctrl,
                                                    const DispatchParams *P, uint32_t batch) {
  __shared__ uint8_t table[kBuckets];
  __shared__ TuneParams tune;
  __shared__ unsigned long long s_e[10], s_w[10], s_b[10];
  const int tid = threadIdx.x;
  table[tid] = ctrl->table[tid];
  if (tid == 0) tune = P->tune;
  if (tid < 10) s_e[tid] = s_w[tid] = s_b[tid] = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + tid; e < g.m; e += stride) {
    const uint32_t bk = g.ebk[e];
    uint32_t slot = 9, words = 0, bytes = 0;
    if (bk != kNoEdgeBucket) {
      const uint32_t a = g.src[e], b = g.col[e] >> 1;
      const uint32_t la = g.off[a], ea = g.off[a + 1];
      const uint32_t db = g.off[b + 1] - g.off[b];
      const uint32_t da = ea - (uint32_t)e - 1u;
      const uint8_t s = bucket_strategy(
          table, tune, ea - la, [&] { return span_of(g.col, la, ea); },
          [&] { return span_of(g.col, g.off[b], g.off[b + 1]); }, (int)bk);
      const bool small = ea - la <= kWaveMaxDeg, small_b = db <= kWaveMaxDeg;
      slot = s == S_SKIP         ? 8u  // left out by a partial strategy table
             : s == S_MERGE_LANE ? 0u
             : s == S_MERGE_WAVE ? 1u
             : s == S_HASH       ? (small ? 2u : 3u)
             : s == S_BITMAP     ? (small ? 4u : 5u)
                                 : (small_b ? 6u : 7u);
      words = (ea - la) + db;
      bytes = s == S_SKIP                              ? 0u
              : (s == S_MERGE_LANE || s == S_MERGE_WAVE) ? 24u + 4u * (da + db) + (batch ? 4u : 0u)
              : s == S_PULL                            ? 4u * da
                                                       : 4u * db;
    }
    atomicAdd(&s_e[slot], 1ull);
    atomicAdd(&s_w[slot], (unsigned long long)words);
    if (bytes) atomicAdd(&s_b[slot], (unsigned long long)bytes);
  }
  __syncthreads();
  if (tid < 10) {
    if (s_e[tid]) atomicAdd(&ctrl->kern_edges[tid], s_e[tid]);
    if (s_w[tid]) atomicAdd(&ctrl->kern_words[tid], s_w[tid]);
    if (s_b[tid]) atomicAdd(&ctrl->kern_bytes[tid], s_b[tid]);
  }
}

// Algorithmic bytes of the vertex-list items (the second part of
// kern_bytes), one thread per list entry:
//   push source a (hash, bitmap; workgroup or wave)  4 + 8 + 8 + 4 d + 12 d:
//       the list entry, off[a], off[a + 1], the list's two ends (its rank
//       span), the staged list N+(a), and per out-edge j its column word and
//       the target's two offsets (every out-edge is visited for the strategy
//       test);
//   pull chunk of target v with c in-edges  8 + 12 + 8 + 4 d + 8 c: the entry
//       and its target, off[v], off[v + 1], tin_off[v + 1], the list's two
//       ends, the staged list N+(v), and one 8-byte in-edge record per
//       in-edge of the chunk;
// plus 4 for the graph_of read of each item in a batch.  List L's entries sit
// at vlist + L vstride (targets of the pull lists 2 strides after them).
__global__ __launch_bounds__(256) void k_list_stats(DevCsr g, Ctrl *ctrl,
                                                    const DispatchParams *P,
                                                    const uint32_t *__restrict__ vlist,
                                                    uint64_t vstride, uint32_t batch) {
  __shared__ unsigned long long s_b[kVLists];
  const int tid = threadIdx.x;
  if (tid < kVLists) s_b[tid] = 0;
  __syncthreads();
  const uint32_t extra = batch ? 4u : 0u;
  for (int L = 0; L < kVLists; L++) {
    const uint32_t len = ctrl->len[2 + L];
    unsigned long long acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + tid; i < len; i += gridDim.x * blockDim.x) {
      if (L < 4) {
        const uint32_t a = vlist[(uint64_t)L * vstride + i];
        const uint32_t d = g.off[a + 1] - g.off[a];
        acc += 20u + 16ull * d + extra;
      } else {
        const uint32_t j0 = vlist[(uint64_t)L * vstride + i];
        const uint32_t v = vlist[(uint64_t)(L + 2) * vstride + i];
        const uint32_t chunk =
            1u << (L == 4 ? P->tune.pull_block_chunk_log2 : P->tune.pull_wave_chunk_log2);
        const uint32_t c = min(j0 + chunk, g.tin_off[v + 1]) - j0;
        const uint32_t d = g.off[v + 1] - g.off[v];
        acc += 28u + 4ull * d + 8ull * c + extra;
      }
    }
    if (acc) atomicAdd(&s_b[L], acc);
  }
  __syncthreads();
  // list -> kernel slot: workgroup hash 3, workgroup bitmap 5, wave hash 2,
  // wave bitmap 4, workgroup pull 7, wave pull 6
  if (tid < kVLists && s_b[tid]) {
    const int slot[kVLists] = {3, 5, 2, 4, 7, 6};
    atomicAdd(&ctrl->kern_bytes[slot[tid]], s_b[tid]);
  }
}

// ---- class accumulation -----------------------------------------------------
// neg = number of negative edges of a found triangle (0..3): class T(3-neg).
// In two-class mode c[0] counts balanced (even neg) and c[1] unbalanced.
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
  // Wave-uniform totals of one edge (u,v) with sign bit sab: e[k] triangles
  // whose other two edges carry k negatives (two-class mode: e[0] even,
  // e[1] odd).
  __device__ __forceinline__ void add_edge(uint32_t sab, uint32_t e0, uint32_t e1, uint32_t e2) {
    if (CLS) {
      // negatives = k + sab -> class T(3 - k - sab)
      if (sab) {
        c[2] += e0;
        c[1] += e1;
        c[0] += e2;
      } else {
        c[3] += e0;
        c[2] += e1;
        c[1] += e2;
      }
    } else {
      c[0] += sab ? e1 : e0;
      c[1] += sab ? e0 : e1;
    }
  }
  __device__ __forceinline__ void flush(Ctrl *ctrl) {
#pragma unroll
    for (int k = 0; k < 4; k++) wave_add(&ctrl->t[blockIdx.x % kTSlots][k], c[k]);
  }
  // Batched counting: add this wave's totals to graph gid's (balanced,
  // unbalanced) pair and restart from zero.
  __device__ __forceinline__ void flush_graph(unsigned long long *per_graph, uint32_t gid) {
    if (CLS) {
      wave_add(&per_graph[2 * gid], c[3] + c[1]);
      wave_add(&per_graph[2 * gid + 1], c[2] + c[0]);
    } else {
      wave_add(&per_graph[2 * gid], c[0]);
      wave_add(&per_graph[2 * gid + 1], c[1]);
    }
    c[0] = c[1] = c[2] = c[3] = 0;
  }
};

// Batched counting: a wave's totals go to its graph's counters only when the
// wave moves on to another graph (and once at the end), not after every item.
// Consecutive items of a queue are neighbouring ranks, mostly of one graph,
// and every wave of an XCD works near the same ranks, so a flush per item put
// two device-scope atomics per item on the same two counters.  to(gid) is
// called before an item of graph gid.  The four-class variant serves one
// graph only and adds its totals to the count's counters at the end.
template <bool CLS>
struct GraphAcc {
  Acc<CLS> acc;
  uint32_t cur = 0xFFFFFFFFu;
  __device__ __forceinline__ void to(unsigned long long *per_graph, uint32_t gid) {
    if constexpr (!CLS) {
      if (gid == cur) return;
      if (cur != 0xFFFFFFFFu) acc.flush_graph(per_graph, cur);
      cur = gid;
    }
  }
  // four-class counting is one graph's: its four counters sit in ctrl->t
  __device__ __forceinline__ void finish(unsigned long long *per_graph, Ctrl *ctrl) {
    if constexpr (CLS) {
      acc.flush(ctrl);
    } else if (cur != 0xFFFFFFFFu) {
      acc.flush_graph(per_graph, cur);
    }
  }
};

// Per-vertex counts (four-class kernels only): credit source vertex a (rank)
// with the classes its triangles added to acc since `before` was taken.  acc
// is wave-uniform; lane 0 adds.
template <bool CLS>
__device__ __forceinline__ void pv_credit(const DevCsr &g, uint32_t a, const Acc<CLS> &acc,
                                          const uint64_t (&before)[4]) {
  if constexpr (CLS) {
    if (g.pv && __lane_id() == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint64_t d = acc.c[k] - before[k];
        if (d) atomicAdd(&g.pv[4ull * a + k], (unsigned long long)d);
      }
    }
  }
}

// Per-lane probe result folded into wave-uniform edge totals.  A probe
// reports a streamed entry x (sign nb of its edge) as found in the staged
// list with the same sign (same: the two edges to x carry 0 or 2 negatives,
// 2 nb of them) or the opposite sign (diff: exactly 1 negative).
template <bool CLS>
__device__ __forceinline__ void tally(uint32_t same, uint32_t diff, uint32_t nb, uint32_t &e0,
                                      uint32_t &e1, uint32_t &e2) {
  if (CLS) {
    e0 += (uint32_t)__popcll(__ballot(same != 0u && nb == 0u));
    e1 += (uint32_t)__popcll(__ballot(diff != 0u));
    e2 += (uint32_t)__popcll(__ballot(same != 0u && nb != 0u));
  } else {
    e0 += (uint32_t)__popcll(__ballot(same != 0u));  // even
    e1 += (uint32_t)__popcll(__ballot(diff != 0u));  // odd
  }
}

// same / diff (0 or 1) of a found staged word x ((rank << 1) | neg) against
// a streamed entry with sign nb.
__device__ __forceinline__ void found_as(bool found, uint32_t x, uint32_t nb, uint32_t &same,
                                         uint32_t &diff) {
  same = found && ((x ^ nb) & 1u) == 0u ? 1u : 0u;
  diff = found && ((x ^ nb) & 1u) != 0u ? 1u : 0u;
}

// ---- merge: one lane per edge ----------------------------------------------
constexpr uint32_t kLaneRegs = 8;  // lists up to this long are merged in registers
// Two-pointer merge of the part of N+(u) after v with N+(v); per-lane class
// counters (two-class mode: c[0] balanced, c[1] unbalanced). REGS adds the
// in-register path for two short lists (kept out of the persistent kernel,
// whose other paths it would crowd).
template <bool CLS, bool REGS = true>
__device__ __forceinline__ void merge_lane_edge(const DevCsr &g, uint32_t e, uint32_t c[4]) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t w = col[e];
  const uint32_t a = g.src[e], b = w >> 1, sab = w & 1u;
  uint32_t pa = e + 1;
  const uint32_t ea = g.off[a + 1];
  uint32_t pb = g.off[b];
  const uint32_t eb = g.off[b + 1];
  if (REGS && ea - pa <= kLaneRegs && eb - pb <= kLaneRegs) {
    // both lists short: load them whole (independent loads, no chain of one
    // load per merge step) and compare every pair; ranks are unique in a
    // list, and the two pads differ from each other and from every rank
    uint32_t A[kLaneRegs], Bv[kLaneRegs];
#pragma unroll
    for (uint32_t i = 0; i < kLaneRegs; i++) {
      A[i] = pa + i < ea ? col[pa + i] : 0xFFFFFFFFu;
      Bv[i] = pb + i < eb ? col[pb + i] : 0xFFFFFFFDu;
    }
#pragma unroll
    for (uint32_t i = 0; i < kLaneRegs; i++)
#pragma unroll
      for (uint32_t j = 0; j < kLaneRegs; j++) {
        if ((A[i] >> 1) == (Bv[j] >> 1)) {
          const uint32_t neg = sab + (A[i] & 1u) + (Bv[j] & 1u);
          if (CLS) {
            c[0] += neg == 3;
            c[1] += neg == 2;
            c[2] += neg == 1;
            c[3] += neg == 0;
          } else {
            c[0] += (neg & 1u) ^ 1u;
            c[1] += neg & 1u;
          }
        }
      }
    return;
  }
  uint32_t x = col[pa], y = col[pb];
  while (true) {
    const uint32_t xr = x >> 1, yr = y >> 1;
    if (xr == yr) {
      const uint32_t neg = sab + (x & 1u) + (y & 1u);
      if (CLS) {
        c[0] += neg == 3;
        c[1] += neg == 2;
        c[2] += neg == 1;
        c[3] += neg == 0;
      } else {
        c[0] += (neg & 1u) ^ 1u;
        c[1] += neg & 1u;
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

template <bool CLS>
__global__ __launch_bounds__(256) void k_merge_lane(DevCsr g, Ctrl *ctrl,
                                                    const uint32_t *__restrict__ list) {
  const uint32_t len = ctrl->len[0];
  uint32_t c[4] = {0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    const uint32_t e = list[i];
    if constexpr (CLS) {
      const uint32_t before[4] = {c[0], c[1], c[2], c[3]};
      merge_lane_edge<CLS>(g, e, c);
      if (g.pv) {  // per-vertex counts: credit the edge's source
        const uint32_t a = g.src[e];
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (c[k] != before[k])
            atomicAdd(&g.pv[4ull * a + k], (unsigned long long)(c[k] - before[k]));
      }
    } else {
      merge_lane_edge<CLS>(g, e, c);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint64_t v = c[k];
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    wave_add(&ctrl->t[blockIdx.x % kTSlots][k], v);
  }
}

// ---- merge: one wave64 per edge, 64x64 tiles -----------------------------
// A = the part of N+(u) after v, B = N+(v).  Each lane holds one element of the
// current A tile and one of the current B tile; a 6-step ds_bpermute binary
// search finds each A element in the B tile; the tile whose last element is
// smaller advances (both on a tie).  The tile after each current tile is
// loaded one step ahead so its latency overlaps the search.  e is
// wave-uniform.
template <bool CLS>
__device__ __forceinline__ void merge_wave_edge(const DevCsr &g, uint32_t e, Acc<CLS> &acc) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t lane = __lane_id();
  const uint32_t w = col[e];
  const uint32_t a = g.src[e], b = w >> 1, sab = w & 1u;
  uint32_t ia = e + 1;
  const uint32_t ea = g.off[a + 1];
  uint32_t ib = g.off[b];
  const uint32_t eb = g.off[b + 1];
  uint32_t nA = min(64u, ea - ia), nB = min(64u, eb - ib);
  uint32_t xa = lane < nA ? col[ia + lane] : kEmpty;
  uint32_t yb = lane < nB ? col[ib + lane] : kEmpty;
  uint32_t ia2 = ia + nA, ib2 = ib + nB;
  uint32_t xa2 = ia2 + lane < ea ? col[ia2 + lane] : kEmpty;
  uint32_t yb2 = ib2 + lane < eb ? col[ib2 + lane] : kEmpty;
  while (true) {
    // lower bound of xa among the 64 B entries (padding sorts last)
    const uint32_t key = xa >> 1;
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t step = 32; step > 0; step >>= 1) {
      uint32_t cand = (uint32_t)__shfl((int)yb, (int)(pos + step - 1), 64);
      if ((cand >> 1) < key) pos += step;
    }
    const uint32_t hit = (uint32_t)__shfl((int)yb, (int)min(pos, 63u), 64);
    const bool found = lane < nA && pos < nB && (hit >> 1) == key;
    acc.add(found, sab + (xa & 1u) + (hit & 1u));
    // tile ends are wave-uniform: keep them (and the advance) in SGPRs so
    // the loop runs as uniform control flow with one exit test
    const uint32_t lastA =
        __builtin_amdgcn_readfirstlane((uint32_t)__shfl((int)xa, (int)(nA - 1), 64) >> 1);
    const uint32_t lastB =
        __builtin_amdgcn_readfirstlane((uint32_t)__shfl((int)yb, (int)(nB - 1), 64) >> 1);
    if (lastA <= lastB) {
      ia = ia2;
      nA = min(64u, ea - ia);
      xa = xa2;
      ia2 = ia + nA;
      xa2 = ia2 + lane < ea ? col[ia2 + lane] : kEmpty;
    }
    if (lastB <= lastA) {
      ib = ib2;
      nB = min(64u, eb - ib);
      yb = yb2;
      ib2 = ib + nB;
      yb2 = ib2 + lane < eb ? col[ib2 + lane] : kEmpty;
    }
    if (ia >= ea || ib >= eb) break;
  }
}

template <bool CLS>
__global__ __launch_bounds__(256) void k_merge_wave(DevCsr g, Ctrl *ctrl,
                                                    const uint32_t *__restrict__ list) {
  const uint32_t len = ctrl->len[1];
  const uint32_t nw = gridDim.x * (blockDim.x >> 6);
  Acc<CLS> acc;
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < len; i += nw) {
    const uint32_t e = __builtin_amdgcn_readfirstlane(list[i]);
    const uint64_t before[4] = {acc.c[0], acc.c[1], acc.c[2], acc.c[3]};
    merge_wave_edge<CLS>(g, e, acc);
    if constexpr (CLS) {
      if (g.pv) pv_credit<CLS>(g, __builtin_amdgcn_readfirstlane(g.src[e]), acc, before);
    }
  }
  acc.flush(ctrl);
}

// ---- vertex-grouped kernels -------------------------------------------------
// Wave-scope LDS ordering for the wave-per-vertex variant (no workgroup
// barrier is needed when one wave owns an LDS region).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kUnroll = 8;  // 64-element loads per unit (a wave keeps two units in flight)
// Word 3 of a raw (stride 0) buffer descriptor on gfx9-family GPUs (gfx950):
// 32-bit data format, range checked against the byte count in word 2.
constexpr int kBufferWord3 = 0x00020000;

// Look up rank yr among the nlow sorted words L[0, nlow) ((rank << 1) | neg)
// by a branch-free lower bound: the matching word, or kEmpty.
__device__ __forceinline__ uint32_t low_lookup(const uint32_t *__restrict__ L, uint32_t nlow,
                                               uint32_t yr) {
  uint32_t pos = 0;  // entries below yr seen so far
  for (uint32_t step = nlow ? 1u << (31 - __clz(nlow)) : 0u; step; step >>= 1)
    if (pos + step <= nlow && (L[pos + step - 1] >> 1) < yr) pos += step;
  if (pos < nlow) {
    const uint32_t x = L[pos];
    if ((x >> 1) == yr) return x;
  }
  return kEmpty;
}

// ---- probes of a staged list ----------------------------------------------------
// probe(y, s, same, diff): whether the streamed entry y = (rank << 1) | nb is
// in the staged list with the sign nb ^ s (same = 1) or the other sign (diff
// = 1).  With s = 0 that is the entry's own sign.  Two-class counting passes
// the sign of the streamed list's own edge as s: a hit then has an even
// number of negative edges in its triangle exactly when same = 1.  y = 0 and
// y = kEmpty always miss.  0 is no streamed entry (every streamed rank lies
// above the vertex whose list holds it), and the streaming loops leave the
// lanes of a buffer load past the list end, which read 0, unmasked.

// LDS bitmap B over the rank window [lo, lo + span) of the staged list, two
// bits per rank, one-hot by sign: rank lo + r with sign s sets bit 2r + s.
// The sorted array L[0, nlow) holds the entries below lo.  With d2 = y - 2 lo
// = 2 (rank - lo) + nb, the word is d2 >> 5, bit d2 & 31 is the same-sign bit
// and bit (d2 ^ 1) & 31 the opposite-sign bit (one bit-field extract each).
// The word index is clamped to zw, the word after the window's last: no bit
// at or past 2 span is ever set (the window ends at the list's highest rank,
// and staging sets no other bits), and word zw is zero too (a word of the
// bitmap past the list, or the guard after a full window), so padding
// (kEmpty) and ranks above or below the window miss without a window test.
// A 64-element group with a
// target below lo (a prefix of the sorted N+(v), so rare) also searches L for
// those lanes.  A target below lo wraps d2 to at least 2^32 - 2 lo > 2^21 >=
// 2 span, and kEmpty - 1 - 2 lo exceeds 2 span too, because the API admits at most
// 2^31 - 2^20 vertices (ranks).  LOW = false is the probe of a list with no
// entries below the window (nlow = 0), which skips the low test entirely.
// No branch sits around the read, so the reads of an unrolled group go out
// together instead of each waiting for the one before.
template <bool LOW>
struct BitmapProbe {
  const uint32_t *B;
  uint32_t lo2, zw;
  const uint32_t *L;
  uint32_t nlow;
  __device__ __forceinline__ void operator()(uint32_t y, uint32_t s, uint32_t &same,
                                             uint32_t &diff) const {
    const uint32_t d2 = y - lo2;
    const uint32_t w = B[min(d2 >> 5, zw)];
    // v_bfe_u32 reads bits [4:0] of its offset operand: d2 & 31 needs no mask;
    // s flips bit 0, the sign half of the rank's two bits
    const uint32_t ds = d2 ^ s;
    same = __builtin_amdgcn_ubfe(w, ds, 1u);
    diff = __builtin_amdgcn_ubfe(w, ds ^ 1u, 1u);
    if constexpr (LOW) {
      const bool below = y - 1u < lo2 - 1u;  // 0 < y < lo2: 0 is no entry
      if (__ballot(below)) {                  // wave-uniform
        if (below) {
          const uint32_t x = low_lookup(L, nlow, y >> 1);
          found_as(x != kEmpty, x, y ^ s, same, diff);
        }
      }
    }
  }
};

// LDS hash table T of N+(u): slot = rank & mask, linear probing, kEmpty free.
struct HashProbe {
  const uint32_t *T;
  uint32_t mask;
  __device__ __forceinline__ void operator()(uint32_t y, uint32_t s, uint32_t &same,
                                             uint32_t &diff) const {
    uint32_t hit = kEmpty;
    if (y != kEmpty && y != 0u) {
      const uint32_t yr = y >> 1;
      uint32_t h = yr & mask;
      while (true) {
        const uint32_t x = T[h];
        if (x == kEmpty) break;
        if ((x >> 1) == yr) {
          hit = x;
          break;
        }
        h = (h + 1) & mask;
      }
    }
    found_as(hit != kEmpty, hit, y ^ s, same, diff);
  }
};

// Stream one list per edge index j = first, first + step, ... (below ea)
// that `keep` accepts against `probe`, adding each edge's triangles to acc.
// meta(j, w, pb, eb, x) gives edge j's column word w (bit 0: the sign of the
// edge between the staged vertex and the list's owner), its list col[pb, eb)
// and the vertex x that owns the edge's triangles when EDGE_OWNER (pull:
// each triangle's lowest-ranked vertex, credited per edge for per-vertex
// counts).  Push streams N+(v) for the staged source's out-edges; pull
// streams the part of N+(u) after v for the staged target's in-edges.
//   Edge metadata moves 64 edges at a time, one edge per lane: the column
// word, the list bounds and the keep flag are loaded and evaluated in
// parallel, and the kept edges are visited through the ballot with
// readlane, so no memory latency sits between consecutive edges.
//   List entries move in units of up to 64 * kUnroll entries of one list.  The
// next unit's loads (the rest of the list, or the head of the next kept
// list) are issued before the current unit is probed, so a wave keeps two
// units of loads in flight across edge boundaries.  A unit is read with
// buffer loads through a descriptor whose base is the unit's first entry and
// whose range ends at min(list end, unit end): lanes past the list end read 0
// without touching memory and are masked to kEmpty when the unit is probed.
// The loads are unconditional, their per-lane offsets never change and only
// the descriptor (SGPRs) moves, so each wait counts exactly the loads it
// needs and no address register is rewritten under a load in flight.
template <bool CLS, bool EDGE_OWNER, class Meta, class Keep, class Probe>
__device__ __forceinline__ void stream_edges(const DevCsr &g, uint32_t first, uint32_t ea,
                                             uint32_t step, const Meta &meta, const Keep &keep,
                                             const Probe &probe, Acc<CLS> &acc) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t lane = __lane_id();
  constexpr uint32_t kUnit = 64u * kUnroll;
  auto load = [&](uint32_t(&y)[kUnroll], uint32_t pos, uint32_t eb) {
    const uint32_t nrec = pos < eb ? min(eb - pos, kUnit) * 4u : 0u;
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void *)(col + pos), (short)0, (int)nrec, kBufferWord3);
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
      y[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)((u * 64u + lane) * 4u), 0, 0);
  };
  uint32_t e0 = 0, e1 = 0, e2 = 0;
  // a unit: column word of its edge (bit 0 = sign), first entry, list end,
  // the vertex owning the edge's triangles (EDGE_OWNER)
  struct Unit {
    uint32_t w, pos, eb, x;
  };
  auto lane_of = [](uint32_t v, int i) { return (uint32_t)__builtin_amdgcn_readlane((int)v, i); };
  auto probe_unit = [&](const uint32_t(&y)[kUnroll], const Unit &c) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      if (c.pos + u * 64u >= c.eb) break;  // wave-uniform
      uint32_t same, diff;
      probe(y[u], 0u, same, diff);  // lanes past the list end read 0, which misses
      tally<CLS>(same, diff, y[u] & 1u, e0, e1, e2);
    }
    // Retire the buffer's loads here (a no-op use), so a group skipped past
    // the list end leaves nothing pending into the loop's next turn.
#pragma unroll
    for (int u = 0; u < kUnroll; u++) asm volatile("" ::"v"(y[u]));
    if (c.pos + kUnit >= c.eb) {  // last unit of its list
      if constexpr (CLS && EDGE_OWNER) {
        const uint64_t before[4] = {acc.c[0], acc.c[1], acc.c[2], acc.c[3]};
        acc.add_edge(c.w & 1u, e0, e1, e2);
        pv_credit<CLS>(g, c.x, acc, before);
      } else {
        acc.add_edge(c.w & 1u, e0, e1, e2);
      }
      e0 = e1 = e2 = 0;
    }
  };
  for (uint32_t j0 = first; j0 < ea; j0 += 64u * step) {
    const uint32_t j = j0 + lane * step;
    uint32_t w = 0, pb = 0, eb = 0, x = 0;
    if (j < ea) meta(j, w, pb, eb, x);
    uint64_t mask = __ballot(j < ea && keep(j, pb, eb));
    // next unit after c: the rest of its list, else the next kept edge's list;
    // when none is left, n repeats c past its end (its loads stay in bounds)
    auto advance = [&](const Unit &c, Unit &n) -> bool {
      if (c.pos + kUnit < c.eb) {
        n = Unit{c.w, c.pos + kUnit, c.eb, c.x};
        return true;
      }
      if (mask) {
        const int i = __builtin_ctzll(mask);
        mask &= mask - 1;
        n = Unit{lane_of(w, i), lane_of(pb, i), lane_of(eb, i),
                 EDGE_OWNER ? lane_of(x, i) : 0u};
        return true;
      }
      n = Unit{c.w, c.pos + kUnit, c.eb, c.x};
      return false;
    };
    if (!mask) continue;
    const int i0 = __builtin_ctzll(mask);
    mask &= mask - 1;
    Unit a{lane_of(w, i0), lane_of(pb, i0), lane_of(eb, i0), EDGE_OWNER ? lane_of(x, i0) : 0u};
    uint32_t ya[kUnroll], yb[kUnroll];
    load(ya, a.pos, a.eb);
    // two buffers in turn: the other buffer's loads are in flight while one
    // is probed
    while (true) {
      Unit b;
      const bool vb = advance(a, b);
      load(yb, b.pos, b.eb);
      probe_unit(ya, a);
      if (!vb) break;
      const bool va = advance(b, a);
      load(ya, a.pos, a.eb);
      probe_unit(yb, b);
      if (!va) break;
    }
    // the last turn's loads past the end: retire them before the next batch
#pragma unroll
    for (int u = 0; u < kUnroll; u++) asm volatile("" ::"v"(ya[u]), "v"(yb[u]));
  }
}

// Row-packed streaming: the kept lists of a 64-edge metadata batch are cut
// into rows of 64 consecutive entries (list i holds ceil(len_i / 64) rows,
// the rows of the batch numbered by an inclusive scan), and a unit is the
// next kRows rows whatever lists they come from.  Each row is one buffer load
// through its own descriptor, built when the load is issued, whose range ends
// at its list's end, so only the last row of a list is part-filled: short
// lists (a workgroup-pull suffix is about 108 entries) no longer leave most of
// a unit's loads empty, and a wave keeps two full units of loads in flight.
// The lanes locate 64 rows at a time: row r belongs to the first lane whose
// inclusive row count exceeds r (a binary search over the scan with lane
// shuffles); each row's triangles are added at once with its edge's sign, so
// no per-list state crosses rows.  Two-class counting folds the row's edge
// sign into the probe and keeps per-lane hit counts (even, odd), summed over
// the wave once per call: the per-row work is then vector instructions only,
// where ballot, popcount and a sign-selected 64-bit add per row kept the
// scalar unit busier than the vector unit.  A lane's count in one call is at
// most the rows it probes, at most 2^26 + 2^13 (pull streams one chunk of at
// most 2^13 in-edges of one target, whose suffixes are disjoint parts of the
// column array of fewer than 2^32 entries).
constexpr uint32_t kRows = kUnroll;
static_assert(64u % kRows == 0u, "a unit of rows never straddles two groups of 64 rows");

template <bool CLS, class Meta, class Keep, class Probe>
__device__ __forceinline__ void stream_rows(const DevCsr &g, uint32_t first, uint32_t ea,
                                            uint32_t step, const Meta &meta, const Keep &keep,
                                            const Probe &probe, Acc<CLS> &acc) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t lane = __lane_id();
  auto lane_of = [](uint32_t v, int i) { return (uint32_t)__builtin_amdgcn_readlane((int)v, i); };
  auto shfl = [](uint32_t v, uint32_t i) { return (uint32_t)__shfl((int)v, (int)i, 64); };
  uint32_t hit_even = 0, hit_odd = 0;  // two classes: this lane's hits
  for (uint32_t j0 = first; j0 < ea; j0 += 64u * step) {
    const uint32_t j = j0 + lane * step;
    uint32_t w = 0, pb = 0, eb = 0, x = 0;
    if (j < ea) meta(j, w, pb, eb, x);
    const bool kept = j < ea && keep(j, pb, eb);
    const uint32_t rows = kept ? (eb - pb + 63u) >> 6 : 0u;
    const uint32_t incl = wave_scan_incl(rows);
    const uint32_t total = lane_of(incl, 63);
    if (total == 0u) continue;
    // Rows are located 64 at a time, one per lane (a unit's rows then come
    // from two readlanes each): row G + lane's first entry and byte count (0
    // past the last row, which then reads nothing), and the rows' edge
    // signs as a bit mask.
    uint32_t posg = 0, nrecg = 0;
    uint64_t sgg = 0;
    auto group = [&](uint32_t G) {
      const uint32_t r = G + lane;
      uint32_t o = 0;  // lanes before the owner: incl <= r
#pragma unroll
      for (uint32_t st = 32; st > 0; st >>= 1)
        if (shfl(incl, o + st - 1u) <= r) o += st;
      o = min(o, 63u);
      const uint32_t ro = shfl(incl, o) - shfl(rows, o);  // the owner's first row
      const uint32_t p = shfl(pb, o) + 64u * (r - ro);
      const uint32_t e = shfl(eb, o);
      // every lane takes part in the shuffles: a lane shuffle reads nothing
      // from a lane that is switched off at that point
      const uint32_t sg = shfl(w, o) & 1u;
      const bool live = r < total;
      posg = live ? p : 0u;
      nrecg = live ? min(e - p, 64u) * 4u : 0u;
      sgg = __ballot(live && sg);
    };
    auto load = [&](uint32_t(&y)[kRows], uint32_t R) -> uint32_t {
      if ((R & 63u) == 0u) group(R);  // wave-uniform
      const uint32_t l0 = R & 63u;
#pragma unroll
      for (uint32_t u = 0; u < kRows; u++) {
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(col + lane_of(posg, (int)(l0 + u))), (short)0,
            (int)lane_of(nrecg, (int)(l0 + u)), kBufferWord3);
        y[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(lane * 4u), 0, 0);
      }
      return (uint32_t)(sgg >> l0) & ((1u << kRows) - 1u);
    };
    auto probe_unit = [&](const uint32_t(&y)[kRows], uint32_t R, uint32_t sabm) {
#pragma unroll
      for (uint32_t u = 0; u < kRows; u++) {
        if (R + u >= total) break;  // wave-uniform
        uint32_t same, diff;
        // lanes past the row's end read 0, which misses
        if constexpr (CLS) {
          probe(y[u], 0u, same, diff);
          uint32_t e0 = 0, e1 = 0, e2 = 0;
          tally<CLS>(same, diff, y[u] & 1u, e0, e1, e2);
          acc.add_edge((sabm >> u) & 1u, e0, e1, e2);
        } else {
          probe(y[u], (sabm >> u) & 1u, same, diff);
          hit_even += same;
          hit_odd += diff;
        }
      }
      // retire the unit's loads here, so none is left pending past its use
#pragma unroll
      for (uint32_t u = 0; u < kRows; u++) asm volatile("" ::"v"(y[u]));
    };
    uint32_t ya[kRows], yb[kRows];
    uint32_t R = 0;
    uint32_t sa = load(ya, R);
    // two units in turn: one's loads are in flight while the other is probed
    while (true) {
      const uint32_t sb = load(yb, R + kRows);
      probe_unit(ya, R, sa);
      if (R + kRows >= total) {
#pragma unroll
        for (uint32_t u = 0; u < kRows; u++) asm volatile("" ::"v"(yb[u]));
        break;
      }
      R += kRows;
      sa = load(ya, R + kRows);
      probe_unit(yb, R, sb);
      if (R + kRows >= total) {
#pragma unroll
        for (uint32_t u = 0; u < kRows; u++) asm volatile("" ::"v"(ya[u]));
        break;
      }
      R += kRows;
    }
  }
  if constexpr (!CLS) {
    acc.c[0] += wave_sum(hit_even);
    acc.c[1] += wave_sum(hit_odd);
  }
}

// Inclusive max-scan over the wave.
__device__ __forceinline__ uint32_t wave_max_scan_incl(uint32_t x) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
    if (lane >= (uint32_t)d) x = max(x, y);
  }
  return x;
}

// Packed streaming, for short lists: the kept lists of a 64-edge metadata
// batch are laid end to end (exclusive scan P of their lengths) and read in
// units of kPackUnit consecutive entries, which may span many lists, so a
// unit's loads are full whatever the list lengths.  Entry f of the packed
// sequence belongs to the last list i with P_i <= f and lies at col[f + pb_i
// - P_i].  The lists are numbered in order (their compact index c, from
// mbcnt over the ballot of non-empty lists), and each list's record
// {pb_i - P_i, its edge's sign} goes to H[c].  Per unit, every list starting
// inside it sets bit f - B of a kPackUnit-bit start mask M (16 LDS words),
// so slot f of group u (64 slots) lies in the list numbered S_u + (number of
// start bits at or below f within the group) - 1, S_u being the lists
// started before the group: one mbcnt over the group's 64-bit mask word,
// with the counts of earlier groups carried in a scalar.  Each entry's
// triangle class follows from the sign bits of its three edges, so no
// per-list reduction is needed (callers that credit triangles per edge owner
// use stream_edges).  H holds kPackWords words: 64 records, then M.
constexpr uint32_t kPackUnit = 512;  // 64 lanes x 8 slots
static_assert(kPackUnit == 64u * kUnroll, "packed unit: 8 slots per lane");
constexpr uint32_t kPackWords = 2u * 64u + kPackUnit / 32u;

template <bool CLS, class Meta, class Keep, class Probe>
__device__ __forceinline__ void stream_packed(const DevCsr &g, uint32_t first, uint32_t ea,
                                              uint32_t step, const Meta &meta, const Keep &keep,
                                              const Probe &probe, uint64_t *H, Acc<CLS> &acc) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t lane = __lane_id();
  uint32_t *const M = (uint32_t *)(H + 64);
  for (uint32_t j0 = first; j0 < ea; j0 += 64u * step) {
    const uint32_t j = j0 + lane * step;
    uint32_t w = 0, pb = 0, eb = 0, x = 0;
    if (j < ea) meta(j, w, pb, eb, x);
    const bool kept = j < ea && keep(j, pb, eb);
    const uint32_t len = kept ? eb - pb : 0u;
    const uint32_t incl = wave_scan_incl(len);
    const uint32_t P = incl - len;
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    if (total == 0u) continue;
    const bool live = len != 0u;
    const uint64_t lm = __ballot(live);
    wave_lds_sync();  // the previous batch's reads of H and M are done
    if (live) {
      const uint32_t c = __builtin_amdgcn_mbcnt_hi((uint32_t)(lm >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)lm, 0u));
      H[c] = ((uint64_t)(w & 1u) << 32) | (uint32_t)(pb - P);  // f + (pb - P) mod 2^32
    }
    // the unit at B: start bits, then each slot's list record and its loads
    // (groups past the end load nothing); n[u] is the list's edge sign
    auto prepare = [&](uint32_t B, uint32_t(&y)[kUnroll], uint32_t(&n)[kUnroll]) {
      if (B) wave_lds_sync();  // the previous unit's reads of M are done
      if (lane < kPackUnit / 32u) M[lane] = 0u;
      wave_lds_sync();
      const uint32_t s = P - B;  // wraps for lists that started before B
      if (live && s < kPackUnit) atomicOr(&M[s >> 5], 1u << (s & 31u));
      wave_lds_sync();
      const uint32_t mw = lane < kPackUnit / 32u ? M[lane] : 0u;
      // lists started before the unit, minus one (wraps to ~0 at B = 0,
      // where slot 0 always starts a list)
      uint32_t S = (uint32_t)__popcll(__ballot(live && P < B)) - 1u;
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        if (B + u * 64u >= total) break;  // wave-uniform
        const uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)mw, 2 * u + 1) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)mw, 2 * u);
        // start bits at or below this lane: bit 0, plus the bits of m >> 1 below it
        const uint64_t m1 = m >> 1;
        const uint32_t c = S + (uint32_t)(m & 1u) +
                           __builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
        S += (uint32_t)__popcll(m);
        // c < 64 on every lane (at most 64 lists start at or below any slot)
        const uint64_t rec = H[c & 63u];
        const uint32_t f = B + u * 64u + lane;
        y[u] = f < total ? col[f + (uint32_t)rec] : kEmpty;
        n[u] = (uint32_t)(rec >> 32);
      }
    };
    auto consume = [&](uint32_t B, const uint32_t(&y)[kUnroll], const uint32_t(&n)[kUnroll]) {
      uint32_t even = 0, odd = 0;
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        if (B + u * 64u >= total) break;  // wave-uniform
        uint32_t same, diff;
        if constexpr (CLS) {
          probe(y[u], 0u, same, diff);
          // negatives among the triangle's three edges: the list's edge, plus
          // 1 (diff) or 2 nb (same) for the two edges to the found vertex
          const uint32_t neg = n[u] + (diff ? 1u : 2u * (y[u] & 1u));
          acc.add((same | diff) != 0u, neg);
        } else {
          // two classes: the list edge's sign folded into the probe, so that
          // "same" reports the hits whose triangle has an even number of
          // negative edges
          probe(y[u], n[u], same, diff);
          even += (uint32_t)__popcll(__ballot(same != 0u));
          odd += (uint32_t)__popcll(__ballot(diff != 0u));
        }
      }
      if constexpr (!CLS) {
        acc.c[0] += even;
        acc.c[1] += odd;
      }
    };
    // One unit at a time: a unit already holds up to kPackUnit entries of many
    // lists, and a second buffer in flight measured no faster while it cost
    // the registers that let the batched kernel run packed pull items.
    uint32_t y[kUnroll], n[kUnroll];
    for (uint32_t B = 0; B < total; B += kPackUnit) {
      prepare(B, y, n);
      consume(B, y, n);
    }
  }
}

// Edge metadata of push streaming: out-edge j of the staged source; its list
// is the target's out-list N+(v).
struct PushMeta {
  const DevCsr &g;
  __device__ __forceinline__ void operator()(uint32_t j, uint32_t &w, uint32_t &pb, uint32_t &eb,
                                             uint32_t &x) const {
    w = g.col[j];
    pb = g.off[w >> 1];
    eb = g.off[(w >> 1) + 1];
    x = 0;
  }
};

// Edge metadata of pull streaming: in-edge j of the staged target v is the
// record of edge e = (u -> v): its list is the part of N+(u) after v,
// col[e + 1, e + 1 + len), with the edge's sign, in one 8-byte load.  u owns
// the edge's triangles; it is read (src[e]) only for per-vertex counts.
struct PullMeta {
  const DevCsr &g;
  __device__ __forceinline__ void operator()(uint32_t j, uint32_t &w, uint32_t &pb, uint32_t &eb,
                                             uint32_t &x) const {
    const InEdge r = g.tin[j];
    w = r.w & 1u;
    pb = r.pos;
    eb = r.pos + in_len(r.w);
    x = g.pv ? g.src[r.pos - 1u] : 0u;
  }
};

// LDS hash table of a source list: `cap` = next_pow2(4 d) slots (load
// <= 1/4), slot = rank & (cap - 1), linear probing.  Out-lists are runs of
// nearby ranks, so the low rank bits spread them with few collisions.
__device__ __forceinline__ int table_log2(uint32_t d, int min_lg) {
  int lg = clog2(4u * d);
  return lg < min_lg ? min_lg : lg;
}

// Stage N+(a) in an LDS hash table and count a's hash edges; called by every
// thread of the workgroup (T holds at least 4 * d slots).
template <bool CLS>
__device__ __forceinline__ void hash_vertex_block(const DevCsr &g, uint32_t a, uint32_t *T,
                                                  const uint8_t *table, const TuneParams &tune,
                                                  Acc<CLS> &acc) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwaves = nthreads >> 6;
  const uint32_t la = __builtin_amdgcn_readfirstlane(g.off[a]);
  const uint32_t ea = __builtin_amdgcn_readfirstlane(g.off[a + 1]), d = ea - la;
  [[maybe_unused]] const uint64_t before[4] = {acc.c[0], acc.c[1], acc.c[2], acc.c[3]};
  const uint32_t span = span_of(col, la, ea);
  const int lg = table_log2(d, 6);
  const uint32_t cap = 1u << lg, mask = cap - 1u;
  for (uint32_t k = tid; k < cap; k += nthreads) T[k] = kEmpty;
  __syncthreads();
  for (uint32_t k = tid; k < d; k += nthreads) {
    uint32_t x = col[la + k];
    uint32_t h = (x >> 1) & mask;
    while (atomicCAS(&T[h], kEmpty, x) != kEmpty) h = (h + 1) & mask;
  }
  __syncthreads();
  stream_edges<CLS, false>(
      g, la + wave, ea, nwaves, PushMeta{g},
      [&](uint32_t j, uint32_t pb, uint32_t eb) {
        return edge_strategy(table, tune, d, span, ea - j - 1, g.col, pb, eb) == S_HASH;
      },
      HashProbe{T, mask}, acc);
  pv_credit<CLS>(g, a, acc, before);
  __syncthreads();
}

// One wave stages N+(a) (at most kWaveMaxDeg entries) in its own table T of
// kWaveCap slots and counts a's hash edges.
constexpr uint32_t kWaveCap = 4 * kWaveMaxDeg;

template <bool CLS>
__device__ __forceinline__ void hash_vertex_wave(const DevCsr &g, uint32_t a, uint32_t *T,
                                                 const uint8_t *table, const TuneParams &tune,
                                                 Acc<CLS> &acc) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t lane = __lane_id();
  const uint32_t la = __builtin_amdgcn_readfirstlane(g.off[a]);
  const uint32_t ea = __builtin_amdgcn_readfirstlane(g.off[a + 1]), d = ea - la;
  [[maybe_unused]] const uint64_t before[4] = {acc.c[0], acc.c[1], acc.c[2], acc.c[3]};
  const uint32_t span = span_of(col, la, ea);
  const int lg = table_log2(d, 4);
  const uint32_t cap = 1u << lg, mask = cap - 1u;
  for (uint32_t k = lane; k < cap; k += 64) T[k] = kEmpty;
  wave_lds_sync();
  if (lane < d) {
    uint32_t x = col[la + lane];
    uint32_t h = (x >> 1) & mask;
    while (atomicCAS(&T[h], kEmpty, x) != kEmpty) h = (h + 1) & mask;
  }
  wave_lds_sync();
  stream_edges<CLS, false>(
      g, la, ea, 1u, PushMeta{g},
      [&](uint32_t j, uint32_t pb, uint32_t eb) {
        return edge_strategy(table, tune, d, span, ea - j - 1, g.col, pb, eb) == S_HASH;
      },
      HashProbe{T, mask}, acc);
  pv_credit<CLS>(g, a, acc, before);
  wave_lds_sync();
}

// Rank window of a source list [lo, hi] for a bitmap of `window` entries: the
// top of the list, where the targets' out-lists concentrate (high ranks are
// high degrees).  Returns the window base.
__device__ __forceinline__ uint32_t window_base(uint32_t lo, uint32_t hi, uint32_t window) {
  return hi - lo + 1u > window ? hi - window + 1u : lo;
}

// A vertex x's out-list N+(x) staged in LDS: a bitmap over the rank window
// [base, hi], one-hot by sign (rank base + r with sign s sets bit 2r + s),
// and the entries below base (a prefix of the sorted list) in the low array
// L[0, nlow).
struct Staged {
  uint32_t la, ea, d, lo, hi, span, base, nlow;
  // f(probe) with the probe of this staging; nlow is wave-uniform, so lists
  // without low entries run a probe without the low test.
  template <class F>
  __device__ __forceinline__ void with_probe(const uint32_t *B, const uint32_t *L,
                                             const F &f) const {
    // zw: the word after the one holding the window's highest bit
    const uint32_t lo2 = base << 1, zw = (((hi - base + 1u) << 1) + 31u) >> 5;
    if (nlow)
      f(BitmapProbe<true>{B, lo2, zw, L, nlow});
    else
      f(BitmapProbe<false>{B, lo2, zw, L, 0u});
  }
};

// Workgroup staging (every thread calls it; at most kLowCap low entries).  B
// must be zero on entry; unstage_block clears the words the list set.
// The list col[la, ea), whose first and last ranks lo and hi the caller read.
__device__ __forceinline__ Staged stage_block_at(const DevCsr &g, uint32_t la, uint32_t ea,
                                                 uint32_t lo, uint32_t hi, uint32_t *B,
                                                 uint32_t *L, uint32_t *s_nlow,
                                                 uint32_t window) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
  Staged st;
  st.la = la;
  st.ea = ea;
  st.d = st.ea - st.la;
  st.lo = lo;
  st.hi = hi;
  st.span = st.hi - st.lo + 1u;
  st.base = window_base(st.lo, st.hi, window);
  for (uint32_t k = tid; k < st.d; k += nthreads) {
    const uint32_t w = col[st.la + k];
    const uint32_t r = w >> 1;
    if (r >= st.base) {
      const uint32_t o = r - st.base;
      atomicOr(&B[o >> 4], 1u << ((o & 15u) * 2u + (w & 1u)));
      if (k == 0) *s_nlow = 0u;
    } else {
      L[k] = w;
      if ((col[st.la + k + 1] >> 1) >= st.base) *s_nlow = k + 1u;  // last low entry (hi >= base)
    }
  }
  __syncthreads();
  st.nlow = __builtin_amdgcn_readfirstlane(*s_nlow);
  return st;
}

__device__ __forceinline__ Staged stage_block(const DevCsr &g, uint32_t x, uint32_t *B,
                                              uint32_t *L, uint32_t *s_nlow, uint32_t window) {
  const uint32_t la = __builtin_amdgcn_readfirstlane(g.off[x]);
  const uint32_t ea = __builtin_amdgcn_readfirstlane(g.off[x + 1]);
  return stage_block_at(g, la, ea, __builtin_amdgcn_readfirstlane(g.col[la] >> 1),
                        __builtin_amdgcn_readfirstlane(g.col[ea - 1] >> 1), B, L, s_nlow,
                        window);
}

__device__ __forceinline__ void unstage_block(const DevCsr &g, const Staged &st, uint32_t *B) {
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < st.d; k += blockDim.x) {
    const uint32_t r = g.col[st.la + k] >> 1;
    if (r >= st.base) B[(r - st.base) >> 4] = 0u;
  }
  __syncthreads();
}

// Wave staging of a list in the wave's own B (zero on entry) and L
// (kWaveMaxDeg words): lane k stages entry k, and the lane's entry is kept
// for unstage_wave.  A list of more than kWaveMaxDeg entries is wave-staged
// only when it spans at most the window (a pull target's list, so chosen by
// the vertex-list pass): its window starts at its first entry, no entry is
// low, and stage_wave_rest sets the bits of entries 64 and up.
struct StagedWave {
  Staged st;
  bool in_w;
  uint32_t o;
};

// Entries [k0, k1) of a wave-staged list longer than a wave (every one in
// the window; the test only keeps a list that breaks that rule inside B).
__device__ __forceinline__ void stage_wave_rest(const uint32_t *__restrict__ col, uint32_t k0,
                                                uint32_t k1, uint32_t base, uint32_t window,
                                                uint32_t *B) {
  for (uint32_t k = k0 + __lane_id(); k < k1; k += 64u) {
    const uint32_t x = col[k];
    const uint32_t o = (x >> 1) - base;
    if (o < window) atomicOr(&B[o >> 4], 1u << ((o & 15u) * 2u + (x & 1u)));
  }
}

// The list col[la, ea) (offsets known to the caller).
__device__ __forceinline__ StagedWave stage_wave_at(const DevCsr &g, uint32_t la, uint32_t ea,
                                                    uint32_t *B, uint32_t *L, uint32_t window) {
  const uint32_t *__restrict__ col = g.col;
  const uint32_t lane = __lane_id();
  StagedWave sw;
  Staged &st = sw.st;
  st.la = la;
  st.ea = ea;
  st.d = st.ea - st.la;
  st.lo = __builtin_amdgcn_readfirstlane(col[st.la] >> 1);
  st.hi = __builtin_amdgcn_readfirstlane(col[st.ea - 1] >> 1);
  st.span = st.hi - st.lo + 1u;
  st.base = window_base(st.lo, st.hi, window);
  const bool mine = lane < st.d;
  uint32_t w = 0;
  sw.o = 0;
  if (mine) {
    w = col[st.la + lane];
    sw.o = (w >> 1) - st.base;  // wraps for entries below the window
  }
  sw.in_w = mine && (w >> 1) >= st.base;
  if (sw.in_w) atomicOr(&B[sw.o >> 4], 1u << ((sw.o & 15u) * 2u + (w & 1u)));
  else if (mine) L[lane] = w;  // the low entries are the lanes [0, nlow)
  if (st.d > kWaveMaxDeg) stage_wave_rest(col, la + 64u, ea, st.base, window, B);
  st.nlow = (uint32_t)__popcll(__ballot(mine && !sw.in_w));
  wave_lds_sync();
  return sw;
}

__device__ __forceinline__ StagedWave stage_wave(const DevCsr &g, uint32_t x, uint32_t *B,
                                                 uint32_t *L, uint32_t window) {
  return stage_wave_at(g, __builtin_amdgcn_readfirstlane(g.off[x]),
                       __builtin_amdgcn_readfirstlane(g.off[x + 1]), B, L, window);
}

__device__ __forceinline__ void unstage_wave(const StagedWave &sw, uint32_t *B) {
  wave_lds_sync();
  if (sw.st.d > kWaveMaxDeg) {
    // a longer list: the words of its whole span (at most the window)
    const uint32_t words = (sw.st.span + 15u) >> 4;
    for (uint32_t k = __lane_id(); k < words; k += 64u) B[k] = 0u;
  } else if (sw.in_w) {
    B[sw.o >> 4] = 0u;
  }
  wave_lds_sync();
}

// Push: stage N+(a) (workgroup: every thread calls it) and count a's bitmap
// edges by streaming each target's out-list.
template <bool CLS>
__device__ __forceinline__ void bitmap_vertex_block(const DevCsr &g, uint32_t a, uint32_t *B,
                                                    uint32_t *L, uint32_t *s_nlow,
                                                    const uint8_t *table,
                                                    const TuneParams &tune, Acc<CLS> &acc) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  [[maybe_unused]] const uint64_t before[4] = {acc.c[0], acc.c[1], acc.c[2], acc.c[3]};
  const Staged st = stage_block(g, a, B, L, s_nlow, 1u << tune.bitmap_span_log2);
  st.with_probe(B, L, [&](const auto &probe) {
    stream_edges<CLS, false>(
        g, st.la + wave, st.ea, nwaves, PushMeta{g},
        [&](uint32_t j, uint32_t pb, uint32_t eb) {
          return edge_strategy(table, tune, st.d, st.span, st.ea - j - 1, g.col, pb, eb) ==
                 S_BITMAP;
        },
        probe, acc);
  });
  pv_credit<CLS>(g, a, acc, before);
  unstage_block(g, st, B);
}

// Push, one wave per source of at most kWaveMaxDeg entries.
template <bool CLS>
__device__ __forceinline__ void bitmap_vertex_wave(const DevCsr &g, uint32_t a, uint32_t *B,
                                                   uint32_t *L, uint32_t window,
                                                   const uint8_t *table,
                                                   const TuneParams &tune, Acc<CLS> &acc) {
  [[maybe_unused]] const uint64_t before[4] = {acc.c[0], acc.c[1], acc.c[2], acc.c[3]};
  const StagedWave sw = stage_wave(g, a, B, L, window);
  const Staged &st = sw.st;
  st.with_probe(B, L, [&](const auto &probe) {
    stream_edges<CLS, false>(
        g, st.la, st.ea, 1u, PushMeta{g},
        [&](uint32_t j, uint32_t pb, uint32_t eb) {
          return edge_strategy(table, tune, st.d, st.span, st.ea - j - 1, g.col, pb, eb) ==
                 S_BITMAP;
        },
        probe, acc);
  });
  pv_credit<CLS>(g, a, acc, before);
  unstage_wave(sw, B);
}

// Pull: stage the target's list N+(v) and count the pull in-edges (u -> v)
// of one chunk of v's in-edge range by streaming the part of N+(u) after v;
// each edge's triangles belong to u.  The chunk starts at in-edge position
// j0 and holds at most `chunk` in-edges; the list entry carries v beside j0,
// so staging need not wait for a load of the chunk's first edge.
// An in-edge is pull iff its bucket (|N+(u) after v|, d+(v)) says pull (v was
// listed only if its list fits).
// Pull streaming.  PACKED: the lists are streamed packed (H: the wave's
// kPackWords-word records and start masks), which pays off for the short lists of the
// wave kernel's targets; the workgroup kernel's targets stream lists long
// enough for the list-at-a-time path, which spends fewer instructions per
// entry.  Per-vertex counts credit each edge's owner and always take the
// list-at-a-time path.
template <bool CLS, bool PACKED, class Probe>
__device__ __forceinline__ void pull_stream(const DevCsr &g, uint32_t first, uint32_t t1,
                                            uint32_t step, uint32_t d, const uint8_t *table,
                                            const Probe &probe, uint64_t *H, Acc<CLS> &acc) {
  auto keep = [&](uint32_t, uint32_t pb, uint32_t eb) {
    return eb > pb && table[bucket_of(eb - pb, d)] == S_PULL;
  };
  if (CLS && g.pv)  // per-vertex counts credit each edge's owner
    stream_edges<CLS, true>(g, first, t1, step, PullMeta{g}, keep, probe, acc);
  else if (PACKED)
    stream_packed<CLS>(g, first, t1, step, PullMeta{g}, keep, probe, H, acc);
  else
    stream_rows<CLS>(g, first, t1, step, PullMeta{g}, keep, probe, acc);
}

// The header of a workgroup-pull chunk: its in-edges [j0, t1) of target v,
// whose list is col[la, ea) with first and last ranks lo and hi, and v's
// graph.  The two-phase kernel gathers the headers of a pop's items in one
// pass (three rounds of loads for up to kBlockPop items), so an item starts
// staging at once instead of after a chain of four dependent loads.
constexpr uint32_t kHdrFields = 8;
enum { kHdrJ0, kHdrV, kHdrLa, kHdrEa, kHdrT1, kHdrLo, kHdrHi, kHdrGraph };

template <bool CLS>
__device__ __forceinline__ void pull_vertex_block_at(const DevCsr &g, const uint32_t *hdr,
                                                     uint32_t *B, uint32_t *L, uint32_t *s_nlow,
                                                     const uint8_t *table,
                                                     const TuneParams &tune, Acc<CLS> &acc) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  auto f = [&](int k) { return __builtin_amdgcn_readfirstlane(hdr[k]); };
  const uint32_t j0 = f(kHdrJ0);
  const uint32_t t1 = min(j0 + (1u << tune.pull_block_chunk_log2), f(kHdrT1));
  const Staged st = stage_block_at(g, f(kHdrLa), f(kHdrEa), f(kHdrLo), f(kHdrHi), B, L, s_nlow,
                                   1u << tune.bitmap_span_log2);
  st.with_probe(B, L, [&](const auto &probe) {
    pull_stream<CLS, false>(g, j0 + wave, t1, nwaves, st.d, table, probe, nullptr, acc);
  });
  unstage_block(g, st, B);
}

// One wave: the in-edges [j0, t1) of a target whose list is col[la, ea) (at
// most kWaveMaxDeg entries, or more over a span within the window).  The
// list's first 64 entries (one per lane; the first and last entries give the
// rank window), the first 64 in-edge records and the
// target class's pull mask are fetched together, so an item starts with one
// memory round trip instead of a chain of four.  The mask has bit c set when
// bucket (c, class of d+(v)) is binned to pull: an in-edge is pull iff its
// suffix length's class c has its bit set.
template <bool CLS>
__device__ __forceinline__ void pull_vertex_wave(const DevCsr &g, uint32_t j0, uint32_t t1,
                                                 uint32_t la, uint32_t ea, uint32_t *B,
                                                 uint32_t *L, uint32_t window, uint64_t *H,
                                                 const uint8_t *table, Acc<CLS> &acc) {
  if (CLS && g.pv) {  // per-vertex counts credit each edge's owner (stream_edges)
    const StagedWave sw = stage_wave_at(g, la, ea, B, L, window);
    const Staged &st = sw.st;
    st.with_probe(B, L, [&](const auto &probe) {
      pull_stream<CLS, true>(g, j0, t1, 1u, st.d, table, probe, H, acc);
    });
    unstage_wave(sw, B);
    return;
  }
  const uint32_t *__restrict__ col = g.col;
  const uint32_t lane = __lane_id();
  const uint32_t d = ea - la;
  // lane 63 of a longer list takes its last entry (the top of the window),
  // and entries 63 .. d - 2 are staged after the first load
  uint32_t w = 0;
  if (lane < d) w = col[lane < 63u || d <= 64u ? la + lane : ea - 1u];
  InEdge r0{0u, 0u};
  if (j0 + lane < t1) r0 = g.tin[j0 + lane];
  const int cb = min(clog2(d), kBucketSide - 1);
  const uint32_t pmask =
      (uint32_t)__ballot(lane < (uint32_t)kBucketSide && table[lane * kBucketSide + cb] == S_PULL);
  // the window from the list's first and last entries
  StagedWave sw;
  Staged &st = sw.st;
  st.la = la;
  st.ea = ea;
  st.d = d;
  st.lo = (uint32_t)__builtin_amdgcn_readlane((int)w, 0) >> 1;
  st.hi = (uint32_t)__builtin_amdgcn_readlane((int)w, (int)min(d, 64u) - 1) >> 1;
  st.span = st.hi - st.lo + 1u;
  st.base = window_base(st.lo, st.hi, window);
  const bool mine = lane < d;
  sw.o = (w >> 1) - st.base;  // wraps for entries below the window
  sw.in_w = mine && (w >> 1) >= st.base;
  if (sw.in_w) atomicOr(&B[sw.o >> 4], 1u << ((sw.o & 15u) * 2u + (w & 1u)));
  else if (mine) L[lane] = w;  // the low entries are the lanes [0, nlow)
  if (d > kWaveMaxDeg) stage_wave_rest(col, la + 63u, ea - 1u, st.base, window, B);
  st.nlow = (uint32_t)__popcll(__ballot(mine && !sw.in_w));
  wave_lds_sync();
  // the first metadata batch's records are the ones fetched above
  const auto meta = [&](uint32_t j, uint32_t &mw, uint32_t &pb, uint32_t &eb, uint32_t &x) {
    const InEdge r = j - j0 < 64u ? r0 : g.tin[j];
    mw = r.w & 1u;
    pb = r.pos;
    eb = r.pos + in_len(r.w);
    x = 0u;
  };
  const auto keep = [&](uint32_t, uint32_t pb, uint32_t eb) {
    return eb > pb && ((pmask >> min(clog2(eb - pb), kBucketSide - 1)) & 1u);
  };
  // a longer list's in-edges carry longer suffixes: rows suit them better
  if (d > kWaveMaxDeg)
    st.with_probe(B, L, [&](const auto &probe) {
      stream_rows<CLS>(g, j0, t1, 1u, meta, keep, probe, acc);
    });
  else
    st.with_probe(B, L, [&](const auto &probe) {
      stream_packed<CLS>(g, j0, t1, 1u, meta, keep, probe, H, acc);
    });
  unstage_wave(sw, B);
}

// Items a workgroup of the two-phase kernel takes per pop in phase 1.
constexpr uint32_t kBlockPop = 4;

// Pop up to `chunk` items for the whole workgroup from an XCD queue.
__device__ __forceinline__ bool block_pop_n(uint32_t (*ctr)[32], uint32_t len, QueueCursor &cur,
                                            uint32_t *s_first, uint32_t *s_count, uint32_t chunk,
                                            uint32_t &first, uint32_t &count) {
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t f = 0, c = 0;
    if (!queue_pop(ctr, len, chunk, cur, f, c)) c = 0;
    *s_first = f;
    *s_count = c;
  }
  __syncthreads();
  first = __builtin_amdgcn_readfirstlane(*s_first);
  count = __builtin_amdgcn_readfirstlane(*s_count);
  return count != 0;
}

// Pop one item for the whole workgroup from an XCD queue.
__device__ __forceinline__ bool block_pop(uint32_t (*ctr)[32], uint32_t len, QueueCursor &cur,
                                          uint32_t *s_first, uint32_t *s_count,
                                          uint32_t &first) {
  uint32_t count;
  return block_pop_n(ctr, len, cur, s_first, s_count, 1u, first, count);
}

// One workgroup per source vertex (source lists longer than one wave).
template <bool CLS>
__global__ __launch_bounds__(kVertexBlock) void k_hash(DevCsr g, Ctrl *ctrl,
                                                       const DispatchParams *P,
                                                       const uint32_t *__restrict__ vlist) {
  extern __shared__ __attribute__((aligned(128))) uint32_t T[];
  __shared__ uint8_t table[kBuckets];
  __shared__ uint32_t s_first, s_count;
  __shared__ TuneParams tune;
  const uint32_t len = ctrl->len[2];
  if (len == 0) return;  // the whole grid: no vertex has workgroup-hash edges
  if (threadIdx.x < kBuckets) table[threadIdx.x] = ctrl->table[threadIdx.x];
  if (threadIdx.x == 0) tune = P->tune;
  QueueCursor cur{blockIdx.x & 7u, 0u};
  Acc<CLS> acc;
  uint32_t first;
  while (block_pop(ctrl->queue[0], len, cur, &s_first, &s_count, first))
    hash_vertex_block<CLS>(g, __builtin_amdgcn_readfirstlane(vlist[first]), T, table, tune, acc);
  acc.flush(ctrl);
}

// One wave per source vertex (source lists of at most kWaveMaxDeg entries).
constexpr uint32_t kWaveChunk = 8;  // the most items a wave takes per queue pop

// Headers of the wave-pull chunks among the items of one queue pop, loaded
// together so that an item does not start with a chain of dependent loads.
// Items skip <= i < cnt are pull-list entries base + (i - skip).  Lane
// kWaveChunk f + i of the result holds field f of item i: 0 the chunk's first
// in-edge, 1 its target v, 2 off[v], 3 off[v + 1], 4 tin_off[v + 1], 5 v's
// graph (0 without graph_of).  head_of reads one field back.
constexpr uint32_t kHeadFields = 6;
static_assert(kWaveChunk * kHeadFields <= 64, "a pop's headers fit one register");

__device__ __forceinline__ uint32_t pull_heads(const DevCsr &g, const uint32_t *starts,
                                               const uint32_t *targets,
                                               const uint32_t *graph_of, uint32_t base,
                                               uint32_t skip, uint32_t cnt) {
  const uint32_t lane = __lane_id(), i = lane % kWaveChunk, f = lane / kWaveChunk;
  uint32_t x = 0;
  if (i >= skip && i < cnt && f < kHeadFields) {
    const uint32_t k = base + (i - skip);
    if (f == 0) {
      x = starts[k];
    } else {
      const uint32_t v = targets[k];
      const uint32_t *p = f == 2               ? g.off + v
                          : f == 3             ? g.off + v + 1
                          : f == 4             ? g.tin_off + v + 1
                          : f == 5 && graph_of ? graph_of + v
                                               : nullptr;
      x = f == 1 ? v : p ? *p : 0u;
    }
  }
  return x;
}

__device__ __forceinline__ uint32_t head_of(uint32_t heads, uint32_t field, uint32_t i) {
  return (uint32_t)__builtin_amdgcn_readlane((int)heads, (int)(field * kWaveChunk + i));
}

template <bool CLS>
__global__ __launch_bounds__(256) void k_hash_wave(DevCsr g, Ctrl *ctrl,
                                                   const DispatchParams *P,
                                                   const uint32_t *__restrict__ vlist) {
  __shared__ uint32_t Tall[4 * kWaveCap];
  __shared__ uint8_t table[kBuckets];
  __shared__ TuneParams tune;
  const uint32_t len = ctrl->len[4];
  if (len == 0) return;  // the whole grid: no vertex has wave-hash edges
  const int wave = threadIdx.x >> 6;
  const uint32_t lane = __lane_id();
  table[threadIdx.x] = ctrl->table[threadIdx.x];
  if (threadIdx.x == 0) tune = P->tune;
  __syncthreads();
  uint32_t *T = Tall + wave * kWaveCap;
  QueueCursor cur{blockIdx.x & 7u, 0u};
  Acc<CLS> acc;
  while (true) {
    uint32_t first = 0, count = 0;
    if (lane == 0 &&
        !queue_pop<true>(ctrl->queue[2], len, (uint32_t)tune.wave_pop, cur, first, count))
      count = 0;
    // lane 0 popped: readfirstlane makes the pop wave-uniform to the
    // compiler, so the item loop runs on scalar registers
    first = __builtin_amdgcn_readfirstlane(first);
    count = __builtin_amdgcn_readfirstlane(count);
    if (count == 0) break;
    for (uint32_t it = 0; it < count; it++)
      hash_vertex_wave<CLS>(g, __builtin_amdgcn_readfirstlane(vlist[first + it]), T, table, tune,
                            acc);
  }
  acc.flush(ctrl);
}

// One workgroup per source vertex with more than kWaveMaxDeg out-edges (push):
// LDS bitmap (zeroed once per workgroup) followed by the kLowCap-word low
// array.  (Workgroup pull chunks are counted by the two-phase kernel.)
template <bool CLS>
__global__ __launch_bounds__(kBitmapBlock) void k_bitmap(DevCsr g, Ctrl *ctrl,
                                                         const DispatchParams *P,
                                                         const uint32_t *__restrict__ vlist,
                                                         uint32_t bitmap_words) {
  extern __shared__ __attribute__((aligned(128))) uint32_t B[];
  uint32_t *L = B + bitmap_words + kGuard;
  __shared__ uint8_t table[kBuckets];
  __shared__ uint32_t s_first, s_count, s_nlow;
  __shared__ TuneParams tune;
  const uint32_t len = ctrl->len[3];
  if (len == 0) return;  // the whole grid: no vertex has this kernel's edges
  if (threadIdx.x < kBuckets) table[threadIdx.x] = ctrl->table[threadIdx.x];
  if (threadIdx.x == 0) tune = P->tune;
  for (uint32_t k = threadIdx.x; k < bitmap_words + kGuard; k += blockDim.x) B[k] = 0u;
  QueueCursor cur{blockIdx.x & 7u, 0u};
  Acc<CLS> acc;
  uint32_t first;
  while (block_pop(ctrl->queue[1], len, cur, &s_first, &s_count, first))
    bitmap_vertex_block<CLS>(g, __builtin_amdgcn_readfirstlane(vlist[first]), B, L, &s_nlow,
                             table, tune, acc);
  acc.flush(ctrl);
}

// One wave per vertex (lists of at most kWaveMaxDeg entries): each wave owns
// window_words bitmap words followed by kWaveMaxDeg low words.  Push stages
// sources, pull the targets of in-edge chunks.
template <bool CLS, bool PULL>
__global__ __launch_bounds__(256) void k_bitmap_wave(DevCsr g, Ctrl *ctrl,
                                                     const DispatchParams *P,
                                                     const uint32_t *__restrict__ vlist,
                                                     const uint32_t *__restrict__ vtarget,
                                                     uint32_t window_words, uint32_t defer) {
  extern __shared__ __attribute__((aligned(128))) uint32_t lds[];
  __shared__ uint8_t table[kBuckets];
  __shared__ TuneParams tune;
  __shared__ uint64_t s_heads[PULL ? 4 * kPackWords / 2 : 1];  // packed streaming, per wave
  const uint32_t len = ctrl->len[PULL ? 7 : 5];
  if (len == 0) return;  // the whole grid: no vertex has this kernel's edges
  // defer (pull): the two-phase launch that follows counts this list with the
  // workgroup chunks whenever there are any
  if (PULL && defer && ctrl->len[6] != 0) return;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = __lane_id();
  table[threadIdx.x] = ctrl->table[threadIdx.x];
  if (threadIdx.x == 0) tune = P->tune;
  uint32_t *B = lds + wave * (window_words + kGuard + kWaveMaxDeg);
  uint32_t *L = B + window_words + kGuard;
  for (uint32_t k = lane; k < window_words + kGuard; k += 64) B[k] = 0u;
  __syncthreads();
  const uint32_t window = window_words * 16u;
  QueueCursor cur{blockIdx.x & 7u, 0u};
  Acc<CLS> acc;
  while (true) {
    uint32_t first = 0, count = 0;
    if (lane == 0 &&
        !queue_pop<true>(ctrl->queue[PULL ? 6 : 4], len, (uint32_t)tune.wave_pop, cur, first,
                         count))
      count = 0;
    // lane 0 popped: readfirstlane makes the pop wave-uniform to the
    // compiler, so the item loop runs on scalar registers
    first = __builtin_amdgcn_readfirstlane(first);
    count = __builtin_amdgcn_readfirstlane(count);
    if (count == 0) break;
    const uint32_t heads =
        PULL ? pull_heads(g, vlist, vtarget, nullptr, first, 0u, count) : 0u;
    for (uint32_t it = 0; it < count; it++) {
      if constexpr (PULL) {
        const uint32_t j0 = head_of(heads, 0, it);
        pull_vertex_wave<CLS>(g, j0, min(j0 + (1u << tune.pull_wave_chunk_log2), head_of(heads, 4, it)),
                              head_of(heads, 2, it), head_of(heads, 3, it), B, L, window,
                              s_heads + wave * (kPackWords / 2u), table, acc);
      } else {
        bitmap_vertex_wave<CLS>(g, __builtin_amdgcn_readfirstlane(vlist[first + it]), B, L,
                                window, table, tune, acc);
      }
    }
  }
  acc.flush(ctrl);
}

// ---- batched counting: one persistent kernel over the union of k graphs ----
// Three instantiations by the item kinds they compile (KIND):
//   kKindAll        a batch with merge, hash or bitmap items (two-class);
//   kKindPullBatch  a batch whose only items are pull chunks (the learned
//                   table's choice for batches of small graphs);
//   kKindPullOne    one graph's pull items, as a batch of one: graph_of null,
//                   queues q1, q2 that no other kernel of that count uses;
//                   two-class totals go to per_graph[0..1] of the class
//                   counters' copy the workgroup picks, the four-class
//                   variant (CLS) adds its classes to ctrl->t.  The two pull
//                   phases then share one launch, with no boundary and tail
//                   between the workgroup and the wave chunks.
// The pull-only instantiations carry no merge, hash or push-bitmap state, so
// they fit the 64-VGPR cap without spilling to scratch (the all-kinds kernel
// spills 36 bytes per lane).  Both batch instantiations are launched; each
// reads the list lengths and the one whose kinds are absent returns at once.
// Two phases per workgroup, each draining one queue:
//   1. workgroup items, popped by the whole workgroup (heaviest kinds first):
//      merge-lane chunks of 2^lane_item_log2 edges (tuned; 512 by default,
//      one per thread), hash vertices,
//      bitmap vertices, workgroup-pull in-edge chunks;
//   2. wave items, popped by each wave on its own (wave_pop per pop):
//      merge-wave edges, wave-hash and wave-bitmap vertices, wave-pull
//      in-edge chunks.  A wave never waits for the others, so short and long
//      items mix without idling the workgroup.
// A batch's queues keep contiguous per-XCD ranges: a batch graph's ranks are
// contiguous, so each small graph stays in one XCD's L2 (dealing the wave
// pops over the XCDs instead made bitmap counting of the 4,096-graph batch
// 109 -> 168 ms).  The workgroup is capped at 64 VGPRs (8 waves per SIMD).
// Every item's triangles are added to the counters of the graph owning its
// vertex (the source, or for pull the target: the graphs are disjoint, so
// its in-edges come from its own graph; graph_of is indexed by rank; each
// graph's ranks are one contiguous range).
//   All LDS structures share one region R, kept zero between items.  Phase 1
// uses the workgroup bitmap R[0, bitmap_words), kGuard zero words and the
// low array after them,
// or the workgroup hash table R[0, hash_slots).  In phase 2 wave w owns the
// slot R[w S, (w + 1) S), S = window_words + kGuard + kWaveMaxDeg + kPackWords:
// its hash table or window bitmap at the start, the window's guard words and
// low array after the window, and its packed-streaming records and start masks at the end
// (they need no clearing: phase 1, which needs zeros there, is over by then,
// and the streaming clears the masks itself).  Bitmaps
// clear what they set; tables and low arrays are cleared after use.
constexpr uint32_t kBatchBlock = 512;
constexpr uint32_t kWavesPerBatchBlock = kBatchBlock / 64;
constexpr int kKindAll = 0, kKindPullBatch = 1, kKindPullOne = 2;

__device__ __forceinline__ void lane_flush_graph(unsigned long long *per_graph, bool valid,
                                                 uint32_t gid, uint32_t c0, uint32_t c1) {
  const uint64_t vm = __ballot(valid);
  if (!vm) return;
  const uint32_t g0 = __shfl(gid, __builtin_ctzll(vm), 64);
  if (__ballot(valid && gid == g0) == vm) {
    uint64_t b = valid ? c0 : 0u, u = valid ? c1 : 0u;
    for (int d = 32; d > 0; d >>= 1) {
      b += __shfl_down(b, d, 64);
      u += __shfl_down(u, d, 64);
    }
    wave_add(&per_graph[2 * g0], b);
    wave_add(&per_graph[2 * g0 + 1], u);
  } else if (valid) {
    if (c0) atomicAdd(&per_graph[2 * gid], (unsigned long long)c0);
    if (c1) atomicAdd(&per_graph[2 * gid + 1], (unsigned long long)c1);
  }
}

template <bool CLS, int KIND>
__global__ __launch_bounds__(kBatchBlock, 8) void k_two_phase(
    DevCsr g, Ctrl *ctrl, const DispatchParams *P, const uint32_t *__restrict__ lane_list,
    const uint32_t *__restrict__ wave_list, const uint32_t *__restrict__ vlist, uint64_t vstride,
    const uint32_t *__restrict__ graph_of,
    unsigned long long *__restrict__ per_graph, uint32_t bitmap_words, uint32_t hash_slots,
    uint32_t window_words, uint32_t region_words, uint32_t q1, uint32_t q2) {
  // four-class counting serves one graph's pull items only: an all-kinds
  // four-class kernel would have to count merge-lane items in four classes
  static_assert(!CLS || KIND == kKindPullOne, "four-class counting is one graph's pull items");
  constexpr bool ALL = KIND == kKindAll;
  // the region starts on a 128-byte boundary whatever the static arrays take:
  // the wave records are 8-byte words, and the start that followed the static
  // arrays (4-byte aligned) once left them misaligned, 13% slower on a batch
  extern __shared__ __attribute__((aligned(128))) uint32_t R[];
  __shared__ uint8_t table[kBuckets];
  __shared__ uint32_t s_first, s_count, s_nlow;
  __shared__ TuneParams tune;
  const uint32_t tid = threadIdx.x, lane = __lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t others = ctrl->len[0] | ctrl->len[1] | ctrl->len[2] | ctrl->len[3] |
                          ctrl->len[4] | ctrl->len[5];
  const uint32_t nl = ALL ? ctrl->len[0] : 0u, nwv = ALL ? ctrl->len[1] : 0u,
                 nh = ALL ? ctrl->len[2] : 0u, nb = ALL ? ctrl->len[3] : 0u,
                 nhw = ALL ? ctrl->len[4] : 0u, nbw = ALL ? ctrl->len[5] : 0u,
                 np = ctrl->len[6], nwp = ctrl->len[7];
  // the whole grid returns when its kinds are absent: a batch runs the
  // all-kinds kernel when it has items other than pull chunks, else the
  // pull-only one; one graph's launch needs workgroup-pull chunks (without
  // them the wave-pull kernel, launched before, has counted the wave chunks)
  if constexpr (KIND == kKindAll) {
    if (others == 0u) return;
  } else if constexpr (KIND == kKindPullBatch) {
    if (others != 0u || (np | nwp) == 0u) return;
  } else {
    if (np == 0u) return;
  }
  if (tid < kBuckets) table[tid] = ctrl->table[tid];
  if (tid == 0) tune = P->tune;
  for (uint32_t k = tid; k < region_words; k += blockDim.x) R[k] = 0u;
  __syncthreads();  // table, tune and the zeroed region are visible to every wave
  // one graph (no graph_of): its counter pair is one copy of the class
  // counters, chosen by workgroup
  if (!graph_of) per_graph += (blockIdx.x % kTSlots) * kTStride;
  GraphAcc<CLS> ga;
  Acc<CLS> &acc = ga.acc;
  // phase 1: workgroup items [lane chunks | hash | bitmap | pull chunks]
  {
    const uint32_t lane_item = 1u << tune.lane_item_log2;
    const uint32_t b1 = (nl + lane_item - 1) / lane_item, b2 = b1 + nh, b3 = b2 + nb,
                   total = b3 + np;
    QueueCursor cur{blockIdx.x & 7u, 0u};
    // kBlockPop consecutive items per pop: a pop is a round trip that every
    // wave of the workgroup waits for (one item per pop: RMAT-22 5.75 ms,
    // two 5.62, four 5.61)
    // the pop's pull-chunk headers: the last words of the region, past every
    // phase-1 structure (phase 2 reuses them once phase 1 is over)
    uint32_t *const Hdr = R + region_words - kBlockPop * kHdrFields;
    uint32_t it0, cnt;
    while (block_pop_n(ctrl->queue[q1], total, cur, &s_first, &s_count, kBlockPop, it0, cnt)) {
      if (it0 + cnt > b3) {  // the pop holds pull chunks: their headers, thread k item k
        if (tid < cnt && it0 + tid >= b3) {
          const uint32_t k = it0 + tid - b3;
          uint32_t *h = Hdr + tid * kHdrFields;
          const uint32_t j0 = vlist[4 * vstride + k], v = vlist[6 * vstride + k];
          const uint32_t la = g.off[v], ea = g.off[v + 1];
          h[kHdrJ0] = j0;
          h[kHdrV] = v;
          h[kHdrLa] = la;
          h[kHdrEa] = ea;
          h[kHdrT1] = g.tin_off[v + 1];
          h[kHdrGraph] = graph_of ? graph_of[v] : 0u;
          h[kHdrLo] = g.col[la] >> 1;
          h[kHdrHi] = g.col[ea - 1] >> 1;
        }
        __syncthreads();
      }
      for (uint32_t it = it0; it < it0 + cnt; it++) {
        if (ALL && it < b1) {  // merge-lane chunk: kBatchBlock edges per round
          const uint32_t end = min((it + 1) * lane_item, nl);
          for (uint32_t i0 = it * lane_item; i0 < end; i0 += kBatchBlock) {
            const uint32_t i = i0 + tid;
            const bool valid = i < end;
            uint32_t c[4] = {0, 0, 0, 0};
            uint32_t gid = 0;
            if (valid) {
              const uint32_t e = lane_list[i];
              gid = graph_of[g.src[e]];
              merge_lane_edge<false, false>(g, e, c);  // batches: two-class
            }
            lane_flush_graph(per_graph, valid, gid, c[0], c[1]);
          }
        } else if (ALL && it < b2) {  // hash vertex
          const uint32_t a = __builtin_amdgcn_readfirstlane(vlist[it - b1]);
          ga.to(per_graph, graph_of[a]);
          hash_vertex_block<CLS>(g, a, R, table, tune, acc);
          for (uint32_t k = tid; k < hash_slots; k += blockDim.x) R[k] = 0u;
        } else if (ALL && it < b3) {  // bitmap vertex
          const uint32_t a = __builtin_amdgcn_readfirstlane(vlist[vstride + it - b2]);
          ga.to(per_graph, graph_of[a]);
          bitmap_vertex_block<CLS>(g, a, R, R + bitmap_words + kGuard, &s_nlow, table, tune,
                                   acc);
          for (uint32_t k = tid; k < kLowCap; k += blockDim.x) R[bitmap_words + kGuard + k] = 0u;
        } else {  // workgroup-pull chunk
          const uint32_t *h = Hdr + (it - it0) * kHdrFields;
          ga.to(per_graph, __builtin_amdgcn_readfirstlane(h[kHdrGraph]));
          pull_vertex_block_at<CLS>(g, h, R, R + bitmap_words + kGuard, &s_nlow, table, tune,
                                    acc);
          for (uint32_t k = tid; k < kLowCap; k += blockDim.x) R[bitmap_words + kGuard + k] = 0u;
        }
      }
    }
  }
  // phase 2: wave items [merge-wave | wave hash | wave bitmap | wave pull]
  {
    const uint32_t slot = window_words + kGuard + kWaveMaxDeg + kPackWords;
    uint32_t *const W = R + wave * slot;
    uint32_t *const LW = W + window_words + kGuard;
    uint64_t *const HW = (uint64_t *)(LW + kWaveMaxDeg);
    const uint32_t window = window_words * 16u;
    const uint32_t w1 = nwv, w2 = w1 + nhw, w3 = w2 + nbw, total = w3 + nwp;
    // A batch's waves start on their own XCD's share (blockIdx mod 8 names the
    // XCD), so each XCD works on one contiguous range of ranks and its small
    // graphs stay in that XCD's L2; starting each wave of a workgroup on
    // another XCD's share cost the 4,096-graph batch 6% (17.89 against
    // 16.85 ms).  One graph's pops are dealt round-robin over the XCDs.
    QueueCursor cur{KIND == kKindPullOne ? (blockIdx.x + wave) & 7u : blockIdx.x & 7u, 0u};
    while (true) {
      uint32_t first = 0, count = 0;
      // a batch keeps each XCD on a contiguous range (its graphs stay in one
      // L2); one graph's pops are dealt round-robin over the XCDs, which
      // spreads the heavy top of the list (as in the single-graph wave kernels)
      if (lane == 0 && !queue_pop<KIND == kKindPullOne>(ctrl->queue[q2], total,
                                                         (uint32_t)tune.wave_pop, cur, first,
                                                         count))
        count = 0;
      // lane 0 popped: readfirstlane makes the pop wave-uniform to the
      // compiler, so the item loop runs on scalar registers
      first = __builtin_amdgcn_readfirstlane(first);
      count = __builtin_amdgcn_readfirstlane(count);
      if (count == 0) break;
      // headers of the pop's wave-pull items (items skip.. of the pop)
      const uint32_t skip = first >= w3 ? 0u : min(w3 - first, count);
      const uint32_t heads =
          skip < count ? pull_heads(g, vlist + 5 * vstride, vlist + 7 * vstride, graph_of,
                                    first + skip - w3, skip, count)
                       : 0u;
      for (uint32_t it = first; it < first + count; it++) {
        if (ALL && it < w1) {  // merge-wave edge
          const uint32_t e = __builtin_amdgcn_readfirstlane(wave_list[it]);
          ga.to(per_graph, graph_of[g.src[e]]);
          merge_wave_edge<CLS>(g, e, acc);
        } else if (ALL && it < w2) {  // wave-hash vertex
          const uint32_t a = __builtin_amdgcn_readfirstlane(vlist[2 * vstride + it - w1]);
          ga.to(per_graph, graph_of[a]);
          hash_vertex_wave<CLS>(g, a, W, table, tune, acc);
#pragma unroll
          for (uint32_t k = lane; k < kWaveCap; k += 64) W[k] = 0u;
        } else if (ALL && it < w3) {  // wave-bitmap vertex
          const uint32_t a = __builtin_amdgcn_readfirstlane(vlist[3 * vstride + it - w2]);
          ga.to(per_graph, graph_of[a]);
          bitmap_vertex_wave<CLS>(g, a, W, LW, window, table, tune, acc);
          LW[lane] = 0u;  // kWaveMaxDeg == 64 low words
        } else {  // wave-pull chunk
          const uint32_t i = it - first;
          const uint32_t j0 = head_of(heads, 0, i);
          ga.to(per_graph, head_of(heads, 5, i));
          pull_vertex_wave<CLS>(g, j0, min(j0 + (1u << tune.pull_wave_chunk_log2), head_of(heads, 4, i)),
                                  head_of(heads, 2, i), head_of(heads, 3, i), W, LW, window,
                                  HW, table, acc);
          LW[lane] = 0u;
        }
      }
    }
  }
  ga.finish(per_graph, ctrl);
}

// Per-vertex counts from rank order to vertex-id order (perm: rank -> id).
__global__ void k_pv_to_ids(const unsigned long long *__restrict__ pv, const uint32_t *__restrict__ perm,
                            uint32_t n, unsigned long long *__restrict__ out) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const uint64_t x = perm[r];
#pragma unroll
    for (int k = 0; k < 4; k++) out[4 * x + k] = pv[4ull * r + k];
  }
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
  return (size_t)4u << std::max(t.hash_max_log2 + 2, 6);
}
size_t bitmap_lds_bytes(const TuneParams &t) {
  return (size_t)((1u << t.bitmap_span_log2) / 4u) + 4u * (kGuard + kLowCap);
}
size_t wave_bitmap_lds_bytes(const TuneParams &t) {
  return 4u * (size_t)(256 / 64) * ((1u << t.wave_window_log2) / 16u + kGuard + kWaveMaxDeg);
}

#define STC_TRY(expr)                \
  do {                               \
    hipError_t e_ = (expr);          \
    if (e_ != hipSuccess) return e_; \
  } while (0)

// Raise the dynamic-LDS limit of the kernels that use more than 64 KB; called
// once per process before any launch (never inside a stream capture).
hipError_t init_kernel_attributes() {
  static bool done = false;
  if (done) return hipSuccess;
  const void *fns[] = {(const void *)k_bitmap<true>,             (const void *)k_bitmap<false>,
                       (const void *)k_bitmap_wave<true, false>, (const void *)k_bitmap_wave<false, false>,
                       (const void *)k_bitmap_wave<true, true>,  (const void *)k_bitmap_wave<false, true>,
                       (const void *)k_hash<true>,               (const void *)k_hash<false>,
                       (const void *)k_two_phase<true, kKindPullOne>,
                       (const void *)k_two_phase<false, kKindPullOne>,
                       (const void *)k_two_phase<false, kKindAll>,
                       (const void *)k_two_phase<false, kKindPullBatch>};
  for (const void *f : fns)
    STC_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
  done = true;
  return hipSuccess;
}

namespace {

struct BinOut {
  uint32_t *lane_list, *wave_list, *vh, *vb, *vw, *vwb, *vp, *vwp;
  uint32_t *vpt, *vwpt;  // the target of each pull list entry
  uint64_t vstride;      // list L (vh, vb, vw, vwb, vp, vwp, vpt, vwpt) at vh + L vstride
};

// Buffers: cb.gcount holds Ctrl; cb.lists the per-workgroup segments and the
// packed merge-lane / merge-wave edge lists; cb.vflag n words kept zero
// between runs; cb.vlist the six vertex lists (vlist_stride words each) and the
// per-tile list counts and offsets.  Every buffer is sized before the first launch, so a second call
// with the same graph allocates nothing (and can be captured in a graph).
// Sizes of the binning pass's buffers for a graph of n vertices and m
// oriented edges, and where its lists sit inside them.
struct BinLayout {
  int bin_grid;
  uint32_t seg, ntiles;
  uint64_t seg_words, vstride;
  size_t lists_bytes, blk_bytes, vflag_bytes, vlist_bytes;
};

BinLayout bin_layout(uint32_t n, uint64_t m, int num_cus) {
  BinLayout b;
  b.bin_grid = grid_for(m, 256, num_cus * 8);
  // per-workgroup segment size: the most edges one workgroup visits
  const uint64_t stride = (uint64_t)b.bin_grid * 256;
  b.seg = (uint32_t)(((m + stride - 1) / stride) * 256);
  b.seg_words = (uint64_t)b.seg * b.bin_grid;
  b.ntiles = (uint32_t)(((uint64_t)n + kVTile - 1) / kVTile);
  b.vstride = vlist_stride(n, m);
  b.lists_bytes = (2 * b.seg_words + 2 * (m ? m : 1)) * 4;
  b.blk_bytes = (4 * (uint64_t)b.bin_grid + 8) * 4;
  b.vflag_bytes = ((uint64_t)n + 1) * 4;
  b.vlist_bytes = ((kVLists + 2) * b.vstride + 2 * kVLists * (uint64_t)b.ntiles + 8) * 4;
  return b;
}

void bin_lists(const BinLayout &b, uint64_t m, uint32_t *lists, uint32_t *vlist, BinOut &out) {
  out.lane_list = lists + 2 * b.seg_words;
  out.wave_list = out.lane_list + (m ? m : 1);
  out.vh = vlist;
  out.vstride = b.vstride;
  out.vb = out.vh + b.vstride;
  out.vw = out.vb + b.vstride;
  out.vwb = out.vw + b.vstride;
  out.vp = out.vwb + b.vstride;
  out.vwp = out.vp + b.vstride;
  out.vpt = out.vwp + b.vstride;  // list L's targets sit 2 strides after it
  out.vwpt = out.vpt + b.vstride;
}

hipError_t enqueue_binning(const DevCsr &g, const DispatchParams *dp, CountBuffers &cb,
                           int num_cus, hipStream_t st, const hipEvent_t *marks, BinOut &out) {
  const uint64_t m = g.m;
  const uint32_t n = g.n;
  const BinLayout lay = bin_layout(n, m, num_cus);
  const int bin_grid = lay.bin_grid;
  const uint32_t seg = lay.seg, ntiles = lay.ntiles;
  const uint64_t seg_words = lay.seg_words, vstride = lay.vstride;
  STC_TRY(cb.gcount.ensure(sizeof(Ctrl)));
  STC_TRY(cb.lists.ensure(lay.lists_bytes));
  STC_TRY(cb.blk.ensure(lay.blk_bytes));
  if (cb.vflag.bytes < lay.vflag_bytes) {
    STC_TRY(cb.vflag.ensure(lay.vflag_bytes));
    hipLaunchKernelGGL(k_fill_words, dim3(grid_for((uint64_t)n + 1, 256, 8192)), dim3(256),
                       0, st, cb.vflag.as<uint32_t>(), cb.vflag.bytes / 4, 0u);
    STC_TRY(hipGetLastError());
  }
  STC_TRY(cb.vlist.ensure(lay.vlist_bytes));
  Ctrl *ctrl = cb.gcount.as<Ctrl>();
  uint32_t *lane_seg = cb.lists.as<uint32_t>();
  uint32_t *wave_seg = lane_seg + seg_words;
  bin_lists(lay, m, cb.lists.as<uint32_t>(), cb.vlist.as<uint32_t>(), out);
  uint32_t *blk_cnt = cb.blk.as<uint32_t>();
  uint32_t *blk_pos = blk_cnt + 2 * bin_grid;
  uint32_t *vf = cb.vflag.as<uint32_t>();
  uint32_t *tile_cnt = out.vwpt + vstride;
  uint32_t *tile_pos = tile_cnt + kVLists * (uint64_t)ntiles;

  if (marks) STC_TRY(hipEventRecord(marks[0], st));
  hipLaunchKernelGGL(k_fill_words, dim3(8), dim3(256), 0, st, (uint32_t *)ctrl,
                     (uint64_t)(sizeof(Ctrl) / 4), 0u);
  STC_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bin_hist, dim3(bin_grid), dim3(256), 0, st, g, ctrl);
  hipLaunchKernelGGL(k_bin_table, dim3(1), dim3(kBuckets), 0, st, ctrl, dp);
  STC_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bin_scatter, dim3(bin_grid), dim3(256), 0, st, g, ctrl, dp,
                     lane_seg, wave_seg, seg, blk_cnt, vf);
  STC_TRY(hipGetLastError());
  STC_TRY(scan_exclusive_u32(blk_cnt, blk_pos, 2 * (uint64_t)bin_grid, cb.scan, st));
  hipLaunchKernelGGL(k_list_compact, dim3(2 * bin_grid), dim3(256), 0, st, lane_seg, wave_seg,
                     seg, blk_cnt, blk_pos, (uint32_t)bin_grid, out.lane_list, out.wave_list,
                     ctrl);
  STC_TRY(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL(k_vcount, dim3(ntiles), dim3(256), 0, st, g, ctrl, dp, vf, tile_cnt);
    STC_TRY(hipGetLastError());
    STC_TRY(scan_exclusive_u32(tile_cnt, tile_pos, kVLists * (uint64_t)ntiles, cb.scan, st));
    hipLaunchKernelGGL(k_vscatter, dim3(ntiles), dim3(256), 0, st, g, dp, vf, tile_cnt,
                       tile_pos, out.vh, vstride, ctrl);
    STC_TRY(hipGetLastError());
  }
  if (marks) STC_TRY(hipEventRecord(marks[1], st));
  return hipSuccess;
}

}  // namespace

hipError_t launch_two_phase(const DevCsr &g, const DispatchParams &hp, const DispatchParams *dp,
                            CountBuffers &cb, const BinOut &L, const uint32_t *graph_of,
                            unsigned long long *per_graph, int num_cus, hipStream_t st,
                            uint32_t q1, uint32_t q2, bool one_graph, bool classes);

// ---- static scratch (the allocation-free counting entry point) -------------
// The counting pass's buffers for callers that cannot allocate device memory
// (code under stream capture, serving loops): a file-scope array that every
// GPU running this library holds until the process ends.  In host code its
// name is only a stand-in; the device address comes from
// static_scratch_address, once per device with that device current, and only
// that address is handed to kernels and fills.  Every call on one GPU shares
// the array, so calls that overlap on one GPU race.
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

hipError_t static_scratch_address(void **ptr) {
  *ptr = nullptr;
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

bool carve_count_buffers(void *base, size_t bytes, uint32_t n, uint64_t m, int num_cus,
                         CountBuffers &cb, size_t *need) {
  const BinLayout lay = bin_layout(n, m, num_cus);
  const uint64_t scan_n[2] = {2 * (uint64_t)lay.bin_grid, (uint64_t)kVLists * lay.ntiles};
  size_t level[4];
  for (int l = 0; l < 4; l++)
    level[l] = std::max<size_t>(256, std::max(scan_level_bytes(scan_n[0], l),
                                              scan_level_bytes(scan_n[1], l)));
  const size_t parts[] = {sizeof(Ctrl),    lay.lists_bytes, lay.blk_bytes, lay.vflag_bytes,
                          lay.vlist_bytes, level[0],        level[1],      level[2],
                          level[3]};
  DevBuf *views[] = {&cb.gcount, &cb.lists,         &cb.blk,           &cb.vflag,
                     &cb.vlist,  &cb.scan.level[0], &cb.scan.level[1], &cb.scan.level[2],
                     &cb.scan.level[3]};
  size_t off = 0;
  for (size_t b : parts) off += (b + 255) & ~(size_t)255;
  if (need) *need = off;
  if (off > bytes) return false;
  off = 0;
  for (size_t i = 0; i < sizeof(parts) / sizeof(parts[0]); i++) {
    views[i]->borrow((uint8_t *)base + off, parts[i]);
    off += (parts[i] + 255) & ~(size_t)255;
  }
  return true;
}

hipError_t clear_vflags(CountBuffers &cb, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_words, dim3(grid_for((uint64_t)n + 1, 256, 8192)), dim3(256), 0, st,
                     cb.vflag.as<uint32_t>(), (uint64_t)n + 1, 0u);
  return hipGetLastError();
}

hipError_t enqueue_count(const DevCsr &g, const DispatchParams &hp,
                         const DispatchParams *dp, CountBuffers &cb, bool classes,
                         int num_cus, hipStream_t st, const hipEvent_t *marks) {
  BinOut L;
  STC_TRY(enqueue_binning(g, dp, cb, num_cus, st, marks, L));
  Ctrl *ctrl = cb.gcount.as<Ctrl>();
  const int grid = num_cus * 8;
  if (g.pv) {  // per-vertex counters start from zero on every run
    hipLaunchKernelGGL(k_fill_words, dim3(grid_for(8ull * g.n, 256, 8192)), dim3(256), 0, st,
                       (uint32_t *)g.pv, 8ull * g.n, 0u);
    STC_TRY(hipGetLastError());
  }
  const size_t hash_lds = hash_lds_bytes(hp.tune);
  const size_t bm_lds = bitmap_lds_bytes(hp.tune);
  const int hash_grid = num_cus * (int)std::max<size_t>(1, std::min<size_t>(8, (128 * 1024) / hash_lds));
  const int bm_grid = num_cus * (int)std::max<size_t>(1, std::min<size_t>(4, (148 * 1024) / bm_lds));
  const uint32_t bm_words = (1u << hp.tune.bitmap_span_log2) / 16u;
  const size_t bw_lds = wave_bitmap_lds_bytes(hp.tune);
  const int bw_grid = num_cus * (int)std::max<size_t>(1, std::min<size_t>(8, (152 * 1024) / bw_lds));
  const uint32_t bw_words = (1u << hp.tune.wave_window_log2) / 16u;
  auto mark = [&](int i) -> hipError_t {
    return marks ? hipEventRecord(marks[i], st) : hipSuccess;
  };
  if (classes) {
    hipLaunchKernelGGL(k_merge_lane<true>, dim3(grid), dim3(256), 0, st, g, ctrl, L.lane_list);
    STC_TRY(mark(2));
    hipLaunchKernelGGL(k_merge_wave<true>, dim3(grid), dim3(256), 0, st, g, ctrl, L.wave_list);
    STC_TRY(mark(3));
    hipLaunchKernelGGL(k_hash_wave<true>, dim3(grid), dim3(256), 0, st, g, ctrl, dp, L.vw);
    STC_TRY(mark(4));
    hipLaunchKernelGGL(k_hash<true>, dim3(hash_grid), dim3(kVertexBlock), hash_lds, st, g, ctrl,
                       dp, L.vh);
    STC_TRY(mark(5));
    hipLaunchKernelGGL((k_bitmap_wave<true, false>), dim3(bw_grid), dim3(256), bw_lds, st, g, ctrl,
                       dp, L.vwb, nullptr, bw_words, 0u);
    STC_TRY(mark(6));
    hipLaunchKernelGGL((k_bitmap<true>), dim3(bm_grid), dim3(kBitmapBlock), bm_lds, st, g, ctrl,
                       dp, L.vb, bm_words);
    STC_TRY(mark(7));
    hipLaunchKernelGGL((k_bitmap_wave<true, true>), dim3(bw_grid), dim3(256), bw_lds, st, g, ctrl,
                       dp, L.vwp, L.vwpt, bw_words, 1u);
    STC_TRY(mark(8));
    STC_TRY(launch_two_phase(g, hp, dp, cb, L, nullptr, (unsigned long long *)ctrl->t, num_cus,
                             st, 5u, 6u, true, true));
    STC_TRY(mark(9));
  } else {
    hipLaunchKernelGGL(k_merge_lane<false>, dim3(grid), dim3(256), 0, st, g, ctrl, L.lane_list);
    STC_TRY(mark(2));
    hipLaunchKernelGGL(k_merge_wave<false>, dim3(grid), dim3(256), 0, st, g, ctrl, L.wave_list);
    STC_TRY(mark(3));
    hipLaunchKernelGGL(k_hash_wave<false>, dim3(grid), dim3(256), 0, st, g, ctrl, dp, L.vw);
    STC_TRY(mark(4));
    hipLaunchKernelGGL(k_hash<false>, dim3(hash_grid), dim3(kVertexBlock), hash_lds, st, g, ctrl,
                       dp, L.vh);
    STC_TRY(mark(5));
    hipLaunchKernelGGL((k_bitmap_wave<false, false>), dim3(bw_grid), dim3(256), bw_lds, st, g, ctrl,
                       dp, L.vwb, nullptr, bw_words, 0u);
    STC_TRY(mark(6));
    hipLaunchKernelGGL((k_bitmap<false>), dim3(bm_grid), dim3(kBitmapBlock), bm_lds, st, g, ctrl,
                       dp, L.vb, bm_words);
    STC_TRY(mark(7));
    // pull: with workgroup chunks, both lists in one two-phase launch (a batch
    // of one graph; its time is the workgroup slot's, the wave slot's about
    // 0); without, the wave kernel alone (the two-phase launch exits at once)
    hipLaunchKernelGGL((k_bitmap_wave<false, true>), dim3(bw_grid), dim3(256), bw_lds, st, g, ctrl,
                       dp, L.vwp, L.vwpt, bw_words, 1u);
    STC_TRY(mark(8));
    STC_TRY(launch_two_phase(g, hp, dp, cb, L, nullptr, (unsigned long long *)ctrl->t, num_cus,
                             st, 5u, 6u, true, false));
    STC_TRY(mark(9));
  }
  return hipGetLastError();
}

hipError_t enqueue_stats(const DevCsr &g, const DispatchParams *dp, CountBuffers &cb,
                         int num_cus, bool batch, hipStream_t st) {
  hipLaunchKernelGGL(k_edge_stats, dim3(grid_for(g.m, 256, num_cus * 8)), dim3(256), 0, st, g,
                     cb.gcount.as<Ctrl>(), dp, batch ? 1u : 0u);
  STC_TRY(hipGetLastError());
  // the vertex lists of the run just counted (the buffers binning filled)
  BinOut L;
  bin_lists(bin_layout(g.n, g.m, num_cus), g.m, cb.lists.as<uint32_t>(), cb.vlist.as<uint32_t>(),
            L);
  hipLaunchKernelGGL(k_list_stats, dim3(num_cus * 4), dim3(256), 0, st, g, cb.gcount.as<Ctrl>(),
                     dp, L.vh, L.vstride, batch ? 1u : 0u);
  return hipGetLastError();
}

hipError_t enqueue_pv_to_ids(const unsigned long long *pv_rank, const uint32_t *perm, uint32_t n,
                             unsigned long long *out, int num_cus, hipStream_t st) {
  hipLaunchKernelGGL(k_pv_to_ids, dim3(grid_for(n, 256, num_cus * 8)), dim3(256), 0, st, pv_rank,
                     perm, n, out);
  return hipGetLastError();
}

// Workgroups of `threads` threads and `lds` dynamic LDS bytes that one CU
// holds at once, cached per kernel and size (the runtime query is made once,
// before any stream capture of the batch, whose eager run comes first).
hipError_t resident_blocks(int *out, const void *kern, int threads, size_t lds) {
  struct Entry {
    const void *kern;
    size_t lds;
    int blocks;
  };
  static std::mutex mu;
  static std::vector<Entry> cache;
  std::lock_guard<std::mutex> lock(mu);
  for (const Entry &e : cache)
    if (e.kern == kern && e.lds == lds) {
      *out = e.blocks;
      return hipSuccess;
    }
  int b = 0;
  STC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, threads, lds));
  cache.push_back(Entry{kern, lds, b});
  *out = b;
  return hipSuccess;
}

// Words of the batch kernel's shared LDS region (the largest structure).
static size_t batch_region_words(const TuneParams &t) {
  const size_t bitmap = (size_t)((1u << t.bitmap_span_log2) / 16u) + kGuard + kLowCap;
  const size_t hash = (size_t)4u << t.hash_max_log2;
  // phase 1: the bitmap or hash table, then the pop's pull-chunk headers
  const size_t phase1 = std::max(bitmap, hash) + kBlockPop * kHdrFields;
  // phase 2: one slot per wave (hash table or window + low array, then heads)
  const size_t wave_slot = std::max<size_t>(kWaveCap, (1u << t.wave_window_log2) / 16u) +
                           kGuard + kWaveMaxDeg + kPackWords;
  const size_t waves = (size_t)kWavesPerBatchBlock * wave_slot;
  return std::max(phase1, waves);
}

size_t batch_lds_bytes(const TuneParams &t) { return 4u * batch_region_words(t); }

hipError_t launch_two_phase(const DevCsr &g, const DispatchParams &hp, const DispatchParams *dp,
                            CountBuffers &cb, const BinOut &L, const uint32_t *graph_of,
                            unsigned long long *per_graph, int num_cus, hipStream_t st,
                            uint32_t q1, uint32_t q2, bool one_graph, bool classes) {
  // four-class counting is one graph's (its pull items); batches are two-class
  if (classes && !one_graph) return hipErrorInvalidValue;
  const size_t lds = batch_lds_bytes(hp.tune);
  auto launch = [&](auto kern) -> hipError_t {
    // The persistent grid is as many workgroups as fit on the CUs together
    // (at most 4 per CU): the runtime's count, which rounds each
    // workgroup's LDS (the region plus the static arrays) to the hardware's
    // allocation unit.  A count from the byte sizes alone once launched a
    // fourth workgroup per CU that waited for a third to finish.
    int blocks_per_cu = 0;
    STC_TRY(resident_blocks(&blocks_per_cu, (const void *)kern, kBatchBlock, lds));
    blocks_per_cu = std::max(1, std::min(4, blocks_per_cu));
    hipLaunchKernelGGL(kern, dim3(num_cus * blocks_per_cu), dim3(kBatchBlock), lds, st, g,
                       cb.gcount.as<Ctrl>(), dp, L.lane_list, L.wave_list, L.vh, L.vstride,
                       graph_of, per_graph, (1u << hp.tune.bitmap_span_log2) / 16u,
                       4u << hp.tune.hash_max_log2, (1u << hp.tune.wave_window_log2) / 16u,
                       (uint32_t)batch_region_words(hp.tune), q1, q2);
    return hipGetLastError();
  };
  if (one_graph)
    return classes ? launch(k_two_phase<true, kKindPullOne>)
                   : launch(k_two_phase<false, kKindPullOne>);
  // a batch: the all-kinds kernel when it has merge, hash or bitmap items,
  // else the pull-only kernel (each returns at once when it is not the one)
  STC_TRY(launch(k_two_phase<false, kKindAll>));
  return launch(k_two_phase<false, kKindPullBatch>);
}

hipError_t enqueue_batch(const DevCsr &g, const DispatchParams &hp, const DispatchParams *dp,
                         CountBuffers &cb, const uint32_t *graph_of,
                         unsigned long long *per_graph, uint32_t k, int num_cus,
                         hipStream_t st, const hipEvent_t *marks) {
  BinOut L;
  STC_TRY(enqueue_binning(g, dp, cb, num_cus, st, marks, L));
  hipLaunchKernelGGL(k_fill_words, dim3(grid_for(4ull * k, 256, 4096)), dim3(256), 0, st,
                     (uint32_t *)per_graph, 4ull * k, 0u);
  STC_TRY(hipGetLastError());
  // one launch counts every item: its time is the last kernel slot's
  for (int i = 2; marks && i < kMarks - 1; i++) STC_TRY(hipEventRecord(marks[i], st));
  STC_TRY(launch_two_phase(g, hp, dp, cb, L, graph_of, per_graph, num_cus, st, 3u, 4u, false,
                           false));
  if (marks) STC_TRY(hipEventRecord(marks[kMarks - 1], st));
  return hipSuccess;
}

hipError_t read_count(CountBuffers &cb, bool classes, CountResult &r, hipStream_t st) {
  Ctrl h;
  STC_TRY(hipMemcpyAsync(&h, cb.gcount.p, offsetof(Ctrl, bucket_part), hipMemcpyDeviceToHost, st));
  STC_TRY(hipStreamSynchronize(st));
  for (int k = 0; k < 4; k++) {
    r.t[k] = 0;
    for (int s = 0; s < kTSlots; s++) r.t[k] += h.t[s][k];
  }
  (void)classes;
  for (int k = 0; k < 10; k++) {
    r.kern_edges[k] = h.kern_edges[k];
    r.kern_words[k] = h.kern_words[k];
    r.kern_bytes[k] = h.kern_bytes[k];
  }
  for (int k = 0; k < kBuckets; k++) {
    r.table[k] = h.table[k];
    r.bucket_cnt[k] = h.bucket_cnt[k];
  }
  return hipSuccess;
}

}  // namespace stc
