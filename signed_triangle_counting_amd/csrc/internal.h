// Internal interface between the host orchestration (api.cpp) and the HIP
// kernels (csr_build.hip, count.hip, mlp.hip, batch.hip).  Not part of the
// public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace stc {

// ---- strategies -----------------------------------------------------------
// Per edge bucket, the dispatcher picks one of these.  Merge runs edge-centric
// (one lane, or one wave64, per oriented edge); hash and bitmap run
// vertex-grouped (one workgroup stages the source's out-list in LDS and
// streams the targets' out-lists against it).
enum Strategy : uint8_t {
  S_SKIP = 0,
  S_MERGE_LANE = 1,  // one lane per edge, two-pointer merge
  S_MERGE_WAVE = 2,  // one wave64 per edge, LDS-tiled cooperative merge
  S_HASH = 3,        // LDS hash table of the source list
  S_BITMAP = 4,      // LDS presence/sign bitmap over a rank window at the top
                     // of the source list (lower part kept as a sorted array)
  S_PULL = 5,        // the same bitmap of the target's list N+(v), probed by the
                     // part of N+(u) after v (grouped by target vertex)
  S_COUNT = 6
};

// Bucket of an oriented edge (u->v): (ceil_log2 of the usable part of
// N+(u), ceil_log2 d+(v)), each clamped to 15.  The usable part of N+(u) is
// the suffix after v: only out-neighbours ranked above v can close a
// triangle with u->v.
constexpr int kBucketSide = 16;
constexpr int kBuckets = kBucketSide * kBucketSide;

// Dispatcher inputs/outputs resident in device memory.
constexpr int kMlpIn = 7;  // la, lb, la - lb, log2 count, log2 n, log2 m, log2 graphs
constexpr int kMlpHidden = 16;
constexpr int kMlpOut = 4;  // merge, hash, bitmap, pull
struct MlpModel {
  float mean[kMlpIn];
  float inv_std[kMlpIn];
  float w1[kMlpHidden][kMlpIn];
  float b1[kMlpHidden];
  float w2[kMlpOut][kMlpHidden];
  float b2[kMlpOut];
};

enum DispatchMode : int {
  D_LEARNED = 0,
  D_FIXED = 1,
  D_FORCE_MERGE = 2,
  D_FORCE_HASH = 3,
  D_FORCE_BITMAP = 4,
  D_FORCE_PULL = 5,
  D_TABLE = 6  // use the strategy table given by the host (training/timing)
};

struct TuneParams {
  // merge lane class: buckets whose longer side is <= 2^wave_min_log2 use one
  // lane per edge, else one wave per edge
  int8_t merge_wave_min_log2[kBucketSide];  // indexed by log2 of shorter side
  int hash_max_log2;      // hash only if the source list has <= 2^x entries
  int bitmap_span_log2;   // rank window of the workgroup bitmap (entries)
  int wave_window_log2;   // rank window of the per-wave bitmap (entries)
  int hash_block;         // threads per workgroup for hash/bitmap
  // The fixed rule, on the bucket's la = ceil_log2 |N+(u) after v| and
  // lb = ceil_log2 d+(v), both <= 15 (so 16 switches a branch off):
  int fixed_merge_log2;   // max(la, lb) <= x: merge (lane or wave); -1: off
  int fixed_pull_log2;    // lb - la >= x: pull (the suffix streamed); -15..16
  int fixed_skew_log2;    // |la - lb| >= x: wave merge (binary search of the longer)
  int fixed_hub_log2;     // la >= x: push bitmap (the long source list staged)
  // Item sizes of the vertex-grouped and batched kernels:
  int pull_block_chunk_log2;  // in-edges per workgroup-pull item (8..14)
  int pull_wave_chunk_log2;   // in-edges per wave-pull item (5..12)
  // pull targets with 64 < d+ <= 2^x entries whose list spans at most the
  // wave window are counted by waves too (6..9; 6: none), in single graphs
  // and in batches of small graphs
  int pull_wave_max_log2;
  int batch_pull_wave_max_log2;
  int wave_pop;               // items a wave takes per queue pop (1..8)
  int lane_item_log2;         // merge-lane edges per batched workgroup item (9..12)
  int allow_pull;         // 1: pull buckets are counted by the pull kernels (single
                          // graphs and batches); 0: they fall back to bitmap
};

struct DispatchParams {
  int mode;
  int have_model;
  MlpModel model;
  TuneParams tune;
  uint8_t host_table[kBuckets];  // D_TABLE
  float graph_feat[3];           // log2 (n/k + 1), log2 (m/k + 1), log2 k for k graphs
                                 // counted together (k = 1: one graph)
  // 1 when every target list of class 11 (1,024 < d+ <= 2,048 entries) spans
  // at most the workgroup bitmap window, so a pull bucket of class <= 11 never
  // demotes an edge (from the prepared graph's largest such span)
  uint32_t pull_c11_fits;
  // the wave-pull limit of this count (pull_wave_max_log2, or its batch value)
  uint32_t wave_pull_max_log2;
};

// One in-edge e = (u -> v) of the transposed index: the part of N+(u) after
// v is col[pos, pos + len) with pos = e + 1.  The word w holds the edge's sign
// in bit 0 and len in the bits above.
struct InEdge {
  uint32_t pos, w;
};
__host__ __device__ inline uint32_t in_len(uint32_t w) { return w >> 1; }
// Edge bucket of an oriented edge that can close no triangle (an empty part
// of N+(u) after v, or an empty N+(v)); every other edge's bucket is
// 16 min(ceil_log2 |N+(u) after v|, 15) + min(ceil_log2 d+(v), 15), 0..255.
constexpr uint16_t kNoEdgeBucket = 0xFFFF;

// Device view of the oriented CSR (rank space).
struct DevCsr {
  uint32_t n = 0;
  uint64_t m = 0;
  const uint32_t *off = nullptr;  // n + 1
  const uint32_t *col = nullptr;  // m: (dst_rank << 1) | neg
  const uint32_t *src = nullptr;  // m: src rank
  const uint32_t *tin_off = nullptr;  // n + 1: in-edges of each target rank
  const InEdge *tin = nullptr;        // m: in-edges grouped by target, ascending e
  const uint8_t *dcls = nullptr;      // n: min(ceil_log2 d+(v), 15), 0xFF when d+(v) = 0
  // m rounded up to 8: each oriented edge's binning bucket, in edge order
  // (kNoEdgeBucket past m)
  const uint16_t *ebk = nullptr;
  // n: bit c set when some in-edge (u -> v) has a non-empty part of N+(u)
  // after v, of length len with min(ceil_log2 len, 15) = c
  const uint32_t *smask = nullptr;
  // Per-vertex four-class counters by rank (4 n words: T0..T3), credited to
  // each triangle's lowest-ranked vertex by the four-class kernels; nullptr
  // when per-vertex counts are not requested.
  unsigned long long *pv = nullptr;
};

// ---- host -> device upload through pinned staging ----------------------------
// One graph's edge arrays (host memory, m lines).
struct EdgePiece {
  const uint32_t *u, *v;
  const int8_t *s;
  uint64_t m;
};

// Two pinned staging buffers of kStageLines lines each (u, then v, then the
// sign bytes) and their copy-done events, kept by a context after first use.
constexpr uint64_t kStageLines = 1ull << 23;     // 72 MiB per buffer
constexpr uint64_t kStageMinLines = 1ull << 20;  // smaller inputs are copied directly
struct UploadRing {
  void *host[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  ~UploadRing() { release(); }
  hipError_t ensure();
  void release();
  // Copy the pieces back to back into du / dv / ds (device) on stream cs and
  // wait for the copies to finish.  After each chunk's copies, stream ks waits
  // for them and on_chunk(l0, l1) (if set) enqueues work on ks for lines
  // [l0, l1), which then overlaps the copies of the next chunks.
  using ChunkFn = std::function<hipError_t(uint64_t, uint64_t)>;
  hipError_t upload(const std::vector<EdgePiece> &pieces, uint32_t *du, uint32_t *dv, int8_t *ds,
                    hipStream_t cs, hipStream_t ks = nullptr, const ChunkFn &on_chunk = nullptr);
};

// ---- device memory --------------------------------------------------------
// A device buffer the context owns (grown by hipMalloc), or a borrowed view
// of `bytes` bytes of memory someone else owns (the static scratch), which
// never allocates: ensure() then fails when the view is too small.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;
  ~DevBuf() { release(); }
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
  }
  void borrow(void *q, size_t n) {
    release();
    p = q;
    bytes = n;
    borrowed = true;
  }
  // Grow to at least `need` bytes (contents not preserved).
  hipError_t ensure(size_t need) {
    if (need == 0) need = 256;
    if (need <= bytes) return hipSuccess;
    if (borrowed) return hipErrorInvalidValue;
    release();
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return e;
    }
    bytes = need;
    return hipSuccess;
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
};

// Scratch for scans and sorts.
struct ScanScratch {
  DevBuf level[4];
};

// ---- launchers (return hipSuccess or the first error) -------------------
hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t n,
                              ScanScratch &scr, hipStream_t s);
// Bytes of ScanScratch level l that a scan of n words needs (0: level unused).
size_t scan_level_bytes(uint64_t n, int l);
// Stable LSD radix sort of n < 2^32 64-bit keys on bits [lo_bit, lo_bit +
// nbits), 8-bit digits; scratch holds the per-tile (4096 keys) digit counts.
// On return *keys points at the sorted data (*tmp at the other buffer).
hipError_t radix_sort_u64(uint64_t **keys, uint64_t **tmp, uint64_t n, int lo_bit, int nbits,
                          DevBuf &scratch, hipStream_t s, bool first_counted);
// Bytes of radix sort scratch for n keys; with first_counted the tile counts
// of the first digit are already in scratch (written by the keys' producer).
size_t radix_scratch_bytes(uint64_t n);

struct BuildBuffers {
  DevBuf in_u, in_v, in_s;
  DevBuf keyA, keyB, flag, pos;
  DevBuf deg, vkeyA, vkeyB, rank, perm, graph_of;
  DevBuf col, src, off, fdeg, tin, tin_off, dcls, smask, ebk;
  DevBuf misc, sort;
  ScanScratch scr;
};

// A device edge list to build from: m_in lines (u, v, sign) over ids < n, or
// for a disjoint union of k graphs (k >= 1) ebase / vbase (device, k + 1
// entries each): graph g owns lines [ebase[g], ebase[g+1]) with ids below
// vbase[g+1] - vbase[g], shifted by vbase[g] into the union's id space.
struct BuildInput {
  const uint32_t *u = nullptr, *v = nullptr;
  const int8_t *s = nullptr;
  uint64_t m_in = 0;
  uint32_t n = 0;
  const uint64_t *ebase = nullptr;
  const uint32_t *vbase = nullptr;
  uint32_t k = 0;
};
struct BuildResult {
  uint32_t n = 0;
  uint64_t m_in = 0, m = 0;
  uint32_t max_deg = 0, max_out = 0;
  uint32_t max_span_c11 = 0;  // largest rank span of an out-list of 1,025..2,048 entries
  double b_alg = 0;
};
// Builds the rank-space oriented CSR (and the transposed index) on stream st
// in three steps: build_begin sizes the buffers, build_canon turns lines
// [l0, l1) into canonical sort keys and counts their first digit (l0 a
// multiple of 4096; called once for all lines, or once per uploaded chunk as
// soon as its copies are done), build_finish does the rest.
hipError_t build_begin(const BuildInput &in, BuildBuffers &bb, std::string &err, hipStream_t st);
hipError_t build_canon(const BuildInput &in, uint64_t l0, uint64_t l1, BuildBuffers &bb,
                       hipStream_t st);
hipError_t build_finish(const BuildInput &in, BuildBuffers &bb, BuildResult &res,
                        std::string &err, hipStream_t st);

struct CountBuffers {
  DevBuf gcount, lists, blk, vflag, vlist, pv, pv_ids;
  ScanScratch scan;
};

struct CountResult {
  uint64_t t[4] = {0, 0, 0, 0};  // t0..t3 when classes, else t[0]=bal t[1]=unbal
  // per counting kernel (merge-lane, merge-wave, wave-hash, workgroup-hash,
  // wave-bitmap, workgroup-bitmap, wave-pull, workgroup-pull; slot 8: edges
  // a partial table leaves out, slot 9: skipped edges): oriented edges and
  // their sum of d+(u) + d+(v)
  uint64_t kern_edges[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t kern_words[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  // bytes each kernel's strategy reads: the staged lists, the streamed lists
  // and the metadata of the edges and items it visits
  uint64_t kern_bytes[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint8_t table[kBuckets];
  uint32_t bucket_cnt[kBuckets];
};

// Enqueue binning + counting on `st` (no host sync inside).  hp is the host
// copy of *dp (launch shapes come from its tune part).  marks (nullable, 8
// events) are recorded before binning, after binning and after each of the
// merge-lane, merge-wave, wave-hash, workgroup-hash, wave-bitmap,
// workgroup-bitmap, wave-pull and workgroup-pull kernels.
constexpr int kMarks = 10;
constexpr int kKernels = 8;
hipError_t enqueue_count(const DevCsr &g, const DispatchParams &hp,
                         const DispatchParams *dp, CountBuffers &cb, bool classes,
                         int num_cus, hipStream_t st, const hipEvent_t *marks);
// Per-kernel edge and word counts of the run just enqueued (not timed: enqueue
// after the run's closing event).
hipError_t enqueue_stats(const DevCsr &g, const DispatchParams *dp, CountBuffers &cb,
                         int num_cus, bool batch, hipStream_t st);
// Per-vertex counts: pv_rank (4 n, by rank) to out (4 n, by vertex id) through
// perm (rank -> vertex id).
hipError_t enqueue_pv_to_ids(const unsigned long long *pv_rank, const uint32_t *perm, uint32_t n,
                             unsigned long long *out, int num_cus, hipStream_t st);
// Read back totals and per-kernel sizes after the stream completes.
hipError_t read_count(CountBuffers &cb, bool classes, CountResult &r,
                      hipStream_t st);
size_t hash_lds_bytes(const TuneParams &t);
// The counting pass's buffers for a graph of n vertices and m oriented edges,
// carved from `bytes` bytes at `base` (256-byte aligned) as borrowed views;
// false when they do not fit.  The vertex-flag words are cleared on st (they
// must start at zero; the flag pass clears only what it set).
constexpr size_t kStaticScratchBytes = (size_t)128 << 20;
bool carve_count_buffers(void *base, size_t bytes, uint32_t n, uint64_t m, int num_cus,
                         CountBuffers &cb, size_t *need);
hipError_t clear_vflags(CountBuffers &cb, uint32_t n, hipStream_t st);
// Device address of this library's static scratch array (kStaticScratchBytes)
// on the current device.
hipError_t static_scratch_address(void **ptr);
size_t bitmap_lds_bytes(const TuneParams &t);
size_t wave_bitmap_lds_bytes(const TuneParams &t);

// Batched counting over the union CSR of k graphs: binning, then one
// persistent kernel; per_graph gets k (balanced, unbalanced) pairs.
// graph_of maps a rank to its graph.
// marks (nullable; never during stream capture): as enqueue_count's, with
// the single batched launch between the last two.
hipError_t enqueue_batch(const DevCsr &g, const DispatchParams &hp, const DispatchParams *dp,
                         CountBuffers &cb, const uint32_t *graph_of,
                         unsigned long long *per_graph, uint32_t k, int num_cus,
                         hipStream_t st, const hipEvent_t *marks);
size_t batch_lds_bytes(const TuneParams &t);
hipError_t init_kernel_attributes();

// MLP training on device (full-batch Adam in one wave64, fixed reduction
// order -> deterministic).  x: nsamp x kMlpIn normalised features; r: nsamp x
// kMlpOut costs normalised to [0, 1] per sample; w: per-sample weights;
// params_io: flat parameters (w1, b1, w2, b2), updated in place.
hipError_t train_mlp_device(const float *x, const float *r, const float *w, int nsamp,
                            int iters, float lr, float *params_io, hipStream_t st);
int mlp_param_count();
void mlp_unpack(const float *params, MlpModel &m);

}  // namespace stc
