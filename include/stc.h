/*
 * stc.h -- C ABI of the MI355X signed triangle counter.
 *
 * Drop-in boundary for the reference's counting path.  The reference offers it
 * two ways, both replaced by the calls below:
 *   (i)  process CLI "<binary> <graph_file>" printing Balanced / Unbalanced /
 *        Total (Signed_TC/Multicore Implementations/tc_merge_sign_enu.cpp:46-154,
 *        tc_hash_sign_enu.cpp:76-177) -- the `stc` program built on this ABI;
 *   (ii) in-process GPU operator
 *        float gunrock::signed_tc::run(G, param_t&, result_t&, context)
 *        with result_t{size_t* total_balanced; size_t* total_unbalanced;}
 *        (Signed_TC/GPU Implementations/GlobalMEM/signed_tc.hxx:17-24,527-542)
 *        -- stc_count() below (and stc_run(), its literal shape).
 *
 * Conventions: return 0 (STC_OK) on success or a negative STC_E_* code; the
 * library never exits the process; a context is used by one host thread at a
 * time; every call is synchronous with respect to its outputs.  Input pointers
 * are read-only during the call and owned by the caller.
 *
 * Counting semantics (identical to the reference's OpenMP merge/hash path):
 * lines are canonicalised to (min(u,v), max(u,v)); u == v lines are dropped;
 * duplicate lines of one undirected edge collapse to one edge whose sign is the
 * sign of its first line; a triangle is balanced iff it has an even number of
 * negative edges (reference: tc_merge_sign_enu.cpp:29-33, tc_serial.cpp:95-112).
 */
#ifndef STC_H
#define STC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef STC_OK
#define STC_OK 0
#define STC_E_ARG (-1)       /* bad argument, or an id >= n */
#define STC_E_TOO_LARGE (-2) /* n > 2^31 - 2^20, or > 2^32 - 2^20 input lines */
#define STC_E_OOM (-3)       /* device allocation failed */
#define STC_E_HIP (-4)       /* HIP runtime error */
#define STC_E_MODEL (-5)     /* dispatcher model missing/invalid: fixed used */
#define STC_E_IO (-6)        /* file could not be read/written */
#define STC_E_NODEV (-7)     /* no usable GPU */
#endif

typedef struct stc_ctx stc_ctx;

/* A signed edge list: m lines (u[i], v[i], sign[i]) with sign -1 = negative
 * and anything else positive; every id must be < n.  For stc_count* these are
 * host pointers; for stc_count_device* they are device pointers. */
typedef struct stc_edges {
  uint32_t n;
  uint64_t m;
  const uint32_t *u;
  const uint32_t *v;
  const int8_t *sign;
} stc_edges;

typedef struct stc_counts {
  uint64_t balanced;
  uint64_t unbalanced;
} stc_counts;

/* Four classes by number of negative edges (reference tc_serial.cpp:11-31):
 * t3 = +++ , t1 = +-- (balanced); t2 = ++- , t0 = --- (unbalanced). */
typedef struct stc_classes {
  uint64_t t0, t1, t2, t3;
} stc_classes;

typedef enum stc_dispatch {
  STC_LEARNED = 0,      /* MLP per edge bucket; fixed rule if no model */
  STC_FIXED = 1,        /* degree-threshold rule */
  STC_FORCE_MERGE = 2,  /* every edge through the merge kernels */
  STC_FORCE_HASH = 3,   /* every edge through the LDS hash kernel */
  STC_FORCE_BITMAP = 4, /* every edge through the hub bitmap kernel (edges
                           whose source list does not fit fall back to hash) */
  STC_FORCE_PULL = 5    /* every edge through the pull kernels (target's list
                           staged, the source's remaining list streamed; a
                           target list of more than 1025 entries, or a batch,
                           falls back to bitmap) */
} stc_dispatch;

typedef struct stc_opts {
  int dispatch;        /* stc_dispatch */
  int repeat;          /* run the counting kernels this many times (>= 1);
                          counts come from the last run */
  int classes;         /* nonzero: also produce the four-class counts */
} stc_opts;

/* Timings (hipEvent, ms) and sizes of the last counting call. */
typedef struct stc_stats {
  uint32_t n;              /* vertex-id bound */
  uint64_t m_in;           /* input lines */
  uint64_t m_oriented;     /* unique non-loop undirected edges = oriented edges */
  uint32_t max_out_degree;
  double b_alg;            /* sum over oriented edges of 4(d+(u)+d+(v)),
                              + 4 m+ + 8 (n+1) bytes */
  float ms_upload;         /* host -> device copy of the edge list (staged copies
                              also canonicalise each chunk on the device as it
                              arrives) */
  float ms_build;          /* on-device CSR build + orientation, from the end of
                              the upload */
  float ms_count;          /* binning + intersection, median over repeats */
  float ms_count_min;
  float ms_bin;            /* binning part of the last run */
  uint64_t edges_merge;    /* oriented edges per strategy in the last run */
  uint64_t edges_hash;
  uint64_t edges_bitmap;
  uint64_t edges_pull;
  uint64_t edges_skipped;  /* edges that cannot close a triangle */
  int dispatch_used;       /* stc_dispatch actually applied */
  float ms_kernel[8];      /* merge-lane, merge-wave, wave-hash, workgroup-hash,
                              wave-bitmap, workgroup-bitmap, wave-pull and
                              workgroup-pull kernel time (hipEvent on the
                              library stream), median over repeats; a
                              count with workgroup-pull edges runs both pull
                              slots in one launch, reported in the last slot
                              (wave-pull then reads about 0); a batch counts
                              every item in one launch, reported in the last
                              slot (timed on a warm eager run when repeat > 1) */
  double bytes_kernel[8];  /* merge-model bytes 4(d+(u)+d+(v)) of the
                              oriented edges each of those kernels counted
                              (the same pull folding) */
  double alg_bytes_kernel[8]; /* bytes each kernel's strategy must read: the
                              staged lists once per item, the streamed list
                              of every counted edge (pull: the 4 da bytes of
                              the source's suffix; hash and push bitmap: the
                              4 d+(v) bytes of the target's list; merge: both),
                              and the metadata of every edge and item visited
                              (the same pull folding) */
} stc_stats;

/* Create a context on HIP device `device`.  model_dir (nullable) holds the
 * dispatcher model "dispatch_mlp.txt" and the autotuner table "tune.txt";
 * NULL uses the files shipped beside the library.  A missing or invalid model
 * is not an error: the context falls back to the fixed rule and
 * stc_model_status() reports STC_E_MODEL. */
int stc_create(int device, const char *model_dir, stc_ctx **out);
void stc_destroy(stc_ctx *ctx);
const char *stc_last_error(const stc_ctx *ctx);
int stc_model_status(const stc_ctx *ctx);

/* Count the signed triangles of one graph given as a host edge list. */
int stc_count(stc_ctx *ctx, const stc_edges *g, const stc_opts *opts,
              stc_counts *out);
/* Same, with u/v/sign already in device memory. */
int stc_count_device(stc_ctx *ctx, const stc_edges *g_dev,
                     const stc_opts *opts, stc_counts *out);
/* Four-class counts of the graph counted last (requires opts.classes). */
int stc_last_classes(const stc_ctx *ctx, stc_classes *out);

/* Two-phase form: stc_prepare uploads and builds the oriented CSR once;
 * stc_count_prepared runs only binning + intersection on it (the timed
 * counting phase). */
int stc_prepare(stc_ctx *ctx, const stc_edges *g, int g_is_device);
int stc_count_prepared(stc_ctx *ctx, const stc_opts *opts, stc_counts *out);

/* stc_count_prepared without any device allocation, for callers that cannot
 * allocate device memory (code under HIP stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * counting pass's buffers (control block, edge and vertex lists, scan
 * levels) come from one static device array of 128 MiB that the library
 * holds on every GPU it runs on until the process ends; that is enough for
 * graphs of about 4 M oriented edges (the 960 x 960 grid needs 72 MB).  A larger graph returns STC_E_TOO_LARGE.  Same counts as
 * stc_count_prepared in every mode, classes included (stc_last_classes);
 * per-kernel edge and byte statistics read zero.  Every call on one GPU
 * shares the array and keeps what the last call left there: calls that
 * overlap on one GPU (two host threads, two contexts) race. */
int stc_count_prepared_static(stc_ctx *ctx, const stc_opts *opts, stc_counts *out);

/* Per-vertex four-class counts of the prepared graph (stc_prepare): cnt has
 * 4 n entries, and cnt[4 v + k] counts the T_k triangles (k = 0..3: t0..t3 of
 * stc_classes) whose lowest-ranked vertex in the (degree, id) orientation is
 * vertex v.  This is the per-vertex array of the reference's four-class
 * program (tc_serial.cpp:141-158), which that program only sums.  out
 * (nullable) receives the balanced / unbalanced totals; the classes are
 * available from stc_last_classes. */
int stc_count_per_vertex(stc_ctx *ctx, const stc_opts *opts, uint64_t *cnt, stc_counts *out);

/* Count k independent graphs in one persistent launch; out has k entries.
 * Equivalent to stc_prepare_batch followed by stc_count_batch_prepared.  The
 * reference counts one graph per program run; this has no counterpart there. */
int stc_count_batch(stc_ctx *ctx, const stc_edges *graphs, uint32_t k,
                    const stc_opts *opts, stc_counts *out);
/* Two-phase batch: upload the k host edge lists and build their union CSR
 * once; then count it (eager run, hipGraph capture, opts.repeat replays)
 * with out receiving the k per-graph counts of the prepared batch. */
int stc_prepare_batch(stc_ctx *ctx, const stc_edges *graphs, uint32_t k);
int stc_count_batch_prepared(stc_ctx *ctx, const stc_opts *opts, stc_counts *out);

/* Statistics of the last counting call.  The per-kernel edge and byte counts
 * are computed by one extra device pass on the first call after a count. */
int stc_get_stats(stc_ctx *ctx, stc_stats *out);

/* Reference operator shape (signed_tc.hxx:527-542): counts one graph and
 * writes the two totals; returns elapsed milliseconds of the counting phase,
 * or a negative value on error. */
float stc_run(stc_ctx *ctx, const stc_edges *g, uint64_t *total_balanced,
              uint64_t *total_unbalanced);

#ifdef __cplusplus
}
#endif
#endif /* STC_H */
