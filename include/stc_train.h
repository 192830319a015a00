/*
 * stc_train.h -- dispatcher training and tuning hooks of the signed triangle
 * counter.  No counterpart in the reference: its only dispatch is a fixed
 * 4x length-skew rule (Signed_TC/GPU Implementations/GlobalMEM/
 * signed_tc.hxx:48,65).  These calls let the counter time its own strategies
 * per edge bucket, train the dispatcher MLP on those timings with a HIP
 * kernel, and apply autotuned launch parameters.  A table or model can only
 * change which kernel counts an edge, never the count.
 *
 * Buckets: oriented edge (u->v) falls in bucket 16*a + b with
 * a = min(15, ceil(log2(#out-neighbours of u ranked above v))),
 * b = min(15, ceil(log2 d+(v))); edges with either part empty cannot close a
 * triangle and are skipped.
 */
#ifndef STC_TRAIN_H
#define STC_TRAIN_H

#include <stdint.h>

#include "stc.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  STC_TABLE_SKIP = 0,
  STC_TABLE_MERGE = 1,
  STC_TABLE_HASH = 2,
  STC_TABLE_BITMAP = 3,
  STC_TABLE_PULL = 4
};
#define STC_BUCKETS 256
#define STC_MLP_FEATURES 7

/* Oriented edges per bucket of the prepared graph, or of the prepared batch's
 * union graph (runs the binning pass). */
int stc_bucket_histogram(stc_ctx *ctx, uint32_t out[STC_BUCKETS]);

/* Count the prepared graph (or the prepared batch, through the batched
 * kernel, with the per-graph counts summed) with an explicit strategy per
 * bucket (STC_TABLE_*); buckets set to STC_TABLE_SKIP are not counted, so the
 * result is partial unless every non-empty bucket has a strategy. */
int stc_count_with_table(stc_ctx *ctx, const uint8_t table[STC_BUCKETS], int repeat,
                         stc_counts *out);

/* Dispatcher features of bucket `bucket` holding `count` edges when k graphs
 * (k = 1: one graph) with n vertices and m oriented edges in total are
 * counted together: la, lb, la - lb, log2(count + 1), log2(n/k + 1),
 * log2(m/k + 1), log2 k. */
void stc_bucket_features(int bucket, uint32_t count, uint32_t n, uint64_t m, uint32_t k,
                         float out[STC_MLP_FEATURES]);

/* Train the dispatcher MLP (7 -> 16 ReLU -> 4) on the device by full-batch
 * Adam over nsamp samples: feat (nsamp x 7) and cost (nsamp x 4, time of the
 * whole count with the sample's bucket under the merge / hash / bitmap / pull
 * strategy).  The loss is the softmax-expected normalised cost, each sample
 * weighted by its relative time at stake (max - min) / min, clamped at 1.
 * Deterministic for a given seed.  Writes the model file (text, checksummed)
 * to out_path and loads it into the context. */
int stc_train_dispatcher(stc_ctx *ctx, const float *feat, const float *cost, int nsamp,
                         int iters, float lr, uint64_t seed, const char *out_path);

/* The strategy per bucket (STC_TABLE_*) that the last count's dispatcher
 * chose on the device -- the fixed rule, the MLP's argmax, a forced mode or a
 * given table -- before the per-edge demotions of lists that do not fit the
 * LDS structures; STC_TABLE_SKIP for empty buckets.  No counterpart in the
 * reference (its 4x rule is decided per edge inside the kernel,
 * signed_tc.hxx:47,65). */
int stc_last_table(const stc_ctx *ctx, uint8_t out[STC_BUCKETS]);

/* The prepared graph's (or batch's) device CSR in rank space, copied to the
 * host (every pointer nullable; n and m from stc_get_stats): off[n + 1] the
 * out-list offsets by rank; col[m] the lists, (rank << 1) | neg ascending by
 * rank; perm[n] the vertex id of each rank; tin_off[n + 1] and tin[2 m] the
 * transposed index, per target rank its in-edges e = (u -> v) in ascending
 * e as (e + 1, (|N+(u) after v| << 1) | neg) pairs; dcls[n] each rank's
 * out-degree class min(ceil_log2 d+, 15) (0xFF for an empty list); smask[n]
 * the suffix classes of its in-edges.  Replaces, for tests, the reference's
 * ForwardCSR (Multicore Implementations/graph_io.h:22-30,221-402), which
 * orients by the same (degree, id) rank but keeps vertex ids. */
int stc_get_csr(stc_ctx *ctx, uint32_t *off, uint32_t *col, uint32_t *perm, uint32_t *tin_off,
                uint32_t *tin, uint8_t *dcls, uint32_t *smask);

/* Reload "dispatch_mlp.txt" / "tune.txt" from dir (NULL: the shipped ones).
 * A directory without a valid tune.txt keeps the context's tune values. */
int stc_load_models(stc_ctx *ctx, const char *dir);

/* Apply tune parameters given in the tune.txt text format. */
int stc_set_tune(stc_ctx *ctx, const char *text);

#ifdef __cplusplus
}
#endif
#endif /* STC_TRAIN_H */
