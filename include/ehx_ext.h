/*
 * ehx_ext.h -- libehx entry points with no counterpart in the reference
 * engine: the tuned form's knobs (set by the autotuner) and the learned layout
 * selector (loaded after training).  Same conventions as ehx.h.
 */
#ifndef EHX_EXT_H
#define EHX_EXT_H

#include <stdint.h>

#include "ehx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Knobs of the degree-binned persistent last level. */
typedef struct {
  int32_t lds_slots;    /* 32-bit LDS slots per workgroup for the staged sets (a quarter per wavefront): 2048, 4096, 8192 */
  int32_t big_deg;      /* child sets at least this long are streamed by the whole wavefront */
  int32_t waves_per_cu; /* persistent grid = CUs x waves_per_cu / 4 four-wave workgroups */
  int32_t grab;         /* queue entries taken per atomic */
  int32_t bitmap_all;   /* stage every set whose word window fits as a bitmap */
  int32_t use_graph;    /* capture the level in a hipGraph and replay it */
  int32_t big_work;     /* prefixes with at least this much estimated work are staged once per workgroup */
  int32_t stream_lanes; /* lanes streaming one long child set together: a power of two in [2, 64] */
  int32_t gap_coded;    /* 1: stream the 16-bit gap-coded copy of a long child list where one exists (plans whose last level has one operand besides x) */
  int32_t split_work;   /* split staged-per-workgroup prefixes into parts of about this much work; 0: never */
  int32_t tiny_work;    /* prefixes below this estimated work (single-operand plans) run up to 64 per wavefront; 0: never */
} ehx_tune_params;

int ehx_tune_get(ehx_tune_params *out);
int ehx_tune_set(const ehx_tune_params *in);

/* Learned layout selector: an 8 -> 32 -> 32 -> 2 MLP, FP16 weights (IEEE
 * half bit patterns), FP32 accumulation.  norm holds 8 feature means then 8
 * inverse standard deviations.  max_words bounds the bitset word window the
 * selector may choose (the staging capacity it was trained for).  Used by
 * trie builds with EHX_POLICY_LEARNED on `device`. */
int ehx_selector_num_params(void);
int ehx_selector_load(int device, const uint16_t *params, uint64_t nparams, const float *norm, float max_words);
int ehx_selector_clear(void);

/* Per-set features of the selector for every child set of a trie (num_heads
 * rows of 8 floats), computed by the device code the trie build evaluates. */
int ehx_trie_set_features(ehx_ctx *ctx, const ehx_trie *t, float *out);

/* Engine-recorded timings: runs the tuned form of `plan` once and writes the
 * shader-clock cycles spent on each prefix (cap entries at most) with its
 * staged set forced to a bitmap (force_mode 0), a hash table (1), or chosen
 * by layout (-1); *nprefix receives the number of prefixes. */
int ehx_join_record(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int force_mode, uint64_t *cycles,
                    uint64_t cap, uint64_t *nprefix, int64_t *result);

/* Engine-recorded timings of the two uint x uint algorithms, for the learned
 * per-pair choice: pair i is child set nodes0[i] of trie a and child set
 * nodes1[i] of trie b (a node equal to the trie's num_heads names its head
 * set).  cycles[2i] receives the shader-clock cycles of counting their
 * intersection by merge path, cycles[2i+1] by galloping the smaller set into
 * the larger, each the fewest of `reps` runs (1..16); both are 0 when the pair
 * is not uint x uint or a set is empty.  Fails when the two counts differ. */
int ehx_pair_record(ehx_ctx *ctx, const ehx_trie *a, const ehx_trie *b, const uint32_t *nodes0, const uint32_t *nodes1,
                    uint64_t n, int reps, uint64_t *cycles);

/* Learned per-pair algorithm of the learned form's uint x uint intersections
 * (the levels it does not run in the persistent kernel): entry
 * s * EHX_PAIR_RATIO_BINS + r is the choice for a smaller set of 2^s to
 * 2^(s+1) - 1 values and a larger/smaller ratio of 2^r to 2^(r+1) - 1 (the last
 * bins take everything beyond): 0 merge path, 1 gallop, 2 the reference's rule
 * (gallop when the ratio exceeds 32, uinteger.hpp:1209-1227).  The table is
 * the selector network's decision, tabulated by its trainer; table NULL
 * unloads it (ehx_selector_clear does too). */
#define EHX_PAIR_SMALL_BINS 24
#define EHX_PAIR_RATIO_BINS 24
int ehx_pair_table_load(int device, const uint8_t *table);

/* Training the selector on the GPU: data (n samples of 8 features, 2 targets,
 * 2 mask weights) lives on the device; ehx_mlp_grad returns the masked mean
 * squared error and its gradient with respect to the FP16-rounded weights
 * (params are FP32 masters, norm as for ehx_selector_load). */
typedef struct ehx_mlp ehx_mlp;
int ehx_mlp_create(ehx_ctx *ctx, const float *x, const float *y, const float *m, uint64_t n, ehx_mlp **out);
int ehx_mlp_destroy(ehx_mlp *h);
int ehx_mlp_grad(ehx_mlp *h, const float *params, const float *norm, float *grad, float *loss);
int ehx_mlp_predict(ehx_ctx *ctx, const float *params, const float *norm, const float *x, uint64_t n, float *out);

/* Page-locked host memory (hipHostMalloc), for callers that stage large
 * batches of host arrays through the entry points above (operand node lists,
 * counts, materialized values): copies from and to it run at the full PCIe
 * rate.  Free with ehx_host_free. */
int ehx_host_alloc(uint64_t bytes, void **out);
int ehx_host_free(void *p);

/* Measurement aid: one launch of a plain 16-byte-per-thread copy kernel that
 * reads and writes `bytes` bytes (the reference for scaling HBM traffic
 * counters); *ms receives its event-timed duration. */
int ehx_copy_calibrate(ehx_ctx *ctx, uint64_t bytes, double *ms);

#ifdef __cplusplus
}
#endif
#endif
