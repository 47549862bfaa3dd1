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
  int32_t lds_slots;    /* 32-bit LDS slots per wavefront for the staged set: 2048, 4096, 8192 */
  int32_t big_deg;      /* child sets at least this long are streamed by the whole wavefront */
  int32_t waves_per_cu; /* persistent grid = CUs x waves_per_cu one-wave workgroups */
  int32_t grab;         /* queue entries taken per atomic */
  int32_t bitmap_all;   /* stage every set whose word window fits as a bitmap */
  int32_t use_graph;    /* capture the level in a hipGraph and replay it */
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

#ifdef __cplusplus
}
#endif
#endif
