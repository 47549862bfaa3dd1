// ehx_selector.h -- learned layout selector: a 8 -> 32 -> 32 -> 2 MLP with FP16
// weights and FP32 accumulation, evaluated inline in the trie build (one
// thread per set) to choose uint or range_bitset per set.  The two outputs
// are predicted log2 costs of the set's intersections with the set staged as
// an LDS bitmap (output 0, the bitset layout) or as an LDS hash table
// (output 1, the uint layout); the lower wins.  Without loaded weights the
// density rule of common.hpp:56-67 decides.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ehx_internal.h"

namespace ehx {

constexpr int SEL_IN = 8, SEL_H = 32, SEL_OUT = 2;
// Parameter packing (halfs): W1[H][IN], b1[H], W2[H][H], b2[H], W3[OUT][H], b3[OUT]
constexpr int SEL_OFF_W1 = 0;
constexpr int SEL_OFF_B1 = SEL_OFF_W1 + SEL_H * SEL_IN;
constexpr int SEL_OFF_W2 = SEL_OFF_B1 + SEL_H;
constexpr int SEL_OFF_B2 = SEL_OFF_W2 + SEL_H * SEL_H;
constexpr int SEL_OFF_W3 = SEL_OFF_B2 + SEL_H;
constexpr int SEL_OFF_B3 = SEL_OFF_W3 + SEL_OUT * SEL_H;
constexpr int SEL_NPARAM = SEL_OFF_B3 + SEL_OUT;

struct SelectorWeights {
  const __half *w;     // SEL_NPARAM halfs (device)
  const float *norm;   // mean[SEL_IN], inv_std[SEL_IN] (device)
  float max_words;     // bitset staging limit (words) the model was trained for
  int valid;
};

SelectorWeights selector_current();

// Features of a set with n elements spanning [first, last] whose elements'
// own child sets hold `work` elements in total (sampled): cardinality, range,
// density, and the degree skew the set's intersections will see.
__host__ __device__ inline void selector_features(uint32_t n, uint64_t span, uint32_t first, uint32_t last,
                                                  uint32_t work, float *f) {
  float fn = (float)n;
  float fs = (float)span + 1.0f;
  float words = (float)((last >> 6) - (first >> 6) + 1);
  float fnn = fn > 0.0f ? fn : 1.0f;
  f[0] = log2f(fn + 1.0f);
  f[1] = log2f(fs);
  f[2] = log2f(fs / fnn);
  f[3] = (n <= 2) ? 1.0f : 0.0f;
  f[4] = log2f(words);
  f[5] = log2f((float)work + 1.0f);
  f[6] = log2f(((float)work + 1.0f) / fnn);  // mean child-set size of the elements
  f[7] = log2f((words * 12.0f + 8.0f) / (4.0f * fnn));  // bitset bytes / uint bytes
}

// Sampled sum of child-set sizes of a set's elements in trie t.
__device__ inline uint32_t set_work(const DTrie &t, const uint32_t *v, uint32_t n) {
  const uint32_t ns = n < 32 ? n : 32;
  uint64_t w = 0;
  for (uint32_t k = 0; k < ns; k++) {
    int32_t nx = node_of(t, v[(uint64_t)k * n / ns]);
    if (nx >= 0) w += t.off[nx + 1] - t.off[nx];
  }
  w = ns ? w * n / ns : 0;
  return w > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)w;
}

__host__ __device__ inline void selector_forward(const __half *w, const float *norm, const float *x_raw,
                                                 float *out) {
  float x[SEL_IN], h1[SEL_H], h2[SEL_H];
  for (int i = 0; i < SEL_IN; i++) x[i] = (x_raw[i] - norm[i]) * norm[SEL_IN + i];
  for (int j = 0; j < SEL_H; j++) {
    float a = __half2float(w[SEL_OFF_B1 + j]);
    for (int i = 0; i < SEL_IN; i++) a += __half2float(w[SEL_OFF_W1 + j * SEL_IN + i]) * x[i];
    h1[j] = a > 0.0f ? a : 0.0f;
  }
  for (int j = 0; j < SEL_H; j++) {
    float a = __half2float(w[SEL_OFF_B2 + j]);
    for (int i = 0; i < SEL_H; i++) a += __half2float(w[SEL_OFF_W2 + j * SEL_H + i]) * h1[i];
    h2[j] = a > 0.0f ? a : 0.0f;
  }
  for (int k = 0; k < SEL_OUT; k++) {
    float a = __half2float(w[SEL_OFF_B3 + k]);
    for (int i = 0; i < SEL_H; i++) a += __half2float(w[SEL_OFF_W3 + k * SEL_H + i]) * h2[i];
    out[k] = a;
  }
}

// Layout for one set: the learned choice where a bitset fits the staging
// limit, the rule's choice otherwise.
__device__ inline int selector_layout(const SelectorWeights &s, uint32_t n, uint64_t span, uint32_t first,
                                      uint32_t last, uint32_t work, int rule) {
  if (n == 0) return rule;
  float words = (float)((last >> 6) - (first >> 6) + 1);
  if (words > s.max_words) return 1;  // EHX_UINTEGER: bitset would not fit the staging area
  float f[SEL_IN], o[SEL_OUT];
  selector_features(n, span, first, last, work, f);
  selector_forward(s.w, s.norm, f, o);
  // outputs: predicted log2 cost staged as bitmap [0] and as hash [1]
  return (o[0] < o[1]) ? 0 : 1;  // 0 = EHX_RANGE_BITSET, 1 = EHX_UINTEGER
}

}  // namespace ehx
