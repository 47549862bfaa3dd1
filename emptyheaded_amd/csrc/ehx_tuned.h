// ehx_tuned.h -- the batched last level of a clique-shaped plan.
//
// Applies when the last level's operands are one child set of x_{K-2} (in
// relation x_rel) plus child sets S_1..S_s of earlier attributes, and level K-2
// intersects head-set filters with either exactly those S_i or one other child
// set D of an earlier attribute.  Then for each prefix p (a tuple of levels
// 0..K-3)
//     count(p) = sum over x in E(p) ∩ heads of |child(x_rel, x) ∩ S(p)|,
// with S(p) = S_1 ∩ .. ∩ S_s and E(p) = S(p) or D(p).  Triangle (S = N(a)),
// 4-clique (S = N(a) ∩ N(b)) and the triangle ordered by its middle vertex
// (E = in-neighbors of b, S = out-neighbors of b) have this shape.  One persistent kernel drains a queue of
// prefixes binned by estimated work (largest first); each wavefront stages S(p)
// in LDS once -- as a bitmap when S(p) is a bitset that fits, as an open-
// addressing hash table otherwise -- and streams every child(x) through it.
#pragma once

#include "ehx_join_common.h"

namespace ehx {

struct TunedPlan {
  int32_t K;
  int32_t ns;
  int32_t s_rel[EHX_MAX_OPS], s_parent[EHX_MAX_OPS];
  int32_t nhf;
  int32_t h_rel[EHX_MAX_OPS];
  int32_t x_rel;
  // nd = 1: level K-2 enumerates x over one driver set D(p) = child(d_rel,
  // prefix[d_parent]) instead of over S(p) itself
  int32_t nd, d_rel, d_parent;
};

constexpr int TUNED_MAX_S = 4;  // operand sets of S(p) the tuned kernel accepts
constexpr int TUNED_MAX_H = 4;  // head-set filters of level K-2

// Kernel arguments.  The plan's operands are bound to fixed slots on the host
// (tuned_bind), so the kernels index them only with unrolled loop counters: a
// table of tries indexed by a runtime relation number would be copied to
// scratch memory.
struct TunedArgs {
  const uint32_t *prefix[EHX_MAX_ATTRS];
  uint64_t nprefix;
  unsigned long long *rec;  // per-prefix shader-clock cycles when recording, else null
  int32_t force_mode;       // -1: staging by layout; 0: bitmap where it fits; 1: hash where it fits
  int32_t nh;               // head filters in use
  SegOut seg;               // per-segment counts by first attribute value (seg.out null: off)
  DTrie ts[TUNED_MAX_S];              // trie of each S operand
  const uint32_t *sp[TUNED_MAX_S];    // prefix column holding each S operand's parent value
  DTrie tx;                           // relation of child(x)
  DTrie td;                           // driver set's relation (nd = 1)
  const uint32_t *dp;                 // prefix column of the driver's parent
  DTrie th[TUNED_MAX_H];              // head-set filters
  // single-S plans: the operands of queue entry i, resolved by the binning pass,
  // as desc[2i] = {S offset, |S|, D offset, |D|}, desc[2i+1] = {first of S,
  // last of S, S's bitset word offset, S's layout}; null: resolved per prefix
  const uint4 *desc;
  // with desc: xlink[e] = x's child list {off, length, first chunk, chunks} in
  // the x trie for the value at position e of E(p)'s trie; null: looked up by value
  const uint4 *xlink;
  // with xlink: xhint[e] = values of that child list below the parent value
  // of E(p) (the head of the set holding e); a child list whose values start
  // below S(p)'s first value is streamed from there when S(p)'s first value is
  // at least that parent value (null: streamed whole)
  const uint32_t *xhint;
  // with xlink: xwork[h] = the sum of the child-list lengths of the elements
  // of set h of E(p)'s trie, the binning's work estimate
  const uint64_t *xwork;
  // with a whole query captured in one graph: the prefix count on the device
  // (level 0's output); nprefix is then its upper bound, which sizes the grids
  // and buffers.  Null: nprefix is the count.
  const uint64_t *nprefix_dev;
};

// Knobs picked by the autotuner (ehx_tune_set); defaults are the measured best.
struct TuneParams {
  int32_t lds_slots;      // 32-bit LDS slots per workgroup (2048, 4096 or 8192), a quarter per wavefront
  int32_t big_deg;        // child sets at least this long are streamed by the whole wave
  int32_t waves_per_cu;   // persistent grid size
  int32_t grab;           // queue items taken per atomic
  int32_t bitmap_all;     // 1: stage any set whose word window fits as a bitmap
  int32_t use_graph;      // capture the level in a hipGraph and replay it
  int32_t big_work;       // prefixes with at least this much estimated work go to the workgroup bin
  int32_t stream_lanes;   // lanes that stream one long child set together (a power of two, 2..64)
  int32_t gap_coded;      // 1: stream the gap-coded copy of a child list where the trie has one (single-S plans)
  int32_t split_work;     // workgroup-bin prefixes are split into parts of about this much work (0: never)
  int32_t tiny_work;      // single-S prefixes below this estimated work run up to 64 per wavefront (0: never)
};

// A request to time every prefix of the tuned level under a forced staging.
struct RecordReq {
  uint64_t *host_out;  // cycles per prefix (host)
  uint64_t cap;
  uint64_t *n_out;     // number of prefixes
  int force_mode;      // -1 layout-driven, 0 bitmap, 1 hash
};

bool tuned_applicable(const ehx_plan *p, TunedPlan *tp);
// Fills A's operand slots from the plan, the tries and A.prefix.
void tuned_bind(const TunedPlan &P, const Tries &T, TunedArgs &A);
// allow_graph = false: enqueue the level's launches on the stream as they are
// (the caller may be capturing them itself).
int tuned_run(ehx_ctx *ctx, const TunedPlan &tp, const TunedArgs &a, int form, unsigned long long *acc,
              bool allow_graph = true);
// Sizes the context buffers tuned_run uses for up to nprefix prefixes, so that a
// later tuned_run of at most that many allocates nothing.
int tuned_reserve(ehx_ctx *ctx, const TunedPlan &tp, uint64_t nprefix);
TuneParams &tune_params();

}  // namespace ehx
