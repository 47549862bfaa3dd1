// ehx_join.hip -- level-synchronous generic join (COUNT) on the device.
//
// The generated code runs the join depth-first per thread, consuming each
// intersection inside host lambdas (EHGenerator.scala:273-520).  Here each
// level runs as device launches over a materialized frontier of prefix tuples:
//   level 0     ParTrieBuilder::build_aggregated_set over the head sets
//               (TrieBuilder.cpp:478-578): flag + scan + compact;
//   level k     TrieBuilder::build_aggregated_set per tuple (TrieBuilder.cpp:
//               242-348): count pass, exclusive scan, write pass (the frontier
//               stays in lexicographic order, the order foreach_aggregate visits);
//   last level  count only, folded into one int64 (par::reducer, parallel.hpp:8-34).
// The plain form gives every tuple one wavefront; the tuned and learned forms
// replace the last level with the degree-binned persistent kernel in
// ehx_tuned.hip.
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "ehx_join_common.h"
#include "ehx_tuned.h"

namespace ehx {

struct Frontier {
  const uint32_t *col[EHX_MAX_ATTRS];
};
struct OutFrontier {
  uint32_t *col[EHX_MAX_ATTRS];
};

enum { MODE_COUNT = 0, MODE_WRITE = 1, MODE_SUM = 2 };

// One wavefront per tuple.  The tuple and its operand sets are wave-uniform and
// indexed at run time (by the plan's parent and operand numbers), so they live
// in the wave's LDS rather than in per-lane private arrays, which the compiler
// would place in scratch memory and re-read on every step.  Lane i < nops
// resolves operand i (its child lookup runs beside the others).
template <int MODE>
__global__ __launch_bounds__(256) void k_level_plain(Tries T, LevelDesc L, Frontier in, int k, uint64_t n,
                                                     uint32_t *__restrict__ counts, const uint64_t *__restrict__ offs,
                                                     OutFrontier outf, unsigned long long *__restrict__ acc,
                                                     SegOut seg, int pick, const uint8_t *__restrict__ ptab) {
  __shared__ unsigned long long s_acc[6];
  __shared__ DSet s_sets[4][EHX_MAX_OPS];
  __shared__ uint32_t s_tuple[4][EHX_MAX_ATTRS];
  if (MODE == MODE_SUM) {
    if (threadIdx.x < 6) s_acc[threadIdx.x] = 0;
    __syncthreads();
  }
  const uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  DSet *S = s_sets[wv];
  uint32_t *tuple = s_tuple[wv];
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long sum = 0, n_merge = 0, n_gallop = 0, n_probe = 0, n_and = 0, n_empty = 0;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n; t += nwaves) {
    if (lane < (uint32_t)k) tuple[lane] = in.col[lane][t];
    wave_sync();
    bool ok = true;
    if (lane < (uint32_t)L.nops) {
      const DTrie &tr = T.t[L.rel[lane]];
      const int par = L.parent[lane];
      if (par == EHX_PARENT_HEAD) {
        S[lane] = set_of(tr, tr.nh);
      } else {
        const int32_t h = node_of(tr, par >= 0 ? tuple[par] : L.value[lane]);  // else EHX_PARENT_CONST
        if (h < 0) ok = false;  // NULL TrieBlock: empty (TrieBuilder.cpp:249-254)
        else S[lane] = set_of(tr, (uint32_t)h);
      }
    }
    const bool resolved = __ballot(!ok) == 0;
    wave_sync();
    uint32_t c = 0;
    int alg = ALG_EMPTY;
    if (resolved) {
      if (MODE == MODE_WRITE) {
        uint64_t o = offs[t];
        c = wave_intersect<true>(S, L.nops, outf.col[k] + o, &alg, pick != 0, ptab);
        for (int j = 0; j < k; j++) {
          const uint32_t v = tuple[j];
          for (uint32_t i = lane; i < c; i += 64) outf.col[j][o + i] = v;
        }
      } else {
        c = wave_intersect<false>(S, L.nops, nullptr, &alg, pick != 0, ptab);
      }
    }
    if (MODE == MODE_COUNT && lane == 0) counts[t] = c;
    if (MODE == MODE_SUM) {
      sum += c;
      n_merge += alg == ALG_MERGE;
      n_gallop += alg == ALG_GALLOP;
      n_probe += alg == ALG_PROBE;
      n_and += alg == ALG_AND;
      n_empty += alg == ALG_EMPTY;
      if (seg.out && lane == 0) seg_add(seg, tuple[0], c);
    }
    wave_sync();  // every lane has read this tuple's sets before the next overwrites them
  }
  if (MODE == MODE_SUM) {
    if (lane == 0) {
      const unsigned long long per[6] = {sum, n_merge, n_gallop, n_probe, n_and, n_empty};
#pragma unroll
      for (int a = 0; a < 6; a++)
        if (per[a]) atomicAdd(&s_acc[a], per[a]);
    }
    __syncthreads();
    if (threadIdx.x < 6 && s_acc[threadIdx.x]) atomicAdd(&acc[threadIdx.x], s_acc[threadIdx.x]);
  }
}

// Last level with one operand: its count is the operand's cardinality
// (build_aggregated_set over one TrieBlock, TrieBuilder.cpp:223-240), so one
// thread per tuple suffices.
__global__ __launch_bounds__(256) void k_level_card(Tries T, LevelDesc L, Frontier in, uint64_t n,
                                                    unsigned long long *__restrict__ acc, SegOut seg) {
  const DTrie &t = T.t[L.rel[0]];
  const int par = L.parent[0];
  unsigned long long sum = 0, nonempty = 0, empty = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c;
    if (par == EHX_PARENT_HEAD) {
      c = t.nh;
    } else {
      const int32_t h = node_of(t, par == EHX_PARENT_CONST ? L.value[0] : in.col[par][i]);
      c = h >= 0 ? t.off[h + 1] - t.off[h] : 0u;
    }
    sum += c;
    if (c) nonempty++;
    else empty++;
    if (seg.out && c) seg_add(seg, in.col[0][i], c);
  }
  // one set of atomics per workgroup: same-address atomics serialize in L2
  __shared__ unsigned long long s_acc[3];
  if (threadIdx.x < 3) s_acc[threadIdx.x] = 0;
  __syncthreads();
  sum = wave_sum(sum);
  nonempty = wave_sum(nonempty);
  empty = wave_sum(empty);
  if (__lane_id() == 0) {
    atomicAdd(&s_acc[0], sum);
    atomicAdd(&s_acc[1], nonempty);
    atomicAdd(&s_acc[2], empty);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_acc[0]) atomicAdd(&acc[0], s_acc[0]);
    if (s_acc[1]) atomicAdd(&acc[1 + ALG_MERGE], s_acc[1]);
    if (s_acc[2]) atomicAdd(&acc[1 + ALG_EMPTY], s_acc[2]);
  }
}

__global__ void k_copy_u64(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst) { *dst = *src; }

// Level 0: flags over the driver's set dv[0, dn), membership in the other
// operands' sets (a head set, or the child set of a selection constant).
__global__ void k_level0_flags(Tries T, LevelDesc L, int driver, const uint32_t *__restrict__ dv, uint64_t dn,
                               uint32_t *__restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dn; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t x = dv[i];
    bool hit = true;
    for (int o = 0; o < L.nops && hit; o++) {
      if (o == driver) continue;
      const DTrie &t = T.t[L.rel[o]];
      if (L.parent[o] == EHX_PARENT_CONST) {
        const int32_t h = node_of(t, L.value[o]);
        hit = h >= 0 && set_contains(set_of(t, (uint32_t)h), x);
      } else {
        hit = node_of(t, x) >= 0;  // membership in a head set = a child exists
      }
    }
    flags[i] = hit ? 1u : 0u;
  }
}

// {offset in vals, length} of the child set of a constant (0, 0 when absent).
__global__ void k_const_set(DTrie t, uint32_t value, uint32_t *__restrict__ out) {
  const int32_t h = node_of(t, value);
  out[0] = h >= 0 ? t.off[h] : 0u;
  out[1] = h >= 0 ? t.off[h + 1] - t.off[h] : 0u;
}

__global__ void k_compact_u32(const uint32_t *__restrict__ v, const uint32_t *__restrict__ flags,
                              const uint64_t *__restrict__ pos, uint64_t n, uint32_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (flags[i]) out[pos[i]] = v[i];
}

static int check_plan(const ehx_plan *p, ehx_trie *const *tries, int *nrels) {
  if (!p || !tries) return fail(EHX_ERR_INVALID, "null plan or tries");
  if (p->num_attrs < 1 || p->num_attrs > EHX_MAX_ATTRS) return fail(EHX_ERR_INVALID, "num_attrs %d out of range", p->num_attrs);
  int mr = -1;
  for (int k = 0; k < p->num_attrs; k++) {
    if (p->num_ops[k] < 1 || p->num_ops[k] > EHX_MAX_OPS)
      return fail(EHX_ERR_INVALID, "level %d: num_ops %d out of range", k, p->num_ops[k]);
    for (int i = 0; i < p->num_ops[k]; i++) {
      const ehx_op &o = p->ops[k][i];
      if (o.rel < 0 || o.rel >= EHX_MAX_RELS) return fail(EHX_ERR_INVALID, "level %d op %d: rel %d", k, i, o.rel);
      if (o.parent >= k) return fail(EHX_ERR_INVALID, "level %d op %d: parent %d not before level", k, i, o.parent);
      if (o.parent < EHX_PARENT_CONST) return fail(EHX_ERR_INVALID, "level %d op %d: parent %d", k, i, o.parent);
      if (o.rel > mr) mr = o.rel;
    }
  }
  for (int r = 0; r <= mr; r++)
    if (!tries[r]) return fail(EHX_ERR_INVALID, "relation %d has no trie", r);
  *nrels = mr + 1;
  return EHX_OK;
}

static LevelDesc level_desc(const ehx_plan *p, int k) {
  LevelDesc L;
  memset(&L, 0, sizeof(L));
  L.nops = p->num_ops[k];
  for (int i = 0; i < L.nops; i++) {
    L.rel[i] = p->ops[k][i].rel;
    L.parent[i] = p->ops[k][i].parent;
    L.value[i] = p->ops[k][i].value;
  }
  return L;
}

static const uint64_t FRONTIER_BUDGET = 96ull << 30;  // bytes per materialized level

// Single-attribute plans: every level-0 value is one result tuple.
__global__ void k_seg_level0(const uint32_t *__restrict__ f0, uint64_t n, SegOut seg) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    seg_add(seg, f0[i], 1ull);
}

// Level 0: ParTrieBuilder::build_aggregated_set over the head sets
// (TrieBuilder.cpp:478-578).  One relation: its head set; several: the driver
// (a selection's child set if there is one, else the smallest head) filtered by
// membership in the others, then compacted.
// n0_dev (may be null): when level 0 is a filtered head set, leave its count on
// the device instead (*n0_dev points at it) and return its upper bound in
// *n0_out, with no host synchronization.
static int level0(ehx_ctx *ctx, const ehx_plan *p, ehx_trie *const *tries, const Tries &T, const uint32_t **f0_out,
                  uint64_t *n0_out, const uint64_t **n0_dev = nullptr) {
  if (n0_dev) *n0_dev = nullptr;
  hipStream_t st = ctx->stream;
  int rc;
  LevelDesc L0 = level_desc(p, 0);
  int driver = -1;
  for (int i = 0; i < L0.nops && driver < 0; i++)
    if (L0.parent[i] == EHX_PARENT_CONST) driver = i;
  const uint32_t *dv = nullptr;
  uint64_t dn = 0;
  if (driver >= 0) {  // the selection's child set: {offset, length} from the device
    DevBuf &cb = ctx->scratch[S_MISC];
    if ((rc = cb.ensure(64))) return rc;
    const ehx_trie *tc = tries[L0.rel[driver]];
    k_const_set<<<1, 1, 0, st>>>(T.t[L0.rel[driver]], L0.value[driver], cb.as<uint32_t>());
    EHX_HIP(hipGetLastError());
    uint32_t ol[2] = {0, 0};
    if ((rc = copy_d2h(ctx, ol, cb.p, 8))) return rc;
    dv = tc->vals + ol[0];
    dn = ol[1];
    if (L0.nops == 1 || dn == 0) {
      *f0_out = dv;
      *n0_out = dn;
      return EHX_OK;
    }
  } else {
    driver = 0;
    for (int i = 1; i < L0.nops; i++)
      if (tries[L0.rel[i]]->nh <= tries[L0.rel[driver]]->nh) driver = i;
    bool single = true;
    for (int i = 0; i < L0.nops; i++)
      if (L0.rel[i] != L0.rel[driver]) single = false;
    const ehx_trie *td = tries[L0.rel[driver]];
    if (single) {
      *f0_out = td->head;  // head ∩ head of one trie is the head itself
      *n0_out = td->nh;
      return EHX_OK;
    }
    dv = td->head;
    dn = td->nh;
  }
  DevBuf &fl = ctx->scratch[S_FLAGS], &ps = ctx->scratch[S_POS], &fb = ctx->scratch[S_F0];
  if ((rc = fl.ensure(dn * 4 + 4)) || (rc = ps.ensure(dn * 8 + 8)) || (rc = fb.ensure(dn * 4 + 4))) return rc;
  k_level0_flags<<<grid_for(dn, 256, 16384), 256, 0, st>>>(T, L0, driver, dv, dn, fl.as<uint32_t>());
  EHX_HIP(hipGetLastError());
  if ((rc = scan_u32_to_u64(ctx, fl.as<uint32_t>(), ps.as<uint64_t>(), dn))) return rc;
  k_compact_u32<<<grid_for(dn, 256, 16384), 256, 0, st>>>(dv, fl.as<uint32_t>(), ps.as<uint64_t>(), dn,
                                                          fb.as<uint32_t>());
  EHX_HIP(hipGetLastError());
  *f0_out = fb.as<uint32_t>();
  if (n0_dev) {
    *n0_dev = ps.as<uint64_t>() + dn;
    *n0_out = dn;
    return EHX_OK;
  }
  return copy_d2h(ctx, n0_out, ps.as<uint64_t>() + dn, 8);
}

// ---- a whole count query in one graph -----------------------------------------
// A plan whose level 0 intersects head sets and whose remaining levels the
// tuned kernel runs (three attributes) has no host synchronization left once
// level 0's count stays on the device: its launches are captured once and the
// graph is replayed.  The key holds every value the launches bake in; context
// buffer addresses are covered by the buffers' generation count.
struct QueryGraph {
  bool valid = false;
  std::vector<unsigned char> key;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int64_t n0_bound = 0;
};
static QueryGraph g_qgraph;
static std::mutex g_qgraph_mu;

static void query_graph_drop() {
  if (g_qgraph.exec) (void)hipGraphExecDestroy(g_qgraph.exec);
  if (g_qgraph.graph) (void)hipGraphDestroy(g_qgraph.graph);
  g_qgraph = QueryGraph();
}

void query_graph_forget() {
  std::lock_guard<std::mutex> lk(g_qgraph_mu);
  query_graph_drop();
}

template <class V>
static void key_add(std::vector<unsigned char> &k, const V &v) {
  const unsigned char *b = reinterpret_cast<const unsigned char *>(&v);
  k.insert(k.end(), b, b + sizeof(V));
}

// Level 0 of p reads head sets only (no selection constant) and is not one
// trie's head alone.
static bool level0_filtered_heads(const ehx_plan *p, ehx_trie *const *tries) {
  const LevelDesc L0 = level_desc(p, 0);
  for (int i = 0; i < L0.nops; i++)
    if (L0.parent[i] != EHX_PARENT_HEAD) return false;
  for (int i = 1; i < L0.nops; i++)
    if (L0.rel[i] != L0.rel[0]) return true;
  return false;
}

static int query_graph_count(ehx_ctx *ctx, const ehx_plan *p, ehx_trie *const *tries, int nrels, const Tries &T,
                             const TunedPlan &tp, int form, DevBuf &ac, ehx_join_stats &S) {
  hipStream_t st = ctx->stream;
  int rc;
  // everything that allocates or synchronizes happens before the capture
  TunedArgs ta;
  memset(&ta, 0, sizeof(ta));
  tuned_bind(tp, T, ta);
  if (tp.ns == 1 &&
      (rc = trie_link(ctx, tries[tp.nd ? tp.d_rel : tp.s_rel[0]], tries[tp.x_rel], &ta.xlink, &ta.xhint, &ta.xwork)))
    return rc;
  const LevelDesc L0 = level_desc(p, 0);
  int driver = 0;
  for (int i = 1; i < L0.nops; i++)
    if (tries[L0.rel[i]]->nh <= tries[L0.rel[driver]]->nh) driver = i;
  const uint64_t dn = tries[L0.rel[driver]]->nh;
  DevBuf &fl = ctx->scratch[S_FLAGS], &ps = ctx->scratch[S_POS], &fb = ctx->scratch[S_F0];
  if ((rc = fl.ensure(dn * 4 + 4)) || (rc = ps.ensure(dn * 8 + 8)) || (rc = fb.ensure(dn * 4 + 4)) ||
      (rc = scan_reserve(ctx, dn)) || (rc = tuned_reserve(ctx, tp, dn)))
    return rc;
  std::vector<unsigned char> key;
  key_add(key, *p);
  key_add(key, form);
  key_add(key, ctx->device);
  key_add(key, st);
  key_add(key, ac.p);
  key_add(key, tune_params());
  key_add(key, pair_table_current());
  key_add(key, devbuf_generation());
  for (int r = 0; r < nrels; r++) {
    key_add(key, tries[r]->uid);
    key_add(key, T.t[r]);
  }
  key_add(key, ta.xlink);
  key_add(key, ta.xhint);
  key_add(key, ta.xwork);
  std::lock_guard<std::mutex> lk(g_qgraph_mu);
  if (!(g_qgraph.valid && g_qgraph.key == key)) {
    query_graph_drop();
    EHX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const uint32_t *f0 = nullptr;
    uint64_t n0 = 0;
    const uint64_t *n0_dev = nullptr;
    rc = fill_u64(ctx, ac.as<uint64_t>(), 0, 8);
    if (!rc) rc = level0(ctx, p, tries, T, &f0, &n0, &n0_dev);
    if (!rc && (!n0_dev || n0 != dn)) rc = fail(EHX_ERR_DEVICE, "level 0 of a captured query");
    if (!rc) {
      for (int j = 0; j < EHX_MAX_ATTRS; j++) ta.prefix[j] = nullptr;
      ta.prefix[0] = f0;
      tuned_bind(tp, T, ta);
      ta.nprefix = dn;
      ta.nprefix_dev = n0_dev;
      ta.force_mode = -1;
      rc = tuned_run(ctx, tp, ta, form, ac.as<unsigned long long>(), false);
    }
    // the level-0 count beside the result, read back with it
    if (!rc) {
      k_copy_u64<<<1, 1, 0, st>>>(n0_dev, ac.as<uint64_t>() + 7);
      if (hipGetLastError() != hipSuccess) rc = fail(EHX_ERR_DEVICE, "launch in a captured query");
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return rc ? rc : fail(EHX_ERR_DEVICE, "query capture: %s", hipGetErrorString(e));
    }
    const hipError_t ie = hipGraphInstantiate(&g_qgraph.exec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      return fail(EHX_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    }
    g_qgraph.graph = g;
    g_qgraph.key = key;
    g_qgraph.valid = true;
  }
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K0], st));
  EHX_HIP(hipGraphLaunch(g_qgraph.exec, st));
  EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K1], st));
  S.frontier[0] = dn;  // replaced by the count read back with the result
  return EHX_OK;
}

// Level k materialized: count pass, exclusive scan, write pass (the frontier
// stays lexicographic, the order foreach_aggregate visits).  The k+1 output
// columns live in scratch slot S_F0 + k.
static int expand_level(ehx_ctx *ctx, const Tries &T, const LevelDesc &L, Frontier &in, int k, uint64_t n, int pick,
                        const uint8_t *ptab, uint64_t *nn_out) {
  hipStream_t st = ctx->stream;
  int rc;
  DevBuf &cb = ctx->scratch[S_COUNTS], &ob = ctx->scratch[S_OFFS];
  if ((rc = cb.ensure(n * 4 + 4)) || (rc = ob.ensure(n * 8 + 8))) return rc;
  unsigned G = grid_for(n * 64, 256, 65536);
  OutFrontier none;
  memset(&none, 0, sizeof(none));
  if (n) {
    k_level_plain<MODE_COUNT><<<G, 256, 0, st>>>(T, L, in, k, n, cb.as<uint32_t>(), nullptr, none, nullptr,
                                                 SegOut{}, pick, ptab);
    EHX_HIP(hipGetLastError());
  }
  if ((rc = scan_u32_to_u64(ctx, cb.as<uint32_t>(), ob.as<uint64_t>(), n))) return rc;
  uint64_t nn = 0;
  if ((rc = copy_d2h(ctx, &nn, ob.as<uint64_t>() + n, 8))) return rc;
  if (nn * 4ull * (k + 1) > FRONTIER_BUDGET)
    return fail(EHX_ERR_UNSUPPORTED, "level %d frontier of %llu tuples exceeds the %llu-byte budget", k,
                (unsigned long long)nn, (unsigned long long)FRONTIER_BUDGET);
  DevBuf &fb = ctx->scratch[S_F0 + k];
  if ((rc = fb.ensure(nn * 4ull * (k + 1) + 4))) return rc;
  OutFrontier out;
  memset(&out, 0, sizeof(out));
  for (int j = 0; j <= k; j++) out.col[j] = fb.as<uint32_t>() + (uint64_t)j * nn;
  if (n && nn) {
    k_level_plain<MODE_WRITE><<<G, 256, 0, st>>>(T, L, in, k, n, nullptr, ob.as<uint64_t>(), out, nullptr, SegOut{},
                                                 pick, ptab);
    EHX_HIP(hipGetLastError());
  }
  for (int j = 0; j <= k; j++) in.col[j] = out.col[j];
  *nn_out = nn;
  return EHX_OK;
}

int join_count(ehx_ctx *ctx, const ehx_plan *p, ehx_trie *const *tries, int form, uint64_t begin, uint64_t end,
               int64_t *result, ehx_join_stats *stats, const RecordReq *rec, const SegOut *seg) {
  int nrels = 0, rc;
  if ((rc = check_plan(p, tries, &nrels))) return rc;
  if (!result) return fail(EHX_ERR_INVALID, "null result");
  if (form < EHX_FORM_PLAIN || form > EHX_FORM_LEARNED) return fail(EHX_ERR_INVALID, "bad form %d", form);
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  ehx_join_stats S;
  memset(&S, 0, sizeof(S));
  Tries T;
  memset(&T, 0, sizeof(T));
  for (int r = 0; r < nrels; r++) {
    if (tries[r]->device != ctx->device) return fail(EHX_ERR_INVALID, "trie %d lives on another device", r);
    T.t[r] = view(tries[r]);
  }
  const int K = p->num_attrs;
  const int pick = form != EHX_FORM_PLAIN;  // per-pair algorithm by cost in the tuned and learned forms
  const uint8_t *ptab = form == EHX_FORM_LEARNED ? pair_table_current() : nullptr;  // learned uint x uint choice
  EHX_HIP(hipEventRecord(ctx->ev[0], st));

  // ---- the whole query as one replayed graph ------------------------------------
  TunedPlan qtp;
  if (tune_params().use_graph && form != EHX_FORM_PLAIN && K == 3 && !rec && !seg && begin == 0 &&
      end == ~0ull && tuned_applicable(p, &qtp) && level0_filtered_heads(p, tries)) {
    DevBuf &ac = ctx->scratch[S_ACC];
    if ((rc = ac.ensure(8 * 8))) return rc;
    if ((rc = query_graph_count(ctx, p, tries, nrels, T, qtp, form, ac, S))) return rc;
    unsigned long long accv[8];
    if ((rc = copy_d2h(ctx, accv, ac.p, 64))) return rc;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev[EHX_EV_K0], ctx->ev[EHX_EV_K1]) == hipSuccess) {
      S.main_kernel_ms = ms;
      S.level_ms[K - 1] = ms;
    }
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[EHX_EV_K1]) == hipSuccess) S.total_ms = ms;
    S.frontier[0] = accv[7];
    for (int a = 0; a < 5; a++) S.pairs_by_alg[a] = accv[1 + a];
    S.form_used = form;
    *result = (int64_t)accv[0];
    if (stats) *stats = S;
    return EHX_OK;
  }

  // ---- level 0 -------------------------------------------------------------
  const uint32_t *f0;
  uint64_t n0;
  if ((rc = level0(ctx, p, tries, T, &f0, &n0))) return rc;
  if (begin > n0) begin = n0;
  if (end > n0) end = n0;
  if (end < begin) end = begin;
  f0 += begin;
  n0 = end - begin;
  S.frontier[0] = n0;
  EHX_HIP(hipEventRecord(ctx->ev[1], st));

  DevBuf &ac = ctx->scratch[S_ACC];
  if ((rc = ac.ensure(8 * 8))) return rc;
  if ((rc = fill_u64(ctx, ac.as<uint64_t>(), 0, 8))) return rc;

  Frontier in;
  memset(&in, 0, sizeof(in));
  in.col[0] = f0;
  uint64_t n = n0;
  int64_t total = 0;
  int form_used = EHX_FORM_PLAIN;
  if (K == 1) {
    total = (int64_t)n0;
    if (seg && n0) {
      k_seg_level0<<<grid_for(n0, 256, 4096), 256, 0, st>>>(f0, n0, *seg);
      EHX_HIP(hipGetLastError());
    }
  } else {
    // ---- middle levels (materialized) ------------------------------------------
    int last_mat = K - 2;  // materialize levels 1..K-2
    bool tuned = false;
    TunedPlan tp;
    if (form != EHX_FORM_PLAIN && tuned_applicable(p, &tp)) {
      tuned = true;
      last_mat = K - 3;  // the tuned kernel enumerates level K-2 itself
    }
    for (int k = 1; k <= last_mat; k++) {
      LevelDesc L = level_desc(p, k);
      uint64_t nn = 0;
      if ((rc = expand_level(ctx, T, L, in, k, n, pick, ptab, &nn))) return rc;
      n = nn;
      S.frontier[k] = nn;
      EHX_HIP(hipEventRecord(ctx->ev[1 + k], st));
    }
    // ---- last level -------------------------------------------------------------
    // events: ev[1 + k] closes materialized level k; ev[K] closes the last level
    if (rec && !tuned) return fail(EHX_ERR_UNSUPPORTED, "recording needs a plan the tuned form runs");
    if (tuned) {
      TunedArgs ta;
      memset(&ta, 0, sizeof(ta));
      for (int j = 0; j < EHX_MAX_ATTRS; j++) ta.prefix[j] = in.col[j];
      tuned_bind(tp, T, ta);
      if (tp.ns == 1) {  // child pointers from E(p)'s trie into x's
        if ((rc = trie_link(ctx, tries[tp.nd ? tp.d_rel : tp.s_rel[0]], tries[tp.x_rel], &ta.xlink, &ta.xhint, &ta.xwork)))
          return rc;
      }
      ta.nprefix = n;
      ta.force_mode = -1;
      if (seg) ta.seg = *seg;
      DevBuf &rb = ctx->scratch[S_IN1];
      if (rec) {
        if ((rc = rb.ensure(n * 8 + 8)) || (rc = fill_u64(ctx, rb.as<uint64_t>(), 0, n))) return rc;
        ta.rec = rb.as<unsigned long long>();
        ta.force_mode = rec->force_mode;
      }
      if ((rc = tuned_run(ctx, tp, ta, form, ac.as<unsigned long long>()))) return rc;
      if (rec) {
        if (rec->n_out) *rec->n_out = n;
        if ((rc = copy_d2h(ctx, rec->host_out, rb.p, 8 * (n < rec->cap ? n : rec->cap)))) return rc;
      }
      form_used = form;
    } else {
      LevelDesc L = level_desc(p, K - 1);
      EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K0], st));
      if (n && L.nops == 1) {
        k_level_card<<<grid_for(n, 256, (unsigned)ctx->num_cus * 8), 256, 0, st>>>(T, L, in, n, ac.as<unsigned long long>(),
                                                              seg ? *seg : SegOut{});
        EHX_HIP(hipGetLastError());
      } else if (n) {
        unsigned G = grid_for(n * 64, 256, 65536);
        OutFrontier none;
        memset(&none, 0, sizeof(none));
        k_level_plain<MODE_SUM><<<G, 256, 0, st>>>(T, L, in, K - 1, n, nullptr, nullptr, none,
                                                   ac.as<unsigned long long>(), seg ? *seg : SegOut{}, pick, ptab);
        EHX_HIP(hipGetLastError());
      }
      EHX_HIP(hipEventRecord(ctx->ev[EHX_EV_K1], st));
    }
    EHX_HIP(hipEventRecord(ctx->ev[K], st));
    unsigned long long accv[8];
    if ((rc = copy_d2h(ctx, accv, ac.p, 64))) return rc;
    total = (int64_t)accv[0];
    for (int a = 0; a < 5; a++) S.pairs_by_alg[a] = accv[1 + a];
    float ms = 0;
    for (int k = 1; k <= last_mat; k++)
      if (hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]) == hipSuccess) S.level_ms[k] = ms;
    if (hipEventElapsedTime(&ms, ctx->ev[last_mat + 1], ctx->ev[K]) == hipSuccess) S.level_ms[K - 1] = ms;
    if (n && hipEventElapsedTime(&ms, ctx->ev[EHX_EV_K0], ctx->ev[EHX_EV_K1]) == hipSuccess) S.main_kernel_ms = ms;
  }
  EHX_HIP(hipEventRecord(ctx->ev[K + 1], st));
  EHX_HIP(hipEventSynchronize(ctx->ev[K + 1]));
  float ms = 0;
  EHX_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  S.level_ms[0] = ms;
  EHX_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[K + 1]));
  S.total_ms = ms;
  S.form_used = form_used;
  *result = total;
  if (stats) *stats = S;
  return EHX_OK;
}

// ---- materialized result -------------------------------------------------------------
// Every result tuple, in lexicographic order: each level is materialized with a
// count pass, a scan and a write pass (what TrieBuilder::build_set, set_level and
// allocate_next build level by level, TrieBuilder.cpp:44-221,376-446, read back in
// Trie::foreach order, Trie.cpp:78-160).
int join_materialize(ehx_ctx *ctx, const ehx_plan *p, ehx_trie *const *tries, int form, uint32_t *const *cols,
                     uint64_t cap, uint64_t *rows, ehx_join_stats *stats) {
  int nrels = 0, rc;
  if (!ctx) return fail(EHX_ERR_INVALID, "null ctx");
  if ((rc = check_plan(p, tries, &nrels))) return rc;
  if (!rows) return fail(EHX_ERR_INVALID, "null rows");
  if (form < EHX_FORM_PLAIN || form > EHX_FORM_LEARNED) return fail(EHX_ERR_INVALID, "bad form %d", form);
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  ehx_join_stats S;
  memset(&S, 0, sizeof(S));
  Tries T;
  memset(&T, 0, sizeof(T));
  for (int r = 0; r < nrels; r++) {
    if (tries[r]->device != ctx->device) return fail(EHX_ERR_INVALID, "trie %d lives on another device", r);
    T.t[r] = view(tries[r]);
  }
  const int K = p->num_attrs;
  EHX_HIP(hipEventRecord(ctx->ev[0], st));
  const uint32_t *f0;
  uint64_t n;
  if ((rc = level0(ctx, p, tries, T, &f0, &n))) return rc;
  S.frontier[0] = n;
  Frontier in;
  memset(&in, 0, sizeof(in));
  in.col[0] = f0;
  const uint8_t *ptab = form == EHX_FORM_LEARNED ? pair_table_current() : nullptr;
  for (int k = 1; k < K; k++) {
    uint64_t nn = 0;
    if ((rc = expand_level(ctx, T, level_desc(p, k), in, k, n, form != EHX_FORM_PLAIN, ptab, &nn))) return rc;
    n = nn;
    S.frontier[k] = nn;
  }
  EHX_HIP(hipEventRecord(ctx->ev[1], st));
  *rows = n;
  if (n > cap) return fail(EHX_ERR_NOMEM, "result has %llu rows, capacity %llu", (unsigned long long)n,
                           (unsigned long long)cap);
  for (int j = 0; j < K && n; j++) {
    if (!cols || !cols[j]) return fail(EHX_ERR_INVALID, "null output column %d", j);
    if ((rc = copy_d2h(ctx, cols[j], in.col[j], n * 4))) return rc;
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) S.total_ms = ms;
  S.form_used = form;  // the per-pair choices differ by form (cost rule, learned table)
  if (stats) *stats = S;
  return EHX_OK;
}

// ---- per-segment counts -------------------------------------------------------------
int seg_begin(ehx_ctx *ctx, const uint32_t *bounds, uint64_t nseg, SegOut *so) {
  if (!bounds || nseg == 0) return fail(EHX_ERR_INVALID, "segments: null bounds or nseg == 0");
  if (nseg >= (1ull << 32)) return fail(EHX_ERR_UNSUPPORTED, "segments: more than 2^32 segments");
  for (uint64_t s = 0; s < nseg; s++)
    if (bounds[s] > bounds[s + 1]) return fail(EHX_ERR_INVALID, "segments: bounds not ascending at %llu",
                                               (unsigned long long)s);
  DevBuf &b = ctx->scratch[S_SEGB], &o = ctx->scratch[S_SEGO];
  int rc;
  if ((rc = b.ensure((nseg + 1) * 4)) || (rc = o.ensure(nseg * 8))) return rc;
  if ((rc = copy_h2d(ctx, b.p, bounds, (nseg + 1) * 4))) return rc;
  if ((rc = fill_u64(ctx, o.as<uint64_t>(), 0, nseg))) return rc;
  so->bounds = b.as<uint32_t>();
  so->out = o.as<unsigned long long>();
  so->nseg = (uint32_t)nseg;
  so->pad = 0;
  return EHX_OK;
}

int seg_end(ehx_ctx *ctx, const SegOut &so, int64_t *counts) {
  if (!counts) return EHX_OK;
  return copy_d2h(ctx, counts, so.out, (size_t)so.nseg * 8);
}

int join_count_segments(ehx_ctx *ctx, const ehx_plan *p, ehx_trie *const *tries, int form, const uint32_t *bounds,
                        uint64_t nseg, int64_t *counts, int64_t *result, ehx_join_stats *stats) {
  if (!ctx) return fail(EHX_ERR_INVALID, "null ctx");
  EHX_HIP(hipSetDevice(ctx->device));
  SegOut so;
  int rc;
  if ((rc = seg_begin(ctx, bounds, nseg, &so))) return rc;
  if ((rc = join_count(ctx, p, tries, form, 0, UINT64_MAX, result, stats, nullptr, &so))) return rc;
  return seg_end(ctx, so, counts);
}

// ---- batched pair intersections ---------------------------------------------------
template <bool EMIT>
__global__ __launch_bounds__(256) void k_pairs(DTrie A, DTrie B, const uint32_t *__restrict__ ka,
                                               const uint32_t *__restrict__ kb, uint64_t n,
                                               unsigned long long *__restrict__ counts,
                                               const unsigned long long *__restrict__ offs, uint32_t *__restrict__ out) {
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n; t += nwaves) {
    int32_t ha = node_of(A, ka[t]), hb = node_of(B, kb[t]);
    uint32_t c = 0;
    int alg;
    if (ha >= 0 && hb >= 0) {
      DSet S[2] = {set_of(A, (uint32_t)ha), set_of(B, (uint32_t)hb)};
      c = wave_intersect<EMIT>(S, 2, EMIT ? out + offs[t] : nullptr, &alg);
    }
    if (!EMIT && __lane_id() == 0) counts[t] = c;
  }
}

int intersect_pairs(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a, const ehx_trie *tb,
                    const uint32_t *keys_b, uint64_t n, uint64_t *counts, uint64_t *offsets, uint32_t *values,
                    uint64_t cap, uint64_t *total) {
  if (!ctx || !ta || !tb || (n && (!keys_a || !keys_b))) return fail(EHX_ERR_INVALID, "null argument");
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (n == 0) {
    if (offsets) offsets[0] = 0;
    if (total) *total = 0;
    return EHX_OK;
  }
  DevBuf &a = ctx->scratch[S_IN0], &b = ctx->scratch[S_IN1], &c = ctx->scratch[S_COUNTS], &o = ctx->scratch[S_OFFS];
  int rc;
  if ((rc = a.ensure(n * 4)) || (rc = b.ensure(n * 4)) || (rc = c.ensure(n * 8 + 8)) || (rc = o.ensure(n * 8 + 8)))
    return rc;
  EHX_HIP(hipMemcpyAsync(a.p, keys_a, n * 4, hipMemcpyHostToDevice, st));
  EHX_HIP(hipMemcpyAsync(b.p, keys_b, n * 4, hipMemcpyHostToDevice, st));
  unsigned G = grid_for(n * 64, 256, 65536);
  k_pairs<false><<<G, 256, 0, st>>>(view(ta), view(tb), a.as<uint32_t>(), b.as<uint32_t>(), n,
                                    c.as<unsigned long long>(), nullptr, nullptr);
  EHX_HIP(hipGetLastError());
  if (counts) return copy_d2h(ctx, counts, c.p, n * 8);
  if ((rc = scan_u64(ctx, c.as<uint64_t>(), o.as<uint64_t>(), n))) return rc;
  uint64_t tot = 0;
  if ((rc = copy_d2h(ctx, &tot, o.as<uint64_t>() + n, 8))) return rc;
  if (total) *total = tot;
  if (tot > cap) return fail(EHX_ERR_NOMEM, "intersection needs %llu values, capacity %llu", (unsigned long long)tot,
                             (unsigned long long)cap);
  DevBuf &v = ctx->scratch[S_MISC];
  if ((rc = v.ensure(tot * 4 + 4))) return rc;
  k_pairs<true><<<G, 256, 0, st>>>(view(ta), view(tb), a.as<uint32_t>(), b.as<uint32_t>(), n, nullptr,
                                   o.as<unsigned long long>(), v.as<uint32_t>());
  EHX_HIP(hipGetLastError());
  if ((rc = copy_d2h(ctx, offsets, o.p, (n + 1) * 8))) return rc;
  return copy_d2h(ctx, values, v.p, tot * 4);
}

// ---- pair intersection counts without allocating ---------------------------------
// The scratch of ehx_intersect_count_batch_noalloc lives in this file-scope
// device array: the staged keys of both sides (u32 each) and the counts (u64),
// for at most EHX_NOALLOC_MAX_PAIRS pairs (16 MiB).  Each device holds its own
// copy until the process ends.  Host code never passes `scratch` itself (its
// name is a host-side stand-in there): the device address of each device's
// copy is taken once with hipGetSymbolAddress, with that device current, and
// kept in g_noalloc_ptr.
constexpr uint64_t NOALLOC_KEYS_B = EHX_NOALLOC_MAX_PAIRS * 4;
constexpr uint64_t NOALLOC_BYTES = EHX_NOALLOC_MAX_PAIRS * 16;
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static uint8_t *g_noalloc_ptr[NOALLOC_MAX_DEVICES];

// The device address of `device`'s copy of scratch (that device current).
static int noalloc_scratch(int device, uint8_t **out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(EHX_ERR_UNSUPPORTED, "device %d", device);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return fail(EHX_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    }
    g_noalloc_ptr[device] = (uint8_t *)ptr;
  }
  *out = g_noalloc_ptr[device];
  return EHX_OK;
}

int intersect_count_noalloc(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a, const ehx_trie *tb,
                            const uint32_t *keys_b, uint64_t n, uint64_t *counts) {
  if (!ctx || !ta || !tb || (n && (!keys_a || !keys_b || !counts))) return fail(EHX_ERR_INVALID, "null argument");
  if (n > EHX_NOALLOC_MAX_PAIRS)
    return fail(EHX_ERR_INVALID, "%llu pairs: the no-allocation batch takes at most %llu", (unsigned long long)n,
                (unsigned long long)EHX_NOALLOC_MAX_PAIRS);
  if (ta->device != ctx->device || tb->device != ctx->device)
    return fail(EHX_ERR_INVALID, "trie lives on another device");
  if (n == 0) return EHX_OK;
  EHX_HIP(hipSetDevice(ctx->device));
  uint8_t *base = nullptr;
  int rc;
  if ((rc = noalloc_scratch(ctx->device, &base))) return rc;
  uint32_t *ka = (uint32_t *)base, *kb = (uint32_t *)(base + NOALLOC_KEYS_B);
  unsigned long long *cnt = (unsigned long long *)(base + 2 * NOALLOC_KEYS_B);
  hipStream_t st = ctx->stream;
  EHX_HIP(hipMemcpyAsync(ka, keys_a, n * 4, hipMemcpyHostToDevice, st));
  EHX_HIP(hipMemcpyAsync(kb, keys_b, n * 4, hipMemcpyHostToDevice, st));
  // every count is written (lane 0 of the pair's wave), so nothing is cleared first
  k_pairs<false><<<grid_for(n * 64, 256, 65536), 256, 0, st>>>(view(ta), view(tb), ka, kb, n, cnt, nullptr, nullptr);
  EHX_HIP(hipGetLastError());
  return copy_d2h(ctx, counts, cnt, n * 8);
}

// ---- batched N-way intersections of trie nodes ------------------------------------
struct NodeOps {
  DTrie t[EHX_MAX_OPS];
  const uint32_t *nodes[EHX_MAX_OPS];
  int32_t nsets;
};

// One wavefront per call; lane s < nsets resolves operand s into the wave's LDS
// (runtime-indexed per-lane arrays would live in scratch memory).
template <bool EMIT>
__global__ __launch_bounds__(256) void k_nodes(NodeOps O, uint64_t n, unsigned long long *__restrict__ counts,
                                               const unsigned long long *__restrict__ offs, uint32_t *__restrict__ out) {
  __shared__ DSet s_sets[4][EHX_MAX_OPS];
  const uint32_t lane = __lane_id();
  DSet *S = s_sets[threadIdx.x >> 6];
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n; t += nwaves) {
    if (lane < (uint32_t)O.nsets) S[lane] = set_of(O.t[lane], O.nodes[lane][t]);
    wave_sync();
    int alg;
    uint32_t c = wave_intersect<EMIT>(S, O.nsets, EMIT ? out + offs[t] : nullptr, &alg);
    if (!EMIT && lane == 0) counts[t] = c;
    wave_sync();  // every lane has read these sets before the next call overwrites them
  }
}

// Engine-recorded timings of the two uint x uint algorithms: for pair t, one
// wavefront counts A ∩ B by merge path, then by galloping the smaller set into
// the larger, `reps` times, and keeps the fewest shader-clock cycles of each
// (cyc[2t], cyc[2t + 1]); both counts must agree (bad[0] = 1 otherwise).  A
// pair that is not uint x uint, or has an empty set, records zeros.
__global__ __launch_bounds__(256) void k_pair_record(NodeOps O, uint64_t n, int reps,
                                                     unsigned long long *__restrict__ cyc, unsigned *__restrict__ bad) {
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n; t += nwaves) {
    const DSet A = set_of(O.t[0], O.nodes[0][t]), B = set_of(O.t[1], O.nodes[1][t]);
    unsigned long long best0 = ~0ull, best1 = ~0ull;
    if (A.lay != EHX_UINTEGER || B.lay != EHX_UINTEGER || A.n == 0 || B.n == 0) {
      best0 = best1 = 0;
    } else {
      const DSet &r = (A.n > B.n) ? B : A;
      const DSet &f = (A.n > B.n) ? A : B;
      for (int k = 0; k < reps; k++) {
        const unsigned long long t0 = __builtin_amdgcn_s_memtime();
        const uint32_t c0 = __builtin_amdgcn_readfirstlane(wave_merge<false>(r.v, r.n, f.v, f.n, nullptr));
        asm volatile("" ::"s"(c0));  // the clock is read after the count is known
        const unsigned long long t1 = __builtin_amdgcn_s_memtime();
        const uint32_t c1 = __builtin_amdgcn_readfirstlane(wave_gallop<false>(r.v, r.n, f.v, f.n, nullptr));
        asm volatile("" ::"s"(c1));
        const unsigned long long t2 = __builtin_amdgcn_s_memtime();
        best0 = min(best0, t1 - t0);
        best1 = min(best1, t2 - t1);
        if (c0 != c1 && __lane_id() == 0) atomicOr(bad, 1u);
      }
    }
    if (__lane_id() == 0) {
      cyc[2 * t] = best0;
      cyc[2 * t + 1] = best1;
    }
  }
}

// Operand node lists checked on the device before any set is read: bad[0] =
// 1 when some node of operand s exceeds lim[s] (node == num_heads names the
// head set).
struct NodeLimits {
  const uint32_t *nodes[EHX_MAX_OPS];
  uint32_t lim[EHX_MAX_OPS];
  int32_t nsets;
};
__global__ __launch_bounds__(256) void k_check_nodes(NodeLimits L, uint64_t n, unsigned *__restrict__ bad) {
  bool over = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    for (int s = 0; s < L.nsets; s++) over |= L.nodes[s][i] > L.lim[s];
  if (__any(over) && __lane_id() == 0) atomicOr(bad, 1u);
}

// Count-only pairs whose smaller set holds at most SMALL_PAIR values, 16
// lanes per pair (four pairs per wavefront): each lane probes its share of the
// smaller set's values in the larger set (binary search in a uint set, a bit
// test in a bitset), and the group sums the hits.  A wavefront per pair would
// leave most of its lanes idle on such sets.  Pairs above the bound are left
// to k_nodes, which skips the small ones.
constexpr uint32_t SMALL_PAIR = 256;
__device__ __forceinline__ bool small_pair(const DSet &A, const DSet &B) { return min(A.n, B.n) <= SMALL_PAIR; }
__global__ __launch_bounds__(256) void k_pairs_small(NodeOps O, uint64_t n, unsigned long long *__restrict__ counts) {
  const uint32_t g = __lane_id() & 15;
  const uint64_t ngroups = ((uint64_t)gridDim.x * blockDim.x) >> 4;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; t < n; t += ngroups) {
    const DSet A = set_of(O.t[0], O.nodes[0][t]), B = set_of(O.t[1], O.nodes[1][t]);
    if (!small_pair(A, B)) continue;  // group-uniform
    const DSet &r = A.n <= B.n ? A : B, &f = A.n <= B.n ? B : A;
    uint32_t hits = 0;
    for (uint32_t i = g; i < r.n; i += 16) hits += set_contains(f, r.v[i]) ? 1u : 0u;
#pragma unroll
    for (int d = 8; d >= 1; d >>= 1) hits += __shfl_xor(hits, d, 16);
    if (g == 0) counts[t] = hits;
  }
}

template <bool EMIT, bool BIG_ONLY>
__global__ __launch_bounds__(256) void k_nodes_big(NodeOps O, uint64_t n, unsigned long long *__restrict__ counts) {
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n; t += nwaves) {
    DSet S[2] = {set_of(O.t[0], O.nodes[0][t]), set_of(O.t[1], O.nodes[1][t])};
    if (BIG_ONLY && small_pair(S[0], S[1])) continue;  // wave-uniform
    int alg;
    const uint32_t c = wave_intersect<false>(S, 2, nullptr, &alg);
    if (__lane_id() == 0) counts[t] = c;
  }
}

int pair_record(ehx_ctx *ctx, const ehx_trie *a, const ehx_trie *b, const uint32_t *nodes0, const uint32_t *nodes1,
                uint64_t n, int reps, uint64_t *cycles) {
  if (!ctx || !a || !b || !nodes0 || !nodes1 || !cycles) return fail(EHX_ERR_INVALID, "null argument");
  if (reps < 1 || reps > 16) return fail(EHX_ERR_INVALID, "reps %d (1..16)", reps);
  if (a->device != ctx->device || b->device != ctx->device) return fail(EHX_ERR_INVALID, "trie lives on another device");
  for (uint64_t i = 0; i < n; i++)
    if (nodes0[i] > a->nh || nodes1[i] > b->nh)
      return fail(EHX_ERR_INVALID, "pair %llu: node out of range", (unsigned long long)i);
  if (n == 0) return EHX_OK;
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  NodeOps O;
  memset(&O, 0, sizeof(O));
  O.nsets = 2;
  DevBuf &nb = ctx->scratch[S_IN0], &c = ctx->scratch[S_COUNTS], &fl = ctx->scratch[S_FLAGS];
  int rc;
  if ((rc = nb.ensure(n * 8)) || (rc = c.ensure(n * 16)) || (rc = fl.ensure(64))) return rc;
  O.t[0] = view(a);
  O.t[1] = view(b);
  O.nodes[0] = nb.as<uint32_t>();
  O.nodes[1] = nb.as<uint32_t>() + n;
  if ((rc = copy_h2d(ctx, (void *)O.nodes[0], nodes0, n * 4)) || (rc = copy_h2d(ctx, (void *)O.nodes[1], nodes1, n * 4)) ||
      (rc = fill_u32(ctx, fl.as<uint32_t>(), 0u, 1)))
    return rc;
  k_pair_record<<<grid_for(n * 64, 256, 65536), 256, 0, st>>>(O, n, reps, c.as<unsigned long long>(),
                                                                fl.as<unsigned>());
  EHX_HIP(hipGetLastError());
  uint32_t bad = 0;
  if ((rc = copy_d2h(ctx, &bad, fl.p, 4))) return rc;
  if (bad) return fail(EHX_ERR_DEVICE, "merge and gallop counts disagree");
  return copy_d2h(ctx, cycles, c.p, n * 16);
}

int intersect_nodes(ehx_ctx *ctx, int nsets, ehx_trie *const *tries, const uint32_t *const *nodes, uint64_t n,
                    uint64_t *counts, uint64_t *offsets, uint32_t *values, uint64_t cap, uint64_t *total) {
  if (!ctx || !tries || !nodes || nsets < 1 || nsets > EHX_MAX_OPS) return fail(EHX_ERR_INVALID, "bad node batch");
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (n == 0) {
    if (offsets) offsets[0] = 0;
    if (total) *total = 0;
    return EHX_OK;
  }
  NodeOps O;
  memset(&O, 0, sizeof(O));
  O.nsets = nsets;
  DevBuf &nb = ctx->scratch[S_IN0], &c = ctx->scratch[S_COUNTS], &o = ctx->scratch[S_OFFS];
  DevBuf &fl = ctx->scratch[S_FLAGS];
  int rc;
  if ((rc = nb.ensure((uint64_t)nsets * n * 4)) || (rc = c.ensure(n * 8 + 8)) || (rc = o.ensure(n * 8 + 8)) ||
      (rc = fl.ensure(64)))
    return rc;
  NodeLimits NL;
  memset(&NL, 0, sizeof(NL));
  NL.nsets = nsets;
  for (int s = 0; s < nsets; s++) {
    if (!tries[s] || !nodes[s]) return fail(EHX_ERR_INVALID, "operand %d: null trie or nodes", s);
    if (tries[s]->device != ctx->device) return fail(EHX_ERR_INVALID, "operand %d lives on another device", s);
    O.t[s] = view(tries[s]);
    O.nodes[s] = nb.as<uint32_t>() + (uint64_t)s * n;
    NL.nodes[s] = O.nodes[s];
    NL.lim[s] = (uint32_t)tries[s]->nh;
    EHX_HIP(hipMemcpyAsync((void *)O.nodes[s], nodes[s], n * 4, hipMemcpyHostToDevice, st));
  }
  // every node in range before a set is read
  if ((rc = fill_u32(ctx, fl.as<uint32_t>(), 0u, 1))) return rc;
  k_check_nodes<<<grid_for(n, 256, 8192), 256, 0, st>>>(NL, n, fl.as<unsigned>());
  EHX_HIP(hipGetLastError());
  uint32_t bad = 0;
  if ((rc = copy_d2h(ctx, &bad, fl.p, 4))) return rc;
  if (bad) return fail(EHX_ERR_INVALID, "an operand node is out of range (above its trie's head count)");
  unsigned G = grid_for(n * 64, 256, 65536);
  if (counts && nsets == 2) {  // count-only pairs: small ones 16 lanes each, the rest a wavefront each
    k_pairs_small<<<grid_for(n * 16, 256, 65536), 256, 0, st>>>(O, n, c.as<unsigned long long>());
    EHX_HIP(hipGetLastError());
    k_nodes_big<false, true><<<G, 256, 0, st>>>(O, n, c.as<unsigned long long>());
  } else {
    k_nodes<false><<<G, 256, 0, st>>>(O, n, c.as<unsigned long long>(), nullptr, nullptr);
  }
  EHX_HIP(hipGetLastError());
  if (counts) return copy_d2h(ctx, counts, c.p, n * 8);
  if ((rc = scan_u64(ctx, c.as<uint64_t>(), o.as<uint64_t>(), n))) return rc;
  uint64_t tot = 0;
  if ((rc = copy_d2h(ctx, &tot, o.as<uint64_t>() + n, 8))) return rc;
  if (total) *total = tot;
  if (tot > cap) return fail(EHX_ERR_NOMEM, "intersection needs %llu values, capacity %llu", (unsigned long long)tot,
                             (unsigned long long)cap);
  DevBuf &v = ctx->scratch[S_MISC];
  if ((rc = v.ensure(tot * 4 + 4))) return rc;
  k_nodes<true><<<G, 256, 0, st>>>(O, n, nullptr, o.as<unsigned long long>(), v.as<uint32_t>());
  EHX_HIP(hipGetLastError());
  if ((rc = copy_d2h(ctx, offsets, o.p, (n + 1) * 8))) return rc;
  return copy_d2h(ctx, values, v.p, tot * 4);
}

}  // namespace ehx
