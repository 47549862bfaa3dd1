// Recursive annotated SUM over a trie's child sets: the reference's PageRank
//   PageRank(x;y) :- Edge(x,z),y:float<-[(1.0/N)].
//   PageRank(x;y)*[i=5] :- Edge(x,z),PageRank(z),InvDegree(z),y:float<-[0.15+0.85*SUM(z;1.0)].
// (test/graph/travis.py:585-598).  The generated code (query_compiler/src/main/
// scala/codegen/query/EHGenerator.scala: the recursion loop of emitNPRR
// :752-764,820-828, emitAnnotationAccessors :398-433, emitFinalAnnotation
// :435-456, emitAnnotationComputation :457-478, emitAnnotationExpression
// :479-502) folds, per head x, over z in the sorted set child(x) ∩ keys(PageRank)
// ∩ keys(W):
//   float acc = 0;  acc += (float)(((double)1 * init * PR(z)) * W(z));
//   PR'(x) = (float)(add + mul * acc);
// with TrieIterator::get_annotation (storage_engine/src/TrieIterator.cpp:44-59)
// reading PR(z) and W(z).  g++ -O3 -march=native -mavx2, which compiles the
// generated code (cython/query/setup.py:10-20), fuses add + mul*acc into one
// double FMA (C++'s default -ffp-contract=fast), so the finish here is
// fma(mul, acc, add).  Here the two unary lookups become one dense array
// of per-value terms (a value absent from either relation gets the term +0.0,
// which leaves every float sum unchanged), computed once per iteration, and
// the fold is a gather over each child list:
//   plain form: one lane per head, terms added in ascending z -- the
//     reference's order, so the result is bit-identical to its fold;
//   tuned form: 16-lane groups per head; heads of more than 512 children are
//     cut into chunks of 4096 children, one wavefront each, whose partial sums
//     are added in chunk order.  Each lane sums a strided subsequence and the
//     lanes combine by a fixed butterfly, so results are deterministic (the
//     cuts depend on the degrees only) but rounded differently.
// Nothing else may be contracted: fp contraction is off in this file, and the
// one FMA is written out.
#include <hip/hip_runtime.h>

#include "ehx_internal.h"
#include "ehx_join_common.h"

#pragma clang fp contract(off)

namespace ehx {

constexpr uint32_t RANK_GROUP = 16;   // lanes per head in the tuned group kernel
constexpr uint32_t RANK_BIG = 512;    // heads with more children are cut into chunks
constexpr uint32_t RANK_CHUNK = 4096; // children per chunk, one wavefront each

// hv[v] = head node of value v, or -1 (the PageRank relation's keys are the heads)
__global__ void k_rank_heads(DTrie t, uint32_t V, int32_t *__restrict__ hv) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
    hv[v] = node_of(t, (uint32_t)v);
}

__global__ void k_rank_zero(float *__restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// wd[key] = annotation of the unary relation W (keys sorted, distinct, < V)
__global__ void k_rank_scatter(const uint32_t *__restrict__ keys, const float *__restrict__ vals, uint64_t n,
                               float *__restrict__ wd) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    wd[keys[i]] = vals[i];
}

// The base case: every head gets (float)base.
__global__ void k_rank_fill(float *__restrict__ pr, uint64_t nh, float base) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += (uint64_t)gridDim.x * blockDim.x)
    pr[i] = base;
}

// term[v] = (float)((init * PR(v)) * W(v)), +0 when v is not a head.
__global__ void k_rank_terms(const int32_t *__restrict__ hv, const float *__restrict__ wd,
                             const float *__restrict__ pr, uint32_t V, double init, float *__restrict__ term) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t h = hv[v];
    float r = 0.0f;
    if (h >= 0) r = (float)__dmul_rn(__dmul_rn(init, (double)pr[h]), (double)wd[v]);
    term[v] = r;
  }
}

__device__ __forceinline__ float rank_finish(float acc, double add, double mul) {
  return (float)__fma_rn(mul, (double)acc, add);
}

// Plain form: one lane per head, children in ascending order.
__global__ __launch_bounds__(256) void k_rank_plain(DTrie t, const float *__restrict__ term, double add, double mul,
                                                    float *__restrict__ out) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < t.nh; h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t b = t.off[h], e = t.off[h + 1];
    float acc = 0.0f;
    for (uint32_t k = b; k < e; k++) acc = __fadd_rn(acc, term[t.vals[k]]);
    out[h] = rank_finish(acc, add, mul);
  }
}

// Tuned form, heads of at most RANK_BIG children: RANK_GROUP lanes per head.
__global__ __launch_bounds__(256) void k_rank_group(DTrie t, const float *__restrict__ term, double add, double mul,
                                                    float *__restrict__ out) {
  const uint32_t g = threadIdx.x & (RANK_GROUP - 1);
  const uint64_t ngroups = ((uint64_t)gridDim.x * blockDim.x) / RANK_GROUP;
  for (uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / RANK_GROUP; h < t.nh; h += ngroups) {
    const uint32_t b = t.off[h], e = t.off[h + 1];
    if (e - b > RANK_BIG) continue;  // uniform across the group
    float a0 = 0.0f, a1 = 0.0f;
    uint32_t k = b + g;
    for (; k + RANK_GROUP < e; k += 2 * RANK_GROUP) {  // two gathers in flight per lane
      const uint32_t z0 = t.vals[k], z1 = t.vals[k + RANK_GROUP];
      a0 = __fadd_rn(a0, term[z0]);
      a1 = __fadd_rn(a1, term[z1]);
    }
    if (k < e) a0 = __fadd_rn(a0, term[t.vals[k]]);
    float acc = __fadd_rn(a0, a1);
#pragma unroll
    for (uint32_t s = RANK_GROUP / 2; s >= 1; s >>= 1) acc = __fadd_rn(acc, __shfl_xor(acc, (int)s, RANK_GROUP));
    if (g == 0) out[h] = rank_finish(acc, add, mul);
  }
}

// Tuned form, heads of more than RANK_BIG children: cnt[i] = chunks of big head i.
__global__ void k_rank_chunk_counts(DTrie t, const uint32_t *__restrict__ bigs, uint32_t nb,
                                    uint32_t *__restrict__ cnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t h = bigs[i];
    cnt[i] = (t.off[h + 1] - t.off[h] + RANK_CHUNK - 1) / RANK_CHUNK;
  }
}

// One wavefront per chunk j (< choff[nb]): the big head i with choff[i] <= j <
// choff[i+1], its children [c*RANK_CHUNK, (c+1)*RANK_CHUNK) with c = j - choff[i],
// four loads in flight per lane; partial[j] = their sum.
__global__ __launch_bounds__(256) void k_rank_chunk(DTrie t, const uint32_t *__restrict__ bigs,
                                                    const uint32_t *__restrict__ choff, uint32_t nb, uint32_t nch,
                                                    const float *__restrict__ term, float *__restrict__ partial) {
  const uint32_t lane = __lane_id();
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t j = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < nch; j += nwaves) {
    uint32_t lo = 0, hi = nb;  // last i with choff[i] <= j
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (choff[mid] <= j) lo = mid; else hi = mid;
    }
    const uint32_t h = bigs[lo];
    const uint32_t b = t.off[h] + ((uint32_t)j - choff[lo]) * RANK_CHUNK;
    const uint32_t e = min(t.off[h + 1], b + RANK_CHUNK);
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    uint32_t k = b + lane;
    for (; k + 192 < e; k += 256) {
      const uint32_t z0 = t.vals[k], z1 = t.vals[k + 64], z2 = t.vals[k + 128], z3 = t.vals[k + 192];
      a0 = __fadd_rn(a0, term[z0]);
      a1 = __fadd_rn(a1, term[z1]);
      a2 = __fadd_rn(a2, term[z2]);
      a3 = __fadd_rn(a3, term[z3]);
    }
    for (; k < e; k += 64) a0 = __fadd_rn(a0, term[t.vals[k]]);
    float acc = __fadd_rn(__fadd_rn(a0, a1), __fadd_rn(a2, a3));
#pragma unroll
    for (uint32_t s = 32; s >= 1; s >>= 1) acc = __fadd_rn(acc, __shfl_xor(acc, (int)s, 64));
    if (lane == 0) partial[j] = acc;
  }
}

// One lane per big head: its chunks' partial sums in chunk order, then the finish.
__global__ void k_rank_combine(const uint32_t *__restrict__ bigs, const uint32_t *__restrict__ choff, uint32_t nb,
                               const float *__restrict__ partial, double add, double mul, float *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (uint64_t)gridDim.x * blockDim.x) {
    float acc = 0.0f;
    for (uint32_t j = choff[i]; j < choff[i + 1]; j++) acc = __fadd_rn(acc, partial[j]);
    out[bigs[i]] = rank_finish(acc, add, mul);
  }
}

// bigs[] = heads with more than RANK_BIG children, in head order.
__global__ __launch_bounds__(256) void k_rank_bigs(DTrie t, uint32_t *__restrict__ bigs, unsigned *__restrict__ nb) {
  const uint32_t lane = __lane_id();
  for (uint64_t h0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ull; h0 < t.nh;
       h0 += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t h = h0 + lane;
    const bool big = h < t.nh && t.off[h + 1] - t.off[h] > RANK_BIG;
    const uint64_t m = __ballot(big);
    unsigned base = 0;
    if (lane == 0 && m) base = atomicAdd(nb, (unsigned)__popcll(m));
    base = __shfl(base, 0, 64);
    if (big) bigs[base + wave_popc_below(m)] = (uint32_t)h;
  }
}

int rank_recursive(ehx_ctx *ctx, const ehx_trie *t, const uint32_t *w_keys, const float *w_vals, uint64_t n_w,
                   const ehx_rank_rule *rule, int form, float *rank_out, double *ms_out) {
  EHX_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int rc;
  if (!rule || rule->iterations < 0) return fail(EHX_ERR_INVALID, "bad rank rule");
  if (form != EHX_FORM_PLAIN && form != EHX_FORM_TUNED && form != EHX_FORM_LEARNED)
    return fail(EHX_ERR_INVALID, "unknown form %d", form);
  if (n_w && (!w_keys || !w_vals)) return fail(EHX_ERR_INVALID, "null annotation arrays");
  for (uint64_t i = 1; i < n_w; i++)
    if (w_keys[i] <= w_keys[i - 1]) return fail(EHX_ERR_INVALID, "annotation keys not sorted and distinct at %llu",
                                                (unsigned long long)i);
  const uint64_t nh = t->nh;
  if (nh == 0) {
    if (ms_out) *ms_out = 0.0;
    return EHX_OK;
  }
  const uint64_t V = (uint64_t)t->info.max_value + 1;
  if (V > (1ull << 31)) return fail(EHX_ERR_UNSUPPORTED, "value domain of %llu exceeds 2^31", (unsigned long long)V);
  uint64_t nw_in = 0;  // keys inside the value domain (larger keys have no edge into them)
  while (nw_in < n_w && w_keys[nw_in] < V) nw_in++;
  DevBuf &bh = ctx->scratch[S_F0], &bw = ctx->scratch[S_F1], &bt = ctx->scratch[S_F2], &bp = ctx->scratch[S_F3],
         &bq = ctx->scratch[S_F4], &bk = ctx->scratch[S_F5], &bv = ctx->scratch[S_F6], &bb = ctx->scratch[S_F7],
         &bc = ctx->scratch[S_MISC];
  if ((rc = bh.ensure(V * 4)) || (rc = bw.ensure(V * 4)) || (rc = bt.ensure(V * 4)) || (rc = bp.ensure(nh * 4)) ||
      (rc = bq.ensure(nh * 4)) || (rc = bk.ensure(nw_in * 4 + 4)) || (rc = bv.ensure(nw_in * 4 + 4)) ||
      (rc = bb.ensure(nh * 4)) || (rc = bc.ensure(64)))
    return rc;
  int32_t *hv = bh.as<int32_t>();
  float *wd = bw.as<float>(), *term = bt.as<float>(), *cur = bp.as<float>(), *nxt = bq.as<float>();
  const DTrie dv = view(t);
  const unsigned gV = grid_for(V, 256, 8192), gH = grid_for(nh, 256, 8192);
  k_rank_heads<<<gV, 256, 0, st>>>(dv, (uint32_t)V, hv);
  k_rank_zero<<<gV, 256, 0, st>>>(wd, V);
  EHX_HIP(hipGetLastError());
  if (nw_in) {
    if ((rc = copy_h2d(ctx, bk.p, w_keys, nw_in * 4)) || (rc = copy_h2d(ctx, bv.p, w_vals, nw_in * 4))) return rc;
    k_rank_scatter<<<grid_for(nw_in, 256, 8192), 256, 0, st>>>(bk.as<uint32_t>(), bv.as<float>(), nw_in, wd);
    EHX_HIP(hipGetLastError());
  }
  uint32_t nb = 0, nch = 0;
  DevBuf &bn = ctx->scratch[S_COUNTS], &bo = ctx->scratch[S_OFFS], &bpart = ctx->scratch[S_ACC];
  if (form != EHX_FORM_PLAIN) {
    if ((rc = fill_u32(ctx, bc.as<uint32_t>(), 0u, 1))) return rc;
    k_rank_bigs<<<gH, 256, 0, st>>>(dv, bb.as<uint32_t>(), bc.as<unsigned>());
    EHX_HIP(hipGetLastError());
    if ((rc = copy_d2h(ctx, &nb, bc.p, 4))) return rc;
    if (nb) {
      if ((rc = bn.ensure((uint64_t)nb * 4 + 4)) || (rc = bo.ensure((uint64_t)nb * 4 + 4))) return rc;
      k_rank_chunk_counts<<<grid_for(nb, 256, 8192), 256, 0, st>>>(dv, bb.as<uint32_t>(), nb, bn.as<uint32_t>());
      EHX_HIP(hipGetLastError());
      if ((rc = scan_u32(ctx, bn.as<uint32_t>(), bo.as<uint32_t>(), nb))) return rc;
      if ((rc = copy_d2h(ctx, &nch, bo.as<uint32_t>() + nb, 4))) return rc;
      if ((rc = bpart.ensure((uint64_t)nch * 4 + 4))) return rc;
    }
  }
  k_rank_fill<<<gH, 256, 0, st>>>(cur, nh, (float)rule->base);
  EHX_HIP(hipGetLastError());
  hipEvent_t e0 = ctx->ev[0], e1 = ctx->ev[1];
  EHX_HIP(hipEventRecord(e0, st));
  for (int it = 0; it < rule->iterations; it++) {
    k_rank_terms<<<gV, 256, 0, st>>>(hv, wd, cur, (uint32_t)V, rule->init, term);
    if (form == EHX_FORM_PLAIN) {
      k_rank_plain<<<gH, 256, 0, st>>>(dv, term, rule->add, rule->mul, nxt);
    } else {
      k_rank_group<<<grid_for(nh * RANK_GROUP, 256, 16384), 256, 0, st>>>(dv, term, rule->add, rule->mul, nxt);
      if (nb) {
        k_rank_chunk<<<grid_for((uint64_t)nch * 64, 256, 16384), 256, 0, st>>>(
            dv, bb.as<uint32_t>(), bo.as<uint32_t>(), nb, nch, term, bpart.as<float>());
        k_rank_combine<<<grid_for(nb, 256, 8192), 256, 0, st>>>(bb.as<uint32_t>(), bo.as<uint32_t>(), nb,
                                                                 bpart.as<float>(), rule->add, rule->mul, nxt);
      }
    }
    EHX_HIP(hipGetLastError());
    float *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  EHX_HIP(hipEventRecord(e1, st));
  EHX_HIP(hipEventSynchronize(e1));
  if (ms_out) {
    float ms = 0.0f;
    EHX_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_out = ms;
  }
  return copy_d2h(ctx, rank_out, cur, nh * 4);
}

}  // namespace ehx
