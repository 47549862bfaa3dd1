// ehx_api.cpp -- extern "C" entry points of libehx (declared in include/ehx.h
// and include/ehx_ext.h), error reporting, contexts.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/ehx.h"
#include "../../include/ehx_ext.h"
#include "ehx_internal.h"
#include "ehx_tuned.h"

namespace ehx {

static thread_local std::string t_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_err = buf;
}

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_err = buf;
  return code;
}

int trie_build(ehx_ctx *, const uint32_t *, const uint32_t *, uint64_t, int, int, ehx_trie **);
int trie_destroy(ehx_trie *);
int trie_export(ehx_ctx *, const ehx_trie *, uint32_t *, uint32_t *);
int trie_node_info(ehx_ctx *, const ehx_trie *, uint8_t *, uint32_t *);
int trie_lookup(ehx_ctx *, const ehx_trie *, const uint32_t *, uint64_t, int64_t *);
struct SegOut;
int join_count(ehx_ctx *, const ehx_plan *, ehx_trie *const *, int, uint64_t, uint64_t, int64_t *, ehx_join_stats *,
               const RecordReq *, const SegOut *);
int join_count_segments(ehx_ctx *, const ehx_plan *, ehx_trie *const *, int, const uint32_t *, uint64_t, int64_t *,
                        int64_t *, ehx_join_stats *);
int cycle4_count(ehx_ctx *, const ehx_trie *, uint64_t, uint64_t, const SegOut *, int64_t *, ehx_join_stats *);
int weighted_pairs(ehx_ctx *, const ehx_trie *, const int64_t *, uint64_t, const int64_t *, uint64_t, int64_t *);
int sssp_unweighted(ehx_ctx *, const ehx_trie *, uint32_t, int, uint32_t *, uint64_t, uint32_t *);
int sssp_directed(ehx_ctx *, const ehx_trie *, const ehx_trie *, uint32_t, int, uint32_t *, uint64_t, uint32_t *, int);
int trie_orient(ehx_ctx *, const ehx_trie *, int, ehx_trie **, ehx_trie **, uint64_t *);
int rank_recursive(ehx_ctx *, const ehx_trie *, const uint32_t *, const float *, uint64_t, const ehx_rank_rule *, int,
                   float *, double *);
int join_materialize(ehx_ctx *, const ehx_plan *, ehx_trie *const *, int, uint32_t *const *, uint64_t, uint64_t *,
                     ehx_join_stats *);
int cycle4_count_segments(ehx_ctx *, const ehx_trie *, const uint32_t *, uint64_t, int64_t *, int64_t *,
                          ehx_join_stats *);
int intersect_count_noalloc(ehx_ctx *, const ehx_trie *, const uint32_t *, const ehx_trie *, const uint32_t *, uint64_t,
                            uint64_t *);
int intersect_pairs(ehx_ctx *, const ehx_trie *, const uint32_t *, const ehx_trie *, const uint32_t *, uint64_t,
                    uint64_t *, uint64_t *, uint32_t *, uint64_t, uint64_t *);
int intersect_nodes(ehx_ctx *, int, ehx_trie *const *, const uint32_t *const *, uint64_t, uint64_t *, uint64_t *,
                    uint32_t *, uint64_t, uint64_t *);
void tuned_forget_graphs();
void query_graph_forget();
int selector_load(int device, const uint16_t *params, const float *norm, float max_words);
int selector_clear();
int copy_calibrate(ehx_ctx *ctx, uint64_t bytes, double *ms);
int pair_record(ehx_ctx *ctx, const ehx_trie *a, const ehx_trie *b, const uint32_t *nodes0, const uint32_t *nodes1,
                uint64_t n, int reps, uint64_t *cycles);

}  // namespace ehx

using namespace ehx;

extern "C" {

int ehx_version(void) { return 1; }

const char *ehx_last_error(void) { return t_err.c_str(); }

int ehx_device_count(int *count) {
  if (!count) return fail(EHX_ERR_INVALID, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return EHX_OK;
}

int ehx_ctx_create(int device, ehx_ctx **out) {
  if (!out) return fail(EHX_ERR_INVALID, "null out");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(EHX_ERR_DEVICE, "no GPU visible to HIP");
  if (device < 0 || device >= n) return fail(EHX_ERR_INVALID, "device %d of %d", device, n);
  EHX_HIP(hipSetDevice(device));
  ehx_ctx *c = new ehx_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(EHX_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e)); }
  for (auto &ev : c->ev) {
    e = hipEventCreate(&ev);
    if (e != hipSuccess) { delete c; return fail(EHX_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(e)); }
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->num_cus = prop.multiProcessorCount;
  *out = c;
  return EHX_OK;
}

int ehx_ctx_destroy(ehx_ctx *ctx) {
  if (!ctx) return EHX_OK;
  EHX_HIP(hipSetDevice(ctx->device));
  EHX_HIP(hipStreamSynchronize(ctx->stream));
  tuned_forget_graphs();
  query_graph_forget();
  for (auto &b : ctx->scratch) b.release();
  upload_release(ctx);
  for (auto &ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return EHX_OK;
}

int ehx_ctx_synchronize(ehx_ctx *ctx) {
  if (!ctx) return fail(EHX_ERR_INVALID, "null ctx");
  EHX_HIP(hipSetDevice(ctx->device));
  EHX_HIP(hipStreamSynchronize(ctx->stream));
  return EHX_OK;
}

int ehx_trie_build(ehx_ctx *ctx, const uint32_t *c0, const uint32_t *c1, uint64_t rows, int policy, int on_device,
                   ehx_trie **out) {
  if (!c0 || !c1) return fail(EHX_ERR_INVALID, "ehx_trie_build: null column");
  return trie_build(ctx, c0, c1, rows, policy, on_device, out);
}

int ehx_trie_destroy(ehx_trie *t) { return trie_destroy(t); }

int ehx_trie_info(const ehx_trie *t, ehx_trie_meta *out) {
  if (!t || !out) return fail(EHX_ERR_INVALID, "null argument");
  *out = t->info;
  return EHX_OK;
}

int ehx_trie_export(ehx_ctx *ctx, const ehx_trie *t, uint32_t *c0, uint32_t *c1) {
  if (!ctx || !t || !c0 || !c1) return fail(EHX_ERR_INVALID, "null argument");
  return trie_export(ctx, t, c0, c1);
}

int ehx_trie_node_info(ehx_ctx *ctx, const ehx_trie *t, uint8_t *types, uint32_t *cards) {
  if (!ctx || !t) return fail(EHX_ERR_INVALID, "null argument");
  return trie_node_info(ctx, t, types, cards);
}

int ehx_trie_lookup(ehx_ctx *ctx, const ehx_trie *t, const uint32_t *keys, uint64_t n, int64_t *nodes) {
  if (!ctx || !t || (n && (!keys || !nodes))) return fail(EHX_ERR_INVALID, "null argument");
  return trie_lookup(ctx, t, keys, n, nodes);
}

int ehx_trie_device_arrays(const ehx_trie *t, const uint32_t **head_vals, const uint32_t **offsets,
                           const uint32_t **values) {
  if (!t) return fail(EHX_ERR_INVALID, "null trie");
  if (head_vals) *head_vals = t->head;
  if (offsets) *offsets = t->off;
  if (values) *values = t->vals;
  return EHX_OK;
}

int ehx_host_alloc(uint64_t bytes, void **out) {
  if (!out) return fail(EHX_ERR_INVALID, "null argument");
  *out = nullptr;
  if (!bytes) return EHX_OK;
  void *p = nullptr;
  const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(EHX_ERR_NOMEM, "hipHostMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
  }
  *out = p;
  return EHX_OK;
}

int ehx_host_free(void *p) {
  if (p) EHX_HIP(hipHostFree(p));
  return EHX_OK;
}

int ehx_intersect_count_batch_noalloc(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a, const ehx_trie *tb,
                                      const uint32_t *keys_b, uint64_t n, uint64_t *counts) {
  return intersect_count_noalloc(ctx, ta, keys_a, tb, keys_b, n, counts);
}

int ehx_intersect_count_batch(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a, const ehx_trie *tb,
                              const uint32_t *keys_b, uint64_t n, uint64_t *counts) {
  if (!counts && n) return fail(EHX_ERR_INVALID, "null counts");
  return intersect_pairs(ctx, ta, keys_a, tb, keys_b, n, counts, nullptr, nullptr, 0, nullptr);
}

int ehx_intersect_batch(ehx_ctx *ctx, const ehx_trie *ta, const uint32_t *keys_a, const ehx_trie *tb,
                        const uint32_t *keys_b, uint64_t n, uint64_t *offsets, uint32_t *values, uint64_t cap,
                        uint64_t *total) {
  if (!offsets || (!values && cap)) return fail(EHX_ERR_INVALID, "null output");
  return intersect_pairs(ctx, ta, keys_a, tb, keys_b, n, nullptr, offsets, values, cap, total);
}

int ehx_intersect_nodes_count(ehx_ctx *ctx, int nsets, ehx_trie *const *tries, const uint32_t *const *nodes,
                              uint64_t n, uint64_t *counts) {
  if (!counts && n) return fail(EHX_ERR_INVALID, "null counts");
  return intersect_nodes(ctx, nsets, tries, nodes, n, counts, nullptr, nullptr, 0, nullptr);
}

int ehx_intersect_nodes(ehx_ctx *ctx, int nsets, ehx_trie *const *tries, const uint32_t *const *nodes, uint64_t n,
                        uint64_t *offsets, uint32_t *values, uint64_t cap, uint64_t *total) {
  if (!offsets || (!values && cap)) return fail(EHX_ERR_INVALID, "null output");
  return intersect_nodes(ctx, nsets, tries, nodes, n, nullptr, offsets, values, cap, total);
}

int ehx_trie_host_arrays(ehx_ctx *ctx, const ehx_trie *t, uint32_t *head_vals, uint32_t *offsets) {
  if (!ctx || !t) return fail(EHX_ERR_INVALID, "null argument");
  EHX_HIP(hipSetDevice(ctx->device));
  int rc;
  if (head_vals && (rc = copy_d2h(ctx, head_vals, t->head, t->nh * 4))) return rc;
  if (offsets && (rc = copy_d2h(ctx, offsets, t->off, (t->nh + 1) * 4))) return rc;
  return EHX_OK;
}

int ehx_join_count(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form, int64_t *result,
                   ehx_join_stats *stats) {
  if (!ctx) return fail(EHX_ERR_INVALID, "null ctx");
  return join_count(ctx, plan, tries, form, 0, UINT64_MAX, result, stats, nullptr, nullptr);
}

int ehx_join_count_range(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form, uint64_t begin,
                         uint64_t end, int64_t *result, ehx_join_stats *stats) {
  if (!ctx) return fail(EHX_ERR_INVALID, "null ctx");
  return join_count(ctx, plan, tries, form, begin, end, result, stats, nullptr, nullptr);
}

int ehx_join_count_segments(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form,
                            const uint32_t *bounds, uint64_t nseg, int64_t *counts, int64_t *result,
                            ehx_join_stats *stats) {
  return join_count_segments(ctx, plan, tries, form, bounds, nseg, counts, result, stats);
}

int ehx_join_materialize(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int form, uint32_t *const *cols,
                         uint64_t cap, uint64_t *rows, ehx_join_stats *stats) {
  return join_materialize(ctx, plan, tries, form, cols, cap, rows, stats);
}

int ehx_weighted_pairs(ehx_ctx *ctx, const ehx_trie *t, const int64_t *wa, uint64_t na, const int64_t *wb,
                       uint64_t nb, int64_t *result) {
  if (!ctx || !t || !result || (na && !wa) || (nb && !wb)) return fail(EHX_ERR_INVALID, "null argument");
  if (t->device != ctx->device) return fail(EHX_ERR_INVALID, "trie lives on another device");
  return weighted_pairs(ctx, t, wa, na, wb, nb, result);
}

int ehx_sssp_unweighted(ehx_ctx *ctx, const ehx_trie *t, uint32_t start, int32_t iterations, uint32_t *dist_out,
                        uint64_t nd, uint32_t *levels) {
  if (!ctx || !t || !dist_out) return fail(EHX_ERR_INVALID, "null argument");
  if (t->device != ctx->device) return fail(EHX_ERR_INVALID, "trie lives on another device");
  return sssp_unweighted(ctx, t, start, iterations, dist_out, nd, levels);
}

int ehx_trie_orient(ehx_ctx *ctx, const ehx_trie *t, int policy, ehx_trie **lower, ehx_trie **reverse,
                    uint64_t *counts) {
  if (!ctx || !t) return fail(EHX_ERR_INVALID, "null argument");
  if (t->device != ctx->device) return fail(EHX_ERR_INVALID, "trie lives on another device");
  return trie_orient(ctx, t, policy, lower, reverse, counts);
}

int ehx_trie_link(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, int *linked) {
  if (!ctx || !src || !dst) return fail(EHX_ERR_INVALID, "null argument");
  if (src->device != ctx->device || dst->device != ctx->device)
    return fail(EHX_ERR_INVALID, "trie lives on another device");
  EHX_HIP(hipSetDevice(ctx->device));
  const uint4 *arr = nullptr;
  int rc = trie_link(ctx, src, dst, &arr);
  if (linked) *linked = arr != nullptr;
  return rc;
}

int ehx_trie_link_hints(ehx_ctx *ctx, ehx_trie *src, const ehx_trie *dst, int32_t route, uint32_t *hints_out,
                        uint64_t n, int *state) {
  if (!ctx || !src || !dst || (n && !hints_out) || !state) return fail(EHX_ERR_INVALID, "null argument");
  if (src->device != ctx->device || dst->device != ctx->device)
    return fail(EHX_ERR_INVALID, "trie lives on another device");
  EHX_HIP(hipSetDevice(ctx->device));
  return trie_link_hints(ctx, src, dst, route, hints_out, n, state);
}

int ehx_sssp_unweighted_dir(ehx_ctx *ctx, const ehx_trie *t, const ehx_trie *rev, uint32_t start, int32_t iterations,
                            uint32_t *dist_out, uint64_t nd, uint32_t *levels, int32_t bottom_up) {
  if (!ctx || !t || !rev || !dist_out) return fail(EHX_ERR_INVALID, "null argument");
  if (t->device != ctx->device || rev->device != ctx->device)
    return fail(EHX_ERR_INVALID, "trie lives on another device");
  if (bottom_up < 0) return fail(EHX_ERR_INVALID, "bottom_up %d", bottom_up);
  return sssp_directed(ctx, t, rev, start, iterations, dist_out, nd, levels, bottom_up);
}

int ehx_pagerank(ehx_ctx *ctx, const ehx_trie *t, const uint32_t *w_keys, const float *w_vals, uint64_t n_w,
                 const ehx_rank_rule *rule, int form, float *rank_out, double *ms) {
  if (!ctx || !t || !rule || !rank_out) return fail(EHX_ERR_INVALID, "null argument");
  if (t->device != ctx->device) return fail(EHX_ERR_INVALID, "trie lives on another device");
  return rank_recursive(ctx, t, w_keys, w_vals, n_w, rule, form, rank_out, ms);
}

int ehx_cycle4_count(ehx_ctx *ctx, const ehx_trie *sym, uint64_t begin, uint64_t end, int64_t *result,
                     ehx_join_stats *stats) {
  return cycle4_count(ctx, sym, begin, end, nullptr, result, stats);
}

int ehx_cycle4_count_segments(ehx_ctx *ctx, const ehx_trie *sym, const uint32_t *bounds, uint64_t nseg,
                              int64_t *counts, int64_t *result, ehx_join_stats *stats) {
  return cycle4_count_segments(ctx, sym, bounds, nseg, counts, result, stats);
}

int ehx_join_record(ehx_ctx *ctx, const ehx_plan *plan, ehx_trie *const *tries, int force_mode, uint64_t *cycles,
                    uint64_t cap, uint64_t *nprefix, int64_t *result) {
  if (!ctx || !cycles || !nprefix) return fail(EHX_ERR_INVALID, "null argument");
  if (force_mode < -1 || force_mode > 1) return fail(EHX_ERR_INVALID, "force_mode must be -1, 0 or 1");
  RecordReq r{cycles, cap, nprefix, force_mode};
  return join_count(ctx, plan, tries, EHX_FORM_TUNED, 0, UINT64_MAX, result, nullptr, &r, nullptr);
}

int ehx_tune_get(ehx_tune_params *out) {
  if (!out) return fail(EHX_ERR_INVALID, "null out");
  const TuneParams &p = tune_params();
  out->lds_slots = p.lds_slots;
  out->big_deg = p.big_deg;
  out->waves_per_cu = p.waves_per_cu;
  out->grab = p.grab;
  out->bitmap_all = p.bitmap_all;
  out->use_graph = p.use_graph;
  out->big_work = p.big_work;
  out->stream_lanes = p.stream_lanes;
  out->gap_coded = p.gap_coded;
  out->split_work = p.split_work;
  out->tiny_work = p.tiny_work;
  return EHX_OK;
}

int ehx_tune_set(const ehx_tune_params *in) {
  if (!in) return fail(EHX_ERR_INVALID, "null params");
  if (in->lds_slots != 2048 && in->lds_slots != 4096 && in->lds_slots != 8192)
    return fail(EHX_ERR_INVALID, "lds_slots must be 2048, 4096 or 8192");
  if (in->big_deg < 1 || in->waves_per_cu < 1 || in->waves_per_cu > 64 || in->grab < 1 || in->grab > 64 ||
      in->big_work < 1 || in->split_work < 0 || in->tiny_work < 0)
    return fail(EHX_ERR_INVALID, "tune parameter out of range");
  if (in->stream_lanes < 2 || in->stream_lanes > 64 || (in->stream_lanes & (in->stream_lanes - 1)))
    return fail(EHX_ERR_INVALID, "stream_lanes must be a power of two in [2, 64]");
  TuneParams &p = tune_params();
  p.lds_slots = in->lds_slots;
  p.big_deg = in->big_deg;
  p.waves_per_cu = in->waves_per_cu;
  p.grab = in->grab;
  p.bitmap_all = in->bitmap_all ? 1 : 0;
  p.use_graph = in->use_graph ? 1 : 0;
  p.big_work = in->big_work;
  p.stream_lanes = in->stream_lanes;
  p.gap_coded = in->gap_coded ? 1 : 0;
  p.split_work = in->split_work;
  p.tiny_work = in->tiny_work;
  return EHX_OK;
}

int ehx_selector_load(int device, const uint16_t *params, uint64_t nparams, const float *norm, float max_words) {
  if (!params || !norm) return fail(EHX_ERR_INVALID, "null weights");
  if (nparams != (uint64_t)ehx_selector_num_params()) return fail(EHX_ERR_INVALID, "selector expects %d params", ehx_selector_num_params());
  return selector_load(device, params, norm, max_words);
}

int ehx_selector_clear(void) {
  int rc = selector_clear();
  int rc2 = pair_table_load(0, nullptr);
  return rc ? rc : rc2;
}

int ehx_pair_record(ehx_ctx *ctx, const ehx_trie *a, const ehx_trie *b, const uint32_t *nodes0, const uint32_t *nodes1,
                    uint64_t n, int reps, uint64_t *cycles) {
  return pair_record(ctx, a, b, nodes0, nodes1, n, reps, cycles);
}

int ehx_pair_table_load(int device, const uint8_t *table) { return pair_table_load(device, table); }

int ehx_copy_calibrate(ehx_ctx *ctx, uint64_t bytes, double *ms) {
  if (!ctx || bytes < 16) return fail(EHX_ERR_INVALID, "bad copy size");
  EHX_HIP(hipSetDevice(ctx->device));
  return copy_calibrate(ctx, bytes, ms);
}

}  // extern "C"
