// dropin.cpp -- runtime behind the drop-in storage-engine API.
//
// Tries are device tries of libehx with a host mirror of the head set (values,
// child offsets, dense value -> node array) so iterators can navigate on the
// host.  Builders record their build_aggregated_set calls per level and chunk
// of level-0 values, and each recorded level is computed in one batched GPU
// launch (ehx_intersect_nodes[_count]); results are replayed to the generated
// code's lambdas in their original call order.
#include <stdio.h>
#include <stdlib.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <memory>
#include <fstream>
#include <mutex>
#include <thread>

#include "../../../include/ehx.h"
#include "../../../include/ehx_ext.h"
#include "../include/emptyheaded.hpp"

typedef TrieBlock<hybrid, ParMemoryBuffer> Block;

namespace ehx_dropin {

struct HostTrie : HostNav {
  ehx_trie *dev = nullptr;
  uint64_t nh = 0, m = 0;
  int ncols = 2;  // 1: a unary relation, kept as the device trie of (x, x)
  std::vector<uint32_t> head, off;
  std::vector<int32_t> dense;
  std::vector<Block> blocks;         // nh + 1 (the last is the head block)
  std::vector<uint8_t> node_layout;  // nh + 1: each block's set layout (type::layout), read without the block
  // annotated relations: one annotation per row in trie order (per head for a
  // unary relation), and the child values on the host to find a row's position
  bool annotated = false;
  std::vector<uint64_t> annos;
  std::vector<uint32_t> child_vals;
  // A relation of three or more attributes is a chain of binary tries: this
  // one holds its first two levels (or, further down, one level keyed by the
  // row position of the prefix in the trie above), `next` the level below,
  // keyed by the position of each row of this trie (its index in `vals`).
  HostTrie *next = nullptr;
};

static std::mutex g_mu;
static ehx_ctx *g_ctx = nullptr;

[[noreturn]] void not_implemented(const char *what) {
  std::cerr << "ERROR: NOT IMPLEMENTED: " << what << std::endl;
  abort();
}

static void check(int rc, const char *what) {
  if (rc != EHX_OK) {
    std::cerr << "ehx error " << rc << " in " << what << ": " << ehx_last_error() << std::endl;
    abort();
  }
}

static ehx_ctx *ctx() {
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ctx) check(ehx_ctx_create(0, &g_ctx), "ehx_ctx_create");
  return g_ctx;
}

HostTrie *build_host_trie(const uint32_t *c0, const uint32_t *c1, size_t rows) {
  HostTrie *t = new HostTrie();
  check(ehx_trie_build(ctx(), c0, c1, rows, EHX_POLICY_RULE, 0, &t->dev), "ehx_trie_build");
  ehx_trie_meta meta;
  check(ehx_trie_info(t->dev, &meta), "ehx_trie_info");
  t->nh = meta.num_heads;
  t->m = meta.num_rows;
  t->head.resize(t->nh);
  t->off.resize(t->nh + 1);
  check(ehx_trie_host_arrays(ctx(), t->dev, t->head.data(), t->off.data()), "ehx_trie_host_arrays");
  std::vector<uint8_t> types(t->nh);
  std::vector<uint32_t> cards(t->nh);
  check(ehx_trie_node_info(ctx(), t->dev, types.data(), cards.data()), "ehx_trie_node_info");
  // dense navigation when the head is dense by (cardinality, max): TrieBlock.hpp:26-35
  const uint32_t maxh = t->nh ? t->head.back() : 0;
  if (t->nh && !common::is_sparse(t->nh, maxh)) {
    t->dense.assign((size_t)maxh + 1, -1);
    for (uint64_t h = 0; h < t->nh; h++) t->dense[t->head[h]] = (int32_t)h;
  }
  t->blocks.resize(t->nh + 1);
  HostNav &nv = *t;
  nv.dense = t->dense.empty() ? nullptr : t->dense.data();
  nv.dense_len = t->dense.size();
  nv.head = t->head.data();
  nv.nh = (uint32_t)t->nh;
  nv.blocks = (char *)t->blocks.data();
  if ((const void *)static_cast<const HostNav *>(t) != (const void *)t) not_implemented("host trie layout");
  for (uint64_t h = 0; h <= t->nh; h++) {
    Block &b = t->blocks[h];
    b.trie = t;
    b.node = (uint32_t)h;
    if (h < t->nh) {
      b.set.cardinality = t->off[h + 1] - t->off[h];
      b.set.type = types[h] == EHX_RANGE_BITSET ? type::RANGE_BITSET : type::UINTEGER;
      b.set.number_of_bytes = 4ull * b.set.cardinality;
    } else {
      b.set.cardinality = (uint32_t)t->nh;
      b.set.range = maxh;
      b.set.values = t->head.data();
      b.set.type = meta.head_layout == EHX_RANGE_BITSET ? type::RANGE_BITSET : type::UINTEGER;
      b.set.number_of_bytes = 4ull * t->nh;
    }
  }
  t->node_layout.resize(t->nh + 1);
  for (uint64_t h = 0; h <= t->nh; h++) t->node_layout[h] = (uint8_t)t->blocks[h].set.type;
  return t;
}

// Rows sorted lexicographically and deduplicated with their annotations (the
// last of equal rows wins); a unary relation is the trie of (x, x).
HostTrie *build_annotated_trie(int ncols, const uint32_t *c0, const uint32_t *c1, const uint64_t *annos,
                               size_t rows) {
  if (ncols != 1 && ncols != 2) not_implemented("tries over relations of arity other than 1 or 2");
  const uint32_t *second = ncols == 1 ? c0 : c1;
  std::vector<uint64_t> idx(rows);
  for (size_t i = 0; i < rows; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
    return c0[a] != c0[b] ? c0[a] < c0[b] : second[a] < second[b];
  });
  std::vector<uint32_t> s0, s1;
  std::vector<uint64_t> an;
  s0.reserve(rows);
  s1.reserve(rows);
  for (size_t k = 0; k < rows; k++) {
    const uint64_t i = idx[k];
    if (!s0.empty() && s0.back() == c0[i] && s1.back() == second[i]) {
      if (annos) an.back() = annos[i];
      continue;
    }
    s0.push_back(c0[i]);
    s1.push_back(second[i]);
    if (annos) an.push_back(annos[i]);
  }
  HostTrie *t = build_host_trie(s0.data(), s1.data(), s0.size());
  t->ncols = ncols;
  if (annos) {
    t->annotated = true;
    t->annos = std::move(an);
    if (ncols == 2) t->child_vals = std::move(s1);
  }
  return t;
}

static const HostTrie *last_level(const HostTrie *t) {
  while (t && t->next) t = t->next;
  return t;
}

// Rows of any arity from columns: sorted lexicographically, equal rows merged
// (the last annotation wins).  Arity 1 and 2 are one device trie; arity k >= 3
// a chain of k-1 device tries: level 0 over the distinct (c0, c1) pairs, level
// j over (position of the (c0..cj) prefix in level j-1, c(j+1)).  Every level
// keeps its child values on the host (navigation, and sets the generated code
// iterates), the last one the annotations.
HostTrie *build_rows_trie(int ncols, const uint32_t *const *cols, size_t rows, const uint64_t *annos) {
  if (ncols < 1 || ncols > 8) not_implemented("tries over relations of arity above 8");
  if (ncols <= 2) {
    if (ncols == 2 && !annos) return build_host_trie(cols[0], cols[1], rows);
    return build_annotated_trie(ncols, cols[0], ncols == 2 ? cols[1] : nullptr, annos, rows);
  }
  std::vector<uint64_t> idx(rows);
  for (size_t i = 0; i < rows; i++) idx[i] = i;
  auto less = [&](uint64_t a, uint64_t b) {
    for (int c = 0; c < ncols; c++)
      if (cols[c][a] != cols[c][b]) return cols[c][a] < cols[c][b];
    return false;
  };
  std::stable_sort(idx.begin(), idx.end(), less);
  std::vector<std::vector<uint32_t>> C(ncols);
  std::vector<uint64_t> an;
  for (size_t k = 0; k < rows; k++) {
    const uint64_t i = idx[k];
    if (k && !less(idx[k - 1], i)) {  // equal to the previous row
      if (annos) an.back() = annos[i];
      continue;
    }
    for (int c = 0; c < ncols; c++) C[c].push_back(cols[c][i]);
    if (annos) an.push_back(annos[i]);
  }
  const size_t n = C[0].size();
  std::vector<uint32_t> pos(n), npos(n);
  HostTrie *first = nullptr, *prev = nullptr;
  for (int lv = 0; lv + 1 < ncols; lv++) {
    std::vector<uint32_t> hk, vk;
    hk.reserve(n);
    vk.reserve(n);
    for (size_t r = 0; r < n; r++) {
      const uint32_t key = lv == 0 ? C[0][r] : pos[r], v = C[lv + 1][r];
      if (hk.empty() || hk.back() != key || vk.back() != v) {
        hk.push_back(key);
        vk.push_back(v);
      }
      npos[r] = (uint32_t)(hk.size() - 1);
    }
    HostTrie *t = build_host_trie(hk.data(), vk.data(), hk.size());
    t->ncols = ncols;
    t->child_vals = std::move(vk);
    for (uint64_t h = 0; h < t->nh; h++) {
      Block &b = t->blocks[h];
      b.set.values = t->child_vals.data() + t->off[h];
      b.set.range = b.set.cardinality ? b.set.values[b.set.cardinality - 1] : 0;
    }
    if (!first) first = t;
    else prev->next = t;
    prev = t;
    pos.swap(npos);
  }
  if (annos) {
    prev->annotated = true;
    prev->annos = std::move(an);
  }
  return first;
}

bool host_trie_annotated(const HostTrie *t) { return last_level(t)->annotated; }
const std::vector<uint64_t> *host_trie_annos(const HostTrie *t) {
  const HostTrie *l = last_level(t);
  return l->annotated ? &l->annos : nullptr;
}
int host_trie_columns(const HostTrie *t) { return t->ncols; }

// The block of `data` in the set of non-head block (t, node): the set one level
// down, found through the row position of `data` (NULL when absent or when
// the set is a relation's last level).
void *child_of(const HostTrie *t, uint32_t node, uint32_t data) {
  if (!t->next) return nullptr;
  const uint32_t lo = t->off[node], hi = t->off[node + 1];
  auto it = std::lower_bound(t->child_vals.begin() + lo, t->child_vals.begin() + hi, data);
  if (it == t->child_vals.begin() + hi || *it != data) return nullptr;
  const long h = head_node(t->next, (uint32_t)(it - t->child_vals.begin()));
  return h < 0 ? nullptr : block_of(t->next, (uint32_t)h);
}

bool block_contains(const HostTrie *t, uint32_t node, uint32_t data) {
  if (node == t->nh) return head_node(t, data) >= 0;
  if (t->child_vals.empty()) not_implemented("membership test on a child set whose values are on the GPU");
  const uint32_t lo = t->off[node], hi = t->off[node + 1];
  return std::binary_search(t->child_vals.begin() + lo, t->child_vals.begin() + hi, data);
}

// Every row in lexicographic order (row-major), with its annotation.
void host_trie_all_rows(const HostTrie *t, std::vector<uint32_t> &rows, std::vector<uint64_t> *annos) {
  rows.clear();
  if (annos) annos->clear();
  if (t->ncols <= 2) {
    std::vector<uint32_t> c0, c1;
    host_trie_tuples(t, c0, c1);
    const std::vector<uint64_t> *an = host_trie_annos(t);
    for (size_t i = 0; i < c0.size(); i++) {
      rows.push_back(c0[i]);
      if (t->ncols == 2) rows.push_back(c1[i]);
      if (annos) annos->push_back(an ? an->at(i) : 0);
    }
    return;
  }
  const int nc = t->ncols;
  std::vector<uint32_t> prefix(nc);
  // depth-first over the chain: (trie, node) at level lv covers prefix[0..lv]
  std::function<void(const HostTrie *, uint32_t, int)> walk = [&](const HostTrie *L, uint32_t node, int lv) {
    for (uint32_t p = L->off[node]; p < L->off[node + 1]; p++) {
      prefix[lv + 1] = L->child_vals[p];
      if (!L->next) {
        rows.insert(rows.end(), prefix.begin(), prefix.end());
        if (annos) annos->push_back(L->annotated ? L->annos[p] : 0);
        continue;
      }
      const long h = head_node(L->next, p);
      if (h >= 0) walk(L->next, (uint32_t)h, lv + 1);
    }
  };
  for (uint32_t h = 0; h < t->nh; h++) {
    prefix[0] = t->head[h];
    walk(t, h, 0);
  }
}

// TrieIterator::get_annotation (src/TrieIterator.cpp:44-59): the annotation of
// `data` in the set of block `b`, 0 when absent.
uint64_t annotation_of(const void *block, uint32_t data) {
  const Block *b = (const Block *)block;
  if (!b || !b->trie) return 0;
  const HostTrie *t = b->trie;
  if (t->next) not_implemented("annotations above a relation's last level");
  if (!t->annotated) not_implemented("TrieIterator::get_annotation on an unannotated relation");
  if (b->node == t->nh) {  // the head set
    if (t->ncols != 1) not_implemented("annotations above a relation's last level");
    const long h = head_node(t, data);
    return h < 0 ? 0 : t->annos[h];
  }
  const uint32_t lo = t->off[b->node], hi = t->off[b->node + 1];
  auto it = std::lower_bound(t->child_vals.begin() + lo, t->child_vals.begin() + hi, data);
  if (it == t->child_vals.begin() + hi || *it != data) return 0;
  return t->annos[it - t->child_vals.begin()];
}

static std::string rows_file(const std::string &path) { return path + "/ehx_rows.bin"; }

static bool exists(const std::string &f) {
  std::ifstream in(f, std::ios::binary);
  return (bool)in;
}

HostTrie *load_host_trie(const std::string &path, int anno_bytes) {
  if (exists(path + "/ram/trieinfo.bin") || exists(path + "/mmap/trieinfo.bin")) {
    // a relation the reference's loader saved (Trie<A,M>::load, src/Trie.cpp:37-75)
    ehx_refdb *db = nullptr;
    check(ehx_refdb_open(path.c_str(), anno_bytes, &db), "ehx_refdb_open");
    uint32_t nc = 0;
    uint64_t rows = 0;
    int32_t annotated = 0;
    check(ehx_refdb_info(db, &nc, &rows, &annotated), "ehx_refdb_info");
    if (nc < 1 || nc > 8) not_implemented("tries over relations of arity above 8");
    std::vector<std::vector<uint32_t>> c(nc, std::vector<uint32_t>(rows));
    std::vector<uint64_t> an(annotated ? rows : 0);
    std::vector<uint32_t *> cols(nc);
    std::vector<const uint32_t *> ccols(nc);
    for (uint32_t k = 0; k < nc; k++) ccols[k] = cols[k] = c[k].data();
    check(ehx_refdb_columns(db, cols.data(), annotated ? an.data() : nullptr), "ehx_refdb_columns");
    ehx_refdb_close(db);
    return build_rows_trie((int)nc, ccols.data(), rows, annotated ? an.data() : nullptr);
  }
  const size_t slash = path.find_last_of('/');
  const std::string parent = slash == std::string::npos ? "." : path.substr(0, slash);
  const std::string leaf = slash == std::string::npos ? path : path.substr(slash + 1);
  if (!exists(rows_file(path)) && exists(parent + "/encoded.bin")) {
    // only the relation's encoded column store: build this ordering's trie from
    // it, as the reference's loader does per ordering (createDB.scala:240-275);
    // the ordering is the folder name's suffix, <relation>_<i>_<j>
    ehx_refdb *db = nullptr;
    check(ehx_encoded_open(parent.c_str(), &db), "ehx_encoded_open");
    uint32_t nc = 0;
    uint64_t rows = 0;
    int32_t annotated = 0;
    check(ehx_refdb_info(db, &nc, &rows, &annotated), "ehx_refdb_info");
    if (nc < 1 || nc > 8) not_implemented("tries over relations of arity above 8");
    std::vector<std::vector<uint32_t>> c(nc, std::vector<uint32_t>(rows));
    std::vector<uint64_t> an(annotated ? rows : 0);
    // the ordering: the folder name's suffix _<i0>_<i1>..., attribute i_k on level k
    std::vector<int> order(nc, 0);
    if (nc > 1) {
      size_t end = leaf.size();
      for (int k = (int)nc - 1; k >= 0; k--) {
        const size_t u = end == 0 ? std::string::npos : leaf.find_last_of('_', end - 1);
        if (u == std::string::npos) not_implemented("relation folder without an _<i>_<j>... ordering suffix");
        order[k] = atoi(leaf.substr(u + 1, end - u - 1).c_str());
        end = u;
      }
      std::vector<int> seen(nc, 0);
      for (int k = 0; k < (int)nc; k++) {
        if (order[k] < 0 || order[k] >= (int)nc || seen[order[k]]++)
          not_implemented("relation folder whose ordering suffix is not a permutation of its attributes");
      }
    }
    // column j of the store lands on the level that holds attribute j
    std::vector<uint32_t *> cols(nc);
    std::vector<const uint32_t *> levels(nc);
    for (uint32_t k = 0; k < nc; k++) {
      cols[order[k]] = c[k].data();
      levels[k] = c[k].data();
    }
    check(ehx_refdb_columns(db, cols.data(), annotated ? an.data() : nullptr), "ehx_refdb_columns");
    ehx_refdb_close(db);
    return build_rows_trie((int)nc, levels.data(), rows, annotated ? an.data() : nullptr);
  }
  std::ifstream in(rows_file(path), std::ios::binary);
  if (!in) {
    std::cerr << "ERROR: cannot open " << rows_file(path) << std::endl;
    abort();
  }
  uint64_t rows = 0;
  in.read((char *)&rows, 8);
  std::vector<uint32_t> c0(rows), c1(rows);
  in.read((char *)c0.data(), 4 * rows);
  in.read((char *)c1.data(), 4 * rows);
  if (!in) {
    std::cerr << "ERROR: truncated " << rows_file(path) << std::endl;
    abort();
  }
  return build_host_trie(c0.data(), c1.data(), rows);
}

void host_trie_tuples(const HostTrie *t, std::vector<uint32_t> &c0, std::vector<uint32_t> &c1) {
  c0.resize(t->m);
  c1.resize(t->m);
  check(ehx_trie_export(ctx(), t->dev, c0.data(), c1.data()), "ehx_trie_export");
}

void save_host_trie(const HostTrie *t, const std::string &path) {
  if (t->ncols > 2) not_implemented("Trie::save of a relation of arity above 2");
  std::vector<uint32_t> c0, c1;
  host_trie_tuples(t, c0, c1);
  std::ofstream out(rows_file(path), std::ios::binary);
  uint64_t rows = c0.size();
  out.write((const char *)&rows, 8);
  out.write((const char *)c0.data(), 4 * rows);
  out.write((const char *)c1.data(), 4 * rows);
}

void free_host_trie(HostTrie *t) {
  while (t) {
    HostTrie *n = t->next;
    if (t->dev) ehx_trie_destroy(t->dev);
    delete t;
    t = n;
  }
}

uint32_t single_source_levels(HostTrie *t, uint32_t start, int iterations, size_t num_distinct,
                              std::vector<uint32_t> &dist) {
  ehx_trie_meta meta;
  size_t nd = num_distinct;
  if (t && t->dev) {
    check(ehx_trie_info(t->dev, &meta), "ehx_trie_info");
    if (t->m && (size_t)meta.max_value + 1 > nd) nd = (size_t)meta.max_value + 1;
  }
  if ((size_t)start + 1 > nd) nd = (size_t)start + 1;
  dist.assign(nd, 0xFFFFFFFFu);
  if (!t || !t->dev || !t->m) {
    dist[start] = 0;
    return 0;
  }
  uint32_t levels = 0;
  uint64_t c[4] = {0, 0, 0, 0};
  check(ehx_trie_orient(ctx(), t->dev, EHX_POLICY_RULE, nullptr, nullptr, c), "ehx_trie_orient");
  if (c[0] == c[1] && c[3] == 0)  // symmetric (loops do not matter here): the trie is its own reverse
    check(ehx_sssp_unweighted_dir(ctx(), t->dev, t->dev, start, iterations, dist.data(), nd, &levels, 32),
          "ehx_sssp_unweighted_dir");
  else
    check(ehx_sssp_unweighted(ctx(), t->dev, start, iterations, dist.data(), nd, &levels), "ehx_sssp_unweighted");
  uint32_t deepest = 0;
  for (uint32_t d : dist)
    if (d != 0xFFFFFFFFu && d > deepest) deepest = d;
  return deepest;
}

size_t host_trie_rows(const HostTrie *t) { return last_level(t)->m; }
uint32_t num_heads(const HostTrie *t) { return (uint32_t)t->nh; }
void *block_of(const HostTrie *t, uint32_t node) { return (void *)&t->blocks[node]; }

long head_node(const HostTrie *t, uint32_t data) {
  if (!t->dense.empty()) return data < t->dense.size() ? t->dense[data] : -1;
  auto it = std::lower_bound(t->head.begin(), t->head.end(), data);
  return (it != t->head.end() && *it == data) ? (long)(it - t->head.begin()) : -1;
}

// ---- batched execution -----------------------------------------------------------

static void run_batch(int nops, const HostTrie *const *tries, std::vector<std::vector<uint32_t>> &nodes, bool count_only,
                      std::vector<uint64_t> &counts, std::vector<uint64_t> &offsets, std::vector<uint32_t> &values) {
  const uint64_t n = nodes.empty() ? 0 : nodes[0].size();
  ehx_trie *dt[8];
  const uint32_t *np[8];
  for (int s = 0; s < nops; s++) {
    dt[s] = tries[s]->dev;
    np[s] = nodes[s].data();
  }
  counts.assign(n, 0);
  if (n) check(ehx_intersect_nodes_count(ctx(), nops, dt, np, n, counts.data()), "ehx_intersect_nodes_count");
  if (count_only) return;
  uint64_t total = 0;
  for (uint64_t c : counts) total += c;
  offsets.assign(n + 1, 0);
  values.assign(total ? total : 1, 0);
  uint64_t got = 0;
  if (n) check(ehx_intersect_nodes(ctx(), nops, dt, np, n, offsets.data(), values.data(), total, &got),
               "ehx_intersect_nodes");
}

// Page-locked host memory, grow-only (ehx_host_alloc): the batches of a level
// cross PCIe at the full rate.  Pool buffers are kept for the life of the
// process (one query at a time runs its passes, so one set per builder level).
struct Pinned {
  void *p = nullptr;
  size_t cap = 0;
  template <class T> T *as() const { return (T *)p; }
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) check(ehx_host_free(p), "ehx_host_free");
    p = nullptr;
    cap = 0;
    const size_t want = bytes + bytes / 4;  // headroom: the next query is often a little larger
    check(ehx_host_alloc(want, &p), "ehx_host_alloc");
    cap = want;
  }
};
struct LevelBuffers {
  Pinned nodes, counts, offsets, values;
};
static std::vector<LevelBuffers *> g_level_pool;
static LevelBuffers &level_buffers(size_t level) {
  while (g_level_pool.size() <= level) g_level_pool.push_back(new LevelBuffers());
  return *g_level_pool[level];
}

// Cardinality of a set named by (trie, node): the bound of any intersection
// that takes it as an operand.
static inline uint64_t card_of(const HostTrie *t, uint32_t node) {
  return node == t->nh ? t->nh : (uint64_t)(t->off[node + 1] - t->off[node]);
}

static inline void set_empty(Set<hybrid> *out) {
  *out = Set<hybrid>(0, 0, 0, type::UINTEGER);
}

static inline const Block *blk(const Operands &ops, int i) { return (const Block *)ops.blocks[i]; }

static type::layout result_type(const Operands &ops) {
  // U x U -> U, U x R -> U, R x R -> R (ops/intersection/set.hpp:53-91)
  for (int i = 0; i < ops.n; i++)
    if (blk(ops, i)->set.type != type::RANGE_BITSET) return type::UINTEGER;
  return type::RANGE_BITSET;
}

// Level-0 results live in the storage of the builder that asked for them.
size_t aggregate_head(const Operands &ops, Set<hybrid> *out, std::vector<uint32_t> *storage) {
  for (int i = 0; i < ops.n; i++)
    if (!blk(ops, i)) { set_empty(out); return 0; }
  bool same = true;
  for (int i = 1; i < ops.n; i++) same = same && ops.blocks[i] == ops.blocks[0];
  const Block *b0 = blk(ops, 0);
  if (same && b0->set.values) {  // head ∩ head of one trie is the head itself
    *out = b0->set;
    return out->cardinality;
  }
  std::vector<std::vector<uint32_t>> nodes(ops.n, std::vector<uint32_t>(1));
  const HostTrie *tries[8];
  for (int i = 0; i < ops.n; i++) {
    tries[i] = blk(ops, i)->trie;
    nodes[i][0] = blk(ops, i)->node;
  }
  std::vector<uint64_t> counts, offsets;
  std::vector<uint32_t> values;
  run_batch(ops.n, tries, nodes, false, counts, offsets, values);
  *storage = std::move(values);
  *out = Set<hybrid>((uint32_t)counts[0], counts[0] ? (*storage)[counts[0] - 1] : 0, 4 * counts[0],
                     result_type(ops));
  out->values = storage->data();
  return counts[0];
}

// Calls one chunk of level-0 values made at one builder level.  Chunks run on
// different threads, so each log has its own cache lines.
struct alignas(128) LevelLog {
  std::vector<uint32_t> nodes;  // call-major, nops per call
  uint64_t ncalls = 0, cursor = 0;
  // the chunk's results in its batch's pinned buffers: a count per call
  // (count-only levels), or offsets[ncalls + 1] into `vals` (materialized)
  const uint64_t *res = nullptr;
  const uint32_t *vals = nullptr;
};

struct LevelBatch {
  int nops = 0;
  const HostTrie *tries[8] = {};
  bool materialized = false;
  bool innermost = false;  // innermost aggregated level with no annotated operand
};

// Worker threads kept for one run of passes: run(fn) calls fn(tid) on every
// worker (tid 0 on the calling thread) and returns once all have finished, so
// the many short parallel phases of a run (chunk ranges, gathers) start no
// threads of their own.
class Workers {
 public:
  explicit Workers(size_t n) : n_(n ? n : 1) {
    for (size_t t = 1; t < n_; t++) th_.emplace_back([this, t]() { loop(t); });
  }
  ~Workers() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto &x : th_) x.join();
  }
  size_t size() const { return n_; }
  void run(const std::function<void(size_t)> &fn) {
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> l(mu_);
    done_.wait(l, [this]() { return pending_ == 0; });
  }

 private:
  void loop(size_t tid) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(size_t)> *f;
      {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&]() { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        f = fn_;
      }
      (*f)(tid);
      std::lock_guard<std::mutex> g(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  size_t n_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t)> *fn_ = nullptr;
  size_t pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct Session {
  size_t num_attrs = 0;
  int record_level = -1;
  std::vector<std::vector<LevelLog>> logs;  // [chunk][level]
  std::vector<LevelBatch> batches;          // [level]
  std::unique_ptr<std::atomic<bool>[]> signed_levels;  // [level]: its LevelBatch signature is set
  std::mutex sign_mu;
  OutRows *out = nullptr;                   // materialized output (par_foreach_builder)
  std::vector<std::vector<uint32_t>> rows;  // [chunk] rows emitted by the replay
  std::vector<std::vector<uint64_t>> annos; // [chunk]
};

static Session *g_sess = nullptr;
// the chunk of level-0 values a worker is running (initial-exec: one load, no call)
static thread_local uint64_t t_chunk __attribute__((tls_model("initial-exec"))) = 0;
// 0: no pass running, 1: recording pass, 2: final replay.  Written only by the
// thread driving the passes while no worker runs (workers start after the
// write and are joined before the next), read inline by reducers and builders.
int g_phase = 0;
static const uint64_t CHUNK = 128;  // par::for_range dynamic chunk (parallel.cpp:65-79)


void emit_rows(const uint32_t *prefix, int plen, const Set<hybrid> &leaf) {
  Session *S = g_sess;
  if (!S || S->record_level >= 0 || !S->out) return;
  if (plen + 1 != S->out->ncols) not_implemented("rows of an arity other than the output trie's");
  if (leaf.cardinality && !leaf.values) not_implemented("leaf set without host values");
  std::vector<uint32_t> &R = S->rows[t_chunk];
  for (uint32_t k = 0; k < leaf.cardinality; k++) {
    R.insert(R.end(), prefix, prefix + plen);
    R.push_back(leaf.values[k]);
  }
}

void emit_row(const uint32_t *row, int ncols, uint64_t anno_bits) {
  Session *S = g_sess;
  if (!S || S->record_level >= 0 || !S->out) return;
  if (ncols != S->out->ncols) not_implemented("rows of an arity other than the output trie's");
  S->rows[t_chunk].insert(S->rows[t_chunk].end(), row, row + ncols);
  S->annos[t_chunk].push_back(anno_bits);
}

// Node of block p when p is one of trie t's node blocks (the block itself is
// not read), else -1.
static inline long node_in(const HostTrie *t, const void *p) {
  const uintptr_t base = (uintptr_t)t->blocks.data(), a = (uintptr_t)p;
  if (a < base) return -1;
  const uintptr_t k = (a - base) / sizeof(Block);
  return (k < t->blocks.size() && a == base + k * sizeof(Block)) ? (long)k : -1;
}

// The level's call signature, from its first recorded call.
static void sign(LevelBatch &B, size_t num_attrs, uint32_t level, const Operands &ops, bool materialize) {
  bool any_annotated = false;
  for (int i = 0; i < ops.n; i++) any_annotated = any_annotated || blk(ops, i)->trie->annotated;
  B.nops = ops.n;
  for (int i = 0; i < ops.n; i++) B.tries[i] = blk(ops, i)->trie;
  B.innermost = level + 2 == num_attrs && !any_annotated;
  B.materialized = materialize || !B.innermost;
}

// The level's signature from the first call that records it (double-checked:
// one acquire load per later call).
static inline void sign_once(Session *S, LevelBatch &B, size_t num_attrs, uint32_t level, const Operands &ops,
                             bool materialize) {
  if (S->signed_levels[level].load(std::memory_order_acquire)) return;
  std::lock_guard<std::mutex> g(S->sign_mu);
  if (S->signed_levels[level].load(std::memory_order_relaxed)) return;
  sign(B, num_attrs, level, ops, materialize);
  S->signed_levels[level].store(true, std::memory_order_release);
}

// A recorded call's result in the replay: the values of a materialized level,
// the cardinality of a count-only one.
static inline size_t replay_result(const LevelBatch &B, LevelLog &L, type::layout t, Set<hybrid> *out) {
  const uint64_t idx = L.cursor++;
  if (B.materialized) {
    const uint64_t a = L.res[idx], b = L.res[idx + 1];
    *out = Set<hybrid>((uint32_t)(b - a), b > a ? L.vals[b - 1] : 0, 4 * (b - a), t);
    out->values = L.vals + a;
  } else {
    *out = Set<hybrid>((uint32_t)L.res[idx], 0, 4 * L.res[idx], t);
  }
  return out->cardinality;
}

size_t aggregate(size_t num_attrs, uint32_t level, const Operands &ops, Set<hybrid> *out, bool materialize) {
  Session *S = g_sess;
  if (!S) not_implemented("TrieBuilder::build_aggregated_set / build_set outside par_foreach_aggregate / builder");
  if (ops.n < 1 || ops.n > 8) not_implemented("build_aggregated_set / build_set arity");
  for (int i = 0; i < ops.n; i++)
    if (!blk(ops, i)) { set_empty(out); return 0; }  // a NULL TrieBlock is empty (TrieBuilder.cpp:249-254)
  const bool known = level < S->logs[t_chunk].size();
  const bool recording = S->record_level == (int)level;
  if (known && ops.n >= 2 && (recording || S->record_level < 0 || (int)level < S->record_level)) {
    // Every operand a node block of the trie the level's signature names: the
    // nodes come from the block addresses and the layouts from node_layout, so
    // no block is read (one random host access saved per call).  Anything
    // else takes the general path below.
    LevelLog &L = S->logs[t_chunk][level];
    LevelBatch &B = S->batches[level];
    if (recording) sign_once(S, B, num_attrs, level, ops, materialize);
    uint32_t nodes[8];
    bool fast = B.nops == ops.n && B.materialized == (materialize || !B.innermost);
    for (int i = 0; fast && i < ops.n; i++) {
      const long h = node_in(B.tries[i], ops.blocks[i]);
      fast = h >= 0;
      nodes[i] = (uint32_t)h;
    }
    if (fast) {
      if (recording) {
        L.nodes.insert(L.nodes.end(), nodes, nodes + ops.n);
        L.ncalls++;
        set_empty(out);
        return 0;
      }
      // U x U -> U, U x R -> U, R x R -> R (ops/intersection/set.hpp:53-91).  A
      // count-only result is never iterated (it holds no values), so its
      // layout is not looked up.
      type::layout t = B.materialized ? type::RANGE_BITSET : type::UINTEGER;
      for (int i = 0; B.materialized && i < ops.n; i++)
        if (B.tries[i]->node_layout[nodes[i]] != (uint8_t)type::RANGE_BITSET) t = type::UINTEGER;
      return replay_result(B, L, t, out);
    }
  }
  // The innermost aggregated level yields a cardinality only, unless an operand
  // is annotated: then the generated code loops over the set to read the
  // annotations (EHGenerator.scala:435-456), so its values are needed.
  bool any_annotated = false;
  for (int i = 0; i < ops.n; i++) any_annotated = any_annotated || blk(ops, i)->trie->annotated;
  const bool innermost = level + 2 == num_attrs && !any_annotated;
  if (ops.n == 1 && ((innermost && !materialize) || blk(ops, 0)->set.values)) {
    *out = blk(ops, 0)->set;  // the set itself (TrieBuilder.cpp:223-240)
    return out->cardinality;
  }
  if (!known) not_implemented("builder level beyond num_attributes");
  LevelLog &L = S->logs[t_chunk][level];
  LevelBatch &B = S->batches[level];
  if (recording) {
    sign_once(S, B, num_attrs, level, ops, materialize);
    if (B.nops != ops.n) not_implemented("two call sites with different arity at one builder level");
    if (B.materialized != (materialize || !innermost))
      not_implemented("build_set and build_aggregated_set at one builder level");
    for (int i = 0; i < ops.n; i++) {
      if (B.tries[i] != blk(ops, i)->trie) not_implemented("two call sites on different tries at one builder level");
      L.nodes.push_back(blk(ops, i)->node);
    }
    L.ncalls++;
    set_empty(out);
    return 0;
  }
  if (S->record_level >= 0 && (int)level > S->record_level) not_implemented("call below the level being recorded");
  return replay_result(B, L, result_type(ops), out);
}

static bool profiling();
static double now_ms();

// The body over the level-0 values of chunks [c_begin, c_end), chunks handed
// out dynamically to the pool's workers.
static void run_chunks(const Set<hybrid> *level0, Workers &pool,
                       const std::function<void(size_t, uint32_t, uint32_t)> &body, uint64_t c_begin,
                       uint64_t c_end) {
  const size_t nthreads = pool.size();
  const uint64_t n0 = level0->cardinality;
  const uint64_t nchunks = c_end;
  std::atomic<uint64_t> next(c_begin);
  const bool prof = profiling();
  std::vector<double> first(nthreads, 0.0), last(nthreads, 0.0);
  const double t0 = prof ? now_ms() : 0.0;
  auto worker = [&](size_t tid) {
    if (prof) first[tid] = now_ms() - t0;
    for (;;) {
      const uint64_t c = next.fetch_add(1);
      if (c >= nchunks) break;
      t_chunk = c;
      const uint64_t e = std::min(n0, (c + 1) * CHUNK);
      for (uint64_t i = c * CHUNK; i < e; i++) body(tid, (uint32_t)i, level0->values[i]);
    }
    if (prof) last[tid] = now_ms() - t0;
  };
  pool.run(worker);
  if (prof) {
    double busy = 0, f_max = 0, l_min = 1e30, l_max = 0;
    for (size_t t = 0; t < nthreads; t++) {
      busy += last[t] - first[t];
      f_max = std::max(f_max, first[t]);
      l_min = std::min(l_min, last[t]);
      l_max = std::max(l_max, last[t]);
    }
    fprintf(stderr, "[dropin]   %zu workers: last start %.2f ms, first done %.2f ms, last done %.2f ms, busy %.1f ms\n",
            nthreads, f_max, l_min, l_max, busy);
  }
}

// EHX_DROPIN_PROFILE=1: per pass, the wall time of the host lambdas, of
// gathering the recorded calls and of the batched GPU launch, on stderr.
static bool profiling() {
  static const bool on = getenv("EHX_DROPIN_PROFILE") && atoi(getenv("EHX_DROPIN_PROFILE")) > 0;
  return on;
}
static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Calls of chunks [c0, c1) of level p, gathered for one GPU batch.
struct Gathered {
  uint64_t c0 = 0, c1 = 0, total = 0, bound = 0;
  LevelBuffers *LB = nullptr;
};

// Operand nodes of the recorded calls of chunks [c0, c1), operand-major, in
// chunk order, gathered from the chunk logs by the pool's threads (each chunk
// fills its own range), with the bound on the results' size when the level is
// materialized.
static Gathered gather_calls(Session &S, size_t p, uint64_t c0, uint64_t c1, LevelBuffers &LB, Workers &pool) {
  LevelBatch &B = S.batches[p];
  Gathered g;
  g.c0 = c0;
  g.c1 = c1;
  g.LB = &LB;
  std::vector<uint64_t> base(c1 - c0);
  for (uint64_t c = c0; c < c1; c++) {
    base[c - c0] = g.total;
    g.total += S.logs[c][p].ncalls;
  }
  if (g.total == 0) return g;
  const int nops = B.nops;
  LB.nodes.ensure((size_t)nops * g.total * 4);
  uint32_t *nodes = LB.nodes.as<uint32_t>();
  std::atomic<uint64_t> next_chunk(c0), bound(0);
  auto worker = [&](size_t) {
    uint64_t local = 0;
    for (;;) {
      const uint64_t c = next_chunk.fetch_add(1);
      if (c >= c1) break;
      LevelLog &L = S.logs[c][p];
      const uint64_t b0 = base[c - c0];
      for (uint64_t i = 0; i < L.ncalls; i++) {
        uint64_t mn = ~0ull;
        for (int s = 0; s < nops; s++) {
          const uint32_t nd = L.nodes[i * nops + s];
          nodes[(uint64_t)s * g.total + b0 + i] = nd;
          if (B.materialized) mn = std::min(mn, card_of(B.tries[s], nd));
        }
        if (B.materialized) local += mn;
      }
      std::vector<uint32_t>().swap(L.nodes);
    }
    bound += local;
  };
  pool.run(worker);
  g.bound = bound.load();
  return g;
}

// The GPU batch of gathered calls; each chunk's log then points at its results.
static void run_batch_gathered(Session &S, size_t p, const Gathered &g) {
  if (g.total == 0) return;
  LevelBatch &B = S.batches[p];
  LevelBuffers &LB = *g.LB;
  const int nops = B.nops;
  ehx_trie *dt[8];
  const uint32_t *np[8];
  for (int s = 0; s < nops; s++) {
    dt[s] = B.tries[s]->dev;
    np[s] = LB.nodes.as<uint32_t>() + (uint64_t)s * g.total;
  }
  const uint64_t *res;
  const uint32_t *vals = nullptr;
  if (B.materialized) {
    LB.offsets.ensure((g.total + 1) * 8);
    LB.values.ensure((g.bound ? g.bound : 1) * 4);
    uint64_t got = 0;
    check(ehx_intersect_nodes(ctx(), nops, dt, np, g.total, LB.offsets.as<uint64_t>(), LB.values.as<uint32_t>(),
                              g.bound, &got),
          "ehx_intersect_nodes");
    res = LB.offsets.as<uint64_t>();
    vals = LB.values.as<uint32_t>();
  } else {
    LB.counts.ensure(g.total * 8);
    check(ehx_intersect_nodes_count(ctx(), nops, dt, np, g.total, LB.counts.as<uint64_t>()),
          "ehx_intersect_nodes_count");
    res = LB.counts.as<uint64_t>();
  }
  uint64_t b = 0;
  for (uint64_t c = g.c0; c < g.c1; c++) {
    LevelLog &L = S.logs[c][p];
    L.res = res + b;
    L.vals = vals;
    b += L.ncalls;
  }
}

// A count-only level is recorded in up to PIECES ranges of chunks: once a
// range is recorded, its calls are gathered and its GPU batch runs on a helper
// thread while the workers record the next range (batches run one at a time:
// the helper of the previous range is joined first).  Materialized levels
// need every offset of one batch, so they record all chunks first.
static const uint64_t PIECES = 8, MIN_CHUNKS_PER_PIECE = 32;

// Chunk bounds of the ranges of pass p, cut at equal shares of an estimate of
// each chunk's calls at level p: the values its calls at level p-1 returned
// (each value runs the loop body that makes level p's calls once), one per
// chunk besides.
static std::vector<uint64_t> piece_bounds(const Session &S, size_t p, uint64_t nchunks, uint64_t pieces) {
  std::vector<uint64_t> w(nchunks, 1);
  if (p > 0 && S.batches[p - 1].materialized)
    for (uint64_t c = 0; c < nchunks; c++) {
      const LevelLog &L = S.logs[c][p - 1];
      if (L.ncalls && L.res) w[c] += L.res[L.ncalls] - L.res[0];
    }
  uint64_t total = 0;
  for (uint64_t x : w) total += x;
  std::vector<uint64_t> b(1, 0);
  uint64_t acc = 0;
  for (uint64_t c = 0; c < nchunks && b.size() < pieces; c++) {
    acc += w[c];
    if (acc * pieces >= total * b.size() && c + 1 - b.back() >= MIN_CHUNKS_PER_PIECE) b.push_back(c + 1);
  }
  if (b.back() != nchunks) b.push_back(nchunks);
  return b;
}

void run_passes(size_t num_attrs, const Set<hybrid> *level0, size_t nthreads,
                const std::function<void(size_t, uint32_t, uint32_t)> &body, OutRows *out) {
  if (g_sess) not_implemented("nested par_foreach_aggregate / par_foreach_builder");
  if (level0->cardinality && !level0->values) not_implemented("level-0 set without host values");
  const uint64_t nchunks = (level0->cardinality + CHUNK - 1) / CHUNK;
  const size_t levels = num_attrs >= 2 ? num_attrs - 1 : 0;  // builder levels 0..num_attrs-2
  Session S;
  S.num_attrs = num_attrs;
  S.logs.assign(nchunks, std::vector<LevelLog>(levels));
  S.batches.resize(levels);
  S.signed_levels.reset(new std::atomic<bool>[levels ? levels : 1]);
  for (size_t l = 0; l < levels; l++) S.signed_levels[l].store(false);
  S.out = out;
  if (out) {
    S.rows.resize(nchunks);
    S.annos.resize(nchunks);
  }
  g_sess = &S;
  const bool prof = profiling();
  Workers pool(nthreads);
  for (size_t p = 0; p < levels; p++) {
    S.record_level = (int)p;
    g_phase = 1;
    for (auto &cl : S.logs)
      for (auto &l : cl) l.cursor = 0;
    LevelBatch &B = S.batches[p];
    double t_lambdas = 0, t_gather = 0, t_wait = 0;
    uint64_t handed = 0, calls = 0;  // chunks whose calls went to a batch
    std::thread gpu;
    // only the innermost builder level can be count-only (LevelBatch::innermost)
    const std::vector<uint64_t> bounds = piece_bounds(S, p, nchunks, p + 1 == levels ? PIECES : 1);
    const uint64_t pieces = bounds.size() - 1;
    for (uint64_t k = 0; k < pieces; k++) {
      const uint64_t c0 = bounds[k], c1 = bounds[k + 1];
      double t = prof ? now_ms() : 0;
      run_chunks(level0, pool, body, c0, c1);
      if (prof) t_lambdas += now_ms() - t;
      // the signature's first call has been recorded by now if the range had one
      const bool count_only = S.signed_levels[p].load() && !B.materialized;
      if (!count_only) continue;
      t = prof ? now_ms() : 0;
      // each range's results stay in its own buffers until the final replay
      LevelBuffers &LB = level_buffers(p * PIECES + k);
      const Gathered g = gather_calls(S, p, handed, c1, LB, pool);
      if (prof) t_gather += now_ms() - t;
      if (gpu.joinable()) {
        const double tw = prof ? now_ms() : 0;
        gpu.join();
        if (prof) t_wait += now_ms() - tw;
      }
      calls += g.total;
      handed = c1;
      gpu = std::thread([&S, p, g]() { run_batch_gathered(S, p, g); });
    }
    if (gpu.joinable()) {
      const double tw = prof ? now_ms() : 0;
      gpu.join();
      if (prof) t_wait += now_ms() - tw;
    }
    if (handed < nchunks) {  // a materialized level, or calls not handed out above: one batch
      const double t = prof ? now_ms() : 0;
      const Gathered g = gather_calls(S, p, handed, nchunks, level_buffers(p * PIECES), pool);
      const double t2 = prof ? now_ms() : 0;
      run_batch_gathered(S, p, g);
      calls += g.total;
      if (prof) {
        t_gather += t2 - t;
        t_wait += now_ms() - t2;
      }
    }
    if (prof)
      fprintf(stderr,
              "[dropin] pass %zu: lambdas %.2f ms, gather %.2f ms, GPU batches of %llu calls (%s) %.2f ms not "
              "overlapped\n",
              p, t_lambdas, t_gather, (unsigned long long)calls, B.materialized ? "materialized" : "counts", t_wait);
  }
  S.record_level = -1;
  g_phase = 2;
  for (auto &cl : S.logs)
    for (auto &l : cl) l.cursor = 0;
  const double t3 = prof ? now_ms() : 0;
  run_chunks(level0, pool, body, 0, nchunks);
  if (prof) fprintf(stderr, "[dropin] replay: lambdas %.2f ms\n", now_ms() - t3);
  g_sess = nullptr;
  g_phase = 0;
  if (out) {  // chunks follow the level-0 set, so their concatenation is lexicographic
    size_t nr = 0, na = 0;
    for (auto &r : S.rows) nr += r.size();
    for (auto &a : S.annos) na += a.size();
    out->rows.reserve(out->rows.size() + nr);
    out->annos.reserve(out->annos.size() + na);
    for (auto &r : S.rows) out->rows.insert(out->rows.end(), r.begin(), r.end());
    for (auto &a : S.annos) out->annos.insert(out->annos.end(), a.begin(), a.end());
  }
}

}  // namespace ehx_dropin

// ---- utils ---------------------------------------------------------------------------

namespace thread_pool {
void initializeThreadPool() { (void)ehx_dropin::ctx(); }
void deleteThreadPool() {}
}  // namespace thread_pool

namespace par {
size_t for_range(const size_t from, const size_t to, const size_t block_size,
                 std::function<void(size_t tid, size_t i)> body) {
  const size_t len = to > from ? to - from : 0;
  const size_t bs = (block_size == 0 || block_size > len) ? 1 : block_size;
  std::atomic<size_t> next(0);
  auto worker = [&](size_t tid) {
    for (;;) {
      const size_t s = next.fetch_add(bs);
      if (s >= len) break;
      const size_t e = std::min(len, s + bs);
      for (size_t j = s; j < e; j++) body(tid, from + j);
    }
  };
  std::vector<std::thread> th;
  for (size_t t = 1; t < NUM_THREADS; t++) th.emplace_back(worker, t);
  worker(0);
  for (auto &x : th) x.join();
  return 1;
}
size_t for_range(const size_t from, const size_t to, std::function<void(size_t tid, size_t i)> body) {
  const size_t len = to > from ? to - from : 0;
  return for_range(from, to, (len + NUM_THREADS - 1) / (NUM_THREADS ? NUM_THREADS : 1), body);
}
}  // namespace par
