// ehx_graph.cpp -- synthetic inputs on the host: counter-based generators and
// undirected clean-up (symmetrize, de-duplicate, relabel, orient).
//
// Random numbers: SplitMix64 finalizer applied to seed + counter * golden
// ratio, so every sample is a pure function of (seed, index) and generation
// splits across threads without changing the output.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <math.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/ehx.h"

namespace ehx {
int fail(int code, const char *fmt, ...);
}

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rnd(uint64_t seed, uint64_t ctr) { return mix64(seed + (ctr + 1) * 0x9E3779B97F4A7C15ull); }
// uniform double in [0,1) from the top 53 bits
inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

unsigned num_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  if (n > 64) n = 64;
  return n;
}

template <class F> void parallel_for(uint64_t n, F f) {
  unsigned T = num_threads();
  if (n < 65536) T = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; t++) {
    uint64_t a = n * t / T, b = n * (t + 1) / T;
    th.emplace_back([=]() { f(a, b); });
  }
  for (auto &x : th) x.join();
}

// Parallel LSD radix sort of 64-bit keys (16-bit digits) on the low `bits`.
void sort_u64(std::vector<uint64_t> &k, int bits) {
  const uint64_t n = k.size();
  if (n < 2) return;
  std::vector<uint64_t> tmp(n);
  unsigned T = num_threads();
  if (n < 65536) T = 1;
  uint64_t *src = k.data(), *dst = tmp.data();
  for (int shift = 0; shift < bits; shift += 16) {
    std::vector<std::vector<uint64_t>> cnt(T, std::vector<uint64_t>(65536, 0));
    {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t]() {
          uint64_t a = n * t / T, b = n * (t + 1) / T;
          auto &c = cnt[t];
          for (uint64_t i = a; i < b; i++) c[(src[i] >> shift) & 0xFFFF]++;
        });
      for (auto &x : th) x.join();
    }
    uint64_t s = 0;
    for (int d = 0; d < 65536; d++)
      for (unsigned t = 0; t < T; t++) {
        uint64_t c = cnt[t][d];
        cnt[t][d] = s;
        s += c;
      }
    {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t]() {
          uint64_t a = n * t / T, b = n * (t + 1) / T;
          auto &c = cnt[t];
          for (uint64_t i = a; i < b; i++) dst[c[(src[i] >> shift) & 0xFFFF]++] = src[i];
        });
      for (auto &x : th) x.join();
    }
    std::swap(src, dst);
  }
  if (src != k.data()) memcpy(k.data(), src, n * 8);
}

int bits_for(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b)) b++;
  return b;
}

}  // namespace

extern "C" int ehx_gen_er(uint64_t nv, uint64_t samples, uint64_t seed, uint32_t *src, uint32_t *dst) {
  if (!src || !dst || nv == 0 || nv > (1ull << 32)) return ehx::fail(EHX_ERR_INVALID, "ehx_gen_er: bad argument");
  parallel_for(samples, [=](uint64_t a, uint64_t b) {
    for (uint64_t i = a; i < b; i++) {
      uint64_t r0 = rnd(seed, 2 * i), r1 = rnd(seed, 2 * i + 1);
      src[i] = (uint32_t)(((unsigned __int128)r0 * nv) >> 64);
      dst[i] = (uint32_t)(((unsigned __int128)r1 * nv) >> 64);
    }
  });
  return EHX_OK;
}

// R-MAT (Chakrabarti et al. 2004): at each of `scale` levels pick a quadrant
// with probabilities (a, b, c, 1-a-b-c); quadrant b sets the column bit,
// c the row bit, d both.  Sample i at level l uses counter i*scale + l.
extern "C" int ehx_gen_rmat(int scale, uint64_t samples, double a, double b, double c, uint64_t seed, uint32_t *src,
                            uint32_t *dst) {
  if (!src || !dst || scale < 1 || scale > 32 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return ehx::fail(EHX_ERR_INVALID, "ehx_gen_rmat: bad argument");
  const double ab = a + b, abc = a + b + c;
  parallel_for(samples, [=](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; i++) {
      uint64_t u = 0, v = 0;
      for (int l = 0; l < scale; l++) {
        double x = unit(rnd(seed, i * (uint64_t)scale + (uint64_t)l));
        uint64_t bit = 1ull << (scale - 1 - l);
        if (x < a) {
        } else if (x < ab) {
          v |= bit;
        } else if (x < abc) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      src[i] = (uint32_t)u;
      dst[i] = (uint32_t)v;
    }
  });
  return EHX_OK;
}

extern "C" int ehx_graph_prepare(const uint32_t *src, const uint32_t *dst, uint64_t n, int relabel, int orient,
                                 uint32_t **out_src, uint32_t **out_dst, uint64_t *out_m, uint32_t *out_nv) {
  if ((!src || !dst) && n) return ehx::fail(EHX_ERR_INVALID, "ehx_graph_prepare: null input");
  if (!out_src || !out_dst || !out_m) return ehx::fail(EHX_ERR_INVALID, "ehx_graph_prepare: null output");
  if (relabel < 0 || relabel > 2 || orient < 0 || orient > 2) return ehx::fail(EHX_ERR_INVALID, "bad relabel/orient");
  uint32_t maxv = 0;
  for (uint64_t i = 0; i < n; i++) maxv = std::max(maxv, std::max(src[i], dst[i]));
  const int vb = bits_for(maxv);
  // 1. both directions of every non-loop edge, sorted and unique
  std::vector<uint64_t> k;
  k.reserve(2 * n);
  for (uint64_t i = 0; i < n; i++) {
    if (src[i] == dst[i]) continue;
    k.push_back(((uint64_t)src[i] << vb) | dst[i]);
    k.push_back(((uint64_t)dst[i] << vb) | src[i]);
  }
  sort_u64(k, 2 * vb);
  k.erase(std::unique(k.begin(), k.end()), k.end());
  const uint64_t mask = vb ? ((1ull << vb) - 1) : 0;
  // 2. degrees and relabel map
  const uint64_t V = (uint64_t)maxv + 1;
  std::vector<uint32_t> deg(V, 0);
  for (uint64_t e : k) deg[e >> vb]++;
  std::vector<uint32_t> lab(V);
  uint32_t nv = 0;
  if (relabel == 0) {
    for (uint64_t v = 0; v < V; v++) lab[v] = (uint32_t)v;
    nv = (uint32_t)V;
  } else if (relabel == 2) {
    for (uint64_t v = 0; v < V; v++) lab[v] = deg[v] ? nv++ : 0xFFFFFFFFu;
  } else {
    // ascending degree, ties by id; isolated vertices dropped
    std::vector<uint64_t> order;
    order.reserve(V);
    const int ib = bits_for(maxv);
    for (uint64_t v = 0; v < V; v++)
      if (deg[v]) order.push_back(((uint64_t)deg[v] << ib) | v);
    sort_u64(order, 32 + ib);
    for (uint64_t v = 0; v < V; v++) lab[v] = 0xFFFFFFFFu;
    for (uint64_t r = 0; r < order.size(); r++) lab[order[r] & ((1ull << ib) - 1)] = (uint32_t)r;
    nv = (uint32_t)order.size();
  }
  // 3. relabel + orient, sort
  const int nb = bits_for(nv ? nv - 1 : 0);
  std::vector<uint64_t> o;
  o.reserve(orient ? k.size() / 2 : k.size());
  for (uint64_t e : k) {
    uint32_t u = lab[e >> vb], v = lab[e & mask];
    if (orient == 1 && !(u < v)) continue;
    if (orient == 2 && !(u > v)) continue;
    o.push_back(((uint64_t)u << nb) | v);
  }
  std::vector<uint64_t>().swap(k);
  sort_u64(o, 2 * nb);
  const uint64_t m = o.size();
  uint32_t *os = (uint32_t *)malloc(4 * (m ? m : 1));
  uint32_t *od = (uint32_t *)malloc(4 * (m ? m : 1));
  if (!os || !od) {
    free(os);
    free(od);
    return ehx::fail(EHX_ERR_NOMEM, "ehx_graph_prepare: out of host memory");
  }
  const uint64_t nmask = nb ? ((1ull << nb) - 1) : 0;
  for (uint64_t i = 0; i < m; i++) {
    os[i] = (uint32_t)(o[i] >> nb);
    od[i] = (uint32_t)(o[i] & nmask);
  }
  *out_src = os;
  *out_dst = od;
  *out_m = m;
  if (out_nv) *out_nv = nv;
  return EHX_OK;
}

// Zipf(s) over k ranks: P(rank r) proportional to 1 / (r + 1)^s, by inverse CDF.
extern "C" int ehx_gen_zipf(uint64_t n, uint32_t k, double s, uint64_t seed, uint32_t *out) {
  if (!out && n) return ehx::fail(EHX_ERR_INVALID, "ehx_gen_zipf: null output");
  if (k == 0 || !(s >= 0.0)) return ehx::fail(EHX_ERR_INVALID, "ehx_gen_zipf: k >= 1 and s >= 0 required");
  std::vector<double> cdf(k);
  double acc = 0.0;
  for (uint32_t r = 0; r < k; r++) cdf[r] = (acc += std::pow((double)(r + 1), -s));
  for (uint32_t r = 0; r < k; r++) cdf[r] /= acc;
  cdf[k - 1] = 1.0;
  parallel_for(n, [&](uint64_t a, uint64_t b) {
    for (uint64_t i = a; i < b; i++) {
      const double u = unit(rnd(seed, i));
      out[i] = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
      if (out[i] >= k) out[i] = k - 1;
    }
  });
  return EHX_OK;
}

// ---- ego-network batch --------------------------------------------------------
namespace {

struct EgoNet {
  std::vector<uint64_t> edges;  // local (u << 32 | v), u < v after relabel
  uint32_t nv = 0;
};

// One ego-network: center c and its neighbors in hash order join one at a time
// (a radius-1 BFS from c); every join adds the edges from the new vertex to the
// vertices already in, in adjacency order, and the net stops at `target` edges.
// Returns the edges in join order as local (earlier, later) pairs, center = 0.
uint64_t grow_ego(const uint64_t *off, const uint32_t *adj, uint32_t c, uint64_t target, uint64_t seed,
                  std::vector<uint32_t> &pos, std::vector<uint32_t> &members, std::vector<uint64_t> &out) {
  out.clear();
  members.clear();
  std::vector<std::pair<uint64_t, uint32_t>> order;
  order.reserve(off[c + 1] - off[c]);
  for (uint64_t i = off[c]; i < off[c + 1]; i++) order.emplace_back(mix64(seed ^ adj[i]), adj[i]);
  std::sort(order.begin(), order.end());
  pos[c] = 0;
  members.push_back(c);
  for (auto &hv : order) {
    if (out.size() >= target) break;
    const uint32_t v = hv.second, k = (uint32_t)members.size();
    pos[v] = k;
    members.push_back(v);
    for (uint64_t i = off[v]; i < off[v + 1] && out.size() < target; i++) {
      const uint32_t u = adj[i];
      if (pos[u] < k) out.push_back(((uint64_t)pos[u] << 32) | k);
    }
  }
  for (uint32_t v : members) pos[v] = 0xFFFFFFFFu;
  return out.size();
}

}  // namespace

extern "C" int ehx_gen_egonets(const uint32_t *src, const uint32_t *dst, uint64_t m, uint32_t num_vertices,
                               uint32_t num_nets, uint32_t min_edges, uint32_t max_edges, uint64_t seed,
                               uint32_t **out_src, uint32_t **out_dst, uint64_t *out_m, uint32_t *net_base,
                               uint64_t *net_off) {
  if ((!src || !dst) && m) return ehx::fail(EHX_ERR_INVALID, "ehx_gen_egonets: null input");
  if (!out_src || !out_dst || !out_m || !net_base || !net_off)
    return ehx::fail(EHX_ERR_INVALID, "ehx_gen_egonets: null output");
  if (min_edges < 1 || max_edges < min_edges) return ehx::fail(EHX_ERR_INVALID, "ehx_gen_egonets: bad edge range");
  const uint32_t n = num_vertices;
  for (uint64_t i = 0; i < m; i++) {
    if (src[i] >= n || dst[i] >= n) return ehx::fail(EHX_ERR_INVALID, "ehx_gen_egonets: vertex id >= num_vertices");
    if (i && (src[i] < src[i - 1] || (src[i] == src[i - 1] && dst[i] <= dst[i - 1])))
      return ehx::fail(EHX_ERR_INVALID, "ehx_gen_egonets: rows must be sorted and unique");
  }
  // CSR of the (symmetric) input
  std::vector<uint64_t> off(n + 1, 0);
  for (uint64_t i = 0; i < m; i++) off[src[i] + 1]++;
  for (uint32_t v = 0; v < n; v++) off[v + 1] += off[v];
  // vertices by ascending degree (ties by id): candidate pools are suffixes
  std::vector<uint32_t> byd(n);
  for (uint32_t v = 0; v < n; v++) byd[v] = v;
  std::stable_sort(byd.begin(), byd.end(),
                   [&](uint32_t a, uint32_t b) { return off[a + 1] - off[a] < off[b + 1] - off[b]; });
  auto deg = [&](uint32_t v) { return off[v + 1] - off[v]; };
  std::vector<EgoNet> nets(num_nets);
  const double lr = std::log((double)max_edges / (double)min_edges);
  unsigned T = num_threads();
  if (T > num_nets) T = num_nets ? num_nets : 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; t++) {
    th.emplace_back([&, t]() {
      std::vector<uint32_t> pos(n, 0xFFFFFFFFu), members;
      std::vector<uint64_t> e;
      for (uint32_t q = t; q < num_nets; q += T) {
        const uint64_t qs = mix64(seed ^ (0xE90ull << 40) ^ q);
        uint64_t target = (uint64_t)std::floor((double)min_edges * std::exp(lr * unit(rnd(qs, 0))));
        target = std::min<uint64_t>(std::max<uint64_t>(target, min_edges), max_edges);
        // pool: vertices of degree >= target / 2; halve the pool towards the hubs on every miss
        uint64_t lo = std::lower_bound(byd.begin(), byd.end(), target / 2,
                                       [&](uint32_t v, uint64_t d) { return deg(v) < d; }) - byd.begin();
        if (lo >= n) lo = n ? n - 1 : 0;
        uint64_t best = 0, best_seed = 0;
        uint32_t best_c = n ? byd[n - 1] : 0;
        for (uint32_t attempt = 0; n; attempt++) {
          const uint32_t c = byd[lo + rnd(qs, 1 + attempt) % (n - lo)];
          const uint64_t hs = rnd(qs, 1000 + attempt);
          const uint64_t got = grow_ego(off.data(), dst, c, target, hs, pos, members, e);
          if (got >= target) break;
          if (got > best) { best = got; best_c = c; best_seed = hs; }
          if (lo == n - 1 || attempt >= 64) {  // nothing larger to draw: keep the largest net seen
            grow_ego(off.data(), dst, best_c, target, best_seed, pos, members, e);
            break;
          }
          lo = lo + (n - lo) / 2;
        }
        // relabel by ascending degree inside the net (ties by join order), orient u < v
        EgoNet &N = nets[q];
        N.nv = (uint32_t)members.size();
        std::vector<uint32_t> d(N.nv, 0), rank(N.nv);
        for (uint64_t x : e) { d[x >> 32]++; d[x & 0xFFFFFFFFu]++; }
        std::vector<uint32_t> ord(N.nv);
        for (uint32_t i = 0; i < N.nv; i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return d[a] < d[b]; });
        for (uint32_t r = 0; r < N.nv; r++) rank[ord[r]] = r;
        N.edges.resize(e.size());
        for (size_t i = 0; i < e.size(); i++) {
          uint32_t u = rank[e[i] >> 32], v = rank[e[i] & 0xFFFFFFFFu];
          if (u > v) std::swap(u, v);
          N.edges[i] = ((uint64_t)u << 32) | v;
        }
        std::sort(N.edges.begin(), N.edges.end());
      }
    });
  }
  for (auto &x : th) x.join();
  uint64_t total = 0, vtot = 0;
  for (uint32_t q = 0; q < num_nets; q++) {
    net_base[q] = (uint32_t)vtot;
    net_off[q] = total;
    vtot += nets[q].nv;
    total += nets[q].edges.size();
    if (vtot > 0xFFFFFFFFull) return ehx::fail(EHX_ERR_UNSUPPORTED, "ehx_gen_egonets: more than 2^32 vertices");
  }
  net_base[num_nets] = (uint32_t)vtot;
  net_off[num_nets] = total;
  uint32_t *os = (uint32_t *)malloc(4 * (total ? total : 1));
  uint32_t *od = (uint32_t *)malloc(4 * (total ? total : 1));
  if (!os || !od) {
    free(os);
    free(od);
    return ehx::fail(EHX_ERR_NOMEM, "ehx_gen_egonets: out of host memory");
  }
  for (uint32_t q = 0; q < num_nets; q++) {
    const uint64_t o = net_off[q];
    for (size_t i = 0; i < nets[q].edges.size(); i++) {
      os[o + i] = net_base[q] + (uint32_t)(nets[q].edges[i] >> 32);
      od[o + i] = net_base[q] + (uint32_t)(nets[q].edges[i] & 0xFFFFFFFFu);
    }
  }
  *out_src = os;
  *out_dst = od;
  *out_m = total;
  return EHX_OK;
}

extern "C" void ehx_free(void *p) { free(p); }
