// SSSP(x;y) :- Edge(w,x),w=0,y:long <- [1].
// SSSP(x;y)*[c=0] :- Edge(w,x),SSSP(w),y:long <- [1+MIN(w;1)].
// (test/graph/travis.py:556-567), and the same rule with an iteration bound:
// SSSPIter2(x;y)*[i=2] :- Edge(w,x),SSSPIter2(w),y:long <- [1+MIN(w;1)].
//
// The query code EmptyHeaded's code generator emits for these programs
// (query_compiler/src/main/scala/codegen/query/EHGenerator.scala: getCode
// :134-153; detectTransitiveClosure :26-44 -- two bags, the base case one
// relation with one selection, the recursion converging on a MIN; the
// single-source branch of run :59-117: only Edge_0_1 is loaded, the output is
// an intermediate trie (emitIntermediateTrie :711-726) filled by
// tc::unweighted_single_source with iterations -1 for *[c=0] and the bound for
// *[i=k], the source through the encoding, and the annotation expression with
// AGG removed; emitOutputs :155-169; emitEndQuery :201-207), restated by hand
// -- no JVM is available to run the generator -- and compiled unchanged
// against the drop-in headers.  Only the database folder is a parameter (the
// generator bakes in an absolute path).
#include "query_io.hpp"
#include "TransitiveClosure.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_SSSP(mymap *input_tries) {
  thread_pool::initializeThreadPool();
  Trie<void *, ParMemoryBuffer> *Trie_Edge_0_1 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_Edge_0_1 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/Edge/Edge_0_1");
    timer::stop_clock("LOADING Trie Edge_0_1", start_time);
  }
  auto e_loading_uint32_t = timer::start_clock();
  Encoding<uint32_t> *Encoding_uint32_t = Encoding<uint32_t>::from_binary(db_folder + "/encodings/uint32_t/");
  (void)Encoding_uint32_t;
  timer::stop_clock("LOADING ENCODINGS uint32_t", e_loading_uint32_t);
  par::reducer<size_t> num_rows_reducer(0, [](size_t a, size_t b) { return a + b; });
  //
  // query plan
  //
  auto query_timer = timer::start_clock();
  Trie<long, ParMemoryBuffer> *Trie_SSSP_0 = new Trie<long, ParMemoryBuffer>(db_folder + "/relations/SSSP", 1, true);
  tc::unweighted_single_source<hybrid, ParMemoryBuffer, long>(
      -1,
      Encoding_uint32_t->value_to_key.at(0),  // from encoding
      Encoding_uint32_t->num_distinct,        // from encoding
      Trie_Edge_0_1,                          // input graph
      Trie_SSSP_0,                            // output vector
      1, [&](long a) { return 1 + a; });
  Trie_SSSP_0->encodings.push_back((void *)Encoding_uint32_t);
  input_tries->insert(std::make_pair("SSSP_0", Trie_SSSP_0));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

void run_SSSPIter2(mymap *input_tries) {
  thread_pool::initializeThreadPool();
  Trie<void *, ParMemoryBuffer> *Trie_Edge_0_1 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_Edge_0_1 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/Edge/Edge_0_1");
    timer::stop_clock("LOADING Trie Edge_0_1", start_time);
  }
  auto e_loading_uint32_t = timer::start_clock();
  Encoding<uint32_t> *Encoding_uint32_t = Encoding<uint32_t>::from_binary(db_folder + "/encodings/uint32_t/");
  (void)Encoding_uint32_t;
  timer::stop_clock("LOADING ENCODINGS uint32_t", e_loading_uint32_t);
  par::reducer<size_t> num_rows_reducer(0, [](size_t a, size_t b) { return a + b; });
  //
  // query plan
  //
  auto query_timer = timer::start_clock();
  Trie<long, ParMemoryBuffer> *Trie_SSSPIter2_0 =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/SSSPIter2", 1, true);
  tc::unweighted_single_source<hybrid, ParMemoryBuffer, long>(
      2,
      Encoding_uint32_t->value_to_key.at(0),  // from encoding
      Encoding_uint32_t->num_distinct,        // from encoding
      Trie_Edge_0_1,                          // input graph
      Trie_SSSPIter2_0,                       // output vector
      1, [&](long a) { return 1 + a; });
  Trie_SSSPIter2_0->encodings.push_back((void *)Encoding_uint32_t);
  input_tries->insert(std::make_pair("SSSPIter2_0", Trie_SSSPIter2_0));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// c_run_<id> (QueryPlan.scala:25-39) plus the PTrie read of the output trie:
// num_rows through the pointer, up to `cap` (x, y) rows copied.
template <void (*RUN)(mymap *)>
static long rows_of(const char *db, const char *name, uint32_t *rows, double *annos, unsigned long cap,
                    unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  RUN(&m);
  Trie<long, ParMemoryBuffer> *out = (Trie<long, ParMemoryBuffer> *)m.at(name);
  if (num_rows) *num_rows = out->num_rows;
  const size_t r = rows ? export_rows(out, rows, annos, cap) : 0;
  delete out;
  return (long)r;
}

extern "C" long c_run_SSSP(const char *db, unsigned long *num_rows) {
  rows_of<run_SSSP>(db, "SSSP_0", nullptr, nullptr, 0, num_rows);
  return 0;
}
extern "C" long c_rows_SSSP(const char *db, uint32_t *rows, double *annos, unsigned long cap,
                            unsigned long *num_rows) {
  return rows_of<run_SSSP>(db, "SSSP_0", rows, annos, cap, num_rows);
}
extern "C" long c_run_SSSPIter2(const char *db, unsigned long *num_rows) {
  rows_of<run_SSSPIter2>(db, "SSSPIter2_0", nullptr, nullptr, 0, num_rows);
  return 0;
}
extern "C" long c_rows_SSSPIter2(const char *db, uint32_t *rows, double *annos, unsigned long cap,
                                 unsigned long *num_rows) {
  return rows_of<run_SSSPIter2>(db, "SSSPIter2_0", rows, annos, cap, num_rows);
}
