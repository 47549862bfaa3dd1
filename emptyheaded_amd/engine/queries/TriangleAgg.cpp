// TriangleAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),z:long<-[COUNT(*)].
//
// The query code EmptyHeaded's code generator emits for this rule
// (query_compiler/src/main/scala/codegen/query/EHGenerator.scala: getCode
// :134-153, emitLoadRelations :171-199, emitNPRR :728-838 with
// nprrRecursiveCall :523-563), restated by hand -- no JVM is available to run
// the generator -- and compiled unchanged against the drop-in headers.  Only
// the database folder is a parameter (the generator bakes in an absolute path).
#include "emptyheaded.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_TriangleAgg(mymap *input_tries) {
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
  Trie<long, ParMemoryBuffer> *Trie_TriangleAgg_ =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/TriangleAgg", 0, true);
  long TriangleAgg;
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_TriangleAgg_, 3);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_b(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_c(Trie_Edge_0_1);
    const size_t count_a = Builders.build_aggregated_set(Iterators_Edge_a_b.head, Iterators_Edge_a_c.head);
    (void)count_a;
    par::reducer<long> annotation_a(0, [&](long a, long b) { return a + b; });
    Builders.par_foreach_aggregate([&](const size_t tid, const uint32_t a_d) {
      TrieBuilder<long, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_b = Iterators_Edge_a_b.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_b_c = Iterators_Edge_b_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_c = Iterators_Edge_a_c.iterators.at(tid);
      Iterator_Edge_a_b->get_next_block(0, a_d);
      Iterator_Edge_a_c->get_next_block(0, a_d);
      const long intermediate_a = (long)1;
      const size_t count_b =
          Builder->build_aggregated_set(Iterator_Edge_a_b->get_block(1), Iterator_Edge_b_c->get_block(0));
      (void)count_b;
      long annotation_b = (long)0;
      Builder->foreach_aggregate([&](const uint32_t b_d) {
        Iterator_Edge_b_c->get_next_block(0, b_d);
        const long intermediate_b = (long)1 * 1 * intermediate_a;
        const size_t count_c =
            Builder->build_aggregated_set(Iterator_Edge_b_c->get_block(1), Iterator_Edge_a_c->get_block(1));
        num_rows_reducer.update(tid, 1);
        long annotation_c = (long)0;
        const long intermediate_c = (long)1 * 1 * 1 * intermediate_b * (count_c * 1);
        annotation_c += intermediate_c;
        annotation_b += annotation_c;
      });
      annotation_a.update(tid, annotation_b);
    });
    Builders.trie->annotation = annotation_a.evaluate(0);
    TriangleAgg = Builders.trie->annotation;
    (void)TriangleAgg;
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG TriangleAgg TIME", bag_timer);
    Trie_TriangleAgg_->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_TriangleAgg_->num_rows = 0;
    Trie_TriangleAgg_->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("TriangleAgg_", Trie_TriangleAgg_));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// Cython's c_run_<id> equivalent (QueryPlan.scala:25-39): runs the query over
// the database folder and returns the bag's annotation; num_rows receives the
// output trie's row count (0 for a scalar bag).
extern "C" long c_run_TriangleAgg(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_TriangleAgg(&m);
  Trie<long, ParMemoryBuffer> *out = (Trie<long, ParMemoryBuffer> *)m.at("TriangleAgg_");
  long v = out->annotation;
  if (num_rows) *num_rows = out->num_rows;
  delete out;
  return v;
}
