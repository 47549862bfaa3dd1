// TriangleCounts(a;z) :- Edge(a,b),Edge(b,c),Edge(a,c),z:long<-[COUNT(b,c)].
//
// The per-vertex triangle count of the reference's parser and optimizer tests
// (python/test_parser.py, python/test_optimizer.py: `Triangle(a;z) :- ...
// z:long<-[COUNT(b,c)]`, renamed so its symbols do not clash with the
// materializing Triangle query).  The query code the generator emits for it
// (EHGenerator.scala: emitHeadBuildCode :565-606 with build_set,
// emitHeadAllocations :643-655 with allocate_annotation, emitHeadParForeach
// :273-300, nprrRecursiveCall :523-563 over the aggregated b and c with
// emitAnnotationAccessors :398-433, emitFinalAnnotation :435-456,
// emitAnnotationComputation :457-478, then emitSetValues :369-380 with
// set_annotation), restated by hand and compiled unchanged against the
// drop-in headers.  Attribute info follows codegen/query/QueryPlan.scala
// :208-245: a materialized with annotation "b" (the attribute after the last
// materialized one); b and c aggregated by COUNT (SUM with init 1,
// parsers/Datalog.scala:116-117), prev(c) = b.
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_TriangleCounts(mymap *input_tries) {
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
  Trie<long, ParMemoryBuffer> *Trie_TriangleCounts_0 =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/TriangleCounts", 1, true);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_TriangleCounts_0, 3);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_b(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_c(Trie_Edge_0_1);
    const size_t count_a = Builders.build_set(Iterators_Edge_a_b.head, Iterators_Edge_a_c.head);
    (void)count_a;
    Builders.allocate_annotation();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<long, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_b = Iterators_Edge_a_b.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_b_c = Iterators_Edge_b_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_c = Iterators_Edge_a_c.iterators.at(tid);
      Iterator_Edge_a_b->get_next_block(0, a_d);
      Iterator_Edge_a_c->get_next_block(0, a_d);
      const size_t count_b =
          Builder->build_aggregated_set(Iterator_Edge_a_b->get_block(1), Iterator_Edge_b_c->get_block(0));
      (void)count_b;
      long annotation_b = (long)0;
      Builder->foreach_aggregate([&](const uint32_t b_d) {
        Iterator_Edge_b_c->get_next_block(0, b_d);
        const long intermediate_b = (long)1 * 1;
        const size_t count_c =
            Builder->build_aggregated_set(Iterator_Edge_b_c->get_block(1), Iterator_Edge_a_c->get_block(1));
        num_rows_reducer.update(tid, 1);
        long annotation_c = (long)0;
        const long intermediate_c = (long)1 * 1 * 1 * intermediate_b * (count_c * 1);
        annotation_c += intermediate_c;
        annotation_b += annotation_c;
      });
      Builder->set_annotation(annotation_b, a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG TriangleCounts TIME", bag_timer);
    Trie_TriangleCounts_0->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_TriangleCounts_0->num_rows = Builders.trie->num_rows;
    Trie_TriangleCounts_0->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("TriangleCounts_0", Trie_TriangleCounts_0));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// c_run_<id> (QueryPlan.scala:25-39) plus the PTrie read: returns num_rows
// through the pointer and copies up to `cap` (a, count) rows.
extern "C" long c_run_TriangleCounts(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_TriangleCounts(&m);
  Trie<long, ParMemoryBuffer> *out = (Trie<long, ParMemoryBuffer> *)m.at("TriangleCounts_0");
  if (num_rows) *num_rows = out->num_rows;
  delete out;
  return 0;
}

extern "C" long c_rows_TriangleCounts(const char *db, uint32_t *rows, double *annos, unsigned long cap,
                                      unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_TriangleCounts(&m);
  Trie<long, ParMemoryBuffer> *out = (Trie<long, ParMemoryBuffer> *)m.at("TriangleCounts_0");
  if (num_rows) *num_rows = out->num_rows;
  const size_t r = export_rows(out, rows, annos, cap);
  delete out;
  return (long)r;
}
