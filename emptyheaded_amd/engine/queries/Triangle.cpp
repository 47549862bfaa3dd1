// Triangle(a,b,c) :- Edge(a,b),Edge(b,c),Edge(a,c).
//
// The query code EmptyHeaded's code generator emits for this materializing
// rule (query_compiler/src/main/scala/codegen/query/EHGenerator.scala:
// emitIntermediateTrie :711-726, emitParallelBuilder :702-710,
// emitHeadBuildCode :565-606 with build_set, emitHeadAllocations :643-655,
// emitHeadParForeach :273-300 with par_foreach_builder, nprrRecursiveCall
// :523-563 with emitBuildCode :302-354, emitAllocateCode :356-367,
// emitForeach :382-396, emitSetValues :369-380; emitNPRR :728-838; the output
// insertion of emitOutputs :152-167), restated by hand -- no JVM is available
// to run the generator -- and compiled unchanged against the drop-in headers.
// Attribute info follows codegen/query/QueryPlan.scala:226-245: a, b, c all
// materialized, nextMaterialized a->b->c, no annotation.
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_Triangle(mymap *input_tries) {
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
  Trie<void *, ParMemoryBuffer> *Trie_Triangle_0_1_2 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/Triangle", 3, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_Triangle_0_1_2, 3);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_b(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_c(Trie_Edge_0_1);
    const size_t count_a = Builders.build_set(Iterators_Edge_a_b.head, Iterators_Edge_a_c.head);
    (void)count_a;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_b = Iterators_Edge_a_b.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_b_c = Iterators_Edge_b_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_c = Iterators_Edge_a_c.iterators.at(tid);
      Iterator_Edge_a_b->get_next_block(0, a_d);
      Iterator_Edge_a_c->get_next_block(0, a_d);
      const size_t count_b =
          Builder->build_set(tid, Iterator_Edge_a_b->get_block(1), Iterator_Edge_b_c->get_block(0));
      (void)count_b;
      Builder->allocate_next(tid);
      Builder->foreach_builder([&](const uint32_t b_i, const uint32_t b_d) {
        Iterator_Edge_b_c->get_next_block(0, b_d);
        const size_t count_c =
            Builder->build_set(tid, Iterator_Edge_b_c->get_block(1), Iterator_Edge_a_c->get_block(1));
        num_rows_reducer.update(tid, count_c);
        Builder->set_level(b_i, b_d);
      });
      Builder->set_level(a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG Triangle TIME", bag_timer);
    Trie_Triangle_0_1_2->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_Triangle_0_1_2->num_rows = Builders.trie->num_rows;
    Trie_Triangle_0_1_2->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("Triangle_0_1_2", Trie_Triangle_0_1_2));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// Cython's c_run_<id> equivalent (QueryPlan.scala:25-39) plus the PTrie read:
// runs the query, returns num_rows, and copies up to `cap` rows (3 columns).
extern "C" long c_run_Triangle(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_Triangle(&m);
  Trie<void *, ParMemoryBuffer> *out = (Trie<void *, ParMemoryBuffer> *)m.at("Triangle_0_1_2");
  if (num_rows) *num_rows = out->num_rows;
  delete out;
  return 0;
}

extern "C" long c_rows_Triangle(const char *db, uint32_t *rows, unsigned long cap, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_Triangle(&m);
  Trie<void *, ParMemoryBuffer> *out = (Trie<void *, ParMemoryBuffer> *)m.at("Triangle_0_1_2");
  if (num_rows) *num_rows = out->num_rows;
  const size_t r = export_rows(out, rows, (double *)nullptr, cap);
  delete out;
  return (long)r;
}
