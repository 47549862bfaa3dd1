// TriangleProj(a,b) :- Edge(a,b),Edge(b,c),Edge(a,c).
// (test/graph/travis.py:282-300: 85,658 rows on facebook_pruned, first (6,5))
//
// The query code EmptyHeaded's code generator emits for this projection,
// restated by hand -- no JVM is available to run the planner or the generator
// -- and compiled unchanged against the drop-in headers.  Only the database
// folder is a parameter.
//
// The plan.  The output (a, b) is covered by Edge(a,b) alone, so the planner
// roots the GHD there and puts Edge(b,c),Edge(a,c) -- the relations that hold
// c -- in a child bag that keeps the attributes it shares with the root
// (solver/GHDSolver.scala:8-42, GHDNode.scala:121-147); the attribute order is
// [a, b, c].  The child bag bag_1_a_b_c_TriangleProj(a,b) projects c away, and
// the generator drops an attribute that is neither materialized, aggregated
// nor selected (codegen/query/QueryPlan.scala:244): its code intersects
// nothing on c, a comes from head(Edge(a,c)) and b from head(Edge(b,c)).  The
// root bag joins Edge(a,b) with it.  (The result: the edges (a, b) whose b has
// a child set -- the reference's 85,658 rows; the edges that close a triangle
// are 84,185.)
//
// The code, per bag (emitNPRR, EHGenerator.scala:728-838): emitIntermediateTrie
// :711-726, emitHeadBuildCode :565-606 with build_set, emitHeadAllocations
// :643-655 with allocate_next, emitHeadParForeach :273-300 with
// par_foreach_builder, nprrRecursiveCall :523-563 with emitBuildCode :302-354,
// emitSetValues :369-380 with set_level; emitOutputs :152-167.
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_TriangleProj(mymap *input_tries) {
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
  Trie<void *, ParMemoryBuffer> *Trie_bag_1_a_b_c_TriangleProj_0_1 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/bag_1_a_b_c_TriangleProj", 2, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_bag_1_a_b_c_TriangleProj_0_1, 2);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_c(Trie_Edge_0_1);
    const size_t count_a = Builders.build_set(Iterators_Edge_a_c.head);
    (void)count_a;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_b_c = Iterators_Edge_b_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_c = Iterators_Edge_a_c.iterators.at(tid);
      Iterator_Edge_a_c->get_next_block(0, a_d);
      const size_t count_b = Builder->build_set(tid, Iterator_Edge_b_c->get_block(0));
      num_rows_reducer.update(tid, count_b);
      Builder->set_level(a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG bag_1_a_b_c_TriangleProj TIME", bag_timer);
    Trie_bag_1_a_b_c_TriangleProj_0_1->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_bag_1_a_b_c_TriangleProj_0_1->num_rows = Builders.trie->num_rows;
    Trie_bag_1_a_b_c_TriangleProj_0_1->encodings = Builders.trie->encodings;
  }
  Trie<void *, ParMemoryBuffer> *Trie_TriangleProj_0_1 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/TriangleProj", 2, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_TriangleProj_0_1, 2);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_b(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_bag_1_a_b_c_TriangleProj_a_b(Trie_bag_1_a_b_c_TriangleProj_0_1);
    const size_t count_a = Builders.build_set(Iterators_Edge_a_b.head, Iterators_bag_1_a_b_c_TriangleProj_a_b.head);
    (void)count_a;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_b = Iterators_Edge_a_b.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_bag_1_a_b_c_TriangleProj_a_b =
          Iterators_bag_1_a_b_c_TriangleProj_a_b.iterators.at(tid);
      Iterator_Edge_a_b->get_next_block(0, a_d);
      Iterator_bag_1_a_b_c_TriangleProj_a_b->get_next_block(0, a_d);
      const size_t count_b = Builder->build_set(tid, Iterator_Edge_a_b->get_block(1),
                                                Iterator_bag_1_a_b_c_TriangleProj_a_b->get_block(1));
      num_rows_reducer.update(tid, count_b);
      Builder->set_level(a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG TriangleProj TIME", bag_timer);
    Trie_TriangleProj_0_1->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_TriangleProj_0_1->num_rows = Builders.trie->num_rows;
    Trie_TriangleProj_0_1->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("TriangleProj_0_1", Trie_TriangleProj_0_1));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// c_run_<id> (QueryPlan.scala:25-39) plus the PTrie read of the output trie.
extern "C" long c_run_TriangleProj(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_TriangleProj(&m);
  Trie<void *, ParMemoryBuffer> *out = (Trie<void *, ParMemoryBuffer> *)m.at("TriangleProj_0_1");
  if (num_rows) *num_rows = out->num_rows;
  delete out;
  return 0;
}

extern "C" long c_rows_TriangleProj(const char *db, uint32_t *rows, unsigned long cap, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_TriangleProj(&m);
  Trie<void *, ParMemoryBuffer> *out = (Trie<void *, ParMemoryBuffer> *)m.at("TriangleProj_0_1");
  if (num_rows) *num_rows = out->num_rows;
  const size_t r = export_rows(out, rows, (double *)nullptr, cap);
  delete out;
  return (long)r;
}
