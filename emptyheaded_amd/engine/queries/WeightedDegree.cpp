// WeightedDegree(x;y) :- Edge(x,z;w),y:float<-[SUM(z;w)]: each x with the sum
// of its edges' float annotations, in ascending z.
//
// Hand-written in the generated code's idiom (as PageRankSteps.cpp is), to
// exercise an annotated binary relation: Trie<float,M>::load of a relation
// whose encoded column store carries a float annotation per row, a
// one-operand aggregated level whose set is iterated, and
// TrieIterator::get_annotation(1, z) on the child level
// (storage_engine/src/TrieIterator.cpp:44-59).
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;

static std::string db_folder;

void run_WeightedDegree(mymap *input_tries) {
  thread_pool::initializeThreadPool();
  Trie<float, ParMemoryBuffer> *Trie_Edge_0_1 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_Edge_0_1 = Trie<float, ParMemoryBuffer>::load(db_folder + "/relations/Edge/Edge_0_1");
    timer::stop_clock("LOADING Trie Edge_0_1", start_time);
  }
  par::reducer<size_t> num_rows_reducer(0, [](size_t a, size_t b) { return a + b; });
  auto query_timer = timer::start_clock();
  Trie<float, ParMemoryBuffer> *Trie_WeightedDegree_0 =
      new Trie<float, ParMemoryBuffer>(db_folder + "/relations/WeightedDegree", 1, true);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<float, ParMemoryBuffer> Builders(Trie_WeightedDegree_0, 2);
    ParTrieIterator<float, ParMemoryBuffer> Iterators_Edge_x_z(Trie_Edge_0_1);
    const size_t count_x = Builders.build_set(Iterators_Edge_x_z.head);
    (void)count_x;
    Builders.allocate_annotation();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t x_i, const uint32_t x_d) {
      TrieBuilder<float, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<float, ParMemoryBuffer> *Iterator_Edge_x_z = Iterators_Edge_x_z.iterators.at(tid);
      Iterator_Edge_x_z->get_next_block(0, x_d);
      const size_t count_z = Builder->build_aggregated_set(Iterator_Edge_x_z->get_block(1));
      (void)count_z;
      num_rows_reducer.update(tid, 1);
      float annotation_z = (float)0;
      Builder->foreach_aggregate([&](const uint32_t z_d) {
        const float intermediate_z = (float)1 * Iterator_Edge_x_z->get_annotation(1, z_d);
        annotation_z += intermediate_z;
      });
      Builder->set_annotation(annotation_z, x_i, x_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    timer::stop_clock("BAG WeightedDegree TIME", bag_timer);
  }
  input_tries->insert(std::make_pair("WeightedDegree_0", Trie_WeightedDegree_0));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

extern "C" long c_run_WeightedDegree(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_WeightedDegree(&m);
  Trie<float, ParMemoryBuffer> *out = (Trie<float, ParMemoryBuffer> *)m.at("WeightedDegree_0");
  if (num_rows) *num_rows = out->num_rows;
  return 0;
}

extern "C" long c_rows_WeightedDegree(const char *db, uint32_t *rows, double *annos, unsigned long cap,
                                      unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_WeightedDegree(&m);
  Trie<float, ParMemoryBuffer> *out = (Trie<float, ParMemoryBuffer> *)m.at("WeightedDegree_0");
  if (num_rows) *num_rows = out->num_rows;
  return (long)export_rows(out, rows, annos, cap);
}
