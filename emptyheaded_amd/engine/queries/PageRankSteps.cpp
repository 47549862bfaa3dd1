// PageRank(x;y) :- Edge(x,z),y:float<-[(1.0/N)].
// PageRank(x;y)*[i=5]:-Edge(x,z),PageRank(z),InvDegree(z),y:float <- [0.15+0.85*SUM(z;1.0)].
// (test/graph/travis.py:585-598, with N(;w) :- Edge(x,y),w:long<-[SUM(x;1)].)
//
// Hand-written, NOT the code generator's output: the generator's emission for
// recursive bags and CONST annotations depends on planner details that cannot
// be run here (no JVM).  The code follows the generator's idiom for the parts
// that are fixed by its templates (EHGenerator.scala: emitLoadRelations
// :171-199, emitHeadBuildCode :565-606, emitHeadAllocations :643-655,
// emitHeadParForeach :273-300, emitBuildCode :302-354 with three accessors,
// emitAnnotationAccessors :398-433 with get_annotation on the annotated
// accessors, emitFinalAnnotation :435-456, emitAnnotationExpression :479-502,
// emitSetValues :369-380 with set_annotation, and the recursion loop :752-764,
// 820-828) so that it exercises the drop-in's annotated-relation calls:
// annotated unary input tries (InvDegree, and each step's PageRank output fed
// back as the next step's input), TrieIterator::get_annotation, a three-way
// build_aggregated_set whose innermost result is iterated, and float
// set_annotation.  The drop-in build turns fp contraction off, so each term is
// rounded to float before it is added and 0.15+0.85*AGG is a double multiply
// then add (the device kernels fuse the latter, as the reference's -O3
// -march=native build of generated code does).
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_PageRankSteps(mymap *input_tries) {
  thread_pool::initializeThreadPool();
  Trie<void *, ParMemoryBuffer> *Trie_Edge_0_1 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_Edge_0_1 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/Edge/Edge_0_1");
    timer::stop_clock("LOADING Trie Edge_0_1", start_time);
  }
  Trie<float, ParMemoryBuffer> *Trie_InvDegree_0 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_InvDegree_0 = Trie<float, ParMemoryBuffer>::load(db_folder + "/relations/InvDegree/InvDegree_0");
    timer::stop_clock("LOADING Trie InvDegree_0", start_time);
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
  // N(;w) :- Edge(x,y),w:long<-[SUM(x;1)]: the number of distinct x
  Trie<long, ParMemoryBuffer> *Trie_N_ = new Trie<long, ParMemoryBuffer>(db_folder + "/relations/N", 0, true);
  long N;
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_N_, 1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_y(Trie_Edge_0_1);
    const size_t count_x = Builders.build_aggregated_set(Iterators_Edge_x_y.head);
    par::reducer<long> annotation_x(0, [&](long a, long b) { return a + b; });
    num_rows_reducer.update(0, count_x);
    const long intermediate_x = (long)1 * (count_x * 1);
    annotation_x.update(0, intermediate_x);
    Builders.trie->annotation = annotation_x.evaluate(0);
    N = Builders.trie->annotation;
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    timer::stop_clock("BAG N TIME", bag_timer);
    Trie_N_->num_rows = 0;
  }
  // base case: PageRank(x;y) :- Edge(x,z),y:float<-[(1.0/N)]
  Trie<float, ParMemoryBuffer> *Trie_PageRank_0 =
      new Trie<float, ParMemoryBuffer>(db_folder + "/relations/PageRank", 1, true);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<float, ParMemoryBuffer> Builders(Trie_PageRank_0, 2);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_z(Trie_Edge_0_1);
    const size_t count_x = Builders.build_set(Iterators_Edge_x_z.head);
    (void)count_x;
    Builders.allocate_annotation();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t x_i, const uint32_t x_d) {
      TrieBuilder<float, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      num_rows_reducer.update(tid, 1);
      float annotation_z = (float)0;
      annotation_z = (1.0 / N);
      Builder->set_annotation(annotation_z, x_i, x_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    timer::stop_clock("BAG PageRank TIME", bag_timer);
  }
  // PageRank(x;y)*[i=5]:-Edge(x,z),PageRank(z),InvDegree(z),y:float <- [0.15+0.85*SUM(z;1.0)]
  {
    size_t num_iterations = 0;
    while (num_iterations < 5) {
      Trie<float, ParMemoryBuffer> *Trie_recursion_0 =
          new Trie<float, ParMemoryBuffer>(db_folder + "/relations/recursion", 1, true);
      auto bag_timer = timer::start_clock();
      num_rows_reducer.clear();
      ParTrieBuilder<float, ParMemoryBuffer> Builders(Trie_recursion_0, 2);
      Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
      ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_z(Trie_Edge_0_1);
      ParTrieIterator<float, ParMemoryBuffer> Iterators_PageRank_z(Trie_PageRank_0);
      ParTrieIterator<float, ParMemoryBuffer> Iterators_InvDegree_z(Trie_InvDegree_0);
      const size_t count_x = Builders.build_set(Iterators_Edge_x_z.head);
      (void)count_x;
      Builders.allocate_annotation();
      Builders.par_foreach_builder([&](const size_t tid, const uint32_t x_i, const uint32_t x_d) {
        TrieBuilder<float, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
        TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_x_z = Iterators_Edge_x_z.iterators.at(tid);
        TrieIterator<float, ParMemoryBuffer> *Iterator_PageRank_z = Iterators_PageRank_z.iterators.at(tid);
        TrieIterator<float, ParMemoryBuffer> *Iterator_InvDegree_z = Iterators_InvDegree_z.iterators.at(tid);
        Iterator_Edge_x_z->get_next_block(0, x_d);
        std::vector<const TrieBlock<hybrid, ParMemoryBuffer> *> z_sets;
        z_sets.push_back(Iterator_Edge_x_z->get_block(1));
        z_sets.push_back(Iterator_PageRank_z->get_block(0));
        z_sets.push_back(Iterator_InvDegree_z->get_block(0));
        const size_t count_z = Builder->build_aggregated_set(&z_sets);
        (void)count_z;
        num_rows_reducer.update(tid, 1);
        float annotation_z = (float)0;
        Builder->foreach_aggregate([&](const uint32_t z_d) {
          const float intermediate_z = (float)1 * 1.0 * Iterator_PageRank_z->get_annotation(0, z_d) *
                                       Iterator_InvDegree_z->get_annotation(0, z_d);
          annotation_z += intermediate_z;
        });
        annotation_z = 0.15 + 0.85 * annotation_z;
        Builder->set_annotation(annotation_z, x_i, x_d);
      });
      Builders.trie->num_rows = num_rows_reducer.evaluate(0);
      timer::stop_clock("BAG recursion TIME", bag_timer);
      Trie_PageRank_0 = Builders.trie;
      num_iterations++;
    }
  }
  input_tries->insert(std::make_pair("PageRank_0", Trie_PageRank_0));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

extern "C" long c_run_PageRankSteps(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_PageRankSteps(&m);
  Trie<float, ParMemoryBuffer> *out = (Trie<float, ParMemoryBuffer> *)m.at("PageRank_0");
  if (num_rows) *num_rows = out->num_rows;
  return 0;
}

extern "C" long c_rows_PageRankSteps(const char *db, uint32_t *rows, double *annos, unsigned long cap,
                                     unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_PageRankSteps(&m);
  Trie<float, ParMemoryBuffer> *out = (Trie<float, ParMemoryBuffer> *)m.at("PageRank_0");
  if (num_rows) *num_rows = out->num_rows;
  return (long)export_rows(out, rows, annos, cap);
}
