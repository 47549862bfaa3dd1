// LollipopAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,x),z:long<-[COUNT(*)].
// BarbellAgg(;w) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,x),Edge(x,y),Edge(y,z),Edge(x,z),w:long<-[COUNT(*)].
// (test/graph/travis.py:14-70)
//
// The query code EmptyHeaded's code generator emits for these two-bag plans,
// restated by hand -- no JVM is available to run the planner or the generator
// -- and compiled unchanged against the drop-in headers.  Only the database
// folder is a parameter (the generator bakes in an absolute path).
//
// The plan.  The GHD planner (query_compiler/src/main/scala/solver/
// GHDSolver.scala, QueryPlanner.scala:64-152) puts Edge(a,x) in the root bag
// and each triangle in a child bag (fractional width 1.5; the lowest tree);
// the attribute order is the root's attributes, then the children's
// (solver/AttrOrderingUtil.scala:33-63): [a, x, b, c] for the lollipop, and
// [a, x, y, z, b, c] for the barbell, whose plan the reference's Query
// Compiler notebook prints (docs/notebooks/Query Compiler.ipynb, cell 4).  A
// child bag keeps the attribute it shares with the root and is annotated by
// the COUNT over the others (solver/GHDNode.scala:121-147, 254-285): the
// triangle COUNT grouped by its first attribute.  The barbell's two triangle
// bags are the same computation, so the IR keeps one, bag_1_x_y_z_BarbellAgg,
// read under both (a, b, c) and (x, y, z) (solver/IROptimizer.scala:22-62).
// The root bag joins Edge(a,x) with the annotated child tries and folds
// COUNT(a, x) times their annotations.
//
// The code.  Per bag, emitNPRR (EHGenerator.scala:728-838): the child bag as
// emitIntermediateTrie :711-726, emitHeadBuildCode :565-606 with build_set,
// emitHeadAllocations :643-655 with allocate_annotation, emitHeadParForeach
// :273-300, nprrRecursiveCall :523-563 over the aggregated attributes with
// emitAnnotationAccessors :398-433, emitFinalAnnotation :435-456 and
// emitAnnotationComputation :457-478, then emitSetValues :369-380 with
// set_annotation; the root bag as emitHeadBuildCode with
// build_aggregated_set, emitInitHeadAnnotations :608-620 (par::reducer),
// emitAnnotationAccessors with get_annotation on the annotated child trie,
// and emitFinalAnnotation, which loops over the last set when an accessor of
// it is annotated (the barbell's x); attribute info per
// codegen/query/QueryPlan.scala:177-245.  Operand order inside a bag follows
// the rule's join list; intersections commute, so it changes no result.
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_LollipopAgg(mymap *input_tries) {
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
  Trie<long, ParMemoryBuffer> *Trie_bag_1_a_b_c_LollipopAgg_0 =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/bag_1_a_b_c_LollipopAgg", 1, true);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_bag_1_a_b_c_LollipopAgg_0, 3);
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
    timer::stop_clock("BAG bag_1_a_b_c_LollipopAgg TIME", bag_timer);
    Trie_bag_1_a_b_c_LollipopAgg_0->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_bag_1_a_b_c_LollipopAgg_0->num_rows = Builders.trie->num_rows;
    Trie_bag_1_a_b_c_LollipopAgg_0->encodings = Builders.trie->encodings;
  }
  Trie<long, ParMemoryBuffer> *Trie_LollipopAgg_ =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/LollipopAgg", 0, true);
  long LollipopAgg;
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_LollipopAgg_, 2);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_x(Trie_Edge_0_1);
    ParTrieIterator<long, ParMemoryBuffer> Iterators_bag_1_a_b_c_LollipopAgg_a(Trie_bag_1_a_b_c_LollipopAgg_0);
    const size_t count_a = Builders.build_aggregated_set(Iterators_Edge_a_x.head, Iterators_bag_1_a_b_c_LollipopAgg_a.head);
    (void)count_a;
    par::reducer<long> annotation_a(0, [&](long a, long b) { return a + b; });
    Builders.par_foreach_aggregate([&](const size_t tid, const uint32_t a_d) {
      TrieBuilder<long, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_x = Iterators_Edge_a_x.iterators.at(tid);
      TrieIterator<long, ParMemoryBuffer> *Iterator_bag_1_a_b_c_LollipopAgg_a =
          Iterators_bag_1_a_b_c_LollipopAgg_a.iterators.at(tid);
      Iterator_Edge_a_x->get_next_block(0, a_d);
      const long intermediate_a = (long)1 * Iterator_bag_1_a_b_c_LollipopAgg_a->get_annotation(0, a_d);
      const size_t count_x = Builder->build_aggregated_set(Iterator_Edge_a_x->get_block(1));
      num_rows_reducer.update(tid, 1);
      long annotation_x = (long)0;
      const long intermediate_x = (long)1 * 1 * intermediate_a * (count_x * 1);
      annotation_x += intermediate_x;
      annotation_a.update(tid, annotation_x);
    });
    Builders.trie->annotation = annotation_a.evaluate(0);
    LollipopAgg = Builders.trie->annotation;
    (void)LollipopAgg;
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG LollipopAgg TIME", bag_timer);
    Trie_LollipopAgg_->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_LollipopAgg_->num_rows = 0;
    Trie_LollipopAgg_->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("LollipopAgg_", Trie_LollipopAgg_));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

void run_BarbellAgg(mymap *input_tries) {
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
  Trie<long, ParMemoryBuffer> *Trie_bag_1_x_y_z_BarbellAgg_0 =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/bag_1_x_y_z_BarbellAgg", 1, true);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_bag_1_x_y_z_BarbellAgg_0, 3);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_y_z(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_z(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_y(Trie_Edge_0_1);
    const size_t count_x = Builders.build_set(Iterators_Edge_x_z.head, Iterators_Edge_x_y.head);
    (void)count_x;
    Builders.allocate_annotation();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t x_i, const uint32_t x_d) {
      TrieBuilder<long, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_y_z = Iterators_Edge_y_z.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_x_z = Iterators_Edge_x_z.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_x_y = Iterators_Edge_x_y.iterators.at(tid);
      Iterator_Edge_x_z->get_next_block(0, x_d);
      Iterator_Edge_x_y->get_next_block(0, x_d);
      const size_t count_y =
          Builder->build_aggregated_set(Iterator_Edge_y_z->get_block(0), Iterator_Edge_x_y->get_block(1));
      (void)count_y;
      long annotation_y = (long)0;
      Builder->foreach_aggregate([&](const uint32_t y_d) {
        Iterator_Edge_y_z->get_next_block(0, y_d);
        const long intermediate_y = (long)1 * 1;
        const size_t count_z =
            Builder->build_aggregated_set(Iterator_Edge_y_z->get_block(1), Iterator_Edge_x_z->get_block(1));
        num_rows_reducer.update(tid, 1);
        long annotation_z = (long)0;
        const long intermediate_z = (long)1 * 1 * 1 * intermediate_y * (count_z * 1);
        annotation_z += intermediate_z;
        annotation_y += annotation_z;
      });
      Builder->set_annotation(annotation_y, x_i, x_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG bag_1_x_y_z_BarbellAgg TIME", bag_timer);
    Trie_bag_1_x_y_z_BarbellAgg_0->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_bag_1_x_y_z_BarbellAgg_0->num_rows = Builders.trie->num_rows;
    Trie_bag_1_x_y_z_BarbellAgg_0->encodings = Builders.trie->encodings;
  }
  Trie<long, ParMemoryBuffer> *Trie_BarbellAgg_ =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/BarbellAgg", 0, true);
  long BarbellAgg;
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_BarbellAgg_, 2);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_x(Trie_Edge_0_1);
    ParTrieIterator<long, ParMemoryBuffer> Iterators_bag_1_x_y_z_BarbellAgg_a(Trie_bag_1_x_y_z_BarbellAgg_0);
    ParTrieIterator<long, ParMemoryBuffer> Iterators_bag_1_x_y_z_BarbellAgg_x(Trie_bag_1_x_y_z_BarbellAgg_0);
    const size_t count_a =
        Builders.build_aggregated_set(Iterators_Edge_a_x.head, Iterators_bag_1_x_y_z_BarbellAgg_a.head);
    (void)count_a;
    par::reducer<long> annotation_a(0, [&](long a, long b) { return a + b; });
    Builders.par_foreach_aggregate([&](const size_t tid, const uint32_t a_d) {
      TrieBuilder<long, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_x = Iterators_Edge_a_x.iterators.at(tid);
      TrieIterator<long, ParMemoryBuffer> *Iterator_bag_1_x_y_z_BarbellAgg_a =
          Iterators_bag_1_x_y_z_BarbellAgg_a.iterators.at(tid);
      TrieIterator<long, ParMemoryBuffer> *Iterator_bag_1_x_y_z_BarbellAgg_x =
          Iterators_bag_1_x_y_z_BarbellAgg_x.iterators.at(tid);
      Iterator_Edge_a_x->get_next_block(0, a_d);
      const long intermediate_a = (long)1 * Iterator_bag_1_x_y_z_BarbellAgg_a->get_annotation(0, a_d);
      const size_t count_x = Builder->build_aggregated_set(Iterator_Edge_a_x->get_block(1),
                                                           Iterator_bag_1_x_y_z_BarbellAgg_x->get_block(0));
      (void)count_x;
      num_rows_reducer.update(tid, 1);
      long annotation_x = (long)0;
      Builder->foreach_aggregate([&](const uint32_t x_d) {
        const long intermediate_x =
            (long)1 * 1 * Iterator_bag_1_x_y_z_BarbellAgg_x->get_annotation(0, x_d) * intermediate_a;
        annotation_x += intermediate_x;
      });
      annotation_a.update(tid, annotation_x);
    });
    Builders.trie->annotation = annotation_a.evaluate(0);
    BarbellAgg = Builders.trie->annotation;
    (void)BarbellAgg;
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG BarbellAgg TIME", bag_timer);
    Trie_BarbellAgg_->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_BarbellAgg_->num_rows = 0;
    Trie_BarbellAgg_->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("BarbellAgg_", Trie_BarbellAgg_));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// Cython's c_run_<id> equivalent (QueryPlan.scala:25-39): runs the query over
// the database folder and returns the bag's annotation; num_rows receives the
// output trie's row count (0 for a scalar bag).
template <void (*RUN)(mymap *)>
static long scalar_of(const char *db, const char *name, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  RUN(&m);
  Trie<long, ParMemoryBuffer> *out = (Trie<long, ParMemoryBuffer> *)m.at(name);
  const long v = out->annotation;
  if (num_rows) *num_rows = out->num_rows;
  delete out;
  return v;
}

extern "C" long c_run_LollipopAgg(const char *db, unsigned long *num_rows) {
  return scalar_of<run_LollipopAgg>(db, "LollipopAgg_", num_rows);
}
extern "C" long c_run_BarbellAgg(const char *db, unsigned long *num_rows) {
  return scalar_of<run_BarbellAgg>(db, "BarbellAgg_", num_rows);
}
