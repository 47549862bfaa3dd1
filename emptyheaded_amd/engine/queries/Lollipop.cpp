// Lollipop(a,b,c,x) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,x).
// Barbell(a,b,c,x,y,z) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,x),Edge(x,y),Edge(y,z),Edge(x,z).
// (test/graph/travis.py:109-172)
//
// The query code EmptyHeaded's code generator emits for these materializing
// multi-bag plans, restated by hand -- no JVM is available to run the planner
// or the generator -- and compiled unchanged against the drop-in headers.
// Only the database folder is a parameter.
//
// The plan (printed for Barbell by docs/notebooks/Query Compiler.ipynb, cell 4;
// the lollipop's follows the same rules): the root bag holds Edge(a,x), the
// triangle bags hang below it and are materialized whole; the attribute order
// is [a, x, b, c] ([a, x, y, z, b, c] for the barbell, whose two triangle bags
// are one computation, solver/IROptimizer.scala:22-62).  Because the root does
// not cover every output attribute, the planner adds a top-down pass
// (solver/GHD.scala:63-99): the root bag (Lollipop_root(a,x), the edges whose a
// and x end triangles; b, c and y, z projected away) joined back with the
// triangle bag(s) in that attribute order.  Three bags run in order: the
// triangle bag (3 attributes), the root, the top-down pass.  The output trie's
// levels follow the attribute order: Lollipop_0_3_1_2 holds (a, x, b, c),
// Barbell_0_3_4_5_1_2 holds (a, x, y, z, b, c) (QueryPlan.scala:95-100,
// 156-170), and the top-down pass reads the triangle trie three levels deep.
//
// The code, per bag (emitNPRR, EHGenerator.scala:728-838): emitIntermediateTrie
// :711-726, emitHeadBuildCode :565-606 with build_set, emitHeadAllocations
// :643-655 with allocate_next, emitHeadParForeach :273-300 with
// par_foreach_builder, nprrRecursiveCall :523-563 with emitBuildCode :302-354,
// emitAllocateCode :356-367, emitForeach :382-396 with foreach_builder,
// emitSetValues :369-380 with set_level; attribute info per
// codegen/query/QueryPlan.scala:177-245 (the root bag's projected attributes
// carry no work and are dropped, :244); emitOutputs :152-167 for the output
// relation only.
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_Lollipop(mymap *input_tries) {
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
  Trie<void *, ParMemoryBuffer> *Trie_bag_1_a_b_c_Lollipop_0_1_2 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/bag_1_a_b_c_Lollipop", 3, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_bag_1_a_b_c_Lollipop_0_1_2, 3);
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
    timer::stop_clock("BAG bag_1_a_b_c_Lollipop TIME", bag_timer);
    Trie_bag_1_a_b_c_Lollipop_0_1_2->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_bag_1_a_b_c_Lollipop_0_1_2->num_rows = Builders.trie->num_rows;
    Trie_bag_1_a_b_c_Lollipop_0_1_2->encodings = Builders.trie->encodings;
  }
  Trie<void *, ParMemoryBuffer> *Trie_Lollipop_root_0_1 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/Lollipop_root", 2, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_Lollipop_root_0_1, 2);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_x(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_bag_1_a_b_c_Lollipop_a_b_c(Trie_bag_1_a_b_c_Lollipop_0_1_2);
    const size_t count_a = Builders.build_set(Iterators_Edge_a_x.head, Iterators_bag_1_a_b_c_Lollipop_a_b_c.head);
    (void)count_a;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_x = Iterators_Edge_a_x.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_bag_1_a_b_c_Lollipop_a_b_c =
          Iterators_bag_1_a_b_c_Lollipop_a_b_c.iterators.at(tid);
      Iterator_Edge_a_x->get_next_block(0, a_d);
      Iterator_bag_1_a_b_c_Lollipop_a_b_c->get_next_block(0, a_d);
      const size_t count_x = Builder->build_set(tid, Iterator_Edge_a_x->get_block(1));
      num_rows_reducer.update(tid, count_x);
      Builder->set_level(a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG Lollipop_root TIME", bag_timer);
    Trie_Lollipop_root_0_1->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_Lollipop_root_0_1->num_rows = Builders.trie->num_rows;
    Trie_Lollipop_root_0_1->encodings = Builders.trie->encodings;
  }
  Trie<void *, ParMemoryBuffer> *Trie_Lollipop_0_3_1_2 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/Lollipop", 4, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_Lollipop_0_3_1_2, 4);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Lollipop_root_a_x(Trie_Lollipop_root_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_bag_1_a_b_c_Lollipop_a_b_c(Trie_bag_1_a_b_c_Lollipop_0_1_2);
    const size_t count_a =
        Builders.build_set(Iterators_Lollipop_root_a_x.head, Iterators_bag_1_a_b_c_Lollipop_a_b_c.head);
    (void)count_a;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Lollipop_root_a_x = Iterators_Lollipop_root_a_x.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_bag_1_a_b_c_Lollipop_a_b_c =
          Iterators_bag_1_a_b_c_Lollipop_a_b_c.iterators.at(tid);
      Iterator_Lollipop_root_a_x->get_next_block(0, a_d);
      Iterator_bag_1_a_b_c_Lollipop_a_b_c->get_next_block(0, a_d);
      const size_t count_x = Builder->build_set(tid, Iterator_Lollipop_root_a_x->get_block(1));
      (void)count_x;
      Builder->allocate_next(tid);
      Builder->foreach_builder([&](const uint32_t x_i, const uint32_t x_d) {
        const size_t count_b = Builder->build_set(tid, Iterator_bag_1_a_b_c_Lollipop_a_b_c->get_block(1));
        (void)count_b;
        Builder->allocate_next(tid);
        Builder->foreach_builder([&](const uint32_t b_i, const uint32_t b_d) {
          Iterator_bag_1_a_b_c_Lollipop_a_b_c->get_next_block(1, b_d);
          const size_t count_c = Builder->build_set(tid, Iterator_bag_1_a_b_c_Lollipop_a_b_c->get_block(2));
          num_rows_reducer.update(tid, count_c);
          Builder->set_level(b_i, b_d);
        });
        Builder->set_level(x_i, x_d);
      });
      Builder->set_level(a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG Lollipop TIME", bag_timer);
    Trie_Lollipop_0_3_1_2->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_Lollipop_0_3_1_2->num_rows = Builders.trie->num_rows;
    Trie_Lollipop_0_3_1_2->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("Lollipop_0_3_1_2", Trie_Lollipop_0_3_1_2));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

void run_Barbell(mymap *input_tries) {
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
  Trie<void *, ParMemoryBuffer> *Trie_bag_1_x_y_z_Barbell_0_1_2 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/bag_1_x_y_z_Barbell", 3, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_bag_1_x_y_z_Barbell_0_1_2, 3);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_y_z(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_z(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_y(Trie_Edge_0_1);
    const size_t count_x = Builders.build_set(Iterators_Edge_x_z.head, Iterators_Edge_x_y.head);
    (void)count_x;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t x_i, const uint32_t x_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_y_z = Iterators_Edge_y_z.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_x_z = Iterators_Edge_x_z.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_x_y = Iterators_Edge_x_y.iterators.at(tid);
      Iterator_Edge_x_z->get_next_block(0, x_d);
      Iterator_Edge_x_y->get_next_block(0, x_d);
      const size_t count_y =
          Builder->build_set(tid, Iterator_Edge_y_z->get_block(0), Iterator_Edge_x_y->get_block(1));
      (void)count_y;
      Builder->allocate_next(tid);
      Builder->foreach_builder([&](const uint32_t y_i, const uint32_t y_d) {
        Iterator_Edge_y_z->get_next_block(0, y_d);
        const size_t count_z =
            Builder->build_set(tid, Iterator_Edge_y_z->get_block(1), Iterator_Edge_x_z->get_block(1));
        num_rows_reducer.update(tid, count_z);
        Builder->set_level(y_i, y_d);
      });
      Builder->set_level(x_i, x_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG bag_1_x_y_z_Barbell TIME", bag_timer);
    Trie_bag_1_x_y_z_Barbell_0_1_2->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_bag_1_x_y_z_Barbell_0_1_2->num_rows = Builders.trie->num_rows;
    Trie_bag_1_x_y_z_Barbell_0_1_2->encodings = Builders.trie->encodings;
  }
  Trie<void *, ParMemoryBuffer> *Trie_Barbell_root_0_1 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/Barbell_root", 2, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_Barbell_root_0_1, 2);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_x(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_bag_1_x_y_z_Barbell_a_b_c(Trie_bag_1_x_y_z_Barbell_0_1_2);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_bag_1_x_y_z_Barbell_x_y_z(Trie_bag_1_x_y_z_Barbell_0_1_2);
    const size_t count_a = Builders.build_set(Iterators_Edge_a_x.head, Iterators_bag_1_x_y_z_Barbell_a_b_c.head);
    (void)count_a;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_x = Iterators_Edge_a_x.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_bag_1_x_y_z_Barbell_a_b_c =
          Iterators_bag_1_x_y_z_Barbell_a_b_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_bag_1_x_y_z_Barbell_x_y_z =
          Iterators_bag_1_x_y_z_Barbell_x_y_z.iterators.at(tid);
      Iterator_Edge_a_x->get_next_block(0, a_d);
      Iterator_bag_1_x_y_z_Barbell_a_b_c->get_next_block(0, a_d);
      const size_t count_x =
          Builder->build_set(tid, Iterator_Edge_a_x->get_block(1), Iterator_bag_1_x_y_z_Barbell_x_y_z->get_block(0));
      num_rows_reducer.update(tid, count_x);
      Builder->set_level(a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG Barbell_root TIME", bag_timer);
    Trie_Barbell_root_0_1->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_Barbell_root_0_1->num_rows = Builders.trie->num_rows;
    Trie_Barbell_root_0_1->encodings = Builders.trie->encodings;
  }
  Trie<void *, ParMemoryBuffer> *Trie_Barbell_0_3_4_5_1_2 =
      new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/Barbell", 6, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_Barbell_0_3_4_5_1_2, 6);
    for (int k = 0; k < 6; k++) Builders.trie->encodings.push_back((void *)Encoding_uint32_t);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Barbell_root_a_x(Trie_Barbell_root_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_bag_1_x_y_z_Barbell_a_b_c(Trie_bag_1_x_y_z_Barbell_0_1_2);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_bag_1_x_y_z_Barbell_x_y_z(Trie_bag_1_x_y_z_Barbell_0_1_2);
    const size_t count_a =
        Builders.build_set(Iterators_Barbell_root_a_x.head, Iterators_bag_1_x_y_z_Barbell_a_b_c.head);
    (void)count_a;
    Builders.allocate_next();
    Builders.par_foreach_builder([&](const size_t tid, const uint32_t a_i, const uint32_t a_d) {
      TrieBuilder<void *, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Barbell_root_a_x = Iterators_Barbell_root_a_x.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_bag_1_x_y_z_Barbell_a_b_c =
          Iterators_bag_1_x_y_z_Barbell_a_b_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_bag_1_x_y_z_Barbell_x_y_z =
          Iterators_bag_1_x_y_z_Barbell_x_y_z.iterators.at(tid);
      Iterator_Barbell_root_a_x->get_next_block(0, a_d);
      Iterator_bag_1_x_y_z_Barbell_a_b_c->get_next_block(0, a_d);
      const size_t count_x = Builder->build_set(tid, Iterator_Barbell_root_a_x->get_block(1),
                                                Iterator_bag_1_x_y_z_Barbell_x_y_z->get_block(0));
      (void)count_x;
      Builder->allocate_next(tid);
      Builder->foreach_builder([&](const uint32_t x_i, const uint32_t x_d) {
        Iterator_bag_1_x_y_z_Barbell_x_y_z->get_next_block(0, x_d);
        const size_t count_y = Builder->build_set(tid, Iterator_bag_1_x_y_z_Barbell_x_y_z->get_block(1));
        (void)count_y;
        Builder->allocate_next(tid);
        Builder->foreach_builder([&](const uint32_t y_i, const uint32_t y_d) {
          Iterator_bag_1_x_y_z_Barbell_x_y_z->get_next_block(1, y_d);
          const size_t count_z = Builder->build_set(tid, Iterator_bag_1_x_y_z_Barbell_x_y_z->get_block(2));
          (void)count_z;
          Builder->allocate_next(tid);
          Builder->foreach_builder([&](const uint32_t z_i, const uint32_t z_d) {
            const size_t count_b = Builder->build_set(tid, Iterator_bag_1_x_y_z_Barbell_a_b_c->get_block(1));
            (void)count_b;
            Builder->allocate_next(tid);
            Builder->foreach_builder([&](const uint32_t b_i, const uint32_t b_d) {
              Iterator_bag_1_x_y_z_Barbell_a_b_c->get_next_block(1, b_d);
              const size_t count_c = Builder->build_set(tid, Iterator_bag_1_x_y_z_Barbell_a_b_c->get_block(2));
              num_rows_reducer.update(tid, count_c);
              Builder->set_level(b_i, b_d);
            });
            Builder->set_level(z_i, z_d);
          });
          Builder->set_level(y_i, y_d);
        });
        Builder->set_level(x_i, x_d);
      });
      Builder->set_level(a_i, a_d);
    });
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG Barbell TIME", bag_timer);
    Trie_Barbell_0_3_4_5_1_2->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_Barbell_0_3_4_5_1_2->num_rows = Builders.trie->num_rows;
    Trie_Barbell_0_3_4_5_1_2->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("Barbell_0_3_4_5_1_2", Trie_Barbell_0_3_4_5_1_2));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// c_run_<id> (QueryPlan.scala:25-39) plus the PTrie read of the output trie:
// num_rows through the pointer, up to `cap` rows copied in the output trie's
// level order ((a, x, b, c) and (a, x, y, z, b, c)).
template <void (*RUN)(mymap *)>
static long rows_of(const char *db, const char *name, uint32_t *rows, unsigned long cap, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  RUN(&m);
  Trie<void *, ParMemoryBuffer> *out = (Trie<void *, ParMemoryBuffer> *)m.at(name);
  if (num_rows) *num_rows = out->num_rows;
  const size_t r = rows ? export_rows(out, rows, (double *)nullptr, cap) : 0;
  delete out;
  return (long)r;
}

extern "C" long c_run_Lollipop(const char *db, unsigned long *num_rows) {
  rows_of<run_Lollipop>(db, "Lollipop_0_3_1_2", nullptr, 0, num_rows);
  return 0;
}
extern "C" long c_rows_Lollipop(const char *db, uint32_t *rows, unsigned long cap, unsigned long *num_rows) {
  return rows_of<run_Lollipop>(db, "Lollipop_0_3_1_2", rows, cap, num_rows);
}
extern "C" long c_run_Barbell(const char *db, unsigned long *num_rows) {
  rows_of<run_Barbell>(db, "Barbell_0_3_4_5_1_2", nullptr, 0, num_rows);
  return 0;
}
extern "C" long c_rows_Barbell(const char *db, uint32_t *rows, unsigned long cap, unsigned long *num_rows) {
  return rows_of<run_Barbell>(db, "Barbell_0_3_4_5_1_2", rows, cap, num_rows);
}
