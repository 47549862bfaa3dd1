// FliqueSelAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,d),Edge(b,d),Edge(c,d),Edge(a,x),x=0,
//                     z:long<-[COUNT(*)].
// (test/graph/travis.py:385-400 on facebook_duplicated: 3,759,972.)
//
// Hand-written in the generated code's idiom, NOT the generator's output (the
// planner's aggregation chain for a selected attribute cannot be run here):
// the selection follows emitSelectionValues / emitContainsSelection /
// emitHeadContainsSelections (EHGenerator.scala:656-700) -- the constant through
// Encoding::value_to_key, then Iterators_Edge_x_a.get_next_block(selection) on
// the trie ordered (x, a) -- and the rest follows emitHeadBuildCode with four
// accessors, emitBuildCode with two and three, and the COUNT fold as in
// FliqueAgg.cpp.
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;

static std::string db_folder;

void run_FliqueSelAgg(mymap *input_tries) {
  thread_pool::initializeThreadPool();
  Trie<void *, ParMemoryBuffer> *Trie_Edge_0_1 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_Edge_0_1 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/Edge/Edge_0_1");
    timer::stop_clock("LOADING Trie Edge_0_1", start_time);
  }
  Trie<void *, ParMemoryBuffer> *Trie_Edge_1_0 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_Edge_1_0 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/Edge/Edge_1_0");
    timer::stop_clock("LOADING Trie Edge_1_0", start_time);
  }
  auto e_loading_uint32_t = timer::start_clock();
  Encoding<uint32_t> *Encoding_uint32_t = Encoding<uint32_t>::from_binary(db_folder + "/encodings/uint32_t/");
  timer::stop_clock("LOADING ENCODINGS uint32_t", e_loading_uint32_t);
  par::reducer<size_t> num_rows_reducer(0, [](size_t a, size_t b) { return a + b; });
  auto query_timer = timer::start_clock();
  Trie<long, ParMemoryBuffer> *Trie_FliqueSelAgg_ =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/FliqueSelAgg", 0, true);
  long FliqueSelAgg;
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_FliqueSelAgg_, 5);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_b(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_d(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_d(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_c_d(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_x_a(Trie_Edge_1_0);
    const uint32_t selection_x_0 = Encoding_uint32_t->value_to_key.at(0);
    Iterators_Edge_x_a.get_next_block(selection_x_0);
    long annotation_x = (long)1;
    std::vector<const TrieBlock<hybrid, ParMemoryBuffer> *> a_sets;
    a_sets.push_back(Iterators_Edge_a_b.head);
    a_sets.push_back(Iterators_Edge_a_c.head);
    a_sets.push_back(Iterators_Edge_a_d.head);
    a_sets.push_back(Iterators_Edge_x_a.head);
    const size_t count_a = Builders.build_aggregated_set(&a_sets);
    (void)count_a;
    par::reducer<long> annotation_a(0, [&](long a, long b) { return a + b; });
    Builders.par_foreach_aggregate([&](const size_t tid, const uint32_t a_d) {
      TrieBuilder<long, ParMemoryBuffer> *Builder = Builders.builders.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_b = Iterators_Edge_a_b.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_b_c = Iterators_Edge_b_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_c = Iterators_Edge_a_c.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_a_d = Iterators_Edge_a_d.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_b_d = Iterators_Edge_b_d.iterators.at(tid);
      TrieIterator<void *, ParMemoryBuffer> *Iterator_Edge_c_d = Iterators_Edge_c_d.iterators.at(tid);
      Iterator_Edge_a_b->get_next_block(0, a_d);
      Iterator_Edge_a_c->get_next_block(0, a_d);
      Iterator_Edge_a_d->get_next_block(0, a_d);
      const long intermediate_a = (long)1 * 1 * annotation_x;
      const size_t count_b =
          Builder->build_aggregated_set(Iterator_Edge_a_b->get_block(1), Iterator_Edge_b_c->get_block(0));
      (void)count_b;
      long annotation_b = (long)0;
      Builder->foreach_aggregate([&](const uint32_t b_d) {
        Iterator_Edge_b_c->get_next_block(0, b_d);
        Iterator_Edge_b_d->get_next_block(0, b_d);
        const long intermediate_b = (long)1 * 1 * intermediate_a;
        std::vector<const TrieBlock<hybrid, ParMemoryBuffer> *> c_sets;
        c_sets.push_back(Iterator_Edge_b_c->get_block(1));
        c_sets.push_back(Iterator_Edge_a_c->get_block(1));
        c_sets.push_back(Iterator_Edge_c_d->get_block(0));
        const size_t count_c = Builder->build_aggregated_set(&c_sets);
        (void)count_c;
        long annotation_c = (long)0;
        Builder->foreach_aggregate([&](const uint32_t c_d) {
          Iterator_Edge_c_d->get_next_block(0, c_d);
          const long intermediate_c = (long)1 * 1 * 1 * intermediate_b;
          std::vector<const TrieBlock<hybrid, ParMemoryBuffer> *> d_sets;
          d_sets.push_back(Iterator_Edge_a_d->get_block(1));
          d_sets.push_back(Iterator_Edge_b_d->get_block(1));
          d_sets.push_back(Iterator_Edge_c_d->get_block(1));
          const size_t count_d = Builder->build_aggregated_set(&d_sets);
          num_rows_reducer.update(tid, 1);
          long annotation_d = (long)0;
          const long intermediate_d = (long)1 * 1 * 1 * 1 * intermediate_c * (count_d * 1);
          annotation_d += intermediate_d;
          annotation_c += annotation_d;
        });
        annotation_b += annotation_c;
      });
      annotation_a.update(tid, annotation_b);
    });
    Builders.trie->annotation = annotation_a.evaluate(0);
    FliqueSelAgg = Builders.trie->annotation;
    (void)FliqueSelAgg;
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    timer::stop_clock("BAG FliqueSelAgg TIME", bag_timer);
    Trie_FliqueSelAgg_->num_rows = 0;
  }
  input_tries->insert(std::make_pair("FliqueSelAgg_", Trie_FliqueSelAgg_));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

extern "C" long c_run_FliqueSelAgg(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_FliqueSelAgg(&m);
  Trie<long, ParMemoryBuffer> *out = (Trie<long, ParMemoryBuffer> *)m.at("FliqueSelAgg_");
  long v = out->annotation;
  if (num_rows) *num_rows = out->num_rows;
  delete out;
  return v;
}
