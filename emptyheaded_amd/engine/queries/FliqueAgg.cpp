// FliqueAgg(;z) :- Edge(a,b),Edge(b,c),Edge(a,c),Edge(a,d),Edge(b,d),Edge(c,d),z:long<-[COUNT(*)].
//
// The query code EmptyHeaded's code generator emits for this rule
// (EHGenerator.scala emitHeadBuildCode :565-606 and emitBuildCode :302-354 with
// three accessors per level), restated by hand and compiled unchanged against
// the drop-in headers; the database folder is a parameter.
#include "emptyheaded.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_FliqueAgg(mymap *input_tries) {
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
  auto query_timer = timer::start_clock();
  Trie<long, ParMemoryBuffer> *Trie_FliqueAgg_ =
      new Trie<long, ParMemoryBuffer>(db_folder + "/relations/FliqueAgg", 0, true);
  long FliqueAgg;
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<long, ParMemoryBuffer> Builders(Trie_FliqueAgg_, 4);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_b(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_c(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_a_d(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_b_d(Trie_Edge_0_1);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_Edge_c_d(Trie_Edge_0_1);
    std::vector<const TrieBlock<hybrid, ParMemoryBuffer> *> a_sets;
    a_sets.push_back(Iterators_Edge_a_b.head);
    a_sets.push_back(Iterators_Edge_a_c.head);
    a_sets.push_back(Iterators_Edge_a_d.head);
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
      const long intermediate_a = (long)1;
      std::vector<const TrieBlock<hybrid, ParMemoryBuffer> *> b_sets;
      b_sets.push_back(Iterator_Edge_a_b->get_block(1));
      b_sets.push_back(Iterator_Edge_b_c->get_block(0));
      b_sets.push_back(Iterator_Edge_b_d->get_block(0));
      const size_t count_b = Builder->build_aggregated_set(&b_sets);
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
    FliqueAgg = Builders.trie->annotation;
    (void)FliqueAgg;
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG FliqueAgg TIME", bag_timer);
    Trie_FliqueAgg_->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_FliqueAgg_->num_rows = 0;
    Trie_FliqueAgg_->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("FliqueAgg_", Trie_FliqueAgg_));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

extern "C" long c_run_FliqueAgg(const char *db, unsigned long *num_rows) {
  db_folder = db;
  mymap m;
  run_FliqueAgg(&m);
  Trie<long, ParMemoryBuffer> *out = (Trie<long, ParMemoryBuffer> *)m.at("FliqueAgg_");
  long v = out->annotation;
  if (num_rows) *num_rows = out->num_rows;
  delete out;
  return v;
}
