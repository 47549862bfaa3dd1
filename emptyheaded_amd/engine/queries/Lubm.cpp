// lubm1(a) :- b='http://www.Department0.University0.edu/GraduateCourse0',
//   c='http://www.lehigh.edu/~zhp2/2004/0401/univ-bench.owl#GraduateStudent',
//   takesCourse(a,b),rdftype(a,c).                                  (4 rows)
// lubm6(a) :- b='http://www.lehigh.edu/~zhp2/2004/0401/univ-bench.owl#UndergraduateStudent',
//   rdftype(a,b).                                                   (5,916 rows)
// (test/rdf/travis.py:8-20, 111-122)
//
// The query code EmptyHeaded's code generator emits for these single-bag
// selections over string-typed relations, restated by hand -- no JVM is
// available to run the planner or the generator -- and compiled unchanged
// against the drop-in headers.  Only the database folder is a parameter.
// The planner puts the equality-selected attributes first (solver/
// AttrOrderingUtil.scala:15-18), so each selected relation is read in its
// reverse ordering (rdftype_1_0, takesCourse_1_0) and every selected attribute
// has one accessor: its level becomes a lookup of the constant's key through
// Encoding<string>::value_to_key (emitSelectionValues, EHGenerator.scala:
// 656-666) followed by ParTrieIterator::get_next_block (emitContainsSelection
// / emitHeadContainsSelections :668-700); the output attribute a is then the
// head build of the selected child sets (emitHeadBuildCode :565-606 with
// build_set), one column, so its count is the bag's row count (emitNPRR
// :766-796).
#include "query_io.hpp"

typedef std::unordered_map<std::string, void *> mymap;
typedef std::string string;

static std::string db_folder;

void run_lubm1(mymap *input_tries) {
  thread_pool::initializeThreadPool();
  Trie<void *, ParMemoryBuffer> *Trie_takesCourse_1_0 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_takesCourse_1_0 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/takesCourse/takesCourse_1_0");
    timer::stop_clock("LOADING Trie takesCourse_1_0", start_time);
  }
  Trie<void *, ParMemoryBuffer> *Trie_rdftype_1_0 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_rdftype_1_0 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/rdftype/rdftype_1_0");
    timer::stop_clock("LOADING Trie rdftype_1_0", start_time);
  }
  auto e_loading_string = timer::start_clock();
  Encoding<string> *Encoding_string = Encoding<string>::from_binary(db_folder + "/encodings/string/");
  (void)Encoding_string;
  timer::stop_clock("LOADING ENCODINGS string", e_loading_string);
  par::reducer<size_t> num_rows_reducer(0, [](size_t a, size_t b) { return a + b; });
  //
  // query plan
  //
  auto query_timer = timer::start_clock();
  Trie<void *, ParMemoryBuffer> *Trie_lubm1_0 = new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/lubm1", 1, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_lubm1_0, 3);
    Builders.trie->encodings.push_back((void *)Encoding_string);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_takesCourse_b_a(Trie_takesCourse_1_0);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_rdftype_c_a(Trie_rdftype_1_0);
    const uint32_t selection_b_0 =
        Encoding_string->value_to_key.at("http://www.Department0.University0.edu/GraduateCourse0");
    const uint32_t selection_c_0 =
        Encoding_string->value_to_key.at("http://www.lehigh.edu/~zhp2/2004/0401/univ-bench.owl#GraduateStudent");
    Iterators_takesCourse_b_a.get_next_block(selection_b_0);
    Iterators_rdftype_c_a.get_next_block(selection_c_0);
    const size_t count_a = Builders.build_set(Iterators_takesCourse_b_a.head, Iterators_rdftype_c_a.head);
    num_rows_reducer.update(0, count_a);
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG lubm1 TIME", bag_timer);
    Trie_lubm1_0->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_lubm1_0->num_rows = Builders.trie->num_rows;
    Trie_lubm1_0->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("lubm1_0", Trie_lubm1_0));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

void run_lubm6(mymap *input_tries) {
  thread_pool::initializeThreadPool();
  Trie<void *, ParMemoryBuffer> *Trie_rdftype_1_0 = NULL;
  {
    auto start_time = timer::start_clock();
    Trie_rdftype_1_0 = Trie<void *, ParMemoryBuffer>::load(db_folder + "/relations/rdftype/rdftype_1_0");
    timer::stop_clock("LOADING Trie rdftype_1_0", start_time);
  }
  auto e_loading_string = timer::start_clock();
  Encoding<string> *Encoding_string = Encoding<string>::from_binary(db_folder + "/encodings/string/");
  (void)Encoding_string;
  timer::stop_clock("LOADING ENCODINGS string", e_loading_string);
  par::reducer<size_t> num_rows_reducer(0, [](size_t a, size_t b) { return a + b; });
  //
  // query plan
  //
  auto query_timer = timer::start_clock();
  Trie<void *, ParMemoryBuffer> *Trie_lubm6_0 = new Trie<void *, ParMemoryBuffer>(db_folder + "/relations/lubm6", 1, false);
  {
    auto bag_timer = timer::start_clock();
    num_rows_reducer.clear();
    ParTrieBuilder<void *, ParMemoryBuffer> Builders(Trie_lubm6_0, 2);
    Builders.trie->encodings.push_back((void *)Encoding_string);
    ParTrieIterator<void *, ParMemoryBuffer> Iterators_rdftype_b_a(Trie_rdftype_1_0);
    const uint32_t selection_b_0 =
        Encoding_string->value_to_key.at("http://www.lehigh.edu/~zhp2/2004/0401/univ-bench.owl#UndergraduateStudent");
    Iterators_rdftype_b_a.get_next_block(selection_b_0);
    const size_t count_a = Builders.build_set(Iterators_rdftype_b_a.head);
    num_rows_reducer.update(0, count_a);
    Builders.trie->num_rows = num_rows_reducer.evaluate(0);
    std::cout << "NUM ROWS: " << Builders.trie->num_rows << " ANNOTATION: " << Builders.trie->annotation << std::endl;
    timer::stop_clock("BAG lubm6 TIME", bag_timer);
    Trie_lubm6_0->memoryBuffers = Builders.trie->memoryBuffers;
    Trie_lubm6_0->num_rows = Builders.trie->num_rows;
    Trie_lubm6_0->encodings = Builders.trie->encodings;
  }
  input_tries->insert(std::make_pair("lubm6_0", Trie_lubm6_0));
  timer::stop_clock("QUERY TIME", query_timer);
  thread_pool::deleteThreadPool();
}

// c_run_<id> (QueryPlan.scala:25-39) plus the PTrie read of the output trie,
// whose rows are keys of the string encoding (the DataFrame decodes them
// through Encoding::key_to_value, codegen/trie/Trie.scala:211-217).
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

extern "C" long c_run_lubm1(const char *db, unsigned long *num_rows) {
  rows_of<run_lubm1>(db, "lubm1_0", nullptr, 0, num_rows);
  return 0;
}
extern "C" long c_rows_lubm1(const char *db, uint32_t *rows, unsigned long cap, unsigned long *num_rows) {
  return rows_of<run_lubm1>(db, "lubm1_0", rows, cap, num_rows);
}
extern "C" long c_run_lubm6(const char *db, unsigned long *num_rows) {
  rows_of<run_lubm6>(db, "lubm6_0", nullptr, 0, num_rows);
  return 0;
}
extern "C" long c_rows_lubm6(const char *db, uint32_t *rows, unsigned long cap, unsigned long *num_rows) {
  return rows_of<run_lubm6>(db, "lubm6_0", rows, cap, num_rows);
}
