// Reference-shaped C++ tests against the facade (include/graphzeppelin_amd.hpp): the same calls the
// reference's test/cc_alg_test.cpp makes, driven through GraphSketchDriver<CCSketchAlg>.
// Usage: driver_test <golden_dir>   (exit 0 = all checks passed)
#include <cstdlib>
#include <iostream>
#include <string>

#include "graphzeppelin_amd.hpp"

static int failures = 0;
#define CHECK(cond)                                                              \
  do {                                                                           \
    if (!(cond)) {                                                               \
      std::cerr << __FILE__ << ":" << __LINE__ << ": CHECK failed: " #cond "\n"; \
      ++failures;                                                                \
    }                                                                            \
  } while (0)

// test/cc_alg_test.cpp:41-57 (SmallGraphConnectivity) for both gutter systems
static void small_graph_connectivity(const std::string& dir, gzamd::GutterSystem sys) {
  AsciiFileStream stream{dir + "/multiples_graph_1024.txt", false};
  CCSketchAlg cc_alg{stream.vertices(), 12345};
  GraphSketchDriver<CCSketchAlg> driver(&cc_alg, &stream, DriverConfiguration().gutter_sys(sys));
  driver.process_stream_until(END_OF_STREAM);
  driver.prep_query(CONNECTIVITY);
  CHECK(cc_alg.connected_components().size() == 78);
}

// test/cc_alg_test.cpp:149-176 (TestPointQuery)
static void point_query(const std::string& dir) {
  AsciiFileStream stream{dir + "/multiples_graph_1024.txt", false};
  CCSketchAlg cc_alg{stream.vertices(), 777};
  GraphSketchDriver<CCSketchAlg> driver(&cc_alg, &stream, DriverConfiguration().gutter_sys(STANDALONE));
  driver.process_stream_until(END_OF_STREAM);
  driver.prep_query(CONNECTIVITY);
  auto sets = cc_alg.connected_components().get_component_sets();
  std::vector<node_id_t> ccid(stream.vertices());
  for (node_id_t i = 0; i < sets.size(); ++i)
    for (node_id_t v : sets[i]) ccid[v] = i;
  for (node_id_t i = 0; i < 10; ++i)
    for (node_id_t j = 0; j < 10; ++j) CHECK(cc_alg.point_query(i, j) == (ccid[i] == ccid[j]));
}

// test/cc_alg_test.cpp:223-263 (EagerDSUTest)
static void eager_dsu() {
  CCSketchAlg cc_alg{100, 5};
  const int steps[6][3] = {{1, 2, INSERT}, {2, 3, INSERT}, {1, 3, INSERT}, {1, 3, DELETE}, {2, 3, DELETE}, {2, 3, INSERT}};
  const bool cached[6] = {true, true, true, true, false, true};
  for (int s = 0; s < 6; ++s) {
    cc_alg.update({{(node_id_t)steps[s][0], (node_id_t)steps[s][1]}, (gzamd::UpdateType)steps[s][2]});
    CHECK(cc_alg.has_cached_query(CONNECTIVITY) == cached[s]);
    cc_alg.connected_components();
  }
  CHECK(cc_alg.connected_components().size() == 98);
}

// test/cc_alg_test.cpp:300-350 (MTStreamWithMultipleQueries): 4 stream threads, 10 queries
static void mt_stream_queries(const std::string& dir, gzamd::GutterSystem sys) {
  BinaryFileStream stream{dir + "/multiples_graph_1024_stream.data"};
  CCSketchAlg cc_alg{stream.vertices(), 4242};
  GraphSketchDriver<CCSketchAlg> driver(&cc_alg, &stream, DriverConfiguration().gutter_sys(sys).worker_threads(4), 4);
  const edge_id_t per = stream.edges() / 10;
  for (int q = 0; q < 9; ++q) {
    driver.process_stream_until(per * (q + 1));
    driver.prep_query(CONNECTIVITY);
    cc_alg.connected_components();
  }
  driver.process_stream_until(END_OF_STREAM);
  driver.prep_query(CONNECTIVITY);
  CHECK(cc_alg.connected_components().size() == 78);
  CHECK(driver.get_total_updates() == 2 * stream.edges());
}

// test/sketch_test.cpp:14-18 and 66-80
static void sketch_known_answers() {
  Sketch s1(10, 99, 1, 7);
  CHECK(s1.sample().result == ZERO);
  s1.update(1);
  bool threw = false;
  try {
    s1.sample();
  } catch (const OutOfSamplesException&) {
    threw = true;
  }
  CHECK(threw);
  Sketch s2(40, 3, 1, 7);
  s2.update(0);
  s2.update(0);
  s2.update(0);
  auto r = s2.sample();
  CHECK(r.result == GOOD && r.idx == 0);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: driver_test <golden_dir>\n";
    return 2;
  }
  const std::string dir = argv[1];
  try {
    small_graph_connectivity(dir, DEVICE);
    small_graph_connectivity(dir, STANDALONE);
    point_query(dir);
    eager_dsu();
    mt_stream_queries(dir, DEVICE);
    mt_stream_queries(dir, STANDALONE);
    sketch_known_answers();
  } catch (const std::exception& e) {
    std::cerr << "exception: " << e.what() << "\n";
    return 1;
  }
  std::cout << (failures ? "FAILED" : "all passed") << " (" << failures << " failures)\n";
  return failures ? 1 : 0;
}
