// gz_process_stream -- end-to-end harness over a 9-byte binary stream file, the counterpart of the
// reference's tools/process_stream.cpp:63-142 on the MI355X path.
//
//   gz_process_stream <stream_file> <num_queries> <graph_workers> <reader_threads> [seed]
//
// Reader threads pull records from the file (BinaryFileStream, pread at reserved offsets), fill
// per-reader host gutters and hand them to the device (GraphSketchDriver with DEVICE gutters).  The
// stream is cut into num_queries evenly spaced parts (process_stream_until); after each part
// prep_query + connected_components() is timed and split into the flush of the device gutters and
// Borůvka, as the reference reports it.  Updates/s = stream updates / (ingest start -> end of the
// final flush), process_stream.cpp:130-136.  graph_workers sizes the host-gutter route, which the
// DEVICE route does not use.  The last line is a JSON summary.
#include <sys/resource.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "graphzeppelin_amd.hpp"

int main(int argc, char** argv) {
  if (argc != 5 && argc != 6) {
    std::cerr << "Arguments: stream_file, num_queries, graph_workers, reader_threads [, seed]" << std::endl;
    return EXIT_FAILURE;
  }
  const std::string stream_file = argv[1];
  const int num_queries = std::atoi(argv[2]);
  if (num_queries < 1 || num_queries > 1000) {
    std::cerr << "ERROR: Invalid number of queries! Must be > 0 and <= 1000" << std::endl;
    return EXIT_FAILURE;
  }
  const int num_threads = std::atoi(argv[3]);
  if (num_threads < 1) {
    std::cerr << "ERROR: Invalid number of graph workers! Must be > 0." << std::endl;
    return EXIT_FAILURE;
  }
  const size_t reader_threads = std::max<long>(1, std::atol(argv[4]));
  const size_t seed = argc == 6 ? std::strtoull(argv[5], nullptr, 10)
                                : (size_t)std::chrono::steady_clock::now().time_since_epoch().count();

  try {
    BinaryFileStream stream(stream_file);
    const node_id_t num_nodes = stream.vertices();
    const size_t num_updates = stream.edges();
    std::cout << "Processing stream: " << stream_file << "\nnodes       = " << num_nodes
              << "\nnum_updates = " << num_updates << "\n" << std::endl;

    auto driver_config = DriverConfiguration().gutter_sys(DEVICE).worker_threads(num_threads);
    auto cc_config = CCAlgConfiguration().batch_factor(1.0);
    CCSketchAlg cc_alg{num_nodes, seed, cc_config};
    GraphSketchDriver<CCSketchAlg> driver{&cc_alg, &stream, driver_config, reader_threads};

    struct Q {
      double total, flush, boruvka;
      size_t cc;
    };
    std::vector<Q> qs;
    auto query = [&]() {
      const auto t0 = std::chrono::steady_clock::now();
      driver.prep_query(CONNECTIVITY);
      const size_t cc = cc_alg.connected_components().size();
      const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      const double flush = std::chrono::duration<double>(driver.flush_end - driver.flush_start).count();
      const double bor = std::chrono::duration<double>(cc_alg.cc_alg_end - cc_alg.cc_alg_start).count();
      qs.push_back(Q{total, flush, bor, cc});
      return qs.back();
    };

    const double query_percent = 1.0 / num_queries;
    const auto ins_start = std::chrono::steady_clock::now();
    for (int q = 0; q + 1 < num_queries; ++q) {
      driver.process_stream_until((edge_id_t)((q + 1) * query_percent * num_updates));
      const Q r = query();
      std::cout << "Query " << q + 1 << "\nTotal CC query latency:       " << r.total
                << "\n  Flush Gutters(sec):           " << r.flush << "\n  Boruvka's Algorithm(sec):     " << r.boruvka
                << "\nConnected Components:         " << r.cc << std::endl;
    }
    driver.process_stream_until(END_OF_STREAM);
    const Q r = query();
    const double insert_time = std::chrono::duration<double>(driver.flush_end - ins_start).count();
    rusage ru{};
    getrusage(RUSAGE_SELF, &ru);
    std::cout << "Total insertion time(sec):    " << insert_time << "\nUpdates per second:           "
              << num_updates / insert_time << "\nTotal CC query latency:       " << r.total
              << "\n  Flush Gutters(sec):           " << r.flush << "\n  Boruvka's Algorithm(sec):     " << r.boruvka
              << "\nConnected Components:         " << r.cc
              << "\nMaximum Memory Usage(MiB):    " << ru.ru_maxrss / 1024.0 << std::endl;
    std::printf("{\"updates\": %zu, \"reader_threads\": %zu, \"insert_time_s\": %.6f, \"updates_per_s\": %.1f, \"queries\": [",
                num_updates, reader_threads, insert_time, num_updates / insert_time);
    for (size_t i = 0; i < qs.size(); ++i)
      std::printf("%s{\"total_s\": %.6f, \"flush_s\": %.6f, \"boruvka_s\": %.6f, \"cc\": %zu}", i ? ", " : "", qs[i].total,
                  qs[i].flush, qs[i].boruvka, qs[i].cc);
    std::printf("]}\n");
  } catch (const std::exception& e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return 0;
}
