// bfs -- direction-optimizing BFS (bench/src/bfs.cc:272-343) on the GPU.
#include "drv_common.h"

int main(int argc, char *argv[]) {
  drv::Session s;
  s.cli.name = "breadth-first search";
  s.cli.kernel = "bfs";  // SetBenchmarkTypeHint (bench/src/bfs.cc main)
  if (!s.cli.Parse(argc, argv, "")) return -1;
  s.MakeGraph(false);
  int64_t n;
  gbk_graph_info(s.g, &n, nullptr, nullptr, nullptr);
  const auto srcs = s.Sources(s.cli.trials);
  std::vector<int32_t> parent(n);
  gbk_bfs_stats st{};
  s.Benchmark(
      "bfs",
      [&](int t) {
        s.Check(gbk_bfs(s.ctx, s.g, srcs[t].second, 0, 0, parent.data(), &st), "bfs");
        return st.device_ms;
      },
      [&](int t) { s.PrintSource(srcs[t]); },
      [&]() {  // PrintBFSStats (graphbrew/bfs_common.h:11-29)
        printf("BFS Tree has %lld nodes and %lld edges\n", static_cast<long long>(st.reached),
               static_cast<long long>(st.edges_reached));
      },
      [&](int t) {
        const auto &h = s.Host();
        return gbh_verify_bfs(n, h.off.data(), h.nbr.data(), h.io(), h.in(), srcs[t].second,
                              parent.data()) == 1;
      });
  return 0;
}
