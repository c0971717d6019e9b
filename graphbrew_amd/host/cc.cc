// cc -- connected components, Afforest (bench/src/cc.cc:333-420) on the GPU.
#include <unordered_map>

#include "drv_common.h"

int main(int argc, char *argv[]) {
  drv::Session s;
  s.cli.name = "connected-components-afforest";
  s.cli.kernel = "cc";  // SetBenchmarkTypeHint (bench/src/cc.cc main)
  if (!s.cli.Parse(argc, argv, "")) return -1;
  s.MakeGraph(false);
  int64_t n;
  int directed;
  gbk_graph_info(s.g, &n, nullptr, &directed, nullptr);
  std::vector<int32_t> comp(n);
  gbk_cc_stats st{};
  s.Benchmark(
      "cc",
      [&](int) {
        s.Check(gbk_cc(s.ctx, s.g, GBK_CC_AFFOREST, comp.data(), &st), "cc");
        return st.device_ms;
      },
      [&](int) {},
      [&]() { drv::PrintCompStats(comp); },
      [&](int) {
        const auto &h = s.Host();
        return gbh_verify_cc(n, h.off.data(), h.nbr.data(), h.io(), h.in(), directed, comp.data()) == 1;
      });
  return 0;
}
