// cc_sv -- connected components, Shiloach-Vishkin (bench/src/cc_sv.cc:205-240) on the GPU.
#include "drv_common.h"

int main(int argc, char *argv[]) {
  drv::Session s;
  s.cli.name = "connected-components";
  s.cli.kernel = "cc_sv";  // SetBenchmarkTypeHint (bench/src/cc_sv.cc main)
  if (!s.cli.Parse(argc, argv, "")) return -1;
  s.MakeGraph(false);
  int64_t n;
  int directed;
  gbk_graph_info(s.g, &n, nullptr, &directed, nullptr);
  std::vector<int32_t> comp(n);
  gbk_cc_stats st{};
  s.Benchmark(
      "cc_sv",
      [&](int) {
        s.Check(gbk_cc(s.ctx, s.g, GBK_CC_SV, comp.data(), &st), "cc_sv");
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
