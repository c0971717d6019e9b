// sssp -- delta-stepping SSSP (bench/src/sssp.cc:326-420) on the GPU.
#include "drv_common.h"

static std::string hex2(const uint64_t h[2]) {
  char buf[40];
  snprintf(buf, sizeof(buf), "%016" PRIx64 "%016" PRIx64, h[0], h[1]);
  return buf;
}

int main(int argc, char *argv[]) {
  drv::Session s;
  s.cli.name = "single-source shortest-path";
  s.cli.kernel = "sssp";  // SetBenchmarkTypeHint (bench/src/sssp.cc main)
  if (!s.cli.Parse(argc, argv, "d:")) return -1;
  s.MakeGraph(true);
  int64_t n;
  gbk_graph_info(s.g, &n, nullptr, nullptr, nullptr);
  const auto &h = s.Host();
  uint64_t wc[2];
  gbh_weight_checksum(n, h.off.data(), h.nbr.data(), h.wgt.data(), h.org_ids.data(), wc);
  drv::PrintLabel("Weight Scheme", s.cli.weight_scheme);
  drv::PrintLabel("Weight Checksum", hex2(wc));
  drv::PrintLabel("Delta", std::to_string(s.cli.delta));
  const auto srcs = s.Sources(s.cli.trials);
  std::vector<int32_t> dist(n);
  gbk_sssp_stats st{};
  s.Benchmark(
      "sssp",
      [&](int t) {
        s.Check(gbk_sssp(s.ctx, s.g, srcs[t].second, s.cli.delta, dist.data(), &st), "sssp");
        return st.device_ms;
      },
      [&](int t) {
        s.PrintSource(srcs[t]);
        uint64_t df[2];
        gbh_distance_fingerprint(n, dist.data(), s.Host().org_ids.data(), df);
        drv::PrintLabel("Distance Fingerprint", hex2(df));
      },
      [&]() { printf("SSSP Tree reaches %lld nodes\n", static_cast<long long>(st.reached)); },
      [&](int t) {
        const auto &hg = s.Host();
        return gbh_verify_sssp(n, hg.off.data(), hg.nbr.data(), hg.wgt.data(), srcs[t].second,
                               dist.data()) == 0;
      });
  return 0;
}
