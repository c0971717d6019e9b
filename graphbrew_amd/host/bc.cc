// bc -- Brandes betweenness centrality (bench/src/bc.cc:329-400) on the GPU.
// Each trial runs -i sources (CLIterApp, default 1); sources come from the
// SourcePicker over trials * iters original ids.
#include "drv_common.h"

int main(int argc, char *argv[]) {
  drv::Session s;
  s.cli.name = "betweenness-centrality";
  s.cli.kernel = "bc";  // SetBenchmarkTypeHint (bench/src/bc.cc main)
  s.cli.max_iters = 1;
  if (!s.cli.Parse(argc, argv, "i:")) return -1;
  if (!s.cli.sources.empty() && s.cli.max_iters != 1)
    drv::Die("Explicit BC source lists require one Brandes source per trial");
  s.MakeGraph(false);
  int64_t n;
  gbk_graph_info(s.g, &n, nullptr, nullptr, nullptr);
  const int iters = s.cli.max_iters;
  const auto srcs = s.Sources(s.cli.sources.empty() ? s.cli.trials * iters : s.cli.trials);
  std::vector<float> scores(n);
  std::vector<int32_t> batch(iters);
  gbk_bc_stats st{};
  s.Benchmark(
      "bc",
      [&](int t) {
        for (int i = 0; i < iters; i++) batch[i] = srcs[t * iters + i].second;
        s.Check(gbk_bc(s.ctx, s.g, batch.data(), iters, scores.data(), &st), "bc");
        return st.device_ms;
      },
      [&](int t) {
        for (int i = 0; i < iters; i++) s.PrintSource(srcs[t * iters + i]);
      },
      [&]() {  // PrintTopScores (bc.cc:240-250): top 5
        std::vector<std::pair<int32_t, float>> kv(n);
        for (int64_t i = 0; i < n; i++) kv[i] = {static_cast<int32_t>(i), scores[i]};
        for (const auto &p : drv::TopK(kv, 5)) printf("%d:%g\n", p.second, p.first);
      },
      [&](int t) {
        const auto &h = s.Host();
        for (int i = 0; i < iters; i++) batch[i] = srcs[t * iters + i].second;
        double worst = 0;
        const bool ok = gbh_verify_bc(n, h.off.data(), h.nbr.data(), batch.data(), iters, scores.data(),
                                      &worst) == 1;
        char buf[32];
        snprintf(buf, sizeof(buf), "%.9e", worst);
        drv::PrintLabel("BC Verify Max Abs Delta", buf);
        return ok;
      });
  return 0;
}
