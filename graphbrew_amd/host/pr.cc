// pr -- PageRank pull, Gauss-Seidel (bench/src/pr.cc:136-179) on the GPU.
#include "drv_common.h"

int main(int argc, char *argv[]) {
  drv::Session s;
  s.cli.name = "pagerank";
  s.cli.kernel = "pr";  // SetBenchmarkTypeHint (bench/src/pr.cc main)
  if (!s.cli.Parse(argc, argv, "i:t:F")) return -1;
  s.MakeGraph(false);
  drv::PrintLabel("PR Mode", s.cli.fixed ? "fixed-work" : "convergence");
  int64_t n;
  gbk_graph_info(s.g, &n, nullptr, nullptr, nullptr);
  std::vector<float> scores(n);
  gbk_pr_stats st{};
  s.Benchmark(
      "pr",
      [&](int) {
        s.Check(gbk_pr(s.ctx, s.g, GBK_PR_GS, s.cli.max_iters, s.cli.tol, s.cli.fixed, scores.data(), &st),
                "pr");
        return st.device_ms;
      },
      [&](int) {
        drv::PrintTime("Iterations", st.iterations);
        printf("%-21s%.17g\n", "Final Error:", st.final_error);
      },
      [&]() {  // PrintTopScores (pr.cc:96-115): original ids, top 100
        const auto &h = s.Host();
        std::vector<std::pair<int32_t, float>> kv(n);
        for (int64_t i = 0; i < n; i++) kv[i] = {h.org_ids[i], scores[i]};
        for (const auto &p : drv::TopK(kv, 100)) printf("%d:%g\n", p.second, p.first);
      },
      [&](int) {
        const auto &h = s.Host();
        const double err = gbh_pr_residual(n, h.off.data(), h.nbr.data(), scores.data());
        drv::PrintTime("Total Error", err);
        return err < s.cli.tol;
      });
  return 0;
}
