// drv_adaptive.h -- "-o 14" (AdaptiveOrder) and the reference CPU reorderers'
// mapping files for the bench drivers.
//
// -o 14[:_:_:_:<model>[:<criterion>[:<reuse>]]] follows ParseDeployable-
// SelectionPolicy (reorder/reorder_adaptive.h:165-237): three reserved empty
// fields, a model, a criterion and a positive reuse count, nothing after.
// Models:
//   perceptron (default)      the deployable learned model slot: here the
//                             MLP trained on this library's own GPU timings
//                             (selector_model.txt, run by gbk_mlp_forward on
//                             gbk_graph_features); without a model file the
//                             all-kernel rule decides
//   allkernel-lowreuse-rule   ShouldUseAllKernelLowReuseRule (:302-333)
//   budgeted-rule             ShouldUseBudgetedLeidenGorder (:240-257)
// The rules take the Tier-0 features, computed on the device
// (gbk_tier0_features), and the kernel's modeled property bytes over the
// host's LLC, as GenerateAdaptiveMappingFullGraphStandalone does
// (:542-567).  A rule's Leiden-Gorder arm is a mapping of the reference's
// GraphBrew reorderer, which stays CPU code: it is taken from a mapping file
// when one is present, else the rule's fallback arm of builds without Rabbit
// order, DBG (-o 5, :372-380), is used and the substitution is printed.
//
// Mapping files: GBK_MAPPING_DIR/<graph>_<tag>.so (binary int32 new_ids),
// <graph> = kron<scale>_ef<degree> for -g, uniform<scale>_ef<degree> for -u,
// else the input file's name without its suffix.  Tags: rcm_bnf (-o 11:bnf),
// leiden_gorder8 (-o 12:leiden:compose:sg_none:comm_size_desc:intra_gorder:gw8),
// leiden_gorder8_lowreuse (the all-kernel rule's arm, that spec with its cost
// controls), leiden_gorder32_budgeted (the budgeted rule's arm).
#pragma once

#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "gbk.h"
#include "gbk_host.h"

namespace drv {

// Reorderer option strings (everything after "12:") and their file tags.
inline const char *kLeidenGorder8 = "leiden:compose:sg_none:comm_size_desc:intra_gorder:gw8";
inline const char *kLowReuseArm =
    "leiden:compose:sg_none:comm_size_desc:intra_gorder:gw8:cd_parallel:sgmb4096:gordf5000:norefine:2:2";
inline const char *kBudgetedArm = "leiden:compose:sg_none:comm_identity:intra_gorder:gw32:gordf500:cd_parallel:1:1";

inline std::string JoinOpts(const std::vector<std::string> &opt) {
  std::string s;
  for (size_t i = 0; i < opt.size(); i++) s += (i ? ":" : "") + opt[i];
  return s;
}

// The file tag of a reference reorderer's ordering, "" when none is kept.
inline std::string ReordererTag(int algo, const std::vector<std::string> &opt) {
  const std::string j = JoinOpts(opt);
  if (algo == 11 && j == "bnf") return "rcm_bnf";
  if (algo == 12 && j == kLeidenGorder8) return "leiden_gorder8";
  if (algo == 12 && j == kLowReuseArm) return "leiden_gorder8_lowreuse";
  if (algo == 12 && j == kBudgetedArm) return "leiden_gorder32_budgeted";
  return "";
}

inline bool FileExists(const std::string &p) { return access(p.c_str(), R_OK) == 0; }

// BenchmarkType codes (reorder/reorder_types.h:1213-1222) of the drivers.
inline int BenchCode(const std::string &name) {
  if (name == "pr") return 1;
  if (name == "bfs") return 2;
  if (name == "cc") return 3;
  if (name == "sssp") return 4;
  if (name == "bc") return 5;
  if (name == "pr_spmv") return 7;
  if (name == "cc_sv") return 8;
  return 0;
}

struct AdaptivePolicy {
  std::string model = "perceptron";
  std::string criterion = "fastest-execution";
  double reuse = 1.0;
  bool reuse_explicit = false;
};

// ParseDeployableSelectionPolicy (reorder/reorder_adaptive.h:165-237) with
// the model and criterion names of reorder_types.h:1927-1958 and the legacy
// mode names of ParseModelCriterion (:2039-2066).
inline AdaptivePolicy ParseAdaptivePolicy(const std::vector<std::string> &opt) {
  auto fail = [](const std::string &m) -> AdaptivePolicy { Die(m); };
  for (size_t i = 0; i < std::min<size_t>(3, opt.size()); i++)
    if (!opt[i].empty() && opt[i] != "_") fail("AdaptiveOrder reserved fields must be empty");
  const std::string ms = opt.size() > 3 ? opt[3] : "", cs = opt.size() > 4 ? opt[4] : "",
                    rs = opt.size() > 5 ? opt[5] : "";
  for (size_t i = 6; i < opt.size(); i++)
    if (!opt[i].empty() && opt[i] != "_")
      fail("AdaptiveOrder no longer accepts graph identity or additional runtime selection fields");
  auto model_of = [&](const std::string &n) -> std::string {
    if (n == "perceptron" || n == "perc" || n == "p") return "perceptron";
    if (n == "budgeted-rule" || n == "budgeted" || n == "one-use-rule") return "budgeted-rule";
    if (n == "allkernel-lowreuse-rule" || n == "allkernel-rule" || n == "lowreuse-rule")
      return "allkernel-lowreuse-rule";
    if (n == "decision-tree" || n == "dt" || n == "tree") return "decision-tree";
    if (n == "hybrid" || n == "model-tree" || n == "hme") return "hybrid";
    if (n == "knn-database" || n == "knn" || n == "database" || n == "db") return "knn-database";
    if (n == "family" || n == "fam") return "family";
    if (n == "topn" || n == "top-n" || n == "top") return "topn";
    if (n == "individual" || n == "ind" || n == "all") return "individual";
    fail("Unknown adaptive selection model: " + n);
    return "";
  };
  auto crit_of = [&](const std::string &n) -> std::string {
    if (n == "fastest-reorder" || n == "reorder" || n == "fr") return "fastest-reorder";
    if (n == "fastest-execution" || n == "execution" || n == "fe") return "fastest-execution";
    if (n == "best-endtoend" || n == "endtoend" || n == "e2e") return "best-endtoend";
    if (n == "best-amortization" || n == "amortization" || n == "amortize") return "best-amortization";
    fail("Unknown adaptive selection criterion: " + n);
    return "";
  };
  AdaptivePolicy p;
  if (!ms.empty() && ms != "_") {
    if (!cs.empty() && cs != "_") {
      p.model = model_of(ms);
      p.criterion = crit_of(cs);
    } else if (ms == "0" || ms == "fastest-reorder" || ms == "reorder") {
      p.criterion = "fastest-reorder";
    } else if (ms == "1" || ms == "fastest-execution" || ms == "execution" || ms == "cache") {
      p.criterion = "fastest-execution";
    } else if (ms == "2" || ms == "best-endtoend" || ms == "endtoend" || ms == "e2e") {
      p.criterion = "best-endtoend";
    } else if (ms == "3" || ms == "best-amortization" || ms == "amortization" || ms == "amortize") {
      p.criterion = "best-amortization";
    } else {
      p.model = model_of(ms);
    }
  } else if (!cs.empty() && cs != "_") {
    fail("AdaptiveOrder criterion requires an explicit model");
  }
  if (!rs.empty() && rs != "_") {
    size_t used = 0;
    double r = 0;
    try {
      r = std::stod(rs, &used);
    } catch (...) {
      used = 0;
    }
    if (used != rs.size() || !std::isfinite(r) || r <= 0.0) fail("AdaptiveOrder reuse count must be positive");
    p.reuse = r;
    p.reuse_explicit = true;
  }
  if (p.model != "perceptron" && p.model != "budgeted-rule" && p.model != "allkernel-lowreuse-rule")
    fail(p.model + " is offline-only and cannot drive deployable AdaptiveOrder");
  if (p.model != "perceptron" && !p.reuse_explicit) fail(p.model + " requires an explicit reuse count");
  if (p.model != "perceptron" && p.criterion != "best-endtoend")
    fail(p.model + " requires best-endtoend criterion");
  return p;
}

// The learned selector's model as selector.py writes it for the drivers.
struct SelectorModel {
  int in_dim = 0, members = 0, nf = 0, norders = 0, nkernels = 0;
  double margin = 0;
  std::vector<std::string> kernels;
  std::vector<double> mean, stdv;
  std::vector<std::vector<float>> params;

  bool Load(const std::string &path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string magic;
    int version = 0;
    f >> magic >> version >> in_dim >> members >> margin >> nf >> norders >> nkernels;
    if (!f || magic != "gbk-selector" || version != 1 || in_dim != nf + norders + nkernels + 1) return false;
    kernels.resize(nkernels);
    for (auto &k : kernels) f >> k;
    mean.resize(nf);
    stdv.resize(nf);
    for (auto &x : mean) f >> x;
    for (auto &x : stdv) f >> x;
    const int64_t P = 64 * in_dim + 64 + 64 * 64 + 64 + 64 + 1;
    params.assign(members, std::vector<float>(P));
    for (auto &p : params)
      for (auto &x : p) f >> x;
    return static_cast<bool>(f) && nf == GBK_NUM_FEATURES;
  }
};

// Where the drivers find the model: GBK_SELECTOR_MODEL, else
// selector_model.txt in the package directory (the drivers live in its bin/).
inline std::string SelectorModelPath() {
  if (const char *p = std::getenv("GBK_SELECTOR_MODEL")) return p;
  char buf[4096];
  const ssize_t k = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (k <= 0) return "selector_model.txt";
  buf[k] = 0;
  std::string dir(buf);
  dir = dir.substr(0, dir.rfind('/'));           // .../bin
  return dir.substr(0, dir.rfind('/')) + "/selector_model.txt";
}

inline const char *kOrderNames[8] = {"original", "random", "sort", "hubsort",
                                     "hubcluster", "dbg", "hubsortdbg", "hubclusterdbg"};

// Predicted end-to-end time of orderings 0..7 relative to the original
// order (1 for the original), averaged over the ensemble's members in log
// space as selector.py does.  Returns false when the kernel is not modelled.
inline bool LearnedPredict(gbk_ctx *ctx, const gbk_graph *g, const SelectorModel &mdl, const std::string &kernel,
                           double reuse, std::vector<double> &pred) {
  int kidx = -1;
  for (int i = 0; i < mdl.nkernels; i++)
    if (mdl.kernels[i] == kernel) kidx = i;
  if (kidx < 0) return false;
  float f[GBK_NUM_FEATURES];
  if (gbk_graph_features(ctx, g, f, GBK_NUM_FEATURES) < 0) return false;
  const int rows = mdl.norders - 1;
  std::vector<float> X(static_cast<size_t>(rows) * mdl.in_dim, 0.0f);
  for (int r = 0; r < rows; r++) {
    float *x = X.data() + static_cast<size_t>(r) * mdl.in_dim;
    for (int i = 0; i < mdl.nf; i++) x[i] = static_cast<float>((f[i] - mdl.mean[i]) / mdl.stdv[i]);
    x[mdl.nf + r + 1] = 1.0f;
    x[mdl.nf + mdl.norders + kidx] = 1.0f;
    x[mdl.in_dim - 1] = static_cast<float>(reuse - 1.5);
  }
  gbk_mlp *net = nullptr;
  if (gbk_mlp_create(ctx, mdl.in_dim, 0, &net) != GBK_OK) return false;
  std::vector<double> acc(rows, 0.0);
  std::vector<float> out(rows);
  for (const auto &p : mdl.params) {
    if (gbk_mlp_set_params(ctx, net, p.data()) != GBK_OK ||
        gbk_mlp_forward(ctx, net, X.data(), rows, out.data()) != GBK_OK) {
      gbk_mlp_destroy(ctx, net);
      return false;
    }
    for (int r = 0; r < rows; r++) acc[r] += out[r];
  }
  gbk_mlp_destroy(ctx, net);
  pred.assign(mdl.norders, 1.0);
  for (int r = 0; r < rows; r++) pred[r + 1] = std::exp(acc[r] / static_cast<double>(mdl.params.size()));
  return true;
}

}  // namespace drv
