// drv_common.h -- shared host side of the seven bench drivers (pr, pr_spmv,
// bfs, cc, cc_sv, sssp, bc).  Mirrors the reference binaries' interface:
//   CLI      getopt flags of CLBase/CLApp/CLPageRank/CLDelta/CLIterApp
//            (gapbs/command_line.h:41, 333, 504-640); -o "id[:opt...]"
//   build    -g/-u synthetic (always symmetrized) or -f .el/.wel/.sg/.wsg, then the
//            orderings of -o in sequence (MakeGraph, gapbs/builder.h:715-1002)
//   harness  BenchmarkKernel (gapbs/benchmark.h:435-545): per-trial
//            "Trial Time", optional analysis / verification, "Average Time",
//            exit 3 on a failed verification
//   output   PrintTime "%-21s%3.5lf" and PrintLabel "%-21s%7s" (gapbs/util.h:60-68)
// The graph lives in HBM (libgbk.so); host copies are made only for the
// verifiers and the analysis printouts, as the reference keeps everything on
// the host.  Orderings produced by the reference's own reorderers (Rabbit,
// GOrder, COrder, RCM, GraphBrew/Leiden, ...) enter through -o 13:<file>, or
// for -o 11:bnf and the GraphBrew specs of drv_adaptive.h as mapping files
// under GBK_MAPPING_DIR; -o 14 is the adaptive selector (drv_adaptive.h).
#pragma once

#include <execinfo.h>
#include <getopt.h>
#include <signal.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "gbk.h"
#include "gbk_host.h"

namespace drv {

inline void PrintLabel(const std::string &label, const std::string &val) {
  printf("%-21s%7s\n", (label + ":").c_str(), val.c_str());
}
inline void PrintTime(const std::string &s, double seconds) {
  printf("%-21s%3.5lf\n", (s + ":").c_str(), seconds);
}

struct Timer {
  std::chrono::high_resolution_clock::time_point t0, t1;
  void Start() { t0 = t1 = std::chrono::high_resolution_clock::now(); }
  void Stop() { t1 = std::chrono::high_resolution_clock::now(); }
  double Seconds() const { return std::chrono::duration<double>(t1 - t0).count(); }
};

[[noreturn]] inline void Die(const std::string &msg, int code = -1) {
  fprintf(stderr, "%s\n", msg.c_str());
  fflush(nullptr);
  std::exit(code);
}

struct Cli {
  std::string name;
  std::string kernel;  // pr, pr_spmv, bfs, cc, cc_sv, sssp or bc: the benchmark type of -o 14
  std::string filename;
  int scale = -1, degree = 16;
  bool symmetrize = false, uniform = false, keep_self = false, logging = false;
  std::vector<std::pair<int, std::vector<std::string>>> orders;
  int trials = 16;
  bool trials_explicit = false;
  bool analysis = false, verify = false;
  std::vector<int64_t> sources;
  int repeats = 1;
  std::string weight_scheme = "unit";
  // kernel extras
  int max_iters = 20;      // PR -i (CLPageRank) / BC -i (CLIterApp, default 1)
  double tol = 1e-4;       // PR -t
  bool fixed = false;      // PR -F
  int delta = 1;           // SSSP -d
  int device = 0;

  // extra: the per-kernel getopt letters beyond CLApp's
  static void OnFault(int sig) {
    void *frames[48];
    const int k = backtrace(frames, 48);
    const char msg[] = "driver: fatal signal, host backtrace:\n";
    (void)!write(2, msg, sizeof(msg) - 1);
    backtrace_symbols_fd(frames, k, 2);
    signal(sig, SIG_DFL);
    raise(sig);
  }

  bool Parse(int argc, char **argv, const std::string &extra) {
    setvbuf(stdout, nullptr, _IOLBF, 0);  // every line reaches a pipe as it is printed
    signal(SIGSEGV, OnFault);
    std::string opts = "f:g:hk:su:m:o:zj:SlD:W:an:r:R:v" + extra;
    int c;
    optind = 1;
    while ((c = getopt(argc, argv, opts.c_str())) != -1) {
      switch (c) {
        case 'f': filename = optarg; break;
        case 'g': scale = atoi(optarg); break;
        case 'u': uniform = true; scale = atoi(optarg); break;
        case 'k': degree = atoi(optarg); break;
        case 's': symmetrize = true; break;
        case 'S': keep_self = true; break;
        case 'l': logging = true; break;
        case 'm': break;  // in-place build: the device builder never needs it
        case 'z': break;  // in-degree for degree orderings: see -o 2
        case 'j': case 'D': break;
        case 'W':
          weight_scheme = optarg;
          if (weight_scheme != "unit" && weight_scheme != "hash") Die("Weight scheme must be unit or hash");
          break;
        case 'o': {
          std::string arg(optarg);
          size_t pos = arg.find(':');
          std::vector<std::string> params;
          const int id = std::stoi(arg.substr(0, pos));
          while (pos != std::string::npos) {
            const size_t next = arg.find(':', pos + 1);
            params.push_back(arg.substr(pos + 1, next == std::string::npos ? std::string::npos : next - pos - 1));
            pos = next;
          }
          orders.emplace_back(id, params);
          break;
        }
        case 'a': analysis = true; break;
        case 'v': verify = true; break;
        case 'n':
          trials = atoi(optarg);
          trials_explicit = true;
          if (trials <= 0) Die("Trial count must be positive");
          break;
        case 'r': {
          std::string s(optarg);
          size_t start = 0;
          while (start <= s.size()) {
            const size_t comma = s.find(',', start);
            const std::string tok = s.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
            if (tok.empty()) Die("Source list contains an empty field");
            sources.push_back(std::stoll(tok));
            if (comma == std::string::npos) break;
            start = comma + 1;
          }
          break;
        }
        case 'R':
          repeats = atoi(optarg);
          if (repeats <= 0) Die("Source repeat count must be positive");
          break;
        case 'i': max_iters = atoi(optarg); break;
        case 't': tol = atof(optarg); break;
        case 'F': fixed = true; break;
        case 'd':
          delta = atoi(optarg);
          if (delta <= 0) Die("Delta must be a positive in-range integer");
          break;
        case 'h':
          printf("%s\n -f <file> -g <scale> -u <scale> -k <degree> -s -o <order> -n <trials>"
                 " -r <ids> -R <repeats> -v -a -l%s\n", name.c_str(), extra.c_str());
          std::exit(0);
        default: return false;
      }
    }
    if (filename.empty() && scale == -1) {
      printf("No graph input specified. (Use -h for help)\n");
      return false;
    }
    if (scale != -1) symmetrize = true;  // command_line.h:114-115
    if (!sources.empty() && (sources.size() > 1 || repeats > 1)) {
      const int expected = static_cast<int>(sources.size()) * repeats;
      if (trials_explicit && trials != expected)
        Die("Explicit source list requires -n to equal source_count * source_repeats");
      trials = expected;
    }
    if (const char *d = std::getenv("GBK_DEVICE")) device = atoi(d);
    return true;
  }
};

}  // namespace drv

#include "drv_adaptive.h"

namespace drv {

// Host view of the device graph (downloaded on demand).
struct HostGraph {
  int64_t n = 0, m = 0;
  bool directed = false, weighted = false;
  std::vector<int64_t> off, in_off;
  std::vector<int32_t> nbr, wgt, in_nbr, org_ids;
  const int64_t *io() const { return directed ? in_off.data() : off.data(); }
  const int32_t *in() const { return directed ? in_nbr.data() : nbr.data(); }
};

struct Session {
  Cli cli;
  gbk_ctx *ctx = nullptr;
  gbk_graph *g = nullptr;
  bool weighted = false;
  HostGraph host;
  bool host_valid = false;

  void Check(int rc, const char *what) {
    if (rc != GBK_OK) Die(std::string(what) + ": " + gbk_last_error(ctx), 1);
  }

  const HostGraph &Host() {
    if (host_valid) return host;
    int64_t n, m;
    int d, w;
    gbk_graph_info(g, &n, &m, &d, &w);
    host.n = n;
    host.m = m;
    host.directed = d;
    host.weighted = w;
    host.off.resize(n + 1);
    host.nbr.resize(m);
    host.org_ids.resize(n);
    if (w) host.wgt.resize(m);
    Check(gbk_graph_download(ctx, g, host.off.data(), host.nbr.data(), w ? host.wgt.data() : nullptr,
                             nullptr, nullptr, nullptr, host.org_ids.data()), "download");
    if (d) {
      host.in_off.resize(n + 1);
      Check(gbk_graph_download(ctx, g, nullptr, nullptr, nullptr, host.in_off.data(), nullptr,
                               nullptr, nullptr), "download");
      host.in_nbr.resize(host.in_off[n]);
      Check(gbk_graph_download(ctx, g, nullptr, nullptr, nullptr, nullptr, host.in_nbr.data(),
                               nullptr, nullptr), "download");
    }
    host_valid = true;
    return host;
  }

  // Mapping files of the reference's CPU reorderers: GBK_MAPPING_DIR/
  // <graph>_<tag>.so, <graph> = kron<scale>_ef<degree>, uniform<scale>_ef<degree>
  // or the input file's name without its suffix ("" when absent).
  std::string GraphKey() const {
    if (!cli.filename.empty()) {
      std::string b = cli.filename.substr(cli.filename.rfind('/') + 1);
      return b.substr(0, b.rfind('.'));
    }
    return std::string(cli.uniform ? "uniform" : "kron") + std::to_string(cli.scale) + "_ef" +
           std::to_string(cli.degree);
  }
  std::string MappingFile(const std::string &tag) const {
    const char *dir = std::getenv("GBK_MAPPING_DIR");
    if (!dir || tag.empty()) return "";
    const std::string p = std::string(dir) + "/" + GraphKey() + "_" + tag + ".so";
    return FileExists(p) ? p : "";
  }

  // -o 14: the learned pick or a frozen rule's (see drv_adaptive.h); returns
  // the ordering to apply in its place.
  std::pair<int, std::vector<std::string>> Adaptive(const std::vector<std::string> &opt) {
    const AdaptivePolicy pol = ParseAdaptivePolicy(opt);
    int64_t n = 0, m = 0;
    int d = 0, w = 0;
    gbk_graph_info(g, &n, &m, &d, &w);
    const int bench = BenchCode(cli.kernel);
    printf("=== Full-Graph Adaptive Mode ===\n");
    printf("Nodes: %" PRId64 ", Edges: %" PRId64 "\n", n, m);
    Timer ft;
    ft.Start();
    double t0[7];
    Check(gbk_tier0_features(ctx, g, 0, 0, t0), "tier0 features");
    const int64_t pb = bench ? gbh_modeled_property_bytes(bench, n, m) : 0;
    const int64_t llc = gbh_llc_bytes();
    const double wsr = llc > 0 ? static_cast<double>(pb) / static_cast<double>(llc) : 0.0;
    ft.Stop();
    PrintTime("Degree Variance", t0[1]);
    PrintTime("Hub Concentration", t0[2]);
    PrintTime("Normalized Edge Span", t0[3]);
    PrintTime("Window Neighbor Overlap", t0[4]);
    PrintTime("Packing Factor", t0[5]);
    PrintTime("Property Working Set Bytes", static_cast<double>(pb));
    PrintTime("LLC Capacity Bytes", static_cast<double>(llc));
    PrintTime("Property WSR LLC", wsr);
    // AdaptiveReuseBucket (reorder_types.h:1276-1284)
    const double r = pol.reuse;
    const int bucket = r <= 1 ? 0 : r <= 5 ? 1 : r <= 10 ? 2 : r <= 20 ? 3 : r <= 50 ? 4 : r <= 100 ? 5 : 6;
    printf("Adaptive Tier0 Features: {\"log10_nodes\":%.17g,\"log10_edges\":%.17g,\"avg_degree\":%.17g,"
           "\"degree_cv\":%.17g,\"hub_concentration\":%.17g,\"normalized_edge_span\":%.17g,"
           "\"window_neighbor_overlap\":%.17g,\"property_wsr_llc\":%.17g,\"kernel_class\":%d,"
           "\"reuse_bucket\":%d}\n",
           std::log10(static_cast<double>(n) + 1.0), std::log10(static_cast<double>(m) + 1.0), t0[0], t0[1],
           t0[2], t0[3], t0[4], wsr, bench, bucket);
    PrintTime("Adaptive Feature Time", ft.Seconds());
    printf("Selection: model=%s criterion=%s\n", pol.model.c_str(), pol.criterion.c_str());
    Timer mt;
    mt.Start();
    std::string rule = pol.model;
    int pick = -1;
    if (pol.model == "perceptron") {
      if (pol.criterion == "fastest-reorder") {
        pick = 0;  // the original order needs no mapping and no relabel
      } else {
        // the model is trained for reuse 1 and 2; kernel-only criteria use 2
        const double reuse = pol.criterion == "best-endtoend" ? std::min(2.0, std::max(1.0, pol.reuse)) : 2.0;
        SelectorModel mdl;
        std::vector<double> pred;
        if (mdl.Load(SelectorModelPath()) && LearnedPredict(ctx, g, mdl, cli.kernel, reuse, pred)) {
          int a = 0;
          for (int i = 1; i < static_cast<int>(pred.size()); i++)
            if (pred[i] < pred[a]) a = i;
          pick = pred[a] < 1.0 - mdl.margin ? a : 0;
          std::string line;
          char buf[64];
          for (size_t i = 0; i < pred.size(); i++) {
            snprintf(buf, sizeof(buf), "%s%s=%.4f", i ? " " : "", kOrderNames[i], pred[i]);
            line += buf;
          }
          printf("Adaptive Learned Prediction: %s\n", line.c_str());
        } else {
          PrintLabel("Adaptive Learned Model", "unavailable");
          rule = "allkernel-lowreuse-rule";
        }
      }
    }
    std::pair<int, std::vector<std::string>> out{5, {}};
    if (pick >= 0) {
      out = {pick, {}};
    } else {
      const bool hit = rule == "budgeted-rule" ? gbh_budgeted_rule(bench, pol.reuse, n, t0[0], t0[1], t0[2])
                                               : gbh_frozen_rule(bench, pol.reuse, n, t0[0], t0[1], wsr);
      PrintLabel("Adaptive Rule", std::string(rule) + (hit ? " match" : " fallback"));
      if (hit) {
        const std::string arm = rule == "budgeted-rule" ? kBudgetedArm : kLowReuseArm;
        std::vector<std::string> o;
        std::stringstream ss(arm);
        std::string tok;
        while (std::getline(ss, tok, ':')) o.push_back(tok);
        PrintLabel("Adaptive Rule Arm", "12:" + arm);
        if (!MappingFile(ReordererTag(12, o)).empty()) {
          out = {12, o};
        } else {
          // the GraphBrew reorderer is CPU code: without its mapping file the
          // rule's fallback arm of Rabbit-less builds, DBG, stands in
          PrintLabel("Adaptive Arm Substitute", "5 (no GraphBrew mapping file)");
        }
      }
    }
    mt.Stop();
    PrintTime("Adaptive Model Time", mt.Seconds());
    std::string spec = std::to_string(out.first);
    for (const auto &p : out.second) spec += ":" + p;
    PrintLabel("Adaptive Selected", spec);
    return out;
  }

  int order_pos = 0;  // index of the -o being applied (mapping files fit the first only)

  std::vector<int32_t> Mapping(int algo, const std::vector<std::string> &opt) {
    if (algo == 14) {
      const auto pick = Adaptive(opt);
      return Mapping(pick.first, pick.second);
    }
    if (algo == 11 || algo == 12) {
      // the reference's RCM and GraphBrew reorderers run on the CPU; their
      // mappings enter as files
      const std::string tag = ReordererTag(algo, opt);
      const std::string path = MappingFile(tag);
      if (path.empty() || order_pos != 0)
        Die("Ordering " + std::to_string(algo) + (opt.empty() ? "" : ":" + JoinOpts(opt)) +
            " is produced by the reference reorderer on the CPU; pass its mapping with -o 13:<file>" +
            (tag.empty() ? std::string() : " (or as GBK_MAPPING_DIR/" + GraphKey() + "_" + tag + ".so, first -o)"));
      int64_t n = 0, m = 0, len = 0;
      int d = 0, w = 0;
      gbk_graph_info(g, &n, &m, &d, &w);
      std::vector<int32_t> ids(n);
      Timer t;
      t.Start();
      if (gbh_mapping_read(path.c_str(), &len, nullptr) != 0 || len != n) Die("Mapping file " + path + " does not fit the graph");
      gbh_mapping_read(path.c_str(), &len, ids.data());
      t.Stop();
      PrintLabel("Reorder Mapping File", path.substr(path.rfind('/') + 1));
      PrintTime("Load Map Time", t.Seconds());
      return ids;
    }
    // the orderings that sort by an out-degree key are made on the device
    static const std::unordered_map<int, const char *> device_made = {
        {0, "Original Map Time"}, {1, "Random Map Time"}, {2, "Sort Map Time"}, {5, "DBG Map Time"},
        {6, "HubSortDBG Map Time"}, {7, "HubClusterDBG Map Time"}};
    const auto dm = device_made.find(algo);
    if (dm != device_made.end() && opt.empty()) {
      int64_t n = 0, m = 0;
      int d = 0, w = 0;
      gbk_graph_info(g, &n, &m, &d, &w);
      std::vector<int32_t> ids(n);
      Timer t;
      t.Start();
      double ms = 0;
      Check(gbk_mapping(ctx, g, algo, ids.data(), &ms), "mapping");
      t.Stop();
      if (algo == 1) PrintLabel("Random Seed", "0");
      PrintTime(dm->second, t.Seconds());
      return ids;
    }
    const HostGraph &h = Host();
    std::vector<int32_t> ids(h.n);
    Timer t;
    t.Start();
    switch (algo) {
      case 0:
        gbh_mapping_identity(h.n, ids.data());
        t.Stop();
        PrintTime("Original Map Time", t.Seconds());
        break;
      case 1:
        gbh_mapping_random(h.n, ids.data());
        PrintLabel("Random Seed", "0");
        t.Stop();
        PrintTime("Random Map Time", t.Seconds());
        break;
      case 2:
        gbh_mapping_sort(h.n, h.off.data(), ids.data());
        t.Stop();
        PrintTime("Sort Map Time", t.Seconds());
        break;
      case 3:
        gbh_mapping_hubsort(h.n, h.off.data(), ids.data());
        t.Stop();
        PrintTime("HubSort Map Time", t.Seconds());
        break;
      case 4:
        gbh_mapping_hubcluster(h.n, h.off.data(), ids.data());
        t.Stop();
        PrintTime("HubCluster Map Time", t.Seconds());
        break;
      case 5:
        gbh_mapping_dbg(h.n, h.off.data(), ids.data());
        t.Stop();
        PrintTime("DBG Map Time", t.Seconds());
        break;
      case 6:
        gbh_mapping_hubsortdbg(h.n, h.off.data(), ids.data());
        t.Stop();
        PrintTime("HubSortDBG Map Time", t.Seconds());
        break;
      case 7:
        gbh_mapping_hubclusterdbg(h.n, h.off.data(), ids.data());
        t.Stop();
        PrintTime("HubClusterDBG Map Time", t.Seconds());
        break;
      case 13: {
        if (opt.empty()) Die("MAP ordering needs a file: -o 13:<mapping.lo|.so>");
        int64_t len = 0;
        if (gbh_mapping_read(opt[0].c_str(), &len, nullptr) != 0) Die("Cannot read mapping " + opt[0]);
        if (len != h.n) Die("Mapping size does not match the graph");
        gbh_mapping_read(opt[0].c_str(), &len, ids.data());
        t.Stop();
        PrintTime("Load Map Time", t.Seconds());
        break;
      }
      default:
        Die("Ordering " + std::to_string(algo) +
            " is produced by the reference reorderer on the CPU; pass its mapping with -o 13:<file>");
    }
    return ids;
  }

  void MakeGraph(bool need_weights) {
    weighted = need_weights;
    if (gbk_ctx_create(cli.device, &ctx) != GBK_OK) Die("No usable GPU (gbk_ctx_create failed)", 1);
    Timer total, rep;
    total.Start();
    rep.Start();
    std::vector<int32_t> u, v, w;
    Timer t;
    const std::string &f = cli.filename;
    const auto has_suffix = [&](const char *suf) {
      const size_t k = std::strlen(suf);
      return f.size() > k && f.compare(f.size() - k, k, suf) == 0;
    };
    if (has_suffix(".sg") || has_suffix(".wsg")) {
      LoadSerialized(need_weights);
      rep.Stop();
      PrintTime("Representation Build Time", rep.Seconds());
      PrintLabel("Topology Analysis", "disabled");
      ApplyOrders(total);
      return;
    }
    if (!cli.filename.empty()) {
      t.Start();
      const bool wel = has_suffix(".wel");
      const bool el = has_suffix(".el");
      if (!wel && !el) Die("Unrecognized suffix: only .el, .wel, .sg and .wsg are read by this driver", -3);
      int64_t m = 0;
      if (gbh_read_el(f.c_str(), wel, &m, nullptr, nullptr, nullptr) != 0) Die("Couldn't open file " + f, -2);
      u.resize(m);
      v.resize(m);
      if (wel) w.resize(m);
      gbh_read_el(f.c_str(), wel, &m, u.data(), v.data(), wel ? w.data() : nullptr);
      t.Stop();
      PrintTime("Read Time", t.Seconds());
      if (need_weights && !wel) {  // .el into a weighted graph: U[1,255] weights
        w.resize(m);
        gbh_insert_weights(m, w.data());
      }
    } else {
      t.Start();
      const int64_t m = (int64_t(1) << cli.scale) * cli.degree;
      u.resize(m);
      v.resize(m);
      if (gbh_generate_el(cli.scale, cli.degree, cli.uniform, u.data(), v.data()) != 0) Die("bad scale", -31);
      t.Stop();
      PrintTime("Generate Time", t.Seconds());
      if (need_weights) {
        w.resize(m);
        gbh_insert_weights(m, w.data());
      }
    }
    if (!need_weights) w.clear();
    t.Start();
    double dev_ms = 0;
    Check(gbk_graph_build(ctx, -1, static_cast<int64_t>(u.size()), u.data(), v.data(),
                          w.empty() ? nullptr : w.data(), cli.symmetrize, cli.keep_self, &g, &dev_ms),
          "build");
    t.Stop();
    PrintTime("Build Time", t.Seconds());
    rep.Stop();
    PrintTime("Representation Build Time", rep.Seconds());
    PrintLabel("Topology Analysis", "disabled");
    ApplyOrders(total);
  }

  // .sg / .wsg (ReadSerializedGraph, gapbs/reader.h:445-578; symmetrize as
  // MakeGraph does for serialized input, gapbs/builder.h:741-790): the CSR
  // is uploaded as stored; -s on a directed unweighted file with identity
  // original ids rebuilds it symmetrized from its edges.
  void LoadSerialized(bool need_weights) {
    Timer t;
    t.Start();
    const std::string &f = cli.filename;
    int directed = 0, file_weighted = 0;
    int64_t n = 0, m = 0;
    if (gbh_sg_info(f.c_str(), &directed, &n, &m, &file_weighted) != 0) Die("Couldn't open file " + f, -6);
    std::vector<int64_t> off(n + 1), in_off(directed ? n + 1 : 0);
    std::vector<int32_t> nbr(m), in_nbr(directed ? m : 0), org(n);
    std::vector<int32_t> w(need_weights ? m : 0), in_w(need_weights && directed ? m : 0);
    const int scheme = cli.weight_scheme == "hash" ? 1 : 0;
    if (gbh_sg_read(f.c_str(), scheme, off.data(), nbr.data(), need_weights ? w.data() : nullptr,
                    directed ? in_off.data() : nullptr, directed ? in_nbr.data() : nullptr,
                    need_weights && directed ? in_w.data() : nullptr, org.data()) != 0)
      Die("Serialized graph " + f + " is truncated or inconsistent", -6);
    t.Stop();
    PrintTime("Read Time", t.Seconds());
    t.Start();
    if (cli.symmetrize && directed) {
      if (need_weights) Die("Symmetrizing weighted serialized graphs is unsupported", 1);
      for (int64_t i = 0; i < n; i++)
        if (org[i] != i) Die("Cannot symmetrize a serialized graph whose original IDs are not the identity", 1);
      std::vector<int32_t> u(m);
      for (int64_t r = 0; r < n; r++)
        for (int64_t e = off[r]; e < off[r + 1]; e++) u[e] = static_cast<int32_t>(r);
      double dev_ms = 0;
      Check(gbk_graph_build(ctx, n, m, u.data(), nbr.data(), nullptr, 1, cli.keep_self, &g, &dev_ms), "build");
    } else {
      Check(gbk_graph_upload(ctx, n, directed, off.data(), nbr.data(), need_weights ? w.data() : nullptr,
                             directed ? in_off.data() : nullptr, directed ? in_nbr.data() : nullptr,
                             need_weights && directed ? in_w.data() : nullptr, org.data(), &g),
            "upload");
    }
    t.Stop();
    PrintTime("Build Time", t.Seconds());
  }

  // The orderings of -o in sequence, then the total preprocessing time.
  void ApplyOrders(Timer &total) {
    for (const auto &o : cli.orders) {
      std::vector<int32_t> ids = Mapping(o.first, o.second);
      order_pos++;
      Timer val;
      val.Start();
      const int bad = gbh_validate_permutation(static_cast<int64_t>(ids.size()), ids.data());
      val.Stop();
      if (bad) Die(bad == -1 ? "generated mapping contains an out-of-range ID"
                             : "generated mapping contains a duplicate ID", 1);
      char fp[32];
      snprintf(fp, sizeof(fp), "%016" PRIx64, gbh_mapping_fingerprint(static_cast<int64_t>(ids.size()), ids.data()));
      PrintLabel("Mapping Fingerprint", fp);
      Timer apply;
      apply.Start();
      gbk_graph *r = nullptr;
      gbk_relabel_stats rs{};
      Check(gbk_relabel(ctx, g, ids.data(), &r, &rs), "relabel");
      apply.Stop();
      PrintTime("Relabel Map Time", rs.device_ms / 1e3);
      gbk_graph_free(ctx, g);
      g = r;
      host_valid = false;
      PrintTime("Reorder Validation Time", val.Seconds());
      PrintTime("Reorder Apply Time", apply.Seconds());
      std::string spec = std::to_string(o.first);
      for (const auto &p : o.second) spec += ":" + (o.first == 13 ? std::string("<mapping>") : p);
      PrintLabel("Resolved Reorder Spec", spec);
    }
    if (!cli.orders.empty()) {
      const HostGraph &h = Host();
      std::vector<int32_t> src_to_final(h.n);
      for (int64_t i = 0; i < h.n; i++) src_to_final[h.org_ids[i]] = static_cast<int32_t>(i);
      char fp[32];
      snprintf(fp, sizeof(fp), "%016" PRIx64, gbh_mapping_fingerprint(h.n, src_to_final.data()));
      PrintLabel("Composed Mapping Fingerprint", fp);
    }
    total.Stop();
    PrintTime("Total Preprocessing Time", total.Seconds());
  }

  void PrintStats() {
    int64_t n, m;
    int d, w;
    gbk_graph_info(g, &n, &m, &d, &w);
    const int64_t edges = d ? m : m / 2;
    std::printf("Graph has %" PRId64 " nodes and %" PRId64 " %sdirected edges for degree: %" PRId64
                " Estimated size: %" PRId64 " MB\n",
                n, edges, d ? "" : "un", n ? edges / n : 0,
                static_cast<int64_t>((n * 4 + (n + 1) * 4 + edges * 4) / (1024 * 1024)));
  }

  // SourcePicker over original ids (gapbs/benchmark.h:92-156).
  std::vector<std::pair<int32_t, int32_t>> Sources(int count) {
    const HostGraph &h = Host();
    std::vector<std::pair<int32_t, int32_t>> out;
    std::vector<int32_t> internal(h.n);
    for (int64_t i = 0; i < h.n; i++) internal[h.org_ids[i]] = static_cast<int32_t>(i);
    if (!cli.sources.empty()) {
      auto resolve = [&](int64_t o) {
        if (o < 0 || o >= h.n) Die("Source original ID is outside the graph");
        const int32_t x = internal[o];
        if (h.off[x + 1] == h.off[x]) Die("Explicit source resolves to an isolated vertex");
        return std::make_pair(static_cast<int32_t>(o), x);
      };
      if (cli.sources.size() == 1) {
        for (int i = 0; i < count; i++) out.push_back(resolve(cli.sources[0]));
      } else {
        for (int64_t s : cli.sources)
          for (int r = 0; r < cli.repeats; r++) out.push_back(resolve(s));
      }
      return out;
    }
    std::vector<int32_t> o(count), in(count);
    if (gbh_source_picker(h.n, h.off.data(), h.org_ids.data(), count, cli.repeats, o.data(), in.data()) != 0)
      Die("Unable to select a non-isolated source vertex");
    for (int i = 0; i < count; i++) out.emplace_back(o[i], in[i]);
    return out;
  }

  void PrintSource(const std::pair<int32_t, int32_t> &s) {
    const HostGraph &h = Host();
    PrintLabel("Source Original", std::to_string(s.first));
    PrintLabel("Source Internal", std::to_string(s.second));
    PrintLabel("Source Out Degree", std::to_string(h.off[s.second + 1] - h.off[s.second]));
  }

  // BenchmarkKernel (gapbs/benchmark.h:435-545).  kernel(trial) runs one
  // trial and returns its device ms; extract/analyse/verify act on its result.
  void Benchmark(const std::string &bench, const std::function<double(int)> &kernel,
                 const std::function<void(int)> &extract, const std::function<void()> &analyse,
                 const std::function<bool(int)> &verify) {
    PrintStats();
    double total = 0;
    for (int it = 0; it < cli.trials; it++) {
      Timer t;
      t.Start();
      const double dev_ms = kernel(it);
      t.Stop();
      PrintTime("Trial Time", t.Seconds());
      PrintTime("Device Time", dev_ms / 1e3);
      total += t.Seconds();
      extract(it);
      if (cli.analysis && it == cli.trials - 1) analyse();
      if (cli.verify) {
        Timer vt;
        vt.Start();
        const bool ok = verify(it);
        PrintLabel("Verification", ok ? "PASS" : "FAIL");
        vt.Stop();
        PrintTime("Verification Time", vt.Seconds());
        if (!ok) {
          PrintLabel("Verification Failure", bench);
          fflush(nullptr);
          std::exit(3);
        }
      }
    }
    PrintTime("Average Time", total / cli.trials);
  }

  ~Session() {
    if (g) gbk_graph_free(ctx, g);
    if (ctx) gbk_ctx_destroy(ctx);
  }
};

// TopK (gapbs/benchmark.h:333-349) on (key, value) pairs, largest values first.
template <typename K, typename V>
std::vector<std::pair<V, K>> TopK(const std::vector<std::pair<K, V>> &kv, size_t k) {
  std::vector<std::pair<V, K>> top;
  for (const auto &p : kv) top.emplace_back(p.second, p.first);
  std::stable_sort(top.begin(), top.end(), std::greater<std::pair<V, K>>());
  if (top.size() > k) top.resize(k);
  return top;
}

inline void PrintCompStats(const std::vector<int32_t> &comp) {
  printf("\n");
  std::unordered_map<int32_t, int32_t> count;
  for (int32_t c : comp) count[c] += 1;
  std::vector<std::pair<int32_t, int32_t>> kv(count.begin(), count.end());
  auto top = TopK(kv, 5);
  printf("%d biggest clusters\n", static_cast<int>(top.size()));
  for (const auto &p : top) printf("%d:%d\n", p.second, p.first);
  printf("There are %zu components\n", count.size());
}

}  // namespace drv
