// cpim_solve: command-line driver of the device engine through the C++
// facade (include/cpim/gmodel.hpp), the counterpart of the reference's
// test/gmodel_solver.cpp and cpim_test_parser output
// (samples/main_new_parser.cpp:239-243).  It builds a seeded random binary
// network, runs the root GAC fixpoint and the device MAC, and prints the
// reference's "MAC stats:" line.
//
//   cpim_solve --n 30 --d 16 --e 120 --p 0.4 --seed 1 [--node_limit K]
//              [--time_limit MS] [--varh dom|domdeg|domwdeg] [--sac] [--dway]
//              [--validate-only]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cpim/gmodel.hpp"

extern "C" {
int cpimw_random_pairs(uint64_t seed, int n, int e, int32_t* scopes);
int cpimw_random_relations(uint64_t seed, int e, int d0, int d1, int words, double p, uint64_t* rel);
}

int main(int argc, char** argv) {
  int n = 30, d = 16, e = 120;
  double p = 0.4;
  unsigned long long seed = 1;
  long long node_limit = 0;
  int time_limit = 0;
  bool sac = false, dway = false, validate_only = false;
  int varh = CPIM_VARH_DOM_MIN;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() { return i + 1 < argc ? argv[++i] : (char*)"0"; };
    if (a == "--n") n = std::atoi(next());
    else if (a == "--d") d = std::atoi(next());
    else if (a == "--e") e = std::atoi(next());
    else if (a == "--p") p = std::atof(next());
    else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
    else if (a == "--node_limit") node_limit = std::atoll(next());
    else if (a == "--time_limit") time_limit = std::atoi(next());
    else if (a == "--varh") {
      const std::string h = next();
      if (h == "dom") varh = CPIM_VARH_DOM_MIN;
      else if (h == "domdeg") varh = CPIM_VARH_DOM_DEG;
      else if (h == "domwdeg") varh = CPIM_VARH_DOM_WDEG;
      else {
        std::fprintf(stderr, "--varh takes dom, domdeg or domwdeg\n");
        return 2;
      }
    } else if (a == "--sac") sac = true;
    else if (a == "--dway") dway = true;
    else if (a == "--validate-only") validate_only = true;
    else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  const int W = (d + 63) / 64;
  std::vector<int32_t> dom(n, d), arity(e, 2), scope(2 * e);
  std::vector<uint64_t> rel((size_t)e * d * W);
  if (cpimw_random_pairs(seed, n, e, scope.data()) != 0 ||
      cpimw_random_relations(seed, e, d, d, W, p, rel.data()) != 0) {
    std::fprintf(stderr, "bad instance parameters\n");
    return 2;
  }
  cpim_model_desc desc{};
  desc.num_vars = n;
  desc.dom_size = dom.data();
  desc.num_cons = e;
  desc.con_arity = arity.data();
  desc.con_scope = scope.data();
  desc.bin_rel = rel.data();
  desc.rel_words = W;
  if (validate_only) {
    // exercise the facade's error mapping without a GPU: a scope naming an
    // unknown variable must raise std::invalid_argument
    scope[0] = n + 7;
    try {
      cpim::model::GModelAdapter::Build(desc);
      std::printf("validate: no exception\n");
      return 1;
    } catch (const std::invalid_argument& ex) {
      std::printf("validate: invalid_argument: %s\n", ex.what());
      return 0;
    }
  }
  cpim::GModel model = cpim::model::GModelAdapter::Build(desc);
  const cpim::GacStats g = model.EnforceGAC_Auto();
  std::printf("root GAC: iterations=%d deletions=%d inconsistent=%s\n", g.iterations, g.deletions,
              g.inconsistent ? "true" : "false");
  if (sac) {
    const cpim_sac_stats s = model.EnforceSAC(true);
    std::printf("SAC: consistent=%s passes=%d tests=%lld removed=%lld\n", s.consistent ? "true" : "false",
                s.passes, (long long)s.singleton_tests, (long long)s.removed);
    return 0;
  }
  cpim::GModelSolver solver(&model, false,
                            dway ? cpim::GModelSolver::Branching::kDway : cpim::GModelSolver::Branching::kBinary);
  if (varh != CPIM_VARH_DOM_MIN) {
    try {
      solver.SetVarHeuristic(varh);
    } catch (const std::invalid_argument& ex) {
      std::fprintf(stderr, "%s\n", ex.what());
      return 2;
    }
  }
  cpim::GpuSearchStatistics st;
  if (node_limit > 0 && !dway) {
    cpim_solve_opts o{};
    o.node_limit = node_limit;
    o.time_limit_ms = time_limit;
    o.var_heuristic = varh;
    cpim_search_stats cs{};
    std::vector<int32_t> sol(n, -1);
    int32_t tl = 0;
    cpim::check(cpim_gmodel_solve(model.handle(), &o, &cs, sol.data(), nullptr, &tl));
    st.num_positive = (int)cs.num_positive;
    st.num_negative = (int)cs.num_negative;
    st.num_solutions = (int)cs.num_solutions;
    st.time_out = cs.time_out != 0;
    st.solve_time = cs.solve_time;
    if (cs.num_solutions) {
      std::printf("MAC solution (canonical indices):");
      for (int v = 0; v < n; ++v) std::printf(" %d", sol[v]);
      std::printf("\n");
    }
  } else {
    st = solver.Solve(time_limit);
    const auto s = solver.GetSolution();
    if (!s.empty()) {
      std::printf("MAC solution (canonical indices):");
      for (int v : s) std::printf(" %d", v);
      std::printf("\n");
    }
  }
  std::printf("MAC stats: time=%d ms, positives=%d, negatives=%d, nodes=%d, timeout=%s\n",
              (int)(st.solve_time * 1000.0), st.num_positive, st.num_negative,
              st.num_positive + st.num_negative, st.time_out ? "true" : "false");
  return 0;
}
