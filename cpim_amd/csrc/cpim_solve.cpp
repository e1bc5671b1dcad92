// cpim_solve: command-line driver of the device engine through the C++
// facade (include/cpim/gmodel.hpp), the counterpart of the reference's
// test/gmodel_solver.cpp and cpim_test_parser output
// (samples/main_new_parser.cpp:239-243).  It builds a seeded random binary
// network, runs the root GAC fixpoint and the device MAC, and prints the
// reference's "MAC stats:" line.
//
//   cpim_solve --n 30 --d 16 --e 120 --p 0.4 --seed 1 [--node_limit K]
//              [--time_limit MS] [--varh dom|domdeg|domwdeg] [--sac] [--dway]
//              [--find-all K] [--cmodel] [--validate-only]
//
// --find-all K: GModelSolver::SetFindAllSolutions(true) + SetMaxSolutions(K);
//   prints every solution kept and the count.
// --cmodel: the legacy CModel entry points (include/cuSAC.cuh:422-447) in
//   turn, printing the domain hash of the level each one leaves: enforceGAC(),
//   CreateNewLevel() + an assignment + enforceGAC(var, 1) + BackLevel(),
//   enforceSAC(), and solve() on a second CModel of the same network.
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

namespace {

// Order-free hash of the domains of one level, the definition of
// cpim_node_record::dom_hash.
uint64_t dom_hash(const std::vector<uint64_t>& dom) {
  uint64_t h = 0;
  for (size_t i = 0; i < dom.size(); ++i) {
    uint64_t z = dom[i] + (uint64_t)i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    h += z ^ (z >> 31);
  }
  return h;
}

int run_cmodel(const cpim_model_desc& desc, int time_limit) {
  cpim::CModel cm(desc);
  const bool gac = cm.enforceGAC();
  std::printf("CModel enforceGAC: consistent=%s hash=%llu\n", gac ? "true" : "false",
              (unsigned long long)dom_hash(cm.gmodel().ReadDomains(0)));
  if (gac) {
    cpim::GModel& g = cm.gmodel();
    const int level = cm.CreateNewLevel();
    int var = 0;
    while (var < g.num_vars() && g.GetDomainSize(var, level) < 2) ++var;
    if (var < g.num_vars()) {
      const int a = g.GetFirstValue(var, level);
      g.AssignValue(var, a, level);
      const bool ok = cm.enforceGAC(var, 1);
      std::printf("CModel decision: level=%d var=%d value=%d consistent=%s hash=%llu\n", level, var, a,
                  ok ? "true" : "false", (unsigned long long)(ok ? dom_hash(g.ReadDomains(level)) : 0ull));
    }
    std::printf("CModel BackLevel: %d\n", cm.BackLevel());
    cm.enforceSAC();
    std::printf("CModel enforceSAC: consistent=%s hash=%llu\n", cm.sac_consistent() ? "true" : "false",
                (unsigned long long)(cm.sac_consistent() ? dom_hash(g.ReadDomains(0)) : 0ull));
  }
  cpim::CModel fresh(desc);
  const cpim::SearchStatistics st = fresh.solve((float)time_limit);
  if (!fresh.solution().empty()) {
    std::printf("CModel solution:");
    for (int v : fresh.solution()) std::printf(" %d", v);
    std::printf("\n");
  }
  std::printf("CModel solve: num_sol=%llu positives=%llu negatives=%llu timeout=%s\n",
              (unsigned long long)st.num_sol, (unsigned long long)st.num_positive,
              (unsigned long long)st.num_negative, st.time_out ? "true" : "false");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  int n = 30, d = 16, e = 120;
  double p = 0.4;
  unsigned long long seed = 1;
  long long node_limit = 0;
  int time_limit = 0;
  bool sac = false, dway = false, validate_only = false, cmodel = false;
  long long find_all = 0;
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
    else if (a == "--cmodel") cmodel = true;
    else if (a == "--find-all") find_all = std::atoll(next());
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
  if (cmodel) return run_cmodel(desc, time_limit);
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
  if (find_all > 0) {
    solver.SetFindAllSolutions(true);
    solver.SetMaxSolutions((int)find_all);
    st = solver.Solve(time_limit);
    for (const auto& sv : solver.GetAllSolutions()) {
      std::printf("solution:");
      for (int v : sv) std::printf(" %d", v);
      std::printf("\n");
    }
    std::printf("solutions found: %d kept: %zu\n", st.num_solutions, solver.GetAllSolutions().size());
  } else if (node_limit > 0 && !dway) {
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
