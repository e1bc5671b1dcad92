// C++17 facade over include/cpim_gpu.h with the shape of the reference's GPU
// entry points:
//
//   cpim::GacStats, cpim::GModel            include/GModel.cuh:17-21, 78-283
//   cpim::model::GModelOptions,
//   cpim::model::GModelAdapter::Build        include/model/gmodel_adapter.h:19-43
//   cpim::GpuSearchStatistics, GModelSolver  include/GModelSolver.h:11-70
//   cpim::CModel (legacy entry points)       include/cuSAC.cuh:422-447
//
// Differences from the reference, on purpose:
//   * Build takes a cpim_model_desc (plain arrays of the normalized model)
//     instead of an IntermediateModel: the format layer stays with the caller.
//   * Device arrays are not public members; ReadDomains() copies a level to
//     the host (the reference exposed cudaMallocManaged pointers).
//   * GModelSolver runs the CPU MAC's binary branching on the device by
//     default; Branching::kDway replays the reference's d-way search through
//     the level API.
// Errors: CPIM_ERR_INVALID_ARGUMENT -> std::invalid_argument, every other
// code -> std::runtime_error, with the library's message.  EnforceGAC never
// throws on inconsistency (it reports it in GacStats), like the reference.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../cpim_gpu.h"

namespace cpim {

inline void check(int code) {
  if (code == CPIM_OK) return;
  const char* msg = cpim_last_error();
  const std::string text = msg ? msg : "cpim error";
  if (code == CPIM_ERR_INVALID_ARGUMENT) throw std::invalid_argument(text);
  throw std::runtime_error(text);
}

struct GacStats {
  int iterations = 0;
  int deletions = 0;
  bool inconsistent = false;
};

struct GpuSearchStatistics {
  int num_positive = 0;
  int num_negative = 0;
  int num_solutions = 0;
  int gac_iterations = 0;
  int gac_deletions = 0;
  bool time_out = false;
  bool unsolvable = false;
  double solve_time = 0.0;  // seconds
  double gac_time = 0.0;    // seconds
};

class GModel {
 public:
  GModel(GModel&& o) noexcept : h_(o.h_), info_(o.info_) { o.h_ = nullptr; }
  GModel& operator=(GModel&& o) noexcept {
    if (this != &o) {
      reset();
      h_ = o.h_;
      info_ = o.info_;
      o.h_ = nullptr;
    }
    return *this;
  }
  GModel(const GModel&) = delete;
  GModel& operator=(const GModel&) = delete;
  ~GModel() { reset(); }

  int num_vars() const { return info_.num_vars; }
  int num_constraints() const { return info_.num_constraints; }
  int max_dom_size() const { return info_.max_dom_size; }
  int bit_dom_int_size() const { return info_.bit_dom_int_size; }
  int max_depth() const { return info_.max_depth; }
  int dom_words() const { return info_.dom_words; }
  cpim_gmodel* handle() const { return h_; }

  GacStats EnforceGAC(bool verbose = true, int assigned_var = -1) {
    return run(assigned_var, CPIM_GAC_PLAIN, verbose, "[GAC-Bitmap]");
  }
  // All rounds in one launch of co-resident workgroups, a grid barrier
  // between rounds.
  GacStats EnforceGAC_Persistent(bool verbose = true, int assigned_var = -1) {
    return run(assigned_var, CPIM_GAC_PERSISTENT, verbose, "[GAC-Persistent]");
  }
  // Deprecated in the reference (include/GModel.cuh:199-201): the event-list
  // fixpoint from every constraint, one launch per round; served by the
  // host-driven rounds, whose closure is the same.
  [[deprecated("use EnforceGAC")]] GacStats EnforceGAC_Legacy(bool verbose = true) {
    return run(-1, CPIM_GAC_PLAIN, verbose, "[GAC-Legacy]");
  }
  // Extensions: one-workgroup LDS fixpoint, and the autotuned choice.
  GacStats EnforceGAC_Block(int assigned_var = -1) { return run(assigned_var, CPIM_GAC_BLOCK, false, ""); }
  GacStats EnforceGAC_Auto(int assigned_var = -1) { return run(assigned_var, CPIM_GAC_AUTO, false, ""); }

  int CreateNewLevel() {
    int32_t l = 0;
    check(cpim_gmodel_create_new_level(h_, &l));
    return l;
  }
  void BackToLevel(int level) { check(cpim_gmodel_back_to_level(h_, level)); }
  int GetCurrentLevel() const {
    int32_t l = 0;
    check(cpim_gmodel_current_level(h_, &l));
    return l;
  }
  bool AssignValue(int var, int value, int level) { return flag(cpim_gmodel_assign_value, var, value, level); }
  bool RemoveValue(int var, int value, int level) { return flag(cpim_gmodel_remove_value, var, value, level); }
  int GetDomainSize(int var, int level) const {
    int32_t s = 0;
    check(cpim_gmodel_domain_size(h_, var, level, &s));
    return s;
  }
  int GetMinDomainVar(int level) const {
    int32_t v = 0;
    check(cpim_gmodel_min_domain_var(h_, level, &v));
    return v;
  }
  bool IsAssigned(int var, int level) const {
    int32_t r = 0;
    check(cpim_gmodel_is_assigned(h_, var, level, &r));
    return r != 0;
  }
  bool IsFullyAssigned(int level) const {
    int32_t r = 0;
    check(cpim_gmodel_is_fully_assigned(h_, level, &r));
    return r != 0;
  }
  int GetAssignedValue(int var, int level) const {
    int32_t v = 0;
    check(cpim_gmodel_assigned_value(h_, var, level, &v));
    return v;
  }
  int GetFirstValue(int var, int level) const {
    int32_t v = 0;
    check(cpim_gmodel_first_value(h_, var, level, &v));
    return v;
  }
  int GetNextValue(int var, int value, int level) const {
    int32_t v = 0;
    check(cpim_gmodel_next_value(h_, var, value, level, &v));
    return v;
  }
  // Host copy of one level, [num_vars][dom_words] 64-bit words.
  std::vector<uint64_t> ReadDomains(int level) const {
    std::vector<uint64_t> out((size_t)info_.num_vars * info_.dom_words);
    check(cpim_gmodel_read_domains(h_, level, out.data()));
    return out;
  }
  // dom/wdeg constraint weights, one per propagated constraint; they start at
  // 1 with the model and persist across solves (Tabular::weight); reset sets
  // them back to 1 after the copy.
  std::vector<float> Weights(bool reset = false) {
    int64_t k = 0;
    check(cpim_gmodel_weights(h_, nullptr, 0, &k, 0));
    std::vector<float> w((size_t)k);
    check(cpim_gmodel_weights(h_, w.data(), k, &k, reset ? 1 : 0));
    return w;
  }
  cpim_sac_stats EnforceSAC(bool batched = true) {
    cpim_sac_stats st{};
    check(cpim_gmodel_enforce_sac(h_, batched ? CPIM_SAC_BATCHED : CPIM_SAC_PLAIN, &st));
    return st;
  }

 private:
  friend class model_adapter_access;
  explicit GModel(cpim_gmodel* h) : h_(h) { check(cpim_gmodel_get_info(h_, &info_)); }
  void reset() {
    if (h_) cpim_gmodel_destroy(h_);
    h_ = nullptr;
  }
  GacStats run(int var, int mode, bool verbose, const char* tag) {
    cpim_gac_stats st{};
    check(cpim_gmodel_enforce_gac(h_, var, mode, &st));
    GacStats g;
    g.iterations = st.iterations;
    g.deletions = st.deletions;
    g.inconsistent = st.inconsistent != 0;
    if (verbose)
      std::printf("%s iterations=%d deletions=%d inconsistent=%s\n", tag, g.iterations, g.deletions,
                  g.inconsistent ? "true" : "false");
    return g;
  }
  template <typename F>
  bool flag(F f, int a, int b, int c) {
    int32_t r = 0;
    check(f(h_, a, b, c, &r));
    return r != 0;
  }
  cpim_gmodel* h_ = nullptr;
  cpim_gmodel_info info_{};
};

class model_adapter_access {
 public:
  static GModel make(cpim_gmodel* h) { return GModel(h); }
};

namespace model {

struct GModelOptions {
  int device_id = 0;
  bool enable_prefetch = false;
  bool skip_non_binary = true;
  bool skip_conflicts = false;
};

class GModelAdapter {
 public:
  static GModel Build(const cpim_model_desc& desc, const GModelOptions& options = {}) {
    cpim_gmodel_options o;
    cpim_gmodel_options_default(&o);
    o.device_id = options.device_id;
    o.enable_prefetch = options.enable_prefetch;
    o.skip_non_binary = options.skip_non_binary;
    o.skip_conflicts = options.skip_conflicts;
    cpim_gmodel* h = nullptr;
    check(cpim_gmodel_create(&desc, &o, &h));
    return model_adapter_access::make(h);
  }
};

}  // namespace model

class GModelSolver {
 public:
  enum class Branching { kBinary, kDway };

  explicit GModelSolver(GModel* model, bool verbose = false, Branching b = Branching::kBinary)
      : model_(model), verbose_(verbose), branching_(b) {
    if (!model_) throw std::runtime_error("[GModelSolver] model is null!");
  }

  // CPIM_VARH_DOM_MIN (default), CPIM_VARH_DOM_DEG or CPIM_VARH_DOM_WDEG;
  // the d-way search orders by smallest domain only.
  void SetVarHeuristic(int h) {
    if (h != CPIM_VARH_DOM_MIN && h != CPIM_VARH_DOM_DEG && h != CPIM_VARH_DOM_WDEG)
      throw std::invalid_argument("unknown variable heuristic");
    if (h != CPIM_VARH_DOM_MIN && branching_ == Branching::kDway)
      throw std::invalid_argument("the d-way search orders variables by smallest domain only");
    var_heuristic_ = h;
  }

  // With SetFindAllSolutions(true) the binary search enumerates, up to
  // SetMaxSolutions (1,000,000 by default then), as the reference's search
  // stops (src/GModelSolver.cu:125-128): each solution is counted, then
  // treated as a failed node.  GetAllSolutions() returns the first ones, as
  // many as fit in kSolutionBytes of device memory; num_solutions counts
  // them all.
  static constexpr long long kSolutionBytes = 256ll << 20;

  GpuSearchStatistics Solve(int time_limit = 0) {
    solutions_.clear();
    if (branching_ == Branching::kDway) return SolveDway(time_limit);
    cpim_solve_opts o{};
    o.time_limit_ms = time_limit;
    o.branching = CPIM_BRANCH_BINARY;
    o.var_heuristic = var_heuristic_;
    const int n = model_->num_vars();
    const long long k = find_all_ ? std::max(1, max_solutions_) : 1;
    const long long cap = std::max(1ll, std::min(k, kSolutionBytes / (4ll * std::max(1, n))));
    o.max_solutions = k;
    o.solution_capacity = cap;
    cpim_search_stats st{};
    std::vector<int32_t> sol((size_t)(n * cap), -1);
    int32_t tl = 0;
    check(cpim_gmodel_solve(model_->handle(), &o, &st, sol.data(), nullptr, &tl));
    GpuSearchStatistics g;
    g.num_positive = (int)st.num_positive;
    g.num_negative = (int)st.num_negative;
    g.num_solutions = (int)st.num_solutions;
    g.gac_iterations = (int)st.gac_iterations;
    g.time_out = st.time_out != 0;
    g.unsolvable = st.unsolvable != 0;
    g.solve_time = st.solve_time;
    g.gac_time = st.gac_time;
    const long long kept = std::min<long long>(st.num_solutions, cap);
    for (long long i = 0; i < kept; ++i) solutions_.emplace_back(sol.begin() + i * n, sol.begin() + (i + 1) * n);
    return g;
  }

  std::vector<int> GetSolution() const { return solutions_.empty() ? std::vector<int>{} : solutions_[0]; }
  std::vector<std::vector<int>> GetAllSolutions() const { return solutions_; }
  void SetFindAllSolutions(bool find_all) {
    find_all_ = find_all;
    if (find_all) max_solutions_ = 1000000;
  }
  void SetMaxSolutions(int m) { max_solutions_ = m; }

 private:
  using clock = std::chrono::steady_clock;
  // d-way search of src/GModelSolver.cu:31-189 over the level API.
  GpuSearchStatistics SolveDway(int time_limit) {
    GpuSearchStatistics st;
    const auto t0 = clock::now();
    GacStats g = model_->EnforceGAC(verbose_);
    st.gac_iterations += g.iterations;
    st.gac_deletions += g.deletions;
    if (g.inconsistent) {
      st.unsolvable = true;
    } else if (model_->IsFullyAssigned(0)) {
      Extract(0);
    } else {
      Search(0, st, time_limit, t0);
    }
    st.num_solutions = (int)solutions_.size();
    st.solve_time = std::chrono::duration<double>(clock::now() - t0).count();
    return st;
  }
  bool Search(int level, GpuSearchStatistics& st, int time_limit, clock::time_point t0) {
    if (time_limit > 0 &&
        std::chrono::duration<double, std::milli>(clock::now() - t0).count() >= time_limit) {
      st.time_out = true;
      return true;
    }
    const int var = model_->GetMinDomainVar(level);
    if (var == -1) {
      if (model_->IsFullyAssigned(level)) {
        Extract(level);
        return !find_all_ || (int)solutions_.size() >= max_solutions_;
      }
      return false;
    }
    for (int value = model_->GetFirstValue(var, level); value != -1;
         value = model_->GetNextValue(var, value, level)) {
      const int nl = model_->CreateNewLevel();
      ++st.num_positive;
      model_->AssignValue(var, value, nl);
      GacStats g = model_->EnforceGAC(false);
      st.gac_iterations += g.iterations;
      st.gac_deletions += g.deletions;
      if (g.inconsistent) {
        model_->BackToLevel(level);
        ++st.num_negative;
        continue;
      }
      const bool stop = Search(nl, st, time_limit, t0);
      model_->BackToLevel(level);
      if (stop) return true;
    }
    return false;
  }
  void Extract(int level) {
    std::vector<int> s(model_->num_vars());
    for (int v = 0; v < model_->num_vars(); ++v) {
      s[v] = model_->GetAssignedValue(v, level);
      if (s[v] == -1)
        throw std::runtime_error("[GModelSolver::ExtractSolution] Variable " + std::to_string(v) +
                                 " is not assigned at level " + std::to_string(level));
    }
    solutions_.push_back(std::move(s));
  }

  GModel* model_;
  bool verbose_;
  Branching branching_;
  int var_heuristic_ = CPIM_VARH_DOM_MIN;
  bool find_all_ = false;
  int max_solutions_ = 1;
  std::vector<std::vector<int>> solutions_;
};

// Legacy CModel entry points served by the same engine.
struct SearchStatistics {
  uint64_t num_sol = 0, num_positive = 0, num_negative = 0, nodes = 0;
  uint64_t build_time = 0, solve_time = 0, total_time = 0;  // milliseconds
  bool time_out = false, pass = false;
};

class CModel {
 public:
  explicit CModel(const cpim_model_desc& desc) : g_(model::GModelAdapter::Build(desc)) {}
  bool enforceGAC() { return !g_.EnforceGAC_Auto().inconsistent; }
  // type 1 = the decision on var was an assignment, 0 = a removal; either
  // way the constraints of var are propagated.
  bool enforceGAC(int var, int /*type*/) { return !g_.EnforceGAC_Auto(var).inconsistent; }
  void enforceSAC() { sac_ = g_.EnforceSAC(true); }
  bool sac_consistent() const { return sac_.consistent != 0; }
  SearchStatistics solve(float time_limit) {
    GModelSolver s(&g_);
    GpuSearchStatistics st = s.Solve((int)time_limit);
    SearchStatistics r;
    r.num_sol = st.num_solutions;
    r.num_positive = st.num_positive;
    r.num_negative = st.num_negative;
    r.solve_time = (uint64_t)(st.solve_time * 1000.0);
    r.time_out = st.time_out;
    solution_ = s.GetSolution();
    return r;
  }
  int CreateNewLevel() { return g_.CreateNewLevel(); }
  int BackLevel() {
    const int l = g_.GetCurrentLevel() - 1;
    g_.BackToLevel(l);
    return l;
  }
  const std::vector<int>& solution() const { return solution_; }
  GModel& gmodel() { return g_; }

 private:
  GModel g_;
  cpim_sac_stats sac_{};
  std::vector<int> solution_;
};

}  // namespace cpim
