// Sanitizer driver for the CPU code: TEST INFRASTRUCTURE ONLY.
//
// Built by tests/test_sanitizers.py with g++ -fsanitize=address,undefined
// together with oracle.cpp and the seeded generators (cpim_amd/csrc/
// workload.cpp), it runs every oracle entry point on seeded random networks
// (binary ones with pre-removed values, and ones with positive tables of
// arity 3 to 5) under every variable ordering and planted fault, and checks
// the restatement against the brute-force closures on the way.  Any
// out-of-bounds access, use after free, leak or undefined behaviour ends the
// run with the sanitizer's report and a non-zero status.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oracle.h"

extern "C" {
int cpimw_random_pairs(uint64_t seed, int n, int e, int32_t* scopes);
int cpimw_random_relations(uint64_t seed, int e, int d0, int d1, int words, double p, uint64_t* rel);
int cpimw_domains_with_removals(uint64_t seed, int n, int d, int words, double frac, uint64_t* out);
int cpimw_random_table(uint64_t seed, int arity, int d, int ntuples, int32_t* tuples);
}

namespace {

int g_fail = 0;

void expect(bool ok, const char* what, int seed) {
  if (!ok) {
    std::printf("MISMATCH %s (seed %d)\n", what, seed);
    g_fail = 1;
  }
}

struct Net {
  int n = 0, d = 0, W = 0;
  std::vector<int32_t> dom, arity, scope, tuples;
  std::vector<int64_t> off;
  std::vector<uint64_t> rel, init;
  cpim_model_desc desc{};
};

// Binary part from the generators; tables (arity 3..5) as support tuples.
void make(Net& N, int n, int d, int e, double p, int seed, double pre, int tables) {
  N.n = n;
  N.d = d;
  N.W = (d + 63) / 64;
  N.dom.assign(n, d);
  N.arity.assign(e, 2);
  N.scope.assign(2 * e, 0);
  N.rel.assign((size_t)e * d * N.W, 0);
  cpimw_random_pairs((uint64_t)seed, n, e, N.scope.data());
  cpimw_random_relations((uint64_t)seed, e, d, d, N.W, p, N.rel.data());
  N.off.clear();
  N.tuples.clear();
  if (tables > 0) {
    N.off.assign(e + 1, 0);
    for (int t = 0; t < tables; ++t) {
      const int ar = 3 + t % 3;
      const int nt = 40 + 10 * t;
      std::vector<int32_t> tup((size_t)nt * ar);
      cpimw_random_table((uint64_t)(seed * 31 + t), ar, d, nt, tup.data());
      N.arity.push_back(ar);
      for (int k = 0; k < ar; ++k) N.scope.push_back((t + 3 * k) % n);  // distinct for n = 8
      N.tuples.insert(N.tuples.end(), tup.begin(), tup.end());
      N.off.push_back((int64_t)N.tuples.size());
    }
  }
  N.init.clear();
  if (pre > 0) {
    N.init.assign((size_t)n * N.W, 0);
    cpimw_domains_with_removals((uint64_t)seed, n, d, N.W, pre, N.init.data());
  }
  cpim_model_desc& D = N.desc;
  std::memset(&D, 0, sizeof(D));
  D.num_vars = n;
  D.dom_size = N.dom.data();
  D.num_cons = (int32_t)N.arity.size();
  D.con_arity = N.arity.data();
  D.con_scope = N.scope.data();
  D.bin_rel = N.rel.data();
  D.rel_words = N.W;
  D.tup_off = N.off.empty() ? nullptr : N.off.data();
  D.tuples = N.tuples.empty() ? nullptr : N.tuples.data();
  D.init_dom = N.init.empty() ? nullptr : N.init.data();
}

void run(Net& N, int algo, int seed, bool brute) {
  const size_t words = (size_t)N.n * N.W;
  std::vector<uint64_t> a(words), b(words);
  orc_stats st;
  for (int varh = 0; varh < 3; ++varh) {
    orc_set_var_heuristic(varh);
    orc_net* o = orc_create(&N.desc, algo);
    const int ok = orc_root_ac(o, a.data(), &st);
    if (brute && varh == 0) {
      const int bok = orc_bf_gac(&N.desc, b.data());
      expect(ok == bok && (!ok || a == b), "root AC vs brute force", seed);
    }
    orc_net* o2 = orc_create(&N.desc, algo);
    const int sok = orc_sac1(o2, a.data(), &st);
    if (brute && varh == 0) {
      const int bok = orc_bf_sac(&N.desc, b.data());
      expect(sok == bok && (!sok || a == b), "SAC1 vs brute force", seed);
    }
    if (ok) {
      std::vector<int32_t> tests;
      for (int v = 0; v < N.n; v += 3)
        for (int x = 0; x < N.d; x += 2) {
          tests.push_back(v);
          tests.push_back(x);
        }
      std::vector<uint8_t> f1(tests.size() / 2), f2(tests.size() / 2);
      orc_singleton_tests(o, tests.data(), (int64_t)f1.size(), f1.data(), 1, &st);
      orc_singleton_tests(o, tests.data(), (int64_t)f2.size(), f2.data(), 3, &st);
      expect(f1 == f2, "singleton tests, 1 vs 3 threads", seed);
    }
    orc_weights(o, nullptr);
    std::vector<float> w(N.desc.num_cons + 1);
    orc_weights(o, w.data());
    orc_net* o3 = orc_create(&N.desc, algo);
    std::vector<int32_t> sol(N.n, -1);
    std::vector<cpim_node_record> tr(512);
    int32_t tl = 0;
    orc_mac(o3, 0, 3000, &st, sol.data(), tr.data(), (int32_t)tr.size(), &tl);
    orc_net* o4 = orc_create(&N.desc, algo);
    std::vector<int32_t> sols((size_t)N.n * 16, -1);
    orc_stats st4;
    orc_mac_all(o4, 0, 3000, 1000000, &st4, sols.data(), 16, tr.data(), (int32_t)tr.size(), &tl);
    if (brute && varh < 2) {
      orc_stats bs;
      std::vector<int32_t> bsols((size_t)N.n * 16, -1);
      std::vector<cpim_node_record> btr(512);
      int32_t btl = 0;
      orc_bf_mac_all(&N.desc, 3000, 1000000, &bs, bsols.data(), 16, btr.data(), (int32_t)btr.size(), &btl);
      expect(bs.num_sol == st4.num_sol && bs.num_positive == st4.num_positive &&
                 bs.num_negative == st4.num_negative && bsols == sols,
             "enumerating MAC vs brute-force MAC", seed);
      std::vector<int32_t> first(N.n, -1);
      orc_bf_count_solutions(&N.desc, 50, first.data());
    }
    (void)orc_dom_hash(a.data(), N.n, N.W);
    orc_destroy(o);
    orc_destroy(o2);
    orc_destroy(o3);
    orc_destroy(o4);
  }
  orc_set_var_heuristic(0);
}

}  // namespace

int main() {
  for (int seed = 1; seed <= 6; ++seed) {
    Net N;
    make(N, 9 + seed, 3 + seed % 4, 14 + 2 * seed, 0.35, seed, seed % 2 ? 0.2 : 0.0, 0);
    run(N, ORC_AC3BIT, seed, true);
    run(N, ORC_AC3, seed, true);
    Net T;
    make(T, 8, 4, 8, 0.3, 100 + seed, 0.0, 3);
    run(T, ORC_AC3, 100 + seed, true);
  }
  // domains of more than one word, and every planted fault
  Net L;
  make(L, 30, 130, 90, 0.5, 77, 0.1, 0);
  run(L, ORC_AC3BIT, 77, false);
  for (int fault = 1; fault <= 5; ++fault) {
    orc_set_fault(fault);
    Net F;
    make(F, 10, 5, 18, 0.35, 200 + fault, 0.0, 0);
    run(F, ORC_AC3BIT, 200 + fault, false);
    orc_set_fault(0);
  }
  std::printf(g_fail ? "sanitize driver: mismatches\n" : "sanitize driver: ok\n");
  return g_fail;
}
