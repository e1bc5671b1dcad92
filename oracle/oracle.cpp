// CPU oracle: TEST INFRASTRUCTURE ONLY (see oracle.h for who may load it and
// for the reference files each part restates).
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

namespace {

int g_fault = 0;
int g_varh = 0;  // 0 = dom-min (VRH_DOM_MIN), 1 = dom/deg (VRH_DOM_DEG_MIN), 2 = dom/wdeg

// Number of distinct other variables sharing a constraint with each variable
// (Network::get_neighbor, src/Network.cpp:467-474).
std::vector<int> neighbour_counts(int n, const std::vector<std::vector<int>>& scope) {
  std::vector<std::vector<int>> nb(n);
  for (const auto& sc : scope)
    for (int x : sc)
      for (int y : sc)
        if (x != y) nb[x].push_back(y);
  std::vector<int> cnt(n, 0);
  for (int v = 0; v < n; ++v) {
    std::sort(nb[v].begin(), nb[v].end());
    cnt[v] = (int)(std::unique(nb[v].begin(), nb[v].end()) - nb[v].begin());
  }
  return cnt;
}

// MAC::select_var (src/MAC.cpp:353-385): the first variable (lowest id) with
// the smallest key among the undecided ones.  dom-min: key = |D(v)|.
// dom/deg: key = |D(v)| / |neighbours(v)| in integer division, or -1 for a
// variable without neighbours.  size(v) gives |D(v)|; skip(v) the decided.
// planted: whether planted fault 2 (last of the tied variables) applies; the
// brute-force driver passes false so that it stays an independent check.
template <typename Size, typename Skip>
int select_var_h(int n, const std::vector<int>& nbr, Size size, Skip skip, bool planted) {
  int var = -1;
  long long best = 0;
  for (int v = 0; v < n; ++v) {
    if (skip(v)) continue;
    const int sz = size(v);
    long long key = sz;
    if (g_varh == 1) key = nbr[v] == 0 ? -1 : sz / nbr[v];
    if (var < 0 || key < best || (planted && g_fault == 2 && key == best)) {
      best = key;
      var = v;
    }
  }
  return var;
}

inline int popc(uint64_t x) { return __builtin_popcountll(x); }
inline int ctz(uint64_t x) { return __builtin_ctzll(x); }

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int max_dom_of(const cpim_model_desc* d) {
  int m = 0;
  for (int v = 0; v < d->num_vars; ++v) m = std::max(m, d->dom_size[v]);
  return m;
}

// ---------------------------------------------------------------------------
// Immutable network description shared by every worker copy.
// Network ctor semantics: src/Network.cpp:358-411 (every constraint of the
// normalized model is a positive table; subscription lists are in
// constraint order; tmp level = n + 2; n + 3 levels).
// ---------------------------------------------------------------------------
struct Shared {
  int n = 0, max_dom = 0, Wm = 0, max_arity = 0, ntabs = 0, algo = ORC_AC3BIT;
  std::vector<int> cap, nw;
  std::vector<std::vector<int>> scope;
  std::vector<std::vector<std::vector<int>>> tuples;  // sorted (Tabular ctor, :299-308)
  std::vector<std::vector<int>> subscription;         // var -> tabs (:402-403)
  std::vector<uint64_t> bitsup;                       // AC3bit rows (AC3bit.cpp:3-27)
  std::vector<uint64_t> init_dom;                     // [n][Wm]
  // Network::GetIntConValIndex (Network.h:337-340)
  size_t row_index(int c, int pos, int a) const {
    return ((size_t)c * max_arity + pos) * max_dom + a;
  }
};

std::shared_ptr<Shared> build_shared(const cpim_model_desc* d, int algo) {
  auto S = std::make_shared<Shared>();
  S->algo = algo;
  S->n = d->num_vars;
  S->max_dom = max_dom_of(d);
  S->Wm = S->max_dom == 0 ? 0 : (S->max_dom + 63) / 64;
  S->cap.assign(d->dom_size, d->dom_size + S->n);
  S->nw.resize(S->n);
  for (int v = 0; v < S->n; ++v) S->nw[v] = std::max(1, (S->cap[v] + 63) / 64);
  S->ntabs = d->num_cons;
  S->scope.resize(S->ntabs);
  S->subscription.resize(S->n);
  int64_t off = 0, bin_k = 0;
  std::vector<int64_t> bin_of(S->ntabs, -1);
  for (int c = 0; c < S->ntabs; ++c) {
    const int ar = d->con_arity[c];
    S->scope[c].assign(d->con_scope + off, d->con_scope + off + ar);
    off += ar;
    S->max_arity = std::max(S->max_arity, ar);
    if (ar == 2) bin_of[c] = bin_k++;
  }
  for (int c = 0; c < S->ntabs; ++c)
    for (int v : S->scope[c]) S->subscription[v].push_back(c);
  const int Wr = d->rel_words;
  // Tuples are needed by the generic AC3 and for non-dense constraints.
  auto dense = [&](int c) { return d->bin_rel != nullptr && bin_of[c] >= 0; };
  auto rel_bit = [&](int c, int a, int b) {
    const uint64_t* row = d->bin_rel + ((size_t)bin_of[c] * S->max_dom + a) * Wr;
    return (row[b >> 6] >> (b & 63)) & 1ull;
  };
  if (algo == ORC_AC3 || algo == ORC_AC3MIX) {
    S->tuples.resize(S->ntabs);
    for (int c = 0; c < S->ntabs; ++c) {
      auto& T = S->tuples[c];
      const int ar = (int)S->scope[c].size();
      if (algo == ORC_AC3MIX && ar == 2) continue;  // revised through bitSup rows
      if (dense(c)) {
        const int x = S->scope[c][0], y = S->scope[c][1];
        for (int a = 0; a < S->cap[x]; ++a)
          for (int b = 0; b < S->cap[y]; ++b)
            if (rel_bit(c, a, b)) T.push_back({a, b});
      } else {
        for (int64_t t = d->tup_off[c]; t + ar <= d->tup_off[c + 1]; t += ar)
          T.emplace_back(d->tuples + t, d->tuples + t + ar);
      }
      std::sort(T.begin(), T.end());
    }
  }
  if (algo == ORC_AC3BIT || algo == ORC_AC3MIX) {
    // AC3bit: bitSup[(c, pos, a)][word] (AC3bit.cpp:10-26).
    S->bitsup.assign((size_t)S->ntabs * S->max_arity * S->max_dom * S->Wm, 0);
    auto set_bit = [&](size_t row, int b) { S->bitsup[row * S->Wm + (b >> 6)] |= 1ull << (b & 63); };
    for (int c = 0; c < S->ntabs; ++c) {
      if (S->scope[c].size() != 2) continue;  // AC3bit handles binary constraints only
      const int x = S->scope[c][0], y = S->scope[c][1];
      if (dense(c)) {
        // every allowed tuple (a, b): bitSup[(c,0,a)] gets b, bitSup[(c,1,b)] gets a
        for (int a = 0; a < S->cap[x]; ++a) {
          const uint64_t* row = d->bin_rel + ((size_t)bin_of[c] * S->max_dom + a) * Wr;
          for (int w = 0; w < S->Wm; ++w) {
            uint64_t bits = row[w];
            if (64 * w + 64 > S->cap[y]) {
              const int k = S->cap[y] - 64 * w;
              bits &= k <= 0 ? 0ull : (k >= 64 ? ~0ull : ((1ull << k) - 1ull));
            }
            S->bitsup[S->row_index(c, 0, a) * S->Wm + w] = bits;
            while (bits) {
              const int b = 64 * w + ctz(bits);
              bits &= bits - 1;
              set_bit(S->row_index(c, 1, b), a);
            }
          }
        }
      } else {
        for (int64_t t = d->tup_off[c]; t + 2 <= d->tup_off[c + 1]; t += 2) {
          const int a = d->tuples[t], b = d->tuples[t + 1];
          set_bit(S->row_index(c, 0, a), b);
          set_bit(S->row_index(c, 1, b), a);
        }
      }
    }
  }
  S->init_dom.assign((size_t)S->n * S->Wm, 0);
  for (int v = 0; v < S->n; ++v) {
    uint64_t* row = &S->init_dom[(size_t)v * S->Wm];
    for (int a = 0; a < S->cap[v]; ++a) {
      const bool keep = d->init_dom == nullptr ||
                        ((d->init_dom[(size_t)v * S->Wm + (a >> 6)] >> (a & 63)) & 1ull);
      if (keep) row[a >> 6] |= 1ull << (a & 63);
    }
  }
  return S;
}

// ---------------------------------------------------------------------------
// Mutable state: IntVar multi-level domains + assigned flags
// (Network.h:142-193, Network.cpp:62-256, :438-455).
// ---------------------------------------------------------------------------
struct Net {
  std::shared_ptr<const Shared> S;
  int n = 0, Wm = 0, levels = 0;
  std::vector<uint64_t> dom;  // [levels][n][Wm]
  std::vector<uint8_t> asg;   // [levels][n]
  std::vector<int> vtop;      // IntVar::top_
  std::vector<float> weight;  // Tabular::weight (wdeg bookkeeping, AC3.cpp:125)

  explicit Net(std::shared_ptr<const Shared> s) : S(std::move(s)) {
    n = S->n;
    Wm = std::max(1, S->Wm);
    levels = n + 3;
    dom.assign((size_t)levels * n * Wm, 0);
    asg.assign((size_t)levels * n, 0);
    vtop.assign(n, 0);
    weight.assign(S->ntabs, 1.0f);
    for (int p = 0; p < levels; ++p)
      for (int v = 0; v < n; ++v)
        std::memcpy(D(v, p), &S->init_dom[(size_t)v * S->Wm], sizeof(uint64_t) * S->Wm);
  }
  int tmp() const { return n + 2; }
  uint64_t* D(int v, int p) { return &dom[((size_t)p * n + v) * Wm]; }
  const uint64_t* D(int v, int p) const { return &dom[((size_t)p * n + v) * Wm]; }
  bool assigned(int v, int p) const { return asg[(size_t)p * n + v] != 0; }
  void assign(int v, bool a, int p) { asg[(size_t)p * n + v] = a; }
  int size(int v, int p) const {
    int s = 0;
    const uint64_t* d = D(v, p);
    for (int w = 0; w < S->nw[v]; ++w) s += popc(d[w]);
    return s;
  }
  int head(int v, int p) const {
    const uint64_t* d = D(v, p);
    for (int w = 0; w < S->nw[v]; ++w)
      if (d[w]) return w * 64 + ctz(d[w]);
    return -1;
  }
  int next(int v, int a, int p) const {  // IntVar::next (Network.cpp:132-145)
    const uint64_t* d = D(v, p);
    int w = a >> 6;
    const int b = a & 63;
    uint64_t rest = b == 63 ? 0 : (d[w] >> b) >> 1;
    if (rest) return a + ctz(rest) + 1;
    for (++w; w < S->nw[v]; ++w)
      if (d[w]) return w * 64 + ctz(d[w]);
    return -1;
  }
  bool have(int v, int a, int p) const {
    if (a < 0) return false;
    return (D(v, p)[a >> 6] >> (a & 63)) & 1ull;
  }
  void remove(int v, int a, int p) { D(v, p)[a >> 6] &= ~(1ull << (a & 63)); }
  void reduce_to(int v, int a, int p) {  // IntVar::ReduceTo sets assigned_[p]
    uint64_t* d = D(v, p);
    for (int w = 0; w < Wm; ++w) d[w] = 0;
    d[a >> 6] = 1ull << (a & 63);
    assign(v, true, p);
  }
  void var_copy(int v, int src, int dst) {
    std::memcpy(D(v, dst), D(v, src), sizeof(uint64_t) * Wm);
    assign(v, assigned(v, src), dst);
  }
  void new_level(int src) {  // Network::NewLevel -> IntVar::new_level
    for (int v = 0; v < n; ++v) {
      var_copy(v, src, src + 1);
      vtop[v] = src;
    }
  }
  void back_to(int dest) {  // IntVar::BackTo clears flags in [dest, top_]
    for (int v = 0; v < n; ++v) {
      for (int i = dest; i <= vtop[v]; ++i) assign(v, false, i);
      vtop[v] = dest;
    }
  }
  void var_copy_from(const Net& o, int v) {  // level 0 of another copy
    std::memcpy(D(v, 0), o.D(v, 0), sizeof(uint64_t) * Wm);
    assign(v, o.assigned(v, 0), 0);
  }
  void copy_level(int src, int dst) {
    if (src == dst) return;
    for (int v = 0; v < n; ++v) var_copy(v, src, dst);
  }
  void clear_level(int p) {
    for (int v = 0; v < n; ++v) assign(v, false, p);
  }
  uint64_t hash(int p) const {
    const int W = S->Wm;
    uint64_t h = 0;
    for (int v = 0; v < n; ++v)
      for (int w = 0; w < W; ++w)
        h += mix64(D(v, p)[w] + (uint64_t)(v * W + w) * 0x9E3779B97F4A7C15ull);
    return h;
  }
};

// ---------------------------------------------------------------------------
// AC3 with time stamps over a FIFO variable queue (AC3.cpp:99-193,
// Solver.cpp:180-218) and the two seek_support variants.
// ---------------------------------------------------------------------------
struct AC3 {
  Net& N;
  std::vector<int64_t> stamp_var, stamp_tab;
  int64_t t = 0;
  std::deque<int> q;
  std::vector<uint8_t> inq;
  std::vector<int> tuple;
  int64_t num_delete = 0, revisions = 0;

  explicit AC3(Net& net) : N(net) {
    stamp_var.assign(N.n, 0);
    stamp_tab.assign(N.S->ntabs, 0);
    inq.assign(N.n, 0);
    tuple.assign(std::max(1, N.S->max_arity), 0);
  }
  void insert(int v) {
    if (!inq[v]) {
      q.push_back(v);
      inq[v] = 1;
    }
    ++t;
    stamp_var[v] = t;
  }
  bool seek_bit(int c, int y, int a, int p) {
    const Shared& S = *N.S;
    const auto& sc = S.scope[c];
    int pos = -1;
    for (int i = (int)sc.size() - 1; i >= 0; --i)
      if (sc[i] == y) { pos = i; break; }  // Tabular::index (Network.cpp:337-341)
    const uint64_t* row = &S.bitsup[S.row_index(c, pos, a) * S.Wm];
    for (int z : sc) {
      if (z == y) continue;
      int words = S.nw[z];
      if (g_fault == 1 && words > 1) --words;
      const uint64_t* dz = N.D(z, p);
      for (int i = 0; i < words; ++i)
        if (row[i] & dz[i]) return true;
    }
    return false;
  }
  bool seek_generic(int c, int y, int a, int p) {
    // GetFirstValidTuple / GetNextValidTuple + binary search (Network.cpp:310-335).
    const auto& sc = N.S->scope[c];
    const auto& T = N.S->tuples[c];
    const int ar = (int)sc.size();
    std::vector<int> cur(ar);
    for (int i = 0; i < ar; ++i) cur[i] = sc[i] != y ? N.head(sc[i], p) : a;
    while (true) {
      if (std::binary_search(T.begin(), T.end(), cur)) return true;
      int i = ar - 1;
      for (; i >= 0; --i) {
        if (sc[i] == y) continue;
        const int nx = N.next(sc[i], cur[i], p);
        if (nx < 0) {
          cur[i] = N.head(sc[i], p);
        } else {
          cur[i] = nx;
          break;
        }
      }
      if (i < 0) return false;
    }
  }
  bool seek(int c, int y, int a, int p) {
    const int algo = N.S->algo;
    if (algo == ORC_AC3BIT || (algo == ORC_AC3MIX && N.S->scope[c].size() == 2)) return seek_bit(c, y, a, p);
    return seek_generic(c, y, a, p);
  }
  bool revise(int c, int y, int p) {
    ++revisions;
    const int before = N.size(y, p);
    int a = N.head(y, p);
    while (a >= 0) {
      const int nxt = N.next(y, a, p);
      const bool skip = g_fault == 3 && nxt < 0;
      if (!skip && !seek(c, y, a, p)) {
        N.remove(y, a, p);
        ++num_delete;
      }
      a = nxt;
    }
    return before != N.size(y, p);
  }
  // Returns consistency (ConsistencyState::state).
  bool enforce(const std::vector<int>& evt, int level) {
    std::fill(stamp_var.begin(), stamp_var.end(), 0);
    std::fill(stamp_tab.begin(), stamp_tab.end(), 0);
    t = 0;
    q.clear();
    std::fill(inq.begin(), inq.end(), 0);
    for (int v : evt) insert(v);
    while (!q.empty()) {
      const int x = q.front();
      q.pop_front();
      inq[x] = 0;
      for (int c : N.S->subscription[x]) {
        if (stamp_var[x] > stamp_tab[c]) {
          for (int y : N.S->scope[c]) {
            if (N.assigned(y, level)) continue;
            bool aa = false;
            for (int z : N.S->scope[c])
              if (z != x && stamp_var[z] > stamp_tab[c]) aa = true;
            if ((y != x) || aa) {
              if (revise(c, y, level)) {
                if (N.size(y, level) == 0) {
                  if (g_fault != 5) N.weight[c] += 1.0f;
                  return false;
                }
                insert(y);
              }
            }
          }
          ++t;
          stamp_tab[c] = t;
        }
      }
    }
    return true;
  }
};

// MAC::select_var VRH_DOM_WDEG_MIN (src/MAC.cpp:386-411): for each
// undecided x, x_w = sum of the weights of the constraints on x with more
// than one undecided variable; key = -1 when |D(x)| = 1 or x_w = 0, else
// |D(x)| / x_w in double; the first variable with the smallest key.  The
// weights are the Tabular::weight counts that AC3 raises on a wipe-out.
int select_var_wdeg(const Net& N, int p) {
  int var = -1;
  double best = 1.7976931348623157e308;  // DBL_MAX
  for (int x = 0; x < N.n; ++x) {
    if (N.assigned(x, p)) continue;
    double xw = 0.0;
    for (int c : N.S->subscription[x]) {
      int cnt = 0;
      for (int y : N.S->scope[c])
        if (!N.assigned(y, p)) ++cnt;
      if (cnt > 1) xw += N.weight[c];
    }
    const int sz = N.size(x, p);
    const double key = (sz == 1 || xw == 0.0) ? -1.0 : sz / xw;
    if (key < best) {
      best = key;
      var = x;
    }
  }
  return var;
}

std::vector<int> all_vars(int n) {
  std::vector<int> v(n);
  for (int i = 0; i < n; ++i) v[i] = i;
  return v;
}

// ---------------------------------------------------------------------------
// Brute force (part 2): explicit relations, definitional closure.
// ---------------------------------------------------------------------------
struct BfModel {
  int n = 0, max_dom = 0, W = 0;
  std::vector<int> cap;
  std::vector<std::vector<int>> scope;
  std::vector<std::vector<uint8_t>> bin;             // binary: [a * max_dom + b]
  std::vector<std::vector<std::vector<int>>> tuples;  // n-ary
  std::vector<std::vector<uint8_t>> init;            // [v][a]
  std::vector<std::vector<int>> by_var;              // constraints per var
};

BfModel bf_model(const cpim_model_desc* d) {
  BfModel M;
  M.n = d->num_vars;
  M.max_dom = max_dom_of(d);
  M.W = (M.max_dom + 63) / 64;
  M.cap.assign(d->dom_size, d->dom_size + M.n);
  M.scope.resize(d->num_cons);
  M.bin.resize(d->num_cons);
  M.tuples.resize(d->num_cons);
  M.by_var.resize(M.n);
  int64_t off = 0, bk = 0;
  for (int c = 0; c < d->num_cons; ++c) {
    const int ar = d->con_arity[c];
    M.scope[c].assign(d->con_scope + off, d->con_scope + off + ar);
    off += ar;
    for (int v : M.scope[c]) M.by_var[v].push_back(c);
    if (ar == 2) {
      auto& B = M.bin[c];
      B.assign((size_t)M.max_dom * M.max_dom, 0);
      if (d->bin_rel) {
        for (int a = 0; a < M.max_dom; ++a)
          for (int b = 0; b < M.max_dom; ++b) {
            const uint64_t* row = d->bin_rel + ((size_t)bk * M.max_dom + a) * d->rel_words;
            B[(size_t)a * M.max_dom + b] = (row[b >> 6] >> (b & 63)) & 1ull;
          }
      } else {
        for (int64_t t = d->tup_off[c]; t + 2 <= d->tup_off[c + 1]; t += 2)
          B[(size_t)d->tuples[t] * M.max_dom + d->tuples[t + 1]] = 1;
      }
      ++bk;
    } else {
      for (int64_t t = d->tup_off[c]; t + ar <= d->tup_off[c + 1]; t += ar)
        M.tuples[c].emplace_back(d->tuples + t, d->tuples + t + ar);
    }
  }
  M.init.assign(M.n, std::vector<uint8_t>(M.max_dom, 0));
  for (int v = 0; v < M.n; ++v)
    for (int a = 0; a < M.cap[v]; ++a)
      M.init[v][a] = d->init_dom == nullptr
                         ? 1
                         : (uint8_t)((d->init_dom[(size_t)v * M.W + (a >> 6)] >> (a & 63)) & 1ull);
  return M;
}

using BfDom = std::vector<std::vector<uint8_t>>;

bool bf_supported(const BfModel& M, const BfDom& D, int c, int pos, int a) {
  const auto& sc = M.scope[c];
  if (sc.size() == 2) {
    const int other = sc[1 - pos];
    for (int b = 0; b < M.cap[other]; ++b) {
      if (!D[other][b]) continue;
      const int x0 = pos == 0 ? a : b, x1 = pos == 0 ? b : a;
      if (M.bin[c][(size_t)x0 * M.max_dom + x1]) return true;
    }
    return false;
  }
  for (const auto& t : M.tuples[c]) {
    if (t[pos] != a) continue;
    bool ok = true;
    for (size_t j = 0; j < sc.size() && ok; ++j) ok = D[sc[j]][t[j]] != 0;
    if (ok) return true;
  }
  return false;
}

// Iterate "keep a iff every constraint on x has a valid support" to the
// fixpoint (samples/verify_gac.cpp:83-118 pattern).  Returns false iff some
// domain is empty.
bool bf_close(const BfModel& M, BfDom& D) {
  bool changed = true;
  while (changed) {
    changed = false;
    for (size_t c = 0; c < M.scope.size(); ++c) {
      const auto& sc = M.scope[c];
      for (size_t pos = 0; pos < sc.size(); ++pos) {
        const int x = sc[pos];
        for (int a = 0; a < M.cap[x]; ++a) {
          if (D[x][a] && !bf_supported(M, D, (int)c, (int)pos, a)) {
            D[x][a] = 0;
            changed = true;
          }
        }
      }
    }
  }
  for (int v = 0; v < M.n; ++v) {
    bool any = false;
    for (int a = 0; a < M.cap[v]; ++a) any |= D[v][a] != 0;
    if (!any) return false;
  }
  return true;
}

void bf_out(const BfModel& M, const BfDom& D, uint64_t* out) {
  if (!out) return;
  std::memset(out, 0, sizeof(uint64_t) * (size_t)M.n * M.W);
  for (int v = 0; v < M.n; ++v)
    for (int a = 0; a < M.cap[v]; ++a)
      if (D[v][a]) out[(size_t)v * M.W + (a >> 6)] |= 1ull << (a & 63);
}

uint64_t bf_hash(const BfModel& M, const BfDom& D) {
  std::vector<uint64_t> w((size_t)M.n * M.W);
  bf_out(M, D, w.data());
  uint64_t h = 0;
  for (size_t i = 0; i < w.size(); ++i) h += mix64(w[i] + (uint64_t)i * 0x9E3779B97F4A7C15ull);
  return h;
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface
// ---------------------------------------------------------------------------
struct orc_net {
  std::shared_ptr<Shared> S;
  std::unique_ptr<Net> N;
  std::vector<int> nbr;  // neighbour counts (Network::neighborhood sizes)
};

extern "C" {

void orc_set_fault(int fault) { g_fault = fault; }
void orc_set_var_heuristic(int varh) { g_varh = varh; }

orc_net* orc_create(const cpim_model_desc* desc, int algo) {
  auto* o = new orc_net;
  o->S = build_shared(desc, algo);
  o->N = std::make_unique<Net>(o->S);
  o->nbr = neighbour_counts(o->S->n, o->S->scope);
  return o;
}

void orc_destroy(orc_net* net) { delete net; }
int orc_num_vars(const orc_net* net) { return net->S->n; }
int orc_dom_words(const orc_net* net) { return net->S->Wm; }
int orc_weights(const orc_net* net, float* out) {
  const auto& w = net->N->weight;
  if (out) std::copy(w.begin(), w.end(), out);
  return (int)w.size();
}

static void copy_out(const Net& N, int level, uint64_t* out) {
  if (!out) return;
  const int W = N.S->Wm;
  for (int v = 0; v < N.n; ++v) std::memcpy(out + (size_t)v * W, N.D(v, level), sizeof(uint64_t) * W);
}

int orc_root_ac(orc_net* net, uint64_t* out, orc_stats* st) {
  Net& N = *net->N;
  AC3 ac(N);
  const bool ok = ac.enforce(all_vars(N.n), 0);
  copy_out(N, 0, out);
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->revisions = ac.revisions;
    st->deletions = ac.num_delete;
    st->fixpoints = 1;
    st->consistent = ok;
  }
  return ok;
}

// SAC1::enforce (SAC1.cpp:57-103).
int orc_sac1(orc_net* net, uint64_t* out, orc_stats* st) {
  Net& N = *net->N;
  AC3 ac(N);
  const int level = 0;
  int64_t fixpoints = 1;
  bool result = ac.enforce(all_vars(N.n), level);
  bool ok = result;
  if (result) {
    bool modified;
    do {
      modified = false;
      for (int x = 0; x < N.n && ok; ++x) {
        for (int a = 0; a < N.S->cap[x]; ++a) {
          if (!N.have(x, a, level)) continue;
          N.copy_level(level, N.tmp());
          N.reduce_to(x, a, N.tmp());
          N.assign(x, true, N.tmp());
          result = ac.enforce({x}, N.tmp());
          ++fixpoints;
          N.assign(x, false, level);
          N.clear_level(N.tmp());
          if (!result) {
            N.remove(x, a, level);
            result = ac.enforce({x}, level);
            ++fixpoints;
            if (!result) {
              ok = false;
              break;
            }
            modified = true;
          }
        }
      }
      if (g_fault == 4) break;
    } while (ok && modified);
  }
  copy_out(N, level, out);
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->revisions = ac.revisions;
    st->deletions = ac.num_delete;
    st->fixpoints = fixpoints;
    st->consistent = ok;
  }
  return ok;
}

int orc_singleton_tests(orc_net* net, const int32_t* tests, int64_t count, uint8_t* fail,
                        int nthreads, orc_stats* st) {
  if (nthreads < 1) nthreads = 1;
  std::vector<int64_t> revs(nthreads, 0);
  auto worker = [&](int tid) {
    Net W(net->S);
    // Start from the current level 0 of the caller's network.
    for (int v = 0; v < W.n; ++v) W.var_copy_from(*net->N, v);
    AC3 ac(W);
    for (int64_t i = tid; i < count; i += nthreads) {
      const int x = tests[2 * i], a = tests[2 * i + 1];
      W.copy_level(0, W.tmp());
      W.reduce_to(x, a, W.tmp());
      W.assign(x, true, W.tmp());
      fail[i] = ac.enforce({x}, W.tmp()) ? 0 : 1;
      W.assign(x, false, 0);
      W.clear_level(W.tmp());
    }
    revs[tid] = ac.revisions;
  };
  if (nthreads == 1) {
    worker(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker, t);
    for (auto& t : th) t.join();
  }
  if (st) {
    std::memset(st, 0, sizeof(*st));
    for (auto r : revs) st->revisions += r;
    st->fixpoints = count;
    st->consistent = 1;
  }
  return 0;
}

// MAC::enforce (MAC.cpp:100-165) with select_var VRH_DOM_MIN (:357-364),
// VRH_DOM_DEG_MIN (:371-384) or VRH_DOM_WDEG_MIN (:386-411) as g_varh says,
// select_val VLH_MIN (:423-425); AssignedStack (Solver.cpp:36-116).  The
// constraint weights live in the network and persist across calls, as
// Tabular::weight does.
//
// Enumeration (max_solutions > 1) is not in the reference's MAC::enforce, which
// stops at the first solution; its commented-out continuation (MAC.cpp:138-139,
// "++sol_count_; consistent_ = false;") is what is done here: a solution is
// counted and stored, then treated as a failed node, so the search backtracks
// and goes on until the tree is exhausted or max_solutions are found.  This
// is the search GModelSolver::SetFindAllSolutions / SetMaxSolutions ask for
// (src/GModelSolver.cu:20-29, 125-128), in the CPU MAC's binary branching.
int orc_mac_all(orc_net* net, int64_t time_limit_ms, int64_t node_limit, int64_t max_solutions,
                orc_stats* st, int32_t* solutions, int64_t sol_cap, cpim_node_record* trace,
                int32_t trace_cap, int32_t* trace_len) {
  Net& N = *net->N;
  AC3 ac(N);
  orc_stats s;
  std::memset(&s, 0, sizeof(s));
  int32_t tl = 0;
  auto record = [&](int pos, int v, int a, bool ok, int level) {
    if (trace && tl < trace_cap) {
      trace[tl].positive = pos;
      trace[tl].var = v;
      trace[tl].value = a;
      trace[tl].consistent = ok;
      trace[tl].dom_hash = ok ? N.hash(level) : 0;
    }
    ++tl;
  };
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed_ms = [&]() {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0)
        .count();
  };
  std::vector<std::pair<int, int>> I;  // AssignedStack::vals_
  bool consistent = ac.enforce(all_vars(N.n), 0);
  ++s.fixpoints;
  bool finished = false;
  if (consistent) {
    while (!finished) {
      if (time_limit_ms > 0 && elapsed_ms() > time_limit_ms) {
        s.time_out = 1;
        break;
      }
      if (node_limit > 0 && s.num_positive + s.num_negative >= node_limit) {
        s.time_out = 1;
        break;
      }
      const int p = (int)I.size();
      const int var = g_varh == 2 ? select_var_wdeg(N, p)
                                  : select_var_h(
                                        N.n, net->nbr, [&](int v) { return N.size(v, p); },
                                        [&](int v) { return N.assigned(v, p); }, true);
      const int val = N.head(var, p);
      // An empty domain that no revision reached (a variable without
      // constraints whose initial domain is empty) is selected first.  The
      // reference would then call ReduceTo with value -1, an access outside
      // the domain; here the search ends without a solution instead, and
      // no node is recorded.
      if (val < 0) {
        finished = true;
        break;
      }
      N.new_level(p);
      I.push_back({var, val});
      N.assign(var, true, (int)I.size());
      ++s.num_positive;
      N.reduce_to(var, val, (int)I.size());
      consistent = ac.enforce({var}, (int)I.size());
      ++s.fixpoints;
      record(1, var, val, consistent, (int)I.size());
      if (consistent && (int)I.size() == N.n) {
        if (solutions && s.num_sol < sol_cap)
          for (int v = 0; v < N.n; ++v) solutions[s.num_sol * N.n + v] = N.head(v, (int)I.size());
        ++s.num_sol;
        if (max_solutions <= 1 || s.num_sol >= max_solutions) {
          finished = true;
          break;
        }
        consistent = false;  // go on: backtrack as after a failed node
      }
      while (!consistent && !I.empty()) {
        const auto va = I.back();
        I.pop_back();
        N.assign(va.first, false, (int)I.size());
        const int L = (int)I.size();
        N.back_to(L);
        N.remove(va.first, va.second, L);
        ++s.num_negative;
        consistent = N.size(va.first, L) != 0 && ac.enforce({va.first}, L);
        if (N.size(va.first, L) != 0) ++s.fixpoints;
        record(0, va.first, va.second, consistent, L);
        for (const auto& e : I) N.assign(e.first, true, L);  // update_model_assigned
      }
      if (!consistent) finished = true;
    }
  }
  s.revisions = ac.revisions;
  s.deletions = ac.num_delete;
  s.consistent = s.num_sol > 0;
  if (st) *st = s;
  if (trace_len) *trace_len = tl;
  return (int)s.num_sol;
}

int orc_mac(orc_net* net, int64_t time_limit_ms, int64_t node_limit, orc_stats* st,
            int32_t* solution, cpim_node_record* trace, int32_t trace_cap, int32_t* trace_len) {
  return orc_mac_all(net, time_limit_ms, node_limit, 1, st, solution, solution ? 1 : 0, trace, trace_cap,
                     trace_len);
}

uint64_t orc_dom_hash(const uint64_t* dom, int n, int words) {
  uint64_t h = 0;
  for (int64_t i = 0; i < (int64_t)n * words; ++i)
    h += mix64(dom[i] + (uint64_t)i * 0x9E3779B97F4A7C15ull);
  return h;
}

int orc_bf_gac(const cpim_model_desc* desc, uint64_t* out) {
  BfModel M = bf_model(desc);
  BfDom D = M.init;
  const bool ok = bf_close(M, D);
  bf_out(M, D, out);
  return ok;
}

int orc_bf_sac(const cpim_model_desc* desc, uint64_t* out) {
  BfModel M = bf_model(desc);
  BfDom D = M.init;
  bool ok = bf_close(M, D);
  bool changed = ok;
  while (ok && changed) {
    changed = false;
    for (int x = 0; x < M.n && ok; ++x) {
      for (int a = 0; a < M.cap[x]; ++a) {
        if (!D[x][a]) continue;
        BfDom T = D;
        std::fill(T[x].begin(), T[x].end(), 0);
        T[x][a] = 1;
        if (!bf_close(M, T)) {
          D[x][a] = 0;
          changed = true;
          if (!bf_close(M, D)) {
            ok = false;
            break;
          }
        }
      }
    }
  }
  bf_out(M, D, out);
  return ok;
}

int64_t orc_bf_count_solutions(const cpim_model_desc* desc, int64_t limit, int32_t* first) {
  BfModel M = bf_model(desc);
  // Constraints checked when their highest-index variable is assigned.
  std::vector<std::vector<int>> check_at(M.n);
  for (size_t c = 0; c < M.scope.size(); ++c) {
    int mx = 0;
    for (int v : M.scope[c]) mx = std::max(mx, v);
    check_at[mx].push_back((int)c);
  }
  std::vector<int> val(M.n, -1);
  int64_t count = 0;
  auto ok_at = [&](int v) {
    for (int c : check_at[v]) {
      const auto& sc = M.scope[c];
      if (sc.size() == 2) {
        if (!M.bin[c][(size_t)val[sc[0]] * M.max_dom + val[sc[1]]]) return false;
      } else {
        std::vector<int> t(sc.size());
        for (size_t j = 0; j < sc.size(); ++j) t[j] = val[sc[j]];
        bool found = false;
        for (const auto& u : M.tuples[c])
          if (u == t) { found = true; break; }
        if (!found) return false;
      }
    }
    return true;
  };
  int v = 0;
  while (v >= 0) {
    if (v == M.n) {
      if (count == 0 && first)
        for (int i = 0; i < M.n; ++i) first[i] = val[i];
      ++count;
      if (limit > 0 && count >= limit) break;
      --v;
      continue;
    }
    int a = val[v] + 1;
    while (a < M.cap[v] && !M.init[v][a]) ++a;
    if (a >= M.cap[v]) {
      val[v] = -1;
      --v;
      continue;
    }
    val[v] = a;
    if (ok_at(v)) ++v;
  }
  return count;
}

int orc_bf_mac_all(const cpim_model_desc* desc, int64_t node_limit, int64_t max_solutions, orc_stats* st,
                   int32_t* solutions, int64_t sol_cap, cpim_node_record* trace, int32_t trace_cap,
                   int32_t* trace_len) {
  // dom/wdeg has no brute-force counterpart: its weights follow AC3's order
  if (g_varh == 2) return -1;
  BfModel M = bf_model(desc);
  const std::vector<int> nbr = neighbour_counts(M.n, M.scope);
  orc_stats s;
  std::memset(&s, 0, sizeof(s));
  int32_t tl = 0;
  auto record = [&](int pos, int v, int a, bool ok, const BfDom& D) {
    if (trace && tl < trace_cap) {
      trace[tl].positive = pos;
      trace[tl].var = v;
      trace[tl].value = a;
      trace[tl].consistent = ok;
      trace[tl].dom_hash = ok ? bf_hash(M, D) : 0;
    }
    ++tl;
  };
  auto dsize = [&](const BfDom& D, int v) {
    int k = 0;
    for (int a = 0; a < M.cap[v]; ++a) k += D[v][a];
    return k;
  };
  BfDom D = M.init;
  bool ok = bf_close(M, D);
  struct Frame { BfDom saved; int var, val; };
  std::vector<Frame> stack;
  std::vector<uint8_t> decided(M.n, 0);
  while (ok) {
    if (node_limit > 0 && s.num_positive + s.num_negative >= node_limit) {
      s.time_out = 1;
      break;
    }
    const int x = select_var_h(
        M.n, nbr, [&](int v) { return dsize(D, v); }, [&](int v) { return decided[v] != 0; }, false);
    int a = 0;
    while (!D[x][a]) ++a;
    stack.push_back({D, x, a});
    decided[x] = 1;
    ++s.num_positive;
    std::fill(D[x].begin(), D[x].end(), 0);
    D[x][a] = 1;
    ok = bf_close(M, D);
    record(1, x, a, ok, D);
    if (ok && (int)stack.size() == M.n) {
      if (solutions && s.num_sol < sol_cap)
        for (int v = 0; v < M.n; ++v) {
          int b = 0;
          while (!D[v][b]) ++b;
          solutions[s.num_sol * M.n + v] = b;
        }
      ++s.num_sol;
      if (max_solutions <= 1 || s.num_sol >= max_solutions) break;
      ok = false;  // go on: backtrack as after a failed node
    }
    while (!ok && !stack.empty()) {
      Frame f = std::move(stack.back());
      stack.pop_back();
      decided[f.var] = 0;
      D = std::move(f.saved);
      D[f.var][f.val] = 0;
      ++s.num_negative;
      ok = dsize(D, f.var) > 0 && bf_close(M, D);
      record(0, f.var, f.val, ok, D);
    }
  }
  s.consistent = s.num_sol > 0;
  if (st) *st = s;
  if (trace_len) *trace_len = tl;
  return (int)s.num_sol;
}

int orc_bf_mac(const cpim_model_desc* desc, int64_t node_limit, orc_stats* st,
               int32_t* solution, cpim_node_record* trace, int32_t trace_cap,
               int32_t* trace_len) {
  return orc_bf_mac_all(desc, node_limit, 1, st, solution, solution ? 1 : 0, trace, trace_cap, trace_len);
}

}  // extern "C"
