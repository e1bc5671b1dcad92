// Seeded synthetic constraint networks for parity tests and benchmarks.
//
// Every generator is pure integer arithmetic (splitmix64 + xoshiro256**), so
// the same seed gives bit-identical instances on every compiler and host.
// Instances are handed to both the device engine and the CPU oracle through
// the same plain arrays, which is what makes bit-exact comparison possible.
//
// Relation encoding (shared with the C-ABI, see include/cpim_gpu.h):
//   rel[c][a][w]  (uint64, W = ceil(max_dom/64) words per row)
//   bit (b % 64) of word (b / 64) is set iff (scope0 = a, scope1 = b) is allowed.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Xoshiro256 {
  uint64_t s[4];
  explicit Xoshiro256(uint64_t seed) {
    uint64_t t = seed;
    for (auto& v : s) v = splitmix64(t);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  inline uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // Uniform integer in [0, bound) by 128-bit multiply (Lemire, no rejection:
  // the bias is below 2^-40 for the bounds used here and is deterministic).
  inline uint64_t below(uint64_t bound) {
    return (uint64_t)(((__uint128_t)next() * bound) >> 64);
  }
};

inline uint64_t threshold_of(double p) {
  if (p <= 0.0) return 0;
  if (p >= 1.0) return ~0ull;
  return (uint64_t)std::ldexp(p, 64);
}

inline uint64_t sub_seed(uint64_t seed, uint64_t k) {
  uint64_t t = seed ^ (0xD1B54A32D192ED03ull * (k + 1));
  return splitmix64(t);
}

int n_threads_for(int64_t work) {
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 1;
  if (hc > 16) hc = 16;
  if (work < (1 << 16)) return 1;
  return (int)hc;
}

}  // namespace

extern "C" {

uint64_t cpimw_splitmix64(uint64_t x) {
  uint64_t t = x;
  return splitmix64(t);
}

uint64_t cpimw_instance_seed(uint64_t base, uint64_t config_id, uint64_t index) {
  uint64_t t = base ^ config_id ^ index;
  return splitmix64(t);
}

// Choose e distinct unordered pairs (x < y) uniformly without replacement
// among the n(n-1)/2 pairs (selection sampling, Knuth TAOCP 3.4.2 Algorithm S),
// emitted in lexicographic order.  Returns 0, or -1 if e is out of range.
int cpimw_random_pairs(uint64_t seed, int n, int e, int32_t* scopes) {
  const int64_t total = (int64_t)n * (n - 1) / 2;
  if (e < 0 || e > total) return -1;
  Xoshiro256 rng(sub_seed(seed, 0xA11));
  int64_t seen = 0, chosen = 0;
  for (int x = 0; x < n && chosen < e; ++x) {
    for (int y = x + 1; y < n && chosen < e; ++y) {
      const uint64_t remaining = (uint64_t)(total - seen);
      const uint64_t need = (uint64_t)(e - chosen);
      if (rng.below(remaining) < need) {
        scopes[2 * chosen] = x;
        scopes[2 * chosen + 1] = y;
        ++chosen;
      }
      ++seen;
    }
  }
  return chosen == e ? 0 : -1;
}

// Dense random relations: every (a, b) in d0 x d1 is forbidden independently
// with probability p_forbid.  Constraint c uses its own xoshiro stream, so the
// result does not depend on the thread count.
int cpimw_random_relations(uint64_t seed, int e, int d0, int d1, int words,
                           double p_forbid, uint64_t* rel) {
  if (e < 0 || d0 <= 0 || d1 <= 0 || words * 64 < d1) return -1;
  const uint64_t thr = threshold_of(p_forbid);
  const int64_t row_words = words;
  const int64_t con_words = (int64_t)d0 * row_words;
  auto work = [&](int c0, int c1) {
    for (int c = c0; c < c1; ++c) {
      Xoshiro256 rng(sub_seed(seed, 0x100000ull + (uint64_t)c));
      uint64_t* base = rel + (int64_t)c * con_words;
      std::memset(base, 0, sizeof(uint64_t) * con_words);
      for (int a = 0; a < d0; ++a) {
        uint64_t* row = base + (int64_t)a * row_words;
        for (int b = 0; b < d1; ++b) {
          const bool forbidden = rng.next() < thr;
          if (!forbidden) row[b >> 6] |= 1ull << (b & 63);
        }
      }
    }
  };
  const int nt = std::max(1, std::min(n_threads_for((int64_t)e * d0 * d1), e));
  if (nt == 1) {
    work(0, e);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
      const int c0 = (int)((int64_t)e * t / nt), c1 = (int)((int64_t)e * (t + 1) / nt);
      th.emplace_back(work, c0, c1);
    }
    for (auto& t : th) t.join();
  }
  return 0;
}

// Domain words for n variables of size d each, with every value removed
// independently with probability frac (seeded).  dom is [n][words].
int cpimw_domains_with_removals(uint64_t seed, int n, int d, int words, double frac,
                                uint64_t* dom) {
  if (words * 64 < d) return -1;
  const uint64_t thr = threshold_of(frac);
  Xoshiro256 rng(sub_seed(seed, 0xD0D0));
  for (int v = 0; v < n; ++v) {
    uint64_t* row = dom + (int64_t)v * words;
    for (int w = 0; w < words; ++w) row[w] = 0;
    for (int a = 0; a < d; ++a) {
      const bool removed = rng.next() < thr;
      if (!removed) row[a >> 6] |= 1ull << (a & 63);
    }
  }
  return 0;
}

// ntuples distinct random tuples of the given arity over [0, d)^arity,
// sorted lexicographically.  Returns -1 if ntuples exceeds d^arity.
int cpimw_random_table(uint64_t seed, int arity, int d, int ntuples, int32_t* tuples) {
  double space = std::pow((double)d, arity);
  if (arity <= 0 || ntuples < 0 || (double)ntuples > space || space > 1e9) return -1;
  const int64_t total = (int64_t)space;
  Xoshiro256 rng(sub_seed(seed, 0x7AB1E));
  // Floyd's sampling of ntuples distinct codes in [0, total).
  std::vector<int64_t> codes;
  codes.reserve(ntuples);
  std::vector<char> taken((size_t)total, 0);
  for (int64_t j = total - ntuples; j < total; ++j) {
    int64_t t = (int64_t)rng.below((uint64_t)(j + 1));
    if (taken[t]) t = j;
    taken[t] = 1;
    codes.push_back(t);
  }
  std::sort(codes.begin(), codes.end());
  for (int i = 0; i < ntuples; ++i) {
    int64_t code = codes[i];
    for (int k = arity - 1; k >= 0; --k) {
      tuples[(int64_t)i * arity + k] = (int32_t)(code % d);
      code /= d;
    }
  }
  return 0;
}

// Uniform integers in [lo, hi] and doubles in [lo, hi), for instance-level
// parameters (sizes, densities) of the uneven batches.
int64_t cpimw_uniform_int(uint64_t seed, uint64_t k, int64_t lo, int64_t hi) {
  Xoshiro256 rng(sub_seed(seed, 0x5000 + k));
  return lo + (int64_t)rng.below((uint64_t)(hi - lo + 1));
}

double cpimw_uniform_real(uint64_t seed, uint64_t k, double lo, double hi) {
  Xoshiro256 rng(sub_seed(seed, 0x6000 + k));
  const double u = (double)(rng.next() >> 11) * (1.0 / 9007199254740992.0);
  return lo + (hi - lo) * u;
}

}  // extern "C"
