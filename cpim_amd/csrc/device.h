// Device-side data layout and wave/workgroup propagation primitives (gfx950).
//
// Layout in HBM (all offsets in 64-bit words):
//   dom  [level][n][W]          domain bitsets, LSB first (value v -> word v/64, bit v%64)
//   size [level][n]             exact popcount of every domain, kept by atomics
//   sup  [c][dir][w][dpad]      binary support planes, word-major: element (w, a) of
//                               plane dir is the word w of the support row of value a
//                               of scope[dir] (bits over scope[1-dir]'s domain).
//                               dpad = 64 W.  Word-major rows make the per-value check
//                               of a revision a coalesced 8-byte-per-lane load.
//   tsup [c][pos][a][tw]        table constraints: bitset over tuple indices of the
//                               tuples whose pos-th value is a.
//   adj  [adj_off[v]..adj_off[v+1])  CSR of the constraints of v in constraint order:
//                               {c, other var (binary) or -1 (table), pos of v,
//                                binary: plane pair index p (planes at sup + p*2*W*dpad),
//                                table: -1}
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;

#define CPIM_MAX_ARITY 8
#define CPIM_MAX_WORDS 16   // domains up to 1024 values

struct ConMeta {
  int kind;  // 0 = binary, 1 = table
  int arity;
  int scope[CPIM_MAX_ARITY];
  int tw;    // table: 64-bit words per tuple bitset
  int pad;
  long long off;  // binary: offset of plane 0 in sup; table: offset in tsup
};

struct DevNet {
  int n;
  int W;
  int dpad;
  int ncons;
  const int* adj_off;
  const int4* adj;
  const ConMeta* con;
  const u64* sup;
  const u64* tsup;
};

// ---------------------------------------------------------------------------
// Wave helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ u64 readlane64(u64 v, int lane) {
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)(v & 0xffffffffull), lane);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
  return ((u64)hi << 32) | lo;
}

// Wave-uniform copies of values read from LDS.  Every value that decides a
// branch or loop exit around a __syncthreads() goes through these, so the
// compiler emits scalar branches: a value it believes divergent turns such a
// loop into an exec-mask loop, which can keep some lanes in the loop while the
// lane that must publish the next work item waits outside it.
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ long long uniform_ll(long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v & 0xffffffffull));
  const unsigned hi = __builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ int mbcnt(u64 mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Row operations over DPP (data-parallel primitives: a VALU operand read
// from another lane of the same 16-lane row), for waves with every lane
// active.  row_shr:n = 0x110 + n (lanes below n of a row read 0 here),
// row_ror:n = 0x120 + n.  Each step costs a few cycles where a shuffle
// through ds_bpermute costs an LDS round trip.
template <int kCtrl>
__device__ __forceinline__ unsigned dpp_u32(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, 0xf, 0xf, true);
}

// OR of v over all 64 lanes, returned to every lane (wave-uniform).
__device__ __forceinline__ u64 wave_or64_dpp(u64 v) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo |= dpp_u32<0x121>(lo);
  hi |= dpp_u32<0x121>(hi);
  lo |= dpp_u32<0x122>(lo);
  hi |= dpp_u32<0x122>(hi);
  lo |= dpp_u32<0x124>(lo);
  hi |= dpp_u32<0x124>(hi);
  lo |= dpp_u32<0x128>(lo);
  hi |= dpp_u32<0x128>(hi);
  // every lane of a row now holds the row's OR
  unsigned rlo = 0, rhi = 0;
#pragma unroll
  for (int r = 0; r < 64; r += 16) {
    rlo |= (unsigned)__builtin_amdgcn_readlane((int)lo, r);
    rhi |= (unsigned)__builtin_amdgcn_readlane((int)hi, r);
  }
  return ((u64)rhi << 32) | rlo;
}

// Word-wise OR for W-word rows spread over the wave: lane l holds word
// l % Wp (Wp = W rounded up to a power of two, at most 16).  Returns, in
// lane w < W, the OR of word w over all lanes, and 0 in the other lanes.
// Rotations by Wp, 2 Wp, ..., 8 within a row combine the lanes of one word;
// then lane w reads its word from each of the four rows.
__device__ __forceinline__ u64 wave_or_words_dpp(u64 v, int Wp, int W) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  if (Wp <= 1) {
    lo |= dpp_u32<0x121>(lo);
    hi |= dpp_u32<0x121>(hi);
  }
  if (Wp <= 2) {
    lo |= dpp_u32<0x122>(lo);
    hi |= dpp_u32<0x122>(hi);
  }
  if (Wp <= 4) {
    lo |= dpp_u32<0x124>(lo);
    hi |= dpp_u32<0x124>(hi);
  }
  if (Wp <= 8) {
    lo |= dpp_u32<0x128>(lo);
    hi |= dpp_u32<0x128>(hi);
  }
  const u64 x = ((u64)hi << 32) | lo;
  const int lane = lane_id();
  u64 out = 0ull;
  for (int w = 0; w < W; ++w) {
    const u64 r = readlane64(x, w) | readlane64(x, 16 + w) | readlane64(x, 32 + w) | readlane64(x, 48 + w);
    if (lane == w) out = r;
  }
  return out;
}

// Minimum and sum of v over all 64 lanes, returned to every lane
// (wave-uniform).  Rotations by 1, 2, 4 and 8 within a row cover each row
// exactly once, so every lane ends with its row's result.
__device__ __forceinline__ u64 dpp_ror_u64_1(u64 v) {
  return ((u64)dpp_u32<0x121>((unsigned)(v >> 32)) << 32) | dpp_u32<0x121>((unsigned)v);
}
__device__ __forceinline__ u64 dpp_ror_u64_2(u64 v) {
  return ((u64)dpp_u32<0x122>((unsigned)(v >> 32)) << 32) | dpp_u32<0x122>((unsigned)v);
}
__device__ __forceinline__ u64 dpp_ror_u64_4(u64 v) {
  return ((u64)dpp_u32<0x124>((unsigned)(v >> 32)) << 32) | dpp_u32<0x124>((unsigned)v);
}
__device__ __forceinline__ u64 dpp_ror_u64_8(u64 v) {
  return ((u64)dpp_u32<0x128>((unsigned)(v >> 32)) << 32) | dpp_u32<0x128>((unsigned)v);
}

__device__ __forceinline__ u64 readlane_rows_min(u64 v) {
  u64 r = readlane64(v, 0);
#pragma unroll
  for (int l = 16; l < 64; l += 16) {
    const u64 o = readlane64(v, l);
    r = o < r ? o : r;
  }
  return r;
}

__device__ __forceinline__ u64 wave_min64_dpp(u64 v) {
  u64 o = dpp_ror_u64_1(v);
  v = o < v ? o : v;
  o = dpp_ror_u64_2(v);
  v = o < v ? o : v;
  o = dpp_ror_u64_4(v);
  v = o < v ? o : v;
  o = dpp_ror_u64_8(v);
  v = o < v ? o : v;
  return readlane_rows_min(v);
}

__device__ __forceinline__ u64 wave_sum64_dpp(u64 v) {
  v += dpp_ror_u64_1(v);
  v += dpp_ror_u64_2(v);
  v += dpp_ror_u64_4(v);
  v += dpp_ror_u64_8(v);
  return readlane64(v, 0) + readlane64(v, 16) + readlane64(v, 32) + readlane64(v, 48);
}

// Inclusive prefix sum over the 64 lanes; *total receives the sum of all
// lanes (wave-uniform).
__device__ __forceinline__ int wave_scan_dpp(int v, int* total) {
  unsigned x = (unsigned)v;
  x += dpp_u32<0x111>(x);
  x += dpp_u32<0x112>(x);
  x += dpp_u32<0x114>(x);
  x += dpp_u32<0x118>(x);
  // lane 16 r + 15 holds the sum of row r
  const unsigned s0 = (unsigned)__builtin_amdgcn_readlane((int)x, 15);
  const unsigned s1 = (unsigned)__builtin_amdgcn_readlane((int)x, 31);
  const unsigned s2 = (unsigned)__builtin_amdgcn_readlane((int)x, 47);
  const unsigned s3 = (unsigned)__builtin_amdgcn_readlane((int)x, 63);
  const int row = lane_id() >> 4;
  x += (row >= 1 ? s0 : 0u) + (row >= 2 ? s1 : 0u) + (row >= 3 ? s2 : 0u);
  *total = (int)(s0 + s1 + s2 + s3);
  return (int)x;
}

// Domain word load.  kCoherent: agent-scope atomic load, which observes every
// update other workgroups (on any XCD) have completed; used by the
// asynchronous fixpoint, which has no barrier between rounds.
template <bool kCoherent>
__device__ __forceinline__ u64 ld_dom(const u64* p);

// Loads and read-modify-writes of a global buffer that only one workgroup
// touches (the bit-sliced test state): workgroup scope is enough, and the
// __syncthreads() barriers between phases order these accesses.  An
// agent-scope fence (__threadfence) would write back and invalidate the whole
// L2 of the XCD instead; an agent-scope load bypasses it.
__device__ __forceinline__ u64 ld_wg(const u64* p) {
  return __hip_atomic_load(const_cast<u64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ u64 or_wg(u64* p, u64 v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ u64 and_wg(u64* p, u64 v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool kCoherent>
__device__ __forceinline__ u64 ld_dom(const u64* p) {
  if constexpr (kCoherent) {
    return __hip_atomic_load(const_cast<u64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    return *p;
  }
}

__device__ __forceinline__ u64 mix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// Arc revision u <- t of a binary constraint, executed by one wave.
//
// Keeps value b of u iff some value of t supports it.  Two ways to get the
// new D(u), chosen by |D(t)|:
//  * source side, when |D(t)| * Wp <= 64 (Wp = W rounded up to a power of
//    two): lane (k, w) loads word w of the row of the k-th live value of t,
//    and an OR across the lanes of each word gives OR_a row_t[a] -- one
//    round of independent loads, however the values of t are spread;
//  * target side otherwise: lanes are 64 values of u; the word of D(t)'s
//    first non-empty word is loaded for every live chunk of u before any is
//    tested, and a value whose support is not in that word scans the later
//    words (rare when D(t) is dense).
// The result mask is applied with atomic AND, the exact number of values
// this wave removed is subtracted from size[u], and the wave that brings
// size[u] to 0 reports the wipe-out.  Concurrent revisions of the same
// domain are safe: domains only shrink, a stale (larger) D(t) can only keep
// too much, and every removal re-enqueues the constraints of u.
// Returns the number of values removed (wave-uniform); *wiped is set when
// this wave emptied D(u).  `off` is the constraint's offset in sup.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pow2_at_least(int w) {
  int p = 1;
  while (p < w) p <<= 1;
  return p;
}

template <bool kCoherent = false>
__device__ __forceinline__ int revise_binary_at(const DevNet& net, u64* dom, int* size, long long off, int u,
                                                int t, int pos_t, bool* wiped) {
  const int W = net.W;
  const int dpad = net.dpad;
  const int lane = lane_id();
  const u64* base = net.sup + off;
  const u64* plane_u = base + (long long)(1 - pos_t) * W * dpad;
  const u64* plane_t = base + (long long)pos_t * W * dpad;
  const u64 dt_l = lane < W ? ld_dom<kCoherent>(&dom[(long long)t * W + lane]) : 0ull;
  const u64 du_l = lane < W ? ld_dom<kCoherent>(&dom[(long long)u * W + lane]) : 0ull;
  int cnt_t = 0;
  int first_w = -1;
  for (int w = 0; w < W; ++w) {
    const u64 x = readlane64(dt_l, w);
    cnt_t += __popcll(x);
    if (first_w < 0 && x) first_w = w;
  }
  const int Wp = pow2_at_least(W);
  u64 mask_l = 0ull;
  if (cnt_t * Wp <= 64) {
    const int w = lane & (Wp - 1);
    int k = lane / Wp;
    u64 row = 0ull;
    if (w < W && k < cnt_t) {
      int a = -1;
      for (int j = 0; j < W && a < 0; ++j) {
        u64 x = readlane64(dt_l, j);
        const int c = __popcll(x);
        if (k < c) {
          for (int i = 0; i < k; ++i) x &= x - 1;
          a = j * 64 + (int)__builtin_ctzll(x);
        } else {
          k -= c;
        }
      }
      row = plane_t[(long long)w * dpad + a];
    }
    mask_l = wave_or_words_dpp(row, Wp, W);
  } else {
    const u64 dt0 = readlane64(dt_l, first_w);
    const u64* plane_u0 = plane_u + (long long)first_w * dpad;
    for (int j0 = 0; j0 < W; j0 += 4) {
      u64 row[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = j0 + q;
        const bool live = j < W && readlane64(du_l, j < W ? j : 0) != 0ull;
        row[q] = live ? plane_u0[j * 64 + lane] : 0ull;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = j0 + q;
        if (j >= W) break;
        const u64 duj = readlane64(du_l, j);
        if (duj == 0ull) continue;
        const int b = j * 64 + lane;
        bool keep = false;
        if ((duj >> lane) & 1ull) {
          keep = (row[q] & dt0) != 0ull;
          for (int w = first_w + 1; w < W && !keep; ++w) {
            const u64 dtw = readlane64(dt_l, w);
            if (dtw && (plane_u[(long long)w * dpad + b] & dtw)) keep = true;
          }
        }
        const u64 mj = __ballot(keep);
        if (lane == j) mask_l = mj;
      }
    }
  }
  int removed_l = 0;
  if (lane < W && (du_l & ~mask_l)) {
    const u64 old = atomicAnd(&dom[(long long)u * W + lane], mask_l);
    removed_l = __popcll(old & ~mask_l);
  }
  int removed = 0;
  for (int w = 0; w < W; ++w) removed += __builtin_amdgcn_readlane(removed_l, w);
  bool wipe = false;
  if (removed > 0) {
    int old_size = 0;
    if (lane == 0) old_size = atomicSub(&size[u], removed);
    old_size = __builtin_amdgcn_readfirstlane(old_size);
    wipe = old_size == removed;
  }
  *wiped = wipe;
  return removed;
}

// Index of the k-th set bit of the W-word domain held by lanes 0..W-1 of
// d_l (k < popcount); lanes call it with their own k.
__device__ __forceinline__ int kth_value(u64 d_l, int W, int k) {
  int a = -1;
  for (int j = 0; j < W && a < 0; ++j) {
    u64 x = readlane64(d_l, j);
    const int c = __popcll(x);
    if (k < c) {
      for (int i = 0; i < k; ++i) x &= x - 1;
      a = j * 64 + (int)__builtin_ctzll(x);
    } else {
      k -= c;
    }
  }
  return a;
}

// P arcs u[p] <- t with the same source t, revised together by one wave:
// the same result as P revise_binary_at calls in a row, with the row loads
// of all P arcs in flight at once (a wave's revisions are bound by the
// latency of those loads).  Targets may repeat (parallel constraints): the
// masks are then computed from the same D(u) and applied one after the
// other, which removes exactly the values either revision removes.
template <int P>
__device__ __forceinline__ void revise_binary_multi(const DevNet& net, u64* dom, int* size, int t,
                                                    const long long (&off)[P], const int (&u)[P],
                                                    const int (&pos_t)[P], int (&removed)[P],
                                                    bool (&wiped)[P]) {
  const int W = net.W;
  const int dpad = net.dpad;
  const int lane = lane_id();
  const u64 dt_l = lane < W ? ld_dom<false>(&dom[(long long)t * W + lane]) : 0ull;
  u64 du_l[P];
#pragma unroll
  for (int p = 0; p < P; ++p) du_l[p] = lane < W ? ld_dom<false>(&dom[(long long)u[p] * W + lane]) : 0ull;
  int cnt_t = 0;
  int first_w = -1;
  for (int w = 0; w < W; ++w) {
    const u64 x = readlane64(dt_l, w);
    cnt_t += __popcll(x);
    if (first_w < 0 && x) first_w = w;
  }
  const int Wp = pow2_at_least(W);
  u64 mask_l[P];
#pragma unroll
  for (int p = 0; p < P; ++p) mask_l[p] = 0ull;
  if (cnt_t * Wp <= 64) {
    // source side: lane (k, w) loads word w of the rows of t's k-th value
    const int w = lane & (Wp - 1);
    const int k = lane / Wp;
    u64 row[P];
#pragma unroll
    for (int p = 0; p < P; ++p) row[p] = 0ull;
    if (w < W && k < cnt_t) {
      const int a = kth_value(dt_l, W, k);
#pragma unroll
      for (int p = 0; p < P; ++p)
        row[p] = net.sup[off[p] + (long long)pos_t[p] * W * dpad + (long long)w * dpad + a];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) mask_l[p] = wave_or_words_dpp(row[p], Wp, W);
  } else {
    // target side: lanes are 64 values of u; the word of D(t)'s first
    // non-empty word is loaded for every arc and live chunk before any test
    const u64 dt0 = readlane64(dt_l, first_w);
    for (int j0 = 0; j0 < W; j0 += 4) {
      u64 row[P][4];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const u64* plane_u0 = net.sup + off[p] + (long long)(1 - pos_t[p]) * W * dpad + (long long)first_w * dpad;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = j0 + q;
          const bool live = j < W && readlane64(du_l[p], j < W ? j : 0) != 0ull;
          row[p][q] = live ? plane_u0[j * 64 + lane] : 0ull;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = j0 + q;
        if (j >= W) break;
        const int b = j * 64 + lane;
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const u64 d = readlane64(du_l[p], j);
          if (d == 0ull) continue;
          const u64* plane_u = net.sup + off[p] + (long long)(1 - pos_t[p]) * W * dpad;
          bool keep = false;
          if ((d >> lane) & 1ull) {
            keep = (row[p][q] & dt0) != 0ull;
            for (int w = first_w + 1; w < W && !keep; ++w) {
              const u64 dtw = readlane64(dt_l, w);
              if (dtw && (plane_u[(long long)w * dpad + b] & dtw)) keep = true;
            }
          }
          const u64 m = __ballot(keep);
          if (lane == j) mask_l[p] = m;
        }
      }
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
    int removed_l = 0;
    if (lane < W && (du_l[p] & ~mask_l[p])) {
      const u64 old = atomicAnd(&dom[(long long)u[p] * W + lane], mask_l[p]);
      removed_l = __popcll(old & ~mask_l[p]);
    }
    int r = 0;
    for (int w = 0; w < W; ++w) r += __builtin_amdgcn_readlane(removed_l, w);
    bool wipe = false;
    if (r > 0) {
      int old_size = 0;
      if (lane == 0) old_size = atomicSub(&size[u[p]], r);
      old_size = __builtin_amdgcn_readfirstlane(old_size);
      wipe = old_size == r;
    }
    removed[p] = r;
    wiped[p] = wipe;
  }
}

// Offset in sup of the planes of a binary arc's constraint.
__device__ __forceinline__ long long arc_plane_off(const DevNet& net, const int4& e) {
  return (long long)e.w * 2 * net.W * net.dpad;
}

template <bool kCoherent = false>
__device__ __forceinline__ int revise_binary(const DevNet& net, u64* dom, int* size, int c, int u,
                                             int t, int pos_t, bool* wiped) {
  return revise_binary_at<kCoherent>(net, dom, size, net.con[c].off, u, t, pos_t, wiped);
}

// ---------------------------------------------------------------------------
// Table constraint revision (positive table, bitset per (pos, value) over the
// tuple indices, tw words per row).  valid = AND over positions of OR over
// live values of tsup[pos][value]; value b of the variable at position p
// keeps iff tsup[p][b] AND valid is not empty.  Positions whose variable is
// `skip_var`, or is assigned (asg bitmap, may be NULL), are not revised; with
// only_var >= 0, only the position of that variable is.  *wiped is set if
// this wave emptied a domain.  on_removed(var, count) is called (by the whole
// wave) for every variable that lost values.  Any number of tuples and up to
// CPIM_MAX_ARITY positions.
//
// One-word domains with at most 512 tuples take a path of their own that
// revises every position at once (below).  Otherwise the tuple words are
// taken in chunks of 64 words (4096 tuples).  In a chunk,
// the OR over live values runs with lanes = (value k, word w) pairs: one round
// of independent row loads per 64 / Tp values (Tp = chunk words rounded up to
// a power of two), then an OR across the lanes of each word.  The per-value
// check loads the row four words at a time.  A table of one chunk computes
// valid once for all positions.  A larger one is scanned per revised
// position: a value supported in one chunk is not checked again, a chunk
// with no valid tuple is skipped, and the scan stops once every value of the
// position has a support.
// ---------------------------------------------------------------------------

template <typename OnRemoved, bool kCoherent = false>
__device__ __forceinline__ void revise_table(const DevNet& net, u64* dom, int* size, int c,
                                             int skip_var, const unsigned* asg,
                                             OnRemoved on_removed, bool* wiped, int only_var = -1) {
  const int W = net.W;
  const int dpad = net.dpad;
  const int lane = lane_id();
  const ConMeta* mp = net.con + c;
  const int arity = mp->arity;
  const u64* ts = net.tsup + mp->off;
  const int tw = mp->tw;
  auto load_dom = [&](int v) -> u64 {
    return lane < W ? ld_dom<kCoherent>(&dom[(long long)v * W + lane]) : 0ull;
  };
  auto revised = [&](int v) -> bool {
    return v != skip_var && (only_var < 0 || v == only_var) && !(asg && ((asg[v >> 5] >> (v & 31)) & 1u));
  };
  // valid tuples of words [c0, c0 + cw), word `lane` (lanes < cw)
  auto chunk_valid = [&](int c0, int cw) -> u64 {
    const int Tp = pow2_at_least(cw);
    const int per = 64 / Tp;  // values per round of loads
    const int w = lane & (Tp - 1);
    u64 valid_l = ~0ull;
    for (int q = 0; q < arity; ++q) {
      const u64 dq_l = load_dom(mp->scope[q]);
      int cnt = 0;
      for (int j = 0; j < W; ++j) cnt += __popcll(readlane64(dq_l, j));
      u64 acc = 0ull;
      for (int k0 = 0; k0 < cnt; k0 += per) {
        const int k = k0 + lane / Tp;
        if (w < cw && k < cnt) {
          const int a = kth_value(dq_l, W, k);
          acc |= ts[((long long)q * dpad + a) * tw + c0 + w];
        }
      }
      for (int o = Tp; o < 64; o <<= 1) acc |= __shfl_xor(acc, o, 64);
      valid_l &= acc;
    }
    return lane < cw ? valid_l : 0ull;
  };
  // values of position p (domain du_l) not in done_l that have a valid tuple
  // in words [c0, c0 + cw); lane j returns word j of the mask
  auto supported = [&](int p, u64 du_l, u64 done_l, u64 valid_l, int c0, int cw) -> u64 {
    u64 mask_l = 0ull;
    for (int j = 0; j < W; ++j) {
      const u64 open = readlane64(du_l, j) & ~readlane64(done_l, j);
      if (open == 0ull) continue;
      const int b = j * 64 + lane;
      bool keep = false;
      if ((open >> lane) & 1ull) {
        const u64* row = ts + ((long long)p * dpad + b) * tw + c0;
        for (int w0 = 0; w0 < cw && !keep; w0 += 4) {
          u64 r[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) r[q] = w0 + q < cw ? row[w0 + q] : 0ull;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (w0 + q < cw && (r[q] & readlane64(valid_l, w0 + q < cw ? w0 + q : 0))) keep = true;
        }
      }
      const u64 mj = __ballot(keep);
      if (lane == j) mask_l = mj;
    }
    return mask_l;
  };
  bool any_wipe = false;
  auto apply = [&](int y, u64 du_l, u64 keep_l) {
    int removed_l = 0;
    if (lane < W && (du_l & ~keep_l)) {
      const u64 old = atomicAnd(&dom[(long long)y * W + lane], keep_l);
      removed_l = __popcll(old & ~keep_l);
    }
    int removed = 0;
    for (int j = 0; j < W; ++j) removed += __builtin_amdgcn_readlane(removed_l, j);
    if (removed > 0) {
      int old_size = 0;
      if (lane == 0) old_size = atomicSub(&size[y], removed);
      old_size = __builtin_amdgcn_readfirstlane(old_size);
      if (old_size == removed) any_wipe = true;
      on_removed(y, removed);
    }
  };
  if (W == 1 && tw <= 8) {
    // One-word domains, at most 512 tuples: every position at once.  Lane
    // (q, w) = (lane / Tp, lane % Tp) ORs word w of the rows of position q's
    // live values, four loads in flight per step; the AND over positions
    // leaves valid word w in lane w.  Lane q < arity holds position q's
    // variable and domain.  Each revised position then tests its values
    // (lane = value) against the valid words, one row of tw words per lane.
    const int Tp = pow2_at_least(tw);
    const int q = lane / Tp, w = lane & (Tp - 1);
    const bool mine = q < arity && w < tw;
    const int gv = mine ? mp->scope[q] : 0;
    const int ov = lane < arity ? mp->scope[lane] : 0;
    u64 x = mine ? ld_dom<kCoherent>(&dom[gv]) : 0ull;
    const u64 own = lane < arity ? ld_dom<kCoherent>(&dom[ov]) : 0ull;
    const u64* tq = ts + (long long)q * dpad * tw + w;
    u64 acc = 0ull;
    while (__ballot(x != 0ull)) {
      int a[4];
      bool v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = x != 0ull;
        a[i] = v[i] ? (int)__builtin_ctzll(x) : 0;
        x &= x - 1;
      }
      u64 r[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = v[i] ? tq[(long long)a[i] * tw] : 0ull;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc |= r[i];
    }
    u64 valid = acc;
    for (int g = 1; g < arity; ++g) valid &= __shfl(acc, g * Tp + w, 64);
    u64 vw[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) vw[i] = i < tw ? readlane64(valid, i) : 0ull;
    for (int p = 0; p < arity; ++p) {
      const int y = __builtin_amdgcn_readlane(ov, p);
      if (!revised(y)) continue;
      const u64 dp = readlane64(own, p);
      bool keep = false;
      if ((dp >> lane) & 1ull) {
        const u64* row = ts + ((long long)p * dpad + lane) * tw;
        u64 r[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = i < tw ? row[i] : 0ull;
#pragma unroll
        for (int i = 0; i < 8; ++i) keep = keep || (r[i] & vw[i]) != 0ull;
      }
      const u64 km = __ballot(keep);
      if (dp & ~km) apply(y, lane == 0 ? dp : 0ull, lane == 0 ? km : 0ull);
    }
  } else if (tw <= 64) {
    // one chunk: valid once, then every revised position
    const u64 valid_l = chunk_valid(0, tw);
    for (int p = 0; p < arity; ++p) {
      const int y = mp->scope[p];
      if (!revised(y)) continue;
      const u64 du_l = load_dom(y);
      apply(y, du_l, supported(p, du_l, 0ull, valid_l, 0, tw));
    }
  } else {
    // chunks of 4096 tuples per revised position, until every value of the
    // position has a support.  A removal leaves the valid tuples unchanged
    // (no valid tuple goes through a value without support), so the later
    // positions may read the reduced domain.
    for (int p = 0; p < arity; ++p) {
      const int y = mp->scope[p];
      if (!revised(y)) continue;
      const u64 du_l = load_dom(y);
      u64 keep_l = 0ull;
      for (int c0 = 0; c0 < tw; c0 += 64) {
        const int cw = min(64, tw - c0);
        const u64 valid_l = chunk_valid(c0, cw);
        if (__ballot(valid_l != 0ull) == 0ull) continue;
        keep_l |= supported(p, du_l, keep_l, valid_l, c0, cw);
        if (__ballot(lane < W && (du_l & ~keep_l) != 0ull) == 0ull) break;
      }
      apply(y, du_l, keep_l);
    }
  }
  *wiped = any_wipe;
}
