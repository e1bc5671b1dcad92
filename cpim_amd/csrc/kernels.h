// Propagation kernels: model build, grid-wide constraint-frontier fixpoint
// (plain and persistent), and the workgroup engine (domains in LDS) used by
// block-mode GAC, batched singleton tests, the device MAC, batches of
// instances and the serving step.
#pragma once

#include <hip/hip_cooperative_groups.h>

#include "device.h"

namespace cg = cooperative_groups;

// ===========================================================================
// Model build: dense rows -> word-major support planes (both directions).
// One wave per (binary constraint k, word w): lanes are the 64 values
// a = 64 w + lane of scope[0]; plane 0 is a transpose of word indices, plane 1
// a bit transpose done with one ballot per value b of scope[1].
// ===========================================================================
__global__ void k_build_planes(const u64* __restrict__ rel, int nb, int max_dom, int rel_words,
                               int W, int dpad, const int* __restrict__ bin_cons,
                               const ConMeta* __restrict__ con, const int* __restrict__ cap,
                               u64* __restrict__ sup) {
  const int lane = lane_id();
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= (long long)nb * W) return;
  const int k = (int)(wave / W);
  const int w = (int)(wave % W);
  const int c = bin_cons[k];
  const ConMeta m = con[c];
  const int cx = cap[m.scope[0]], cy = cap[m.scope[1]];
  u64* plane0 = sup + m.off;
  u64* plane1 = plane0 + (long long)W * dpad;
  const u64* rows = rel + (long long)k * max_dom * rel_words;
  // plane 0: element (w, a) = word w of row a, masked to y's capacity
  const int ybits = cy - 64 * w;
  const u64 ymask = ybits >= 64 ? ~0ull : (ybits <= 0 ? 0ull : ((1ull << ybits) - 1ull));
  for (int a0 = 0; a0 < dpad; a0 += 64) {
    const int a = a0 + lane;
    u64 v = 0ull;
    if (a < cx && w < rel_words) v = rows[(long long)a * rel_words + w] & ymask;
    plane0[(long long)w * dpad + a] = v;
  }
  // plane 1: element (w, b) = bits a in [64w, 64w+64) with (a, b) allowed
  const int a = 64 * w + lane;
  u64 acc = 0ull;
  for (int j = 0; j < W; ++j) {
    const u64 rowword = (a < cx && j < rel_words) ? rows[(long long)a * rel_words + j] : 0ull;
    for (int bb = 0; bb < 64; ++bb) {
      const int b = 64 * j + bb;
      const bool bit = b < cy && ((rowword >> bb) & 1ull);
      const u64 word = __ballot(bit);
      if (lane == bb) acc = word;
    }
    plane1[(long long)w * dpad + 64 * j + lane] = acc;
  }
}

// ===========================================================================
// Grid engine: rounds over a bitmap frontier of constraints.
// Three frontier buffers rotate: in round r, F[r%3] is read, F[(r+1)%3]
// collects the next frontier, F[(r+2)%3] (read in round r-1) is cleared.
// Termination uses epochs that only grow (last_push_round, wipe_round), so
// every block takes the same decision after a barrier without any reset.
// ===========================================================================
struct GacCtl {
  unsigned scanner[3];
  int last_push_round;  // highest round that enqueued a constraint
  int wipe_round;       // lowest round that emptied a domain
  int rounds;           // rounds executed (written at the end)
  int converged;
  int pad;
  unsigned long long deletions;
  unsigned long long revisions;
};

__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void push_var(const DevNet& net, int v, unsigned* B, GacCtl* ctl, int r) {
  const int lane = lane_id();
  const int b0 = net.adj_off[v], b1 = net.adj_off[v + 1];
  for (int i = b0 + lane; i < b1; i += 64) {
    const int c = net.adj[i].x;
    atomicOr(&B[c >> 5], 1u << (c & 31));
  }
  if (lane == 0) atomicMax(&ctl->last_push_round, r);
}

__device__ __forceinline__ void grid_process_constraint(const DevNet& net, u64* dom, int* size,
                                                        int c, unsigned* B, GacCtl* ctl, int r,
                                                        unsigned long long& revs,
                                                        unsigned long long& dels) {
  const ConMeta* mp = net.con + c;
  bool wiped = false;
  if (mp->kind == 0) {
    const int x = mp->scope[0], y = mp->scope[1];
    int rx = revise_binary(net, dom, size, c, x, y, 1, &wiped);
    ++revs;
    if (wiped) atomicMin(&ctl->wipe_round, r);
    if (rx) {
      dels += rx;
      push_var(net, x, B, ctl, r);
    }
    int ry = revise_binary(net, dom, size, c, y, x, 0, &wiped);
    ++revs;
    if (wiped) atomicMin(&ctl->wipe_round, r);
    if (ry) {
      dels += ry;
      push_var(net, y, B, ctl, r);
    }
  } else {
    revise_table(net, dom, size, c, -1, nullptr,
                 [&](int v, int k) {
                   dels += k;
                   push_var(net, v, B, ctl, r);
                 },
                 &wiped);
    ++revs;
    if (wiped) atomicMin(&ctl->wipe_round, r);
  }
}

__device__ __forceinline__ void grid_round(const DevNet& net, u64* dom, int* size, unsigned* F,
                                           int nfw, GacCtl* ctl, int r) {
  unsigned* A = F + (long long)(r % 3) * nfw;
  unsigned* B = F + (long long)((r + 1) % 3) * nfw;
  unsigned* Cl = F + (long long)((r + 2) % 3) * nfw;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicExch(&ctl->scanner[(r + 1) % 3], 0u);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nfw;
       i += (long long)gridDim.x * blockDim.x)
    Cl[i] = 0u;
  const int lane = lane_id();
  unsigned long long revs = 0, dels = 0;
  while (true) {
    int wi = 0;
    if (lane == 0) wi = (int)atomicAdd(&ctl->scanner[r % 3], 1u);
    wi = __shfl(wi, 0, 64);
    if (wi >= nfw) break;
    if (ld_agent(&ctl->wipe_round) <= r) break;
    unsigned bits = A[wi];
    while (bits) {
      const int b = __ffs(bits) - 1;
      bits &= bits - 1u;
      grid_process_constraint(net, dom, size, wi * 32 + b, B, ctl, r, revs, dels);
    }
  }
  if (lane == 0 && (revs | dels)) {
    atomicAdd(&ctl->revisions, revs);
    atomicAdd(&ctl->deletions, dels);
  }
}

// One round per launch (host-driven fixpoint).
__global__ void k_gac_round(DevNet net, u64* dom, int* size, unsigned* F, int nfw, GacCtl* ctl,
                            int r) {
  grid_round(net, dom, size, F, nfw, ctl, r);
}

// Every round in one cooperative launch.
__global__ void k_gac_persistent(DevNet net, u64* dom, int* size, unsigned* F, int nfw,
                                 GacCtl* ctl, int max_rounds) {
  cg::grid_group grid = cg::this_grid();
  int r = 0;
  int converged = 0;
  for (; r < max_rounds; ++r) {
    grid_round(net, dom, size, F, nfw, ctl, r);
    grid.sync();
    const int wr = ld_agent(&ctl->wipe_round);
    const int lp = ld_agent(&ctl->last_push_round);
    if (wr <= r) break;
    if (lp < r) {
      converged = 1;
      break;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->rounds = r < max_rounds ? r + 1 : max_rounds;
    ctl->converged = converged;
  }
}

__global__ void k_seed_all(unsigned* F, int nfw, int ncons) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfw) return;
  const int lo = (int)(i * 32);
  const int k = ncons - lo;
  F[i] = k >= 32 ? 0xffffffffu : (k <= 0 ? 0u : ((1u << k) - 1u));
}

__global__ void k_seed_var(unsigned* F, const int* adj_off, const int4* adj, int var) {
  for (int i = adj_off[var] + threadIdx.x; i < adj_off[var + 1]; i += blockDim.x) {
    const int c = adj[i].x;
    atomicOr(&F[c >> 5], 1u << (c & 31));
  }
}

// Recompute exact sizes of one level (after host-side writes).
__global__ void k_sizes(const u64* dom, int* size, int n, int W) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int s = 0;
  for (int w = 0; w < W; ++w) s += __popcll(dom[(long long)v * W + w]);
  size[v] = s;
}

// ===========================================================================
// Workgroup engine.  One workgroup owns one domain copy (in LDS when it fits,
// else in a per-workgroup global scratch region) and runs an AC fixpoint over
// a variable frontier: every round revises all arcs u <- t for the variables
// t that changed in the previous round, spread over the waves of the
// workgroup (one arc at a time per wave, taken from an LDS counter).
// Assigned variables (asg bitmap) are never revised, as in AC3::enforce
// (src/AC3.cpp:114); the closure is the same with or without that skip.
// ===========================================================================
struct WgMem {
  u64* dom;       // [n * W]
  int* size;      // [n]
  unsigned* cur;  // [nvw]
  unsigned* nxt;  // [nvw]
  unsigned* asg;  // [nvw]
  int* list;      // [n]
  int* pre;       // [n + 1]
};

struct WgCtl {
  int arc_ctr;
  int wipe;
  int nchanged[2];
  int m;
  int total;
  int sel_var;
  int sel_val;
  unsigned long long revs;
  u64 red[16];  // one slot per wave for block reductions
  // MAC control
  int status;
  int depth;
  long long npos;
  long long nneg;
  long long nodes_here;
  int trace_len;
  int ok;
};

__host__ __device__ inline long long wg_mem_bytes(int n, int W) {
  const long long nvw = (n + 31) / 32;
  long long b = 8ll * n * W;
  b += 4ll * n;          // size
  b += 4ll * nvw * 3;    // cur, nxt, asg
  b += 4ll * n;          // list
  b += 4ll * (n + 1);    // pre
  return (b + 15) & ~15ll;
}

__device__ __forceinline__ WgMem wg_carve(char* base, int n, int W) {
  const int nvw = (n + 31) / 32;
  WgMem m;
  m.dom = (u64*)base;
  char* p = base + 8ll * n * W;
  m.size = (int*)p;
  p += 4ll * n;
  m.cur = (unsigned*)p;
  p += 4ll * nvw;
  m.nxt = (unsigned*)p;
  p += 4ll * nvw;
  m.asg = (unsigned*)p;
  p += 4ll * nvw;
  m.list = (int*)p;
  p += 4ll * n;
  m.pre = (int*)p;
  return m;
}

__device__ __forceinline__ bool bit_of(const unsigned* bm, int v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

__device__ __forceinline__ void wg_mark(WgMem& mem, WgCtl* ctl, int v, int slot) {
  const unsigned bit = 1u << (v & 31);
  const unsigned old = atomicOr(&mem.nxt[v >> 5], bit);
  if (!(old & bit)) atomicAdd(&ctl->nchanged[slot], 1);
}

// AC fixpoint from the seed in mem.cur.  both_dirs_round0: in the first round
// also revise t <- u for every arc of a seed variable t (constraint seeding,
// as GModel::EnforceGAC(assigned_var) does).  Returns false on wipe-out.
// Must be called by every thread of the workgroup.
__device__ bool wg_fixpoint(const DevNet& net, WgMem mem, WgCtl* ctl, bool both_dirs_round0) {
  const int n = net.n;
  const int nvw = (n + 31) / 32;
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int wave = tid >> 6;
  const int lane = lane_id();
  unsigned* cur = mem.cur;
  unsigned* nxt = mem.nxt;
  if (tid == 0) ctl->wipe = 0;
  unsigned long long revs = 0;
  bool result = true;
  for (int round = 0;; ++round) {
    const int slot = round & 1;
    if (wave == 0) {
      int base = 0, dbase = 0;
      for (int v0 = 0; v0 < n; v0 += 64) {
        const int v = v0 + lane;
        const bool f = v < n && bit_of(cur, v);
        const u64 mask = __ballot(f);
        const int rank = mbcnt(mask);
        const int deg = f ? net.adj_off[v + 1] - net.adj_off[v] : 0;
        const int inc = wave_inclusive_scan(deg);
        if (f) {
          mem.list[base + rank] = v;
          mem.pre[base + rank] = dbase + inc - deg;
        }
        base += __popcll(mask);
        dbase += __shfl(inc, 63, 64);
      }
      if (lane == 0) {
        mem.pre[base] = dbase;
        ctl->m = base;
        ctl->total = dbase;
        ctl->arc_ctr = 0;
        ctl->nchanged[slot] = 0;
      }
    } else {
      for (int i = tid - 64; i < nvw; i += nthreads - 64) nxt[i] = 0u;
    }
    if (nthreads == 64)
      for (int i = lane; i < nvw; i += 64) nxt[i] = 0u;
    __syncthreads();
    const int total = ctl->total;
    const int m = ctl->m;
    const bool both = both_dirs_round0 && round == 0;
    while (true) {
      int idx = 0;
      if (lane == 0) idx = atomicAdd(&ctl->arc_ctr, 1);
      idx = __shfl(idx, 0, 64);
      if (idx >= total) break;
      if (ctl->wipe) break;
      int lo = 0, hi = m - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mem.pre[mid] <= idx) lo = mid;
        else hi = mid - 1;
      }
      const int t = mem.list[lo];
      const int4 e = net.adj[net.adj_off[t] + (idx - mem.pre[lo])];
      const int c = e.x;
      bool wiped = false;
      if (net.con[c].kind == 0) {
        const int u = e.y;
        if (!bit_of(mem.asg, u)) {
          const int r = revise_binary(net, mem.dom, mem.size, c, u, t, e.z, &wiped);
          ++revs;
          if (r && lane == 0) wg_mark(mem, ctl, u, slot);
          if (wiped && lane == 0) ctl->wipe = 1;
        }
        if (both && !bit_of(mem.asg, t)) {
          const int r = revise_binary(net, mem.dom, mem.size, c, t, u, 1 - e.z, &wiped);
          ++revs;
          if (r && lane == 0) wg_mark(mem, ctl, t, slot);
          if (wiped && lane == 0) ctl->wipe = 1;
        }
      } else {
        revise_table(net, mem.dom, mem.size, c, both ? -1 : t, mem.asg,
                     [&](int v, int) {
                       if (lane == 0) wg_mark(mem, ctl, v, slot);
                     },
                     &wiped);
        ++revs;
        if (wiped && lane == 0) ctl->wipe = 1;
      }
    }
    __syncthreads();
    if (ctl->wipe) {
      result = false;
      break;
    }
    if (ctl->nchanged[slot] == 0) break;
    unsigned* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (lane == 0 && revs) atomicAdd(&ctl->revs, revs);
  __syncthreads();
  return result;
}

__device__ __forceinline__ void wg_clear_bitmap(unsigned* bm, int nvw) {
  for (int i = threadIdx.x; i < nvw; i += blockDim.x) bm[i] = 0u;
}

__device__ __forceinline__ void wg_fill_bitmap(unsigned* bm, int n) {
  const int nvw = (n + 31) / 32;
  for (int i = threadIdx.x; i < nvw; i += blockDim.x) {
    const int k = n - 32 * i;
    bm[i] = k >= 32 ? 0xffffffffu : ((1u << k) - 1u);
  }
}

__device__ __forceinline__ void wg_recount(WgMem mem, int n, int W) {
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    int s = 0;
    for (int w = 0; w < W; ++w) s += __popcll(mem.dom[(long long)v * W + w]);
    mem.size[v] = s;
  }
}

__device__ __forceinline__ void wg_copy_words(u64* dst, const u64* src, long long count) {
  for (long long i = threadIdx.x; i < count; i += blockDim.x) dst[i] = src[i];
}

// Block-wide reduction helpers (all threads call; result valid in every thread).
__device__ __forceinline__ u64 wg_reduce_min_u64(WgCtl* ctl, u64 v) {
  for (int o = 32; o > 0; o >>= 1) {
    const u64 other = __shfl_xor(v, o, 64);
    v = other < v ? other : v;
  }
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  if (lane_id() == 0) ctl->red[wave] = v;
  __syncthreads();
  u64 r = ctl->red[0];
  for (int i = 1; i < nwaves; ++i) r = ctl->red[i] < r ? ctl->red[i] : r;
  __syncthreads();
  return r;
}

__device__ __forceinline__ u64 wg_reduce_sum_u64(WgCtl* ctl, u64 v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  if (lane_id() == 0) ctl->red[wave] = v;
  __syncthreads();
  u64 r = 0;
  for (int i = 0; i < nwaves; ++i) r += ctl->red[i];
  __syncthreads();
  return r;
}

__device__ __forceinline__ u64 wg_dom_hash(WgCtl* ctl, const u64* dom, int n, int W) {
  u64 h = 0;
  for (long long i = threadIdx.x; i < (long long)n * W; i += blockDim.x)
    h += mix64(dom[i] + (u64)i * 0x9E3779B97F4A7C15ull);
  return wg_reduce_sum_u64(ctl, h);
}

// ---------------------------------------------------------------------------
// Block-mode GAC on one level (one workgroup).
// ---------------------------------------------------------------------------
struct BlockGacOut {
  int consistent;
  int rounds;
  unsigned long long revisions;
  long long deletions;
};

template <bool kLds>
__global__ void k_block_gac(DevNet net, u64* dom, int* size, int seed_var, char* scratch,
                            BlockGacOut* out) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WgCtl ctl;
  const int n = net.n, W = net.W;
  char* base = kLds ? smem : scratch;
  WgMem mem = wg_carve(base, n, W);
  if (threadIdx.x == 0) ctl.revs = 0;
  wg_copy_words(mem.dom, dom, (long long)n * W);
  for (int v = threadIdx.x; v < n; v += blockDim.x) mem.size[v] = size[v];
  const int nvw = (n + 31) / 32;
  wg_clear_bitmap(mem.asg, nvw);
  if (seed_var < 0) {
    wg_fill_bitmap(mem.cur, n);
  } else {
    wg_clear_bitmap(mem.cur, nvw);
  }
  __syncthreads();
  if (seed_var >= 0 && threadIdx.x == 0) mem.cur[seed_var >> 5] |= 1u << (seed_var & 31);
  __syncthreads();
  long long before = 0;
  for (int v = threadIdx.x; v < n; v += blockDim.x) before += mem.size[v];
  const bool ok = wg_fixpoint(net, mem, &ctl, seed_var >= 0);
  long long after = 0;
  for (int v = threadIdx.x; v < n; v += blockDim.x) after += mem.size[v];
  const u64 delta = wg_reduce_sum_u64(&ctl, (u64)(before - after));
  wg_copy_words(dom, mem.dom, (long long)n * W);
  for (int v = threadIdx.x; v < n; v += blockDim.x) size[v] = mem.size[v];
  if (threadIdx.x == 0) {
    out->consistent = ok;
    out->revisions = ctl.revs;
    out->deletions = (long long)delta;
  }
}

// ---------------------------------------------------------------------------
// Batched singleton tests: every workgroup repeatedly takes a test (x, a)
// from a global counter, copies the base level into its domain copy, reduces
// x to {a}, marks x assigned and runs the fixpoint from {x}; fail[i] records
// a wipe-out.  The copies never touch HBM when they fit in LDS.
// ---------------------------------------------------------------------------
template <bool kLds>
__global__ void k_singleton_tests(DevNet net, const u64* __restrict__ base_dom,
                                  const int* __restrict__ base_size, const int2* __restrict__ tests,
                                  long long ntests, unsigned char* fail, unsigned long long* counter,
                                  unsigned long long* revs_out, char* scratch, long long scratch_stride) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WgCtl ctl;
  __shared__ long long unit_sh;
  const int n = net.n, W = net.W;
  const int nvw = (n + 31) / 32;
  char* base = kLds ? smem : scratch + (long long)blockIdx.x * scratch_stride;
  WgMem mem = wg_carve(base, n, W);
  if (threadIdx.x == 0) ctl.revs = 0;
  while (true) {
    if (threadIdx.x == 0) unit_sh = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const long long unit = unit_sh;
    if (unit >= ntests) break;
    const int2 tst = tests[unit];
    const int x = tst.x, a = tst.y;
    wg_copy_words(mem.dom, base_dom, (long long)n * W);
    for (int v = threadIdx.x; v < n; v += blockDim.x) mem.size[v] = base_size[v];
    wg_clear_bitmap(mem.asg, nvw);
    wg_clear_bitmap(mem.cur, nvw);
    __syncthreads();
    if (threadIdx.x < W) mem.dom[(long long)x * W + threadIdx.x] = (threadIdx.x == (a >> 6)) ? (1ull << (a & 63)) : 0ull;
    if (threadIdx.x == 0) {
      mem.size[x] = 1;
      mem.asg[x >> 5] |= 1u << (x & 31);
      mem.cur[x >> 5] |= 1u << (x & 31);
    }
    __syncthreads();
    const bool ok = wg_fixpoint(net, mem, &ctl, false);
    if (threadIdx.x == 0) fail[unit] = ok ? 0 : 1;
  }
  if (threadIdx.x == 0 && ctl.revs) atomicAdd(revs_out, ctl.revs);
}

// ---------------------------------------------------------------------------
// Device MAC (binary branching, dom-min over undecided variables with the
// lowest id on ties, min value), CPU MAC::enforce semantics
// (src/MAC.cpp:100-165).  The trail holds the domain of every depth: a
// positive decision at depth L saves the current domain in trail[L]; on
// failure the domain is restored from trail[L], the value removed, and the
// negative decision propagated at depth L.  The state can be suspended after
// `step_budget` nodes and resumed by a later launch (serving).
// ---------------------------------------------------------------------------
struct MacState {
  int status;  // 0 init, 1 running, 2 solution, 3 unsat, 4 limit reached
  int depth;
  long long npos;
  long long nneg;
  long long revisions;
  int trace_len;
  int steps;
};

struct NodeRec {
  int positive;
  int var;
  int value;
  int consistent;
  u64 dom_hash;
};

struct MacArgs {
  MacState* state;
  u64* trail;          // [n + 1][n][W]; trail[0] holds the start domains
  int2* stack;         // [n]
  int* solution;       // [n] or NULL
  NodeRec* trace;      // [trace_cap] or NULL
  int trace_cap;
  long long node_limit;   // total (0 = none)
  long long step_budget;  // per launch (0 = none)
  u64 deadline;           // wall_clock64 ticks (0 = none)
  u64* root_out;          // [n][W] root-fixpoint domains, or NULL
};

enum { MAC_INIT = 0, MAC_RUNNING = 1, MAC_SAT = 2, MAC_UNSAT = 3, MAC_LIMIT = 4 };

__device__ void wg_trace(const DevNet& net, WgMem mem, WgCtl* ctl, const MacArgs& A, int positive,
                         int var, int val, bool ok) {
  if (!A.trace) return;
  const int tl = ctl->trace_len;
  if (tl < A.trace_cap) {
    const u64 h = ok ? wg_dom_hash(ctl, mem.dom, net.n, net.W) : 0ull;
    if (threadIdx.x == 0) {
      NodeRec r;
      r.positive = positive;
      r.var = var;
      r.value = val;
      r.consistent = ok ? 1 : 0;
      r.dom_hash = h;
      A.trace[tl] = r;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) ctl->trace_len = tl + 1;
  __syncthreads();
}

__device__ void wg_mac_run(const DevNet& net, WgMem mem, WgCtl* ctl, const MacArgs& A) {
  const int n = net.n, W = net.W;
  const int nvw = (n + 31) / 32;
  const long long nw = (long long)n * W;
  if (threadIdx.x == 0) {
    ctl->status = A.state->status;
    ctl->depth = A.state->depth;
    ctl->npos = A.state->npos;
    ctl->nneg = A.state->nneg;
    ctl->trace_len = A.state->trace_len;
    ctl->nodes_here = 0;
    ctl->revs = 0;
  }
  __syncthreads();
  if (ctl->status == MAC_INIT) {
    wg_copy_words(mem.dom, A.trail, nw);
    wg_clear_bitmap(mem.asg, nvw);
    __syncthreads();
    wg_recount(mem, n, W);
    wg_fill_bitmap(mem.cur, n);
    __syncthreads();
    const bool ok = wg_fixpoint(net, mem, ctl, false);
    if (ok) wg_copy_words(A.trail, mem.dom, nw);
    if (A.root_out) wg_copy_words(A.root_out, mem.dom, nw);
    if (threadIdx.x == 0) {
      ctl->status = ok ? MAC_RUNNING : MAC_UNSAT;
      ctl->depth = 0;
    }
    __syncthreads();
  } else if (ctl->status == MAC_RUNNING) {
    const int depth = ctl->depth;
    wg_copy_words(mem.dom, A.trail + depth * nw, nw);
    wg_clear_bitmap(mem.asg, nvw);
    __syncthreads();
    wg_recount(mem, n, W);
    if (threadIdx.x == 0)
      for (int i = 0; i < depth; ++i) mem.asg[A.stack[i].x >> 5] |= 1u << (A.stack[i].x & 31);
    __syncthreads();
  }
  while (ctl->status == MAC_RUNNING) {
    if (A.node_limit > 0 && ctl->npos + ctl->nneg >= A.node_limit) {
      __syncthreads();
      if (threadIdx.x == 0) ctl->status = MAC_LIMIT;
      __syncthreads();
      wg_copy_words(A.trail + ctl->depth * nw, mem.dom, nw);
      break;
    }
    if (A.deadline && wall_clock64() > A.deadline) {
      __syncthreads();
      if (threadIdx.x == 0) ctl->status = MAC_LIMIT;
      __syncthreads();
      wg_copy_words(A.trail + ctl->depth * nw, mem.dom, nw);
      break;
    }
    if (A.step_budget > 0 && ctl->nodes_here >= A.step_budget) {
      wg_copy_words(A.trail + ctl->depth * nw, mem.dom, nw);
      break;
    }
    // select: min |D| over undecided variables, lowest id on ties
    u64 key = ~0ull;
    for (int v = threadIdx.x; v < n; v += blockDim.x)
      if (!bit_of(mem.asg, v)) {
        const u64 k = ((u64)(unsigned)mem.size[v] << 32) | (unsigned)v;
        key = k < key ? k : key;
      }
    key = wg_reduce_min_u64(ctl, key);
    const int x = (int)(key & 0xffffffffu);
    int depth = ctl->depth;
    if (threadIdx.x == 0) {
      int a = -1;
      for (int w = 0; w < W; ++w) {
        const u64 dw = mem.dom[(long long)x * W + w];
        if (dw) {
          a = w * 64 + (int)__builtin_ctzll(dw);
          break;
        }
      }
      ctl->sel_var = x;
      ctl->sel_val = a;
    }
    wg_copy_words(A.trail + depth * nw, mem.dom, nw);
    __syncthreads();
    const int a = ctl->sel_val;
    if (threadIdx.x < W) mem.dom[(long long)x * W + threadIdx.x] = (threadIdx.x == (a >> 6)) ? (1ull << (a & 63)) : 0ull;
    wg_clear_bitmap(mem.cur, nvw);
    __syncthreads();
    if (threadIdx.x == 0) {
      A.stack[depth] = make_int2(x, a);
      mem.size[x] = 1;
      mem.asg[x >> 5] |= 1u << (x & 31);
      mem.cur[x >> 5] |= 1u << (x & 31);
      ctl->depth = depth + 1;
      ctl->npos += 1;
      ctl->nodes_here += 1;
    }
    __syncthreads();
    bool ok = wg_fixpoint(net, mem, ctl, false);
    wg_trace(net, mem, ctl, A, 1, x, a, ok);
    depth = ctl->depth;
    if (ok && depth == n) {
      if (A.solution)
        for (int v = threadIdx.x; v < n; v += blockDim.x) {
          int s = -1;
          for (int w = 0; w < W; ++w) {
            const u64 dw = mem.dom[(long long)v * W + w];
            if (dw) {
              s = w * 64 + (int)__builtin_ctzll(dw);
              break;
            }
          }
          A.solution[v] = s;
        }
      __syncthreads();
      if (threadIdx.x == 0) ctl->status = MAC_SAT;
      __syncthreads();
      break;
    }
    while (!ok && depth > 0) {
      --depth;
      const int2 d = A.stack[depth];
      wg_copy_words(mem.dom, A.trail + depth * nw, nw);
      __syncthreads();
      wg_recount(mem, n, W);
      __syncthreads();
      if (threadIdx.x == 0) {
        mem.asg[d.x >> 5] &= ~(1u << (d.x & 31));
        mem.dom[(long long)d.x * W + (d.y >> 6)] &= ~(1ull << (d.y & 63));
        mem.size[d.x] -= 1;
        ctl->depth = depth;
        ctl->nneg += 1;
        ctl->nodes_here += 1;
      }
      wg_clear_bitmap(mem.cur, nvw);
      __syncthreads();
      if (mem.size[d.x] == 0) {
        ok = false;
      } else {
        if (threadIdx.x == 0) mem.cur[d.x >> 5] |= 1u << (d.x & 31);
        __syncthreads();
        ok = wg_fixpoint(net, mem, ctl, false);
      }
      wg_trace(net, mem, ctl, A, 0, d.x, d.y, ok);
    }
    if (!ok) {
      __syncthreads();
      if (threadIdx.x == 0) ctl->status = MAC_UNSAT;
      __syncthreads();
      break;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    A.state->status = ctl->status;
    A.state->depth = ctl->depth;
    A.state->npos = ctl->npos;
    A.state->nneg = ctl->nneg;
    A.state->revisions += (long long)ctl->revs;
    A.state->trace_len = ctl->trace_len;
    A.state->steps += 1;
  }
}

template <bool kLds>
__global__ void k_mac(DevNet net, MacArgs args, char* scratch) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WgCtl ctl;
  char* base = kLds ? smem : scratch;
  WgMem mem = wg_carve(base, net.n, net.W);
  wg_mac_run(net, mem, &ctl, args);
}

// ---------------------------------------------------------------------------
// Many independent MAC runs (uneven instances, serving slots): one workgroup
// per unit, units taken from a global counter (persistent) or fixed by
// `only` (one unit per launch).
// ---------------------------------------------------------------------------
struct UnitDesc {
  DevNet net;
  MacArgs args;
  long long scratch_off;  // per-unit global scratch when the domain does not fit in LDS
  int use_lds;
  int pad;
};

__global__ void k_mac_units(const UnitDesc* __restrict__ units, int count, unsigned* counter,
                            int only, char* scratch) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WgCtl ctl;
  __shared__ int unit_sh;
  while (true) {
    if (threadIdx.x == 0) unit_sh = only >= 0 ? only : (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int u = unit_sh;
    __syncthreads();
    if (u >= count) break;
    const UnitDesc d = units[u];
    char* base = d.use_lds ? smem : scratch + d.scratch_off;
    WgMem mem = wg_carve(base, d.net.n, d.net.W);
    wg_mac_run(d.net, mem, &ctl, d.args);
    if (only >= 0) break;
  }
}
