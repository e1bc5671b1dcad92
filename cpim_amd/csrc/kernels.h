// Propagation kernels: model build, grid-wide constraint-frontier fixpoint
// (plain and persistent), and the workgroup engine (domains in LDS) used by
// block-mode GAC, batched singleton tests, the device MAC, batches of
// instances and the serving step.
#pragma once

#include "device.h"

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
// Statistics counters are spread over kStatSlots addresses (by block), and
// the grid rounds add one workgroup total per round: device-scope atomics on
// one address serialize, and one from every wave of an 8,192-wave round on
// 32 addresses took 80 of the 100 us of a C2 round.
constexpr int kStatSlots = 256;

struct GacCtl {
  unsigned scanner[3];
  int last_push_round;  // highest round that enqueued a constraint
  int wipe_round;       // lowest round that emptied a domain
  int rounds;           // rounds executed (written at the end)
  int converged;
  int pad;
  // grid barrier of the persistent kernel: arrivals in this generation, the
  // generation, and the error flag set by a block whose wait ran out of time
  unsigned long long bar_ticket;  // grid barrier arrivals so far (only grows)
  int err;
  int pad2;
  unsigned long long deletions[kStatSlots];
  unsigned long long revisions[kStatSlots];
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
  // the epoch only grows: skip the atomic once another wave has raised it
  if (lane == 0 && ld_agent(&ctl->last_push_round) < r) atomicMax(&ctl->last_push_round, r);
}

template <bool kCoherent = false>
__device__ __forceinline__ void grid_process_constraint(const DevNet& net, u64* dom, int* size,
                                                        int c, unsigned* B, GacCtl* ctl, int r,
                                                        unsigned long long& revs,
                                                        unsigned long long& dels) {
  const ConMeta* mp = net.con + c;
  bool wiped = false;
  if (mp->kind == 0) {
    const int x = mp->scope[0], y = mp->scope[1];
    int rx = revise_binary<kCoherent>(net, dom, size, c, x, y, 1, &wiped);
    ++revs;
    if (wiped) atomicMin(&ctl->wipe_round, r);
    if (rx) {
      dels += rx;
      push_var(net, x, B, ctl, r);
    }
    int ry = revise_binary<kCoherent>(net, dom, size, c, y, x, 0, &wiped);
    ++revs;
    if (wiped) atomicMin(&ctl->wipe_round, r);
    if (ry) {
      dels += ry;
      push_var(net, y, B, ctl, r);
    }
  } else {
    auto on_removed = [&](int v, int k) {
      dels += k;
      push_var(net, v, B, ctl, r);
    };
    revise_table<decltype(on_removed), kCoherent>(net, dom, size, c, -1, nullptr, on_removed, &wiped);
    ++revs;
    if (wiped) atomicMin(&ctl->wipe_round, r);
  }
}

// Binary constraints over one-word domains (W = 1), up to kPairBatch at a
// time, by one wave.  Both arcs of a constraint c = (x, y) are revised from
// the same D(x) and D(y): value a of x keeps iff row a of plane 0 meets D(y),
// value b of y keeps iff row b of plane 1 meets D(x).  With W = 1 the rows of
// every value are loaded whatever the domains, so the domain words (lane 2p:
// x of constraint p, lane 2p + 1: its y) and the 2 x 64 row words of every
// constraint of the batch are loaded together, one memory round trip for the
// batch.  Revising an arc from a domain that the other arc shrinks in the
// same step is a chaotic-iteration step like any other: the shrunk variable
// is pushed, so c is revised again in the next round, and the closure is the
// same.  The removal masks are applied with atomic AND by lanes 2p / 2p + 1;
// for a one-word domain the returned word gives the exact removals and the
// wipe-out (a non-empty word this AND empties), so the size update needs no
// returned value.
constexpr int kPairBatch = 4;
// The asynchronous fixpoint takes eight: its waves are few (two per CU) and
// bound by the latency of their own loads, so more rows in flight per wave
// pay (3.6-3.75 ms against 3.85-4.1 ms on the C2 cascade), where the round
// kernels lose occupancy at eight.
constexpr int kAsyncPairBatch = 8;

// The revisions of one batch; returns the mask of lanes whose variable (in
// myv) lost values, the values removed in the batch, and whether some domain
// was emptied by it.
template <bool kCoherent, int P = kPairBatch>
__device__ __forceinline__ u64 revise_pairs_w1(const DevNet& net, u64* dom, int* size, const int (&cs)[P],
                                               int k, int& myv, int& removed_total, bool& wiped_any) {
  const int lane = lane_id();
  int xs[P], ys[P];
  long long offs[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const ConMeta* mp = net.con + (p < k ? cs[p] : cs[0]);
    xs[p] = mp->scope[0];
    ys[p] = mp->scope[1];
    offs[p] = mp->off;
  }
  // lane 2p: D(x_p), lane 2p + 1: D(y_p)
  myv = -1;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    if (p < k && lane == 2 * p) myv = xs[p];
    if (p < k && lane == 2 * p + 1) myv = ys[p];
  }
  const u64 dl = myv >= 0 ? ld_dom<kCoherent>(&dom[myv]) : 0ull;
  u64 r0[P], r1[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const u64* plane0 = net.sup + offs[p];
    r0[p] = p < k ? plane0[lane] : 0ull;
    r1[p] = p < k ? plane0[net.dpad + lane] : 0ull;
  }
  u64 keep_l = ~0ull;  // lane 2p: new mask of x_p, lane 2p + 1: of y_p
#pragma unroll
  for (int p = 0; p < P; ++p) {
    if (p >= k) break;
    const u64 dx = readlane64(dl, 2 * p), dy = readlane64(dl, 2 * p + 1);
    const u64 kx = __ballot(((dx >> lane) & 1ull) && (r0[p] & dy) != 0ull);
    const u64 ky = __ballot(((dy >> lane) & 1ull) && (r1[p] & dx) != 0ull);
    if (lane == 2 * p) keep_l = kx;
    if (lane == 2 * p + 1) keep_l = ky;
  }
  int removed_l = 0;
  bool wiped_l = false;
  if (myv >= 0 && (dl & ~keep_l)) {
    const u64 old = atomicAnd(&dom[myv], keep_l);
    removed_l = __popcll(old & ~keep_l);
    wiped_l = old != 0ull && (old & keep_l) == 0ull;
    if (removed_l) atomicSub(&size[myv], removed_l);
  }
  wiped_any = __ballot(wiped_l) != 0ull;
  removed_total = (int)wave_sum64_dpp((u64)removed_l);
  return __ballot(removed_l > 0);
}

template <bool kCoherent>
__device__ __forceinline__ void grid_pairs_w1(const DevNet& net, u64* dom, int* size, const int (&cs)[kPairBatch],
                                              int k, unsigned* B, GacCtl* ctl, int r,
                                              unsigned long long& revs, unsigned long long& dels) {
  int myv = -1, removed = 0;
  bool wiped = false;
  u64 changed = revise_pairs_w1<kCoherent>(net, dom, size, cs, k, myv, removed, wiped);
  revs += 2 * k;
  dels += removed;
  if (wiped && lane_id() == 0) atomicMin(&ctl->wipe_round, r);
  while (changed) {
    const int l = (int)__builtin_ctzll(changed);
    changed &= changed - 1ull;
    push_var(net, __builtin_amdgcn_readlane(myv, l), B, ctl, r);
  }
}

// Work units are (frontier word, part) pairs: part p of word w covers the
// constraint bits [p * 32 / parts, (p + 1) * 32 / parts) of A[w], so a round
// over few words still spreads over many waves; waves take units in a fixed
// stride.  kCoherent: domains and frontier words are read with agent-scope
// loads (rounds of one launch, written by other XCDs since the last barrier).
// With one-word domains, binary constraints go kPairBatch at a time through
// grid_pairs_w1.
// revs / dels: the calling wave's statistics, accumulated.
template <bool kCoherent = false>
__device__ __forceinline__ void grid_round(const DevNet& net, u64* dom, int* size, unsigned* F,
                                           int nfw, GacCtl* ctl, int r, int parts,
                                           unsigned long long& revs, unsigned long long& dels) {
  unsigned* A = F + (long long)(r % 3) * nfw;
  unsigned* B = F + (long long)((r + 1) % 3) * nfw;
  unsigned* Cl = F + (long long)((r + 2) % 3) * nfw;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicExch(&ctl->scanner[(r + 1) % 3], 0u);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nfw;
       i += (long long)gridDim.x * blockDim.x) {
    if (kCoherent) __hip_atomic_store(&Cl[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else Cl[i] = 0u;
  }
  const long long units = (long long)nfw * parts;
  const int pbits = 32 / parts;
  // static stride over the units: one counter for tens of thousands of units
  // would serialize on its atomic
  const long long nwaves_grid = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long ui = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); ui < units;
       ui += nwaves_grid) {
    if (ld_agent(&ctl->wipe_round) <= r) break;
    const int wi = (int)(ui / parts);
    const int part = ui - wi * parts;
    const unsigned pmask = pbits == 32 ? ~0u : (((1u << pbits) - 1u) << (part * pbits));
    const unsigned aw = kCoherent ? (unsigned)ld_agent((const int*)&A[wi]) : A[wi];
    unsigned bits = (unsigned)uniform_int((int)(aw & pmask));
    if (net.W == 1) {
      while (bits) {
        int cs[kPairBatch] = {0, 0, 0, 0};
        int k = 0;
        while (bits && k < kPairBatch) {
          const int b = __ffs(bits) - 1;
          bits &= bits - 1u;
          const int c = wi * 32 + b;
          if (net.con[c].kind != 0) {
            grid_process_constraint<kCoherent>(net, dom, size, c, B, ctl, r, revs, dels);
            continue;
          }
#pragma unroll
          for (int p = 0; p < kPairBatch; ++p)
            if (p == k) cs[p] = c;
          ++k;
        }
        if (k) grid_pairs_w1<kCoherent>(net, dom, size, cs, k, B, ctl, r, revs, dels);
      }
      continue;
    }
    while (bits) {
      const int b = __ffs(bits) - 1;
      bits &= bits - 1u;
      grid_process_constraint<kCoherent>(net, dom, size, wi * 32 + b, B, ctl, r, revs, dels);
    }
  }
}

// Adds the waves' statistics to the grid's totals: one workgroup total per
// call, on the workgroup's slot.  blk: two LDS words, zero on entry.  Every
// thread of the workgroup calls it.
__device__ __forceinline__ void flush_grid_stats(GacCtl* ctl, unsigned long long* blk, unsigned long long revs,
                                                 unsigned long long dels) {
  if (lane_id() == 0 && (revs | dels)) {
    atomicAdd(&blk[0], revs);
    atomicAdd(&blk[1], dels);
  }
  __syncthreads();
  if (threadIdx.x == 0 && (blk[0] | blk[1])) {
    atomicAdd(&ctl->revisions[blockIdx.x % kStatSlots], blk[0]);
    atomicAdd(&ctl->deletions[blockIdx.x % kStatSlots], blk[1]);
  }
}

// One round per launch (host-driven fixpoint).
__global__ void k_gac_round(DevNet net, u64* dom, int* size, unsigned* F, int nfw, GacCtl* ctl,
                            int r, int parts) {
  __shared__ unsigned long long blk[2];
  if (threadIdx.x == 0) {
    blk[0] = 0ull;
    blk[1] = 0ull;
  }
  __syncthreads();
  unsigned long long revs = 0, dels = 0;
  grid_round(net, dom, size, F, nfw, ctl, r, parts, revs, dels);
  flush_grid_stats(ctl, blk, revs, dels);
}

// ===========================================================================
// Persistent fixpoint: every round in one launch, rounds separated by a grid
// barrier.  The grid is sized by the host so that every workgroup is resident
// at once (occupancy x CUs).
//
// The barrier is a ticket count that only grows: thread 0 of each workgroup
// takes a ticket (an atomic add that returns the count) and waits until the
// count reaches the end of its ticket's round, nblocks tickets per round.
// Every wait is bounded: a waiter whose device wall clock passes the launch's
// budget, or that sees the error flag another block raised, stops waiting and
// its block leaves, so the launch always ends (an error, never a hang), even
// if the workgroups were not all resident.  The decision taken after the
// barrier (next round, wipe-out, fixpoint, error) is made by thread 0 and
// broadcast through LDS; every wave reads it back as a wave-uniform value
// (readfirstlane), so the round loop leaves through a scalar branch.
// The round decision uses epochs that only grow (last_push_round,
// wipe_round): a block that already runs round r + 1 can raise them, but
// never so that a slower block would decide differently for round r.
// ===========================================================================
enum : int { kRoundNext = 0, kRoundWipe = 1, kRoundFixpoint = 2, kRoundError = 3 };

// Every write a round makes that another workgroup reads in a later round
// (domain words, sizes, frontier bits, epochs, the cleared frontier) is an
// agent-scope atomic, and every such read is an agent-scope atomic load, so
// the data never waits in a cache that another XCD cannot see.  Before the
// barrier each wave therefore only waits until its own memory operations have
// completed (s_waitcnt vmcnt(0): atomics with and without return, and
// stores), and the barrier's own accesses are relaxed: no fence, no L2
// write-back or invalidation.  (A release fence in every wave made a barrier
// cost 37 us at 1,024 waves; with one fence-pair per workgroup a 21-round
// cascade took 1.41 ms, fence-free 1.06 ms.)  The
// asynchronous fixpoint relies on the same coherence of agent-scope
// accesses, with no barrier at all.
__device__ __forceinline__ void wait_memory_ops() {
  // vmcnt(0); expcnt and lgkmcnt left at their maxima (no wait)
  __builtin_amdgcn_s_waitcnt(0x0F70);
}

__device__ __forceinline__ int grid_barrier_decide(GacCtl* ctl, unsigned nblocks, int r,
                                                   unsigned long long t0, unsigned long long max_ticks,
                                                   int* s_decision) {
  wait_memory_ops();
  __syncthreads();
  if (threadIdx.x == 0) {
    // every workgroup takes one ticket per round, so round r's tickets are
    // r * nblocks .. (r + 1) * nblocks - 1; the round is complete once the
    // count reaches the end of this ticket's round
    const unsigned long long ticket = atomicAdd(&ctl->bar_ticket, 1ull);
    const unsigned long long target = (ticket / nblocks + 1) * nblocks;
    int st = kRoundNext;
    while (__hip_atomic_load(&ctl->bar_ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        st = kRoundError;
        break;
      }
      if (wall_clock64() - t0 > max_ticks) {
        atomicExch(&ctl->err, 1);
        st = kRoundError;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (st == kRoundNext && __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      st = kRoundError;
    if (st == kRoundNext) {
      const int wr = ld_agent(&ctl->wipe_round);
      const int lp = ld_agent(&ctl->last_push_round);
      st = wr <= r ? kRoundWipe : (lp < r ? kRoundFixpoint : kRoundNext);
    }
    *s_decision = st;
  }
  __syncthreads();
  return uniform_int(*s_decision);
}

constexpr int kPersistentThreads = 1024;

__global__ void __launch_bounds__(kPersistentThreads) k_gac_persistent(DevNet net, u64* dom, int* size, unsigned* F, int nfw,
                                                        GacCtl* ctl, int max_rounds, int parts,
                                                        unsigned long long max_ticks) {
  __shared__ int s_decision;
  __shared__ unsigned long long blk[2];
  if (threadIdx.x == 0) {
    blk[0] = 0ull;
    blk[1] = 0ull;
  }
  __syncthreads();
  const unsigned long long t0 = wall_clock64();
  int r = 0;
  int decision = kRoundNext;
  unsigned long long revs = 0, dels = 0;  // the wave's, over every round
  for (; r < max_rounds; ++r) {
    grid_round<true>(net, dom, size, F, nfw, ctl, r, parts, revs, dels);
    decision = grid_barrier_decide(ctl, gridDim.x, r, t0, max_ticks, &s_decision);
    if (decision != kRoundNext) break;
  }
  flush_grid_stats(ctl, blk, revs, dels);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->rounds = r < max_rounds ? r + 1 : max_rounds;
    ctl->converged = decision == kRoundFixpoint ? 1 : 0;
  }
}

// Start of a host-driven fixpoint in one launch: the control block, the three
// frontier buffers cleared, and with seed_all every constraint in the first.
__global__ void k_gac_init(GacCtl* ctl, unsigned* F, int nfw, int ncons, int seed_all) {
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0) {
    for (int k = 0; k < 3; ++k) ctl->scanner[k] = 0u;
    ctl->last_push_round = -1;
    ctl->wipe_round = INT_MAX;
    ctl->rounds = 0;
    ctl->converged = 0;
    ctl->pad = 0;
    ctl->bar_ticket = 0ull;
    ctl->err = 0;
    ctl->pad2 = 0;
  }
  for (long long i = i0; i < kStatSlots; i += (long long)gridDim.x * blockDim.x) {
    ctl->deletions[i] = 0ull;
    ctl->revisions[i] = 0ull;
  }
  for (long long i = i0; i < 3ll * nfw; i += (long long)gridDim.x * blockDim.x) {
    unsigned v = 0u;
    if (seed_all && i < nfw) {
      const int k = ncons - (int)(i * 32);
      v = k >= 32 ? 0xffffffffu : (k <= 0 ? 0u : ((1u << k) - 1u));
    }
    F[i] = v;
  }
}

__global__ void k_seed_var(unsigned* F, const int* adj_off, const int4* adj, int var) {
  for (int i = adj_off[var] + threadIdx.x; i < adj_off[var + 1]; i += blockDim.x) {
    const int c = adj[i].x;
    atomicOr(&F[c >> 5], 1u << (c & 31));
  }
}

// ===========================================================================
// Asynchronous fixpoint in one launch, without any grid barrier.
//
// The frontier is a bitmap of constraints.  A wave claims a whole frontier
// word with an atomic exchange, revises the claimed constraints and pushes
// the constraints of every variable that lost values.  `pending` counts the
// queued plus the claimed-but-unfinished constraints; it is raised before a
// push sets its bits (once per revision batch, for every constraint the batch
// pushes) and lowered only after the claimed constraints are finished and
// for the pushed bits that were already set.  The lowering is kept per wave
// (`owed`) and published when the wave finds no work in a frontier chunk:
// pending never drops below the outstanding work, and it reaches 0 once the
// fixpoint is reached and every wave has gone idle.  (One atomic on pending
// per push and per claim made that word the bottleneck: 11 ms against 4 ms on
// a 23-round cascade.)
// No wave waits for another: an idle wave keeps scanning until pending is 0,
// the wipe-out flag is set, or its time budget runs out (then it sets `err`
// and leaves), so every wave exits and co-residency is not required.
// Domain and frontier words are read with agent-scope atomic loads, so a
// revision always sees removals completed by other XCDs before the push that
// queued it.
// ===========================================================================
struct AsyncCtl {
  unsigned long long pending;
  int wipe;
  int err;
  unsigned long long deletions[kStatSlots];
  unsigned long long revisions[kStatSlots];
};

// counted: pending was already raised by the caller for all of v's constraints
__device__ __forceinline__ void push_var_async(const DevNet& net, int v, unsigned* F, AsyncCtl* ctl, int& owed,
                                               bool counted = false) {
  const int lane = lane_id();
  const int b0 = net.adj_off[v], b1 = net.adj_off[v + 1];
  for (int i0 = b0; i0 < b1; i0 += 64) {
    const int k = min(64, b1 - i0);
    if (!counted && lane == 0) atomicAdd(&ctl->pending, (unsigned long long)k);
    bool already = false;
    if (lane < k) {
      const int c = net.adj[i0 + lane].x;
      const unsigned bit = 1u << (c & 31);
      already = (atomicOr(&F[c >> 5], bit) & bit) != 0u;
    }
    owed += __popcll(__ballot(already));
  }
}

__device__ __forceinline__ void async_process_constraint(const DevNet& net, u64* dom, int* size, int c,
                                                         unsigned* F, AsyncCtl* ctl,
                                                         unsigned long long& revs,
                                                         unsigned long long& dels, int& owed) {
  const ConMeta* mp = net.con + c;
  bool wiped = false;
  if (mp->kind == 0) {
    const int x = mp->scope[0], y = mp->scope[1];
    const int rx = revise_binary<true>(net, dom, size, c, x, y, 1, &wiped);
    ++revs;
    if (wiped && lane_id() == 0) atomicExch(&ctl->wipe, 1);
    if (rx) {
      dels += rx;
      push_var_async(net, x, F, ctl, owed);
    }
    const int ry = revise_binary<true>(net, dom, size, c, y, x, 0, &wiped);
    ++revs;
    if (wiped && lane_id() == 0) atomicExch(&ctl->wipe, 1);
    if (ry) {
      dels += ry;
      push_var_async(net, y, F, ctl, owed);
    }
  } else {
    auto on_removed = [&](int v, int k) {
      dels += k;
      push_var_async(net, v, F, ctl, owed);
    };
    revise_table<decltype(on_removed), true>(net, dom, size, c, -1, nullptr, on_removed, &wiped);
    ++revs;
    if (wiped && lane_id() == 0) atomicExch(&ctl->wipe, 1);
  }
}

__global__ void k_gac_async(DevNet net, u64* dom, int* size, unsigned* F, int nfw, AsyncCtl* ctl,
                            unsigned long long max_ticks) {
  const int lane = lane_id();
  const long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned long long t0 = wall_clock64();
  const int nchunks = (nfw + 63) / 64;
  int chunk = (int)(gw % nchunks);
  int seen = 0;  // chunks scanned since the last claim
  int owed = 0;  // finished claims and duplicate pushes not yet taken off pending
  unsigned long long revs = 0, dels = 0;
  while (true) {
    if (__hip_atomic_load(&ctl->wipe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
        __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      break;
    const int wi = chunk * 64 + lane;
    unsigned bits = 0u;
    if (wi < nfw) bits = __hip_atomic_load(&F[wi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const u64 nz = __ballot(bits != 0u);
    if (nz) {
      const int l = (int)__builtin_ctzll(nz);
      unsigned claimed = 0u;
      if (lane == 0) claimed = atomicExch(&F[chunk * 64 + l], 0u);
      claimed = __shfl(claimed, 0, 64);
      const int nclaimed = __popc(claimed);
      const int base_c = (chunk * 64 + l) * 32;
      if (net.W == 1) {
        // one-word domains: binary constraints kAsyncPairBatch at a time
        claimed = (unsigned)uniform_int((int)claimed);
        while (claimed) {
          int cs[kAsyncPairBatch] = {};
          int k = 0;
          while (claimed && k < kAsyncPairBatch) {
            const int b = __ffs(claimed) - 1;
            claimed &= claimed - 1u;
            const int c = base_c + b;
            if (net.con[c].kind != 0) {
              async_process_constraint(net, dom, size, c, F, ctl, revs, dels, owed);
              continue;
            }
#pragma unroll
            for (int p = 0; p < kAsyncPairBatch; ++p)
              if (p == k) cs[p] = c;
            ++k;
          }
          if (k) {
            int myv = -1, removed = 0;
            bool wiped = false;
            u64 changed = revise_pairs_w1<true, kAsyncPairBatch>(net, dom, size, cs, k, myv, removed, wiped);
            revs += 2 * k;
            dels += removed;
            if (wiped && lane == 0) atomicExch(&ctl->wipe, 1);
            if (changed) {
              // raise pending once for every constraint the changed
              // variables push, before any of their bits is set
              int tot = 0;
              if ((changed >> lane) & 1ull) tot = net.adj_off[myv + 1] - net.adj_off[myv];
              for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
              if (lane == 0) atomicAdd(&ctl->pending, (unsigned long long)tot);
            }
            while (changed) {
              const int cl = (int)__builtin_ctzll(changed);
              changed &= changed - 1ull;
              push_var_async(net, __builtin_amdgcn_readlane(myv, cl), F, ctl, owed, true);
            }
          }
        }
      } else {
        while (claimed) {
          const int b = __ffs(claimed) - 1;
          claimed &= claimed - 1u;
          async_process_constraint(net, dom, size, base_c + b, F, ctl, revs, dels, owed);
        }
      }
      owed += nclaimed;
      seen = 0;
      continue;
    }
    chunk = chunk + 1 == nchunks ? 0 : chunk + 1;
    if (owed) {
      if (lane == 0) atomicSub(&ctl->pending, (unsigned long long)owed);
      owed = 0;
    }
    if (++seen >= nchunks) {
      seen = 0;
      const unsigned long long p =
          __hip_atomic_load(&ctl->pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (p == 0ull) break;
      if (wall_clock64() - t0 > max_ticks) {
        if (lane == 0) atomicExch(&ctl->err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
  }
  if (lane == 0 && (revs | dels)) {
    atomicAdd(&ctl->revisions[blockIdx.x % kStatSlots], revs);
    atomicAdd(&ctl->deletions[blockIdx.x % kStatSlots], dels);
  }
}

// Seed the asynchronous frontier: all active constraints, or those of var.
__global__ void k_seed_async(unsigned* F, int nfw, int ncons, const int* adj_off, const int4* adj, int var,
                             AsyncCtl* ctl) {
  if (var < 0) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nfw) {
      const int k = ncons - (int)(i * 32);
      F[i] = k >= 32 ? 0xffffffffu : (k <= 0 ? 0u : ((1u << k) - 1u));
    }
    if (i == 0) ctl->pending = (unsigned long long)ncons;
    return;
  }
  if (blockIdx.x != 0) return;
  for (int i = adj_off[var] + threadIdx.x; i < adj_off[var + 1]; i += blockDim.x) {
    const int c = adj[i].x;
    const unsigned bit = 1u << (c & 31);
    if (!(atomicOr(&F[c >> 5], bit) & bit)) atomicAdd(&ctl->pending, 1ull);
  }
}

// Streaming copy (16-byte lanes, grid-stride): the known-traffic kernel used
// to calibrate hardware byte counters and to measure achievable bandwidth.
__global__ void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, long long n16) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
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
  const int* aoff;   // adjacency offsets (net.adj_off)
  const int4* adj;   // adjacency entries (net.adj)
};

struct WgCtl {
  int arc_ctr;
  int wipe;
  int nchanged[2];
  int m;
  int total;
  int k;  // arcs per work item this round
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

__device__ __forceinline__ WgMem wg_carve(char* base, const DevNet& net) {
  const int n = net.n, W = net.W;
  const int nvw = (n + 31) / 32;
  WgMem m;
  m.aoff = net.adj_off;
  m.adj = net.adj;
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

// Record that v changed in this round: set its bit in the round's next
// frontier (the swapped pointer, not mem.nxt) and count it once.
__device__ __forceinline__ void wg_mark(unsigned* nxt, WgCtl* ctl, int v, int slot) {
  const unsigned bit = 1u << (v & 31);
  const unsigned old = atomicOr(&nxt[v >> 5], bit);
  if (!(old & bit)) atomicAdd(&ctl->nchanged[slot], 1);
}

// AC fixpoint from the seed in mem.cur.  both_dirs_round0: in the first round
// also revise t <- u for every arc of a seed variable t (constraint seeding,
// as GModel::EnforceGAC(assigned_var) does).  Returns false on wipe-out.
// Must be called by every thread of the workgroup.  Work items are runs of
// up to kWgArcs arcs of one changed variable: the lanes of a wave fetch the
// arcs' metadata together, then the wave revises kP binary arcs at a time
// (revise_binary_multi keeps the row loads of all in flight), or one by one
// for tables and assigned targets.  kP = 4 for the MAC kernels (uneven batch
// 47 ms against 50 ms at kP = 2); 2 elsewhere, where kP = 4 costs occupancy.
constexpr int kWgArcs = 8;  // most arcs per work item

template <int kP = 2>
__device__ __forceinline__ bool wg_fixpoint(const DevNet& net, WgMem mem, WgCtl* ctl, bool both_dirs_round0) {
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
      // item size: at most kWgArcs arcs, small enough for about two items
      // per wave when the round has few arcs
      int arcs = 0;
      for (int v0 = 0; v0 < n; v0 += 64) {
        const int v = v0 + lane;
        if (v < n && bit_of(cur, v)) arcs += mem.aoff[v + 1] - mem.aoff[v];
      }
      (void)wave_scan_dpp(arcs, &arcs);
      const int nwaves = nthreads >> 6;
      const int K = uniform_int(max(1, min(kWgArcs, (arcs + 2 * nwaves - 1) / (2 * nwaves))));
      int base = 0, ibase = 0;
      for (int v0 = 0; v0 < n; v0 += 64) {
        const int v = v0 + lane;
        const bool f = v < n && bit_of(cur, v);
        const u64 mask = __ballot(f);
        const int rank = mbcnt(mask);
        const int items = f ? (mem.aoff[v + 1] - mem.aoff[v] + K - 1) / K : 0;
        int round_items = 0;
        const int inc = wave_scan_dpp(items, &round_items);
        if (f) {
          mem.list[base + rank] = v;
          mem.pre[base + rank] = ibase + inc - items;
        }
        base += __popcll(mask);
        ibase += round_items;
      }
      if (lane == 0) {
        mem.pre[base] = ibase;
        ctl->m = base;
        ctl->total = ibase;
        ctl->k = K;
        ctl->arc_ctr = 0;
        ctl->nchanged[slot] = 0;
      }
    } else {
      for (int i = tid - 64; i < nvw; i += nthreads - 64) nxt[i] = 0u;
    }
    if (nthreads == 64)
      for (int i = lane; i < nvw; i += 64) nxt[i] = 0u;
    __syncthreads();
    const int total = uniform_int(ctl->total);
    const int m = uniform_int(ctl->m);
    const int K = uniform_int(ctl->k);
    const bool both = both_dirs_round0 && round == 0;
    while (true) {
      int idx = 0;
      if (lane == 0) idx = atomicAdd(&ctl->arc_ctr, 1);
      idx = uniform_int(__shfl(idx, 0, 64));
      if (idx >= total) break;
      if (uniform_int(ctl->wipe)) break;
      int lo = 0, hi = m - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mem.pre[mid] <= idx) lo = mid;
        else hi = mid - 1;
      }
      const int t = uniform_int(mem.list[lo]);
      const int s0 = mem.aoff[t] + (idx - uniform_int(mem.pre[lo])) * K;
      const int cnt = uniform_int(min(s0 + K, mem.aoff[t + 1]) - s0);
      int4 e = make_int4(0, 0, 0, 0);
      if (lane < cnt) e = mem.adj[s0 + lane];
      for (int i = 0; i < cnt; ++i) {
        const int c = __builtin_amdgcn_readlane(e.x, i);
        const int u = __builtin_amdgcn_readlane(e.y, i);
        bool wiped = false;
        // up to kP binary arcs with live targets: revised together
        if (u >= 0 && !both) {
          int np = 1;
          while (np < kP && i + np < cnt) {
            const int un = __builtin_amdgcn_readlane(e.y, i + np);
            if (un < 0 || uniform_int((int)bit_of(mem.asg, un))) break;
            ++np;
          }
          if (np >= 2 && !uniform_int((int)bit_of(mem.asg, u))) {
            if (kP >= 4 && np >= 4) {
              long long o[4];
              int uu[4], pt[4], r[4];
              bool w[4];
#pragma unroll
              for (int p = 0; p < 4; ++p) {
                o[p] = (long long)__builtin_amdgcn_readlane(e.w, i + p) * 2 * net.W * net.dpad;
                uu[p] = __builtin_amdgcn_readlane(e.y, i + p);
                pt[p] = __builtin_amdgcn_readlane(e.z, i + p);
              }
              revise_binary_multi<4>(net, mem.dom, mem.size, t, o, uu, pt, r, w);
              revs += 4;
#pragma unroll
              for (int p = 0; p < 4; ++p) {
                if (r[p] && lane == 0) wg_mark(nxt, ctl, uu[p], slot);
                if (w[p] && lane == 0) ctl->wipe = 1;
              }
              i += 3;
              continue;
            }
            long long o[2];
            int uu[2], pt[2], r[2];
            bool w[2];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
              o[p] = (long long)__builtin_amdgcn_readlane(e.w, i + p) * 2 * net.W * net.dpad;
              uu[p] = __builtin_amdgcn_readlane(e.y, i + p);
              pt[p] = __builtin_amdgcn_readlane(e.z, i + p);
            }
            revise_binary_multi<2>(net, mem.dom, mem.size, t, o, uu, pt, r, w);
            revs += 2;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
              if (r[p] && lane == 0) wg_mark(nxt, ctl, uu[p], slot);
              if (w[p] && lane == 0) ctl->wipe = 1;
            }
            ++i;
            continue;
          }
        }
        if (u >= 0) {  // binary
          const int pos_t = __builtin_amdgcn_readlane(e.z, i);
          const long long o = (long long)__builtin_amdgcn_readlane(e.w, i) * 2 * net.W * net.dpad;
          if (!bit_of(mem.asg, u)) {
            const int r = revise_binary_at(net, mem.dom, mem.size, o, u, t, pos_t, &wiped);
            ++revs;
            if (r && lane == 0) wg_mark(nxt, ctl, u, slot);
            if (wiped && lane == 0) ctl->wipe = 1;
          }
          if (both && !bit_of(mem.asg, t)) {
            const int r = revise_binary_at(net, mem.dom, mem.size, o, t, u, 1 - pos_t, &wiped);
            ++revs;
            if (r && lane == 0) wg_mark(nxt, ctl, t, slot);
            if (wiped && lane == 0) ctl->wipe = 1;
          }
        } else {
          revise_table(net, mem.dom, mem.size, c, both ? -1 : t, mem.asg,
                       [&](int v, int) {
                         if (lane == 0) wg_mark(nxt, ctl, v, slot);
                       },
                       &wiped);
          ++revs;
          if (wiped && lane == 0) ctl->wipe = 1;
        }
      }
    }
    __syncthreads();
    if (uniform_int(ctl->wipe)) {
      result = false;
      break;
    }
    if (uniform_int(ctl->nchanged[slot]) == 0) break;
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
// Workgroup reductions, called by every thread: each wave reduces over DPP,
// lane 0 publishes the wave's result, and every wave reduces the published
// results (lane i reads wave i's) the same way.
__device__ __forceinline__ u64 wg_reduce_min_u64(WgCtl* ctl, u64 v) {
  v = wave_min64_dpp(v);
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, lane = lane_id();
  if (lane == 0) ctl->red[wave] = v;
  __syncthreads();
  const u64 r = wave_min64_dpp(lane < nwaves ? ctl->red[lane] : ~0ull);
  __syncthreads();
  return (u64)uniform_ll((long long)r);
}

__device__ __forceinline__ u64 wg_reduce_sum_u64(WgCtl* ctl, u64 v) {
  v = wave_sum64_dpp(v);
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, lane = lane_id();
  if (lane == 0) ctl->red[wave] = v;
  __syncthreads();
  const u64 r = wave_sum64_dpp(lane < nwaves ? ctl->red[lane] : 0ull);
  __syncthreads();
  return (u64)uniform_ll((long long)r);
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
  WgMem mem = wg_carve(base, net);
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
// Batched singleton tests: workgroup `first + blockIdx.x` takes test (x, a),
// copies the base level into its domain copy, reduces x to {a}, marks x
// assigned and runs the fixpoint from {x}; fail[i] records a wipe-out.  The
// copies never touch HBM when they fit in LDS.  One test per workgroup: the
// kernel has no loop around its barriers, and the hardware dispatcher keeps
// every CU busy as workgroups finish at different times.
// ---------------------------------------------------------------------------
template <bool kLds>
__global__ void k_singleton_tests(DevNet net, const u64* __restrict__ base_dom,
                                  const int* __restrict__ base_size, const int2* __restrict__ tests,
                                  long long ntests, long long first, unsigned char* fail,
                                  unsigned long long* revs_out, char* scratch, long long scratch_stride) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WgCtl ctl;
  const long long unit = first + blockIdx.x;
  if (unit >= ntests) return;
  const int n = net.n, W = net.W;
  const int nvw = (n + 31) / 32;
  char* base = kLds ? smem : scratch + (long long)blockIdx.x * scratch_stride;
  WgMem mem = wg_carve(base, net);
  const int2 tst = tests[unit];
  const int x = uniform_int(tst.x), a = uniform_int(tst.y);
  wg_copy_words(mem.dom, base_dom, (long long)n * W);
  for (int v = threadIdx.x; v < n; v += blockDim.x) mem.size[v] = base_size[v];
  wg_clear_bitmap(mem.asg, nvw);
  wg_clear_bitmap(mem.cur, nvw);
  if (threadIdx.x == 0) ctl.revs = 0;
  __syncthreads();
  if (threadIdx.x < W) mem.dom[(long long)x * W + threadIdx.x] = (threadIdx.x == (a >> 6)) ? (1ull << (a & 63)) : 0ull;
  if (threadIdx.x == 0) {
    mem.size[x] = 1;
    mem.asg[x >> 5] |= 1u << (x & 31);
    mem.cur[x >> 5] |= 1u << (x & 31);
  }
  __syncthreads();
  const bool ok = wg_fixpoint(net, mem, &ctl, false);
  if (threadIdx.x == 0) {
    fail[unit] = ok ? 0 : 1;
    if (ctl.revs) atomicAdd(&revs_out[blockIdx.x % kStatSlots], ctl.revs);
  }
}

// ---------------------------------------------------------------------------
// Bit-sliced singleton tests (binary networks).  A workgroup runs up to 64
// singleton tests together: sl[v][b] is the 64-bit mask of the tests in
// which value b of v is alive.  Revising u <- t for all of them at once,
// value b of u keeps test k iff some value a of t with R(b, a) has bit k in
// sl[t][a].  One support row serves 64 tests, and each value stops scanning
// its supports as soon as every test that holds it has found one.
//
// Every test's bits evolve as a chaotic AC iteration of that test alone
// (a revision is monotone and idempotent, and a revision caused by another
// test's change is merely redundant), and a round ends with no change only
// when every arc is consistent in every test.  So each test reaches the GAC
// closure of its own network (unique), and test k failed iff some variable
// holds no value with bit k.  The tests' variables are not skipped as
// assigned: with x reduced to {a} the closure is the same either way.
//
// The slices live in global memory (n * dpad words per workgroup), read with
// agent-scope loads so that no stale L1 line is ever used; the source slice
// of the current work item is staged in wave-private LDS.
// ---------------------------------------------------------------------------
constexpr int kSliceArcs = 16;  // most arcs of one variable per work item

// Staged source slices: word w of t's domain occupies kStageWord entries: a
// zero entry, the masks of values 64w..64w+31, a zero entry, the masks of
// values 64w+32..64w+63.  A half-word scan reads entry ffbl(bits) of its
// half, and v_ffbl_b32 of an exhausted (zero) half is -1: the zero entry.
constexpr int kStageWord = 66;

struct SliceCtl {
  int item_ctr;
  int nchanged[2];
  int m;
  int total;
  int k;  // arcs per work item this round
  u64 act;  // tests not wiped out
  unsigned long long revs;
};

__host__ __device__ inline long long slice_lds_bytes(int n, int dpad, int nwaves) {
  const long long nvw = (n + 31) / 32;
  long long b = 4ll * n + 4ll * (n + 1) + 12ll * nvw;
  b = (b + 15) & ~15ll;
  b += 8ll * nwaves * (kStageWord * (dpad / 64) + CPIM_MAX_WORDS);
  return b;
}

__device__ __forceinline__ void slice_mark(unsigned* nxt, SliceCtl* ctl, int v, int slot) {
  const unsigned bit = 1u << (v & 31);
  const unsigned old = atomicOr(&nxt[v >> 5], bit);
  if (!(old & bit)) atomicAdd(&ctl->nchanged[slot], 1);
}


__device__ __forceinline__ int stage_pos(int v) {
  const int b = v & 63;
  return (v >> 6) * kStageWord + 1 + b + (b >> 5);
}

// Lowest set bit of x, -1 for 0 (the hardware result of v_ffbl_b32).
__device__ __forceinline__ int ffbl_raw(unsigned x) {
  int r;
  asm("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// pend &= ~(st[a] for every set bit a of bits), four bits per step; st[-1]
// is the zero entry.
__device__ __forceinline__ u64 clear_supported(u64 pend, unsigned bits, const u64* st) {
  while (bits && pend) {
    const int a0 = ffbl_raw(bits);
    bits &= bits - 1u;
    const int a1 = ffbl_raw(bits);
    bits &= bits - 1u;
    const int a2 = ffbl_raw(bits);
    bits &= bits - 1u;
    const int a3 = ffbl_raw(bits);
    bits &= bits - 1u;
    pend &= ~((st[a0] | st[a1]) | (st[a2] | st[a3]));
  }
  return pend;
}

// Revise u <- t for every live test (act).  plane_u: rows of the values of u
// over t's domain (element (w, b) = word w of the row of b); slu: slices of
// u; stage: slices of t in the kStageWord layout; atw[w]: values of t alive
// in some live test.  A variable no test has changed yet (touched_u false)
// still holds `all` on its base values, so its slices are derived from the
// base domain words base_u instead of being loaded; a stale "untouched"
// only yields a superset of the live tests, which costs work, not accuracy.  Returns (wave-uniform) whether some bit of u was cleared
// by this wave.
__device__ __forceinline__ bool slice_revise(const u64* __restrict__ plane_u, u64* slu,
                                             const u64* stage, const u64* atw, int W, int dpad,
                                             u64 act, bool touched_u, const u64* __restrict__ base_u,
                                             u64 all) {
  const int lane = lane_id();
  bool changed = false;
  for (int j0 = 0; j0 < W; j0 += 4) {
    u64 need[4], row0[4];
    if (touched_u) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = j0 + q;
        need[q] = j < W ? ld_wg(&slu[j * 64 + lane]) : 0ull;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = j0 + q;
        need[q] = (j < W && ((base_u[j < W ? j : 0] >> lane) & 1ull)) ? all : 0ull;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = j0 + q;
      row0[q] = j < W ? plane_u[j * 64 + lane] : 0ull;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int b = (j0 + q) * 64 + lane;
      u64 pend = need[q] & act;  // live tests holding b that still need a support
      if (pend) {
        const u64 r = row0[q] & atw[0];
        pend = clear_supported(pend, (unsigned)r, stage + 1);
        pend = clear_supported(pend, (unsigned)(r >> 32), stage + 34);
      }
      for (int w = 1; w < W && pend; ++w) {
        const u64 r = plane_u[(long long)w * dpad + b] & atw[w];
        const u64* st = stage + w * kStageWord;
        pend = clear_supported(pend, (unsigned)r, st + 1);
        pend = clear_supported(pend, (unsigned)(r >> 32), st + 34);
      }
      if (pend) {
        const u64 old = and_wg(&slu[b], ~pend);
        if (old & pend) changed = true;
      }
    }
  }
  return __ballot(changed) != 0ull;
}

// Workgroup `blockIdx.x` runs test group g (tests [gfirst[g], gfirst[g+1])
// of the list sorted by variable) in its slice region; fail[perm[i]]
// (fail[i] when perm is null) receives the outcome of sorted test i.  Blocks are mapped to groups so
// that each XCD takes a contiguous run of groups (same variables, shared
// support rows in its L2).
__global__ void k_sliced_tests(DevNet net, const u64* __restrict__ base_dom,
                               const int2* __restrict__ tests, const int* __restrict__ gfirst,
                               long long g0, const int* __restrict__ perm, unsigned char* fail,
                               unsigned long long* revs_out, u64* slices) {
  extern __shared__ __align__(16) char smem[];
  __shared__ SliceCtl ctl;
  const int n = net.n, W = net.W, dpad = net.dpad;
  const int nvw = (n + 31) / 32;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int lane = lane_id();
  const int wave = uniform_int(tid >> 6), nwaves = nthreads >> 6;
  const int B = gridDim.x, bid = blockIdx.x;
  const int xcd = bid & 7, q = B >> 3, rem = B & 7;
  const long long g = g0 + (long long)xcd * q + (xcd < rem ? xcd : rem) + (bid >> 3);
  const int tf = gfirst[g];
  const int tc = uniform_int(gfirst[g + 1] - tf);
  const u64 all = tc >= 64 ? ~0ull : ((1ull << tc) - 1ull);
  u64* sl = slices + (long long)bid * n * dpad;

  int* list = (int*)smem;
  int* pre = list + n;
  unsigned* cur = (unsigned*)(pre + n + 1);
  unsigned* nxt = cur + nvw;
  unsigned* touched = nxt + nvw;  // variables some test has changed
  const long long head = (4ll * n + 4ll * (n + 1) + 12ll * nvw + 15) & ~15ll;
  u64* stage = (u64*)(smem + head) + (long long)wave * (kStageWord * W + CPIM_MAX_WORDS);
  u64* atw = stage + kStageWord * W;
  if (lane < 2 * W) stage[(lane >> 1) * kStageWord + 33 * (lane & 1)] = 0ull;  // zero entries

  // Slices of the base level: every test holds every base value.
  for (int v = wave; v < n; v += nwaves) {
    for (int b = lane; b < dpad; b += 64) {
      const bool in = (base_dom[(long long)v * W + (b >> 6)] >> (b & 63)) & 1ull;
      sl[(long long)v * dpad + b] = in ? all : 0ull;
    }
  }
  for (int i = tid; i < nvw; i += nthreads) {
    cur[i] = 0u;
    nxt[i] = 0u;
    touched[i] = 0u;
  }
  if (tid == 0) {
    ctl.act = all;
    ctl.revs = 0;
  }
  __syncthreads();
  // Test k reduces its variable x_k to {a_k}.
  for (int k = wave; k < tc; k += nwaves) {
    const int2 tk = tests[tf + k];
    for (int b = lane; b < dpad; b += 64) {
      u64* p = &sl[(long long)tk.x * dpad + b];
      if (b == tk.y) or_wg(p, 1ull << k);
      else and_wg(p, ~(1ull << k));
    }
  }
  if (tid < tc) {
    const int x = tests[tf + tid].x;
    atomicOr(&cur[x >> 5], 1u << (x & 31));
    atomicOr(&touched[x >> 5], 1u << (x & 31));
  }
  __syncthreads();

  unsigned long long revs = 0;
  for (int round = 0;; ++round) {
    const int slot = round & 1;
    if (wave == 0) {
      // item size: at most kSliceArcs arcs, small enough for about four
      // items per wave when the round has few arcs
      int arcs = 0;
      for (int v0 = 0; v0 < n; v0 += 64) {
        const int v = v0 + lane;
        if (v < n && bit_of(cur, v)) arcs += net.adj_off[v + 1] - net.adj_off[v];
      }
      for (int d = 32; d > 0; d >>= 1) arcs += __shfl_xor(arcs, d, 64);
      const int K = uniform_int(max(1, min(kSliceArcs, (arcs + 4 * nwaves - 1) / (4 * nwaves))));
      int base = 0, ibase = 0;
      for (int v0 = 0; v0 < n; v0 += 64) {
        const int v = v0 + lane;
        const bool f = v < n && bit_of(cur, v);
        const u64 mask = __ballot(f);
        const int rank = mbcnt(mask);
        const int items = f ? (net.adj_off[v + 1] - net.adj_off[v] + K - 1) / K : 0;
        const int inc = wave_inclusive_scan(items);
        if (f) {
          list[base + rank] = v;
          pre[base + rank] = ibase + inc - items;
        }
        base += __popcll(mask);
        ibase += __shfl(inc, 63, 64);
      }
      if (lane == 0) {
        pre[base] = ibase;
        ctl.m = base;
        ctl.total = ibase;
        ctl.k = K;
        ctl.item_ctr = 0;
        ctl.nchanged[slot] = 0;
      }
    } else {
      for (int i = tid - 64; i < nvw; i += nthreads - 64) nxt[i] = 0u;
    }
    if (nthreads == 64)
      for (int i = lane; i < nvw; i += 64) nxt[i] = 0u;
    __syncthreads();
    const int total = uniform_int(ctl.total);
    const int m = uniform_int(ctl.m);
    const int K = uniform_int(ctl.k);
    const u64 act = (u64)uniform_ll((long long)ctl.act);
    while (true) {
      int idx = 0;
      if (lane == 0) idx = atomicAdd(&ctl.item_ctr, 1);
      idx = uniform_int(__shfl(idx, 0, 64));
      if (idx >= total) break;
      int lo = 0, hi = m - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= idx) lo = mid;
        else hi = mid - 1;
      }
      const int t = uniform_int(list[lo]);
      const int s0 = net.adj_off[t] + (idx - uniform_int(pre[lo])) * K;
      const int s1 = min(s0 + K, net.adj_off[t + 1]);
      const int cnt = uniform_int(s1 - s0);
      // stage the slices of t and the values alive in some live test
      const u64* slt = sl + (long long)t * dpad;
      for (int i = lane; i < dpad; i += 64) {
        const u64 v = ld_wg(&slt[i]);
        stage[stage_pos(i)] = v;
        const u64 alive = __ballot((v & act) != 0ull);
        if (lane == 0) atw[i >> 6] = alive;
      }
      __builtin_amdgcn_wave_barrier();
      int4 e = make_int4(0, 0, 0, 0);
      if (lane < cnt) e = net.adj[s0 + lane];
      for (int i = 0; i < cnt; ++i) {
        const int u = __builtin_amdgcn_readlane(e.y, i);
        const int pos_t = __builtin_amdgcn_readlane(e.z, i);
        const long long o = (long long)__builtin_amdgcn_readlane(e.w, i) * 2 * W * dpad;
        const u64* plane_u = net.sup + o + (long long)(1 - pos_t) * W * dpad;
        const bool tu = uniform_int((int)bit_of(touched, u));
        const bool ch = slice_revise(plane_u, sl + (long long)u * dpad, stage, atw, W, dpad, act, tu,
                                     base_dom + (long long)u * W, all);
        ++revs;
        if (ch && lane == 0) slice_mark(nxt, &ctl, u, slot);
      }
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // Exact wipe-out check: a test wiped out at v only if v changed.
    u64 alive_all = ~0ull;
    for (int v = wave; v < n; v += nwaves) {
      if (!uniform_int((int)bit_of(nxt, v))) continue;
      u64 o = 0ull;
      for (int i = lane; i < dpad; i += 64) o |= ld_wg(&sl[(long long)v * dpad + i]);
      for (int d = 32; d > 0; d >>= 1) o |= __shfl_xor(o, d, 64);
      alive_all &= o;
    }
    if (lane == 0 && alive_all != ~0ull) atomicAnd(&ctl.act, alive_all);
    for (int i = tid; i < nvw; i += nthreads) touched[i] |= nxt[i];
    __syncthreads();
    const u64 act_now = (u64)uniform_ll((long long)ctl.act);
    if (uniform_int(ctl.nchanged[slot]) == 0 || act_now == 0ull) break;
    unsigned* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (lane == 0 && revs) atomicAdd(&ctl.revs, revs);
  __syncthreads();
  if (tid < tc) fail[perm ? perm[tf + tid] : tf + tid] = ((ctl.act >> tid) & 1ull) ? 0 : 1;
  if (tid == 0 && ctl.revs) atomicAdd(&revs_out[blockIdx.x % kStatSlots], ctl.revs);
}

// ---------------------------------------------------------------------------
// Wide bit-sliced singleton tests: 64 M tests per workgroup (M = 1, 2, 4).
// Within a round the workgroup walks the changed variables t one at a time:
// all threads stage t's masks into a shared, double-buffered LDS area (one
// barrier per t), then the waves take t's arcs u <- t from a shared counter.
// One support entry serves 64 M tests; the revision is the one of
// k_sliced_tests applied to M masks at once (same fixpoint argument, same
// exact end-of-round wipe-out check).
//
// The slices hold changes from the base level, not survivors: bit k of
// sl[(v * dpad + b) * M + m] flips the base state of value b of v in test
// 64 m + k.  For a base value a set bit means the test removed it; for a
// value outside the base domain a set bit means the test holds it, which
// only the test's own value can be (a test (x, a) reduces x to {a} whether
// or not a is in the base domain).  So value b is alive in test k iff the
// bit differs from "b in the base domain".  The region of a workgroup is all
// zero when the kernel starts (every test holds exactly the base domain), a
// variable's slice is written only when some test changes one of its values,
// and the workgroup zeroes the slices of the variables it wrote (its touched
// set) before it ends, so the buffer is all zero again for the next launch.
// The caller zeroes the buffer once, when it allocates it.
// ---------------------------------------------------------------------------
template <int M>
struct WideCtl {
  int m;
  int nchanged[2];
  int pad;
  int arc_ctr[2];  // next arc of the staged variable, per stage buffer
  u64 act[M];
  unsigned long long revs;
};

// Words of one stage buffer: kStageWord * W entries of M masks, or with
// nibble tables (nib) 16 W nibbles x 16 subsets of M masks.  With nb byte
// tables the first nb bytes of word 0 have one 256-subset table each in
// place of their 2 nb nibble tables.
__host__ __device__ inline long long wide_stage_words(int W, int M, bool nib, int nb = 0) {
  if (!nib) return (long long)kStageWord * W * M;
  return 256ll * nb * M + 16ll * (16 * W - 2 * nb) * M;
}

// Stage buffers: two (double buffered) without byte tables, one with them.
__host__ __device__ inline int wide_stage_buffers(int nb) { return nb > 0 ? 1 : 2; }

// LDS bytes: list [n], cur/nxt/touched bitmaps, up to 255 bytes that align
// the stage buffers to 256 bytes, the stage buffers and as many alive-value
// rows of W words.  Every nibble-table buffer is a multiple of 256 bytes, so
// each buffer and each half-word table in it stays aligned.
__host__ __device__ inline long long wide_lds_bytes(int n, int W, int M, bool nib = false, int nb = 0) {
  const long long nvw = (n + 31) / 32;
  long long b = 4ll * n + 12ll * nvw;
  b = ((b + 15) & ~15ll) + 256;
  const int nbuf = wide_stage_buffers(nb);
  b += nbuf * 8ll * wide_stage_words(W, M, nib, nb) + nbuf * 8ll * W;
  return b;
}

template <int M>
__device__ __forceinline__ bool any_mask(const u64 (&x)[M]) {
  u64 o = 0ull;
#pragma unroll
  for (int m = 0; m < M; ++m) o |= x[m];
  return o != 0ull;
}

// pend[m] &= ~(st[a][m] for every set bit a of bits), four bits per step;
// entry -1 of st is zero (exhausted bits).  Entries hold the M masks of one
// value together (two 16-byte reads for M = 4).
template <int M>
__device__ __forceinline__ void clear_supported_wide(u64 (&pend)[M], unsigned bits, const u64* st) {
  while (bits && any_mask<M>(pend)) {
    const int a0 = ffbl_raw(bits);
    bits &= bits - 1u;
    const int a1 = ffbl_raw(bits);
    bits &= bits - 1u;
    const int a2 = ffbl_raw(bits);
    bits &= bits - 1u;
    const int a3 = ffbl_raw(bits);
    bits &= bits - 1u;
    const u64* e0 = st + (long long)a0 * M;
    const u64* e1 = st + (long long)a1 * M;
    const u64* e2 = st + (long long)a2 * M;
    const u64* e3 = st + (long long)a3 * M;
#pragma unroll
    for (int m = 0; m < M; ++m) pend[m] &= ~((e0[m] | e1[m]) | (e2[m] | e3[m]));
  }
}

// Nibble tables ("four Russians"): entry (q, s) of a half-word table holds,
// for the 4 values 4q..4q+3 of the half, the OR of the masks of the values
// in subset s (s = 0 is zero).  One lookup covers every support of a row
// nibble.  The nibbles are taken in order, kNibStep per step, by every lane
// of the wave together: the lanes of a step read the same 16-entry blocks,
// which keeps LDS bank conflicts low, and the loads of a step are
// independent.  A lane stops once its row bits or its pending tests are
// exhausted.  Four per step measured 8.2 ms per 128,000-test pass against
// 8.8 ms for two and 8.4 ms for eight (and 10.0 ms for lanes that each jump
// to their own first non-empty nibble, the earlier form).
constexpr int kNibStep = 4;
constexpr unsigned kNibStepMask = kNibStep >= 8 ? 0xffffffffu : ((1u << (4 * kNibStep)) - 1u);

// Word of mask mm of subset `sub` in the table of nibble q.  With four masks
// an entry (32 bytes) is split in two 16-byte halves kept in separate
// 256-byte planes of the nibble's block: plane h holds masks 2h, 2h + 1 of
// the 16 subsets.  The 16 entries a wave's lanes pick from in one read then
// sit on distinct banks (one 16-byte slot each) instead of colliding every
// fourth entry.
template <int M>
__device__ __forceinline__ int nib_word(int q, int sub, int mm) {
  if constexpr (M == 4) return q * 64 + (mm >> 1) * 32 + sub * 2 + (mm & 1);
  else return (q * 16 + sub) * M + mm;
}

// p & ~(a | b | c | d): one three-input OR and one gfx950 three-input logic
// op (truth table 0x10: first operand set, the other two clear).
__device__ __forceinline__ unsigned andnot_or4(unsigned p, unsigned a, unsigned b, unsigned c, unsigned d) {
  return __builtin_amdgcn_bitop3_b32(p, a | b | c, d, 0x10);
}

// (a ^ b) & c, one three-input op per 32-bit half (truth table 0x28).
__device__ __forceinline__ u64 xor_and64(u64 a, u64 b, u64 c) {
  const unsigned lo = __builtin_amdgcn_bitop3_b32((unsigned)a, (unsigned)b, (unsigned)c, 0x28);
  const unsigned hi = __builtin_amdgcn_bitop3_b32((unsigned)(a >> 32), (unsigned)(b >> 32), (unsigned)(c >> 32), 0x28);
  return (u64)lo | ((u64)hi << 32);
}

// a & b as one op per 32-bit half with b uniform (a select of b would first
// copy b out of its scalar register).
__device__ __forceinline__ u64 and64(u64 a, u64 b) {
  const unsigned lo = __builtin_amdgcn_bitop3_b32((unsigned)a, (unsigned)b, 0u, 0xC0);
  const unsigned hi = __builtin_amdgcn_bitop3_b32((unsigned)(a >> 32), (unsigned)(b >> 32), 0u, 0xC0);
  return (u64)lo | ((u64)hi << 32);
}

// 32-bit LDS byte address of a pointer into shared memory.
__device__ __forceinline__ unsigned lds_addr(const void* p) {
  return (unsigned)(size_t)(__attribute__((address_space(3))) const char*)p;
}

typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const v4u lds_v4u;

__device__ __forceinline__ void split_masks(const u64 (&pend)[4], unsigned (&p)[8]) {
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    p[2 * m] = (unsigned)pend[m];
    p[2 * m + 1] = (unsigned)(pend[m] >> 32);
  }
}

__device__ __forceinline__ void join_masks(const unsigned (&p)[8], u64 (&pend)[4]) {
#pragma unroll
  for (int m = 0; m < 4; ++m) pend[m] = (u64)p[2 * m] | ((u64)p[2 * m + 1] << 32);
}

// Four masks, nibbles from 0, on the eight 32-bit halves p of the pending
// masks: the lookups of clear_supported_nib with the table addresses and the
// OR-and-clear written for the gfx950 ALU.  tab must be 256-byte aligned in
// LDS, so that the subset offset (16 bytes per subset) is ORed into the base
// with one three-input op ((x & vF0) | base, truth table 0xEA, vF0 = 0xF0)
// and the nibble's block offset rides in the read's immediate.  Per step of
// four lookups: 7 address ops and 16 mask ops, where the generic form takes
// 11 and 24.  live is the lane's any-pending flag: with kCheck a lane runs a
// step only while live (a live lane with no bits left in this half only
// looks up the zero subset) and live is recomputed after each step; without
// kCheck both steps run and live is computed once at the end (for the first
// half-word of a chunk, where some test of nearly every lane still lacks a
// support after 16 values).
template <bool kCheck>
__device__ __forceinline__ void nib4_run(unsigned (&p)[8], unsigned bits, const u64* tab, bool& live,
                                         unsigned vF0) {
  const unsigned base = lds_addr(tab);
#pragma unroll
  for (int q = 0; q < 8; q += 4) {
    if (kCheck && !live) break;
    v4u lo[4], hi[4];
    unsigned a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int sh = 4 * (q + i) - 4;
      const unsigned x = sh >= 0 ? (bits >> sh) : (bits << 4);
      a[i] = __builtin_amdgcn_bitop3_b32(x, vF0, base, 0xEA) + (q + i) * 512;
      lo[i] = ((lds_v4u*)(size_t)a[i])[0];  // masks 0, 1 of the subset
    }
    p[0] = andnot_or4(p[0], lo[0].x, lo[1].x, lo[2].x, lo[3].x);
    p[1] = andnot_or4(p[1], lo[0].y, lo[1].y, lo[2].y, lo[3].y);
    p[2] = andnot_or4(p[2], lo[0].z, lo[1].z, lo[2].z, lo[3].z);
    p[3] = andnot_or4(p[3], lo[0].w, lo[1].w, lo[2].w, lo[3].w);
    // the second plane's reads after the first plane's ORs: eight reads in
    // flight at once hold 32 registers against the 64-register budget
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) hi[i] = ((lds_v4u*)(size_t)a[i])[16];  // masks 2, 3, 256 bytes on
    p[4] = andnot_or4(p[4], hi[0].x, hi[1].x, hi[2].x, hi[3].x);
    p[5] = andnot_or4(p[5], hi[0].y, hi[1].y, hi[2].y, hi[3].y);
    p[6] = andnot_or4(p[6], hi[0].z, hi[1].z, hi[2].z, hi[3].z);
    p[7] = andnot_or4(p[7], hi[0].w, hi[1].w, hi[2].w, hi[3].w);
    bits &= ~(0xffffu << (4 * q));
    if (kCheck) live = ((p[0] | p[1] | p[2] | p[3]) | (p[4] | p[5] | p[6] | p[7])) != 0u;
  }
  if (!kCheck) live = ((p[0] | p[1] | p[2] | p[3]) | (p[4] | p[5] | p[6] | p[7])) != 0u;
}

// nib4_run on the 64-bit pending masks (used by the byte-table variants for
// their upper half-words).
template <bool kCheck>
__device__ __forceinline__ void clear_supported_nib4(u64 (&pend)[4], unsigned bits, const u64* tab) {
  unsigned p[8];
  split_masks(pend, p);
  bool live = any_mask<4>(pend);
  nib4_run<kCheck>(p, bits, tab, live, 0xF0u);
  join_masks(p, pend);
}

template <int M, int Q0 = 0, bool kCheck = true>
__device__ __forceinline__ void clear_supported_nib(u64 (&pend)[M], unsigned bits, const u64* tab) {
  if constexpr (M == 4 && Q0 == 0 && kNibStep == 4) {
    clear_supported_nib4<kCheck>(pend, bits, tab);
    return;
  }
  for (int q = Q0; q < 8 && bits && any_mask<M>(pend); q += kNibStep) {
    u64 acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = 0ull;
#pragma unroll
    for (int i = 0; i < kNibStep; ++i) {
      const unsigned s = (bits >> (4 * (q + i))) & 15u;
      const u64* e = tab + nib_word<M>(q + i, (int)s, 0);
#pragma unroll
      for (int m = 0; m < M; ++m) acc[m] |= e[nib_word<M>(0, 0, m)];
    }
    bits &= ~(kNibStepMask << (4 * q));
#pragma unroll
    for (int m = 0; m < M; ++m) pend[m] &= ~acc[m];
  }
}

// Byte tables: entry (p, s) of byte p of word 0 holds the OR of the masks of
// the values 8p + i for the set bits i of s, so that one lookup covers 8
// values where a nibble table covers 4.  A support is usually found among the
// first few dozen values of a row (at tightness 0.5 a test keeps about half
// the values of each changed neighbour), so the tables cover the start of
// word 0 only; nibble tables cover the rest.  With four masks an entry is
// split in two 16-byte halves in separate planes, as in the nibble tables.
template <int M>
__device__ __forceinline__ int byte_word(int p, int sub, int mm) {
  if constexpr (M == 4) return p * 1024 + (mm >> 1) * 512 + sub * 2 + (mm & 1);
  else return (p * 256 + sub) * M + mm;
}

// Byte lookups per step: 3 for 6 tables, else 2.
__host__ __device__ constexpr int byte_step(int nb) { return nb % 3 == 0 ? 3 : 2; }

// pend[m] &= ~(masks of the supports among values 0 .. 8 NB - 1 of word 0);
// returns the row bits above them.
template <int M, int NB>
__device__ __forceinline__ u64 clear_supported_bytes(u64 (&pend)[M], u64 r, const u64* btab) {
  constexpr int S = byte_step(NB);
#pragma unroll
  for (int p = 0; p < NB; p += S) {
    if (!(r >> (8 * p)) || !any_mask<M>(pend)) break;
    u64 acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = 0ull;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      if (p + i < NB) {
        const int sub = (int)((r >> (8 * (p + i))) & 255ull);
        const u64* e = btab + byte_word<M>(p + i, sub, 0);
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] |= e[byte_word<M>(0, 0, m)];
      }
    }
#pragma unroll
    for (int m = 0; m < M; ++m) pend[m] &= ~acc[m];
  }
  return r & ~((1ull << (8 * NB)) - 1ull);
}

// Nibble tables of a stage buffer with NB byte tables in front: nibble q of
// word 0 (q >= 2 NB) and every nibble of the later words sit where the
// nibble-only layout would put them, shifted back by the LDS the byte tables
// do not use.
template <int M, int NB>
__device__ __forceinline__ const u64* nib_base(const u64* stage) {
  return stage + (256ll * NB - 32ll * NB) * M;
}

// The M masks of one (variable, value) with plain vector loads (32 bytes for
// M = 4), issued together.
template <int M>
__device__ __forceinline__ void load_masks(const u64* p, u64 (&out)[M]) {
  if constexpr (M % 2 == 0) {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
#pragma unroll
    for (int h = 0; h < M / 2; ++h) {
      const ulonglong2 v = q[h];
      out[2 * h] = v.x;
      out[2 * h + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int m = 0; m < M; ++m) out[m] = p[m];
  }
}

template <int M, bool kNib, int NB>
__device__ __forceinline__ bool wide_revise(const u64* __restrict__ plane_u, u64* slu, const u64* stage,
                                            const u64* atw, int W, int dpad, const u64 (&act)[M],
                                            bool touched_u, const u64* __restrict__ base_u,
                                            const u64 (&all)[M]) {
  const int lane = lane_id();
  bool changed = false;
  u64 aa[M];  // live tests of the group
#pragma unroll
  for (int m = 0; m < M; ++m) aa[m] = all[m] & act[m];
  const bool aa_any = any_mask<M>(aa);
  // the subset mask of the table addresses in a vector register: a
  // three-input op reads one scalar operand, and the table base is one
  unsigned vF0;
  asm volatile("v_mov_b32 %0, 0xf0" : "=v"(vF0));
  const u64 atw0 = atw[0];
  const int lane_hi = lane & 32;
  // word 0 of the next chunk's rows is loaded one chunk ahead, so that its
  // latency overlaps this chunk's lookups
  u64 row_next = plane_u[lane];
  for (int j = 0; j < W; ++j) {
    const int b = j * 64 + lane;
    const u64 row0 = row_next;
    if (j + 1 < W) row_next = plane_u[b + 64];
    // flip: all ones when b is a base value (a set slice bit flips the base
    // state), as one sign-extending bit-field extract
    const int flip32 = __builtin_amdgcn_sbfe((int)(unsigned)(base_u[j] >> lane_hi), lane & 31, 1);
    const bool in = flip32 != 0;
    const u64 flip = (u64)(long long)flip32;
    u64 pend[M];
    bool has;  // some test still holds b
    if (touched_u) {
      // plain loads, all M masks at once: another wave may be removing
      // values of u at the same time, and a load that misses those removals
      // only makes pend a superset of the live tests (the alive sets only
      // shrink), which costs lookups but not exactness, since a removal takes
      // the change from the atomic's returned word
      u64 sv[M];
      load_masks<M>(&slu[(long long)b * M], sv);
#pragma unroll
      for (int m = 0; m < M; ++m) pend[m] = xor_and64(sv[m], flip, aa[m]);
      has = any_mask<M>(pend);
    } else {
      // pend = aa for a base value, nothing otherwise; the four-mask path
      // starts from aa itself, inside the lanes that hold b
#pragma unroll
      for (int m = 0; m < M; ++m) pend[m] = (M == 4 && kNib && NB == 0) ? aa[m] : and64(flip, aa[m]);
      has = in && aa_any;
    }
    const u64* nib = nib_base<M, NB>(stage);
    bool live = has;  // some lane still holds a test without a support
    if constexpr (M == 4 && kNib && NB == 0) {
      // the four-mask path keeps live current after every step, so that
      // each step's test is computed once
      if (has) {
        unsigned p[8];
        split_masks(pend, p);
        const u64 r = row0 & atw0;
        nib4_run<false>(p, (unsigned)r, stage, live, vF0);
        nib4_run<true>(p, (unsigned)(r >> 32), stage + 128ll * M, live, vF0);
        for (int w = 1; w < W && live; ++w) {
          const u64 rw = plane_u[(long long)w * dpad + b] & atw[w];
          const u64* tb = stage + 256ll * w * M;
          nib4_run<true>(p, (unsigned)rw, tb, live, vF0);
          nib4_run<true>(p, (unsigned)(rw >> 32), tb + 128ll * M, live, vF0);
        }
        join_masks(p, pend);
      }
    } else if (has) {
      const u64 r = row0 & atw0;
      if (NB > 0) {
        const u64 rest = clear_supported_bytes<M, NB>(pend, r, stage);
        if (NB < 4) clear_supported_nib<M, 2 * NB>(pend, (unsigned)rest, nib);
        clear_supported_nib<M, (NB > 4 ? 2 * NB - 8 : 0)>(pend, (unsigned)(rest >> 32), nib + 128ll * M);
      } else if (kNib) {
        clear_supported_nib<M, 0, false>(pend, (unsigned)r, stage);
        clear_supported_nib<M>(pend, (unsigned)(r >> 32), stage + 128ll * M);
      } else {
        clear_supported_wide<M>(pend, (unsigned)r, stage + (long long)1 * M);
        clear_supported_wide<M>(pend, (unsigned)(r >> 32), stage + (long long)34 * M);
      }
    }
    if constexpr (!(M == 4 && kNib && NB == 0)) {
      for (int w = 1; w < W && any_mask<M>(pend); ++w) {
        const u64 r = plane_u[(long long)w * dpad + b] & atw[w];
        if (kNib) {
          const u64* tb = nib + 256ll * w * M;
          clear_supported_nib<M>(pend, (unsigned)r, tb);
          clear_supported_nib<M>(pend, (unsigned)(r >> 32), tb + 128ll * M);
        } else {
          const u64* st = stage + (long long)w * kStageWord * M;
          clear_supported_wide<M>(pend, (unsigned)r, st + (long long)1 * M);
          clear_supported_wide<M>(pend, (unsigned)(r >> 32), st + (long long)34 * M);
        }
      }
      live = any_mask<M>(pend);
    }
    if (live) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (pend[m]) {
          // removal: set the bit of a base value, clear that of a value a
          // test made present
          if (in) {
            const u64 old = or_wg(&slu[(long long)b * M + m], pend[m]);
            if (pend[m] & ~old) changed = true;
          } else {
            const u64 old = and_wg(&slu[(long long)b * M + m], ~pend[m]);
            if (pend[m] & old) changed = true;
          }
        }
      }
    }
  }
  return __ballot(changed) != 0ull;
}

// Most threads per workgroup of k_sliced_wide, and the waves per SIMD its
// register budget is set for: 8 (64 VGPRs) keeps two 1024-thread workgroups
// on every CU.  With 4 masks this spills a few registers and is still the
// fastest choice measured (11.5 ms against 13.6 ms at 512 threads and 80
// VGPRs for the 128,000-test pass).
constexpr int kWideMaxThreads = 1024;
constexpr int kWideWavesPerSimd = 8;

template <int M, bool kNib, int NB>
__global__ void __launch_bounds__(kWideMaxThreads, kWideWavesPerSimd) k_sliced_wide(DevNet net, const u64* __restrict__ base_dom,
                              const int2* __restrict__ tests, const int* __restrict__ gfirst,
                              const int* __restrict__ gorder, long long g0, const int* __restrict__ perm,
                              unsigned char* fail, unsigned long long* revs_out, u64* slices) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WideCtl<M> ctl;
  const int n = net.n, W = net.W, dpad = net.dpad;
  const int nvw = (n + 31) / 32;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int lane = lane_id();
  const int wave = uniform_int(tid >> 6), nwaves = nthreads >> 6;
  const int B = gridDim.x, bid = blockIdx.x;
  const int xcd = bid & 7, q = B >> 3, rem = B & 7;
  // slot of this workgroup (consecutive groups on one XCD), and its group
  const long long gs = g0 + (long long)xcd * q + (xcd < rem ? xcd : rem) + (bid >> 3);
  const long long g = gorder ? gorder[gs] : gs;
  const int tf = gfirst[g];
  const int tc = uniform_int(gfirst[g + 1] - tf);
  u64 all[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int k = tc - 64 * m;
    all[m] = k >= 64 ? ~0ull : (k > 0 ? ((1ull << k) - 1ull) : 0ull);
  }
  u64* sl = slices + (long long)bid * n * dpad * M;

  int* list = (int*)smem;
  unsigned* cur = (unsigned*)(list + n);
  unsigned* nxt = cur + nvw;
  unsigned* touched = nxt + nvw;
  // the stage buffers start 256-byte aligned in LDS: the four-mask nibble
  // lookups OR the subset offset into the table base
  const unsigned smem0 = lds_addr(smem);
  const unsigned head_raw = (unsigned)((4ll * n + 12ll * nvw + 15) & ~15ll);
  const long long head = (long long)(((smem0 + head_raw + 255u) & ~255u) - smem0);
  u64* stage0 = (u64*)(smem + head);
  static_assert(NB == 0 || (kNib && NB <= 6), "byte tables go with nibble tables");
  const long long stage_words = wide_stage_words(W, M, kNib, NB);
  constexpr int nbuf = NB > 0 ? 1 : 2;  // double buffer without byte tables
  u64* atw0 = stage0 + nbuf * stage_words;

  for (int i = tid; i < nvw; i += nthreads) {
    cur[i] = 0u;
    nxt[i] = 0u;
    touched[i] = 0u;
  }
  // zero entries of every stage buffer (single-value layout only)
  for (int i = tid; !kNib && i < nbuf * 2 * W * M; i += nthreads) {
    const int m = i % M, k = i / M;  // k = buffer * 2W + entry
    const int buf = k / (2 * W), e = k % (2 * W);
    stage0[buf * stage_words + ((long long)(e >> 1) * kStageWord + 33 * (e & 1)) * M + m] = 0ull;
  }
  if (tid == 0) {
#pragma unroll
    for (int m = 0; m < M; ++m) ctl.act[m] = all[m];
    ctl.revs = 0;
  }
  __syncthreads();
  // test k reduces its variable x_k to {a_k}: flip every base value but a_k,
  // and a_k if it is not a base value.  The tests of a group come in runs of
  // one variable (the list is sorted by variable), so word (x, b, m) of the
  // slices is built whole by one thread from the 64 tests of mask m and
  // stored once: no atomics (one per flipped value and test, 64 to a word,
  // took a sixth of the pass).
  for (int w = tid; w < dpad * M; w += nthreads) {
    const int m = (w >> 6) % M;  // uniform: nthreads is a multiple of 64
    const int b = ((w >> 6) / M) * 64 + (w & 63);
    const int kend = min(tc, 64 * (m + 1));
    int x = -1;
    bool in = false;
    u64 word = 0ull;
    for (int k = 64 * m; k < kend; ++k) {
      const int2 tk = tests[tf + k];
      if (tk.x != x) {
        if (word) sl[((long long)x * dpad + b) * M + m] = word;
        x = tk.x;
        word = 0ull;
        in = (base_dom[(long long)x * W + (b >> 6)] >> (b & 63)) & 1ull;
      }
      if (in != (b == tk.y)) word |= 1ull << (k & 63);
    }
    if (word) sl[((long long)x * dpad + b) * M + m] = word;
  }
  if (tid < tc) {
    const int x = tests[tf + tid].x;
    atomicOr(&cur[x >> 5], 1u << (x & 31));
    atomicOr(&touched[x >> 5], 1u << (x & 31));
  }
  __syncthreads();

  unsigned long long revs = 0;
  for (int round = 0;; ++round) {
    const int slot = round & 1;
    if (wave == 0) {
      int base = 0;
      for (int v0 = 0; v0 < n; v0 += 64) {
        const int v = v0 + lane;
        const bool f = v < n && bit_of(cur, v);
        const u64 mask = __ballot(f);
        if (f) list[base + mbcnt(mask)] = v;
        base += __popcll(mask);
      }
      if (lane == 0) {
        ctl.m = base;
        ctl.nchanged[slot] = 0;
      }
    } else {
      for (int i = tid - 64; i < nvw; i += nthreads - 64) nxt[i] = 0u;
    }
    if (nthreads == 64)
      for (int i = lane; i < nvw; i += 64) nxt[i] = 0u;
    __syncthreads();
    const int mcount = uniform_int(ctl.m);
    u64 act[M];
#pragma unroll
    for (int m = 0; m < M; ++m) act[m] = (u64)uniform_ll((long long)ctl.act[m]);
    for (int i = 0; i < mcount; ++i) {
      const int t = uniform_int(list[i]);
      // one buffer: every wave finishes the previous variable's arcs first
      if (nbuf == 1 && i > 0) __syncthreads();
      u64* stage = stage0 + (i & (nbuf - 1)) * stage_words;
      u64* atw = atw0 + (i & (nbuf - 1)) * W;
      // stage t: thread v holds value v's M masks; one wave covers one word
      const u64* slt = sl + (long long)t * dpad * M;
      const u64* base_t = base_dom + (long long)t * W;
      // masks of value v of t: the live tests that still hold it.  Plain
      // loads, issued together: a wave still revising the previous
      // variable's arcs may remove values of t meanwhile, and a stage that
      // misses such a removal only keeps a support too many; t is then
      // marked changed and staged again next round.
      auto alive_of = [&](int v, int mm) -> u64 {
        const u64 flip = ((base_t[v >> 6] >> (v & 63)) & 1ull) ? ~0ull : 0ull;
        const int kk = tc - 64 * mm;
        const u64 all_m = kk >= 64 ? ~0ull : (kk > 0 ? ((1ull << kk) - 1ull) : 0ull);
        return all_m & (slt[(long long)v * M + mm] ^ flip);
      };
      // the same from a slice word already loaded
      auto alive_bits = [&](u64 raw, int v, int mm) -> u64 {
        const u64 flip = ((base_t[v >> 6] >> (v & 63)) & 1ull) ? ~0ull : 0ull;
        const int kk = tc - 64 * mm;
        const u64 all_m = kk >= 64 ? ~0ull : (kk > 0 ? ((1ull << kk) - 1ull) : 0ull);
        return all_m & (raw ^ flip);
      };
      // every wave finished the arcs of this buffer's previous variable
      // before the last barrier
      if (tid == 0) ctl.arc_ctr[i & 1] = 0;
      if (kNib) {
        // nibble tables: item (q, m) ORs the 16 subsets of values 4q..4q+3
        // (nibbles from 2 NB on); byte tables: item (p, m, h) fills the 16
        // subsets of byte p whose high nibble is h
        u64* nib = stage + (256ll * NB - 32ll * NB) * M;
        const int nib_items = (16 * W - 2 * NB) * M;
        for (int it = tid; it < nib_items + NB * M * 16; it += nthreads) {
          if (it < nib_items) {
            const int q = 2 * NB + it / M, mm = it % M;
            u64 x[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = slt[(long long)(4 * q + i) * M + mm];
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = alive_bits(x[i], 4 * q + i, mm);
#pragma unroll
            for (int sub = 0; sub < 16; ++sub) {
              u64 o = 0ull;
#pragma unroll
              for (int i = 0; i < 4; ++i)
                if (sub & (1 << i)) o |= x[i];
              nib[nib_word<M>(q, sub, mm)] = o;
            }
          } else {
            const int k = it - nib_items;
            const int h = k & 15, mm = (k >> 4) % M, p = (k >> 4) / M;
            u64 x[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) x[i] = alive_of(8 * p + i, mm);
            u64 hi = 0ull;
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (h & (1 << i)) hi |= x[4 + i];
#pragma unroll
            for (int lo = 0; lo < 16; ++lo) {
              u64 o = hi;
#pragma unroll
              for (int i = 0; i < 4; ++i)
                if (lo & (1 << i)) o |= x[i];
              stage[byte_word<M>(p, (h << 4) | lo, mm)] = o;
            }
          }
        }
      }
      // the alive-any words from the waves after the table items when there
      // are enough of them, so that the two sets of slice reads overlap
      const int items = kNib ? (16 * W - 2 * NB) * M + NB * M * 16 : 0;
      const int abase = ((items + 63) & ~63) + dpad <= nthreads ? ((items + 63) & ~63) : 0;
      for (int v = tid - abase; tid >= abase && v < dpad; v += nthreads) {
        u64 alive = 0ull;
        u64* dst = stage + (long long)stage_pos(v) * M;
        u64 sv[M];
        load_masks<M>(&slt[(long long)v * M], sv);
        const u64 flip = ((base_t[v >> 6] >> (v & 63)) & 1ull) ? ~0ull : 0ull;
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const u64 x = all[m] & (sv[m] ^ flip);
          if (!kNib) dst[m] = x;
          alive |= x & act[m];
        }
        const u64 word = __ballot(alive != 0ull);
        if (lane == 0) atw[v >> 6] = word;
      }
      __syncthreads();
      const int a0 = net.adj_off[t], deg = net.adj_off[t + 1] - a0;
      // waves take t's arcs one at a time from a shared counter
      while (true) {
        int j = 0;
        if (lane == 0) j = atomicAdd(&ctl.arc_ctr[i & 1], 1);
        j = __builtin_amdgcn_readfirstlane(j);
        if (j >= deg) break;
        const int4 e = net.adj[a0 + j];
        const int u = uniform_int(e.y);
        const int pos_t = uniform_int(e.z);
        const long long o = (long long)uniform_int(e.w) * 2 * W * dpad;
        const u64* plane_u = net.sup + o + (long long)(1 - pos_t) * W * dpad;
        const bool tu = uniform_int((int)bit_of(touched, u));
        const bool ch = wide_revise<M, kNib, NB>(plane_u, sl + (long long)u * dpad * M, stage, atw, W, dpad, act,
                                             tu, base_dom + (long long)u * W, all);
        ++revs;
        if (ch && lane == 0) {
          const unsigned bit = 1u << (u & 31);
          const unsigned old = atomicOr(&nxt[u >> 5], bit);
          if (!(old & bit)) atomicAdd(&ctl.nchanged[slot], 1);
        }
      }
    }
    __syncthreads();
    // exact wipe-out check: a test wiped out at v only if v changed
    u64 alive_all[M];
#pragma unroll
    for (int m = 0; m < M; ++m) alive_all[m] = ~0ull;
    for (int v = wave; v < n; v += nwaves) {
      if (!uniform_int((int)bit_of(nxt, v))) continue;
      // a lane reads the M masks of its values together, every read of the
      // variable in flight at once
      u64 o[M];
#pragma unroll
      for (int m = 0; m < M; ++m) o[m] = 0ull;
      for (int b = lane; b < dpad; b += 64) {
        const u64 flip = ((base_dom[(long long)v * W + (b >> 6)] >> (b & 63)) & 1ull) ? ~0ull : 0ull;
        u64 sv[M];
        load_masks<M>(&sl[((long long)v * dpad + b) * M], sv);
#pragma unroll
        for (int m = 0; m < M; ++m) o[m] |= all[m] & (sv[m] ^ flip);
      }
#pragma unroll
      for (int m = 0; m < M; ++m) alive_all[m] &= wave_or64_dpp(o[m]);
    }
    if (lane == 0) {
#pragma unroll
      for (int m = 0; m < M; ++m)
        if (alive_all[m] != ~0ull) atomicAnd(&ctl.act[m], alive_all[m]);
    }
    for (int i = tid; i < nvw; i += nthreads) touched[i] |= nxt[i];
    __syncthreads();
    u64 act_any = 0ull;
#pragma unroll
    for (int m = 0; m < M; ++m) act_any |= (u64)uniform_ll((long long)ctl.act[m]);
    if (uniform_int(ctl.nchanged[slot]) == 0 || act_any == 0ull) break;
    unsigned* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (lane == 0 && revs) atomicAdd(&ctl.revs, revs);
  __syncthreads();
  if (tid < tc)
    fail[perm ? perm[tf + tid] : tf + tid] = ((ctl.act[tid >> 6] >> (tid & 63)) & 1ull) ? 0 : 1;
  if (tid == 0 && ctl.revs) atomicAdd(&revs_out[blockIdx.x % kStatSlots], ctl.revs);
  // leave the region all zero: clear the slices of every variable written
  for (int v = wave; v < n; v += nwaves) {
    if (!uniform_int((int)bit_of(touched, v))) continue;
    u64* p = sl + (long long)v * dpad * M;
    for (int i = lane; i < dpad * M; i += 64) p[i] = 0ull;
  }
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
  int status;  // 0 init, 1 running, 2 solution, 3 unsat or enumeration done, 4 limit reached
  int depth;
  long long npos;
  long long nneg;
  long long revisions;
  int trace_len;
  int steps;
  long long nsol;  // solutions found so far
};

struct NodeRec {
  int positive;
  int var;
  int value;
  int consistent;
  u64 dom_hash;
};

// A copy of a network laid out as it sits in LDS behind the working memory:
// adjacency offsets, adjacency, constraint metadata, support planes, tables.
// The value stride of the planes and tables is cut from dpad (64 W) to dc,
// the largest domain, and the constraint offsets follow that stride.  The
// planes are followed by 64 W words of zeros: a target-side revision loads
// whole 64-value chunks of a row, and the lanes past dc (values that are
// never live) must still read inside the image.  bytes = 0: no image.
struct LdsNet {
  const uint4* src;  // device copy of the image
  int bytes;         // multiple of 16
  int dc;
  int o_adj, o_con, o_sup, o_tsup;  // byte offsets in the image
};

// Copies the image behind the working memory (at dst) and returns the
// network whose arrays are those LDS copies.  Called by every thread.
__device__ __forceinline__ DevNet wg_load_net(const DevNet& g, const LdsNet& L, char* dst) {
  uint4* d4 = (uint4*)dst;
  for (int i = threadIdx.x; i < (L.bytes >> 4); i += blockDim.x) d4[i] = L.src[i];
  __syncthreads();
  DevNet ln = g;
  ln.dpad = L.dc;
  ln.adj_off = (const int*)dst;
  ln.adj = (const int4*)(dst + L.o_adj);
  ln.con = (const ConMeta*)(dst + L.o_con);
  ln.sup = (const u64*)(dst + L.o_sup);
  ln.tsup = (const u64*)(dst + L.o_tsup);
  return ln;
}

struct MacArgs {
  MacState* state;
  u64* trail;          // [n + 1][n][W]; trail[0] holds the start domains
  int2* stack;         // [n]
  int* solution;       // [sol_cap][n] or NULL: solution k at solution + k n
  long long sol_cap;   // solutions the buffer holds (0 counts as 1)
  // 0 or 1: stop at the first solution; k > 1: count every solution, treat it
  // as a failed node and go on, until the tree is exhausted or k are found
  long long max_solutions;
  NodeRec* trace;      // [trace_cap] or NULL
  int trace_cap;
  long long node_limit;   // total (0 = none)
  long long step_budget;  // per launch (0 = none)
  u64 deadline;           // wall_clock64 ticks (0 = none)
  u64* root_out;          // [n][W] root-fixpoint domains, or NULL
  const int* nbr;         // [n] neighbour counts: dom/deg ordering; NULL = dom-min
  // dom/wdeg ordering (all NULL otherwise); fixpoints then follow AC3's
  // schedule so that the weights grow as the reference's do
  float* weight;          // [ncons] constraint weights, start at 1
  long long* stamp_tab;   // [ncons] constraint stamps (absolute times), start at 0
  long long* clock;       // [1] stamp clock, starts at 1
};

enum { MAC_INIT = 0, MAC_RUNNING = 1, MAC_SAT = 2, MAC_UNSAT = 3, MAC_LIMIT = 4 };

__device__ __forceinline__ void wg_trace(const DevNet& net, WgMem mem, WgCtl* ctl, const MacArgs& A, int positive,
                         int var, int val, bool ok) {
  if (!A.trace) return;
  const int tl = uniform_int(ctl->trace_len);
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

// ---------------------------------------------------------------------------
// AC3 in the reference's schedule (AC3::enforce, src/AC3.cpp:99-141, with the
// duplicate-free FIFO var_que, src/Solver.cpp:180-218), run by one wave over
// the workgroup's domains.  The closure equals that of any other schedule;
// what the order fixes is which constraint empties a domain first, and that
// constraint's weight is raised (AC3.cpp:122-126): the input of the dom/wdeg
// ordering.
//
// The queue (ring of n + 1), the queued flags and the variable stamps use the
// workgroup's arc-list arrays (pre, nxt, list), which this schedule does not
// otherwise need.  Variable stamps restart at 0 on each call, as in the
// reference.  Constraint stamps stay in global memory as absolute times: a
// call starts at time base = *clock, reads a stamp s as s - base when
// s > base and as 0 otherwise (the reference's reset to 0, without touching
// every constraint), and leaves *clock past its last stamp.
// Lane 0 does every load and store of the control state and broadcasts what
// it read, so the same code is correct whether the arrays are in LDS or in
// global scratch; control values are wave-uniform.
// seed: one variable, or -1 for all variables in id order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane0_load(const int* p) {
  int v = 0;
  if (lane_id() == 0) v = *p;
  return __builtin_amdgcn_readlane(v, 0);
}

__device__ __forceinline__ bool wave_ac3(const DevNet& net, WgMem mem, const MacArgs& A, int seed,
                         unsigned long long* revs) {
  const int n = net.n;
  const int lane = lane_id();
  const int nvw = (n + 31) / 32;
  const int qcap = n + 1;
  int* q = mem.pre;
  unsigned* inq = mem.nxt;
  int* sv = mem.list;
  long long base = 0;
  if (lane == 0) base = *A.clock;
  base = (long long)readlane64((u64)base, 0);
  int front = 0, rear = 0, t = 0;
  if (seed >= 0) {
    for (int v = lane; v < n; v += 64) sv[v] = 0;
    for (int i = lane; i < nvw; i += 64) inq[i] = 0u;
  } else {
    for (int v = lane; v < n; v += 64) {
      q[v] = v;
      sv[v] = v + 1;
    }
    for (int i = lane; i < nvw; i += 64) {
      const int k = n - 32 * i;
      inq[i] = k >= 32 ? 0xffffffffu : ((1u << k) - 1u);
    }
    rear = n;
    t = n;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
  auto insert = [&](int v) {
    const int word = lane0_load((const int*)&inq[v >> 5]);
    if (!((word >> (v & 31)) & 1)) {
      if (lane == 0) {
        q[rear] = v;
        inq[v >> 5] = (unsigned)word | (1u << (v & 31));
      }
      rear = rear + 1 == qcap ? 0 : rear + 1;
    }
    ++t;
    if (lane == 0) sv[v] = t;
  };
  if (seed >= 0) insert(seed);
  bool ok = true;
  unsigned long long nrev = 0;
  while (ok && front != rear) {
    const int x = lane0_load(&q[front]);
    front = front + 1 == qcap ? 0 : front + 1;
    if (lane == 0) inq[x >> 5] &= ~(1u << (x & 31));
    const int k1 = mem.aoff[x + 1];
    for (int k = mem.aoff[x]; k < k1; ++k) {
      const int c = uniform_int(mem.adj[k].x);
      long long sc = 0;
      if (lane == 0) sc = A.stamp_tab[c];
      sc = (long long)readlane64((u64)sc, 0);
      const int stc = sc > base ? (int)(sc - base) : 0;
      if (!(lane0_load(&sv[x]) > stc)) continue;
      const ConMeta* mp = net.con + c;
      const int arity = mp->arity;
      for (int i = 0; i < arity; ++i) {
        const int y = mp->scope[i];
        if (bit_of(mem.asg, y)) continue;
        bool aa = false;
        for (int j = 0; j < arity; ++j) {
          const int z = mp->scope[j];
          if (z != x && lane0_load(&sv[z]) > stc) aa = true;
        }
        if (y == x && !aa) continue;
        bool wiped = false;
        int removed = 0;
        if (mp->kind == 0) {
          removed = revise_binary_at(net, mem.dom, mem.size, mp->off, y, mp->scope[1 - i], 1 - i, &wiped);
        } else {
          revise_table(net, mem.dom, mem.size, c, -1, nullptr, [&](int, int r) { removed += r; }, &wiped, y);
        }
        ++nrev;
        if (removed > 0) {
          if (wiped) {
            if (lane == 0) A.weight[c] += 1.0f;
            ok = false;
            break;
          }
          insert(y);
        }
      }
      if (!ok) break;
      ++t;
      if (lane == 0) A.stamp_tab[c] = base + t;
    }
  }
  if (lane == 0) *A.clock = base + t + 1;
  *revs += nrev;
  return ok;
}

// AC3-scheduled fixpoint for the whole workgroup: wave 0 runs it, the other
// waves wait at the barrier.
__device__ __forceinline__ bool wg_ac3(const DevNet& net, WgMem mem, WgCtl* ctl, const MacArgs& A, int seed) {
  __syncthreads();
  if (threadIdx.x < 64) {
    unsigned long long revs = 0;
    const bool ok = wave_ac3(net, mem, A, seed, &revs);
    if (threadIdx.x == 0) {
      ctl->ok = ok ? 1 : 0;
      ctl->revs += revs;
    }
  }
  __syncthreads();
  return uniform_int(ctl->ok) != 0;
}

// MAC::select_var VRH_DOM_WDEG_MIN (src/MAC.cpp:386-411): for each undecided
// x, x_w = the sum of the weights of the constraints on x that have another
// undecided variable; key = -1 when |D(x)| = 1 or x_w = 0, else |D(x)| / x_w
// in double (IEEE division, as on the host); the lowest id among the smallest
// keys.  Non-negative doubles order as their bit patterns, -1 maps to 0.
__device__ __forceinline__ int wg_select_wdeg(const DevNet& net, WgMem mem, WgCtl* ctl, const MacArgs& A) {
  const int n = net.n;
  u64 best = ~0ull;
  int bv = n;
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    if (bit_of(mem.asg, v)) continue;
    double xw = 0.0;
    const int k1 = mem.aoff[v + 1];
    for (int k = mem.aoff[v]; k < k1; ++k) {
      const int4 e = mem.adj[k];
      int cnt = 0;
      if (e.y >= 0) {
        cnt = 1 + (bit_of(mem.asg, e.y) ? 0 : 1);
      } else {
        const ConMeta* mp = net.con + e.x;
        for (int i = 0; i < mp->arity; ++i) cnt += bit_of(mem.asg, mp->scope[i]) ? 0 : 1;
      }
      if (cnt > 1) xw += (double)A.weight[e.x];
    }
    const int sz = mem.size[v];
    const double key = (sz == 1 || xw == 0.0) ? -1.0 : (double)sz / xw;
    const u64 k = key < 0.0 ? 0ull : (u64)__double_as_longlong(key);
    if (k < best) {
      best = k;
      bv = v;
    }
  }
  const u64 m = wg_reduce_min_u64(ctl, best);
  return (int)wg_reduce_min_u64(ctl, best == m ? (u64)bv : ~0ull);
}

__device__ __forceinline__ void wg_mac_run(const DevNet& net, WgMem mem, WgCtl* ctl, const MacArgs& A) {
  const int n = net.n, W = net.W;
  const int nvw = (n + 31) / 32;
  const long long nw = (long long)n * W;
  // Search control lives in registers, identical in every thread (all values
  // below are wave-uniform); LDS only carries what thread 0 publishes.
  int status = A.state->status;
  int depth = A.state->depth;
  long long npos = A.state->npos;
  long long nneg = A.state->nneg;
  long long nsol = A.state->nsol;
  long long nodes_here = 0;
  if (threadIdx.x == 0) {
    ctl->trace_len = A.state->trace_len;
    ctl->revs = 0;
  }
  status = uniform_int(status);
  depth = uniform_int(depth);
  npos = uniform_ll(npos);
  nneg = uniform_ll(nneg);
  nsol = uniform_ll(nsol);
  __syncthreads();
  if (status == MAC_INIT) {
    wg_copy_words(mem.dom, A.trail, nw);
    wg_clear_bitmap(mem.asg, nvw);
    __syncthreads();
    wg_recount(mem, n, W);
    wg_fill_bitmap(mem.cur, n);
    __syncthreads();
    const bool ok = A.weight ? wg_ac3(net, mem, ctl, A, -1) : wg_fixpoint<4>(net, mem, ctl, false);
    if (ok) wg_copy_words(A.trail, mem.dom, nw);
    if (A.root_out) wg_copy_words(A.root_out, mem.dom, nw);
    status = ok ? MAC_RUNNING : MAC_UNSAT;
    depth = 0;
    __syncthreads();
  } else if (status == MAC_RUNNING) {
    wg_copy_words(mem.dom, A.trail + depth * nw, nw);
    wg_clear_bitmap(mem.asg, nvw);
    __syncthreads();
    wg_recount(mem, n, W);
    if (threadIdx.x == 0)
      for (int i = 0; i < depth; ++i) mem.asg[A.stack[i].x >> 5] |= 1u << (A.stack[i].x & 31);
    __syncthreads();
  }
  while (status == MAC_RUNNING) {
    const bool over_nodes = A.node_limit > 0 && npos + nneg >= A.node_limit;
    const bool over_time = A.deadline && uniform_ll((long long)wall_clock64()) > (long long)A.deadline;
    if (over_nodes || over_time || (A.step_budget > 0 && nodes_here >= A.step_budget)) {
      // suspend (step budget) or stop (limits) at a consistent state of depth `depth`
      wg_copy_words(A.trail + depth * nw, mem.dom, nw);
      if (over_nodes || over_time) status = MAC_LIMIT;
      __syncthreads();
      break;
    }
    // select: min key over undecided variables, lowest id on ties.  dom-min:
    // key = |D|; dom/deg: |D| / neighbours (integer), -1 without neighbours,
    // biased by one so that the packed key stays unsigned.
    int x;
    if (A.weight) {
      x = wg_select_wdeg(net, mem, ctl, A);
    } else {
      u64 key = ~0ull;
      for (int v = threadIdx.x; v < n; v += blockDim.x)
        if (!bit_of(mem.asg, v)) {
          unsigned hi = (unsigned)mem.size[v];
          if (A.nbr) {
            const int g = A.nbr[v];
            hi = g == 0 ? 0u : (unsigned)(mem.size[v] / g) + 1u;
          }
          const u64 k = ((u64)hi << 32) | (unsigned)v;
          key = k < key ? k : key;
        }
      key = wg_reduce_min_u64(ctl, key);
      x = (int)(key & 0xffffffffu);
    }
    // min value of x (every thread scans the same LDS words)
    int a = -1;
    for (int w = 0; w < W && a < 0; ++w) {
      const u64 dw = mem.dom[(long long)x * W + w];
      if (dw) a = w * 64 + (int)__builtin_ctzll(dw);
    }
    a = uniform_int(a);
    if (a < 0) {
      // x has an empty domain that no revision reached (no constraints, an
      // empty initial domain): the search ends without a solution, with no
      // node recorded, as the oracle's restatement does
      status = MAC_UNSAT;
      __syncthreads();
      break;
    }
    wg_copy_words(A.trail + depth * nw, mem.dom, nw);
    __syncthreads();
    if (threadIdx.x < W) mem.dom[(long long)x * W + threadIdx.x] = (threadIdx.x == (a >> 6)) ? (1ull << (a & 63)) : 0ull;
    wg_clear_bitmap(mem.cur, nvw);
    __syncthreads();
    if (threadIdx.x == 0) {
      A.stack[depth] = make_int2(x, a);
      mem.size[x] = 1;
      mem.asg[x >> 5] |= 1u << (x & 31);
      mem.cur[x >> 5] |= 1u << (x & 31);
    }
    ++depth;
    ++npos;
    ++nodes_here;
    __syncthreads();
    bool ok = A.weight ? wg_ac3(net, mem, ctl, A, x) : wg_fixpoint<4>(net, mem, ctl, false);
    wg_trace(net, mem, ctl, A, 1, x, a, ok);
    if (ok && depth == n) {
      if (A.solution && nsol < (A.sol_cap > 0 ? A.sol_cap : 1))
        for (int v = threadIdx.x; v < n; v += blockDim.x) {
          int sv = -1;
          for (int w = 0; w < W && sv < 0; ++w) {
            const u64 dw = mem.dom[(long long)v * W + w];
            if (dw) sv = w * 64 + (int)__builtin_ctzll(dw);
          }
          A.solution[nsol * n + v] = sv;
        }
      ++nsol;
      if (A.max_solutions <= 1 || nsol >= A.max_solutions) {
        status = MAC_SAT;
        __syncthreads();
        break;
      }
      ok = false;  // enumeration: go on as after a failed node
      __syncthreads();
    }
    while (!ok && depth > 0) {
      --depth;
      const int2 d = A.stack[depth];
      const int dx = uniform_int(d.x), dv = uniform_int(d.y);
      wg_copy_words(mem.dom, A.trail + depth * nw, nw);
      __syncthreads();
      wg_recount(mem, n, W);
      __syncthreads();
      if (threadIdx.x == 0) {
        mem.asg[dx >> 5] &= ~(1u << (dx & 31));
        mem.dom[(long long)dx * W + (dv >> 6)] &= ~(1ull << (dv & 63));
        mem.size[dx] -= 1;
      }
      ++nneg;
      ++nodes_here;
      wg_clear_bitmap(mem.cur, nvw);
      __syncthreads();
      if (uniform_int(mem.size[dx]) == 0) {
        ok = false;
      } else {
        if (threadIdx.x == 0) mem.cur[dx >> 5] |= 1u << (dx & 31);
        __syncthreads();
        ok = A.weight ? wg_ac3(net, mem, ctl, A, dx) : wg_fixpoint<4>(net, mem, ctl, false);
      }
      wg_trace(net, mem, ctl, A, 0, dx, dv, ok);
    }
    if (!ok) {
      status = MAC_UNSAT;
      __syncthreads();
      break;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    A.state->status = status;
    A.state->depth = depth;
    A.state->npos = npos;
    A.state->nneg = nneg;
    A.state->nsol = nsol;
    A.state->revisions += (long long)ctl->revs;
    A.state->trace_len = ctl->trace_len;
    A.state->steps += 1;
  }
}

// lnet.bytes > 0 (kLds only): the network image is copied into LDS behind
// the working memory and the search reads the network from there.
template <bool kLds>
__global__ void k_mac(DevNet net, MacArgs args, char* scratch, LdsNet lnet) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WgCtl ctl;
  if (kLds && lnet.bytes > 0) {
    const DevNet ln = wg_load_net(net, lnet, smem + wg_mem_bytes(net.n, net.W));
    wg_mac_run(ln, wg_carve(smem, ln), &ctl, args);
  } else {
    char* base = kLds ? smem : scratch;
    wg_mac_run(net, wg_carve(base, net), &ctl, args);
  }
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
  int use_lds;            // 1: working memory in LDS; 2: network image in LDS as well
  int reset;              // 1: start the search over: level 0 = init, state = MAC_INIT
  const u64* init;        // the unit's initial domains, [n][W]
  LdsNet lnet;
};

// states_out[u] receives a copy of unit u's state when its launch ends, so
// the host reads every unit's outcome with one copy.
__global__ void k_mac_units(const UnitDesc* __restrict__ units, int count, unsigned* counter,
                            int only, char* scratch, MacState* states_out) {
  extern __shared__ __align__(16) char smem[];
  __shared__ WgCtl ctl;
  __shared__ int unit_sh;
  while (true) {
    if (threadIdx.x == 0) unit_sh = only >= 0 ? only : (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int u = uniform_int(unit_sh);
    __syncthreads();
    if (u >= count) break;
    const UnitDesc d = units[u];
    if (uniform_int(d.reset)) {
      // a new run of the batch: the workgroup restores level 0 and the state
      // itself, which spares the host two copies per unit
      const long long nw = (long long)d.net.n * d.net.W;
      for (long long i = threadIdx.x; i < nw; i += blockDim.x) d.args.trail[i] = d.init[i];
      if (threadIdx.x == 0) {
        MacState s0{};
        s0.status = MAC_INIT;
        *d.args.state = s0;
      }
      __syncthreads();
    }
    // One inlined copy of the search per place of the unit's memory (network
    // image and working memory in LDS, working memory in LDS, global
    // scratch), so that LDS accesses compile to ds_read / ds_write: one copy
    // over a pointer that may be either compiles them to flat instructions.
    const int where = uniform_int(d.use_lds);
    if (where == 2) {
      const DevNet ln = wg_load_net(d.net, d.lnet, smem + wg_mem_bytes(d.net.n, d.net.W));
      wg_mac_run(ln, wg_carve(smem, ln), &ctl, d.args);
    } else if (where == 1) {
      wg_mac_run(d.net, wg_carve(smem, d.net), &ctl, d.args);
    } else {
      wg_mac_run(d.net, wg_carve(scratch + d.scratch_off, d.net), &ctl, d.args);
    }
    if (threadIdx.x == 0) states_out[u] = *d.args.state;
    if (only >= 0) break;
  }
}
