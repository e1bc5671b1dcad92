// C-ABI implementation of include/cpim_gpu.h for MI355X (gfx950).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cpim_gpu.h"
#include "kernels.h"

#define CPIM_VERSION "cpim_amd 0.1 (gfx950)"

namespace {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
thread_local std::string g_last_error;

struct CpimFailure {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw CpimFailure{code, msg}; }

#define HIPC(expr)                                                                        \
  do {                                                                                    \
    hipError_t e__ = (expr);                                                              \
    if (e__ != hipSuccess)                                                                \
      fail(e__ == hipErrorOutOfMemory ? CPIM_ERR_OUT_OF_MEMORY : CPIM_ERR_HIP,            \
           std::string(#expr) + ": " + hipGetErrorString(e__));                           \
  } while (0)

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return CPIM_OK;
  } catch (const CpimFailure& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return CPIM_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CPIM_ERR_RUNTIME;
  }
}

#define REQUIRE(cond, msg) \
  do {                     \
    if (!(cond)) fail(CPIM_ERR_INVALID_ARGUMENT, msg); \
  } while (0)

// ---------------------------------------------------------------------------
// Device properties (cached)
// ---------------------------------------------------------------------------
struct DevProps {
  int cus = 0;
  int max_lds = 0;
  int coop = 0;
  long long wall_khz = 0;
  long long mem = 0;
  std::string arch;
};

DevProps props_of(int dev) {
  static std::mutex mu;
  static std::vector<std::unique_ptr<DevProps>> cache;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1);
  if (!cache[dev]) {
    auto p = std::make_unique<DevProps>();
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, dev));
    p->cus = prop.multiProcessorCount;
    p->max_lds = (int)prop.sharedMemPerBlock;
    p->mem = (long long)prop.totalGlobalMem;
    p->arch = prop.gcnArchName;
    int coop = 0;
    HIPC(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    p->coop = coop;
    int khz = 0;
    HIPC(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    p->wall_khz = khz;
    cache[dev] = std::move(p);
  }
  return *cache[dev];
}

// LDS used by one workgroup-engine domain copy; kept below 64 KB so that
// several workgroups share a CU.
constexpr long long kLdsBudget = 60 * 1024;

// Device wall-clock budget of one asynchronous fixpoint launch (ms); a
// launch that exceeds it returns an error instead of running on.
constexpr long long kAsyncBudgetMs = 20000;
// Waves per CU of the asynchronous fixpoint's grid.
constexpr int kAsyncWavesPerCu = 2;

// MAC nodes one launch may run before it returns to the host.
constexpr long long kNodesPerLaunch = 1 << 14;

// ---------------------------------------------------------------------------
// Model image: validated host-side layout of one network
// ---------------------------------------------------------------------------
struct ModelImage {
  int n = 0, max_dom = 0, W = 0, dpad = 0;
  int num_cons = 0;       // given
  int num_active = 0;     // propagated
  std::vector<int> cap;
  std::vector<ConMeta> con;
  std::vector<int> adj_off;
  std::vector<int4> adj;
  std::vector<int> bin_cons;          // active indices of binary constraints (plane build order)
  std::vector<u64> rel_dense;         // owned dense rows when the desc has none
  const u64* rel = nullptr;           // [nb][max_dom][rel_words]
  int rel_words = 0;
  std::vector<u64> tsup;
  long long sup_words = 0;
  std::vector<u64> init_dom;          // [n][W]
  std::vector<int> nbr;               // [n] distinct neighbours over the active constraints
};

ModelImage make_image(const cpim_model_desc* d, bool skip_non_binary, bool own_rel = false) {
  REQUIRE(d != nullptr, "model description is null");
  REQUIRE(d->num_vars >= 1, "model has no variables");
  REQUIRE(d->dom_size != nullptr, "dom_size is null");
  REQUIRE(d->num_cons >= 0, "negative constraint count");
  REQUIRE(d->num_cons == 0 || (d->con_arity && d->con_scope), "constraint arrays are null");
  ModelImage im;
  im.n = d->num_vars;
  im.cap.assign(d->dom_size, d->dom_size + im.n);
  for (int v = 0; v < im.n; ++v) {
    REQUIRE(im.cap[v] >= 1, "domain of variable " + std::to_string(v) + " is empty");
    REQUIRE(im.cap[v] <= 64 * CPIM_MAX_WORDS, "domain size above 1024 is not supported");
    im.max_dom = std::max(im.max_dom, im.cap[v]);
  }
  im.W = (im.max_dom + 63) / 64;
  im.dpad = 64 * im.W;
  im.num_cons = d->num_cons;
  const long long nw = (long long)im.n * im.W;
  // scopes
  std::vector<long long> scope_off(d->num_cons + 1, 0);
  int nbin_total = 0;
  for (int c = 0; c < d->num_cons; ++c) {
    const int ar = d->con_arity[c];
    REQUIRE(ar >= 2, "constraint " + std::to_string(c) + " has arity below 2");
    scope_off[c + 1] = scope_off[c] + ar;
    if (ar == 2) ++nbin_total;
  }
  for (long long i = 0; i < scope_off[d->num_cons]; ++i)
    REQUIRE(d->con_scope[i] >= 0 && d->con_scope[i] < im.n, "scope names an unknown variable");
  const bool any_nary = nbin_total != d->num_cons;
  if (any_nary && !skip_non_binary) REQUIRE(d->tup_off && d->tuples, "tables need tup_off/tuples");
  if (!d->bin_rel) REQUIRE(nbin_total == 0 || (d->tup_off && d->tuples), "binary relations missing");
  if (d->bin_rel) REQUIRE(d->rel_words >= im.W, "rel_words is below ceil(max_dom/64)");
  // active constraints
  long long sup_off = 0, tsup_off = 0;
  int bin_k = 0;
  std::vector<int> bin_ordinal_of_active;
  for (int c = 0; c < d->num_cons; ++c) {
    const int ar = d->con_arity[c];
    const int* sc = d->con_scope + scope_off[c];
    if (ar == 2) {
      REQUIRE(sc[0] != sc[1], "binary constraint " + std::to_string(c) + " repeats a variable");
      ConMeta m{};
      m.kind = 0;
      m.arity = 2;
      m.scope[0] = sc[0];
      m.scope[1] = sc[1];
      m.off = sup_off;
      sup_off += 2ll * im.W * im.dpad;
      im.bin_cons.push_back((int)im.con.size());
      bin_ordinal_of_active.push_back(bin_k);
      im.con.push_back(m);
      ++bin_k;
    } else {
      if (skip_non_binary) continue;
      REQUIRE(ar <= CPIM_MAX_ARITY, "tables of arity above 8 are not supported");
      const long long t0 = d->tup_off[c], t1 = d->tup_off[c + 1];
      REQUIRE(t1 >= t0 && (t1 - t0) % ar == 0, "tup_off does not delimit whole tuples");
      const long long T = (t1 - t0) / ar;
      REQUIRE(T < (1ll << 31), "tables of 2^31 tuples or more are not supported");
      for (int i = 0; i < ar; ++i)
        for (int j = i + 1; j < ar; ++j) REQUIRE(sc[i] != sc[j], "table scope repeats a variable");
      ConMeta m{};
      m.kind = 1;
      m.arity = ar;
      for (int i = 0; i < ar; ++i) m.scope[i] = sc[i];
      m.tw = std::max<int>(1, (int)((T + 63) / 64));
      m.off = tsup_off;
      const long long words = (long long)ar * im.dpad * m.tw;
      REQUIRE(tsup_off + words <= (1ll << 32), "table supports above 32 GiB are not supported");
      im.tsup.resize(tsup_off + words, 0ull);
      for (long long t = 0; t < T; ++t) {
        const int32_t* tup = d->tuples + t0 + t * ar;
        bool ok = true;
        for (int i = 0; i < ar; ++i) ok = ok && tup[i] >= 0 && tup[i] < im.cap[sc[i]];
        REQUIRE(ok, "table tuple value outside its domain");
        for (int i = 0; i < ar; ++i)
          im.tsup[tsup_off + ((long long)i * im.dpad + tup[i]) * m.tw + (t >> 6)] |= 1ull << (t & 63);
      }
      tsup_off += words;
      bin_ordinal_of_active.push_back(-1);
      im.con.push_back(m);
    }
  }
  im.num_active = (int)im.con.size();
  im.sup_words = std::max<long long>(sup_off, 1);
  if (im.tsup.empty()) im.tsup.assign(1, 0ull);
  // dense binary rows
  if (d->bin_rel) {
    im.rel = reinterpret_cast<const u64*>(d->bin_rel);
    im.rel_words = d->rel_words;
    if (own_rel) {
      im.rel_dense.assign(im.rel, im.rel + (size_t)nbin_total * im.max_dom * d->rel_words);
      im.rel = im.rel_dense.data();
    }
  } else if (!im.bin_cons.empty()) {
    im.rel_words = im.W;
    im.rel_dense.assign((size_t)im.bin_cons.size() * im.max_dom * im.W, 0ull);
    int k = 0;
    for (int c = 0; c < d->num_cons; ++c) {
      if (d->con_arity[c] != 2) continue;
      const int* sc = d->con_scope + scope_off[c];
      REQUIRE(d->tup_off[c + 1] >= d->tup_off[c] && (d->tup_off[c + 1] - d->tup_off[c]) % 2 == 0,
              "tup_off does not delimit whole tuples");
      for (long long t = d->tup_off[c]; t < d->tup_off[c + 1]; t += 2) {
        const int a = d->tuples[t], b = d->tuples[t + 1];
        REQUIRE(a >= 0 && a < im.cap[sc[0]] && b >= 0 && b < im.cap[sc[1]],
                "binary tuple value outside its domain");
        im.rel_dense[((size_t)k * im.max_dom + a) * im.W + (b >> 6)] |= 1ull << (b & 63);
      }
      ++k;
    }
    im.rel = im.rel_dense.data();
  }
  // CSR (constraint order per variable, like Network's subscription lists)
  std::vector<int> deg(im.n, 0);
  for (const auto& m : im.con)
    for (int i = 0; i < m.arity; ++i) ++deg[m.scope[i]];
  im.adj_off.assign(im.n + 1, 0);
  for (int v = 0; v < im.n; ++v) im.adj_off[v + 1] = im.adj_off[v] + deg[v];
  im.adj.resize(im.adj_off[im.n]);
  std::vector<int> pos(im.adj_off.begin(), im.adj_off.end() - 1);
  for (int c = 0; c < im.num_active; ++c) {
    const auto& m = im.con[c];
    for (int i = 0; i < m.arity; ++i) {
      const int v = m.scope[i];
      const int other = m.kind == 0 ? m.scope[1 - i] : -1;
      // binary: index of the constraint's plane pair (sup offset / (2 W dpad))
      const int pair = m.kind == 0 ? (int)(m.off / (2ll * im.W * im.dpad)) : -1;
      im.adj[pos[v]++] = make_int4(c, other, i, pair);
    }
  }
  // neighbour counts (Network::get_neighbor, src/Network.cpp:467-474)
  im.nbr.assign(im.n, 0);
  {
    std::vector<int> seen(im.n, -1);
    for (int v = 0; v < im.n; ++v)
      for (int k = im.adj_off[v]; k < im.adj_off[v + 1]; ++k) {
        const auto& m = im.con[im.adj[k].x];
        for (int i = 0; i < m.arity; ++i) {
          const int y = m.scope[i];
          if (y != v && seen[y] != v) {
            seen[y] = v;
            ++im.nbr[v];
          }
        }
      }
  }
  // level 0
  im.init_dom.assign(nw, 0ull);
  for (int v = 0; v < im.n; ++v)
    for (int a = 0; a < im.cap[v]; ++a) {
      const bool keep = d->init_dom == nullptr ||
                        ((d->init_dom[(long long)v * im.W + (a >> 6)] >> (a & 63)) & 1ull);
      if (keep) im.init_dom[(long long)v * im.W + (a >> 6)] |= 1ull << (a & 63);
    }
  return im;
}

// ---------------------------------------------------------------------------
// Device copy of a network
// ---------------------------------------------------------------------------
// Device memory: owned (freed on destruction) or a view into an Arena.
struct DevAlloc {
  void* p = nullptr;
  bool owned = true;
  DevAlloc() = default;
  explicit DevAlloc(size_t bytes) {
    if (bytes == 0) bytes = 16;
    HIPC(hipMalloc(&p, bytes));
  }
  static DevAlloc view(void* q) {
    DevAlloc d;
    d.p = q;
    d.owned = false;
    return d;
  }
  ~DevAlloc() {
    if (p && owned) (void)hipFree(p);
  }
  DevAlloc(const DevAlloc&) = delete;
  DevAlloc& operator=(const DevAlloc&) = delete;
  DevAlloc(DevAlloc&& o) noexcept : p(o.p), owned(o.owned) { o.p = nullptr; }
  DevAlloc& operator=(DevAlloc&& o) noexcept {
    if (this != &o) {
      if (p && owned) (void)hipFree(p);
      p = o.p;
      owned = o.owned;
      o.p = nullptr;
    }
    return *this;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// One device allocation carved into the arrays of one network (and its
// search state): a served request costs one hipMalloc instead of one per
// array.  plan mode (block == nullptr) only adds up the sizes.
struct Arena {
  DevAlloc block;
  long long cap = 0, used = 0;
  static long long round(long long b) { return (std::max<long long>(b, 16) + 255) & ~255ll; }
  void reserve(long long bytes) {
    block = DevAlloc((size_t)std::max<long long>(bytes, 256));
    cap = bytes;
    used = 0;
  }
  DevAlloc take(long long bytes) {
    const long long b = round(bytes);
    if (!block.p) {  // plan mode
      used += b;
      return DevAlloc();
    }
    if (used + b > cap) fail(CPIM_ERR_RUNTIME, "arena overflow (size plan and build disagree)");
    DevAlloc d = DevAlloc::view(static_cast<char*>(block.p) + used);
    used += b;
    return d;
  }
};

struct DevModel {
  int n = 0, W = 0, dpad = 0, max_dom = 0, num_active = 0;
  DevAlloc adj_off, adj, con, sup, tsup, cap, nbr;
  DevAlloc lnet_img;  // the network image for LDS (LdsNet), when small enough
  LdsNet lnet{};
  long long bytes = 0;
  DevNet net() const {
    DevNet d;
    d.n = n;
    d.W = W;
    d.dpad = dpad;
    d.ncons = num_active;
    d.adj_off = adj_off.as<int>();
    d.adj = adj.as<int4>();
    d.con = con.as<ConMeta>();
    d.sup = sup.as<u64>();
    d.tsup = tsup.as<u64>();
    return d;
  }
};

// One device array: a view of `arena` when given, else its own allocation.
DevAlloc dev_array(Arena* arena, long long bytes) {
  return arena ? arena->take(bytes) : DevAlloc((size_t)std::max<long long>(bytes, 16));
}

template <typename T>
void upload(DevAlloc& dst, const std::vector<T>& src, hipStream_t s, long long& bytes, Arena* arena = nullptr) {
  const size_t b = std::max<size_t>(sizeof(T) * src.size(), 16);
  dst = dev_array(arena, (long long)b);
  if (!src.empty()) HIPC(hipMemcpyAsync(dst.p, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, s));
  bytes += (long long)b;
}

// Largest network image copied into LDS by the device MAC (LdsNet).
constexpr long long kLdsNetMax = 128 * 1024;

long long align16(long long b) { return (b + 15) & ~15ll; }

// The LdsNet image of a network: the arrays of DevNet in one block, planes
// and tables with value stride dc = max_dom instead of dpad.  Built after the
// planes (the plane rows are copied from the device planes); skipped when the
// image would exceed kLdsNetMax.  The copies are queued on s from `h`,
// which the caller keeps until s is synchronized.
struct LdsLayout {
  long long o_adj, o_con, o_sup, o_tsup, bytes;
};

LdsLayout lds_layout(const ModelImage& im) {
  const int n = im.n, W = im.W, dc = std::max(1, im.max_dom);
  const long long npairs = (long long)im.bin_cons.size();
  long long tsup_c = 0;
  for (const ConMeta& m : im.con)
    if (m.kind == 1) tsup_c += (long long)m.arity * dc * m.tw;
  LdsLayout L;
  L.o_adj = align16(4ll * (n + 1));
  L.o_con = L.o_adj + align16(16ll * (long long)im.adj.size());
  L.o_sup = L.o_con + align16((long long)sizeof(ConMeta) * (long long)im.con.size());
  const long long sup_words = npairs * 2 * W * dc + 64ll * W;
  L.o_tsup = L.o_sup + align16(8 * sup_words);
  L.bytes = L.o_tsup + align16(8 * std::max<long long>(1, tsup_c));
  return L;
}

void build_lds_image(const ModelImage& im, DevModel& dm, hipStream_t s, std::vector<char>& h,
                     Arena* arena = nullptr) {
  const int n = im.n, W = im.W, dpad = im.dpad, dc = std::max(1, im.max_dom);
  const long long npairs = (long long)im.bin_cons.size();
  const LdsLayout L = lds_layout(im);
  const long long o_adj = L.o_adj, o_con = L.o_con, o_sup = L.o_sup, o_tsup = L.o_tsup, bytes = L.bytes;
  if (bytes > kLdsNetMax) return;
  h.assign((size_t)bytes, 0);
  std::memcpy(h.data(), im.adj_off.data(), 4 * (size_t)(n + 1));
  if (!im.adj.empty()) std::memcpy(h.data() + o_adj, im.adj.data(), 16 * im.adj.size());
  ConMeta* con = (ConMeta*)(h.data() + o_con);
  u64* ts = (u64*)(h.data() + o_tsup);
  long long toff = 0;
  for (size_t c = 0; c < im.con.size(); ++c) {
    ConMeta m = im.con[c];
    if (m.kind == 0) {
      m.off = m.off / (2ll * W * dpad) * (2ll * W * dc);
    } else {
      for (int q = 0; q < m.arity; ++q)
        for (int a = 0; a < dc; ++a)
          std::memcpy(ts + toff + ((long long)q * dc + a) * m.tw,
                      im.tsup.data() + m.off + ((long long)q * dpad + a) * m.tw, 8 * (size_t)m.tw);
      m.off = toff;
      toff += (long long)m.arity * dc * m.tw;
    }
    con[c] = m;
  }
  dm.lnet_img = dev_array(arena, bytes);
  HIPC(hipMemcpyAsync(dm.lnet_img.p, h.data(), (size_t)bytes, hipMemcpyHostToDevice, s));
  if (npairs > 0)
    HIPC(hipMemcpy2DAsync((char*)dm.lnet_img.p + o_sup, 8 * (size_t)dc, dm.sup.p, 8 * (size_t)dpad,
                          8 * (size_t)dc, (size_t)(npairs * 2 * W), hipMemcpyDeviceToDevice, s));
  dm.bytes += bytes;
  dm.lnet.src = dm.lnet_img.as<uint4>();
  dm.lnet.bytes = (int)bytes;
  dm.lnet.dc = dc;
  dm.lnet.o_adj = (int)o_adj;
  dm.lnet.o_con = (int)o_con;
  dm.lnet.o_sup = (int)o_sup;
  dm.lnet.o_tsup = (int)o_tsup;
}

// Device bytes build_device_model takes from an arena for im (the same
// arrays, each rounded as the arena rounds it).
long long model_arena_bytes(const ModelImage& im) {
  auto r = [](long long b) { return Arena::round(std::max<long long>(b, 16)); };
  long long b = r(4ll * (long long)im.adj_off.size()) + r(16ll * (long long)im.adj.size()) +
                r((long long)sizeof(ConMeta) * (long long)im.con.size()) + r(8ll * (long long)im.tsup.size()) +
                r(4ll * (long long)im.cap.size()) + r(4ll * (long long)im.nbr.size()) + r(8ll * im.sup_words);
  const long long img = lds_layout(im).bytes;
  if (img <= kLdsNetMax) b += r(img);
  return b;
}

// Staging of the dense rows on their way to the planes: bins (the binary
// constraints' indices) and slices of rows, reused across builds when the
// caller keeps it.
struct PlaneStaging {
  DevAlloc buf;
  long long bytes = 0;
  char* get(long long b) {
    if (b > bytes) {
      buf = DevAlloc((size_t)b);
      bytes = b;
    }
    return buf.as<char>();
  }
};

// The device arrays of a network; with an arena they are views into it (one
// allocation in all), with staging the row slices reuse its buffer.  Returns
// after the planes and the image are complete on stream s.
void build_device_model(const ModelImage& im, DevModel& dm, hipStream_t s, Arena* arena = nullptr,
                        PlaneStaging* staging = nullptr) {
  dm.n = im.n;
  dm.W = im.W;
  dm.dpad = im.dpad;
  dm.max_dom = im.max_dom;
  dm.num_active = im.num_active;
  upload(dm.adj_off, im.adj_off, s, dm.bytes, arena);
  upload(dm.adj, im.adj, s, dm.bytes, arena);
  upload(dm.con, im.con, s, dm.bytes, arena);
  upload(dm.tsup, im.tsup, s, dm.bytes, arena);
  upload(dm.cap, im.cap, s, dm.bytes, arena);
  upload(dm.nbr, im.nbr, s, dm.bytes, arena);
  dm.sup = dev_array(arena, (long long)sizeof(u64) * im.sup_words);
  dm.bytes += sizeof(u64) * im.sup_words;
  const int nb = (int)im.bin_cons.size();
  if (nb > 0) {
    // Upload the dense rows in slices and build the planes on the device.
    const long long row_words = (long long)im.max_dom * im.rel_words;
    const int slice = (int)std::max<long long>(1, std::min<long long>(nb, (256ll << 20) / (8 * row_words)));
    const long long bins_bytes = Arena::round(4ll * nb);
    PlaneStaging local;
    PlaneStaging& st = staging ? *staging : local;
    char* base = st.get(bins_bytes + 8ll * row_words * slice);
    int* bins = reinterpret_cast<int*>(base);
    u64* rel_dev = reinterpret_cast<u64*>(base + bins_bytes);
    HIPC(hipMemcpyAsync(bins, im.bin_cons.data(), 4ll * nb, hipMemcpyHostToDevice, s));
    for (int k0 = 0; k0 < nb; k0 += slice) {
      const int k1 = std::min(nb, k0 + slice);
      HIPC(hipMemcpyAsync(rel_dev, im.rel + (long long)k0 * row_words,
                          sizeof(u64) * row_words * (k1 - k0), hipMemcpyHostToDevice, s));
      const long long waves = (long long)(k1 - k0) * im.W;
      const int threads = 256;
      const long long blocks = (waves * 64 + threads - 1) / threads;
      hipLaunchKernelGGL(k_build_planes, dim3((unsigned)blocks), dim3(threads), 0, s,
                         rel_dev, k1 - k0, im.max_dom, im.rel_words, im.W, im.dpad,
                         bins + k0, dm.con.as<ConMeta>(), dm.cap.as<int>(), dm.sup.as<u64>());
      HIPC(hipGetLastError());
      if (k1 < nb) HIPC(hipStreamSynchronize(s));  // the next slice overwrites rel_dev
    }
  }
  std::vector<char> lds_image;  // host side of the image copy, until the sync below
  build_lds_image(im, dm, s, lds_image, arena);
  HIPC(hipStreamSynchronize(s));
}

struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    HIPC(hipGetDevice(&prev));
    if (prev != dev) HIPC(hipSetDevice(dev));
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

long long now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

// Device buffers of the singleton-test paths, kept by the model between
// calls (grown on demand) so that a SAC pass allocates nothing once warm.
struct SacScratch {
  DevAlloc tests, fails, revs;
  long long cap = 0;
  DevAlloc stests, perm, gfirst, gorder;  // bit-sliced path: sorted tests, their order, group starts,
                                          // group of each workgroup slot
  long long scap = 0, gcap = 0;
  // removal slices of the wide bit-sliced kernel: zeroed when allocated, and
  // left all zero by every launch of that kernel
  DevAlloc dead;
  long long dead_bytes = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  SacScratch() = default;
  SacScratch(const SacScratch&) = delete;
  SacScratch& operator=(const SacScratch&) = delete;
  ~SacScratch() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  void events() {
    for (hipEvent_t& e : ev)
      if (!e) HIPC(hipEventCreate(&e));
  }
};

// ---------------------------------------------------------------------------
// GModel
// ---------------------------------------------------------------------------
struct cpim_gmodel {
  int device = 0;
  hipStream_t stream = nullptr;
  ModelImage im;  // host layout (rel pointer not kept beyond build)
  DevModel dm;
  int max_depth = 0;
  int current_level = 0;
  int assigned_size = 0;
  DevAlloc dom;    // [max_depth][n][W]
  DevAlloc size;   // [max_depth][n]
  DevAlloc front;  // 3 x nfw
  DevAlloc ctl;    // GacCtl
  GacCtl* ctl_host = nullptr;  // pinned copy of the control block (host rounds)
  DevAlloc block_out;
  DevAlloc async_ctl;
  DevAlloc scratch;  // workgroup-engine scratch when a copy does not fit in LDS
  long long scratch_bytes = 0;
  int nfw = 0;
  std::vector<int2> assigned;  // d_assigned mirror: (var, value) per level
  DevProps props;
  SacScratch sac;
  hipEvent_t gev[2] = {nullptr, nullptr};  // kernel timing of the fixpoint entry points
  // dom/wdeg bookkeeping of the network (Tabular::weight, Network.cpp:303):
  // weights start at 1 with the model and persist across solves, as the
  // reference's constraint weights do; the stamps and clock order AC3's
  // revisions across calls
  DevAlloc wdeg_weight, wdeg_stamp, wdeg_clock;

  // the weights, stamps and clock, allocated and set to their first values on
  // first use or on reset
  void wdeg_init() {
    const int nc = std::max(1, dm.num_active);
    if (!wdeg_weight.p) {
      wdeg_weight = DevAlloc(sizeof(float) * nc);
      wdeg_stamp = DevAlloc(sizeof(long long) * nc);
      wdeg_clock = DevAlloc(sizeof(long long));
    }
    const std::vector<float> ones(nc, 1.0f);
    const long long one = 1;
    HIPC(hipMemcpyAsync(wdeg_weight.p, ones.data(), sizeof(float) * nc, hipMemcpyHostToDevice, stream));
    HIPC(hipMemsetAsync(wdeg_stamp.p, 0, sizeof(long long) * nc, stream));
    HIPC(hipMemcpyAsync(wdeg_clock.p, &one, sizeof(one), hipMemcpyHostToDevice, stream));
    HIPC(hipStreamSynchronize(stream));  // `ones` and `one` are host stack memory
  }

  void gac_events() {
    for (hipEvent_t& e : gev)
      if (!e) HIPC(hipEventCreate(&e));
  }
  // device time between gev[0] and gev[1] (both recorded, gev[1] complete)
  double gac_elapsed_ms() {
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, gev[0], gev[1]));
    return ms;
  }

  u64* level_dom(int level) const { return dom.as<u64>() + (long long)level * dm.n * dm.W; }
  int* level_size(int level) const { return size.as<int>() + (long long)level * dm.n; }
  long long level_words() const { return (long long)dm.n * dm.W; }

  char* get_scratch(long long bytes) {
    if (bytes > scratch_bytes) {
      scratch = DevAlloc((size_t)bytes);
      scratch_bytes = bytes;
    }
    return scratch.as<char>();
  }
  void sync() { HIPC(hipStreamSynchronize(stream)); }
};

extern "C" {

void cpim_gmodel_options_default(cpim_gmodel_options* o) {
  if (!o) return;
  o->device_id = 0;
  o->enable_prefetch = 0;
  o->skip_non_binary = 1;
  o->skip_conflicts = 0;
}

const char* cpim_last_error(void) { return g_last_error.c_str(); }
const char* cpim_version(void) { return CPIM_VERSION; }

int cpim_get_device_info(int32_t device_id, cpim_device_info* out) {
  return guarded([&] {
    REQUIRE(out != nullptr, "out is null");
    int count = 0;
    HIPC(hipGetDeviceCount(&count));
    REQUIRE(device_id >= 0 && device_id < count, "invalid device id");
    const DevProps p = props_of(device_id);
    out->compute_units = p.cus;
    out->lds_bytes_per_block = p.max_lds;
    out->cooperative_launch = p.coop;
    out->wavefront_size = 64;
    out->global_mem_bytes = p.mem;
    std::snprintf(out->arch, sizeof(out->arch), "%s", p.arch.c_str());
  });
}

int cpim_measure_copy_bandwidth(int32_t device_id, int64_t bytes, int32_t reps, double* gbps) {
  return guarded([&] {
    REQUIRE(gbps && bytes >= 16 && reps >= 1, "bad argument");
    DeviceGuard g(device_id);
    const DevProps p = props_of(device_id);
    const long long n16 = bytes / 16;
    DevAlloc a(16 * n16), b(16 * n16);
    hipStream_t s;
    HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPC(hipMemsetAsync(a.p, 1, 16 * n16, s));
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    const int threads = 256;
    const long long blocks = std::min<long long>((n16 + threads - 1) / threads, 8ll * p.cus * 4);
    double best = 0;
    for (int r = 0; r < reps; ++r) {
      HIPC(hipEventRecord(e0, s));
      hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(threads), 0, s, a.as<uint4>(), b.as<uint4>(), n16);
      HIPC(hipGetLastError());
      HIPC(hipEventRecord(e1, s));
      HIPC(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, e0, e1));
      if (ms > 0) best = std::max(best, 2.0 * 16 * n16 / (ms * 1e-3) / 1e9);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    *gbps = best;
  });
}

int cpim_gmodel_create(const cpim_model_desc* desc, const cpim_gmodel_options* opts,
                       cpim_gmodel** out) {
  return guarded([&] {
    REQUIRE(out != nullptr, "out is null");
    *out = nullptr;
    cpim_gmodel_options o;
    cpim_gmodel_options_default(&o);
    if (opts) o = *opts;
    // validate the model before touching the device
    ModelImage im = make_image(desc, o.skip_non_binary != 0);
    int count = 0;
    HIPC(hipGetDeviceCount(&count));
    REQUIRE(o.device_id >= 0 && o.device_id < count, "invalid device id");
    DeviceGuard g(o.device_id);
    auto m = std::make_unique<cpim_gmodel>();
    m->device = o.device_id;
    m->props = props_of(o.device_id);
    HIPC(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->im = std::move(im);
    build_device_model(m->im, m->dm, m->stream);
    m->im.rel = nullptr;
    m->im.rel_dense.clear();
    m->im.rel_dense.shrink_to_fit();
    const int n = m->dm.n;
    m->max_depth = n + 1;
    const long long lw = m->level_words();
    m->dom = DevAlloc(sizeof(u64) * lw * m->max_depth);
    m->size = DevAlloc(sizeof(int) * (long long)n * m->max_depth);
    HIPC(hipMemcpyAsync(m->dom.p, m->im.init_dom.data(), sizeof(u64) * lw, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_sizes, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->level_dom(0),
                       m->level_size(0), n, m->dm.W);
    HIPC(hipGetLastError());
    m->nfw = std::max(1, (m->dm.num_active + 31) / 32);
    m->front = DevAlloc(sizeof(unsigned) * 3ll * m->nfw);
    m->ctl = DevAlloc(sizeof(GacCtl));
    m->block_out = DevAlloc(sizeof(BlockGacOut));
    m->assigned.assign(m->max_depth, make_int2(-1, -1));
    m->dm.bytes += sizeof(u64) * lw * m->max_depth + sizeof(int) * (long long)n * m->max_depth +
                   sizeof(unsigned) * 3ll * m->nfw + sizeof(GacCtl);
    m->sync();
    *out = m.release();
  });
}

int cpim_gmodel_destroy(cpim_gmodel* m) {
  return guarded([&] {
    if (!m) return;
    DeviceGuard g(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    hipStream_t s = m->stream;
    if (m->ctl_host) (void)hipHostFree(m->ctl_host);
    for (hipEvent_t e : m->gev)
      if (e) (void)hipEventDestroy(e);
    delete m;
    if (s) (void)hipStreamDestroy(s);
  });
}

int cpim_gmodel_get_info(const cpim_gmodel* m, cpim_gmodel_info* out) {
  return guarded([&] {
    REQUIRE(m && out, "null argument");
    out->num_vars = m->dm.n;
    out->num_constraints = m->im.num_cons;
    out->num_active = m->dm.num_active;
    out->max_dom_size = m->dm.max_dom;
    out->dom_words = m->dm.W;
    out->bit_dom_int_size = (m->dm.max_dom + 31) / 32;
    out->max_depth = m->max_depth;
    out->subscription_size = m->im.adj_off.empty() ? 0 : m->im.adj_off.back();
    out->device_bytes = m->dm.bytes;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// GAC on the current level
// ---------------------------------------------------------------------------
namespace {

bool fits_lds(const cpim_gmodel* m) { return wg_mem_bytes(m->dm.n, m->dm.W) <= kLdsBudget; }


int knob(const char* name);

// Device wall-clock ticks a one-launch fixpoint (persistent or asynchronous)
// may run before it reports an error: the fixpoint_budget_us knob, or
// kAsyncBudgetMs.
unsigned long long fixpoint_budget_ticks(const cpim_gmodel* m) {
  const long long khz = std::max<long long>(m->props.wall_khz, 1);
  const long long us = knob("fixpoint_budget_us");
  return us > 0 ? (unsigned long long)std::max<long long>(1, khz * us / 1000)
                : (unsigned long long)(khz * kAsyncBudgetMs);
}

// Workgroups per CU of the persistent fixpoint (kPersistentThreads each): one
// workgroup of 16 waves per CU is what the kernel's ~105 VGPRs allow, and it
// keeps the barrier at one arrival per CU.
constexpr int kPersistentBlocksPerCu = 1;

// Parts per frontier word of a grid round: the smallest power of two (at
// most 32) that gives about `per_wave` work units per wave of `waves`.  One
// word-domain networks revise kPairBatch constraints per wave step, so one
// unit of several constraints per wave suffices there; otherwise four units
// of fewer constraints.
int frontier_parts(int nfw, long long waves, int W) {
  const long long per_wave = W == 1 ? 1 : 4;
  int parts = 1;
  while (parts < 32 && (long long)nfw * parts < per_wave * waves) parts *= 2;
  return parts;
}

unsigned long long slot_sum(const unsigned long long (&v)[kStatSlots]) {
  unsigned long long t = 0;
  for (unsigned long long x : v) t += x;
  return t;
}

// Start of a frontier fixpoint: control block and frontiers cleared, the
// first frontier seeded with every constraint (assigned_var < 0) or with the
// constraints of assigned_var.
void gac_start(cpim_gmodel* m, int assigned_var, unsigned* F, GacCtl* ctl) {
  const long long words = 3ll * m->nfw;
  const int blocks = (int)std::max<long long>(1, std::min<long long>((words + 255) / 256, 1024));
  hipLaunchKernelGGL(k_gac_init, dim3(blocks), dim3(256), 0, m->stream, ctl, F, m->nfw, m->dm.num_active,
                     assigned_var < 0 ? 1 : 0);
  if (assigned_var >= 0)
    hipLaunchKernelGGL(k_seed_var, dim3(1), dim3(256), 0, m->stream, F, m->dm.adj_off.as<int>(),
                       m->dm.adj.as<int4>(), assigned_var);
  HIPC(hipGetLastError());
}

// The persistent fixpoint kernel on the current level, after gac_start.
// Every workgroup must be resident for the grid barrier: at most the
// occupancy limit per CU (and at most kPersistentBlocksPerCu) on every CU.
// If some were not resident, the barrier's bounded wait turns that into the
// time-budget error the caller reports (ctl->err).
void persistent_launch(cpim_gmodel* m, unsigned* F, GacCtl* ctl) {
  const int L = m->current_level;
  const int threads = kPersistentThreads;
  int per_cu = 0;
  HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gac_persistent, threads, 0));
  per_cu = std::max(1, std::min(per_cu, kPersistentBlocksPerCu));
  const int waves_per_block = threads / 64;
  const int nfw = m->nfw;
  const int parts = frontier_parts(nfw, (long long)per_cu * waves_per_block * m->props.cus, m->dm.W);
  const long long units = (long long)nfw * parts;
  const int grid = (int)std::max<long long>(
      1, std::min<long long>((units + waves_per_block - 1) / waves_per_block, (long long)per_cu * m->props.cus));
  const long long max_rounds = (long long)m->dm.n * m->dm.max_dom + 2;
  const int mr = (int)std::min<long long>(max_rounds, INT_MAX);
  const unsigned long long max_ticks = fixpoint_budget_ticks(m);
  hipLaunchKernelGGL(k_gac_persistent, dim3(grid), dim3(threads), 0, m->stream, m->dm.net(), m->level_dom(L),
                     m->level_size(L), F, nfw, ctl, mr, parts, max_ticks);
  HIPC(hipGetLastError());
}

void fill_gac_stats(const GacCtl& res, int rounds, double kms, cpim_gac_stats* st) {
  st->iterations = rounds;
  st->deletions = (int32_t)slot_sum(res.deletions);
  st->inconsistent = res.wipe_round != INT_MAX;
  st->converged = st->inconsistent ? 0 : res.converged;
  st->revisions = (int64_t)slot_sum(res.revisions);
  st->kernel_ms = kms;
}

void gac_grid(cpim_gmodel* m, int assigned_var, bool persistent, cpim_gac_stats* st) {
  const int L = m->current_level;
  u64* dom = m->level_dom(L);
  int* size = m->level_size(L);
  unsigned* F = m->front.as<unsigned>();
  GacCtl* ctl = m->ctl.as<GacCtl>();
  const int nfw = m->nfw;
  if (!m->ctl_host) HIPC(hipHostMalloc((void**)&m->ctl_host, sizeof(GacCtl), hipHostMallocDefault));
  m->gac_events();
  double kms = 0.0;
  HIPC(hipEventRecord(m->gev[0], m->stream));
  gac_start(m, assigned_var, F, ctl);
  int rounds = 0;
  GacCtl res{};
  if (persistent) {
    persistent_launch(m, F, ctl);
    HIPC(hipEventRecord(m->gev[1], m->stream));
    HIPC(hipMemcpyAsync(m->ctl_host, ctl, sizeof(res), hipMemcpyDeviceToHost, m->stream));
    m->sync();
    kms = m->gac_elapsed_ms();
    res = *m->ctl_host;
    if (res.err) fail(CPIM_ERR_RUNTIME, "persistent fixpoint exceeded its time budget at a grid barrier");
    rounds = res.rounds;
  } else {
    // about four work units per resident wave (32 per CU)
    DevNet net = m->dm.net();
    const int threads = 256;
    const long long max_rounds = (long long)m->dm.n * m->dm.max_dom + 2;
    const int parts = frontier_parts(nfw, 32ll * m->props.cus, m->dm.W);
    const int waves_per_block = threads / 64;
    const long long units = (long long)nfw * parts;
    const int grid = (int)std::max<long long>(1, std::min<long long>((units + waves_per_block - 1) / waves_per_block,
                                                                      8ll * m->props.cus));
    for (int r = 0; r < max_rounds; ++r) {
      if (r > 0) HIPC(hipEventRecord(m->gev[0], m->stream));
      hipLaunchKernelGGL(k_gac_round, dim3(grid), dim3(threads), 0, m->stream, net, dom, size, F, nfw,
                         ctl, r, parts);
      HIPC(hipGetLastError());
      HIPC(hipEventRecord(m->gev[1], m->stream));
      HIPC(hipMemcpyAsync(m->ctl_host, ctl, sizeof(res), hipMemcpyDeviceToHost, m->stream));
      m->sync();
      kms += m->gac_elapsed_ms();
      res = *m->ctl_host;
      rounds = r + 1;
      if (res.wipe_round <= r) break;
      if (res.last_push_round < r) {
        res.converged = 1;
        break;
      }
    }
  }
  fill_gac_stats(res, rounds, kms, st);
}

void gac_async(cpim_gmodel* m, int assigned_var, cpim_gac_stats* st) {
  const int L = m->current_level;
  u64* dom = m->level_dom(L);
  int* size = m->level_size(L);
  unsigned* F = m->front.as<unsigned>();
  const int nfw = m->nfw;
  if (!m->async_ctl.p) m->async_ctl = DevAlloc(sizeof(AsyncCtl));
  AsyncCtl* ctl = m->async_ctl.as<AsyncCtl>();
  m->gac_events();
  HIPC(hipEventRecord(m->gev[0], m->stream));
  AsyncCtl init{};
  HIPC(hipMemcpyAsync(ctl, &init, sizeof(init), hipMemcpyHostToDevice, m->stream));
  HIPC(hipMemsetAsync(F, 0, sizeof(unsigned) * nfw, m->stream));
  hipLaunchKernelGGL(k_seed_async, dim3(assigned_var < 0 ? (nfw + 255) / 256 : 1), dim3(256), 0, m->stream, F, nfw,
                     m->dm.num_active, m->dm.adj_off.as<int>(), m->dm.adj.as<int4>(), assigned_var, ctl);
  HIPC(hipGetLastError());
  // two waves per CU: on the C2 cascade 1024 waves took 5.0-6.0 ms, 512
  // waves 3.8-4.4 ms and 256 waves 4.4-4.9 ms (more waves, more contention on
  // the frontier words and the pending count)
  const int threads = 256;
  const int waves = std::max(1, std::min(nfw, kAsyncWavesPerCu * m->props.cus));
  const int blocks = (waves + threads / 64 - 1) / (threads / 64);
  const unsigned long long max_ticks = fixpoint_budget_ticks(m);
  DevNet net = m->dm.net();
  hipLaunchKernelGGL(k_gac_async, dim3(blocks), dim3(threads), 0, m->stream, net, dom, size, F, nfw, ctl, max_ticks);
  HIPC(hipGetLastError());
  HIPC(hipEventRecord(m->gev[1], m->stream));
  AsyncCtl res{};
  HIPC(hipMemcpyAsync(&res, ctl, sizeof(res), hipMemcpyDeviceToHost, m->stream));
  m->sync();
  st->kernel_ms = m->gac_elapsed_ms();
  if (res.err) fail(CPIM_ERR_RUNTIME, "asynchronous fixpoint exceeded its time budget");
  st->iterations = 1;
  st->deletions = (int32_t)slot_sum(res.deletions);
  st->inconsistent = res.wipe != 0;
  st->converged = !res.wipe && res.pending == 0;
  st->revisions = (int64_t)slot_sum(res.revisions);
}

void gac_block(cpim_gmodel* m, int assigned_var, cpim_gac_stats* st) {
  const int L = m->current_level;
  DevNet net = m->dm.net();
  const long long bytes = wg_mem_bytes(m->dm.n, m->dm.W);
  const int threads = 1024;
  BlockGacOut* out = m->block_out.as<BlockGacOut>();
  m->gac_events();
  HIPC(hipEventRecord(m->gev[0], m->stream));
  if (bytes <= kLdsBudget) {
    hipLaunchKernelGGL(k_block_gac<true>, dim3(1), dim3(threads), (size_t)bytes, m->stream, net,
                       m->level_dom(L), m->level_size(L), assigned_var, (char*)nullptr, out);
  } else {
    char* scratch = m->get_scratch(bytes);
    hipLaunchKernelGGL(k_block_gac<false>, dim3(1), dim3(threads), 0, m->stream, net, m->level_dom(L),
                       m->level_size(L), assigned_var, scratch, out);
  }
  HIPC(hipGetLastError());
  HIPC(hipEventRecord(m->gev[1], m->stream));
  BlockGacOut res{};
  HIPC(hipMemcpyAsync(&res, out, sizeof(res), hipMemcpyDeviceToHost, m->stream));
  m->sync();
  st->kernel_ms = m->gac_elapsed_ms();
  st->iterations = 1;
  st->deletions = (int32_t)res.deletions;
  st->inconsistent = !res.consistent;
  st->converged = res.consistent;
  st->revisions = (int64_t)res.revisions;
}

int auto_gac_mode(const cpim_gmodel* m, int assigned_var) {
  // Small networks: one workgroup with the domains in LDS (one launch, no
  // grid barrier).  Large ones seeded with every constraint: host-driven
  // rounds, each spread over the whole grid (few rounds at the benchmark
  // tightness; 0.21 ms against 0.52 ms asynchronous for one fixpoint of
  // 80,000 revisions).  Large ones seeded from one variable: the
  // asynchronous one-launch fixpoint, which has no per-round host sync.
  if (fits_lds(m) && m->dm.num_active <= 4096) return CPIM_GAC_BLOCK;
  return assigned_var < 0 ? CPIM_GAC_PLAIN : CPIM_GAC_ASYNC;
}

void gac(cpim_gmodel* m, int assigned_var, int mode, cpim_gac_stats* st) {
  std::memset(st, 0, sizeof(*st));
  if (mode == CPIM_GAC_AUTO) mode = auto_gac_mode(m, assigned_var);
  if (m->dm.num_active == 0) {
    st->converged = 1;
    return;
  }
  switch (mode) {
    case CPIM_GAC_PLAIN: gac_grid(m, assigned_var, false, st); break;
    case CPIM_GAC_PERSISTENT: gac_grid(m, assigned_var, true, st); break;
    case CPIM_GAC_BLOCK: gac_block(m, assigned_var, st); break;
    case CPIM_GAC_ASYNC: gac_async(m, assigned_var, st); break;
    default: fail(CPIM_ERR_INVALID_ARGUMENT, "unknown GAC mode " + std::to_string(mode));
  }
}

void check_var(const cpim_gmodel* m, int var, const char* fn) {
  if (var < 0 || var >= m->dm.n)
    fail(CPIM_ERR_RUNTIME, std::string("[GModel::") + fn + "] Invalid var: " + std::to_string(var));
}

void check_level(const cpim_gmodel* m, int level, const char* fn) {
  if (level < 0 || level >= m->max_depth)
    fail(CPIM_ERR_RUNTIME, std::string("[GModel::") + fn + "] Invalid level: " + std::to_string(level));
}

std::vector<u64> read_var(const cpim_gmodel* m, int var, int level) {
  std::vector<u64> w(m->dm.W);
  HIPC(hipMemcpyAsync(w.data(), m->level_dom(level) + (long long)var * m->dm.W, sizeof(u64) * m->dm.W,
                      hipMemcpyDeviceToHost, m->stream));
  HIPC(hipStreamSynchronize(m->stream));
  return w;
}

int read_size(const cpim_gmodel* m, int var, int level) {
  int s = 0;
  HIPC(hipMemcpyAsync(&s, m->level_size(level) + var, sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIPC(hipStreamSynchronize(m->stream));
  return s;
}

}  // namespace

extern "C" {

int cpim_gmodel_enforce_gac(cpim_gmodel* m, int32_t assigned_var, int32_t mode, cpim_gac_stats* st) {
  return guarded([&] {
    REQUIRE(m && st, "null argument");
    if (assigned_var >= 0) check_var(m, assigned_var, "EnforceGAC");
    DeviceGuard g(m->device);
    gac(m, assigned_var, mode, st);
  });
}


int cpim_gmodel_create_new_level(cpim_gmodel* m, int32_t* new_level) {
  return guarded([&] {
    REQUIRE(m && new_level, "null argument");
    if (m->current_level >= m->max_depth - 1)
      fail(CPIM_ERR_RUNTIME, "[GModel::CreateNewLevel] Max depth reached: " + std::to_string(m->max_depth));
    DeviceGuard g(m->device);
    const int L = ++m->current_level;
    HIPC(hipMemcpyAsync(m->level_dom(L), m->level_dom(L - 1), sizeof(u64) * m->level_words(),
                        hipMemcpyDeviceToDevice, m->stream));
    HIPC(hipMemcpyAsync(m->level_size(L), m->level_size(L - 1), sizeof(int) * m->dm.n,
                        hipMemcpyDeviceToDevice, m->stream));
    m->sync();
    *new_level = L;
  });
}

int cpim_gmodel_back_to_level(cpim_gmodel* m, int32_t level) {
  return guarded([&] {
    REQUIRE(m, "null argument");
    if (level < 0 || level >= m->max_depth)
      fail(CPIM_ERR_RUNTIME, "[GModel::BackToLevel] Invalid level: " + std::to_string(level));
    if (level > m->current_level)
      fail(CPIM_ERR_RUNTIME, "[GModel::BackToLevel] Cannot back to future level: " + std::to_string(level) +
                                 " (current=" + std::to_string(m->current_level) + ")");
    m->current_level = level;
    while (m->assigned_size > level) --m->assigned_size;
  });
}

int cpim_gmodel_current_level(const cpim_gmodel* m, int32_t* level) {
  return guarded([&] {
    REQUIRE(m && level, "null argument");
    *level = m->current_level;
  });
}

int cpim_gmodel_assign_value(cpim_gmodel* m, int32_t var, int32_t value, int32_t level, int32_t* ok) {
  return guarded([&] {
    REQUIRE(m && ok, "null argument");
    check_var(m, var, "AssignValue");
    check_level(m, level, "AssignValue");
    if (value < 0 || value >= m->dm.max_dom)
      fail(CPIM_ERR_RUNTIME, "[GModel::AssignValue] Invalid value: " + std::to_string(value));
    DeviceGuard g(m->device);
    std::vector<u64> w(m->dm.W, 0ull);
    w[value >> 6] = 1ull << (value & 63);
    const int one = 1;
    HIPC(hipMemcpyAsync(m->level_dom(level) + (long long)var * m->dm.W, w.data(), sizeof(u64) * m->dm.W,
                        hipMemcpyHostToDevice, m->stream));
    HIPC(hipMemcpyAsync(m->level_size(level) + var, &one, sizeof(int), hipMemcpyHostToDevice, m->stream));
    m->sync();
    m->assigned[level] = make_int2(var, value);
    m->assigned_size = std::max(m->assigned_size, level + 1);
    *ok = 1;
  });
}

int cpim_gmodel_remove_value(cpim_gmodel* m, int32_t var, int32_t value, int32_t level, int32_t* ok) {
  return guarded([&] {
    REQUIRE(m && ok, "null argument");
    check_var(m, var, "RemoveValue");
    check_level(m, level, "RemoveValue");
    *ok = 0;
    if (value < 0 || value >= m->dm.max_dom) return;
    DeviceGuard g(m->device);
    std::vector<u64> w = read_var(m, var, level);
    const u64 bit = 1ull << (value & 63);
    if (!(w[value >> 6] & bit)) return;
    w[value >> 6] &= ~bit;
    int s = 0;
    for (u64 x : w) s += __builtin_popcountll(x);
    HIPC(hipMemcpyAsync(m->level_dom(level) + (long long)var * m->dm.W + (value >> 6), &w[value >> 6],
                        sizeof(u64), hipMemcpyHostToDevice, m->stream));
    HIPC(hipMemcpyAsync(m->level_size(level) + var, &s, sizeof(int), hipMemcpyHostToDevice, m->stream));
    m->sync();
    *ok = s > 0;
  });
}

int cpim_gmodel_domain_size(const cpim_gmodel* m, int32_t var, int32_t level, int32_t* size) {
  return guarded([&] {
    REQUIRE(m && size, "null argument");
    check_var(m, var, "GetDomainSize");
    check_level(m, level, "GetDomainSize");
    DeviceGuard g(m->device);
    *size = read_size(m, var, level);
  });
}

int cpim_gmodel_min_domain_var(const cpim_gmodel* m, int32_t level, int32_t* var) {
  return guarded([&] {
    REQUIRE(m && var, "null argument");
    check_level(m, level, "GetMinDomainVar");
    DeviceGuard g(m->device);
    std::vector<int> s(m->dm.n);
    HIPC(hipMemcpyAsync(s.data(), m->level_size(level), sizeof(int) * m->dm.n, hipMemcpyDeviceToHost, m->stream));
    HIPC(hipStreamSynchronize(m->stream));
    int best = -1, best_size = m->dm.max_dom + 1;
    for (int v = 0; v < m->dm.n; ++v)
      if (s[v] > 1 && s[v] < best_size) {
        best_size = s[v];
        best = v;
      }
    *var = best;
  });
}

int cpim_gmodel_is_assigned(const cpim_gmodel* m, int32_t var, int32_t level, int32_t* res) {
  return guarded([&] {
    REQUIRE(m && res, "null argument");
    check_var(m, var, "IsAssigned");
    check_level(m, level, "IsAssigned");
    DeviceGuard g(m->device);
    *res = read_size(m, var, level) == 1;
  });
}

int cpim_gmodel_is_fully_assigned(const cpim_gmodel* m, int32_t level, int32_t* res) {
  return guarded([&] {
    REQUIRE(m && res, "null argument");
    check_level(m, level, "IsFullyAssigned");
    DeviceGuard g(m->device);
    std::vector<int> s(m->dm.n);
    HIPC(hipMemcpyAsync(s.data(), m->level_size(level), sizeof(int) * m->dm.n, hipMemcpyDeviceToHost, m->stream));
    HIPC(hipStreamSynchronize(m->stream));
    *res = std::all_of(s.begin(), s.end(), [](int x) { return x == 1; });
  });
}

static int first_from(const std::vector<u64>& w, int start, int max_dom) {
  for (int a = start; a < 64 * (int)w.size(); ++a) {
    const u64 word = w[a >> 6] >> (a & 63);
    if (word) {
      const int v = a + (int)__builtin_ctzll(word);
      return v < max_dom ? v : -1;
    }
    a |= 63;
  }
  return -1;
}

int cpim_gmodel_assigned_value(const cpim_gmodel* m, int32_t var, int32_t level, int32_t* value) {
  return guarded([&] {
    REQUIRE(m && value, "null argument");
    check_var(m, var, "GetAssignedValue");
    check_level(m, level, "GetAssignedValue");
    DeviceGuard g(m->device);
    if (read_size(m, var, level) != 1) {
      *value = -1;
      return;
    }
    *value = first_from(read_var(m, var, level), 0, m->dm.max_dom);
  });
}

int cpim_gmodel_first_value(const cpim_gmodel* m, int32_t var, int32_t level, int32_t* value) {
  return guarded([&] {
    REQUIRE(m && value, "null argument");
    check_var(m, var, "GetFirstValue");
    check_level(m, level, "GetFirstValue");
    DeviceGuard g(m->device);
    *value = first_from(read_var(m, var, level), 0, m->dm.max_dom);
  });
}

int cpim_gmodel_next_value(const cpim_gmodel* m, int32_t var, int32_t value, int32_t level, int32_t* next) {
  return guarded([&] {
    REQUIRE(m && next, "null argument");
    check_var(m, var, "GetNextValue");
    check_level(m, level, "GetNextValue");
    if (value < 0 || value >= m->dm.max_dom - 1) {
      *next = -1;
      return;
    }
    DeviceGuard g(m->device);
    *next = first_from(read_var(m, var, level), value + 1, m->dm.max_dom);
  });
}

int cpim_gmodel_read_domains(const cpim_gmodel* m, int32_t level, uint64_t* out) {
  return guarded([&] {
    REQUIRE(m && out, "null argument");
    check_level(m, level, "ReadDomains");
    DeviceGuard g(m->device);
    HIPC(hipMemcpyAsync(out, m->level_dom(level), sizeof(u64) * m->level_words(), hipMemcpyDeviceToHost,
                        m->stream));
    HIPC(hipStreamSynchronize(m->stream));
  });
}

int cpim_gmodel_weights(cpim_gmodel* m, float* out, int64_t capacity, int64_t* count, int32_t reset) {
  return guarded([&] {
    REQUIRE(m && count, "null argument");
    REQUIRE(capacity >= 0 && (capacity == 0 || out), "null argument");
    DeviceGuard g(m->device);
    const int nc = m->dm.num_active;
    *count = nc;
    if (out && capacity > 0 && nc > 0) {
      const long long k = std::min<long long>(capacity, nc);
      if (m->wdeg_weight.p) {
        HIPC(hipMemcpyAsync(out, m->wdeg_weight.p, sizeof(float) * k, hipMemcpyDeviceToHost, m->stream));
        HIPC(hipStreamSynchronize(m->stream));
      } else {
        for (long long i = 0; i < k; ++i) out[i] = 1.0f;  // no dom/wdeg solve yet
      }
    }
    if (reset && m->wdeg_weight.p) m->wdeg_init();
  });
}

int cpim_gmodel_write_domains(cpim_gmodel* m, int32_t level, const uint64_t* in) {
  return guarded([&] {
    REQUIRE(m && in, "null argument");
    check_level(m, level, "WriteDomains");
    DeviceGuard g(m->device);
    // mask to the domain capacities so that sizes stay exact
    const u64* inw = reinterpret_cast<const u64*>(in);
    std::vector<u64> w(inw, inw + m->level_words());
    for (int v = 0; v < m->dm.n; ++v)
      for (int j = 0; j < m->dm.W; ++j) {
        const int k = m->im.cap[v] - 64 * j;
        const u64 mask = k >= 64 ? ~0ull : (k <= 0 ? 0ull : ((1ull << k) - 1ull));
        w[(long long)v * m->dm.W + j] &= mask;
      }
    HIPC(hipMemcpyAsync(m->level_dom(level), w.data(), sizeof(u64) * w.size(), hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_sizes, dim3((m->dm.n + 255) / 256), dim3(256), 0, m->stream, m->level_dom(level),
                       m->level_size(level), m->dm.n, m->dm.W);
    HIPC(hipGetLastError());
    m->sync();
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Singleton tests and SAC
// ---------------------------------------------------------------------------
namespace {

// Launch k_singleton_tests over `count` tests: one workgroup per test.  With
// domains in LDS the whole list is one launch; with global scratch the list
// is cut into launches of kScratchBlocks workgroups.
constexpr long long kScratchBlocks = 2048;

void run_singleton_batched(cpim_gmodel* m, const int2* tests_host, long long count, unsigned char* fail_host,
                           cpim_sac_stats* st, SacScratch& S) {
  if (count == 0) return;
  if (count > S.cap) {
    S.tests = DevAlloc(sizeof(int2) * count);
    S.fails = DevAlloc(count);
    S.cap = count;
  }
  if (!S.revs.p) S.revs = DevAlloc(sizeof(unsigned long long) * kStatSlots);  // per-slot counters
  HIPC(hipMemcpyAsync(S.tests.p, tests_host, sizeof(int2) * count, hipMemcpyHostToDevice, m->stream));
  HIPC(hipMemsetAsync(S.revs.p, 0, sizeof(unsigned long long) * kStatSlots, m->stream));
  const int L = m->current_level;
  DevNet net = m->dm.net();
  const long long bytes = wg_mem_bytes(m->dm.n, m->dm.W);
  const int threads = 256;
  S.events();
  HIPC(hipEventRecord(S.ev[0], m->stream));
  if (bytes <= kLdsBudget) {
    for (long long first = 0; first < count; first += INT_MAX) {
      const long long blocks = std::min<long long>(count - first, INT_MAX);
      hipLaunchKernelGGL(k_singleton_tests<true>, dim3((unsigned)blocks), dim3(threads), (size_t)bytes,
                         m->stream, net, m->level_dom(L), m->level_size(L), S.tests.as<int2>(), count, first,
                         S.fails.as<unsigned char>(), S.revs.as<unsigned long long>(), (char*)nullptr, 0ll);
      HIPC(hipGetLastError());
    }
  } else {
    char* scratch = m->get_scratch(bytes * kScratchBlocks);
    for (long long first = 0; first < count; first += kScratchBlocks) {
      const long long blocks = std::min<long long>(count - first, kScratchBlocks);
      hipLaunchKernelGGL(k_singleton_tests<false>, dim3((unsigned)blocks), dim3(threads), 0, m->stream, net,
                         m->level_dom(L), m->level_size(L), S.tests.as<int2>(), count, first,
                         S.fails.as<unsigned char>(), S.revs.as<unsigned long long>(), scratch, bytes);
      HIPC(hipGetLastError());
    }
  }
  HIPC(hipEventRecord(S.ev[1], m->stream));
  HIPC(hipMemcpyAsync(fail_host, S.fails.p, count, hipMemcpyDeviceToHost, m->stream));
  unsigned long long revs_slots[kStatSlots];
  HIPC(hipMemcpyAsync(revs_slots, S.revs.p, sizeof(revs_slots), hipMemcpyDeviceToHost, m->stream));
  m->sync();
  const unsigned long long revs = slot_sum(revs_slots);
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
  st->kernel_ms += ms;
  st->revisions += (int64_t)revs;
  st->singleton_tests += count;
}

// Bit-sliced path: tests sorted by variable and cut into groups of at most
// 64 (a variable with more values is split into balanced groups; small
// variables share a group), one workgroup per group.  The slices of the
// resident groups take n * dpad words each, within kSliceBudgetBytes.
// ---------------------------------------------------------------------------
// Tuning knobs (cpim_tune_set / cpim_tune_get), process-wide; 0 = automatic.
// Their first values come from the environment.
// ---------------------------------------------------------------------------
struct Knob {
  const char* name;
  const char* env;
  std::atomic<int> value;
  bool loaded;
};
Knob g_knobs[] = {
    {"sac_slice_masks", "CPIM_SLICE_MASKS", {0}, false},
    {"sac_slice_threads", "CPIM_SLICE_THREADS", {0}, false},
    {"sac_slice_kernel", "CPIM_SLICE_KERNEL", {0}, false},
    {"unit_threads", "CPIM_UNIT_THREADS", {0}, false},
    {"sac_slice_nibbles", "CPIM_SLICE_NIBBLES", {0}, false},
    {"mac_lds_net", "CPIM_MAC_LDS_NET", {0}, false},
    {"fixpoint_budget_us", "CPIM_FIXPOINT_BUDGET_US", {0}, false},
    {"sac_slice_bytes", "CPIM_SLICE_BYTES", {0}, false},
};
std::mutex g_knob_mu;

bool knob_valid(const char* name, long long v) {
  const std::string k(name);
  if (k == "sac_slice_masks") return v == 0 || v == 1 || v == 2 || v == 4;
  if (k == "sac_slice_kernel") return v == 0 || v == 1;
  if (k == "sac_slice_nibbles") return v == 0 || v == 1 || v == 2;
  if (k == "sac_slice_bytes") return v == 0 || v == 1 || v == 4 || v == 6;
  if (k == "mac_lds_net") return v == 0 || v == 1;
  if (k == "fixpoint_budget_us") return v >= 0 && v <= 600000000ll;
  return v == 0 || (v >= 64 && v <= 1024 && v % 64 == 0);
}

Knob* find_knob(const char* name) {
  for (Knob& k : g_knobs)
    if (std::strcmp(k.name, name) == 0) {
      std::lock_guard<std::mutex> lock(g_knob_mu);
      if (!k.loaded) {
        const char* e = std::getenv(k.env);
        const long long v = e ? std::atoll(e) : 0;
        k.value = knob_valid(name, v) ? (int)v : 0;
        k.loaded = true;
      }
      return &k;
    }
  return nullptr;
}

int knob(const char* name) { return find_knob(name)->value.load(); }

// Bit-sliced singleton tests.  Two kernels: k_sliced_wide<M> (default: 64 M
// tests per workgroup, one shared stage per changed variable) and
// k_sliced_tests (64 tests, one stage per wave; sac_slice_kernel = 1).
constexpr int kSliceThreads = 512;

int slice_threads() {
  const int t = knob("sac_slice_threads");
  return t ? t : kSliceThreads;
}
constexpr long long kSliceBudgetBytes = 4ll << 30;

// Threads per workgroup for this network: the configured size, halved while
// the per-wave staging areas exceed the LDS budget (0 if even one wave does not fit).
int slice_threads_for(const cpim_gmodel* m) {
  int t = slice_threads();
  while (t > 64 && slice_lds_bytes(m->dm.n, m->dm.dpad, t / 64) > kLdsBudget) t /= 2;
  return slice_lds_bytes(m->dm.n, m->dm.dpad, t / 64) <= kLdsBudget ? t : 0;
}

// Masks per (variable, value) of the wide kernel: sac_slice_masks, or 4 when
// the tests fill a workgroup of 256 per CU, 2 when 128 do, else 1; halved
// while the LDS image does not fit (0 if M = 1 does not).
int wide_masks_for(const cpim_gmodel* m, long long count) {
  int M = knob("sac_slice_masks");
  if (M == 0) {
    const long long cus = std::max(1, m->props.cus);
    M = count >= 256 * cus ? 4 : (count >= 128 * cus ? 2 : 1);
  }
  while (M > 1 && wide_lds_bytes(m->dm.n, m->dm.W, M) > kLdsBudget) M /= 2;
  return wide_lds_bytes(m->dm.n, m->dm.W, M) <= kLdsBudget ? M : 0;
}

bool use_wide_kernel() { return knob("sac_slice_kernel") == 0; }

// Nibble tables (sac_slice_nibbles: 0 auto = on when their LDS image stays
// within 80 KB, two workgroups per CU; 1 on when it fits; 2 off).
bool use_nibble_tables(const cpim_gmodel* m, int M) {
  const int k = knob("sac_slice_nibbles");
  if (k == 2) return false;
  const long long b = wide_lds_bytes(m->dm.n, m->dm.W, M, true);
  return b <= (k == 1 ? (long long)m->props.max_lds : std::min<long long>(m->props.max_lds, 80 * 1024));
}

// Byte tables in front of the nibble tables (sac_slice_bytes: 4 or 6 = that
// many whenever the single stage buffer fits; 0 and 1 = none).
int byte_tables_for(const cpim_gmodel* m, int M, bool nib) {
  const int k = knob("sac_slice_bytes");
  if (!nib || k <= 1) return 0;
  const long long limit = k ? (long long)m->props.max_lds : std::min<long long>(m->props.max_lds, 80 * 1024);
  for (int nb : {6, 4})
    if ((k == 0 || k == nb) && wide_lds_bytes(m->dm.n, m->dm.W, M, true, nb) <= limit) return nb;
  return 0;
}

bool sliced_eligible(const cpim_gmodel* m) {
  if ((int)m->im.bin_cons.size() != m->im.num_active) return false;
  return use_wide_kernel() ? wide_masks_for(m, 1) > 0 : slice_threads_for(m) > 0;
}

void run_singleton_sliced(cpim_gmodel* m, const int2* tests_host, long long count, unsigned char* fail_host,
                          cpim_sac_stats* st, SacScratch& S) {
  if (count == 0) return;
  REQUIRE(count <= INT_MAX, "too many singleton tests for one call");
  // Tests are grouped by variable: a list sorted by variable (every SAC pass)
  // needs no sort.  The check has no early exit, so that it vectorizes.
  int descents = 0;
  for (long long i = 1; i < count; ++i) descents |= tests_host[i].x < tests_host[i - 1].x;
  bool sorted_in = descents == 0;
  std::vector<int> order;
  std::vector<int2> sorted_buf;
  const int2* sorted = tests_host;
  if (!sorted_in) {
    order.resize((size_t)count);
    for (long long i = 0; i < count; ++i) order[(size_t)i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return tests_host[a].x < tests_host[b].x; });
    sorted_buf.resize((size_t)count);
    for (long long i = 0; i < count; ++i) sorted_buf[(size_t)i] = tests_host[order[(size_t)i]];
    sorted = sorted_buf.data();
  }
  const bool wide = use_wide_kernel();
  const int M = wide ? wide_masks_for(m, count) : 1;
  const bool nib = wide && use_nibble_tables(m, M);
  const int nb = wide ? byte_tables_for(m, M, nib) : 0;
  REQUIRE(M > 0, "bit-sliced singleton tests do not fit in LDS");
  const long long cap = 64ll * M;  // tests per workgroup
  std::vector<int> gfirst;
  std::vector<long long> w;  // expected work of each group: summed degree of its variables
  gfirst.reserve((size_t)(count / 32 + 2));
  w.reserve((size_t)(count / 32 + 2));
  int open = -1;  // start of the group being filled, -1 if none
  for (long long i = 0; i < count;) {
    // end j of the run of x: gallop, then bisect (the list is sorted)
    const int x = sorted[i].x;
    long long lo = i, step = 1;
    while (lo + step < count && sorted[lo + step].x == x) {
      lo += step;
      step *= 2;
    }
    long long hi = std::min(count, lo + step);  // sorted[hi].x != x, or hi == count
    while (hi - lo > 1) {
      const long long mid = lo + (hi - lo) / 2;
      if (sorted[mid].x == x) lo = mid;
      else hi = mid;
    }
    const long long j = hi;
    const long long len = j - i;
    const long long dx = m->im.adj_off[x + 1] - m->im.adj_off[x];
    if (len > cap) {
      open = -1;
      const long long k = (len + cap - 1) / cap;
      for (long long p = 0; p < k; ++p) {
        gfirst.push_back((int)(i + len * p / k));
        w.push_back(dx);
      }
    } else if (open < 0 || i + len - open > cap) {
      gfirst.push_back((int)i);
      w.push_back(dx);
      open = (int)i;
    } else {
      w.back() += dx;
    }
    i = j;
  }
  const long long groups = (long long)gfirst.size();
  gfirst.push_back((int)count);
  // Place the groups on workgroups by expected work.  The pass runs as one
  // wave of workgroups, two per CU; workgroups bid and bid + cus share a CU
  // (dispatch fills one slot per CU, then the second).  The heaviest groups
  // (by the summed degree of their variables) get the CUs that hold a single
  // workgroup, and every other heavy group shares its CU with a light one, so
  // that the kernel does not wait on two heavy groups sharing a CU.  With more
  // groups than slots, heavy groups go first.  (C3: 8.1 -> 7.1 ms per pass.)
  // The kernel reads the group of each slot from gorder; the test list keeps
  // its order.
  std::vector<int> gorder;
  if (wide && groups > 1 && groups <= kSliceBudgetBytes / (8ll * m->dm.n * m->dm.dpad * M)) {
    const int cus = std::max(1, m->props.cus);
    std::vector<int> rank((size_t)groups);
    for (long long g = 0; g < groups; ++g) rank[(size_t)g] = (int)g;
    std::stable_sort(rank.begin(), rank.end(), [&](int a, int b) { return w[(size_t)a] > w[(size_t)b]; });
    const int B = (int)groups;
    std::vector<int> at_bid(rank);  // more than two per CU, or one: heaviest first
    if (B > cus && B <= 2 * cus) {
      int next = 0;
      for (int k = B - cus; k < cus; ++k) at_bid[(size_t)k] = rank[(size_t)next++];
      for (int k = 0; k < B - cus; ++k) {
        at_bid[(size_t)k] = rank[(size_t)(next + k)];
        at_bid[(size_t)(k + cus)] = rank[(size_t)(B - 1 - k)];
      }
    }
    // the kernel takes slot gi = xcd * q + min(xcd, rem) + (bid >> 3) for bid
    const int q = B >> 3, rem = B & 7;
    gorder.resize((size_t)B);
    for (int bid = 0; bid < B; ++bid) {
      const int xcd = bid & 7;
      gorder[(size_t)(xcd * q + (xcd < rem ? xcd : rem) + (bid >> 3))] = at_bid[(size_t)bid];
    }
  }
  if (count > S.scap) {
    S.stests = DevAlloc(sizeof(int2) * count);
    S.perm = DevAlloc(sizeof(int) * count);
    S.scap = count;
  }
  if (count > S.cap) {
    S.fails = DevAlloc(count);
    S.tests = DevAlloc(sizeof(int2) * count);
    S.cap = count;
  }
  if (groups + 1 > S.gcap) {
    S.gfirst = DevAlloc(sizeof(int) * (groups + 1));
    S.gorder = DevAlloc(sizeof(int) * (groups + 1));
    S.gcap = groups + 1;
  }
  if (!S.revs.p) S.revs = DevAlloc(sizeof(unsigned long long) * kStatSlots);  // per-slot counters
  S.events();
  HIPC(hipMemcpyAsync(S.stests.p, sorted, sizeof(int2) * count, hipMemcpyHostToDevice, m->stream));
  if (!sorted_in)
    HIPC(hipMemcpyAsync(S.perm.p, order.data(), sizeof(int) * count, hipMemcpyHostToDevice, m->stream));
  HIPC(hipMemcpyAsync(S.gfirst.p, gfirst.data(), sizeof(int) * (groups + 1), hipMemcpyHostToDevice, m->stream));
  if (!gorder.empty())
    HIPC(hipMemcpyAsync(S.gorder.p, gorder.data(), sizeof(int) * gorder.size(), hipMemcpyHostToDevice, m->stream));
  const int* go = gorder.empty() ? nullptr : S.gorder.as<int>();
  HIPC(hipMemsetAsync(S.revs.p, 0, sizeof(unsigned long long) * kStatSlots, m->stream));
  const long long slice_bytes = 8ll * m->dm.n * m->dm.dpad * M;
  const long long per_launch = std::max<long long>(1, std::min<long long>(groups, kSliceBudgetBytes / slice_bytes));
  u64* slices = nullptr;
  if (wide) {
    if (slice_bytes * per_launch > S.dead_bytes) {
      S.dead = DevAlloc((size_t)(slice_bytes * per_launch));
      S.dead_bytes = slice_bytes * per_launch;
      HIPC(hipMemsetAsync(S.dead.p, 0, (size_t)S.dead_bytes, m->stream));
    }
    slices = S.dead.as<u64>();
  } else {
    slices = reinterpret_cast<u64*>(m->get_scratch(slice_bytes * per_launch));
  }
  const int L = m->current_level;
  DevNet net = m->dm.net();
  int threads = 0;
  size_t lds = 0;
  if (wide) {
    // 16 waves from 128 tests up, 8 below; at least one thread per test, at
    // most the kernel's launch bound
    threads = knob("sac_slice_threads") ? knob("sac_slice_threads") : (M >= 2 ? 1024 : 512);
    threads = std::min(std::max(threads, 64 * M), kWideMaxThreads);
    lds = (size_t)wide_lds_bytes(m->dm.n, m->dm.W, M, nib, nb);
  } else {
    threads = slice_threads_for(m);
    lds = (size_t)slice_lds_bytes(m->dm.n, m->dm.dpad, threads / 64);
  }
  const int* perm = sorted_in ? nullptr : S.perm.as<int>();
  HIPC(hipEventRecord(S.ev[0], m->stream));
  for (long long g0 = 0; g0 < groups; g0 += per_launch) {
    const long long blocks = std::min(per_launch, groups - g0);
    const dim3 grid((unsigned)blocks), block(threads);
    const u64* base = m->level_dom(L);
    const int2* tl = S.stests.as<int2>();
    const int* gf = S.gfirst.as<int>();
    unsigned char* fl = S.fails.as<unsigned char>();
    unsigned long long* rv = S.revs.as<unsigned long long>();
    if (!wide) hipLaunchKernelGGL(k_sliced_tests, grid, block, lds, m->stream, net, base, tl, gf, g0, perm, fl, rv, slices);
#define CPIM_WIDE(MM, NIB, NB) \
  hipLaunchKernelGGL((k_sliced_wide<MM, NIB, NB>), grid, block, lds, m->stream, net, base, tl, gf, go, g0, perm, fl, rv, slices)
#define CPIM_WIDE_M(MM)                                \
  {                                                    \
    if (nb == 6) CPIM_WIDE(MM, true, 6);               \
    else if (nb == 4) CPIM_WIDE(MM, true, 4);          \
    else if (nib) CPIM_WIDE(MM, true, 0);              \
    else CPIM_WIDE(MM, false, 0);                      \
  }
    else if (M == 1) CPIM_WIDE_M(1)
    else if (M == 2) CPIM_WIDE_M(2)
    else CPIM_WIDE_M(4)
#undef CPIM_WIDE_M
#undef CPIM_WIDE
    HIPC(hipGetLastError());
  }
  HIPC(hipEventRecord(S.ev[1], m->stream));
  HIPC(hipMemcpyAsync(fail_host, S.fails.p, count, hipMemcpyDeviceToHost, m->stream));
  unsigned long long revs_slots[kStatSlots];
  HIPC(hipMemcpyAsync(revs_slots, S.revs.p, sizeof(revs_slots), hipMemcpyDeviceToHost, m->stream));
  m->sync();
  const unsigned long long revs = slot_sum(revs_slots);
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
  st->kernel_ms += ms;
  st->revisions += (int64_t)revs;
  st->singleton_tests += count;
}

void run_singleton_plain(cpim_gmodel* m, const int2* tests, long long count, unsigned char* fail_host,
                         cpim_sac_stats* st);

// CPIM_SAC_BATCHED picks the bit-sliced kernel when the network allows it.
void run_singleton_device(cpim_gmodel* m, int mode, const int2* tests, long long count, unsigned char* fail_out,
                          cpim_sac_stats* st, SacScratch& S) {
  if (mode == CPIM_SAC_BATCHED) mode = sliced_eligible(m) ? CPIM_SAC_SLICED : CPIM_SAC_WORKGROUP;
  if (mode == CPIM_SAC_SLICED) {
    if (!sliced_eligible(m))
      fail(CPIM_ERR_UNSUPPORTED, "bit-sliced singleton tests need a binary network whose work lists fit in LDS");
    run_singleton_sliced(m, tests, count, fail_out, st, S);
  } else if (mode == CPIM_SAC_WORKGROUP) {
    run_singleton_batched(m, tests, count, fail_out, st, S);
  } else {
    run_singleton_plain(m, tests, count, fail_out, st);
  }
}

// One host-driven fixpoint per test on a scratch level: copy the level,
// reduce x to {a}, seed the constraints of x, rounds with a sync each.
void run_singleton_plain(cpim_gmodel* m, const int2* tests, long long count, unsigned char* fail_host,
                         cpim_sac_stats* st) {
  const int L = m->current_level;
  if (L + 1 >= m->max_depth) fail(CPIM_ERR_RUNTIME, "no free level for singleton tests");
  const int T = L + 1;
  const long long t0 = now_ns();
  std::vector<u64> word(m->dm.W);
  const int one = 1;
  for (long long i = 0; i < count; ++i) {
    const int x = tests[i].x, a = tests[i].y;
    HIPC(hipMemcpyAsync(m->level_dom(T), m->level_dom(L), sizeof(u64) * m->level_words(),
                        hipMemcpyDeviceToDevice, m->stream));
    HIPC(hipMemcpyAsync(m->level_size(T), m->level_size(L), sizeof(int) * m->dm.n, hipMemcpyDeviceToDevice,
                        m->stream));
    std::fill(word.begin(), word.end(), 0ull);
    word[a >> 6] = 1ull << (a & 63);
    HIPC(hipMemcpyAsync(m->level_dom(T) + (long long)x * m->dm.W, word.data(), sizeof(u64) * m->dm.W,
                        hipMemcpyHostToDevice, m->stream));
    HIPC(hipMemcpyAsync(m->level_size(T) + x, &one, sizeof(int), hipMemcpyHostToDevice, m->stream));
    m->current_level = T;
    cpim_gac_stats gs;
    std::memset(&gs, 0, sizeof(gs));
    try {
      gac_grid(m, x, false, &gs);
    } catch (...) {
      m->current_level = L;
      throw;
    }
    m->current_level = L;
    fail_host[i] = gs.inconsistent ? 1 : 0;
    st->revisions += gs.revisions;
  }
  st->kernel_ms += (now_ns() - t0) * 1e-6;
  st->singleton_tests += count;
}

std::vector<int2> all_tests(cpim_gmodel* m, const std::vector<u64>& dom) {
  std::vector<int2> t;
  for (int v = 0; v < m->dm.n; ++v)
    for (int a = 0; a < m->im.cap[v]; ++a)
      if ((dom[(long long)v * m->dm.W + (a >> 6)] >> (a & 63)) & 1ull) t.push_back(make_int2(v, a));
  return t;
}

std::vector<u64> read_level(cpim_gmodel* m, int level) {
  std::vector<u64> d(m->level_words());
  HIPC(hipMemcpyAsync(d.data(), m->level_dom(level), sizeof(u64) * d.size(), hipMemcpyDeviceToHost, m->stream));
  m->sync();
  return d;
}

}  // namespace

extern "C" {

int cpim_tune_set(const char* name, int64_t value) {
  return guarded([&] {
    REQUIRE(name, "null argument");
    Knob* k = find_knob(name);
    REQUIRE(k, "unknown tuning knob");
    REQUIRE(knob_valid(name, value), "value out of range for this knob");
    k->value = (int)value;
  });
}

int cpim_tune_get(const char* name, int64_t* value) {
  return guarded([&] {
    REQUIRE(name && value, "null argument");
    Knob* k = find_knob(name);
    REQUIRE(k, "unknown tuning knob");
    *value = k->value.load();
  });
}

int cpim_gmodel_singleton_tests(cpim_gmodel* m, const int32_t* tests, int64_t count, int32_t mode, uint8_t* fail_out,
                                cpim_sac_stats* st) {
  return guarded([&] {
    REQUIRE(m && st && (count == 0 || (tests && fail_out)), "null argument");
    REQUIRE(count >= 0, "negative test count");
    DeviceGuard g(m->device);
    std::memset(st, 0, sizeof(*st));
    const int2* t = reinterpret_cast<const int2*>(tests);
    for (long long i = 0; i < count; ++i) {
      REQUIRE(t[i].x >= 0 && t[i].x < m->dm.n, "test names an unknown variable");
      REQUIRE(t[i].y >= 0 && t[i].y < m->im.cap[t[i].x], "test value outside the domain");
    }
    REQUIRE(mode >= CPIM_SAC_BATCHED && mode <= CPIM_SAC_SLICED, "unknown singleton-test mode");
    SacScratch& S = m->sac;
    run_singleton_device(m, mode, t, count, fail_out, st, S);
    st->consistent = 1;
    st->passes = 1;
  });
}

// SAC closure by passes: every singleton test of the current domains (one
// batched launch), removal of the failed values, GAC from the variables that
// lost values, until a pass removes nothing.  The closure is unique, so the
// result equals SAC1::enforce (src/SAC1.cpp:57-103).
int cpim_gmodel_enforce_sac(cpim_gmodel* m, int32_t mode, cpim_sac_stats* st) {
  return guarded([&] {
    REQUIRE(m && st, "null argument");
    REQUIRE(mode >= CPIM_SAC_BATCHED && mode <= CPIM_SAC_SLICED, "unknown SAC mode");
    DeviceGuard g(m->device);
    std::memset(st, 0, sizeof(*st));
    const int L = m->current_level;
    cpim_gac_stats gs;
    gac(m, -1, CPIM_GAC_AUTO, &gs);
    st->revisions += gs.revisions;
    if (gs.inconsistent) {
      st->consistent = 0;
      return;
    }
    SacScratch& S = m->sac;
    while (true) {
      std::vector<u64> dom = read_level(m, L);
      std::vector<int2> tests = all_tests(m, dom);
      std::vector<unsigned char> fl(tests.size(), 0);
      run_singleton_device(m, mode, tests.data(), (long long)tests.size(), fl.data(), st, S);
      ++st->passes;
      long long removed = 0;
      for (size_t i = 0; i < tests.size(); ++i)
        if (fl[i]) {
          dom[(long long)tests[i].x * m->dm.W + (tests[i].y >> 6)] &= ~(1ull << (tests[i].y & 63));
          ++removed;
        }
      if (removed == 0) break;
      st->removed += removed;
      // a variable whose every value failed its singleton test has an empty
      // domain: that is a wipe-out even if no constraint reaches it
      bool empty = false;
      for (int v = 0; v < m->dm.n && !empty; ++v) {
        u64 any = 0;
        for (int w = 0; w < m->dm.W; ++w) any |= dom[(long long)v * m->dm.W + w];
        empty = any == 0;
      }
      HIPC(hipMemcpyAsync(m->level_dom(L), dom.data(), sizeof(u64) * dom.size(), hipMemcpyHostToDevice, m->stream));
      hipLaunchKernelGGL(k_sizes, dim3((m->dm.n + 255) / 256), dim3(256), 0, m->stream, m->level_dom(L),
                         m->level_size(L), m->dm.n, m->dm.W);
      HIPC(hipGetLastError());
      m->sync();
      if (empty) {
        st->consistent = 0;
        return;
      }
      gac(m, -1, CPIM_GAC_AUTO, &gs);
      st->revisions += gs.revisions;
      if (gs.inconsistent) {
        st->consistent = 0;
        return;
      }
    }
    st->consistent = 1;
  });
}

// ---------------------------------------------------------------------------
// Device MAC
// ---------------------------------------------------------------------------
int cpim_gmodel_solve(cpim_gmodel* m, const cpim_solve_opts* opts, cpim_search_stats* st, int32_t* solution,
                      cpim_node_record* trace, int32_t* trace_len) {
  return guarded([&] {
    REQUIRE(m && st, "null argument");
    cpim_solve_opts o{};
    if (opts) o = *opts;
    REQUIRE(o.branching == CPIM_BRANCH_BINARY, "the device solver implements binary branching only");
    REQUIRE(o.var_heuristic == CPIM_VARH_DOM_MIN || o.var_heuristic == CPIM_VARH_DOM_DEG ||
                o.var_heuristic == CPIM_VARH_DOM_WDEG,
            "unknown variable heuristic");
    REQUIRE(o.reserved == 0, "cpim_solve_opts.reserved must be 0");
    REQUIRE(o.trace_capacity >= 0, "negative trace capacity");
    REQUIRE(o.trace_capacity == 0 || trace, "trace buffer is null");
    REQUIRE(o.max_solutions >= 0 && o.solution_capacity >= 0, "negative solution count");
    REQUIRE(m->current_level == 0, "the device solver starts from level 0");
    DeviceGuard g(m->device);
    std::memset(st, 0, sizeof(*st));
    const int n = m->dm.n;
    const long long t0 = now_ns();
    const long long sol_cap = std::max<long long>(1, o.solution_capacity);
    DevAlloc state(sizeof(MacState)), stack(sizeof(int2) * n), sol(sizeof(int) * n * sol_cap);
    // the search keeps its depth-0 trail in level 0 and rewrites it when it
    // backtracks at the root; level 0 gets the root closure back afterwards,
    // as the reference leaves it (GModelSolver::Solve runs EnforceGAC on
    // level 0 and searches on the levels above)
    DevAlloc root(sizeof(u64) * m->level_words());
    DevAlloc tr(sizeof(NodeRec) * std::max(1, o.trace_capacity));
    MacState s0{};
    s0.status = MAC_INIT;
    HIPC(hipMemcpyAsync(state.p, &s0, sizeof(s0), hipMemcpyHostToDevice, m->stream));
    MacArgs A{};
    A.state = state.as<MacState>();
    A.trail = m->level_dom(0);
    A.stack = stack.as<int2>();
    A.solution = sol.as<int>();
    A.sol_cap = sol_cap;
    A.max_solutions = o.max_solutions;
    A.trace = o.trace_capacity > 0 ? tr.as<NodeRec>() : nullptr;
    A.trace_cap = o.trace_capacity;
    A.node_limit = o.node_limit;
    A.step_budget = 0;
    A.deadline = 0;
    A.root_out = root.as<u64>();
    A.nbr = o.var_heuristic == CPIM_VARH_DOM_DEG ? m->dm.nbr.as<int>() : nullptr;
    if (o.var_heuristic == CPIM_VARH_DOM_WDEG) {
      if (!m->wdeg_weight.p) m->wdeg_init();
      A.weight = m->wdeg_weight.as<float>();
      A.stamp_tab = m->wdeg_stamp.as<long long>();
      A.clock = m->wdeg_clock.as<long long>();
    }
    DevNet net = m->dm.net();
    const long long bytes = wg_mem_bytes(n, m->dm.W);
    const int threads = 1024;
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, m->stream));
    // Every launch runs a bounded number of nodes (so no launch can run for
    // long); the search state stays on the device between launches and the
    // host checks the time limit in between.
    A.step_budget = kNodesPerLaunch;
    MacState s{};
    // The network image goes to LDS too when it fits beside the domains.
    LdsNet lnet{};
    if (knob("mac_lds_net") == 0 && m->dm.lnet.bytes > 0 &&
        bytes + m->dm.lnet.bytes <= (long long)m->props.max_lds - 2048)
      lnet = m->dm.lnet;
    while (true) {
      if (bytes <= kLdsBudget) {
        hipLaunchKernelGGL(k_mac<true>, dim3(1), dim3(threads), (size_t)(bytes + lnet.bytes), m->stream, net,
                           A, (char*)nullptr, lnet);
      } else {
        char* scratch = m->get_scratch(bytes);
        hipLaunchKernelGGL(k_mac<false>, dim3(1), dim3(threads), 0, m->stream, net, A, scratch, LdsNet{});
      }
      HIPC(hipGetLastError());
      HIPC(hipMemcpyAsync(&s, state.p, sizeof(s), hipMemcpyDeviceToHost, m->stream));
      m->sync();
      if (s.status != MAC_RUNNING) break;
      if (o.time_limit_ms > 0 && (now_ns() - t0) / 1000000 > o.time_limit_ms) {
        st->time_out = 1;
        break;
      }
    }
    HIPC(hipEventRecord(e1, m->stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st->num_positive = s.npos;
    st->num_negative = s.nneg;
    st->num_solutions = s.nsol;
    st->unsolvable = s.status == MAC_UNSAT && s.nsol == 0;
    if (s.status == MAC_LIMIT) st->time_out = 1;
    st->gac_iterations = s.npos + s.nneg + 1;
    st->solve_time = (now_ns() - t0) * 1e-9;
    st->gac_time = ms * 1e-3;
    if (solution) {
      const long long kept = std::min<long long>(s.nsol, sol_cap);
      if (kept > 0)
        HIPC(hipMemcpyAsync(solution, sol.p, sizeof(int) * n * kept, hipMemcpyDeviceToHost, m->stream));
      for (long long i = kept * n; i < n * (o.solution_capacity > 0 ? sol_cap : 1); ++i) solution[i] = -1;
    }
    const int tl = std::min(s.trace_len, o.trace_capacity);
    if (tl > 0) {
      std::vector<NodeRec> h(tl);
      HIPC(hipMemcpyAsync(h.data(), tr.p, sizeof(NodeRec) * tl, hipMemcpyDeviceToHost, m->stream));
      m->sync();
      for (int i = 0; i < tl; ++i) {
        trace[i].positive = h[i].positive;
        trace[i].var = h[i].var;
        trace[i].value = h[i].value;
        trace[i].consistent = h[i].consistent;
        trace[i].dom_hash = h[i].dom_hash;
      }
    }
    if (trace_len) *trace_len = s.trace_len;
    HIPC(hipMemcpyAsync(m->level_dom(0), root.p, sizeof(u64) * m->level_words(), hipMemcpyDeviceToDevice,
                        m->stream));
    hipLaunchKernelGGL(k_sizes, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->level_dom(0), m->level_size(0), n,
                       m->dm.W);
    HIPC(hipGetLastError());
    m->sync();
  });
}

}  // extern "C"

#include "batch.inc"

#include "noalloc.inc"
