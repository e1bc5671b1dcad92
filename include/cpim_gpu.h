/*
 * cpim_gpu.h - C-ABI of the MI355X (gfx950) GAC/SAC propagation engine.
 *
 * This is the drop-in boundary for the GPU propagation path of cpim.  Every
 * entry point takes plain pointers and sizes; no C++ or HIP types cross it.
 * All calls are synchronous (they return after the device work has finished),
 * like the reference, and each handle owns one HIP stream.  A handle must not
 * be used from two host threads at once.
 *
 * Error convention: every function returns CPIM_OK (0) or a CPIM_ERR_* code;
 * cpim_last_error() returns the message of the last failure on the calling
 * thread.  The C++ facade (include/cpim/gmodel.hpp) turns CPIM_ERR_INVALID_ARGUMENT
 * into std::invalid_argument and every other code into std::runtime_error,
 * with the same message text as the reference.
 *
 * Value encoding: values are canonical indices 0..dom_size-1 (the reference
 * normalizes every model that way, src/model/model_normalizer.cpp:174-316).
 * Domains are bitsets of 64-bit words, LSB first: value v is bit (v % 64) of
 * word (v / 64).  A 64-bit word w holds the reference's 32-bit words 2w (low
 * half) and 2w+1 (high half), so both layouts agree bit for bit.
 */
#ifndef CPIM_GPU_H_
#define CPIM_GPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPIM_OK 0
#define CPIM_ERR_INVALID_ARGUMENT 1
#define CPIM_ERR_RUNTIME 2
#define CPIM_ERR_HIP 3
#define CPIM_ERR_OUT_OF_MEMORY 4
#define CPIM_ERR_UNSUPPORTED 5
#define CPIM_ERR_BUSY 6

/* ------------------------------------------------------------------------ */
/* Model description (normalized constraint network, plain arrays)           */
/* ------------------------------------------------------------------------ */

/*
 * A normalized extensional constraint network.  Replaces the reference's
 * IntermediateModel as the input of GModelAdapter::Build
 * (include/model/gmodel_adapter.h:42-43, src/model/gmodel_adapter.cu:44-312):
 * only the normalized content that Build reads is passed (domain sizes,
 * scopes and support relations); the format layer stays with the caller.
 *
 * Relations are positive (supports).  A binary constraint takes its relation
 * from bin_rel when bin_rel is not NULL, else from the tuple list; a
 * constraint of arity >= 3 always takes it from the tuple list.
 */
typedef struct cpim_model_desc {
  int32_t num_vars;
  const int32_t* dom_size;  /* [num_vars], each >= 1 */
  int32_t num_cons;
  const int32_t* con_arity; /* [num_cons], each >= 2 */
  const int32_t* con_scope; /* concatenated scopes, sum(con_arity) variable ids */
  /* Dense rows of the binary constraints, in their order among the binary
   * constraints: bin_rel[k][a][0..rel_words) is the set of values b of
   * scope[1] such that (scope[0] = a, scope[1] = b) is allowed.  The row
   * count per constraint is max_dom = max(dom_size). */
  const uint64_t* bin_rel;
  int32_t rel_words;        /* words per bin_rel row, >= ceil(max_dom / 64) */
  /* Support tuples: tuples[tup_off[c] .. tup_off[c+1]) holds the tuples of
   * constraint c, con_arity[c] ints per tuple (offsets in int32 units). */
  const int64_t* tup_off;   /* [num_cons + 1] or NULL */
  const int32_t* tuples;
  /* Optional initial domains [num_vars][ceil(max_dom/64)]; NULL = full. */
  const uint64_t* init_dom;
} cpim_model_desc;

/* GModelOptions (include/model/gmodel_adapter.h:19-28). */
typedef struct cpim_gmodel_options {
  int32_t device_id;       /* default 0 */
  int32_t enable_prefetch; /* accepted, no effect: all memory is device memory */
  int32_t skip_non_binary; /* default 1: constraints of arity >= 3 are ignored
                              (reference behaviour); 0: they are propagated as
                              bitset tables (the reference throws instead) */
  int32_t skip_conflicts;  /* accepted, no effect (conflicts are normalized away) */
} cpim_gmodel_options;

void cpim_gmodel_options_default(cpim_gmodel_options* opts);

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

const char* cpim_last_error(void);
const char* cpim_version(void);

typedef struct cpim_device_info {
  int32_t compute_units;
  int32_t lds_bytes_per_block;
  int32_t cooperative_launch;
  int32_t wavefront_size;
  int64_t global_mem_bytes;
  char arch[64];
} cpim_device_info;

int cpim_get_device_info(int32_t device_id, cpim_device_info* out);

/* Measurement utility (no counterpart in the reference): copy `bytes` bytes
 * between two device buffers with a streaming kernel, `reps` times, and return
 * the best achieved bandwidth (read + write bytes per second, GB/s).  It is
 * the reference kernel of known traffic for scaling hardware counters and the
 * measured stream-copy peak reported next to the 8 TB/s vendor figure. */
int cpim_measure_copy_bandwidth(int32_t device_id, int64_t bytes, int32_t reps, double* gbps);

/* ------------------------------------------------------------------------ */
/* GModel: device-resident model, domain level stack and propagation         */
/* ------------------------------------------------------------------------ */

typedef struct cpim_gmodel cpim_gmodel;

/* GModelAdapter::Build (src/model/gmodel_adapter.cu:44-312).  Copies the
 * model to the device: support planes sup[c][dir][word][value], the
 * variable->constraint CSR, and max_depth = num_vars + 1 domain levels with
 * level 0 = init_dom (or full domains). */
int cpim_gmodel_create(const cpim_model_desc* desc, const cpim_gmodel_options* opts,
                       cpim_gmodel** out);
/* ~GModel (src/GModel.cu:122-268). NULL is accepted. */
int cpim_gmodel_destroy(cpim_gmodel* model);

typedef struct cpim_gmodel_info {
  int32_t num_vars;
  int32_t num_constraints;   /* all constraints given */
  int32_t num_active;        /* constraints propagated (binary + kept tables) */
  int32_t max_dom_size;
  int32_t dom_words;         /* 64-bit words per domain */
  int32_t bit_dom_int_size;  /* 32-bit words per domain (reference layout) */
  int32_t max_depth;         /* num_vars + 1 */
  int32_t subscription_size; /* CSR entries */
  int64_t device_bytes;      /* device memory held by the model */
} cpim_gmodel_info;

int cpim_gmodel_get_info(const cpim_gmodel* model, cpim_gmodel_info* out);

/* GacStats (include/GModel.cuh:17-21).  deletions is the popcount delta of
 * the level (never double counted).  revisions counts executed arc revisions.
 * kernel_ms is the device time of the fixpoint's kernels, measured with HIP
 * events on the model's stream (summed over the rounds of host-driven rounds,
 * so the host round trips between them are not in it). */
typedef struct cpim_gac_stats {
  int32_t iterations;
  int32_t deletions;
  int32_t inconsistent;
  int32_t converged;
  int64_t revisions;
  double kernel_ms;
} cpim_gac_stats;

#define CPIM_GAC_PLAIN 0      /* host-driven rounds, one launch + sync per round */
#define CPIM_GAC_PERSISTENT 1 /* all rounds in one launch of co-resident workgroups, a grid
                                 barrier between rounds; every barrier wait is bounded by the
                                 fixpoint_budget_us knob (CPIM_ERR_RUNTIME when it runs out) */
#define CPIM_GAC_BLOCK 2      /* one workgroup, domains in LDS, one launch */
#define CPIM_GAC_AUTO 3       /* autotuned choice among the modes that run */
#define CPIM_GAC_ASYNC 4      /* one launch, whole GPU, asynchronous work queue,
                                 no grid barrier */

/* GModel::EnforceGAC / EnforceGAC_Persistent (include/GModel.cuh:191-197,
 * src/GModel.cu:1110-1391) on the current level.  assigned_var < 0 seeds all
 * constraints; assigned_var >= 0 seeds the constraints of that variable.
 * Never fails on inconsistency: that is reported in stats->inconsistent. */
int cpim_gmodel_enforce_gac(cpim_gmodel* model, int32_t assigned_var, int32_t mode,
                            cpim_gac_stats* stats);

/* No counterpart in the reference.  The persistent mode of
 * cpim_gmodel_enforce_gac for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): it
 * allocates nothing.  Its scratch (control block and frontier bitmaps) is a
 * file-scope device array of 776 KiB per GPU, held until the process ends, so
 * it takes networks of at most 2,097,152 propagated constraints
 * (CPIM_ERR_INVALID_ARGUMENT above that).  Calls that overlap on one GPU race
 * on that array.  stats->kernel_ms is 0 (no events are created). */
int cpim_gmodel_enforce_gac_noalloc(cpim_gmodel* model, int32_t assigned_var, cpim_gac_stats* stats);

/* Level / trail API (include/GModel.cuh:219-253, src/GModel.cu:1397-1701),
 * same return conventions: results are written to *out. */
int cpim_gmodel_create_new_level(cpim_gmodel* model, int32_t* new_level);
int cpim_gmodel_back_to_level(cpim_gmodel* model, int32_t level);
int cpim_gmodel_current_level(const cpim_gmodel* model, int32_t* level);
int cpim_gmodel_assign_value(cpim_gmodel* model, int32_t var, int32_t value, int32_t level,
                             int32_t* ok);
int cpim_gmodel_remove_value(cpim_gmodel* model, int32_t var, int32_t value, int32_t level,
                             int32_t* ok);
int cpim_gmodel_domain_size(const cpim_gmodel* model, int32_t var, int32_t level, int32_t* size);
int cpim_gmodel_min_domain_var(const cpim_gmodel* model, int32_t level, int32_t* var);
int cpim_gmodel_is_assigned(const cpim_gmodel* model, int32_t var, int32_t level, int32_t* res);
int cpim_gmodel_is_fully_assigned(const cpim_gmodel* model, int32_t level, int32_t* res);
int cpim_gmodel_assigned_value(const cpim_gmodel* model, int32_t var, int32_t level,
                               int32_t* value);
int cpim_gmodel_first_value(const cpim_gmodel* model, int32_t var, int32_t level,
                            int32_t* value);
int cpim_gmodel_next_value(const cpim_gmodel* model, int32_t var, int32_t value, int32_t level,
                           int32_t* next);

/* Host view of one level (replaces the reference's host dereference of
 * managed bitDom, e.g. test/gmodel_solver.cpp:126-140). */
int cpim_gmodel_read_domains(const cpim_gmodel* model, int32_t level, uint64_t* out);
int cpim_gmodel_write_domains(cpim_gmodel* model, int32_t level, const uint64_t* in);

/* ------------------------------------------------------------------------ */
/* SAC (the reference's CModel::enforceSAC, src/cuSAC.cu:335, is empty; the   */
/* semantics are those of SAC1::enforce, src/SAC1.cpp:57-103)                */
/* ------------------------------------------------------------------------ */

typedef struct cpim_sac_stats {
  int32_t consistent;
  int32_t passes;
  int64_t singleton_tests;
  int64_t removed;
  int64_t revisions;
  double kernel_ms;
} cpim_sac_stats;

/* Tuning knobs of the engine, process-wide; 0 = automatic.  No counterpart in
 * the reference.  Names and values:
 *   "sac_slice_masks"   masks per (variable, value) of the bit-sliced SAC kernel: 0, 1, 2, 4
 *                       (64 tests per mask)
 *   "sac_slice_threads" threads per workgroup of that kernel: 0 or 64..1024 (multiple of 64)
 *   "sac_slice_kernel"  0 = one shared stage per changed variable, 1 = one stage per wave
 *   "sac_slice_nibbles" nibble (four-Russians) tables in the shared stage: 0 = when they fit
 *                       in 80 KB of LDS, 1 = whenever they fit, 2 = never
 *   "sac_slice_bytes"   256-subset tables for the first bytes of each staged row in front of
 *                       the nibble tables (one stage buffer instead of two): 4 or 6 = that
 *                       many whenever they fit, 0 and 1 = none
 *   "unit_threads"      threads per workgroup of the batch / serving MAC kernel
 *   "mac_lds_net"       device MAC: 0 = copy the network (adjacency, planes, tables) into LDS
 *                       when it fits, 1 = read it from device memory
 *   "fixpoint_budget_us" device wall-clock budget of one persistent or asynchronous fixpoint
 *                       launch in microseconds (0 = 20 s); a launch that exceeds it stops and
 *                       the call returns CPIM_ERR_RUNTIME
 * First values come from CPIM_SLICE_MASKS, CPIM_SLICE_THREADS, CPIM_SLICE_KERNEL,
 * CPIM_SLICE_NIBBLES, CPIM_UNIT_THREADS, CPIM_MAC_LDS_NET, CPIM_FIXPOINT_BUDGET_US,
 * CPIM_SLICE_BYTES. */
int cpim_tune_set(const char* name, int64_t value);
int cpim_tune_get(const char* name, int64_t* value);

#define CPIM_SAC_BATCHED 0   /* every singleton test of a pass on the device: SLICED when the network
                                allows it (binary constraints only), else WORKGROUP */
#define CPIM_SAC_PLAIN 1     /* one host-driven fixpoint per singleton test */
#define CPIM_SAC_WORKGROUP 2 /* one workgroup per singleton test, domain copy in LDS */
#define CPIM_SAC_SLICED 3    /* 64 singleton tests per workgroup, bit-sliced; CPIM_ERR_UNSUPPORTED
                                for networks with table constraints */

/* SAC closure of the current level, in place.  stats->consistent = 0 on
 * wipe-out (the level then holds no meaningful domains). */
int cpim_gmodel_enforce_sac(cpim_gmodel* model, int32_t mode, cpim_sac_stats* stats);

/* Singleton tests (x = a) on copies of the current level: fail[i] = 1 iff the
 * GAC closure of the copy with var tests[2i] reduced to tests[2i+1] is empty.
 * The current level is not modified. */
int cpim_gmodel_singleton_tests(cpim_gmodel* model, const int32_t* tests, int64_t count,
                                int32_t mode, uint8_t* fail, cpim_sac_stats* stats);

/* ------------------------------------------------------------------------ */
/* Search (GModelSolver::Solve, src/GModelSolver.cu:31-102; node semantics of */
/* the CPU MAC::enforce, src/MAC.cpp:100-165)                                 */
/* ------------------------------------------------------------------------ */

#define CPIM_BRANCH_BINARY 0 /* CPU MAC: x = a / x != a, dom-min, min value */

/* Variable ordering, MAC::select_var (src/MAC.cpp:353-411): the undecided
 * variable with the smallest key, the lowest id on ties.
 *   DOM_MIN   key = |D(x)|                        (Heuristic::VRH_DOM_MIN)
 *   DOM_DEG   key = |D(x)| / |neighbours(x)|, in integer division, or -1 for
 *             a variable that shares no constraint (Heuristic::VRH_DOM_DEG_MIN)
 *   DOM_WDEG  key = |D(x)| / x_w in double, x_w = the summed weights of the
 *             constraints on x with another undecided variable; -1 when
 *             |D(x)| = 1 or x_w = 0 (Heuristic::VRH_DOM_WDEG_MIN).  Weights
 *             start at 1 with the model, persist across solves (read and
 *             reset them with cpim_gmodel_weights) and grow by one on the
 *             constraint that empties a domain first in AC3's order
 *             (src/AC3.cpp:122-126);
 *             with this ordering every fixpoint of the search runs in that
 *             order, on one wave, so it is slower than the other two.
 * Neighbours are the distinct other variables of the constraints on x
 * (Network::get_neighbor, src/Network.cpp:467-474), counted over the
 * constraints the model propagates. */
#define CPIM_VARH_DOM_MIN 0
#define CPIM_VARH_DOM_DEG 1
#define CPIM_VARH_DOM_WDEG 2

/* dom/wdeg constraint weights (Tabular::weight, include/Network.h:243): one
 * float per propagated constraint, in constraint order; they start at 1 with
 * the model (Network.cpp:303) and persist across cpim_gmodel_solve calls,
 * as the reference's weights persist in its network; a solve under
 * CPIM_VARH_DOM_WDEG raises them (src/AC3.cpp:122-126).  Copies up to
 * `capacity` weights into out (out may be NULL when capacity is 0) and sets
 * *count to the number of propagated constraints; reset != 0 then sets every
 * weight back to 1.  No counterpart in the reference's GPU interface (its
 * weights are a public member of each constraint). */
int cpim_gmodel_weights(cpim_gmodel* model, float* out, int64_t capacity, int64_t* count, int32_t reset);

typedef struct cpim_solve_opts {
  int64_t time_limit_ms; /* 0 = unlimited */
  int64_t node_limit;    /* 0 = unlimited; checked where the time limit is */
  int32_t branching;     /* CPIM_BRANCH_BINARY */
  int32_t trace_capacity;/* records kept in trace (0 = none) */
  int32_t var_heuristic; /* CPIM_VARH_DOM_MIN (0), _DOM_DEG or _DOM_WDEG */
  int32_t reserved;      /* 0 */
  /* GModelSolver::SetMaxSolutions / SetFindAllSolutions (include/GModelSolver.h:48-54):
   * 0 or 1 = stop at the first solution; k > 1 = every solution found is
   * counted and stored, then treated as a failed node, until the search tree
   * is exhausted or k solutions are found (find-all: a large k) */
  int64_t max_solutions;
  /* solutions the `solution` buffer of cpim_gmodel_solve holds, num_vars ints
   * each, in the order found (0 = 1); stats->num_solutions counts them all */
  int64_t solution_capacity;
} cpim_solve_opts;

/* GpuSearchStatistics (include/GModelSolver.h:11-21) + SearchStatistics
 * (include/Solver.h:71-82).  Times are in seconds. */
typedef struct cpim_search_stats {
  int64_t num_solutions;
  int64_t num_positive;
  int64_t num_negative;
  int64_t gac_iterations;
  int64_t gac_deletions;
  int32_t time_out;
  int32_t unsolvable;
  double solve_time;
  double gac_time;
} cpim_search_stats;

/* One search node: a positive (x = a) or negative (x != a) decision and the
 * outcome of the propagation that followed it.  dom_hash is the order-free
 * hash of all domains after a consistent propagation (0 otherwise):
 *   sum over (v, w) of mix64(dom[v][w] + (v * dom_words + w) * 0x9E3779B97F4A7C15)
 * with mix64 the splitmix64 finalizer. */
typedef struct cpim_node_record {
  int32_t positive;
  int32_t var;
  int32_t value;
  int32_t consistent;
  uint64_t dom_hash;
} cpim_node_record;

/* Runs the whole MAC search on the device from the current level (one
 * launch; the level stack is the trail).  solution: [num_vars] or NULL.
 * trace: [opts->trace_capacity] or NULL; *trace_len receives the number of
 * nodes recorded. */
int cpim_gmodel_solve(cpim_gmodel* model, const cpim_solve_opts* opts, cpim_search_stats* stats,
                      int32_t* solution, cpim_node_record* trace, int32_t* trace_len);

/* ------------------------------------------------------------------------ */
/* Batches of uneven instances (root GAC + MAC nodes, one launch)             */
/* ------------------------------------------------------------------------ */

typedef struct cpim_batch cpim_batch;

typedef struct cpim_batch_result {
  int32_t root_consistent;
  int32_t status; /* 0 = node limit reached, 1 = solution, 2 = unsat */
  int64_t num_positive;
  int64_t num_negative;
  int32_t trace_len;
  int32_t pad;
} cpim_batch_result;

int cpim_batch_create(const cpim_model_desc* models, int32_t count,
                      const cpim_gmodel_options* opts, cpim_batch** out);
int cpim_batch_destroy(cpim_batch* batch);
/* mode: CPIM_GAC_PERSISTENT = every instance in one persistent launch;
 * CPIM_GAC_PLAIN = one launch per instance. */
int cpim_batch_run(cpim_batch* batch, int64_t node_limit, int32_t mode,
                   cpim_batch_result* results, cpim_node_record* traces, int32_t trace_cap,
                   double* kernel_ms);
/* Root-fixpoint domains of instance index ([num_vars][dom_words]). */
int cpim_batch_read_root(const cpim_batch* batch, int32_t index, uint64_t* out);

/* Units of the last cpim_batch_run (persistent mode, first launch) that ran
 * in a launch of their own with 1024 threads per unit, beside the others:
 * units whose estimated work (arcs x words x largest domain) exceeds 1.5
 * times the mean work per resident workgroup, in batches of more units than
 * CUs.  No counterpart in the reference. */
int cpim_batch_heavy_units(const cpim_batch* batch, int32_t* count);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of MAC solve requests                        */
/* ------------------------------------------------------------------------ */

typedef struct cpim_server cpim_server;

typedef struct cpim_server_opts {
  int32_t device_id;
  int32_t max_slots;        /* requests resident on the device at once */
  int64_t hbm_budget_bytes; /* admission limit on resident footprint */
  int64_t nodes_per_step;   /* MAC nodes each resident request runs per step */
} cpim_server_opts;

typedef struct cpim_request_result {
  int32_t done;
  int32_t status; /* 0 = node limit, 1 = solution, 2 = unsat */
  int64_t num_positive;
  int64_t num_negative;
  int64_t footprint_bytes;
  int32_t steps;
  int32_t pad;
} cpim_request_result;

int cpim_server_create(const cpim_server_opts* opts, cpim_server** out);
int cpim_server_destroy(cpim_server* server);
/* Queue a request; the model is copied, so desc may be freed on return. */
int cpim_server_submit(cpim_server* server, const cpim_model_desc* desc, int64_t node_limit,
                       int64_t* request_id);
/* Advance every resident request by nodes_per_step nodes in one launch and
 * retire the finished ones.  While that launch runs, queued requests are
 * admitted (FIFO, while a slot and hbm_budget_bytes allow; a request that
 * finds the device empty is admitted whatever its size) on a second stream;
 * they join the next step.  A request is charged exactly the device bytes it
 * takes: one allocation holding its network and search state.  *active
 * receives the number still resident or queued. */
int cpim_server_step(cpim_server* server, int32_t* active);
/* Ids of the requests finished since the last poll, oldest first (at most
 * cap; the rest stay for the next poll).  No counterpart in the reference. */
int cpim_server_poll(cpim_server* server, int64_t* ids, int32_t cap, int32_t* count);
/* Result of a request; solution [num_vars] may be NULL.  Returns
 * CPIM_ERR_BUSY while the request is not finished. */
int cpim_server_result(cpim_server* server, int64_t request_id, cpim_request_result* out,
                       int32_t* solution);
/* Drop a finished request's result (its solution and record); the server
 * keeps every finished result until it is released.  CPIM_ERR_BUSY while the
 * request is not finished. */
int cpim_server_release(cpim_server* server, int64_t request_id);
/* Device bytes charged to resident requests now and at most so far, and the
 * numbers of queued and resident requests (any pointer may be NULL). */
int cpim_server_memory(const cpim_server* server, int64_t* resident_bytes, int64_t* peak_bytes,
                       int32_t* queued, int32_t* resident);

#ifdef __cplusplus
}
#endif

#endif /* CPIM_GPU_H_ */
