/*
 * CPU oracle for the GAC/SAC/MAC path: TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline.  The product
 * (cpim_amd, libcpim_gpu.so) never links or calls it.
 *
 * Part 1 restates the reference's CPU algorithms (semantics, not code):
 *   IntVar / Tabular / Network        include/Network.h:142-380, src/Network.cpp:62-455
 *   AC3::enforce / revise / seek      src/AC3.cpp:3-14, 99-209
 *   AC3bit ctor + seek_support        src/AC3bit.cpp:3-38
 *   SAC1::enforce                     src/SAC1.cpp:57-103
 *   MAC::enforce / get_solution /
 *     select_var (dom-min, dom/deg) / select_val (min)   src/MAC.cpp:100-165, 256-271, 353-438
 *   Network::get_neighbor               src/Network.cpp:467-474
 *   AssignedStack push/pop/update     src/Solver.cpp:36-116
 *   var_que (FIFO, no duplicates)     src/Solver.cpp:180-218
 * Part 2 is independent brute force used to pin part 1 (the verify_gac /
 * verify_search patterns of samples/verify_gac.cpp:83-164 and
 * samples/verify_search.cpp:26-93): definitional GAC and SAC closures,
 * exhaustive solution enumeration, and a MAC driver whose per-node
 * propagation is the definitional closure.
 *
 * Pins (tests/test_oracle.py): the reference's known answers (test.xml,
 * queens-4, composed-25-1-2-2), brute-force agreement on every fixture and on
 * seeded random networks, and planted faults (orc_set_fault) that each pin
 * must catch.
 */
#ifndef CPIM_ORACLE_H_
#define CPIM_ORACLE_H_

#include <stdint.h>

#include "../include/cpim_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_net orc_net;

#define ORC_AC3BIT 0  /* AC3 + AC3bit::seek_support (binary networks) */
#define ORC_AC3 1     /* AC3 + generic tuple seek_support (any arity) */
#define ORC_AC3MIX 2  /* AC3 + AC3bit::seek_support on binary constraints and the generic
                         tuple search on the others.  A support is a support however it is
                         found, so the closures, and the MAC nodes and constraint weights
                         that follow them, are those of ORC_AC3; this mode exists because
                         ORC_AC3 lists every allowed tuple of every binary relation, which
                         does not fit in memory for 20,000 relations of 256 x 256 values.
                         Not a reference combination: AC3bit refuses tables
                         (src/AC3bit.cpp:4-6). */

typedef struct orc_stats {
  int64_t num_positive;
  int64_t num_negative;
  int64_t num_sol;
  int64_t revisions;   /* AC3::revise calls */
  int64_t deletions;   /* values removed by revise */
  int64_t fixpoints;   /* AC3::enforce calls */
  int32_t time_out;
  int32_t consistent;
} orc_stats;

orc_net* orc_create(const cpim_model_desc* desc, int algo);
void orc_destroy(orc_net* net);
int orc_num_vars(const orc_net* net);
int orc_dom_words(const orc_net* net);
/* Tabular::weight of every propagated constraint (wdeg bookkeeping) into out
 * (may be NULL); returns the number of constraints. */
int orc_weights(const orc_net* net, float* out);

/* Root AC (all variables queued, level 0).  Domains of level 0 are written
 * to out ([n][dom_words]) when out != NULL.  Returns 1 if consistent. */
int orc_root_ac(orc_net* net, uint64_t* out, orc_stats* st);

/* SAC1::enforce(vars, 0). out receives the level-0 domains. */
int orc_sac1(orc_net* net, uint64_t* out, orc_stats* st);

/* Singleton tests on level 0 as it stands (run orc_root_ac first): fail[i]=1
 * iff AC on a copy with tests[2i] reduced to tests[2i+1] wipes out.  Spread
 * over nthreads worker copies of the network. */
int orc_singleton_tests(orc_net* net, const int32_t* tests, int64_t count, uint8_t* fail,
                        int nthreads, orc_stats* st);

/* MAC::enforce with dom-min / min-value; time_limit_ms as the reference
 * (checked once per outer iteration), node_limit (0 = none) checked at the
 * same place.  solution [n] or NULL; trace [trace_cap] records. */
int orc_mac(orc_net* net, int64_t time_limit_ms, int64_t node_limit, orc_stats* st,
            int32_t* solution, cpim_node_record* trace, int32_t trace_cap, int32_t* trace_len);

/* orc_mac with enumeration: every solution is counted (and stored in
 * solutions[k * n .. k * n + n) while k < sol_cap), then treated as a failed
 * node so the search goes on, until the tree is exhausted or max_solutions
 * are found (max_solutions <= 1: stop at the first, as orc_mac). */
int orc_mac_all(orc_net* net, int64_t time_limit_ms, int64_t node_limit, int64_t max_solutions,
                orc_stats* st, int32_t* solutions, int64_t sol_cap, cpim_node_record* trace,
                int32_t trace_cap, int32_t* trace_len);

/* Order-free domain hash of one level (same definition as cpim_node_record). */
uint64_t orc_dom_hash(const uint64_t* dom, int n, int words);

/* ---- Part 2: brute force (independent of part 1) ---- */

/* Definitional GAC closure starting from desc (init_dom or full).  Returns 1
 * if no domain is empty; out [n][ceil(max_dom/64)]. */
int orc_bf_gac(const cpim_model_desc* desc, uint64_t* out);
/* Definitional SAC closure. */
int orc_bf_sac(const cpim_model_desc* desc, uint64_t* out);
/* Count solutions up to limit by plain backtracking over full domains;
 * first receives the lexicographically first solution. */
int64_t orc_bf_count_solutions(const cpim_model_desc* desc, int64_t limit, int32_t* first);
/* MAC with binary branching, dom-min over non-decided variables, min value,
 * and the definitional closure at every node. */
int orc_bf_mac(const cpim_model_desc* desc, int64_t node_limit, orc_stats* st,
               int32_t* solution, cpim_node_record* trace, int32_t trace_cap,
               int32_t* trace_len);
/* orc_bf_mac with the enumeration of orc_mac_all. */
int orc_bf_mac_all(const cpim_model_desc* desc, int64_t node_limit, int64_t max_solutions, orc_stats* st,
                   int32_t* solutions, int64_t sol_cap, cpim_node_record* trace, int32_t trace_cap,
                   int32_t* trace_len);

/* Planted faults for pin tests (0 = none):
 *  1: AC3bit seek_support ignores the last word of the other domain
 *  2: dom-min breaks ties towards the highest variable id
 *  3: revise skips the last value of the domain
 *  4: SAC1 does not repeat passes (one pass only)
 *  5: AC3 does not raise the weight of the constraint that wiped a domain */
void orc_set_fault(int fault);
/* Variable ordering of orc_mac and orc_bf_mac, process-wide: 0 = dom-min
 * (VRH_DOM_MIN), 1 = dom/deg (VRH_DOM_DEG_MIN, src/MAC.cpp:371-385),
 * 2 = dom/wdeg (VRH_DOM_WDEG_MIN, src/MAC.cpp:386-411, weights raised by AC3
 * on a wipe-out, src/AC3.cpp:122-129; orc_bf_mac returns -1 for it). */
void orc_set_var_heuristic(int varh);

#ifdef __cplusplus
}
#endif

#endif /* CPIM_ORACLE_H_ */
