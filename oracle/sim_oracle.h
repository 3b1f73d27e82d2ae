/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the SI-GP route simulator (engine/sim_core.hpp,
 * engine/sim_terms.hpp, engine/sim_host.hpp of the reference), kept in the
 * reference's arithmetic types and operation order so that its results are
 * bit-identical to the reference CPU build (g++ -O2, no FP contraction).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product never links or calls it.
 *
 * Parity pins (what checks this file):
 *   - tests/golden/sim.txt: 240 rewards (ns11 / ext29 / full61 masks on
 *     c101, r101, rc101, pr01) and 120 z-statistics, bitwise;
 *   - tests/golden/gp.txt.gz: 6 x 1,024 fitness values (incl. rmod), bitwise;
 *   - the reference's own simulator compiled from its sources
 *     (oracle/_ref, see oracle/Makefile) on random rules and instances;
 *   - per candidate, against the same reference build (tests/
 *     test_oracle_terminals.py): all 61 slot values of the last feasible
 *     candidate as the reference's own pick_best leaves them in its terms
 *     buffer, and every scan terminal of engine/sim_terms.hpp called
 *     directly for every feasible candidate, bitwise, over about 2,500
 *     decision states on six instances;
 *   - sim_one_rr against the reference's sim_one_rr, bitwise;
 *   - closed forms: t_rand, slots 0-9/24-27/36-40/42/58, REACH1<=REACH3<=REACH5,
 *     MAXN<=NS, brute-force greedy paths on tiny instances.
 */
#ifndef SIGP_SIM_ORACLE_H
#define SIGP_SIM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_N_TERMS 61
#define ORC_MAX_STACK 256

typedef struct {
    int n;              /* customers; POI id = index + 1, depot = 0 */
    int m;              /* vehicles */
    double tmax;        /* depot close time */
    const double *score, *dur, *open, *close;   /* n each */
    const double *dist;                         /* (n+1)^2 row-major */
    const double *d2d;                          /* n: dist to depot */
} orc_instance;

/* engine/sim_core.hpp:11-41 */
double orc_eval_postfix(const int32_t *ops, const int32_t *args, int len,
                        const double *terms, double *stack);

/* engine/sim_core.hpp:43-204.  need == NULL reproduces the reference
 * (every masked slot computed); need != NULL computes only the slots whose
 * need[slot] != 0 (used by the op-counting build to model pruning).
 * pri_out (nullable, n entries) receives each candidate's priority
 * (-INFINITY for infeasible ones). */
int orc_pick_best(const orc_instance *I, int current, double ctime, int remaining,
                  const int32_t *ops, const int32_t *args, int len,
                  const unsigned char *visited, const int64_t *mask, const double *zp,
                  const unsigned char *need, double *pri_out);

/* engine/sim_core.hpp:207-238 */
double orc_sim_one(const orc_instance *I, const int32_t *ops, const int32_t *args, int len,
                   const double *sampled, const int64_t *mask, const double *zp,
                   const unsigned char *need);

/* engine/sim_core.hpp:241-274 (no caller in the reference) */
double orc_sim_one_rr(const orc_instance *I, const int32_t *ops, const int32_t *args, int len,
                      const double *sampled, const int64_t *mask, const double *zp);

/* engine/sim_core.hpp:277-312 + engine/sim_host.hpp:96-105 */
void orc_zstats(const orc_instance *I, double zp[10]);

/* engine/sim_host.hpp:67-83 */
void orc_sim_batch(const orc_instance *I, const int32_t *ops, const int32_t *args, int len,
                   const double *all_sampled, int n_samples, const int64_t *mask,
                   const double *zp, double *out);

/* engine/factory.cuh:50-52,73-76,98-102 index maths over a packed rule set:
 * rule p reads scenario row scen_base[p] + s (row length n) and mask row
 * mask_row[p]; out[p*N + s].  Parallel over (p, s) with OpenMP when
 * n_threads > 1.  need_rows (nullable): P x 61 pruning masks. */
void orc_evaluate(const orc_instance *I, const int32_t *ops, const int32_t *args,
                  const int32_t *offsets, const int32_t *lengths, int P,
                  const double *sampled, int N, const int64_t *scen_base,
                  const int64_t *masks, const int32_t *mask_row, const double *zp,
                  const unsigned char *need_rows, int n_threads, double *out);

/* All 61 slot values pick_best computes for candidate `cand` under `mask`
 * (engine/sim_core.hpp:73-198; masked slots 0.0) into out; returns 1, or 0
 * when the candidate is visited or infeasible (out untouched). */
int orc_terms_at(const orc_instance *I, int cand, int current, double ctime, int remaining,
                 const unsigned char *visited, const int64_t *mask, const double *zp,
                 double *out);

/* The same for every candidate of one decision: out row i (n x 61) holds
 * candidate i's slot values where pri_out[i] (n entries) is finite, i.e.
 * where the candidate is unvisited and feasible; returns their number. */
int orc_terms_all(const orc_instance *I, int current, double ctime, int remaining,
                  const unsigned char *visited, const int64_t *mask, const double *zp,
                  double *out, double *pri_out);

/* Single terminal values at a given state (tests of closed forms). */
double orc_terminal(const orc_instance *I, int slot, int cand, int current, double ctime,
                    int remaining, const unsigned char *visited, const double *zp);

/* Counting build only (liboracle_count.so): FP64 operations executed since
 * the last reset, summed over the calling thread. */
unsigned long long orc_ops_read(void);
void orc_ops_reset(void);

#ifdef __cplusplus
}
#endif
#endif
