/*
 * sigp.h -- C ABI of the MI355X (gfx950) rule x scenario evaluator for SI-GP
 * dispatching rules.  Drop-in for the reference's CUDA boundary
 * (engine/factory.cuh, struct FactoryContext) and its CPU runner
 * (engine/sim_host.hpp: sim_batch, zp_of): the same arrays go in, the same
 * per-(rule, scenario) rewards come out, bit-identical.
 *
 * Plain C: pointers and sizes only.  Every call returns a sigp_status; the
 * last error text of the calling thread is available from sigp_last_error().
 * Host pointers are caller-owned and only read during the call; the library
 * owns every device buffer it allocates.  A sigp_instance may be used from
 * one thread at a time (the reference is single-threaded too).
 */
#ifndef SIGP_H
#define SIGP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIGP_N_TERMS 61     /* terminal slots (engine/vocab.hpp:16) */
#define SIGP_MAX_N 512      /* customers per instance (engine/vocab.hpp:17) */
#define SIGP_MAX_STACK 32   /* postfix stack depth (engine/vocab.hpp:19) */

typedef enum {
    SIGP_OK = 0,
    SIGP_INVALID_ARG = 1,   /* sizes out of range, malformed postfix code, bad slot */
    SIGP_DEVICE_ERROR = 2,  /* a HIP call failed */
    SIGP_OOM = 3,           /* device allocation failed */
    SIGP_NO_DEVICE = 4      /* no gfx950 device visible */
} sigp_status;

/* One STOPTW instance as the reference's SimData holds it
 * (engine/sim_host.hpp:12-40): customer i = POI id i+1, depot = POI 0.
 * dist is the (n+1)x(n+1) row-major travel-time matrix, dist_to_depot[i] =
 * dist[(i+1)*(n+1)].  zp are the 10 z-statistics (5 means, 5 stds) the Z
 * terminals read (engine/sim_host.hpp:96-105); pass NULL to have the library
 * compute them (sigp_zstats). */
typedef struct {
    int32_t n;
    int32_t m;
    double tmax;
    const double *scores, *durations, *open_t, *close_t;
    const double *dist;
    const double *dist_to_depot;
    const double *zp;
} sigp_instance_desc;

/* Packed postfix rules, the reference's PackedTrees (engine/factory.cuh:20-40):
 * rule p is ops[offsets[p] .. offsets[p]+lengths[p]) with two args per node;
 * ops: -1 terminal (args[2k] = slot), -2 constant (args[2k] = round(v*1000)),
 * 0..5 = + - * / min max, 6 = rmod (engine/vocab.hpp:88-120). */
typedef struct {
    const int32_t *ops;
    const int32_t *args;
    const int32_t *offsets;
    const int32_t *lengths;
    int32_t n_rules;
    int64_t n_nodes;
} sigp_rules;

typedef struct sigp_instance sigp_instance;
typedef struct sigp_batch sigp_batch;

/* Text of the last error on the calling thread ("" if none). */
const char *sigp_last_error(void);

/* Device information: name into name_buf, compute-unit count and clock (GHz). */
sigp_status sigp_device_info(char *name_buf, size_t buf_len, int32_t *n_cu, double *clock_ghz);

/* Replaces FactoryContext::FactoryContext (engine/factory.cuh:129-137):
 * uploads the instance once and builds its derived tables (transposed
 * distances, score/distance ratios, the ATRISK pow(var,0.5) table). */
sigp_status sigp_instance_create(const sigp_instance_desc *desc, sigp_instance **out);

/* Replaces FactoryContext::~FactoryContext (engine/factory.cuh:138-141). */
sigp_status sigp_instance_destroy(sigp_instance *inst);

/* Replaces zp_of / zstats (engine/sim_host.hpp:96-105, engine/sim_core.hpp:277-312),
 * computed on the device; zp_out receives 10 doubles. */
sigp_status sigp_zstats(const sigp_instance_desc *desc, double *zp_out);

/* Replaces FactoryContext::evaluate / evaluate_grouped / evaluate_grouped_masks
 * (engine/factory.cuh:144-164, 167-189, 193-216) in one entry point:
 *   rule p reads scenario rows (p / group_size) * T + s, s in [0, T), of
 *   `sampled` (row length n; n_rows rows in total) and terminal-mask row
 *   p / mask_group of `masks` (n_masks rows of SIGP_N_TERMS int64 0/1).
 *   evaluate(pk, sampled, N, mask)            = group_size >= P, T = N, mask_group >= P
 *   evaluate_grouped(pk, s, G, T, gs, mask)   = group_size = gs, mask_group >= P
 *   evaluate_grouped_masks(pk, s, T, gs, mg, masks) as is.
 * out_pairs (nullable) receives P*T rewards, out[p*T + s]; out_mean
 * (nullable) receives the P per-rule means, summed over s in order and
 * divided by T as the reference's callers do (engine/gp_engine.cu:76-81).
 * All pointers are host pointers; the call is synchronous. */
sigp_status sigp_evaluate(sigp_instance *inst, const sigp_rules *rules,
                          const double *sampled, int64_t n_rows, int32_t T,
                          int32_t group_size, const int64_t *masks, int32_t n_masks,
                          int32_t mask_group, double *out_pairs, double *out_mean);

/* Device-resident form of sigp_evaluate for callers that keep scenarios and
 * results in HBM (no counterpart in the reference, which copies per call;
 * engine/factory.cuh:147-162).  sigp_batch_create validates and uploads the
 * rules and masks once; sigp_batch_run launches on `stream` (a hipStream_t,
 * NULL = default stream) and returns without waiting.  sampled_dev, out_pairs_dev
 * and out_mean_dev are device pointers (out_* nullable). */
sigp_status sigp_batch_create(sigp_instance *inst, const sigp_rules *rules, int32_t T,
                              int32_t group_size, const int64_t *masks, int32_t n_masks,
                              int32_t mask_group, sigp_batch **out);
sigp_status sigp_batch_run(sigp_batch *batch, const double *sampled_dev, int64_t n_rows,
                           double *out_pairs_dev, double *out_mean_dev, void *stream);
sigp_status sigp_batch_destroy(sigp_batch *batch);

/* Many instances, the same rules (no counterpart in the reference, which
 * evaluates one instance per FactoryContext and launch; its GA and
 * redeploy tools loop over instances, s4_selection/redeploy.cu:40).  Every
 * rule is simulated on every instance over that instance's own T scenario
 * rows.  Instances of different sizes run in one launch per size class
 * (n <= 128, <= 288, <= 512), each class with its own longest-first order
 * and, after sigp_batch_autotune, its own mapping.  The instances must
 * outlive the batch.  With sigp_batch_run, sampled_dev holds the instances'
 * row blocks back to back (instance k: T rows of its n_k doubles), n_rows is
 * their total row count (n_inst * T), out_pairs receives n_inst * P * T
 * rewards, out[(k * P + p) * T + s], and out_mean n_inst * P means, k * P + p.
 * Rule p reads mask row p / mask_group. */
sigp_status sigp_batch_create_multi(sigp_instance *const *insts, int32_t n_inst,
                                    const sigp_rules *rules, int32_t T, const int64_t *masks,
                                    int32_t n_masks, int32_t mask_group, sigp_batch **out);

/* sigp_batch_create_multi + sigp_batch_run on host arrays, synchronous:
 * `sampled` holds the instances' row blocks back to back as above. */
sigp_status sigp_evaluate_multi(sigp_instance *const *insts, int32_t n_inst, const sigp_rules *rules,
                                const double *sampled, int32_t T, const int64_t *masks,
                                int32_t n_masks, int32_t mask_group, double *out_pairs,
                                double *out_mean);

/* sigp_evaluate without device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under HIP graph capture, tight serving loops and harnesses that forbid
 * hipMalloc.  Same arguments and results as sigp_evaluate, but the scenario
 * rows (sampled_dev) and the results (out_pairs_dev, required; out_mean_dev,
 * nullable) are device pointers, and the work goes on `stream` (a
 * hipStream_t, NULL = default stream); the call returns when the results
 * are written.  The block sigp_batch_create would allocate (compiled rules,
 * read sets, the launch order: about 28 bytes per rule plus 12 per postfix
 * node) lives in a fixed device array of SIGP_NOALLOC_MAX_BYTES per GPU,
 * held until the process ends; rule sets that need more are rejected with
 * SIGP_INVALID_ARG.  No scenario order, repeated-rule skipping or tuning
 * (results are identical without them).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other threads, other streams) race
 * and must be serialised by the caller.  Host-side rule validation uses
 * ordinary heap memory. */
#define SIGP_NOALLOC_MAX_BYTES (8 << 20)
sigp_status sigp_evaluate_noalloc(sigp_instance *inst, const sigp_rules *rules,
                                  const double *sampled_dev, int64_t n_rows, int32_t T,
                                  int32_t group_size, const int64_t *masks, int32_t n_masks,
                                  int32_t mask_group, double *out_pairs_dev, double *out_mean_dev,
                                  void *stream);

/* Launch mapping of the simulation kernel (no counterpart in the reference,
 * whose kernels always run one thread per (rule, scenario) pair).  Every
 * mapping gives bit-identical results.
 *   SIGP_MAP_SCENARIOS:  a workgroup simulates one rule on a block of 64
 *     scenarios, one scenario per lane; its waves_per_item wavefronts split
 *     each decision's candidates between them.
 *   SIGP_MAP_CANDIDATES: a workgroup simulates one (rule, scenario) pair;
 *     its 64 * waves_per_item threads split each decision's candidates and
 *     take the first maximum together.  For few scenarios per rule, such as
 *     the reference's GP loop, which evaluates one per generation
 *     (engine/gp.hpp:32).
 *   SIGP_MAP_CANDIDATES_2: candidate lanes in two segments -- one 64-lane
 *     wavefront simulates 2 scenarios of a rule, 32 lanes splitting each
 *     one's candidates; waves_per_item is ignored.  For n <= 288 customers
 *     (larger instances run as SIGP_MAP_CANDIDATES); pays where most
 *     decisions have few feasible candidates and rules are long (evolved GP
 *     populations at n = 100).
 *   SIGP_MAP_AUTO: candidate lanes when T < 64 or when scenario lanes would
 *     make fewer than 512 workgroups, scenario lanes otherwise.
 * Each wavefront keeps the candidates' terminal values and the postfix stack
 * in LDS, (distinct slots read + deepest stack - 1) * 512 bytes, and a
 * workgroup stays within 64 KiB of LDS: a batch whose rules need more runs
 * with fewer wavefronts than requested (autotune skips such mappings). */
#define SIGP_MAP_AUTO 0
#define SIGP_MAP_SCENARIOS 1
#define SIGP_MAP_CANDIDATES 2
#define SIGP_MAP_CANDIDATES_2 3
typedef struct {
    int32_t waves_per_item;   /* 1, 2 or 4; 0 = default of the mapping */
    int32_t mapping;          /* SIGP_MAP_* */
} sigp_launch_config;

/* Scenario order of scenario lanes (no counterpart in the reference; results
 * are identical either way).  A 64-lane block of scenario lanes simulates 64
 * scenarios of one rule in lockstep; the first decision is the same in every
 * scenario, and the first service time sets the clock after it, so blocks
 * whose scenarios are adjacent by the duration of the rule's first customer
 * keep their routes together longer (so do the two simulations of a
 * SIGP_MAP_CANDIDATES_2 wavefront).  Enabled by default where the batch
 * qualifies (64 <= T <= 4096, orders within 256 MiB): each pair's first
 * customer is found at creation (one kernel on the default stream; the
 * creating call then waits for the device, as its uploads already do) and
 * sigp_batch_run sorts the rows on its stream before the simulation.
 * enable = 0 turns it off; *active (nullable) reports whether it is on. */
sigp_status sigp_batch_order_scenarios(sigp_batch *batch, int32_t enable, int32_t *active);

/* Mapping of one batch. */
sigp_status sigp_batch_configure(sigp_batch *batch, const sigp_launch_config *cfg);

/* Times every candidate configuration (every mapping, 1, 2 and 4 wavefronts
 * where they fit), twice each, on `stream` (device pointers as in
 * sigp_batch_run) and keeps the fastest by the second run, per size class.
 * Each class is timed on an evenly strided sample of its (instance, rule)
 * pairs, about 2^19 simulations, so tuning costs a fraction of a run of a
 * large batch; when the best scenario-lane and candidate-lane
 * configurations come within 25 % of each other, the two are timed once
 * more on a 4x larger sample, which decides.  cfg_out (nullable) receives
 * the configuration of the class with the most work.  Synchronous; the
 * results it writes are scratch. */
sigp_status sigp_batch_autotune(sigp_batch *batch, const double *sampled_dev, int64_t n_rows,
                                void *stream, sigp_launch_config *cfg_out);

/* Repeated rules (no counterpart in the reference): when enabled, batches
 * created later on this instance, including the ones sigp_evaluate creates,
 * simulate each distinct rule of a scenario group once.  A rule is a repeat
 * when an earlier rule of its group has the same postfix code and mask
 * effect; its rewards are copied from that rule, bit for bit the values it
 * would have computed.  Populations bred by GP repeat many trees.  Off by
 * default: detection hashes every rule on the host at batch creation, which
 * costs more than it saves for callers that create a batch per call around
 * a host-side GP step (the reference's GA measured 3-10 % slower with it).
 * A multi-instance batch skips repeats only when every one of its instances
 * has this enabled (its launch-mapping defaults come from its first
 * instance). */
sigp_status sigp_instance_skip_repeats(sigp_instance *inst, int32_t enable);

/* Default mapping for batches created later on this instance, including the
 * ones sigp_evaluate creates internally. */
sigp_status sigp_instance_configure(sigp_instance *inst, const sigp_launch_config *cfg);

/* One launch for a multi-instance batch (no counterpart in the reference,
 * which launches per instance: engine/factory.cuh:42-112 behind each
 * FactoryContext).  When every instance of a batch created later by
 * sigp_batch_create_multi / sigp_evaluate_multi has this enabled, all
 * (instance, rule) pairs run in one launch of the widest size class's
 * kernels, in one longest-first order, instead of one launch per size class
 * on concurrent streams.  Results are identical either way; off by default
 * (the per-class launches measured as fast, and the narrower classes' kernels
 * hold fewer registers). */
sigp_status sigp_instance_one_launch(sigp_instance *inst, int32_t enable);

/* Simulation protocol.  SIGP_PROTOCOL_VEHICLES is sim_one
 * (engine/sim_core.hpp:207-238, the reference's V0 and its only GPU path):
 * vehicle 0 routes until pick_best finds no candidate, then vehicle 1, ...
 * SIGP_PROTOCOL_ROUND_ROBIN is sim_one_rr (engine/sim_core.hpp:241-274, the
 * reference's V2 ablation arm, CPU only there): the open vehicles take one
 * decision each in turn, 0 .. m-1, a vehicle closing when it finds no
 * candidate; pick_best still sees "vehicles left" = m - v - 1 for vehicle v.
 * The round-robin arm needs m <= SIGP_RR_MAX_M on every instance of the
 * batch (each lane keeps every vehicle's route state in LDS).
 * sigp_instance_protocol sets the protocol of batches created later on the
 * instance, including the ones sigp_evaluate and sigp_evaluate_noalloc
 * create; a multi-instance batch needs its instances to agree.
 * sigp_batch_protocol switches an existing batch.  The device GP loop
 * (sigp_gp_*) evaluates with SIGP_PROTOCOL_VEHICLES, as the reference's GP
 * engine does. */
#define SIGP_PROTOCOL_VEHICLES 0
#define SIGP_PROTOCOL_ROUND_ROBIN 1
#define SIGP_RR_MAX_M 8
sigp_status sigp_instance_protocol(sigp_instance *inst, int32_t protocol);
sigp_status sigp_batch_protocol(sigp_batch *batch, int32_t protocol);

/* Number of rules and (rule, scenario) pairs in a batch. */
sigp_status sigp_batch_shape(const sigp_batch *batch, int32_t *n_rules, int64_t *n_pairs);

/* Device self-test hook for the transcendentals the simulator uses, each
 * bitwise equal to the host C library's (glibc 2.35, x86-64 with FMA):
 * out[k] = f(x[k]) computed on the device.  which: 0 = erf, 1 = erfc,
 * 2 = exp, 3 = sqrt, 4 = pow(x[k], y[k]) (the rmod operator's pow,
 * engine/sim_core.hpp:35). */
sigp_status sigp_device_math(int32_t which, const double *x, const double *y, int64_t count,
                             double *out);

/* ---- phenotype (no counterpart in the reference) --------------------------
 * Decision situations: per situation the state of a route under
 * construction (current POI, clock, vehicles left after the current one,
 * visited customers) and n_cand customers to rank. */
typedef struct {
    int32_t n_sit, n_cand;        /* 1 <= n_sit * n_cand <= 1024 */
    const int32_t *cur;           /* n_sit: current POI, 0 = depot */
    const double *ct;             /* n_sit: current time */
    const int32_t *remaining;     /* n_sit: vehicles after the current one */
    const uint8_t *visited;       /* n_sit x n: 1 = customer visited */
    const int32_t *cand;          /* n_sit x n_cand: customers to rank (feasible ones) */
} sigp_situations;

/* A rule's phenotype: for rule p, situation s and candidate slot j,
 * out_ranks[(p * n_sit + s) * n_cand + j] = the candidate's rank by the
 * rule's priority (0 = the one pick_best would choose), computed on the
 * device by the simulator's own per-candidate evaluation
 * (engine/sim_core.hpp:73-203).  Priorities at or below -1e300 (and NaN)
 * count as -1e300; equal priorities rank by candidate slot.  A candidate
 * that is visited or infeasible in its situation counts as -1e300.
 * Masks as in sigp_evaluate.  Synchronous. */
sigp_status sigp_phenotype(sigp_instance *inst, const sigp_rules *rules, const int64_t *masks,
                           int32_t n_masks, int32_t mask_group, const sigp_situations *sit,
                           int32_t *out_ranks);

/* ---- surrogate network (no counterpart in the reference) -----------------
 * A 512 -> 256 -> 256 -> 1 ReLU MLP that predicts a rule's mean reward from
 * its phenotype (sigp_phenotype), trained and run on the device: FP32
 * parameters and Adam state, BF16 copies of the weight matrices and BF16
 * activations on the matrix cores (v_mfma_f32_16x16x32_bf16), FP32
 * accumulation, FP32 pre-activations, outputs and gradients.  Features are
 * row-major float rows of SIGP_MLP_IN values; parameters are laid out as
 * W1[512][256], b1[256], W2[256][256], b2[256], w3[256], b3 (SIGP_MLP_NPARAMS
 * floats).  Loss: mean squared error over the batch. */
#define SIGP_MLP_IN 512
#define SIGP_MLP_HIDDEN 256
#define SIGP_MLP_NPARAMS 197377
typedef struct sigp_mlp sigp_mlp;

/* Glorot-uniform weights and zero biases from a splitmix64 stream of seed. */
sigp_status sigp_mlp_create(uint64_t seed, sigp_mlp **out);
sigp_status sigp_mlp_destroy(sigp_mlp *net);
sigp_status sigp_mlp_get_params(sigp_mlp *net, float *out);
/* Replaces the parameters and resets the Adam state. */
sigp_status sigp_mlp_set_params(sigp_mlp *net, const float *in);
/* Loss gradient of one batch (1..1024 rows) with respect to every parameter,
 * without updating; loss_out (nullable) receives the batch loss. */
sigp_status sigp_mlp_gradients(sigp_mlp *net, const float *x, const float *t, int32_t rows,
                               float *grad_out, float *loss_out);
/* `epochs` passes of Adam (lr; beta 0.9 / 0.999; eps 1e-8) over `rows`
 * examples in batches of `batch` (1..1024), shuffled per epoch from
 * shuffle_seed; epoch_loss (nullable) receives each epoch's mean loss. */
sigp_status sigp_mlp_train(sigp_mlp *net, const float *x, const float *t, int64_t rows,
                           int32_t epochs, int32_t batch, float lr, uint64_t shuffle_seed,
                           float *epoch_loss);
sigp_status sigp_mlp_predict(sigp_mlp *net, const float *x, int64_t rows, float *y_out);

/* ------------------------------------------------------------------------
 * Scenario rows on the device (engine/sim_host.hpp:43-58 gen_durations,
 * whose rows the reference samples on the host).  Row r draws from
 * Mt19937(base_seed + r): customer i gets max(0, gauss(mu_i, sigma mu_i))
 * (engine/rng.hpp:135-147) or, in the Gamma mode (no counterpart in the
 * reference), Gamma(shape, mu_i / shape) by Marsaglia & Tsang (shape >= 1);
 * customers with mu_i <= 0 draw nothing.  The uniform stream is the
 * reference's bit for bit; log/sin/cos are correctly rounded, where glibc's
 * (the reference's) are not in about 0.1 % of calls, so a duration can
 * differ from the reference's host sampler in the last place. */
#define SIGP_SAMPLE_GAUSS 0
#define SIGP_SAMPLE_GAMMA 1
typedef struct {
    int32_t mode;     /* SIGP_SAMPLE_GAUSS or SIGP_SAMPLE_GAMMA */
    double sigma;     /* Gauss: sd = sigma * expected (the reference's 0.2) */
    double shape;     /* Gamma: shape, scale = expected / shape */
} sigp_sampler;

/* n_rows x n durations into host memory. */
sigp_status sigp_sample_durations(sigp_instance *inst, const sigp_sampler *smp, int64_t base_seed,
                                  int64_t n_rows, double *out);
/* The same into device memory on `stream` (hipStream_t), without waiting. */
sigp_status sigp_sample_durations_async(sigp_instance *inst, const sigp_sampler *smp, int64_t base_seed,
                                        int64_t n_rows, double *out_dev, void *stream);

/* ------------------------------------------------------------------------
 * The GP loop on the device.  Replaces the host loop of the reference's
 * generation engine -- evolve_one (engine/gp_engine.cu:66-84) and the
 * seed-batched exhaust_batched (engine/gp_engine.cu:131-183) -- whose step
 * is gp_step (engine/gp.hpp:170-194) over ramped_half_and_half
 * (engine/gp.hpp:67-77).  Populations stay on the device as prefix-order
 * token arrays; every random choice comes from one CPython-compatible
 * Mersenne Twister per run seeded like Mt19937(seed) (engine/rng.hpp), drawn
 * in the reference's order, so the populations are the reference's.
 *
 * Trees are at most SIGP_GP_MAX_NODES tokens (depth <= 8): a terminal is its
 * index in terminal_slots, a function SIGP_GP_TOK_FUNC | its index in
 * function_ops, an ERC constant SIGP_GP_TOK_CONST | k for the value k/1000
 * (the reference's "%.3f" label as compile reads it back). */
#define SIGP_GP_MAX_NODES 512
#define SIGP_GP_TOK_FUNC 0x4000
#define SIGP_GP_TOK_CONST 0x2000

typedef struct sigp_gp sigp_gp;

/* TreeConfig + EvolutionConfig (engine/gp.hpp:13-34). */
typedef struct {
    int32_t n_runs;                  /* independent runs evolved side by side (the reference's seeds) */
    int32_t population;              /* EvolutionConfig::population_size, <= 4096 */
    int32_t offspring;               /* offspring per run and step: 0 = population; a multiple k >= 2 of it
                                        screens them with sigp_gp_set_surrogate's network */
    int32_t n_terminals;             /* TreeConfig::terminals, <= 128 */
    const int32_t *terminal_slots;   /* slot of each terminal (0..60), -1 = "ERC" */
    int32_t n_functions;             /* TreeConfig::functions, <= 8 */
    const int32_t *function_ops;     /* opcode of each function (0..6, engine/vocab.hpp:79-88) */
    int32_t min_depth, max_depth;    /* TreeConfig: initialisation and mutation depths (mutants deeper
                                        than max_depth are rejected) */
    int32_t depth_limit;             /* EvolutionConfig::max_depth: crossover children deeper are rejected (<= 8) */
    int32_t tournament_size;         /* <= 64 */
    int32_t elitism;
    double crossover_rate, mutation_rate;
    int32_t T;                       /* training scenarios per run and generation */
    sigp_launch_config launch;       /* evaluator lane mapping; SIGP_MAP_AUTO: two simulations per
                                        wavefront (SIGP_MAP_CANDIDATES_2) for n <= 128 and T >= 2,
                                        else sigp_batch's choice */
} sigp_gp_config;

/* seeds: n_runs Mt19937 seeds; term_mask: int64[SIGP_N_TERMS] of the
 * evaluation (engine/gp_engine.cu:93).  A screened run (offspring = k x
 * population) breeds k gp_steps per generation, step s on the stream
 * Mt19937(seed + s * 2^32) (s = 0: the reference's stream). */
sigp_status sigp_gp_create(sigp_instance *inst, const sigp_gp_config *cfg, const uint64_t *seeds,
                           const int64_t *term_mask, sigp_gp **out);
sigp_status sigp_gp_destroy(sigp_gp *gp);
/* ramped_half_and_half of every run (engine/gp.hpp:67-77). */
sigp_status sigp_gp_init(sigp_gp *gp);
/* Fitness of the population (engine/gp_engine.cu:70-79): run r's rules read
 * host scenario rows r*T .. r*T+T-1 (n_rows = n_runs*T rows of n durations);
 * fitness (nullable): n_runs*population means.  Trees of a run with the same
 * tokens are simulated once (their results are the same bit for bit), here
 * and in sigp_gp_run. */
sigp_status sigp_gp_evaluate(sigp_gp *gp, const double *sampled, int64_t n_rows, double *fitness);
/* One gp_step of every run (engine/gp.hpp:170-194) with the given host
 * fitness (n_runs*population), or NULL for the last sigp_gp_evaluate. */
sigp_status sigp_gp_step(sigp_gp *gp, const double *fitness);
/* Surrogate pre-screening (no counterpart in the reference) for a run
 * created with offspring = k x population: each generation's k x population
 * offspring are compiled on the device, their phenotypes taken on `sit`
 * (n_sit x n_cand = SIGP_MLP_IN, as sigp_phenotype) and scored by `net`
 * (not owned; it must outlive the run); the next parents are the elites
 * and then population - elitism offspring outside the elites' slots:
 * phenotypes not seen at an earlier slot of the run first (phenotype =
 * FNV-1a 64-bit hash of the ranks, one byte each), then higher score, then
 * lower slot.  Until a network is attached, and after net = NULL (sit may
 * then be NULL) detaches it, such a run takes plain steps: the next parents
 * are sub-step 0's offspring, the gp_step of the run's own stream. */
sigp_status sigp_gp_set_surrogate(sigp_gp *gp, sigp_mlp *net, const sigp_situations *sit);
/* Screened initial population (no counterpart in the reference, whose
 * initial population is one ramped_half_and_half, engine/gp.hpp:67-77): for
 * a run with offspring = k x population and a network attached, each of its
 * k streams draws a ramped_half_and_half of population trees (stream 0's are
 * what sigp_gp_init would give), and the screen of a step with no elites
 * keeps population of the k x population: distinct phenotypes first, then
 * higher score, then lower slot.  Replaces sigp_gp_init for such runs. */
sigp_status sigp_gp_init_screened(sigp_gp *gp);
/* Phenotypes of the current population (no counterpart in the reference):
 * ranks (n_runs*population x n_sit x n_cand int32, as sigp_phenotype) of
 * every tree on `sit`, computed from the device's own compiled trees.  With
 * sigp_gp_init and sigp_gp_evaluate it builds a surrogate's labelled set --
 * random rules, their exact mean rewards and their phenotypes -- without the
 * rules leaving the device. */
sigp_status sigp_gp_phenotype(sigp_gp *gp, const sigp_situations *sit, int32_t *ranks);
/* The whole loop on the device: for g < generations, evaluate on rows
 * g*n_runs*T .. (n_rows = generations*n_runs*T) and step, except after the
 * last unless flags has SIGP_GP_STEP_LAST (a loop run in several calls).
 * SIGP_GP_GRAPH captures one generation as a HIP graph and replays it.
 * Outputs (each nullable), numbered from this call's first generation:
 * fitness_history, generations x n_runs x
 * population; best_tokens / best_lengths, each generation's and run's best
 * rule (first maximum of the fitness), generations x n_runs x
 * SIGP_GP_MAX_NODES tokens and lengths; generation_ms, the device time from
 * the loop's start to the end of each generation. */
#define SIGP_GP_GRAPH 1
#define SIGP_GP_STEP_LAST 2
sigp_status sigp_gp_run(sigp_gp *gp, const double *sampled, int64_t n_rows, int32_t generations,
                        int32_t flags, double *fitness_history, int16_t *best_tokens,
                        int32_t *best_lengths, double *generation_ms);
/* Scenario rows sampled on the device inside the loop: with this set,
 * sigp_gp_run takes sampled = NULL (n_rows = 0) and generation g's row t of
 * run r draws from Mt19937(run_seed[r] + (gen_offset + g) * gen_stride + t)
 * -- the reference's schedule is run_seed[r] = seed_r * 10000, gen_stride = 1
 * (engine/gp_engine.cu:69); gen_offset counts generations of earlier calls. */
sigp_status sigp_gp_set_sampler(sigp_gp *gp, const sigp_sampler *smp, const int64_t *run_seed,
                                int64_t gen_stride, int64_t gen_offset);
/* The population: tokens (n_runs*population*SIGP_GP_MAX_NODES) and lengths. */
sigp_status sigp_gp_population(const sigp_gp *gp, int16_t *tokens, int32_t *lengths);
/* The population compiled as pack_trees would (engine/factory.cuh:20-40),
 * tree i at i*SIGP_GP_MAX_NODES: ops (n_runs*population*SIGP_GP_MAX_NODES),
 * args (2x that), lengths. */
sigp_status sigp_gp_packed(const sigp_gp *gp, int32_t *ops, int32_t *args, int32_t *lengths);
/* Generator state of one run's sub-step stream (0 = the reference's): 624
 * words and the position, as CPython's random.getstate()[1]. */
sigp_status sigp_gp_rng_state(const sigp_gp *gp, int32_t run, int32_t sub, uint32_t *state);

#ifdef __cplusplus
}
#endif
#endif
