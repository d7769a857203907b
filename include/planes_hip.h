/* planes_hip.h -- C ABI of the MI355X fitness-evaluation library (libplanes_hip.so).
 *
 * The reference keeps its GPU path behind seven C++ entry points (src/sim.h:6-13).
 * This header is the plain-C layer underneath their drop-in replacements
 * (include/planes_sim.h): plain pointers, sizes and int status codes, no C++ types,
 * so it binds from ctypes/cffi/cgo/JNI alike.  Every entry point returns 0 on success
 * and a non-zero code (a hipError_t value, or -1 for an argument error) otherwise;
 * pr_last_error() gives the message of the calling thread's last failure.
 *
 * Pointers named d_* are device (HBM) pointers; everything else is host memory.
 * One context = one HIP stream; calls on one context must come from one thread at a
 * time (the reference's single-thread, one-stream model, sim.cu:214-248).
 */
#ifndef PLANES_HIP_H
#define PLANES_HIP_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pr_ctx pr_ctx;

/* Scenario preset and flight settings.  The reference hard-codes them as #defines
 * (common.h:389-421, 470-478; sim.cu:241): preset 0 with max_steps 7200, chunk 600. */
typedef struct pr_sim_params {
    int preset;          /* 0: reference hash preset (common.h:389-421); 1: parametric config preset */
    float gust_rms_max;  /* preset 1: gust field 3-D RMS ~ U(0, gust_rms_max) [m/s]   */
    float warp_max;      /* preset 1: print warp pitch/roll ~ U(-warp_max, warp_max) [rad] */
    float wind_sd;       /* preset 1: steady wind ~ N(0, wind_sd) per axis [m/s]       */
    int max_steps;       /* flight cap in 600 Hz steps (SIM_MAX_STEPS, common.h:471)    */
    int chunk;           /* steps per launch; the clock restarts at step0*dt (sim.cu:113,241) */
    float spread;        /* showcase launch grid spacing [m], 0 for evaluation (sim.cu:71-78) */
} pr_sim_params;

#define PR_MAX_SLOTS 16
#define PR_MAX_CHUNKS 64
typedef struct pr_stats {
    unsigned long long flight_steps; /* integrated 600 Hz steps of the last pr_fly (= plane-flight-s * 600) */
    unsigned long long panel_steps;  /* sum over those steps of the plane's triangle count */
    double fly_ms;                   /* GPU time of all flight-kernel launches (HIP events on the context stream) */
    int fly_launches;
    double prep_ms;                  /* GPU time of the last pr_prepare */
    int n_slots;                     /* slot 2b: bucket b's many-flights-per-wave kernel, 2b+1: its tail kernel */
    int s_nt_max[PR_MAX_SLOTS], s_L[PR_MAX_SLOTS], s_P[PR_MAX_SLOTS], s_launches[PR_MAX_SLOTS];
    unsigned long long s_flight_steps[PR_MAX_SLOTS], s_panel_steps[PR_MAX_SLOTS];
    double s_fly_ms[PR_MAX_SLOTS];   /* own-launch time of a slot (0 for slots the fused kernel flies) */
    int s_fused[PR_MAX_SLOTS];       /* 1: the slot is flown inside the fused kernel */
    double fused_ms;                 /* GPU time of the fused flight-kernel launches */
    int fused_launches;
    /* per 1-s chunk (chunks past PR_MAX_CHUNKS - 1 accumulate in the last entry): flights
     * entering it, their integrated steps and panel-steps, flight-kernel time */
    int n_chunks;
    int c_flights[PR_MAX_CHUNKS];
    unsigned long long c_flight_steps[PR_MAX_CHUNKS], c_panel_steps[PR_MAX_CHUNKS];
    double c_ms[PR_MAX_CHUNKS];
    /* of the totals, the steps flown by the small-chunk kernel (pr_set_split) */
    unsigned long long split_flight_steps, split_panel_steps;
} pr_stats;

/* Replaces panelbuf_alloc/panelbuf_free (sim.h:9-10, sim.cu:220-235).
 * Capacities: genomes per prepare, packed panels (sum of triangle counts), scenarios,
 * flights per pr_fly, and the largest mesh resolution (ch, sp) any genome will use
 * (0 = the reference's CH=5, SP=8). */
int pr_create(pr_ctx** out, int max_genomes, long long max_panels, int max_scenarios, int max_flights,
              int max_ch, int max_sp);
void pr_destroy(pr_ctx* ctx);
const char* pr_last_error(void);

/* Replaces gpu_prepare (sim.h:11, sim.cu:214-218; k_prepare sim.cu:25-49): meshes n
 * genomes (d_genomes [n*35] floats, the Genome byte layout), computes mass properties,
 * the print-volume check and packed panel records.  d_res: optional [n*2] (ch, sp)
 * per genome (null = reference resolution); a resolution outside 1..max_ch x 1..max_sp of
 * pr_create is not meshed (no triangles: the genome flies as unprintable, -100).
 * Asynchronous on the context stream. */
int pr_prepare(pr_ctx* ctx, const float* d_genomes, const int* d_res, int n);

/* Arithmetic of pr_prepare: 32 (default) reproduces the reference's float meshing and
 * mass_props bit for bit; 64 meshes and integrates in double and rounds the results
 * once to float (mass/CG/inertia then agree with an FP64 evaluation to float rounding). */
int pr_set_prepare_precision(pr_ctx* ctx, int bits);

/* Scenario descriptors for n seeds (d_seeds device int32) -- scenario_desc, common.h:389. */
int pr_scenarios(pr_ctx* ctx, const int* d_seeds, int n, const pr_sim_params* p);

/* Flies every flight and writes d_fitness / d_steps (integrated steps) per flight.
 * layout 0: flight f = genome f / n_scen under scenario f % n_scen (n_genomes*n_scen
 *           flights; the GA's seed batch and the holdout batch, sim.cu:434-456);
 * layout 1: flight f = genome f under scenario f (n_genomes == n_scen flights; gpu_sim
 *           with vary_scenarios=1, sim.cu:64).
 * Uses the genomes of the last pr_prepare and the scenarios of the last pr_scenarios.
 * Returns when d_fitness is complete (like gpu_sim, sim.cu:238-248). */
int pr_fly(pr_ctx* ctx, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
           int* d_steps);
/* pr_fly in two halves: pr_fly_async queues the flight batch on the context's stream and
 * returns; pr_fly_wait waits for it and gathers pr_stats. Results (d_fitness, d_steps,
 * readbacks, stats) are valid after pr_fly_wait; one batch per context at a time. */
int pr_fly_async(pr_ctx* ctx, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
                 int* d_steps);
int pr_fly_wait(pr_ctx* ctx);

/* Only the flight set-up of pr_fly: per-flight constants and launch states (sim.cu:63-109),
 * fitness -100 for unprintable genomes.  pr_read_states then returns the launch states. */
int pr_init_flights(pr_ctx* ctx, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
                    int* d_steps);

/* Replaces gpu_sim (sim.h:12, sim.cu:238-248) one for one: flies the prepared genomes
 * from seed (vary = 0: every plane the same scenario; 1: plane p flies seed + p * 7919,
 * 32-bit wrap), launch-grid offset spread, pose recording into d_rec [pop*rec_cap*8]
 * (null: none) every rec_every >= 1 steps; d_fitness device [pop]; scenario preset and
 * flight cap from p. Synchronous, like the reference. */
int pr_sim(pr_ctx* ctx, int pop, int seed, float* d_fitness, float* d_rec, int rec_every, int rec_cap, float spread,
           int vary, const pr_sim_params* p);

/* pr_fly with pose recording (gpu_sim's d_rec path, sim.cu:117-128): d_rec device
 * [n_flights * rec_cap * 8] floats, frame k of flight f at (f*rec_cap + k)*8 =
 * {pos.xyz, q.wxyz, alive}; frames every rec_every steps; a flight that has ended keeps
 * its final pose with alive = 0. */
int pr_fly_record(pr_ctx* ctx, int layout, int n_genomes, int n_scen, const pr_sim_params* p, float* d_fitness,
                  int* d_steps, float* d_rec, int rec_every, int rec_cap);

/* Tail-aware robust score per genome (robust_scores, sim.cu:340-357):
 * 0.65 mean + 0.25 CVaR(worst quarter) + 0.10 worst, from d_samples [pop*nseeds]. */
int pr_scores(pr_ctx* ctx, const float* d_samples, int pop, int nseeds, float* d_scores);

int pr_stats_get(pr_ctx* ctx, pr_stats* out);

/* 1: pr_fly captures its whole launch sequence (flight set-up + the chain of 1-s chunk
 * launches) as a hipGraph on first use and
 * replays it while the launch parameters stay the same; per-launch timings are then
 * replaced by one whole-graph time.  0 (default): direct launches. */
int pr_set_graph(pr_ctx* ctx, int on);

/* Flight-kernel mapping per triangle-count bucket: flights whose count is <= nt_max[i]
 * (and above the previous bucket's) are flown by groups of L[i] lanes holding P[i]
 * panels per lane (L*P >= nt_max).  Available (L, P) pairs: pr_fly_variants. */
int pr_set_buckets(pr_ctx* ctx, int n, const int* nt_max, const int* L, const int* P);
/* 1: each bucket also queues a one-flight-per-wave mapping (64 lanes, P = nt/64)
 * that takes over once the bucket's remaining flights fit one per wave -- the long-flight
 * tail is then latency- rather than throughput-bound.  2 (default): the same, except that
 * in the fused kernel a bucket switches only when the buckets it drains after it hold
 * fewer flights than there are waves, and a chunk too small to fill the waves at a
 * bucket's group width flies that bucket 2x or 4x wider.  0: one mapping throughout. */
int pr_set_tail_mapping(pr_ctx* ctx, int on);
/* 1 (default): every bucket whose mappings are compiled into the fused flight kernel is
 * flown by it -- one persistent launch per 1-s chunk drains all those buckets' work lists,
 * largest triangle counts first, each with its own lane mapping.  0: one launch per bucket
 * and mapping per chunk (per-mapping timings in pr_stats, used by autotuning). */
int pr_set_fused(pr_ctx* ctx, int on);
/* 1 (default): chunk 0's work lists hold each bucket's flights with their genomes in a
 * stable order of mesh size (triangle count / 8), so the groups of a wave -- which take
 * neighbouring list positions -- fly meshes of nearly equal size; 0: flight order.  Either
 * order is a function of the data alone (evaluations stay bit-reproducible).  No
 * reference counterpart (the reference flies one plane per block). */
int pr_set_list_order(pr_ctx* ctx, int sorted);
/* 1 (default): a chunk with at most one flight per two SIMDs (the long-flight tail of a
 * generation, small batches) is flown one flight per workgroup of 4 waves (at most one
 * flight per 4 SIMDs) or 2 waves, the flight's panels spread over all their lanes and the
 * partial force sums added in LDS each step: a step's panel work is 4 (2) times shorter
 * than with one wave per flight.  0: such chunks fly like the others.  The choice depends
 * on the chunk's flight count only (evaluations stay bit-reproducible); the two settings
 * sum the panels in different orders.  No reference counterpart (the reference flies one
 * plane per block). */
int pr_set_split(pr_ctx* ctx, int on);
int pr_fly_variants(int* n, int* L, int* P, int cap);

/* Readback of intermediate results (tests and tools). */
int pr_read_prepared(pr_ctx* ctx, int n, float* mass20, int* nt, int* nv, int* off);
int pr_read_panels(pr_ctx* ctx, long long first, int count, float* out7); /* cx cy cz A nx ny nz rows */
int pr_read_scenarios(pr_ctx* ctx, int n, float* out26);
/* genome order of the last flight batch's chunk-0 lists (pr_set_list_order 1) */
int pr_read_genome_order(pr_ctx* ctx, int n, int* out);
int pr_read_extent(pr_ctx* ctx, int n, float* ext3);                      /* mesh bounding-box extents */
int pr_read_states(pr_ctx* ctx, int first, int n, float* out16);       /* FlightState rows, 16 words */

/* ---------------------------------------------------------------- GA host side
 * Genome bounds (common.h:98-103), the GA's random stream and operators with the
 * reference's libstdc++ call order (sim.cu:315-325, 398, 429-430, 440-441, 472-490). */
void pr_gene_bounds(float* lo35, float* hi35);
void* pr_ga_new(unsigned seed);                       /* std::mt19937(seed) + the GA's distributions */
void pr_ga_free(void* ga);
void pr_ga_random(void* ga, float* genomes, int n);   /* grand(): uniform in bounds */
void pr_ga_sanitize(float* genomes, int n);           /* clamp to bounds */
void pr_ga_rank(const float* acc, int pop, int* idx); /* std::sort by score, descending */
void pr_ga_breed(void* ga, const float* gs, const float* acc, const int* idx, int pop, int gen, int gens,
                 float* next);
float pr_robust_score_host(const float* row, int n);

/* One generation's evaluation (sim.cu:432-456), batched: every genome under every seed in
 * one pr_fly, robust scores, std::sort ranking, then the top min(holdout_k, pop) genomes on
 * the 128-scenario holdout suite (seeds 100003 + p*7919) in one batch; champion = first
 * strict maximum of the holdout score.  d_genomes on the device; everything else host. */
typedef struct pr_gen pr_gen;
int pr_gen_create(pr_gen** out, int pop, int nseeds_max, int max_ch, int max_sp);
void pr_gen_destroy(pr_gen* e);
int pr_gen_evaluate(pr_gen* e, const float* d_genomes, const int* d_res, int pop, const int* seeds, int nseeds,
                    int holdout_k, const pr_sim_params* p, float* acc, int* idx, float* samples,
                    float* holdout, int* champion);
int pr_gen_stats(pr_gen* e, pr_stats* main_stats, pr_stats* holdout_stats);
/* Robustness statistics of one design's n scenario distances (evaluate_design's summary,
 * sim.cu:291-309): out[PR_NSTAT] indexed by PR_ST_*. Host function. */
enum {
    PR_ST_MIN, PR_ST_P05, PR_ST_P10, PR_ST_P25, PR_ST_MEDIAN, PR_ST_MEAN, PR_ST_SD, PR_ST_P75, PR_ST_P90,
    PR_ST_P95, PR_ST_MAX, PR_ST_CVAR10, PR_ST_FAILURES, PR_NSTAT
};
int pr_design_stats(const float* samples, int n, float* out);

/* Batched robustness audit -- evaluate_design (sim.cu:258-312) for n_genomes designs in one
 * flight batch: each flies the same n_scen scenarios (seed base_seed + s * 7919, by scenario
 * index).  d_genomes / d_res as pr_prepare; d_fitness, d_steps: device [n_genomes * n_scen];
 * stats: host [n_genomes * PR_NSTAT]; nthreads host threads for the statistics (0 = all). */
int pr_audit(pr_ctx* ctx, const float* d_genomes, const int* d_res, int n_genomes, int n_scen, int base_seed,
             const pr_sim_params* p, float* d_fitness, int* d_steps, float* stats, int nthreads);

/* The evaluator's flight contexts (0: population flights, 1 and 2: the two holdout slots),
 * so that the pr_set_* options apply to it; owned by the evaluator. */
pr_ctx* pr_gen_context(pr_gen* e, int which);

/* Pipelined generation evaluation: the population part of pr_gen_evaluate (acc, idx,
 * samples valid on return), with the top-8 holdout batch left flying on its own stream;
 * pr_gen_collect waits for the oldest such batch and returns its holdout score and
 * champion. Two batches may be outstanding, so generation g+1's population flights overlap
 * generation g's holdout. No reference counterpart: run_ga's holdout (sim.cu:446-455)
 * never feeds breeding, so deferring it changes no result. -2: both slots outstanding. */
int pr_gen_evaluate_async(pr_gen* e, const float* d_genomes, const int* d_res, int pop, const int* seeds,
                          int nseeds, int holdout_k, const pr_sim_params* p, float* acc, int* idx, float* samples);
int pr_gen_collect(pr_gen* e, float* holdout_out, int* champion_out);
/* Flight distances of genomes [first, first + count) of the last population batch (row-major,
 * one row of nseeds per genome), read after the fact: they stay on the device until the next
 * evaluation, so a caller can check a timed generation's own results without slowing it.
 * No reference counterpart (run_ga keeps its samples on the host, sim.cu:436-439). */
int pr_gen_read_samples(pr_gen* e, int first, int count, float* out);

/* ---------------------------------------------------------------- surrogate MLP
 * BF16 MLP 35 -> 256 -> 256 -> 256 -> 4 (ReLU) on the matrix cores, FP32 accumulation,
 * FP32 master weights + Adam (no reference counterpart: a pre-screen for candidate pools;
 * the flight simulation alone decides).  Targets per genome: mean distance, CVaR of the
 * worst quarter, worst distance [m], oversize flag (0/1).  Parameter layout (FP32):
 * W1[64x256] b1[256] W2[256x256] b2[256] W3[256x256] b3[256] W4[256x32] b4[32]. */
typedef struct pr_mlp pr_mlp;
int pr_mlp_param_count(void);
int pr_mlp_create(pr_mlp** out, int max_batch, unsigned seed);
void pr_mlp_destroy(pr_mlp* m);
const char* pr_mlp_last_error(void);
int pr_mlp_get_params(pr_mlp* m, float* out, int cap);
int pr_mlp_set_params(pr_mlp* m, const float* in, int n);
/* d_pred [n x 5]: mean, CVaR25, worst [m], oversize probability, predicted robust score */
int pr_mlp_predict(pr_mlp* m, const float* d_genomes, int n, float* d_pred);
/* mean loss + gradient of one batch (n <= max_batch); grads host, may be null */
int pr_mlp_loss_grad(pr_mlp* m, const float* d_genomes, const float* d_targets, int n, float* loss, float* grads);
/* Largest batch of the network (rounded up to 32); pr_pool trains with batches of this size. */
int pr_mlp_max_batch(pr_mlp* m);
/* Adam over contiguous mini-batches of (d_genomes [n x 35], d_targets [n x 4]) */
int pr_mlp_train(pr_mlp* m, const float* d_genomes, const float* d_targets, int n, int epochs, int batch, float lr,
                 float* last_loss);

/* ---------------------------------------------------------------- candidate-pool GA
 * Generation: pool = elite + Gaussian mutants (sigma = sigma_frac x gene range) of random
 * elite parents; fly the whole pool (model = null) or the elite + the candidates with the
 * best surrogate-predicted score (n_fly in total); new elite = best simulated scores;
 * top-8 holdout champion; with a model, the flown pairs join a replay buffer of buf_cap
 * and the network trains on it.  The simulation alone ranks the elite and the champion.
 * pr_pool_generation out[17]: best score, elite mean, holdout champion score, wall ms,
 * flight ms, surrogate ms, flights, plane-flight-s, training loss, fly-set size, best
 * holdout score so far; on the fly set's exploration sample (uniform draws from the pool):
 * Spearman rho of predicted vs simulated scores, mean absolute error (m) over the members
 * both call flyable, accuracy of the predicted oversize flag, and the least-squares weights
 * of the mutants' simulated scores on (parent's simulated score, mutant's predicted score,
 * parent's predicted score) (NaN until a generation with the elite's scores; the key of
 * pr_pool_set_parent_delta 2). */
typedef struct pr_pool pr_pool;
int pr_pool_create(pr_pool** out, int n_elite, int n_pool, int n_fly, int nseeds, int buf_cap, float sigma_frac,
                   unsigned seed);
void pr_pool_destroy(pr_pool* p);
const char* pr_pool_last_error(void);
int pr_pool_set_elite(pr_pool* p, const float* genomes, int n);
int pr_pool_get_elite(pr_pool* p, float* genomes, int n);
int pr_pool_generation(pr_pool* p, int gen, const pr_sim_params* prm, pr_mlp* model, int train_epochs, float lr,
                       double* out);
int pr_pool_add_training(pr_pool* p, const float* d_genomes, const float* samples, int n);
/* 1: plain generations also add their flown (genome, targets) pairs to the replay buffer
 * (bootstrapping the surrogate); 0 (default): only surrogate generations do. */
int pr_pool_set_collect(pr_pool* p, int on);
/* Fraction of the surrogate fly set (beyond the elite) drawn at random from the pool. */
int pr_pool_set_explore(pr_pool* p, float frac);
/* 1 (default): once a generation has simulated the elite, a mutant's selection key is its
 * parent's simulated score plus the network's predicted change from parent to mutant
 * (the error the network makes on both cancels); 2: a fitted combination a * parent's
 * simulated score + b * mutant's predicted score + c * parent's predicted score, (a, b, c)
 * from the least squares of the previous generation's exploration sample's simulated scores
 * on those three ((1, 1, -1) until a fit exists); 0, and before the first generation: the
 * predicted score itself.  No reference counterpart (the reference has no surrogate). */
int pr_pool_set_parent_delta(pr_pool* p, int on);
/* Weights (a, b, c) of mode 1's key a * parent's simulated score + b * mutant's predicted
 * score + c * parent's predicted score; (1, 1, -1) by default.  b = -c = lambda < 1 lets the
 * parent's simulated score choose the parents and the network mostly order each parent's
 * mutants.  No reference counterpart (the reference has no surrogate). */
int pr_pool_set_key_weights(pr_pool* p, float w_parent, float w_pred, float w_parent_pred);
/* Generations run without a model (model = null): 0 (default) fly the whole pool; 1 fly the
 * elite + a uniform random sample of the mutants; 2 fly the elite + the mutants of the
 * best-scoring elite parents (the parent's simulated score as the key; a uniform sample
 * before the first generation).  Modes 1 and 2 fly n_fly genomes chosen without the
 * network: the baselines against which the surrogate's selection is measured.  No reference
 * counterpart (the reference has no candidate pool). */
int pr_pool_set_selection(pr_pool* p, int mode);
/* 1: the regression targets (mean, CVaR, worst) of each collected batch -- a generation's
 * fly set, or one pr_pool_add_training call -- are stored relative to the batch's medians
 * over its flyable designs, which removes the scenario set's common effect on every
 * distance of that batch (default); 0: absolute targets.  pr_pool_last returns the stored
 * targets.  No reference counterpart. */
int pr_pool_set_target_norm(pr_pool* p, int on);
int pr_pool_train(pr_pool* p, pr_mlp* model, int epochs, float lr, float* loss);
int pr_pool_champion(pr_pool* p, float* genome35, float* holdout);
/* The last generation's pool [n_pool x 35], fly set [n_fly x 35] in flight order, its
 * simulated robust scores [n_fly] and, if it collected training pairs, their targets
 * [n_fly x 4]; null outputs are skipped; *n_fly = fly-set size. */
int pr_pool_last(pr_pool* p, float* pool, float* fly, float* scores, float* targets, int* n_fly);

/* ---------------------------------------------------------------- allocation-free evaluation
 * No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP stream capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * evaluation as pr_prepare + pr_scenarios + pr_fly with layout 0 (flight f = genome f / n_seeds
 * under seed f % n_seeds; results bit-identical to that path on a fresh context), at the
 * reference mesh resolution, queued on `stream` (a hipStream_t; null = the null stream) and
 * returning without waiting: d_fitness / d_steps are complete when the stream is.  Every
 * intermediate buffer lives in one 32 MiB device array inside the library, which each device
 * holds until the process ends; the call allocates nothing.  Limits: n_genomes <= 2048,
 * n_seeds <= 16 (-1 and a message beyond them).  d_genomes [n_genomes*35], d_seeds
 * [n_seeds] (device int32), d_fitness [n_genomes*n_seeds], d_steps (may be null) on the
 * device.  Calls that overlap on one device race for the array: serialise them (one stream,
 * or one host thread per device). */
int pr_eval_noalloc(const float* d_genomes, int n_genomes, const int* d_seeds, int n_seeds, const pr_sim_params* p,
                    float* d_fitness, int* d_steps, void* stream);

/* Device memory helpers so that hosts without a GPU array library can drive the API. */
int pr_dev_alloc(void** p, size_t bytes);
int pr_dev_free(void* p);
int pr_h2d(void* d_dst, const void* src, size_t bytes);
int pr_d2h(void* dst, const void* d_src, size_t bytes);
int pr_device_sync(void);
int pr_device_name(char* buf, int cap);

#ifdef __cplusplus
}
#endif
#endif /* PLANES_HIP_H */
