/*
 * mcscram_amd — MI355X-native Monte Carlo fault-tree engine, C ABI.
 *
 * Drop-in for the DirectEval Monte Carlo path of mcSCRAM (src/mc): the same
 * PDAG input, the same probability-analysis settings (trials, burn-in, sample
 * shape = batch_size x bitpacks_per_batch), the same Philox4x32-10 counter
 * layout and threshold conversion, and bit-identical per-node tallies.
 *
 * Plain C types only; every call returns an mcs_status and leaves a message
 * for mcs_last_error() on failure (the reference throws or exits instead:
 * src/mc/queue/layer_manager.tpp:43,55,245; kernel_builder.h:341,359,713).
 * One engine owns its device memory and one HIP stream. Calls on the same
 * engine are not re-entrant; distinct engines are independent.
 */
#ifndef MCSCRAM_AMD_H
#define MCSCRAM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mcs_status;
#define MCS_OK 0
#define MCS_ERR_INVALID_ARGUMENT 1 /* null pointer, bad size, p outside [0,1] */
#define MCS_ERR_GRAPH 2            /* duplicate/unknown index, root not a gate, cycle */
#define MCS_ERR_DEVICE 3           /* HIP runtime error */
#define MCS_ERR_OUT_OF_MEMORY 4    /* device allocation would not fit */
#define MCS_ERR_STATE 5            /* call not valid in the engine's state */

/* Connective numbering of the reference PDAG (src/pdag.h:265-275). */
enum mcs_connective {
    MCS_AND = 0, MCS_OR = 1, MCS_ATLEAST = 2, MCS_XOR = 3,
    MCS_NOT = 4, MCS_NAND = 5, MCS_NOR = 6, MCS_NULL = 7
};

/*
 * Flattened PDAG (replaces core::Pdag as the engine's input, src/pdag.h:298-397).
 * Variables: PDAG index (>= 2 in SCRAM, src/pdag.h:804) and probability.
 * Gates: PDAG index, connective, min number (ATLEAST only; stored as 8 bits
 * as the reference does, src/mc/event/node.h:676), and signed argument
 * indices (negative = complement) in insertion order: gate arguments are
 * visited in the order they appear in this list, as are variable arguments,
 * which is the order of the reference's two argument maps
 * (src/pdag.h:311-321). Arrays are copied at create time.
 */
typedef struct mcs_pdag {
    int32_t n_vars;
    const int32_t *var_index;   /* [n_vars] */
    const double *var_p;        /* [n_vars], each in [0, 1] */
    int32_t n_gates;
    const int32_t *gate_index;  /* [n_gates] */
    const uint8_t *gate_type;   /* [n_gates], enum mcs_connective */
    const int32_t *gate_k;      /* [n_gates] */
    const int64_t *arg_offsets; /* [n_gates + 1], CSR offsets into args */
    const int32_t *args;        /* [arg_offsets[n_gates]] */
    int32_t root;               /* PDAG index of the top gate */
} mcs_pdag;

/*
 * Sample shape (src/mc/event/sample_shape.h:103-130): one pass evaluates
 * batch_size * bitpacks_per_batch uint64 words = 64x that many trials per
 * node. The reference derives it from device limits
 * (src/mc/queue/sample_shaper.h:59-147); here it is an explicit input.
 */
typedef struct mcs_sample_shape {
    uint64_t batch_size;
    uint64_t bitpacks_per_batch;
} mcs_sample_shape;

/* Philox key. The reference hard-codes (382307844, 293830103)
 * (src/mc/prng/philox128.h:39-42); any other key is a non-reference mode. */
#define MCS_REFERENCE_KEY0 382307844u
#define MCS_REFERENCE_KEY1 293830103u

typedef struct mcs_engine mcs_engine;

/* Per-node tally in the reference's record layout (src/mc/stats/tally.h:18-33):
 * ci = {lower95, upper95, lower99, upper99}. */
typedef struct mcs_tally {
    int32_t index;
    uint64_t num_one_bits;
    uint64_t total_bits;
    double mean;
    double std_err;
    double ci[4];
} mcs_tally;

typedef struct mcs_engine_info {
    uint64_t num_bitpacks;     /* words per node per pass */
    uint64_t row_stride;       /* words between node rows in device memory */
    uint32_t iteration;        /* passes submitted so far (reference counter) */
    int32_t num_events;        /* basic events reachable from the root */
    int32_t num_gates;         /* gates reachable from the root */
    int32_t num_layers;        /* layers including layer 0 (basic events) */
    int32_t num_tallies;
    uint64_t num_args;         /* total gate inputs (sum of fan-in) */
    uint64_t device_bytes;     /* device memory owned by the engine */
    int32_t event_buffers;     /* 2: the next pass is sampled during this one */
    uint32_t launches_per_pass;/* kernel launches of a pass in steady state */
    int32_t fused_gates;       /* gates evaluated inside their only consumer's
                                  task: their words never reach device memory */
} mcs_engine_info;

/* Message of the last failed call on this thread ("" if none). */
const char *mcs_last_error(void);

/*
 * Build an engine: layered topological order, event positions, device
 * blocks and tallies. Replaces
 *   queue::layer_manager(core::Pdag*, num_trials, TallyNodeMap, overhead)
 *   (src/mc/queue/layer_manager.h:284, layer_manager.tpp:344-375)
 * with the sample shape given explicitly. tally_index lists the PDAG
 * indices to tally (the reference's TallyNodeMap keys); duplicates collapse.
 */
mcs_status mcs_engine_create(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                             const int32_t *tally_index, int32_t n_tally,
                             mcs_engine **out);

/* As mcs_engine_create with a non-reference Philox key and an explicit HIP
 * device ordinal (-1 = current device). */
mcs_status mcs_engine_create_ex(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                                const int32_t *tally_index, int32_t n_tally,
                                uint32_t key0, uint32_t key1, int32_t device,
                                mcs_engine **out);

/* Release all device memory (replaces ~layer_manager, layer_manager.tpp:281-301). */
mcs_status mcs_engine_destroy(mcs_engine *engine);

/* Enqueue n_passes passes without waiting: each advances the 1-based
 * iteration counter and runs sampler -> gate layers -> tallies
 * (replaces layer_manager::pass / single_pass, layer_manager.tpp:202-216;
 * queueable.h:725-731 for the counter). */
mcs_status mcs_engine_pass(mcs_engine *engine, uint32_t n_passes);

/* Block until the engine's stream is idle (the reference's wait_and_throw,
 * layer_manager.tpp:259). */
mcs_status mcs_engine_wait(mcs_engine *engine);

/* Enqueue total_passes passes, waiting every passes_between_waits passes
 * (0 = only at the end), then collect (layer_manager.tpp:225-267). */
mcs_status mcs_engine_pass_wait_collect(mcs_engine *engine, uint32_t total_passes,
                                        uint32_t passes_between_waits,
                                        mcs_tally *out, int32_t n_out);

/* Copy tallies out with host-side point estimates (collect_tallies and
 * stats::populate_point_estimates: layer_manager.tpp:270-279,
 * src/mc/stats/ci_utils.h:167-201). n_out must equal the number of distinct
 * tallied nodes; order = first appearance in tally_index. Waits first. */
mcs_status mcs_engine_collect(mcs_engine *engine, mcs_tally *out, int32_t n_out);

/* Point estimates of a tally from its raw counters: mean, standard error
 * and the normal-approximation 95 % / 99 % intervals, clamped to [0, 1]
 * (stats::populate_point_estimates, src/mc/stats/ci_utils.h:167-201).
 * Reads num_one_bits and total_bits, writes mean, std_err and ci. */
void mcs_tally_point_estimates(mcs_tally *t);

/* Engine geometry and counters. */
mcs_status mcs_engine_get_info(mcs_engine *engine, mcs_engine_info *out);

/* PDAG index of each basic event in event_idx order (the Philox counter's
 * second word is event_idx + 1; src/mc/kernel/basic_event.h:112-122).
 * n_out must be >= num_events. */
mcs_status mcs_engine_event_order(mcs_engine *engine, int32_t *pdag_index_out, int32_t n_out);

/* Host-only: the basic events' PDAG indices in event_idx order for a PDAG,
 * without a device (the same plan mcs_engine_create builds). *n_events
 * receives the count; pdag_index_out may be null to query it. */
mcs_status mcs_pdag_event_order(const mcs_pdag *pdag, int32_t *pdag_index_out, int32_t n_out,
                                int32_t *n_events);

/* Host-only: depth (layer) of each gate slot of the input arrays, -1 for
 * gates not reachable from the root. */
mcs_status mcs_pdag_gate_depths(const mcs_pdag *pdag, int32_t *depth_out, int32_t n_out);

/* Copy the words of one node's row from the last pass to host memory
 * (num_bitpacks words). For parity checks and debugging. */
mcs_status mcs_engine_read_words(mcs_engine *engine, int32_t pdag_index, uint64_t *out,
                                 uint64_t n_out);

/* Run n_passes passes with a HIP event recorded on the engine's stream
 * before and after every kernel launch, then wait. kernel_ms receives the
 * summed device time of the launches, launches their number, pass_ms the
 * device time from the first launch's start to the last launch's end.
 * Measurement hook; it has no counterpart in the reference (whose
 * throughput line is host wall-clock, src/mc/scheduler/progressbar.h:525-645). */
mcs_status mcs_engine_time_passes(mcs_engine *engine, uint32_t n_passes, double *kernel_ms,
                                  uint32_t *launches, double *pass_ms);

/* Checkpoint / resume: set the iteration counter and the raw one-bit counts
 * (order as in collect). The Philox counters are a pure function of
 * (iteration, word), so resuming at iteration i continues the same stream. */
mcs_status mcs_engine_restore(mcs_engine *engine, uint32_t iteration,
                              const uint64_t *num_one_bits, int32_t n);

/*
 * Probability-analysis settings used by the DirectEval entry point
 * (src/settings.h:439-464 defaults; src/mc/scheduler/convergence_controller.h).
 */
typedef struct mcs_settings {
    uint64_t num_trials;       /* 0 = run until the early-stop rule (auto) */
    uint64_t burn_in_trials;   /* reference default 2^20 */
    double ci_confidence;      /* reference default 0.99 */
    double ci_rel_margin;      /* relative half-width target, default 0.001 */
    uint32_t max_passes;       /* safety cap for auto mode (0 = 1<<20) */
    uint32_t key0, key1;       /* Philox key (reference key by default) */
    uint32_t ci_policy;        /* early-stop rule: MCS_CI_BAYES (reference
                                  default) or MCS_CI_WALD (src/settings.h:51, 426) */
    uint32_t early_stop;       /* 1: stop once converged (src/settings.h:370-372,
                                  453; default 1). num_trials == 0 implies it
                                  (src/settings.cc:105-112); with num_trials > 0
                                  the run stops at whichever of the trial count
                                  and convergence comes first, as the CLI's
                                  "-N n --early-stop" does
                                  (targets/scram/utils/ConstructSettings.h:103-109).
                                  "-N n" alone is early_stop = 0. */
} mcs_settings;

/* Stopping rules of the reference's convergence policy
 * (src/mc/scheduler/convergence_policy.h; numbering of src/settings.h:51). */
enum { MCS_CI_BAYES = 0, MCS_CI_WALD = 1 };

/* Convergence state of one tally: current and target half-widths in linear
 * and log10 space, and the trials each target needs. Mirrors the linear and
 * log10 fields the reference's policy writes into stats::tally. */
typedef struct mcs_convergence {
    double epsilon, target_epsilon;
    uint64_t target_trials;
    double epsilon_log10, target_epsilon_log10;
    uint64_t target_trials_log10;
} mcs_convergence;

/* One policy update for a tally of ones true trials out of total at normal
 * quantile z and relative margin rel_margin. Replaces
 * detail::update_wald / detail::update_bayes
 * (src/mc/scheduler/convergence_policy.h:28-61, 66-143). */
mcs_status mcs_convergence_update(uint32_t ci_policy, uint64_t ones, uint64_t total,
                                  double rel_margin, double z, mcs_convergence *out);

/* Quantile of the Beta(a, b) distribution: x with I_x(a, b) = q. The
 * reference calls Boost.Math's quantile(beta_distribution) for the Bayes
 * policy (src/mc/scheduler/convergence_policy.h:97-99); this one is written
 * from the regularized incomplete beta function. Returns NaN for a, b <= 0
 * or q outside [0, 1]. */
double mcs_beta_quantile(double a, double b, double q);

/* Two-sided standard-normal quantile z with P(|Z| <= z) = confidence
 * (Acklam's inverse normal; the reference's src/mc/stats/ci_utils.h:30-92). */
double mcs_normal_quantile_two_sided(double confidence);

/* Fill settings with the reference defaults. */
void mcs_settings_default(mcs_settings *s);

/* Reference trial rounding and pass count: trials rounded to a multiple of
 * 64 (ties down), passes = ceil(trials / (64 * words per pass))
 * (src/mc/queue/sample_shaper.h:73-83,139-146). */
mcs_status mcs_plan_passes(uint64_t num_trials, const mcs_sample_shape *shape,
                           uint64_t *rounded_trials, uint64_t *passes);

/* The sample shape the reference's sample_shaper picks on a device with the
 * given limits (src/mc/queue/sample_shaper.h:59-147 on a non-NVIDIA device,
 * with compute_closest_sample_shape_for_bitpacks, :274-334): trials rounded
 * as in mcs_plan_passes (0 = no limit: 2^26 trials size the shape and
 * *passes is 0); words per pass = the power-of-two product batch x bitpacks,
 * batch <= max_batch and bitpacks <= max_bitpacks (the device's work-item
 * size limits), closest to the rounded trials / 64 capped by
 * max_alloc_bytes * 8 / num_nodes bits (ties to the smaller), collapsed to
 * batch 1. Pass the reference device's limits to reproduce its shape, and
 * with it its bits. */
mcs_status mcs_sample_shape_for(uint64_t num_trials, uint64_t num_nodes, uint64_t max_alloc_bytes,
                                uint64_t max_batch, uint64_t max_bitpacks,
                                mcs_sample_shape *shape, uint64_t *rounded_trials,
                                uint64_t *passes);

/* As mcs_sample_shape_for, with the reference shaper's vendor rule
 * (src/mc/queue/sample_shaper.h:120-134): when nvidia_gpu is non-zero, words
 * per pass = floor((max_alloc_bytes / 8) / (num_nodes * (1 + overhead_ratio)))
 * with overhead_ratio clamped to [0, 1e6] (the reference's default is 0.05,
 * src/settings.h:462), rounded down to a multiple of cache_line_bytes / 8
 * words, whatever the trial count; batch 1. nvidia_gpu = 0 is
 * mcs_sample_shape_for. The reference's only published run used an NVIDIA
 * device, so this is the rule that reproduces its shape. Fails if no whole
 * cache line fits per node (the reference would divide by zero). */
mcs_status mcs_sample_shape_for_device(uint64_t num_trials, uint64_t num_nodes,
                                       uint64_t max_alloc_bytes, uint64_t max_batch,
                                       uint64_t max_bitpacks, int32_t nvidia_gpu,
                                       double overhead_ratio, uint64_t cache_line_bytes,
                                       mcs_sample_shape *shape, uint64_t *rounded_trials,
                                       uint64_t *passes);

/*
 * Entry point of the path: build the engine, run burn-in passes until
 * burn_in_trials are done, then passes until the fixed trial count is
 * reached or (with early stop) every tally's half-width meets ci_rel_margin
 * under ci_policy in linear and log10 space after a pass, whichever comes
 * first; collect tallies and return the root's mean. Convergence is only
 * checked after a post-burn-in pass, so early stop always runs at least one
 * pass after burn-in (convergence_controller.h:294-328, 351-365).
 * Replaces ProbabilityAnalyzer<mc::DirectEval>::CalculateTotalProbability /
 * ComputeTallies
 * (src/mc/core/probability_analysis.cc:88-209) and
 * convergence_controller::run_to_convergence (convergence_controller.h:371-397).
 * observe lists extra PDAG indices to tally besides the root; out receives
 * root first, then the observed nodes (n_out >= 1 + distinct observed).
 */
mcs_status mcs_probability_analysis(const mcs_pdag *pdag, const mcs_settings *settings,
                                    const mcs_sample_shape *shape,
                                    const int32_t *observe, int32_t n_observe,
                                    mcs_tally *out, int32_t n_out, double *p_total,
                                    uint32_t *passes_run);

/*
 * mcs_probability_analysis without any device allocation. No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (serving loops and harnesses that forbid hipMalloc). The engine's device
 * memory (words of every node, descriptors, argument rows, one-bit counters)
 * lives in a fixed per-device array of MCS_NOALLOC_SCRATCH_BYTES, which each
 * GPU the call runs on holds until the process ends; a PDAG and shape whose
 * engine does not fit fail with MCS_ERR_OUT_OF_MEMORY and a message giving
 * both sizes (one event buffer at 1 + events + gates rows of words per pass,
 * 8 bytes a word, plus its side arrays). The call neither allocates nor frees
 * device memory and creates no stream: it launches on `stream` (a
 * hipStream_t of the current device; null = the null stream) and waits on
 * it. Results equal mcs_probability_analysis's bit for bit. Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other threads,
 * other streams) race and must be serialised by the caller.
 */
#define MCS_NOALLOC_SCRATCH_BYTES (64u << 20)
mcs_status mcs_probability_analysis_noalloc(const mcs_pdag *pdag, const mcs_settings *settings,
                                            const mcs_sample_shape *shape,
                                            const int32_t *observe, int32_t n_observe,
                                            mcs_tally *out, int32_t n_out, double *p_total,
                                            uint32_t *passes_run, void *stream);


/*
 * Batches of independent trees evaluated by one persistent launch per pass.
 * No counterpart in the reference, which builds one layer_manager per PDAG
 * (src/mc/queue/layer_manager.tpp:344-375); each tree of a batch gives the
 * same tallies as its own engine (same Philox key, counters and sample
 * shape). A tree is evaluated whole in LDS: it must fit at one word per node
 * in 160 KiB (about 13,600 nodes); larger trees use mcs_engine_create. Every tree advances by one iteration per pass.
 */
typedef struct mcs_batch mcs_batch;

typedef struct mcs_batch_config {
    uint32_t lds_bytes;       /* LDS per workgroup of the fused launch (0 = 64 KiB) */
    uint32_t max_tile_words;  /* words per tile, power of two <= 64 (0 = 64) */
    uint32_t item_words;      /* largest work item in words (0 = 4096); smaller
                                 when the batch has too few tiles to fill the GPU */
    uint32_t use_graph;       /* 1: from the second pass on, capture a whole pass (the
                                 counter kernel, the fused launch, every layered
                                 pipeline's launches, their fork and join) as a
                                 hipGraph, one per event-area pattern (two at
                                 most), and replay it: one graph launch per pass */
    uint32_t key0, key1;      /* Philox key (the reference key by default) */
    uint32_t workgroup_lanes; /* 64..1024, a multiple of 64 (0 = 1024) */
    uint32_t min_fused_tile_words; /* narrowest LDS tile for the fused launch
                                      (0 = 8; 1 = every tree that fits; 128 = none) */
    uint32_t layered_streams; /* layered pipelines per sample shape, each with its
                                 own words arena and HIP stream, run side by side
                                 and beside the fused launch (0 = 2; at most 8) */
} mcs_batch_config;

typedef struct mcs_batch_info {
    int32_t num_trees;
    int32_t num_fused_trees;  /* trees in the persistent LDS launch */
    int32_t num_layered_trees;/* trees in layered pipelines */
    uint32_t layered_launches;/* layered kernel launches per pass */
    uint32_t num_items;       /* work items per pass */
    uint32_t grid;            /* persistent workgroups per launch */
    uint32_t lds_bytes;       /* LDS per workgroup */
    uint32_t workgroup_lanes;
    uint32_t iteration;       /* passes run */
    uint64_t device_bytes;
} mcs_batch_info;

/* Fill a config with the defaults. */
void mcs_batch_config_default(mcs_batch_config *cfg);

/* pdags[n_trees] with shapes[n_trees]; tallies of tree t are
 * tally_index[tally_offsets[t] .. tally_offsets[t+1]) (an empty range tallies
 * the root). cfg may be null (defaults). */
mcs_status mcs_batch_create(const mcs_pdag *pdags, int32_t n_trees,
                            const mcs_sample_shape *shapes, const int32_t *tally_offsets,
                            const int32_t *tally_index, const mcs_batch_config *cfg,
                            mcs_batch **out);
mcs_status mcs_batch_destroy(mcs_batch *batch);
/* Enqueue n_passes passes (one launch each) without waiting. */
mcs_status mcs_batch_pass(mcs_batch *batch, uint32_t n_passes);
mcs_status mcs_batch_wait(mcs_batch *batch);
/* Tallies of one tree, in the order of its tally list (waits first). */
mcs_status mcs_batch_collect(mcs_batch *batch, int32_t tree, mcs_tally *out, int32_t n_out);
mcs_status mcs_batch_get_info(mcs_batch *batch, mcs_batch_info *out);
/* Tile width (words) chosen for a tree; 0 for a tree in a layered pipeline. */
mcs_status mcs_batch_tile_words(mcs_batch *batch, int32_t tree, uint32_t *out);
/* Run n_passes with HIP events around each pass on the batch's stream;
 * kernel_ms receives the summed device time. Measurement hook. */
mcs_status mcs_batch_time_passes(mcs_batch *batch, uint32_t n_passes, double *kernel_ms);

/*
 * Uncertainty analysis: one tree, K parameter samples (basic-event
 * probability vectors), each an independent plain Monte Carlo run. The
 * reference's outer loop (UncertaintyAnalyzer::Sample,
 * src/uncertainty_analysis.h:154-169) calls the probability calculator once
 * per sample; its DirectEval engine ignores the sampled probabilities
 * (src/mc/queue/kernel_builder.h:179-181), so there is no reference Monte
 * Carlo counterpart. Here the samples share the tree's gate descriptors and
 * run in groups: one launch evaluates a gate layer for every sample of a
 * group, and a group's gate launches sample the basic events of the next
 * group. Sample k draws from Philox key (key0 + sample_id[k], key1), or from
 * (key0, key1) for every sample with common_random_numbers. The grouping
 * changes no bit: each sample's tallies equal a one-sample run of the same
 * key.
 */
typedef struct mcs_ua mcs_ua;

typedef struct mcs_ua_config {
    uint32_t key0, key1;             /* base key (the reference key by default) */
    uint32_t common_random_numbers;  /* 1: every sample draws the same stream */
    uint32_t samples_per_launch;     /* samples per group (0: about 16,384 tasks per launch,
                                        at most 64 samples, arenas within 96 GiB) */
    uint32_t pipelines;              /* pipelines on their own streams (0: up to 2) */
    uint32_t chunk_words;            /* words per task: 512, 1024, 2048 (0: the longest the
                                        words per pass fill) */
} mcs_ua_config;

void mcs_ua_config_default(mcs_ua_config *cfg);

/* p_samples[k * pdag->n_vars + i] is variable i's probability (the order of
 * pdag->var_index) in sample k; values outside [0, 1] are clamped as the
 * reference clamps sampled deviates (src/uncertainty_analysis.cc:108-120).
 * sample_ids may be null (sample k has id k). tally_index lists the nodes to
 * count in every sample (n_tally = 0 counts the root). */
mcs_status mcs_ua_create(const mcs_pdag *pdag, const mcs_sample_shape *shape, int32_t n_samples,
                         const double *p_samples, const uint32_t *sample_ids,
                         const int32_t *tally_index, int32_t n_tally, const mcs_ua_config *cfg,
                         mcs_ua **out);
mcs_status mcs_ua_destroy(mcs_ua *ua);
/* Replaces the engine's parameter samples with a new set (same layout and
 * clamping as mcs_ua_create; sample_ids may be null), keeping its device
 * arenas: the engine then behaves as one created with these samples
 * (iteration 0, zero counters), bit for bit. Large arenas are costly to
 * allocate again after a free, so a caller running many sets of samples
 * through one tree loads each set into one engine. Waits for the engine's
 * earlier work. After a failure the engine refuses to run until a set loads. */
mcs_status mcs_ua_set_samples(mcs_ua *ua, int32_t n_samples, const double *p_samples,
                              const uint32_t *sample_ids);
/* Every sample advances by n_passes iterations (enqueued, not waited for). */
mcs_status mcs_ua_run(mcs_ua *ua, uint32_t n_passes);
/* ones[k * n_tally + j] = one bits of tally j in sample k; *total_bits the
 * trials per sample so far. Waits first. */
mcs_status mcs_ua_collect(mcs_ua *ua, uint64_t *ones, int64_t n_out, uint64_t *total_bits);
/* Device time of n_passes over every sample (measurement hook). */
mcs_status mcs_ua_time_run(mcs_ua *ua, uint32_t n_passes, double *kernel_ms);
/* Samples per launch and pipelines chosen at create (fewer than asked when
 * the grid or device memory requires), words per task, and the device bytes
 * held. Any output may be null. */
mcs_status mcs_ua_info(const mcs_ua *ua, uint32_t *samples_per_launch, uint32_t *pipelines,
                       uint32_t *chunk_words, uint64_t *device_bytes);

/*
 * Independence approximation of the top event for many parameter samples
 * (no counterpart in the reference): the closed-form baseline whose residual
 * the uncertainty-analysis surrogate learns. Each gate's probability with
 * its inputs treated as independent (AND: product; OR: one minus the
 * product of complements; XOR: parity recursion; ATLEAST: Poisson-binomial
 * tail, with min(k, n - k + 1) at most 16; NOT/NULL: the last input;
 * NAND/NOR/NOT complemented), in FP64 on the GPU, one thread per sample.
 * p_samples[k * pdag->n_vars + i] is variable i's probability (the order
 * of pdag->var_index) in sample k; logit_out[k] receives
 * ln(p_top / (1 - p_top)) with p_top clamped to [1e-300, 1 - 2^-52].
 */
mcs_status mcs_independence_logits(const mcs_pdag *pdag, const float *p_samples,
                                   int64_t n_samples, double *logit_out);

/*
 * Lognormal parameter samples of the basic events, drawn on the GPU. Replaces
 * the host-side draws of the reference's outer loop (LognormalDeviate,
 * src/expression/random_deviate.cc:91-108, one call per event and sample in
 * UncertaintyAnalyzer::Sample, src/uncertainty_analysis.h:154-169, clamped
 * to [0, 1] as in src/uncertainty_analysis.cc:108-120); the stream is this
 * engine's own (counter-based, so a row does not depend on how rows are
 * split over calls), not the reference's std::mt19937 one.
 * Row r (first_row <= r < first_row + n_rows), variable i:
 * z = standard normal by Box-Muller from Philox4x32-10 with key
 * (seed mod 2^32, seed >> 32) and counter (i / 2, r mod 2^32, r >> 32,
 * 0x4C4E524D): u1 = (top 53 bits of (y0, y1) + 1/2) / 2^53, u2 = top 53 bits
 * of (y2, y3) / 2^53, z = sqrt(-2 ln u1) cos(2 pi u2) for even i and
 * sqrt(-2 ln u1) sin(2 pi u2) for odd i; p = min(1, exp(mu_i + sigma_i z))
 * in FP64. z_out (FP32) and p_out are host arrays
 * [n_rows * n_vars] (either may be null).
 */
mcs_status mcs_lognormal_samples(const double *mu, const double *sigma, int32_t n_vars,
                                 uint64_t seed, int64_t first_row, int64_t n_rows, float *z_out,
                                 double *p_out);

/*
 * Surrogate network for uncertainty analysis (no counterpart in the
 * reference): an MLP n_in -> hidden x 3 (GELU) -> 1 with BF16 weights and
 * activations, FP32 accumulation on the matrix cores, mean-squared-error
 * loss and Adam on FP32 master weights. Host arrays are FP32, row-major.
 * Parameter order: W1 [n_in x hidden], b1, W2 [hidden x hidden], b2, W3, b3,
 * w4 [hidden], b4 (the layer product is x W).
 */
typedef struct mcs_mlp mcs_mlp;

/* hidden must be a multiple of 64; weights He-uniform from a SplitMix64 seed. */
mcs_status mcs_mlp_create(int32_t n_in, int32_t hidden, uint64_t seed, mcs_mlp **out);
mcs_status mcs_mlp_destroy(mcs_mlp *net);
mcs_status mcs_mlp_num_params(mcs_mlp *net, int64_t *n);
mcs_status mcs_mlp_get_params(mcs_mlp *net, float *out, int64_t n);
/* Sets the FP32 master weights (and their BF16 copy); resets Adam's state. */
mcs_status mcs_mlp_set_params(mcs_mlp *net, const float *in, int64_t n);
/* Fixed (not trained) input map applied to every input column before the
 * first layer: u_i = min(cap_i, exp(scale_i * x_i + shift_i)), in FP32, then
 * rounded to BF16. With x_i the standard-normal score of ln p_i, scale_i =
 * sigma_i and shift_i = mu_i - ln s_i, u_i is p_i / s_i (the probability on a
 * per-event scale s_i), which the top event depends on far more linearly
 * than on ln p_i. Each array has n_in entries; all three null clears the
 * map (inputs then go to the first layer as given). */
mcs_status mcs_mlp_set_input_map(mcs_mlp *net, const float *scale, const float *shift,
                                 const float *cap);
mcs_status mcs_mlp_predict(mcs_mlp *net, const float *x, int64_t n_rows, float *y_out);
/* Mean squared error of one batch (n_rows a multiple of 64) and, if grad is
 * not null, its gradient in parameter order. */
mcs_status mcs_mlp_loss_grad(mcs_mlp *net, const float *x, const float *y, int64_t n_rows,
                             float *loss_out, float *grad_out);
/* Decoupled weight decay of the Adam steps (AdamW: each step also takes
 * lr * weight_decay * w off every parameter w); 0, the default, is plain
 * Adam. */
mcs_status mcs_mlp_set_weight_decay(mcs_mlp *net, float weight_decay);
/* Adam over shuffled minibatches (batch a multiple of 64) with a cosine
 * learning-rate schedule: epoch e uses lr * (1 + cos(pi e / epochs)) / 2.
 * epoch_loss[e] receives the mean training loss of epoch e (may be null). */
mcs_status mcs_mlp_fit(mcs_mlp *net, const float *x, const float *y, int64_t n_rows, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float *epoch_loss);
/* Test hook for the matrix-core product: c = A x B with A, B rounded to BF16;
 * a is [M x K] (or [K x M] read transposed), b is [K x N] (or [N x K]).
 * M, N multiples of 64, K of 32. */
mcs_status mcs_mlp_gemm_test(const float *a, const float *b, int32_t M, int32_t N, int32_t K,
                             int32_t trans_a, int32_t trans_b, float *c);

#ifdef __cplusplus
}
#endif

#endif /* MCSCRAM_AMD_H */
