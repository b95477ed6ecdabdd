// Uncertainty-analysis engine: K parameter samples of one tree, G samples
// per launch (kernel: pass_layer in kernels.hip).
//
// The samples share the tree's plan, schedule, gate descriptors and
// argument rows; each has its own event descriptors (thresholds) and Philox
// key. The samples are split into up to two pipelines, each with its own
// HIP stream and words arena, so one pipeline's launch tails leave the CUs
// to the other. A pipeline runs its samples in groups of G: one launch
// evaluates a gate layer for every sample of the group (the sample is a task
// dimension of pass_layer), so a pass of K samples is about K / G times the
// tree's launches instead of K times. The arena holds G sample slots, each
// with two event areas (which successive groups alternate between) and one
// gate area, `stride` words per row. Group g's gate launches carry the
// sampler tasks of group g + 1, so Philox work and gate traffic overlap
// across groups; the pipeline's last group of a pass samples its first group
// of the next pass.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "launch_plan.h"
#include "layering.h"
#include "mcscram_amd.h"
#include "status.h"

using mcs::fail;

namespace {

constexpr uint32_t kUaSplitFanIn = 64;  // as the single-tree engine
constexpr uint32_t kUaTaskRows = 64;
constexpr uint32_t kUaPipes = 2;        // pipelines when the memory allows
constexpr uint32_t kUaMaxGroup = 64;    // samples per launch at most, by default
constexpr uint64_t kUaLaunchTasks = 16384;  // tasks per launch the default grouping aims at
constexpr uint64_t kUaArenaCap = uint64_t(96) << 30;  // default bound on all arenas

// Words per workgroup task: the longest of 2048, 1024 and 512 that the row
// fills. Long tasks amortise a task's set-up (descriptor and summary loads,
// the tally's wave reduction, the chunk summary) over more words; the
// parallelism short tasks would give comes from grouping samples instead.
// (Measured on the 25,000-event tree at 2^14 words per sample: 9.26 ms per
// sample with 512-word tasks, 7.40 with 1024, 6.92 with 2048.)
uint32_t pick_chunk_words(uint64_t n_bp) {
    if (n_bp >= 2048) return 2048;
    if (n_bp >= 1024) return 1024;
    return 512;
}

struct UaPipe {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;           // recorded after a run's launches
    uint64_t *d_words = nullptr;         // the pipeline's words arena (G slots)
    uint8_t *d_sum = nullptr;            // chunk summaries of the arena's rows
    uint32_t k_begin = 0, k_end = 0;     // its samples
    uint64_t seq = 0;                    // group-passes run: event area = seq & 1
    int64_t primed = -1;                 // group (first sample) sampled into area seq & 1 ...
    uint32_t primed_iter = 0;            // ... for this iteration
};

}  // namespace

struct mcs_ua {
    int device = 0;
    hipStream_t stream = nullptr;
    mcs_ua_config cfg{};
    mcs::Plan plan;
    mcs::Schedule sch;
    mcs::TreeLaunchPlan lp;
    mcs::LaunchGeometry geo{};
    uint32_t n_ev = 0, gate_rows = 0;
    int32_t n_samples = 0;
    uint32_t group = 1;                  // samples per launch
    uint64_t slot_rows = 0;              // rows of one sample slot of an arena
    std::vector<int32_t> tally_nodes;    // plan node ids
    uint64_t slots = 0;                  // counters per sample (gates, partials, events)
    uint32_t iteration = 0;
    std::vector<UaPipe> pipes;
    hipEvent_t began = nullptr;          // recorded before a run's launches
    mcs::EventDesc *d_events = nullptr;  // n_samples x n_ev
    uint32_t *d_keys = nullptr;          // n_samples Philox key0 words
    mcs::GateDesc *d_gates = nullptr;
    uint32_t *d_args[2] = {nullptr, nullptr};
    unsigned long long *d_ones = nullptr;  // pad of 2 n_ev, then n_samples x slots
    int32_t capacity = 0;                // samples d_events, d_keys and d_ones hold
    int32_t n_vars = 0;                  // variables per row of p_samples
    std::vector<int32_t> var_of_event;   // p_samples column of each event
    uint64_t device_bytes = 0;

    ~mcs_ua() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        for (UaPipe &q : pipes)
            if (q.stream) (void)hipStreamSynchronize(q.stream);
        for (void *p : {static_cast<void *>(d_events), static_cast<void *>(d_keys),
                        static_cast<void *>(d_gates), static_cast<void *>(d_args[0]),
                        static_cast<void *>(d_args[1]), static_cast<void *>(d_ones)})
            if (p) (void)hipFree(p);
        for (UaPipe &q : pipes) {
            if (q.d_words) (void)hipFree(q.d_words);
            if (q.d_sum) (void)hipFree(q.d_sum);
            if (q.done) (void)hipEventDestroy(q.done);
            if (q.stream) (void)hipStreamDestroy(q.stream);
        }
        if (began) (void)hipEventDestroy(began);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// The per-sample strides of a launch over a group of samples in one arena.
mcs::PassLaunch group_launch(const mcs_ua *u, const UaPipe &q) {
    mcs::PassLaunch pl{};
    pl.key1 = u->cfg.key1;
    pl.state = q.d_sum;
    pl.state_stride = mcs::state_stride_for(u->geo.n_chunks);
    pl.sample_words = u->slot_rows * u->geo.stride;
    pl.sample_state = u->slot_rows * pl.state_stride;
    pl.sample_ones = u->slots;
    return pl;
}

// Sampler tasks of group (k0, size gk) for iteration it into event area `area`.
void set_sampling(const mcs_ua *u, mcs::PassLaunch &pl, uint32_t k0, uint32_t it, uint32_t area) {
    pl.events = u->d_events + uint64_t(k0) * u->n_ev;
    pl.keys0 = u->d_keys + k0;
    pl.samp_events = u->n_ev;
    pl.samp_iteration = it;
    pl.samp_row0 = area * u->n_ev;
}

// Loads a set of parameter samples into an engine whose arenas exist: event
// descriptors and keys (arrays grown when the set is larger than they
// hold), zero counters, each pipeline's share of the samples, and the state
// of a fresh engine (iteration 0, no group sampled, every chunk summary
// "mixed"). Waits for the engine's work first and for the uploads after.
mcs_status load_samples(mcs_ua *u, int32_t n_samples, const double *p_samples,
                        const uint32_t *sample_ids) {
    // No samples are loaded until this call succeeds: runs are refused.
    u->n_samples = 0;
    std::vector<mcs::EventDesc> ev(size_t(n_samples) * u->n_ev);
    for (int32_t k = 0; k < n_samples; ++k) {
        const double *p = p_samples + size_t(k) * size_t(u->n_vars);
        for (uint32_t e = 0; e < u->n_ev; ++e) {
            double pe = p[u->var_of_event[e]];
            if (!(pe == pe)) return fail(MCS_ERR_INVALID_ARGUMENT, "sampled probability is NaN");
            pe = pe > 1.0 ? 1.0 : (pe < 0.0 ? 0.0 : pe);
            ev[size_t(k) * u->n_ev + e] = mcs::make_event_desc(u->plan.event_pdag_index[e], e, pe);
        }
    }
    std::vector<uint32_t> keys(static_cast<size_t>(n_samples));
    for (int32_t k = 0; k < n_samples; ++k) {
        const uint32_t id = sample_ids ? sample_ids[k] : uint32_t(k);
        keys[size_t(k)] = u->cfg.common_random_numbers ? u->cfg.key0 : u->cfg.key0 + id;
    }
    MCS_HIP(hipStreamSynchronize(u->stream));
    for (UaPipe &q : u->pipes) MCS_HIP(hipStreamSynchronize(q.stream));
    const uint64_t ones_n = 2ull * u->n_ev + uint64_t(n_samples) * u->slots;
    if (n_samples > u->capacity) {
        for (void **q : {reinterpret_cast<void **>(&u->d_events), reinterpret_cast<void **>(&u->d_keys),
                         reinterpret_cast<void **>(&u->d_ones)})
            if (*q) {
                MCS_HIP(hipFree(*q));
                *q = nullptr;
            }
        u->capacity = 0;
        MCS_HIP(hipMalloc(&u->d_events, ev.size() * sizeof(mcs::EventDesc)));
        MCS_HIP(hipMalloc(&u->d_keys, keys.size() * 4));
        MCS_HIP(hipMalloc(&u->d_ones, ones_n * 8));
        u->capacity = n_samples;
    }
    MCS_HIP(hipMemcpyAsync(u->d_events, ev.data(), ev.size() * sizeof(mcs::EventDesc),
                           hipMemcpyHostToDevice, u->stream));
    MCS_HIP(hipMemcpyAsync(u->d_keys, keys.data(), keys.size() * 4, hipMemcpyHostToDevice, u->stream));
    MCS_HIP(hipMemsetAsync(u->d_ones, 0, ones_n * 8, u->stream));
    const uint64_t sum_bytes = uint64_t(u->group) * u->slot_rows * mcs::state_stride_for(u->geo.n_chunks);
    const size_t n_pipes = u->pipes.size();
    for (size_t i = 0; i < n_pipes; ++i) {
        UaPipe &q = u->pipes[i];
        q.k_begin = uint32_t(uint64_t(n_samples) * i / n_pipes);
        q.k_end = uint32_t(uint64_t(n_samples) * (i + 1) / n_pipes);
        q.seq = 0;
        q.primed = -1;
        q.primed_iter = 0;
        MCS_HIP(hipMemsetAsync(q.d_sum, mcs::kChunkMixed, sum_bytes, u->stream));
    }
    MCS_HIP(hipStreamSynchronize(u->stream));
    u->n_samples = n_samples;
    u->iteration = 0;
    return MCS_OK;
}

mcs_status enqueue_run(mcs_ua *u, uint32_t n_passes) {
    if (u->n_samples <= 0) return fail(MCS_ERR_STATE, "no parameter samples loaded");
    const uint64_t samp_total = uint64_t(u->n_ev) * u->geo.n_chunks;  // per sample
    const uint32_t G = u->group;
    // Sampler tasks per gate launch, for a group of each size that occurs.
    std::vector<std::vector<uint64_t>> split(G + 1);
    auto split_for = [&](uint32_t gk) -> const std::vector<uint64_t> & {
        if (split[gk].empty()) split[gk] = mcs::split_in_proportion(u->lp.launch_cost, gk * samp_total);
        return split[gk];
    };
    // Fork: the pipelines start after what the engine's stream holds.
    MCS_HIP(hipEventRecord(u->began, u->stream));
    for (UaPipe &q : u->pipes) MCS_HIP(hipStreamWaitEvent(q.stream, u->began, 0));
    // One group's launches on its pipeline's stream: the group's own
    // sampling unless the previous group's launches already drew it, then
    // its gate launches, which carry the sampling of the pipeline's next
    // group (or, for its last group, of its first group's next pass).
    auto enqueue_group = [&](UaPipe &q, uint32_t k0, uint32_t it) -> mcs_status {
        const uint32_t gk = std::min(G, q.k_end - k0);
        const uint32_t area = uint32_t(q.seq & 1u);
        const mcs::PassLaunch base = group_launch(u, q);
        if (!(q.primed == int64_t(k0) && q.primed_iter == it)) {
            mcs::PassLaunch pl = base;
            set_sampling(u, pl, k0, it, area);
            pl.samp_end = gk * samp_total;
            MCS_HIP(mcs::launch_pass(q.stream, pl, q.d_words, u->d_ones, u->geo));
        }
        const bool last = k0 + G >= q.k_end;
        const uint32_t nk0 = last ? q.k_begin : k0 + G;
        const uint32_t nit = last ? it + 1 : it;
        const uint32_t ngk = std::min(G, q.k_end - nk0);
        const bool prefetch = nit != 0;  // iteration counter not wrapping
        const std::vector<uint64_t> &sp = split_for(ngk);
        unsigned long long *ones = u->d_ones + uint64_t(k0) * u->slots;  // + pad - 2 n_ev
        for (size_t l = 0; l + 1 < u->lp.launch_begin.size(); ++l) {
            mcs::PassLaunch pl = base;
            pl.gates = u->d_gates + u->lp.launch_begin[l];
            pl.tree_gates = u->d_gates;
            pl.arg_rows = u->d_args[area];
            pl.n_gates = u->lp.launch_begin[l + 1] - u->lp.launch_begin[l];
            pl.n_samples = gk;
            if (prefetch) {
                set_sampling(u, pl, nk0, nit, area ^ 1u);
                pl.samp_begin = sp[l];
                pl.samp_end = sp[l + 1];
            }
            MCS_HIP(mcs::launch_pass(q.stream, pl, q.d_words, ones, u->geo));
        }
        q.primed = prefetch ? int64_t(nk0) : -1;
        q.primed_iter = nit;
        ++q.seq;
        return MCS_OK;
    };
    std::vector<uint32_t> next(u->pipes.size());
    for (uint32_t p = 0; p < n_passes; ++p) {
        if (u->iteration == 0xFFFFFFFFu) return fail(MCS_ERR_STATE, "iteration counter exhausted");
        const uint32_t it = u->iteration + 1;
        // Groups are dealt to the pipelines in turn, one group each, so every
        // stream is fed from the start: submitting one pipeline's whole run
        // before the next would block the host on that stream's full queue
        // while the other pipeline sat idle. Each pipeline's own sequence of
        // launches is the same either way.
        for (size_t i = 0; i < u->pipes.size(); ++i) next[i] = u->pipes[i].k_begin;
        for (bool any = true; any;) {
            any = false;
            for (size_t i = 0; i < u->pipes.size(); ++i) {
                UaPipe &q = u->pipes[i];
                if (next[i] >= q.k_end) continue;
                const mcs_status st = enqueue_group(q, next[i], it);
                if (st != MCS_OK) return st;
                next[i] += G;
                any = true;
            }
        }
        u->iteration = it;
    }
    // Join: the engine's stream waits for every pipeline.
    for (UaPipe &q : u->pipes) {
        MCS_HIP(hipEventRecord(q.done, q.stream));
        MCS_HIP(hipStreamWaitEvent(u->stream, q.done, 0));
    }
    return MCS_OK;
}

}  // namespace

extern "C" {

void mcs_ua_config_default(mcs_ua_config *cfg) {
    if (!cfg) return;
    cfg->key0 = MCS_REFERENCE_KEY0;
    cfg->key1 = MCS_REFERENCE_KEY1;
    cfg->common_random_numbers = 0;
    cfg->samples_per_launch = 0;
    cfg->pipelines = 0;
    cfg->chunk_words = 0;
}

mcs_status mcs_ua_create(const mcs_pdag *pdag, const mcs_sample_shape *shape, int32_t n_samples,
                         const double *p_samples, const uint32_t *sample_ids,
                         const int32_t *tally_index, int32_t n_tally, const mcs_ua_config *cfg,
                         mcs_ua **out) {
    if (!pdag || !shape || !p_samples || !out || n_samples <= 0 || n_tally < 0 ||
        (n_tally > 0 && !tally_index))
        return fail(MCS_ERR_INVALID_ARGUMENT, "null argument or no samples");
    *out = nullptr;
    auto u = std::make_unique<mcs_ua>();
    mcs_ua_config_default(&u->cfg);
    if (cfg) u->cfg = *cfg;
    const uint64_t n_bp = shape->batch_size * shape->bitpacks_per_batch;
    if (shape->batch_size == 0 || shape->bitpacks_per_batch == 0 ||
        n_bp / shape->batch_size != shape->bitpacks_per_batch || n_bp > (uint64_t(1) << 32))
        return fail(MCS_ERR_INVALID_ARGUMENT, "sample shape must be positive and at most 2^32 words");
    std::string err;
    if (!mcs::build_plan(*pdag, u->plan, err)) return fail(MCS_ERR_GRAPH, err);
    const mcs::Plan &plan = u->plan;
    u->sch = mcs::build_schedule(plan, kUaSplitFanIn);
    mcs::fuse_single_consumer(u->sch, plan, kUaTaskRows);
    u->n_ev = uint32_t(plan.n_events());
    u->gate_rows = uint32_t(plan.n_gates() + u->sch.n_aux);
    u->n_samples = n_samples;

    // Tallies.
    std::unordered_map<int32_t, int32_t> node_of_index;
    for (int32_t n = 0; n < plan.n_nodes(); ++n) node_of_index[plan.node_pdag_index[n]] = n;
    std::vector<uint8_t> tallied(size_t(plan.n_nodes()), 0);
    std::vector<int32_t> want(tally_index, tally_index + n_tally);
    if (want.empty()) want.push_back(pdag->root);
    for (int32_t idx : want) {
        auto it = node_of_index.find(idx);
        if (it == node_of_index.end())
            return fail(MCS_ERR_GRAPH, "tally index " + std::to_string(idx) + " is not reachable from the root");
        tallied[size_t(it->second)] = 1;
        u->tally_nodes.push_back(it->second);  // duplicates kept: one column each
    }

    // Geometry and launch plan (the rows of a sample-pass).
    mcs::LaunchGeometry &geo = u->geo;
    const uint32_t cw = u->cfg.chunk_words;
    if (cw && cw != 512 && cw != 1024 && cw != 2048)
        return fail(MCS_ERR_INVALID_ARGUMENT, "chunk_words must be 0, 512, 1024 or 2048");
    geo.chunk_words = cw ? cw : pick_chunk_words(n_bp);
    geo.n_chunks = uint32_t((n_bp + geo.chunk_words - 1) / geo.chunk_words);
    geo.stride = uint64_t(geo.n_chunks) * geo.chunk_words;
    geo.n_bp = n_bp;
    const mcs::RowMap rm{{0u, u->n_ev}, 2u * u->n_ev, 2u * u->n_ev + u->gate_rows};
    mcs::build_launch_plan(plan, u->sch, tallied, rm, u->lp);
    u->slots = uint64_t(u->gate_rows) + u->n_ev;
    // Every launch of a group of g samples must fit the grid (its workgroups
    // x 256 lanes below 2^32).
    auto grid_fits = [&](uint32_t g) {
        for (size_t l = 0; l + 1 < u->lp.launch_begin.size(); ++l) {
            mcs::PassLaunch pl{};
            pl.n_gates = u->lp.launch_begin[l + 1] - u->lp.launch_begin[l];
            pl.n_samples = g;
            pl.samp_events = u->n_ev;
            pl.samp_end = uint64_t(g) * u->n_ev * geo.n_chunks;
            if (mcs::pass_grid_blocks(pl, geo) * 256ull >= (uint64_t(1) << 32)) return false;
        }
        return true;
    };
    if (!grid_fits(1)) return fail(MCS_ERR_INVALID_ARGUMENT, "grid too large: reduce words per pass");

    // Event descriptors of every sample, in event_idx order: the column of
    // p_samples each event reads.
    std::unordered_map<int32_t, int32_t> slot_of_var;
    for (int32_t i = 0; i < pdag->n_vars; ++i) slot_of_var[pdag->var_index[i]] = i;
    u->n_vars = pdag->n_vars;
    u->var_of_event.resize(u->n_ev);
    for (uint32_t e = 0; e < u->n_ev; ++e) u->var_of_event[e] = slot_of_var.at(plan.event_pdag_index[e]);
    for (int32_t k = 0; k < n_samples; ++k)
        for (uint32_t e = 0; e < u->n_ev; ++e) {
            const double pe = p_samples[size_t(k) * size_t(pdag->n_vars) + size_t(u->var_of_event[e])];
            if (!(pe == pe)) return fail(MCS_ERR_INVALID_ARGUMENT, "sampled probability is NaN");
        }

    // Device.
    MCS_HIP(hipGetDevice(&u->device));
    MCS_HIP(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
    u->slot_rows = 2ull * u->n_ev + u->gate_rows;
    const uint64_t rows = u->slot_rows;
    const uint64_t slot_sum = rows * mcs::state_stride_for(geo.n_chunks);
    const uint64_t slot_bytes = rows * geo.stride * 8 + slot_sum;  // words and summaries
    const uint64_t ones_n = 2ull * u->n_ev + uint64_t(n_samples) * u->slots;
    size_t free_b = 0, total_b = 0;
    MCS_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t fixed = uint64_t(n_samples) * u->n_ev * sizeof(mcs::EventDesc) + uint64_t(n_samples) * 4 +
                           u->lp.gates.size() * sizeof(mcs::GateDesc) +
                           (u->lp.args[0].size() + u->lp.args[1].size()) * 4 + ones_n * 8;
    const uint64_t reserve = uint64_t(256) << 20;
    // Pipelines (one arena each) and samples per launch (slots per arena):
    // as asked, or two pipelines and enough samples per launch for about
    // 16,384 tasks in an average launch (a large tree's single sample has
    // that many; the 200-event tree at 2^14 words needs 32 samples, which
    // took it from 0.094 to 0.056 ms per sample). Then fewer samples per
    // launch while a launch's grid would not fit or (unless a count was
    // asked for) the arenas would pass 96 GiB, fewer pipelines while they
    // would not fit in free memory, and fewer samples per launch again.
    uint32_t n_pipes = u->cfg.pipelines ? u->cfg.pipelines : kUaPipes;
    n_pipes = std::min<uint32_t>(n_pipes, uint32_t(n_samples));
    const uint64_t n_launch = std::max<uint64_t>(1, u->lp.launch_begin.size() - 1);
    const uint64_t tasks = (uint64_t(u->n_ev) + u->lp.launch_begin.back()) * geo.n_chunks / n_launch;
    uint32_t G = u->cfg.samples_per_launch
                     ? u->cfg.samples_per_launch
                     : uint32_t(std::min<uint64_t>(kUaMaxGroup, (kUaLaunchTasks + tasks - 1) / std::max<uint64_t>(tasks, 1)));
    G = std::max<uint32_t>(1, std::min<uint32_t>(G, uint32_t((uint64_t(n_samples) + n_pipes - 1) / n_pipes)));
    while (G > 1 && !grid_fits(G)) --G;
    if (!u->cfg.samples_per_launch)
        while (G > 1 && uint64_t(n_pipes) * G * slot_bytes > kUaArenaCap) --G;
    auto fits = [&](uint32_t np, uint32_t g) { return fixed + uint64_t(np) * g * slot_bytes + reserve <= free_b; };
    while (n_pipes > 1 && !fits(n_pipes, G) && !u->cfg.pipelines) --n_pipes;
    while (G > 1 && !fits(n_pipes, G)) --G;
    u->group = G;
    const uint64_t sum_bytes = uint64_t(G) * slot_sum;
    const uint64_t word_bytes = uint64_t(G) * slot_bytes;
    const uint64_t need = fixed + n_pipes * word_bytes;
    if (need + (uint64_t(256) << 20) > free_b)
        return fail(MCS_ERR_OUT_OF_MEMORY, "uncertainty engine needs " + std::to_string(need >> 20) +
                                               " MiB, " + std::to_string(free_b >> 20) + " MiB free");
    MCS_HIP(hipEventCreateWithFlags(&u->began, hipEventDisableTiming));
    u->pipes.resize(n_pipes);
    for (uint32_t i = 0; i < n_pipes; ++i) {
        UaPipe &q = u->pipes[i];
        MCS_HIP(hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking));
        MCS_HIP(hipEventCreateWithFlags(&q.done, hipEventDisableTiming));
        MCS_HIP(hipMalloc(&q.d_words, word_bytes - sum_bytes));
        MCS_HIP(hipMalloc(&q.d_sum, sum_bytes));
    }
    MCS_HIP(hipMalloc(&u->d_gates, std::max<size_t>(u->lp.gates.size(), 1) * sizeof(mcs::GateDesc)));
    for (int a = 0; a < 2; ++a)
        MCS_HIP(hipMalloc(&u->d_args[a], std::max<size_t>(u->lp.args[a].size(), 1) * 4));
    if (!u->lp.gates.empty())
        MCS_HIP(hipMemcpyAsync(u->d_gates, u->lp.gates.data(), u->lp.gates.size() * sizeof(mcs::GateDesc),
                               hipMemcpyHostToDevice, u->stream));
    for (int a = 0; a < 2; ++a)
        if (!u->lp.args[a].empty())
            MCS_HIP(hipMemcpyAsync(u->d_args[a], u->lp.args[a].data(), u->lp.args[a].size() * 4,
                                   hipMemcpyHostToDevice, u->stream));
    const mcs_status st = load_samples(u.get(), n_samples, p_samples, sample_ids);
    if (st != MCS_OK) return st;
    u->device_bytes = need;
    // The counters of sample k are at d_ones + 2 n_ev + k * slots; the kernel
    // indexes them by row, and gate rows start at 2 n_ev.
    *out = u.release();
    return MCS_OK;
}

mcs_status mcs_ua_set_samples(mcs_ua *ua, int32_t n_samples, const double *p_samples,
                              const uint32_t *sample_ids) {
    if (!ua || !p_samples || n_samples <= 0)
        return fail(MCS_ERR_INVALID_ARGUMENT, "null argument or no samples");
    for (int32_t k = 0; k < n_samples; ++k)
        for (uint32_t e = 0; e < ua->n_ev; ++e) {
            const double pe = p_samples[size_t(k) * size_t(ua->n_vars) + size_t(ua->var_of_event[e])];
            if (!(pe == pe)) return fail(MCS_ERR_INVALID_ARGUMENT, "sampled probability is NaN");
        }
    MCS_HIP(hipSetDevice(ua->device));
    return load_samples(ua, n_samples, p_samples, sample_ids);
}

mcs_status mcs_ua_info(const mcs_ua *ua, uint32_t *samples_per_launch, uint32_t *pipelines,
                       uint32_t *chunk_words, uint64_t *device_bytes) {
    if (!ua) return fail(MCS_ERR_INVALID_ARGUMENT, "null engine");
    if (samples_per_launch) *samples_per_launch = ua->group;
    if (chunk_words) *chunk_words = ua->geo.chunk_words;
    if (pipelines) *pipelines = uint32_t(ua->pipes.size());
    if (device_bytes) *device_bytes = ua->device_bytes;
    return MCS_OK;
}

mcs_status mcs_ua_destroy(mcs_ua *ua) {
    delete ua;
    return MCS_OK;
}

mcs_status mcs_ua_run(mcs_ua *ua, uint32_t n_passes) {
    if (!ua) return fail(MCS_ERR_INVALID_ARGUMENT, "null engine");
    MCS_HIP(hipSetDevice(ua->device));
    return enqueue_run(ua, n_passes);
}

mcs_status mcs_ua_collect(mcs_ua *ua, uint64_t *ones, int64_t n_out, uint64_t *total_bits) {
    if (!ua || !ones || !total_bits) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    const int64_t T = int64_t(ua->tally_nodes.size());
    if (n_out != int64_t(ua->n_samples) * T)
        return fail(MCS_ERR_INVALID_ARGUMENT, "collect needs n_samples x n_tally counters");
    MCS_HIP(hipSetDevice(ua->device));
    const uint64_t pad = 2ull * ua->n_ev;
    std::vector<unsigned long long> all(size_t(ua->n_samples) * ua->slots);
    MCS_HIP(hipMemcpyAsync(all.data(), ua->d_ones + pad, all.size() * 8, hipMemcpyDeviceToHost,
                           ua->stream));
    MCS_HIP(hipStreamSynchronize(ua->stream));
    for (int32_t k = 0; k < ua->n_samples; ++k) {
        for (int64_t j = 0; j < T; ++j) {
            const int32_t node = ua->tally_nodes[size_t(j)];
            const uint64_t slot = node < int32_t(ua->n_ev) ? uint64_t(ua->gate_rows) + uint64_t(node)
                                                           : uint64_t(node) - ua->n_ev;
            ones[size_t(k) * size_t(T) + size_t(j)] = all[size_t(k) * ua->slots + slot];
        }
    }
    *total_bits = uint64_t(ua->iteration) * ua->geo.n_bp * 64ull;
    return MCS_OK;
}

mcs_status mcs_ua_time_run(mcs_ua *ua, uint32_t n_passes, double *kernel_ms) {
    if (!ua || !kernel_ms) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    MCS_HIP(hipSetDevice(ua->device));
    hipEvent_t a = nullptr, z = nullptr;
    MCS_HIP(hipEventCreate(&a));
    hipError_t e = hipEventCreate(&z);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        MCS_HIP(e);
    }
    e = hipEventRecord(a, ua->stream);
    mcs_status st = e == hipSuccess ? enqueue_run(ua, n_passes) : MCS_OK;
    if (e == hipSuccess && st == MCS_OK) e = hipEventRecord(z, ua->stream);
    if (e == hipSuccess && st == MCS_OK) e = hipEventSynchronize(z);
    float ms = 0;
    if (e == hipSuccess && st == MCS_OK) e = hipEventElapsedTime(&ms, a, z);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    if (st != MCS_OK) return st;
    MCS_HIP(e);
    *kernel_ms = ms;
    return MCS_OK;
}

}  // extern "C"
