// Uncertainty-analysis engine: K parameter samples of one tree as one
// layered pipeline (kernel: pass_layer in kernels.hip).
//
// The samples share the tree's plan, schedule, gate descriptors and
// argument rows; each has its own event descriptors (thresholds) and Philox
// key. The samples are split into up to two pipelines, each with its own
// HIP stream and words arena (two event areas, which its samples alternate
// between, and one gate area, `stride` words per row), so one pipeline's
// short launches and launch tails leave the CUs to the other. Within a
// pipeline, sample k's gate launches carry the sampler tasks of sample k + 1,
// so Philox work and gate traffic overlap across samples; the pipeline's last
// sample of a pass samples its first sample of the next pass.
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

uint32_t pick_chunk_words(uint64_t n_bp) {
    if (n_bp >= (uint64_t(1) << 18)) return 2048;
    if (n_bp >= (uint64_t(1) << 16)) return 1024;
    return 512;
}

struct UaPipe {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;           // recorded after a run's launches
    uint64_t *d_words = nullptr;         // the pipeline's words arena
    uint8_t *d_sum = nullptr;            // chunk summaries of the arena's rows
    uint32_t k_begin = 0, k_end = 0;     // its samples
    uint64_t seq = 0;                    // sample-passes run: event area = seq & 1
    int64_t primed = -1;                 // sample sampled into area seq & 1 ...
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
    std::vector<uint32_t> key0;          // per sample
    std::vector<int32_t> tally_nodes;    // plan node ids
    uint64_t slots = 0;                  // counters per sample (gates, partials, events)
    uint32_t iteration = 0;
    std::vector<UaPipe> pipes;
    hipEvent_t began = nullptr;          // recorded before a run's launches
    mcs::EventDesc *d_events = nullptr;  // n_samples x n_ev
    mcs::GateDesc *d_gates = nullptr;
    uint32_t *d_args[2] = {nullptr, nullptr};
    unsigned long long *d_ones = nullptr;  // pad of 2 n_ev, then n_samples x slots
    uint64_t device_bytes = 0;

    ~mcs_ua() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        for (UaPipe &q : pipes)
            if (q.stream) (void)hipStreamSynchronize(q.stream);
        for (void *p : {static_cast<void *>(d_events), static_cast<void *>(d_gates),
                        static_cast<void *>(d_args[0]), static_cast<void *>(d_args[1]),
                        static_cast<void *>(d_ones)})
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

mcs_status enqueue_run(mcs_ua *u, uint32_t n_passes) {
    const uint64_t samp_total = uint64_t(u->n_ev) * u->geo.n_chunks;
    const std::vector<uint64_t> split = mcs::split_in_proportion(u->lp.launch_cost, samp_total);
    // Fork: the pipelines start after what the engine's stream holds.
    MCS_HIP(hipEventRecord(u->began, u->stream));
    for (UaPipe &q : u->pipes) MCS_HIP(hipStreamWaitEvent(q.stream, u->began, 0));
    for (uint32_t p = 0; p < n_passes; ++p) {
        if (u->iteration == 0xFFFFFFFFu) return fail(MCS_ERR_STATE, "iteration counter exhausted");
        const uint32_t it = u->iteration + 1;
        for (UaPipe &q : u->pipes) {
            for (uint32_t k = q.k_begin; k < q.k_end; ++k) {
                const uint32_t area = uint32_t(q.seq & 1u);
                mcs::PassLaunch base{};
                base.key1 = u->cfg.key1;
                base.state = q.d_sum;
                base.state_stride = mcs::state_stride_for(u->geo.n_chunks);
                if (!(q.primed == int64_t(k) && q.primed_iter == it)) {
                    mcs::PassLaunch pl = base;
                    pl.events = u->d_events + uint64_t(k) * u->n_ev;
                    pl.key0 = u->key0[k];
                    pl.samp_iteration = it;
                    pl.samp_row0 = area * u->n_ev;
                    pl.samp_end = samp_total;
                    MCS_HIP(mcs::launch_pass(q.stream, pl, q.d_words, u->d_ones, u->geo));
                }
                const uint32_t nk = k + 1 < q.k_end ? k + 1 : q.k_begin;
                const uint32_t nit = k + 1 < q.k_end ? it : it + 1;
                const bool prefetch = nit != 0;  // iteration counter not wrapping
                unsigned long long *ones = u->d_ones + uint64_t(k) * u->slots;  // + pad - 2 n_ev
                for (size_t l = 0; l + 1 < u->lp.launch_begin.size(); ++l) {
                    mcs::PassLaunch pl = base;
                    pl.gates = u->d_gates + u->lp.launch_begin[l];
                    pl.tree_gates = u->d_gates;
                    pl.arg_rows = u->d_args[area];
                    pl.n_gates = u->lp.launch_begin[l + 1] - u->lp.launch_begin[l];
                    if (prefetch) {
                        pl.events = u->d_events + uint64_t(nk) * u->n_ev;
                        pl.key0 = u->key0[nk];
                        pl.samp_iteration = nit;
                        pl.samp_row0 = (area ^ 1u) * u->n_ev;
                        pl.samp_begin = split[l];
                        pl.samp_end = split[l + 1];
                    }
                    MCS_HIP(mcs::launch_pass(q.stream, pl, q.d_words, ones, u->geo));
                }
                q.primed = prefetch ? int64_t(nk) : -1;
                q.primed_iter = nit;
                ++q.seq;
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
    geo.chunk_words = pick_chunk_words(n_bp);
    geo.n_chunks = uint32_t((n_bp + geo.chunk_words - 1) / geo.chunk_words);
    geo.stride = uint64_t(geo.n_chunks) * geo.chunk_words;
    geo.n_bp = n_bp;
    const mcs::RowMap rm{{0u, u->n_ev}, 2u * u->n_ev, 2u * u->n_ev + u->gate_rows};
    mcs::build_launch_plan(plan, u->sch, tallied, rm, u->lp);
    u->slots = uint64_t(u->gate_rows) + u->n_ev;
    for (size_t l = 0; l + 1 < u->lp.launch_begin.size(); ++l) {
        mcs::PassLaunch pl{};
        pl.n_gates = u->lp.launch_begin[l + 1] - u->lp.launch_begin[l];
        pl.samp_end = uint64_t(u->n_ev) * geo.n_chunks;
        if (mcs::pass_grid_blocks(pl, geo) * 256ull >= (uint64_t(1) << 32))
            return fail(MCS_ERR_INVALID_ARGUMENT, "grid too large: reduce words per pass");
    }

    // Event descriptors of every sample, in event_idx order.
    std::unordered_map<int32_t, int32_t> slot_of_var;
    for (int32_t i = 0; i < pdag->n_vars; ++i) slot_of_var[pdag->var_index[i]] = i;
    std::vector<int32_t> var_of_event(u->n_ev);
    for (uint32_t e = 0; e < u->n_ev; ++e) var_of_event[e] = slot_of_var.at(plan.event_pdag_index[e]);
    std::vector<mcs::EventDesc> ev(size_t(n_samples) * u->n_ev);
    for (int32_t k = 0; k < n_samples; ++k) {
        const double *p = p_samples + size_t(k) * size_t(pdag->n_vars);
        for (uint32_t e = 0; e < u->n_ev; ++e) {
            const int32_t vi = var_of_event[e];
            double pe = p[vi];
            if (!(pe == pe)) return fail(MCS_ERR_INVALID_ARGUMENT, "sampled probability is NaN");
            pe = pe > 1.0 ? 1.0 : (pe < 0.0 ? 0.0 : pe);
            ev[size_t(k) * u->n_ev + e] = mcs::make_event_desc(plan.event_pdag_index[e], e, pe);
        }
    }
    u->key0.resize(size_t(n_samples));
    for (int32_t k = 0; k < n_samples; ++k) {
        const uint32_t id = sample_ids ? sample_ids[k] : uint32_t(k);
        u->key0[size_t(k)] = u->cfg.common_random_numbers ? u->cfg.key0 : u->cfg.key0 + id;
    }

    // Device.
    MCS_HIP(hipGetDevice(&u->device));
    MCS_HIP(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
    const uint64_t rows = 2ull * u->n_ev + u->gate_rows;
    const uint64_t sum_bytes = rows * mcs::state_stride_for(geo.n_chunks);
    const uint64_t word_bytes = rows * geo.stride * 8 + sum_bytes;  // words and summaries
    const uint64_t ones_n = 2ull * u->n_ev + uint64_t(n_samples) * u->slots;
    size_t free_b = 0, total_b = 0;
    MCS_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t fixed = ev.size() * sizeof(mcs::EventDesc) + u->lp.gates.size() * sizeof(mcs::GateDesc) +
                           (u->lp.args[0].size() + u->lp.args[1].size()) * 4 + ones_n * 8;
    // Two pipelines (one arena each) when both fit, else one.
    uint32_t n_pipes = std::min<uint32_t>(kUaPipes, uint32_t(n_samples));
    while (n_pipes > 1 && fixed + n_pipes * word_bytes + (uint64_t(256) << 20) > free_b) --n_pipes;
    const uint64_t need = fixed + n_pipes * word_bytes;
    if (need + (uint64_t(256) << 20) > free_b)
        return fail(MCS_ERR_OUT_OF_MEMORY, "uncertainty engine needs " + std::to_string(need >> 20) +
                                               " MiB, " + std::to_string(free_b >> 20) + " MiB free");
    MCS_HIP(hipEventCreateWithFlags(&u->began, hipEventDisableTiming));
    u->pipes.resize(n_pipes);
    for (uint32_t i = 0; i < n_pipes; ++i) {
        UaPipe &q = u->pipes[i];
        q.k_begin = uint32_t(uint64_t(n_samples) * i / n_pipes);
        q.k_end = uint32_t(uint64_t(n_samples) * (i + 1) / n_pipes);
        MCS_HIP(hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking));
        MCS_HIP(hipEventCreateWithFlags(&q.done, hipEventDisableTiming));
        MCS_HIP(hipMalloc(&q.d_words, word_bytes - sum_bytes));
        MCS_HIP(hipMalloc(&q.d_sum, sum_bytes));
        MCS_HIP(hipMemsetAsync(q.d_sum, mcs::kChunkMixed, sum_bytes, u->stream));
    }
    MCS_HIP(hipMalloc(&u->d_events, ev.size() * sizeof(mcs::EventDesc)));
    MCS_HIP(hipMalloc(&u->d_gates, std::max<size_t>(u->lp.gates.size(), 1) * sizeof(mcs::GateDesc)));
    for (int a = 0; a < 2; ++a)
        MCS_HIP(hipMalloc(&u->d_args[a], std::max<size_t>(u->lp.args[a].size(), 1) * 4));
    MCS_HIP(hipMalloc(&u->d_ones, ones_n * 8));
    MCS_HIP(hipMemcpyAsync(u->d_events, ev.data(), ev.size() * sizeof(mcs::EventDesc),
                           hipMemcpyHostToDevice, u->stream));
    if (!u->lp.gates.empty())
        MCS_HIP(hipMemcpyAsync(u->d_gates, u->lp.gates.data(), u->lp.gates.size() * sizeof(mcs::GateDesc),
                               hipMemcpyHostToDevice, u->stream));
    for (int a = 0; a < 2; ++a)
        if (!u->lp.args[a].empty())
            MCS_HIP(hipMemcpyAsync(u->d_args[a], u->lp.args[a].data(), u->lp.args[a].size() * 4,
                                   hipMemcpyHostToDevice, u->stream));
    MCS_HIP(hipMemsetAsync(u->d_ones, 0, ones_n * 8, u->stream));
    MCS_HIP(hipStreamSynchronize(u->stream));
    u->device_bytes = need;
    // The counters of sample k are at d_ones + 2 n_ev + k * slots; the kernel
    // indexes them by row, and gate rows start at 2 n_ev.
    *out = u.release();
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
