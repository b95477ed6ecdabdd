// Engine: device memory, pass scheduling, tallies and the C ABI.
//
// One engine = one compiled PDAG on one device with one HIP stream. Device
// memory: a words matrix (rows = nodes, see kernels.h), event and gate
// descriptors, argument rows, and one uint64 one-bit counter per row.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "layering.h"
#include "mcscram_amd.h"

namespace {

thread_local std::string g_error;

mcs_status fail(mcs_status code, std::string msg) {
    g_error = std::move(msg);
    return code;
}

#define MCS_HIP(expr)                                                                   \
    do {                                                                                \
        const hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? MCS_ERR_OUT_OF_MEMORY : MCS_ERR_DEVICE, \
                        std::string(#expr " failed: ") + hipGetErrorString(e_));        \
    } while (0)

}  // namespace

namespace mcs {
void set_error(const char *msg) { g_error = msg; }
}  // namespace mcs

struct mcs_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t key0 = MCS_REFERENCE_KEY0, key1 = MCS_REFERENCE_KEY1;
    uint32_t iteration = 0;  // passes submitted (reference iterable_queueable counter)
    mcs::Plan plan;
    mcs::LaunchGeometry geo{};
    std::vector<int32_t> tally_nodes;  // node id per tally, first-appearance order
    std::vector<int32_t> tally_index;  // pdag index per tally
    std::unordered_map<int32_t, int32_t> node_of_index;
    uint64_t n_args = 0;
    uint64_t device_bytes = 0;
    // device
    uint64_t *d_words = nullptr;
    mcs::EventDesc *d_events = nullptr;
    mcs::GateDesc *d_gates = nullptr;
    uint32_t *d_args = nullptr;
    unsigned long long *d_ones = nullptr;

    ~mcs_engine() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        if (d_words) (void)hipFree(d_words);
        if (d_events) (void)hipFree(d_events);
        if (d_gates) (void)hipFree(d_gates);
        if (d_args) (void)hipFree(d_args);
        if (d_ones) (void)hipFree(d_ones);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Words per workgroup task: 512 (256 lanes x 2 words) times a repeat count
// that keeps each grid large enough to fill 256 CUs but amortises per-task
// set-up on long rows.
uint32_t pick_chunk_words(uint64_t n_bp) {
    if (n_bp >= (uint64_t(1) << 18)) return 2048;
    if (n_bp >= (uint64_t(1) << 16)) return 1024;
    return 512;
}

mcs_status build_engine(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                        const int32_t *tally_index, int32_t n_tally, uint32_t key0,
                        uint32_t key1, int32_t device, mcs_engine **out) {
    if (!pdag || !shape || !out || (n_tally > 0 && !tally_index) || n_tally < 0)
        return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint64_t n_bp = shape->batch_size * shape->bitpacks_per_batch;
    if (shape->batch_size == 0 || shape->bitpacks_per_batch == 0 ||
        n_bp / shape->batch_size != shape->bitpacks_per_batch)
        return fail(MCS_ERR_INVALID_ARGUMENT, "sample shape must be positive");
    if (n_bp > (uint64_t(1) << 32))
        return fail(MCS_ERR_INVALID_ARGUMENT,
                    "more than 2^32 words per pass (the sample index is 32-bit)");

    auto eng = std::make_unique<mcs_engine>();
    std::string err;
    if (!mcs::build_plan(*pdag, eng->plan, err)) return fail(MCS_ERR_GRAPH, err);
    const mcs::Plan &plan = eng->plan;
    for (int32_t n = 0; n < plan.n_nodes(); ++n) eng->node_of_index[plan.node_pdag_index[n]] = n;

    std::vector<uint8_t> tallied(plan.n_nodes(), 0);
    for (int32_t t = 0; t < n_tally; ++t) {
        auto it = eng->node_of_index.find(tally_index[t]);
        if (it == eng->node_of_index.end())
            return fail(MCS_ERR_GRAPH, "tally index " + std::to_string(tally_index[t]) +
                                           " is not a node reachable from the root");
        if (tallied[it->second]) continue;
        tallied[it->second] = 1;
        eng->tally_nodes.push_back(it->second);
        eng->tally_index.push_back(tally_index[t]);
    }

    // Geometry.
    mcs::LaunchGeometry geo{};
    geo.chunk_words = pick_chunk_words(n_bp);
    geo.n_chunks = uint32_t((n_bp + geo.chunk_words - 1) / geo.chunk_words);
    geo.stride = uint64_t(geo.n_chunks) * geo.chunk_words;
    geo.n_bp = n_bp;
    uint32_t max_layer = 0;
    for (int32_t l = 0; l < plan.n_gate_layers(); ++l)
        max_layer = std::max(max_layer, plan.layer_begin[l + 1] - plan.layer_begin[l]);
    const uint64_t max_rows = std::max<uint64_t>(uint64_t(plan.n_events()), max_layer);
    if (max_rows * geo.n_chunks * 256ull >= (uint64_t(1) << 32))
        return fail(MCS_ERR_INVALID_ARGUMENT, "grid too large: reduce words per pass");
    eng->geo = geo;

    // Descriptors.
    std::vector<mcs::EventDesc> ev(plan.n_events());
    for (int32_t e = 0; e < plan.n_events(); ++e) {
        const double p = plan.event_p[e];
        const uint64_t thr = static_cast<uint64_t>(p * 4294967296.0);
        uint32_t mode = mcs::kEventDraw;
        if (thr == 0) mode = mcs::kEventNever;
        if (thr >= (uint64_t(1) << 32)) mode = mcs::kEventAlways;
        ev[e].ctr0 = uint32_t(plan.event_pdag_index[e]) + 1u;
        ev[e].ctr1 = uint32_t(e) + 1u;
        ev[e].threshold = uint32_t(std::min<uint64_t>(thr, 0xFFFFFFFFull));
        ev[e].mode_tally = mode | (uint32_t(tallied[e]) << 2);
    }
    std::vector<mcs::GateDesc> gd(plan.n_gates());
    std::vector<uint32_t> arg_rows;
    for (int32_t g = 0; g < plan.n_gates(); ++g) {
        const mcs::PlannedGate &pg = plan.gates[g];
        const int32_t node = plan.n_events() + g;
        gd[g].out_row = uint32_t(node);
        gd[g].arg_begin = uint32_t(arg_rows.size());
        const uint32_t n_in = uint32_t(pg.pos_in.size() + pg.neg_in.size());
        gd[g].counts = n_in | uint32_t(pg.pos_in.size()) << 16;
        gd[g].type_k_tally = uint32_t(pg.type) | (uint32_t(uint8_t(pg.k)) << 8) |
                             (uint32_t(tallied[node]) << 16);
        for (int32_t i : pg.pos_in) arg_rows.push_back(uint32_t(i));
        for (int32_t i : pg.neg_in) arg_rows.push_back(uint32_t(i));
    }
    eng->n_args = arg_rows.size();

    // Device allocation.
    eng->device = device;
    if (device < 0) MCS_HIP(hipGetDevice(&eng->device));
    MCS_HIP(hipSetDevice(eng->device));
    MCS_HIP(hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking));
    const uint64_t rows = uint64_t(plan.n_nodes());
    const uint64_t word_bytes = rows * geo.stride * sizeof(uint64_t);
    size_t free_b = 0, total_b = 0;
    MCS_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t need = word_bytes + ev.size() * sizeof(mcs::EventDesc) +
                          gd.size() * sizeof(mcs::GateDesc) + arg_rows.size() * 4 + rows * 8;
    if (need + (uint64_t(256) << 20) > free_b)
        return fail(MCS_ERR_OUT_OF_MEMORY,
                    "engine needs " + std::to_string(need >> 20) + " MiB of device memory, " +
                        std::to_string(free_b >> 20) + " MiB free: reduce words per pass");
    MCS_HIP(hipMalloc(&eng->d_words, word_bytes));
    MCS_HIP(hipMalloc(&eng->d_ones, rows * sizeof(unsigned long long)));
    if (!ev.empty()) MCS_HIP(hipMalloc(&eng->d_events, ev.size() * sizeof(mcs::EventDesc)));
    MCS_HIP(hipMalloc(&eng->d_gates, gd.size() * sizeof(mcs::GateDesc)));
    MCS_HIP(hipMalloc(&eng->d_args, std::max<size_t>(arg_rows.size(), 1) * 4));
    eng->device_bytes = need;
    if (!ev.empty())
        MCS_HIP(hipMemcpyAsync(eng->d_events, ev.data(), ev.size() * sizeof(mcs::EventDesc),
                               hipMemcpyHostToDevice, eng->stream));
    MCS_HIP(hipMemcpyAsync(eng->d_gates, gd.data(), gd.size() * sizeof(mcs::GateDesc),
                           hipMemcpyHostToDevice, eng->stream));
    if (!arg_rows.empty())
        MCS_HIP(hipMemcpyAsync(eng->d_args, arg_rows.data(), arg_rows.size() * 4,
                               hipMemcpyHostToDevice, eng->stream));
    MCS_HIP(hipMemsetAsync(eng->d_ones, 0, rows * sizeof(unsigned long long), eng->stream));
    MCS_HIP(hipStreamSynchronize(eng->stream));
    eng->key0 = key0;
    eng->key1 = key1;
    *out = eng.release();
    return MCS_OK;
}

mcs_status enqueue_passes(mcs_engine *e, uint32_t n) {
    MCS_HIP(hipSetDevice(e->device));
    const mcs::Plan &plan = e->plan;
    for (uint32_t i = 0; i < n; ++i) {
        if (e->iteration == 0xFFFFFFFFu) return fail(MCS_ERR_STATE, "iteration counter exhausted");
        const uint32_t iteration = ++e->iteration;
        MCS_HIP(mcs::launch_sample_events(e->stream, e->d_events, e->d_words, e->d_ones, e->geo,
                                          uint32_t(plan.n_events()), iteration, e->key0, e->key1));
        for (int32_t l = 0; l < plan.n_gate_layers(); ++l) {
            const uint32_t b = plan.layer_begin[l], c = plan.layer_begin[l + 1] - b;
            MCS_HIP(mcs::launch_eval_gates(e->stream, e->d_gates + b, e->d_args, e->d_words,
                                           e->d_ones, e->geo, c));
        }
    }
    return MCS_OK;
}

void fill_estimates(mcs_tally &t) {
    if (t.total_bits == 0) {
        t.mean = t.std_err = 0.0;
        for (double &c : t.ci) c = 0.0;
        return;
    }
    // Normal-approximation interval of src/mc/stats/ci_utils.h:167-201.
    const double n = double(t.total_bits);
    const double p = double(t.num_one_bits) / n;
    t.mean = p;
    t.std_err = std::sqrt(p * (1.0 - p) / n);
    const double z95 = 1.959963984540054, z99 = 2.5758293035489004;
    t.ci[0] = std::clamp(p - t.std_err * z95, 0.0, 1.0);
    t.ci[1] = std::clamp(p + t.std_err * z95, 0.0, 1.0);
    t.ci[2] = std::clamp(p - t.std_err * z99, 0.0, 1.0);
    t.ci[3] = std::clamp(p + t.std_err * z99, 0.0, 1.0);
}

mcs_status collect(mcs_engine *e, mcs_tally *out, int32_t n_out) {
    if (!out || n_out != int32_t(e->tally_nodes.size()))
        return fail(MCS_ERR_INVALID_ARGUMENT,
                    "collect needs exactly " + std::to_string(e->tally_nodes.size()) + " records");
    MCS_HIP(hipSetDevice(e->device));
    std::vector<unsigned long long> ones(e->plan.n_nodes());
    MCS_HIP(hipMemcpyAsync(ones.data(), e->d_ones, ones.size() * 8, hipMemcpyDeviceToHost, e->stream));
    MCS_HIP(hipStreamSynchronize(e->stream));
    for (size_t t = 0; t < e->tally_nodes.size(); ++t) {
        mcs_tally &r = out[t];
        r.index = e->tally_index[t];
        r.num_one_bits = ones[e->tally_nodes[t]];
        // total = iteration * N_bp * 64 (src/mc/kernel/tally.h:345-352).
        r.total_bits = uint64_t(e->iteration) * e->geo.n_bp * 64ull;
        fill_estimates(r);
    }
    return MCS_OK;
}

}  // namespace

extern "C" {

const char *mcs_last_error(void) { return g_error.c_str(); }

mcs_status mcs_engine_create(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                             const int32_t *tally_index, int32_t n_tally, mcs_engine **out) {
    return build_engine(pdag, shape, tally_index, n_tally, MCS_REFERENCE_KEY0,
                        MCS_REFERENCE_KEY1, -1, out);
}

mcs_status mcs_engine_create_ex(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                                const int32_t *tally_index, int32_t n_tally, uint32_t key0,
                                uint32_t key1, int32_t device, mcs_engine **out) {
    return build_engine(pdag, shape, tally_index, n_tally, key0, key1, device, out);
}

mcs_status mcs_engine_destroy(mcs_engine *engine) {
    delete engine;
    return MCS_OK;
}

mcs_status mcs_engine_pass(mcs_engine *engine, uint32_t n_passes) {
    if (!engine) return fail(MCS_ERR_INVALID_ARGUMENT, "null engine");
    return enqueue_passes(engine, n_passes);
}

mcs_status mcs_engine_wait(mcs_engine *engine) {
    if (!engine) return fail(MCS_ERR_INVALID_ARGUMENT, "null engine");
    MCS_HIP(hipSetDevice(engine->device));
    MCS_HIP(hipStreamSynchronize(engine->stream));
    return MCS_OK;
}

mcs_status mcs_engine_pass_wait_collect(mcs_engine *engine, uint32_t total_passes,
                                        uint32_t passes_between_waits, mcs_tally *out,
                                        int32_t n_out) {
    if (!engine) return fail(MCS_ERR_INVALID_ARGUMENT, "null engine");
    if (total_passes == 0)
        return fail(MCS_ERR_INVALID_ARGUMENT, "pass_wait_collect: total_passes must be greater than 0");
    const uint32_t interval = passes_between_waits == 0 ? total_passes : passes_between_waits;
    uint32_t remaining = total_passes;
    while (remaining > 0) {
        const uint32_t batch = std::min(interval, remaining);
        mcs_status st = enqueue_passes(engine, batch);
        if (st != MCS_OK) return st;
        st = mcs_engine_wait(engine);
        if (st != MCS_OK) return st;
        remaining -= batch;
    }
    return collect(engine, out, n_out);
}

mcs_status mcs_engine_collect(mcs_engine *engine, mcs_tally *out, int32_t n_out) {
    if (!engine) return fail(MCS_ERR_INVALID_ARGUMENT, "null engine");
    return collect(engine, out, n_out);
}

mcs_status mcs_engine_get_info(mcs_engine *engine, mcs_engine_info *out) {
    if (!engine || !out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    out->num_bitpacks = engine->geo.n_bp;
    out->row_stride = engine->geo.stride;
    out->iteration = engine->iteration;
    out->num_events = engine->plan.n_events();
    out->num_gates = engine->plan.n_gates();
    out->num_layers = engine->plan.n_gate_layers() + 1;
    out->num_tallies = int32_t(engine->tally_nodes.size());
    out->num_args = engine->n_args;
    out->device_bytes = engine->device_bytes;
    return MCS_OK;
}

mcs_status mcs_engine_event_order(mcs_engine *engine, int32_t *pdag_index_out, int32_t n_out) {
    if (!engine || (!pdag_index_out && n_out > 0)) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n_out < engine->plan.n_events()) return fail(MCS_ERR_INVALID_ARGUMENT, "output too small");
    std::copy(engine->plan.event_pdag_index.begin(), engine->plan.event_pdag_index.end(),
              pdag_index_out);
    return MCS_OK;
}

mcs_status mcs_engine_read_words(mcs_engine *engine, int32_t pdag_index, uint64_t *out,
                                 uint64_t n_out) {
    if (!engine || !out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    auto it = engine->node_of_index.find(pdag_index);
    if (it == engine->node_of_index.end())
        return fail(MCS_ERR_GRAPH, "unknown node " + std::to_string(pdag_index));
    if (n_out < engine->geo.n_bp) return fail(MCS_ERR_INVALID_ARGUMENT, "output too small");
    MCS_HIP(hipSetDevice(engine->device));
    MCS_HIP(hipMemcpyAsync(out, engine->d_words + uint64_t(it->second) * engine->geo.stride,
                           engine->geo.n_bp * 8, hipMemcpyDeviceToHost, engine->stream));
    MCS_HIP(hipStreamSynchronize(engine->stream));
    return MCS_OK;
}

mcs_status mcs_engine_time_passes(mcs_engine *engine, uint32_t n_passes, double *sampler_ms,
                                  double *gates_ms, double *pass_ms) {
    if (!engine || !sampler_ms || !gates_ms || !pass_ms)
        return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n_passes == 0) return fail(MCS_ERR_INVALID_ARGUMENT, "n_passes must be positive");
    MCS_HIP(hipSetDevice(engine->device));
    const mcs::Plan &plan = engine->plan;
    // Three events per pass: before the sampler, between sampler and gates,
    // after the gates.
    std::vector<hipEvent_t> ev(size_t(n_passes) * 3, nullptr);
    auto release = [&]() {
        for (hipEvent_t x : ev)
            if (x) (void)hipEventDestroy(x);
    };
    for (hipEvent_t &x : ev) {
        const hipError_t e = hipEventCreate(&x);
        if (e != hipSuccess) {
            release();
            return fail(MCS_ERR_DEVICE, std::string("hipEventCreate failed: ") + hipGetErrorString(e));
        }
    }
    mcs_status st = MCS_OK;
    for (uint32_t p = 0; p < n_passes && st == MCS_OK; ++p) {
        if (engine->iteration == 0xFFFFFFFFu) {
            st = fail(MCS_ERR_STATE, "iteration counter exhausted");
            break;
        }
        const uint32_t iteration = ++engine->iteration;
        hipError_t e = hipEventRecord(ev[3 * p], engine->stream);
        if (e == hipSuccess)
            e = mcs::launch_sample_events(engine->stream, engine->d_events, engine->d_words,
                                          engine->d_ones, engine->geo, uint32_t(plan.n_events()),
                                          iteration, engine->key0, engine->key1);
        if (e == hipSuccess) e = hipEventRecord(ev[3 * p + 1], engine->stream);
        for (int32_t l = 0; l < plan.n_gate_layers() && e == hipSuccess; ++l) {
            const uint32_t b = plan.layer_begin[l], c = plan.layer_begin[l + 1] - b;
            e = mcs::launch_eval_gates(engine->stream, engine->d_gates + b, engine->d_args,
                                       engine->d_words, engine->d_ones, engine->geo, c);
        }
        if (e == hipSuccess) e = hipEventRecord(ev[3 * p + 2], engine->stream);
        if (e != hipSuccess) st = fail(MCS_ERR_DEVICE, std::string("pass launch failed: ") + hipGetErrorString(e));
    }
    if (st == MCS_OK) {
        const hipError_t e = hipStreamSynchronize(engine->stream);
        if (e != hipSuccess) st = fail(MCS_ERR_DEVICE, std::string("hipStreamSynchronize failed: ") + hipGetErrorString(e));
    }
    if (st == MCS_OK) {
        double s = 0, g = 0;
        float ms = 0;
        for (uint32_t p = 0; p < n_passes; ++p) {
            (void)hipEventElapsedTime(&ms, ev[3 * p], ev[3 * p + 1]);
            s += ms;
            (void)hipEventElapsedTime(&ms, ev[3 * p + 1], ev[3 * p + 2]);
            g += ms;
        }
        (void)hipEventElapsedTime(&ms, ev[0], ev[3 * (n_passes - 1) + 2]);
        *sampler_ms = s;
        *gates_ms = g;
        *pass_ms = ms;
    }
    release();
    return st;
}

mcs_status mcs_engine_restore(mcs_engine *engine, uint32_t iteration, const uint64_t *num_one_bits,
                              int32_t n) {
    if (!engine || (n > 0 && !num_one_bits)) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n != int32_t(engine->tally_nodes.size()))
        return fail(MCS_ERR_INVALID_ARGUMENT, "restore needs one count per tally");
    MCS_HIP(hipSetDevice(engine->device));
    std::vector<unsigned long long> ones(engine->plan.n_nodes(), 0);
    for (int32_t t = 0; t < n; ++t) ones[engine->tally_nodes[t]] = num_one_bits[t];
    MCS_HIP(hipStreamSynchronize(engine->stream));
    MCS_HIP(hipMemcpy(engine->d_ones, ones.data(), ones.size() * 8, hipMemcpyHostToDevice));
    engine->iteration = iteration;
    return MCS_OK;
}

mcs_status mcs_pdag_event_order(const mcs_pdag *pdag, int32_t *pdag_index_out, int32_t n_out,
                                int32_t *n_events) {
    if (!pdag || !n_events) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    mcs::Plan plan;
    std::string err;
    if (!mcs::build_plan(*pdag, plan, err)) return fail(MCS_ERR_GRAPH, err);
    *n_events = plan.n_events();
    if (!pdag_index_out) return MCS_OK;
    if (n_out < plan.n_events()) return fail(MCS_ERR_INVALID_ARGUMENT, "output too small");
    std::copy(plan.event_pdag_index.begin(), plan.event_pdag_index.end(), pdag_index_out);
    return MCS_OK;
}

mcs_status mcs_pdag_gate_depths(const mcs_pdag *pdag, int32_t *depth_out, int32_t n_out) {
    if (!pdag || !depth_out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n_out < pdag->n_gates) return fail(MCS_ERR_INVALID_ARGUMENT, "output too small");
    mcs::Plan plan;
    std::string err;
    if (!mcs::build_plan(*pdag, plan, err)) return fail(MCS_ERR_GRAPH, err);
    std::unordered_map<int32_t, int32_t> layer_of;
    for (const mcs::PlannedGate &g : plan.gates) layer_of[g.pdag_index] = g.layer;
    for (int32_t i = 0; i < pdag->n_gates; ++i) {
        auto it = layer_of.find(pdag->gate_index[i]);
        depth_out[i] = it == layer_of.end() ? -1 : it->second;
    }
    return MCS_OK;
}

void mcs_settings_default(mcs_settings *s) {
    if (!s) return;
    s->num_trials = 0;
    s->burn_in_trials = uint64_t(1) << 20;
    s->ci_confidence = 0.99;
    s->ci_rel_margin = 0.001;
    s->max_passes = 0;
    s->key0 = MCS_REFERENCE_KEY0;
    s->key1 = MCS_REFERENCE_KEY1;
}

mcs_status mcs_plan_passes(uint64_t num_trials, const mcs_sample_shape *shape,
                           uint64_t *rounded_trials, uint64_t *passes) {
    if (!shape || !rounded_trials || !passes) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    const uint64_t n_bp = shape->batch_size * shape->bitpacks_per_batch;
    if (n_bp == 0) return fail(MCS_ERR_INVALID_ARGUMENT, "sample shape must be positive");
    // Nearest multiple of 64, ties down (src/mc/queue/sample_shaper.h:73-83).
    const uint64_t rem = num_trials % 64;
    uint64_t t = num_trials;
    if (rem != 0) t = rem <= 32 ? num_trials - rem : num_trials + 64 - rem;
    *rounded_trials = t;
    const uint64_t per = n_bp * 64;
    *passes = (t + per - 1) / per;
    return MCS_OK;
}

}  // extern "C"
