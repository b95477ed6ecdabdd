// Engine: device memory, pass scheduling, tallies and the C ABI.
//
// One engine = one compiled PDAG on one device with one HIP stream. Device
// memory: the words matrix (layout in kernels.h), event and gate
// descriptors, the argument rows of each event buffer, and one uint64
// one-bit counter per row.
//
// A pass is one launch of pass_layer per gate layer. With two event buffers
// the launches of pass i also sample pass i + 1 into the other buffer, a
// slice per launch in proportion to the layer's gate traffic, so the Philox
// work (VALU-bound) runs beside the gate evaluation (HBM-bound). Tallies of
// basic events are tally-only tasks in the first launch of their pass.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "launch_plan.h"
#include "layering.h"
#include "mcscram_amd.h"
#include "status.h"

namespace {
thread_local std::string g_error;
}  // namespace

namespace mcs {
mcs_status fail(mcs_status code, std::string msg) {
    g_error = std::move(msg);
    return code;
}
void set_error(const char *msg) { g_error = msg; }
}  // namespace mcs

using mcs::fail;

struct mcs_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_memory = true;  // false: buffers carved from caller-provided memory
    bool owns_stream = true;
    uint32_t key0 = MCS_REFERENCE_KEY0, key1 = MCS_REFERENCE_KEY1;
    uint32_t iteration = 0;  // passes submitted (reference iterable_queueable counter)
    uint32_t primed = 0;     // iteration already sampled into its buffer (0 = none)
    mcs::Plan plan;
    mcs::LaunchGeometry geo{};
    uint32_t n_buf = 1;                     // event buffers
    std::vector<uint32_t> launch_gate_begin;// per launch: first gate desc (size L+1)
    std::vector<mcs::GateDesc> fused_desc;  // by node id: descriptor of a fused gate
    std::vector<uint8_t> fused;             // by node id: 1 if evaluated inside its consumer
    std::vector<uint64_t> launch_samp_begin;// per launch: first sampler task (size L+1)
    std::vector<int32_t> tally_nodes;       // node id per tally, first-appearance order
    std::vector<int32_t> tally_index;       // pdag index per tally
    std::unordered_map<int32_t, int32_t> node_of_index;
    uint64_t n_args = 0;
    uint64_t n_rows = 0;
    uint64_t arg_span = 0;                  // entries per argument-row array
    uint64_t device_bytes = 0;
    // device
    uint64_t *d_words = nullptr;
    mcs::EventDesc *d_events = nullptr;
    mcs::GateDesc *d_gates = nullptr;
    mcs::GateDesc *d_probe = nullptr;       // one descriptor: a fused gate stored on demand
    mcs::GateDesc probe_host{};             // its host copy (source of the async upload)
    uint32_t *d_args = nullptr;             // n_buf arrays of arg_span entries
    unsigned long long *d_ones = nullptr;
    uint8_t *d_state = nullptr;             // chunk summaries [n_rows][state_stride]
    uint32_t state_stride = 0;

    uint64_t gate_row(int32_t node) const {
        return uint64_t(n_buf) * plan.n_events() + uint64_t(node - plan.n_events());
    }
    // Counter slot of a tallied node: event id for events, row for gates.
    uint64_t tally_slot(int32_t node) const {
        return node < plan.n_events() ? uint64_t(node) : gate_row(node);
    }
    uint32_t parity(uint32_t it) const { return n_buf == 2 ? (it & 1u) : 0u; }
    uint32_t n_launches() const { return uint32_t(launch_gate_begin.size() - 1); }

    ~mcs_engine() {
        if (stream || !owns_stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        if (owns_memory) {
            if (d_words) (void)hipFree(d_words);
            if (d_events) (void)hipFree(d_events);
            if (d_gates) (void)hipFree(d_gates);
            if (d_probe) (void)hipFree(d_probe);
            if (d_args) (void)hipFree(d_args);
            if (d_ones) (void)hipFree(d_ones);
            if (d_state) (void)hipFree(d_state);
        }
        if (stream && owns_stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Wide AND/OR/XOR family gates are split into partial folds of at most
// kSplitFanIn inputs; single-consumer gates are then fused into their
// consumers' tasks up to kTaskRows rows read per task. (A narrower split
// leaves more room in the partial folds for fused gates: at 16 inputs the
// large synthetic tree fuses 13,800 gates instead of 6,700 and moves 10 %
// fewer bytes, but its pass is 2 % slower, being bound by the sampler's
// VALU work rather than by HBM.)
constexpr uint32_t kSplitFanIn = 64;
constexpr uint32_t kTaskRows = 64;

// Words per workgroup task: 512 (256 lanes x 2 words) times 1, 2 or 4,
// keeping grids far above the 256 CUs while amortising per-task set-up on
// long rows.
uint32_t pick_chunk_words(uint64_t n_bp) {
    if (n_bp >= (uint64_t(1) << 18)) return 2048;
    if (n_bp >= (uint64_t(1) << 16)) return 1024;
    return 512;
}

// Bump allocator over caller-provided memory: 256-byte aligned pieces.
struct Carver {
    uint8_t *base;
    uint64_t bytes, used = 0;
    template <class T>
    T *take(uint64_t n) {
        used = (used + 255) & ~uint64_t(255);
        T *p = base ? reinterpret_cast<T *>(base + used) : nullptr;
        used += std::max<uint64_t>(n, 1) * sizeof(T);
        return p;
    }
};

mcs_status build_engine(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                        const int32_t *tally_index, int32_t n_tally, uint32_t key0,
                        uint32_t key1, int32_t device, const mcs::ExternalMemory *ext,
                        mcs_engine **out) {
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
    const int32_t n_ev = plan.n_events();
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
    eng->geo = geo;

    // Device and memory budget: two event buffers when they fit.
    eng->device = device;
    if (device < 0) MCS_HIP(hipGetDevice(&eng->device));
    MCS_HIP(hipSetDevice(eng->device));
    size_t free_b = 0, total_b = 0;
    if (ext) {
        free_b = size_t(ext->bytes);  // the whole region, nothing held back
    } else {
        MCS_HIP(hipMemGetInfo(&free_b, &total_b));
    }
    const uint64_t row_bytes = geo.stride * sizeof(uint64_t);
    const uint64_t reserve = ext ? 0 : uint64_t(256) << 20;
    const uint64_t one_buf = (uint64_t(plan.n_nodes()) * row_bytes);
    const uint64_t two_buf = one_buf + uint64_t(n_ev) * row_bytes;
    // Execution schedule (wide gates split into partial folds, each partial
    // one more row).
    mcs::Schedule sch = mcs::build_schedule(plan, kSplitFanIn);
    mcs::fuse_single_consumer(sch, plan, kTaskRows);
    const uint64_t aux_bytes = uint64_t(sch.n_aux) * row_bytes;
    eng->n_buf = (n_ev > 0 && two_buf + aux_bytes + reserve <= free_b) ? 2u : 1u;

    // Launch plan: one launch per schedule layer; launch 0 also carries the
    // tally-only tasks of tallied basic events (slot = event id). With
    // caller memory the second event buffer is kept only if every carved
    // array still fits; the plan's row numbers depend on the buffer count,
    // so it is rebuilt for one buffer otherwise.
    mcs::TreeLaunchPlan lp;
    mcs::RowMap rows{};
    for (;;) {
        eng->n_rows = uint64_t(eng->n_buf) * n_ev + plan.n_gates() + uint64_t(sch.n_aux);
        rows = mcs::RowMap{{0u, eng->n_buf == 2 ? uint32_t(n_ev) : 0u},
                           uint32_t(eng->n_buf) * uint32_t(n_ev), 0u};
        lp = mcs::TreeLaunchPlan{};
        mcs::build_launch_plan(plan, sch, tallied, rows, lp);
        if (!ext || eng->n_buf == 1) break;
        Carver c{nullptr, ext->bytes};
        c.take<uint64_t>(eng->n_rows * geo.stride);
        c.take<unsigned long long>(eng->n_rows);
        c.take<mcs::EventDesc>(uint64_t(n_ev));
        c.take<mcs::GateDesc>(lp.gates.size());
        c.take<mcs::GateDesc>(1);
        c.take<uint32_t>(std::max<uint64_t>(lp.args[0].size(), 1) * eng->n_buf);
        c.take<uint8_t>(eng->n_rows * mcs::state_stride_for(geo.n_chunks));
        if (c.used <= c.bytes) break;
        eng->n_buf = 1;
    }
    const std::vector<mcs::GateDesc> &gd = lp.gates;
    const std::vector<uint32_t> &args0 = lp.args[0], &args1 = lp.args[1];
    eng->launch_gate_begin = lp.launch_begin;
    // Fused gates, by node: their descriptors follow the task descriptors,
    // and each one's word is recomputed on demand by mcs_engine_read_words.
    eng->fused.assign(size_t(plan.n_nodes()), 0);
    eng->fused_desc.assign(size_t(plan.n_nodes()), mcs::GateDesc{});
    for (size_t i = lp.launch_begin.back(); i < gd.size(); ++i) {
        const int32_t node = int32_t(gd[i].out_row - rows.gate_row0) + n_ev;
        eng->fused[size_t(node)] = 1;
        eng->fused_desc[size_t(node)] = gd[i];
    }
    eng->n_args = args0.size();
    eng->arg_span = std::max<uint64_t>(args0.size(), 1);

    // Sampler tasks of the next pass, split over the launches in proportion
    // to each launch's gate traffic.
    const uint64_t samp_total = uint64_t(n_ev) * geo.n_chunks;
    eng->launch_samp_begin = mcs::split_in_proportion(lp.launch_cost, samp_total);
    for (uint32_t l = 0; l < eng->n_launches(); ++l) {
        mcs::PassLaunch pl{};
        pl.n_gates = eng->launch_gate_begin[l + 1] - eng->launch_gate_begin[l];
        pl.samp_begin = eng->launch_samp_begin[l];
        pl.samp_end = eng->launch_samp_begin[l + 1];
        if (mcs::pass_grid_blocks(pl, geo) * 256ull >= (uint64_t(1) << 32))
            return fail(MCS_ERR_INVALID_ARGUMENT, "grid too large: reduce words per pass");
    }
    {
        mcs::PassLaunch pl{};
        pl.samp_end = samp_total;
        if (mcs::pass_grid_blocks(pl, geo) * 256ull >= (uint64_t(1) << 32))
            return fail(MCS_ERR_INVALID_ARGUMENT, "grid too large: reduce words per pass");
    }

    // Event descriptors.
    std::vector<mcs::EventDesc> ev(n_ev);
    for (int32_t e = 0; e < n_ev; ++e)
        ev[e] = mcs::make_event_desc(plan.event_pdag_index[e], uint32_t(e), plan.event_p[e]);

    // Device memory: allocated, or carved from the caller's region.
    eng->state_stride = mcs::state_stride_for(geo.n_chunks);
    const uint64_t state_bytes = eng->n_rows * eng->state_stride;
    const uint64_t word_bytes = eng->n_rows * row_bytes;
    const uint64_t need = word_bytes + ev.size() * sizeof(mcs::EventDesc) +
                          gd.size() * sizeof(mcs::GateDesc) + eng->arg_span * 4 * eng->n_buf +
                          eng->n_rows * 8 + state_bytes;
    if (ext) {
        // Not the engine's to free, even if the carve below does not fit.
        eng->owns_memory = false;
        eng->owns_stream = false;
        eng->stream = ext->stream;
        Carver c{static_cast<uint8_t *>(ext->base), ext->bytes};
        eng->d_words = c.take<uint64_t>(word_bytes / 8);
        eng->d_ones = c.take<unsigned long long>(eng->n_rows);
        eng->d_events = ev.empty() ? nullptr : c.take<mcs::EventDesc>(ev.size());
        eng->d_gates = c.take<mcs::GateDesc>(gd.size());
        eng->d_probe = c.take<mcs::GateDesc>(1);
        eng->d_args = c.take<uint32_t>(eng->arg_span * eng->n_buf);
        eng->d_state = c.take<uint8_t>(state_bytes);
        if (c.used > c.bytes)
            return fail(MCS_ERR_OUT_OF_MEMORY,
                        "engine needs " + std::to_string(c.used) + " bytes of device memory, " +
                            std::string(ext->name ? ext->name : "the caller's region") +
                            " holds " + std::to_string(c.bytes) +
                            ": reduce words per pass or the tree");
        eng->device_bytes = c.used;
    } else {
        if (need + reserve > free_b)
            return fail(MCS_ERR_OUT_OF_MEMORY,
                        "engine needs " + std::to_string(need >> 20) + " MiB of device memory, " +
                            std::to_string(free_b >> 20) + " MiB free: reduce words per pass");
        MCS_HIP(hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking));
        MCS_HIP(hipMalloc(&eng->d_words, word_bytes));
        MCS_HIP(hipMalloc(&eng->d_ones, eng->n_rows * sizeof(unsigned long long)));
        if (!ev.empty()) MCS_HIP(hipMalloc(&eng->d_events, ev.size() * sizeof(mcs::EventDesc)));
        MCS_HIP(hipMalloc(&eng->d_gates, std::max<size_t>(gd.size(), 1) * sizeof(mcs::GateDesc)));
        MCS_HIP(hipMalloc(&eng->d_probe, sizeof(mcs::GateDesc)));
        MCS_HIP(hipMalloc(&eng->d_args, eng->arg_span * 4 * eng->n_buf));
        MCS_HIP(hipMalloc(&eng->d_state, std::max<uint64_t>(state_bytes, 1)));
        eng->device_bytes = need;
    }
    if (!ev.empty())
        MCS_HIP(hipMemcpyAsync(eng->d_events, ev.data(), ev.size() * sizeof(mcs::EventDesc),
                               hipMemcpyHostToDevice, eng->stream));
    if (!gd.empty())
        MCS_HIP(hipMemcpyAsync(eng->d_gates, gd.data(), gd.size() * sizeof(mcs::GateDesc),
                               hipMemcpyHostToDevice, eng->stream));
    if (!args0.empty()) {
        MCS_HIP(hipMemcpyAsync(eng->d_args, args0.data(), args0.size() * 4, hipMemcpyHostToDevice,
                               eng->stream));
        if (eng->n_buf == 2)
            MCS_HIP(hipMemcpyAsync(eng->d_args + eng->arg_span, args1.data(), args1.size() * 4,
                                   hipMemcpyHostToDevice, eng->stream));
    }
    MCS_HIP(hipMemsetAsync(eng->d_ones, 0, eng->n_rows * sizeof(unsigned long long), eng->stream));
    // A summary no task has written yet says "mixed": its words are loaded.
    MCS_HIP(hipMemsetAsync(eng->d_state, mcs::kChunkMixed, state_bytes, eng->stream));
    MCS_HIP(hipStreamSynchronize(eng->stream));
    eng->key0 = key0;
    eng->key1 = key1;
    *out = eng.release();
    return MCS_OK;
}

mcs::PassLaunch base_launch(const mcs_engine *e) {
    mcs::PassLaunch pl{};
    pl.events = e->d_events;
    pl.key0 = e->key0;
    pl.key1 = e->key1;
    pl.state = e->d_state;
    pl.state_stride = e->state_stride;
    return pl;
}

// Launch hook: called around every kernel launch (for timing); null = none.
struct LaunchHook {
    virtual hipError_t before() = 0;
    virtual hipError_t after() = 0;
    virtual ~LaunchHook() = default;
};

hipError_t launch(mcs_engine *e, const mcs::PassLaunch &pl, LaunchHook *hook) {
    if (mcs::pass_grid_blocks(pl, e->geo) == 0) return hipSuccess;
    hipError_t st = hook ? hook->before() : hipSuccess;
    if (st == hipSuccess) st = mcs::launch_pass(e->stream, pl, e->d_words, e->d_ones, e->geo);
    if (st == hipSuccess && hook) st = hook->after();
    return st;
}

mcs_status enqueue_passes(mcs_engine *e, uint32_t n, LaunchHook *hook = nullptr) {
    MCS_HIP(hipSetDevice(e->device));
    const uint32_t n_ev = uint32_t(e->plan.n_events());
    const uint64_t samp_total = uint64_t(n_ev) * e->geo.n_chunks;
    for (uint32_t i = 0; i < n; ++i) {
        if (e->iteration == 0xFFFFFFFFu) return fail(MCS_ERR_STATE, "iteration counter exhausted");
        const uint32_t it = ++e->iteration;
        const uint32_t par = e->parity(it);
        if (e->primed != it) {  // sample this pass on its own
            mcs::PassLaunch pl = base_launch(e);
            pl.samp_iteration = it;
            pl.samp_row0 = par * n_ev;
            pl.samp_begin = 0;
            pl.samp_end = samp_total;
            MCS_HIP(launch(e, pl, hook));
        }
        const bool prefetch = e->n_buf == 2 && it != 0xFFFFFFFFu;
        for (uint32_t l = 0; l < e->n_launches(); ++l) {
            mcs::PassLaunch pl = base_launch(e);
            pl.gates = e->d_gates + e->launch_gate_begin[l];
            pl.tree_gates = e->d_gates;
            pl.arg_rows = e->d_args + uint64_t(par) * e->arg_span;
            pl.n_gates = e->launch_gate_begin[l + 1] - e->launch_gate_begin[l];
            if (prefetch) {
                pl.samp_iteration = it + 1;
                pl.samp_row0 = (par ^ 1u) * n_ev;
                pl.samp_begin = e->launch_samp_begin[l];
                pl.samp_end = e->launch_samp_begin[l + 1];
            }
            MCS_HIP(launch(e, pl, hook));
        }
        e->primed = prefetch ? it + 1 : 0;
    }
    return MCS_OK;
}

void fill_estimates_impl(mcs_tally &t) {
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
    std::vector<unsigned long long> ones(e->n_rows);
    MCS_HIP(hipMemcpyAsync(ones.data(), e->d_ones, ones.size() * 8, hipMemcpyDeviceToHost, e->stream));
    MCS_HIP(hipStreamSynchronize(e->stream));
    for (size_t t = 0; t < e->tally_nodes.size(); ++t) {
        mcs_tally &r = out[t];
        r.index = e->tally_index[t];
        r.num_one_bits = ones[e->tally_slot(e->tally_nodes[t])];
        // total = iteration * N_bp * 64 (src/mc/kernel/tally.h:345-352).
        r.total_bits = uint64_t(e->iteration) * e->geo.n_bp * 64ull;
        mcs_tally_point_estimates(&r);
    }
    return MCS_OK;
}

struct EventTimer : LaunchHook {
    hipStream_t stream;
    std::vector<hipEvent_t> ev;
    explicit EventTimer(hipStream_t s) : stream(s) {}
    hipError_t record() {
        hipEvent_t x = nullptr;
        hipError_t st = hipEventCreate(&x);
        if (st != hipSuccess) return st;
        ev.push_back(x);
        return hipEventRecord(x, stream);
    }
    hipError_t before() override { return record(); }
    hipError_t after() override { return record(); }
    ~EventTimer() override {
        for (hipEvent_t x : ev) (void)hipEventDestroy(x);
    }
};

}  // namespace

mcs_status mcs::create_engine(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                              const int32_t *tally_index, int32_t n_tally, uint32_t key0,
                              uint32_t key1, int32_t device, const ExternalMemory *ext,
                              mcs_engine **out) {
    return build_engine(pdag, shape, tally_index, n_tally, key0, key1, device, ext, out);
}

extern "C" {

const char *mcs_last_error(void) { return g_error.c_str(); }

void mcs_tally_point_estimates(mcs_tally *t) {
    if (t) fill_estimates_impl(*t);
}

mcs_status mcs_engine_create(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                             const int32_t *tally_index, int32_t n_tally, mcs_engine **out) {
    return build_engine(pdag, shape, tally_index, n_tally, MCS_REFERENCE_KEY0,
                        MCS_REFERENCE_KEY1, -1, nullptr, out);
}

mcs_status mcs_engine_create_ex(const mcs_pdag *pdag, const mcs_sample_shape *shape,
                                const int32_t *tally_index, int32_t n_tally, uint32_t key0,
                                uint32_t key1, int32_t device, mcs_engine **out) {
    return build_engine(pdag, shape, tally_index, n_tally, key0, key1, device, nullptr, out);
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
    out->event_buffers = int32_t(engine->n_buf);
    out->launches_per_pass = engine->n_launches() + (engine->n_buf == 2 ? 0u : 1u);
    out->fused_gates = int32_t(std::count(engine->fused.begin(), engine->fused.end(), uint8_t(1)));
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
    if (engine->iteration == 0) return fail(MCS_ERR_STATE, "no pass has run yet");
    const int32_t node = it->second;
    const uint32_t par = engine->parity(engine->iteration);
    const uint64_t row = node < engine->plan.n_events()
                             ? uint64_t(node) + uint64_t(par) * engine->plan.n_events()
                             : engine->gate_row(node);
    MCS_HIP(hipSetDevice(engine->device));
    if (engine->fused[size_t(node)]) {
        // A fused gate's word lives only in registers during a pass: store
        // it into its own (otherwise unused) row from the last pass's rows,
        // which stay intact until the next pass, without tallying it again.
        mcs::GateDesc &d = engine->probe_host;
        d = engine->fused_desc[size_t(node)];
        d.type_k_flags &= ~(mcs::kGateNoStore | mcs::kGateTallied);
        d.type_k_flags |= mcs::kGateForceStore;
        MCS_HIP(hipMemcpyAsync(engine->d_probe, &d, sizeof(d), hipMemcpyHostToDevice,
                               engine->stream));
        mcs::PassLaunch pl = base_launch(engine);
        pl.gates = engine->d_probe;
        pl.tree_gates = engine->d_gates;
        pl.arg_rows = engine->d_args + uint64_t(par) * engine->arg_span;
        pl.n_gates = 1;
        MCS_HIP(launch(engine, pl, nullptr));
    }
    MCS_HIP(hipMemcpyAsync(out, engine->d_words + row * engine->geo.stride, engine->geo.n_bp * 8,
                           hipMemcpyDeviceToHost, engine->stream));
    std::vector<uint8_t> summary(engine->geo.n_chunks);
    MCS_HIP(hipMemcpyAsync(summary.data(), engine->d_state + row * engine->state_stride,
                           summary.size(), hipMemcpyDeviceToHost, engine->stream));
    MCS_HIP(hipStreamSynchronize(engine->stream));
    // Chunks whose summary is constant were not stored: their words are the
    // constant (fused gates were just stored whole).
    if (!engine->fused[size_t(node)]) {
        const uint64_t cw = engine->geo.chunk_words;
        for (uint32_t c = 0; c < engine->geo.n_chunks; ++c) {
            if (summary[c] == mcs::kChunkMixed) continue;
            const uint64_t w = summary[c] == mcs::kChunkOnes ? ~uint64_t(0) : 0;
            const uint64_t end = std::min<uint64_t>((c + 1) * cw, engine->geo.n_bp);
            std::fill(out + c * cw, out + end, w);
        }
    }
    return MCS_OK;
}

mcs_status mcs_engine_time_passes(mcs_engine *engine, uint32_t n_passes, double *kernel_ms,
                                  uint32_t *launches, double *pass_ms) {
    if (!engine || !kernel_ms || !launches || !pass_ms)
        return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n_passes == 0) return fail(MCS_ERR_INVALID_ARGUMENT, "n_passes must be positive");
    MCS_HIP(hipSetDevice(engine->device));
    EventTimer timer(engine->stream);
    mcs_status st = enqueue_passes(engine, n_passes, &timer);
    if (st != MCS_OK) return st;
    MCS_HIP(hipStreamSynchronize(engine->stream));
    double sum = 0;
    float ms = 0;
    for (size_t i = 0; i + 1 < timer.ev.size(); i += 2) {
        MCS_HIP(hipEventElapsedTime(&ms, timer.ev[i], timer.ev[i + 1]));
        sum += ms;
    }
    *kernel_ms = sum;
    *launches = uint32_t(timer.ev.size() / 2);
    *pass_ms = 0;
    if (timer.ev.size() >= 2) {
        MCS_HIP(hipEventElapsedTime(&ms, timer.ev.front(), timer.ev.back()));
        *pass_ms = ms;
    }
    return MCS_OK;
}

mcs_status mcs_engine_restore(mcs_engine *engine, uint32_t iteration, const uint64_t *num_one_bits,
                              int32_t n) {
    if (!engine || (n > 0 && !num_one_bits)) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n != int32_t(engine->tally_nodes.size()))
        return fail(MCS_ERR_INVALID_ARGUMENT, "restore needs one count per tally");
    MCS_HIP(hipSetDevice(engine->device));
    std::vector<unsigned long long> ones(engine->n_rows, 0);
    for (int32_t t = 0; t < n; ++t) ones[engine->tally_slot(engine->tally_nodes[t])] = num_one_bits[t];
    MCS_HIP(hipStreamSynchronize(engine->stream));
    MCS_HIP(hipMemcpy(engine->d_ones, ones.data(), ones.size() * 8, hipMemcpyHostToDevice));
    engine->iteration = iteration;
    engine->primed = 0;
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
    s->ci_policy = MCS_CI_BAYES;  // src/settings.h:426
    s->early_stop = 1;            // src/settings.h:453
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

mcs_status mcs_sample_shape_for(uint64_t num_trials, uint64_t num_nodes, uint64_t max_alloc_bytes,
                                uint64_t max_batch, uint64_t max_bitpacks,
                                mcs_sample_shape *shape, uint64_t *rounded_trials,
                                uint64_t *passes) {
    return mcs_sample_shape_for_device(num_trials, num_nodes, max_alloc_bytes, max_batch,
                                       max_bitpacks, 0, 0.0, 0, shape, rounded_trials, passes);
}

mcs_status mcs_sample_shape_for_device(uint64_t num_trials, uint64_t num_nodes,
                                       uint64_t max_alloc_bytes, uint64_t max_batch,
                                       uint64_t max_bitpacks, int32_t nvidia_gpu,
                                       double overhead_ratio, uint64_t cache_line_bytes,
                                       mcs_sample_shape *shape, uint64_t *rounded_trials,
                                       uint64_t *passes) {
    if (!shape || !rounded_trials || !passes) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (num_nodes == 0 || max_alloc_bytes == 0 || max_batch == 0 || max_bitpacks == 0)
        return fail(MCS_ERR_INVALID_ARGUMENT, "nodes, allocation size and work-item limits must be positive");
    if (max_alloc_bytes > (~uint64_t(0) >> 3))
        return fail(MCS_ERR_INVALID_ARGUMENT, "allocation size too large");
    if (nvidia_gpu && (cache_line_bytes < 8 || !(overhead_ratio == overhead_ratio)))
        return fail(MCS_ERR_INVALID_ARGUMENT,
                    "the NVIDIA rule needs a cache line of at least one word and a numeric overhead ratio");
    // No trial limit: 2^26 trials size the shape (sample_shaper.h:67-72).
    const bool unlimited = num_trials == 0;
    const uint64_t requested = unlimited ? (uint64_t(1) << 26) : num_trials;
    const uint64_t rem = requested % 64;  // nearest multiple of 64, ties down (:74-83)
    const uint64_t total_bits = rem == 0 ? requested : (rem <= 32 ? requested - rem : requested + 64 - rem);
    // Bits per node per iteration the largest allocation allows (:89).
    const uint64_t cap_bits = (max_alloc_bytes * 8) / num_nodes;
    const uint64_t target_bits = std::min(total_bits, cap_bits);  // :105
    // Whole bitpacks (:340-344).
    const uint64_t target = (target_bits + 63) / 64;
    const uint64_t capacity = cap_bits / 64;
    // Closest power-of-two product batch x bitpacks, ties to the smaller
    // product (:274-334; power-of-two batch sizes always meet its sub-group
    // alignment rule).
    uint64_t best = 0, best_diff = ~uint64_t(0);
    for (uint64_t bs = 1; bs <= max_batch && bs <= capacity; bs <<= 1) {
        const uint64_t max_ss = std::min(max_bitpacks, capacity / bs);
        for (uint64_t ss = 1; ss <= max_ss; ss <<= 1) {
            const uint64_t prod = bs * ss;
            const uint64_t diff = prod > target ? prod - target : target - prod;
            if (diff < best_diff || (diff == best_diff && prod < best)) {
                best_diff = diff;
                best = prod;
                if (diff == 0) break;
            }
            if (ss > (~uint64_t(0) >> 1)) break;
        }
        if (best_diff == 0 || bs > (~uint64_t(0) >> 1)) break;
    }
    if (best == 0) best = 1;  // no legal shape under the cap: the smallest one
    if (nvidia_gpu) {
        // NVIDIA GPUs take the largest allocation whatever the trial count:
        // bitpacks = floor((max_alloc / 8) / (nodes (1 + overhead))), the
        // overhead clamped to [0, 1e6], rounded down to whole cache lines
        // (:120-134).
        const double max_bitpacks_d = double(max_alloc_bytes) / 8.0;
        const double effective_nodes = double(num_nodes) * (1.0 + std::clamp(overhead_ratio, 0.0, 1e6));
        const uint64_t bitpacks = uint64_t(std::floor(max_bitpacks_d / effective_nodes));
        const uint64_t line = uint64_t(double(cache_line_bytes) / 8.0);
        best = bitpacks - bitpacks % line;
        if (best == 0)
            return fail(MCS_ERR_INVALID_ARGUMENT,
                        "the NVIDIA rule leaves no whole cache line per node: the reference would divide by zero");
    }
    shape->batch_size = 1;    // all the work in the last dimension (:109-113)
    shape->bitpacks_per_batch = best;
    *rounded_trials = total_bits;
    // Passes = ceil(total bits / bits per pass) in double precision, as the
    // reference computes it (:139-146); exact below 2^53 bits.
    const double per = double(best) * 64.0;
    *passes = unlimited ? 0 : uint64_t(std::ceil(double(total_bits) / per));
    return MCS_OK;
}

}  // extern "C"
