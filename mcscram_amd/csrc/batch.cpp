// Batch engine: many independent trees per pass.
//
// Each tree goes one of two ways, decided by the LDS tile it would get:
//
//  * fused: the tree fits the LDS budget at min_fused_tile_words or more
//    words per node. All such trees share one persistent launch of
//    fused_trees (fused.hip): work items of whole tiles, largest first,
//    every tile sampled and evaluated in LDS; only tallies reach HBM.
//  * layered: larger trees, grouped by sample shape into pipelines through
//    pass_layer (kernels.hip), layered_streams pipelines per shape with the
//    trees dealt to them by cost. The trees of a pipeline reuse one words
//    arena: two event areas and one gate area. Tree t's gate launches carry
//    the sampler tasks of the next tree (into the other event area), so the
//    Philox work of one tree runs beside the gate traffic of another; the
//    last tree of a pass samples the first tree of the next pass.
//
// A pass starts with a tiny kernel that resets the fused work queue and
// advances the device iteration counter read by both kernels. Each pipeline
// has its own stream: after that kernel the pipelines fork from the batch's
// stream, run beside the fused launch and beside each other (one pipeline's
// short launches and launch tails leave the CUs to the others), and join it
// again at the end of the pass. From the second pass on, a pass is the same
// launches every time up to the event areas the pipelines alternate
// between, so with use_graph the whole pass (the counter kernel, the fork,
// the fused launch, every pipeline's launches and the join) is captured
// once per area pattern as a hipGraph (two at most) and replayed with one
// graph launch per pass.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "fused.h"
#include "kernels.h"
#include "launch_plan.h"
#include "layering.h"
#include "mcscram_amd.h"
#include "status.h"

using mcs::fail;

namespace {

constexpr uint32_t kBatchMaxFanIn = 64;

struct TreeHost {
    mcs::Plan plan;
    mcs::Schedule sch;
    int32_t n_nodes = 0;                // plan nodes + partial folds
    std::vector<int32_t> tally_nodes;   // plan node ids, first-appearance order
    std::vector<int32_t> tally_index;   // pdag indices
    std::vector<uint8_t> tallied;       // per plan node
    uint64_t n_bp = 0;
    bool fused = false;
    // fused
    std::vector<int32_t> slot_of_node;  // plan node id -> LDS node slot
    uint32_t ones_begin = 0;            // in the fused counters
    uint32_t tile_words = 0;
    // layered
    int32_t group = -1;
    uint64_t group_ones_begin = 0;      // in the group's counters (gates, partials, events)
};

// One layered pipeline: trees with the same words per pass.
struct LayeredGroup {
    uint64_t n_bp = 0;
    mcs::LaunchGeometry geo{};
    uint32_t max_ev = 0;
    std::vector<int32_t> trees;  // batch tree ids, in pipeline order
    hipStream_t stream = nullptr;  // the pipeline's launches
    hipEvent_t done = nullptr;     // recorded after a pass's launches
    struct Tree {
        mcs::TreeLaunchPlan lp;
        uint64_t gate_begin = 0, args_begin = 0, ev_begin = 0;
        uint32_t n_ev = 0;
    };
    std::vector<Tree> t;
    uint64_t *d_words = nullptr;
    mcs::EventDesc *d_events = nullptr;
    mcs::GateDesc *d_gates = nullptr;
    uint32_t *d_args[2] = {nullptr, nullptr};
    unsigned long long *d_ones = nullptr;  // leading pad of 2 * max_ev, then the trees
    uint8_t *d_sum = nullptr;              // chunk summaries of the arena's rows
    uint32_t sum_stride = 0;
    uint64_t seq = 0;                      // tree-passes run: event area = seq & 1
    int64_t primed_pos = -1;               // tree position sampled into area seq & 1
    uint32_t primed_iter = 0;
};

template <typename T>
hipError_t upload(T **dst, const std::vector<T> &src, hipStream_t s) {
    const size_t n = std::max<size_t>(src.size(), 1);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), n * sizeof(T));
    if (e != hipSuccess || src.empty()) return e;
    return hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

uint32_t pick_chunk_words(uint64_t n_bp) {
    if (n_bp >= (uint64_t(1) << 18)) return 2048;
    if (n_bp >= (uint64_t(1) << 16)) return 1024;
    return 512;
}

}  // namespace

struct mcs_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipGraph_t pass_graph[2] = {nullptr, nullptr};  // steady passes, by area pattern
    hipGraphExec_t pass_exec[2] = {nullptr, nullptr};
    uint64_t pass_key[2] = {~uint64_t(0), ~uint64_t(0)};
    mcs_batch_config cfg{};
    std::vector<TreeHost> trees;
    std::vector<LayeredGroup> groups;
    uint32_t iteration = 0;
    uint32_t n_items = 0;
    uint32_t grid = 0;
    uint32_t lds_bytes = 0;
    uint32_t lanes = 0;
    uint32_t n_fused = 0;
    uint64_t n_ones = 0;
    uint64_t device_bytes = 0;
    uint32_t *d_state = nullptr;  // [0] fused work-queue head, [1] iteration
    hipEvent_t began = nullptr;   // recorded after a pass's first kernel
    mcs::FusedLaunch launch{};
    std::vector<void *> allocations;

    ~mcs_batch() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        for (LayeredGroup &g : groups)
            if (g.stream) (void)hipStreamSynchronize(g.stream);
        for (int i = 0; i < 2; ++i) {
            if (pass_exec[i]) (void)hipGraphExecDestroy(pass_exec[i]);
            if (pass_graph[i]) (void)hipGraphDestroy(pass_graph[i]);
        }
        for (void *p : allocations) (void)hipFree(p);
        for (LayeredGroup &g : groups) {
            if (g.done) (void)hipEventDestroy(g.done);
            if (g.stream) (void)hipStreamDestroy(g.stream);
        }
        if (began) (void)hipEventDestroy(began);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// The layered pipelines' launches of one pass. The device iteration counter
// already holds this pass's iteration (it).
mcs_status enqueue_layered(mcs_batch *b, uint32_t it) {
    for (LayeredGroup &g : b->groups) {
        const uint32_t T = uint32_t(g.trees.size());
        for (uint32_t pos = 0; pos < T; ++pos) {
            const LayeredGroup::Tree &tr = g.t[pos];
            const uint32_t area = uint32_t(g.seq & 1u);
            mcs::PassLaunch base{};
            base.key0 = b->cfg.key0;
            base.key1 = b->cfg.key1;
            base.iteration_ptr = b->d_state + 1;
            base.state = g.d_sum;
            base.state_stride = g.sum_stride;
            if (!(g.primed_pos == int64_t(pos) && g.primed_iter == it)) {
                mcs::PassLaunch pl = base;
                pl.events = g.d_events + tr.ev_begin;
                pl.samp_iteration = 0;  // this pass
                pl.samp_row0 = area * g.max_ev;
                pl.samp_end = uint64_t(tr.n_ev) * g.geo.n_chunks;
                MCS_HIP(mcs::launch_pass(g.stream, pl, g.d_words, g.d_ones, g.geo));
            }
            // The next tree of the pipeline: the following one, or the first
            // one of the next pass.
            const uint32_t npos = pos + 1 < T ? pos + 1 : 0;
            const uint32_t nit_off = pos + 1 < T ? 0u : 1u;
            const LayeredGroup::Tree &nx = g.t[npos];
            const std::vector<uint64_t> split =
                mcs::split_in_proportion(tr.lp.launch_cost, uint64_t(nx.n_ev) * g.geo.n_chunks);
            unsigned long long *ones = g.d_ones + b->trees[size_t(g.trees[pos])].group_ones_begin;
            for (size_t l = 0; l + 1 < tr.lp.launch_begin.size(); ++l) {
                mcs::PassLaunch pl = base;
                pl.gates = g.d_gates + tr.gate_begin + tr.lp.launch_begin[l];
                pl.tree_gates = g.d_gates + tr.gate_begin;
                pl.arg_rows = g.d_args[area] + tr.args_begin;
                pl.n_gates = tr.lp.launch_begin[l + 1] - tr.lp.launch_begin[l];
                pl.events = g.d_events + nx.ev_begin;
                pl.samp_iteration = nit_off;
                pl.samp_row0 = (area ^ 1u) * g.max_ev;
                pl.samp_begin = split[l];
                pl.samp_end = split[l + 1];
                // Gate rows start at 2 max_ev; counters of this tree start at
                // ones, so the kernel's slot (row) is offset by 2 max_ev.
                MCS_HIP(mcs::launch_pass(g.stream, pl, g.d_words, ones - 2u * g.max_ev, g.geo));
            }
            g.primed_pos = npos;
            g.primed_iter = it + nit_off;
            ++g.seq;
        }
    }
    return MCS_OK;
}

// The launches of pass it, and the host bookkeeping they imply.
mcs_status enqueue_pass_launches(mcs_batch *b, uint32_t it) {
    MCS_HIP(mcs::launch_fused_begin_pass(b->stream, b->d_state));
    // Fork: the pipelines start once the iteration counter has advanced.
    if (!b->groups.empty()) MCS_HIP(hipEventRecord(b->began, b->stream));
    for (LayeredGroup &g : b->groups) MCS_HIP(hipStreamWaitEvent(g.stream, b->began, 0));
    if (b->n_fused) MCS_HIP(mcs::launch_fused(b->stream, b->launch, b->grid, b->lanes, b->lds_bytes));
    const mcs_status st = enqueue_layered(b, it);
    if (st != MCS_OK) return st;
    // Join: the batch's stream (and the next pass's first kernel, which
    // advances the counter the pipelines read) waits for every pipeline.
    for (LayeredGroup &g : b->groups) {
        MCS_HIP(hipEventRecord(g.done, g.stream));
        MCS_HIP(hipStreamWaitEvent(b->stream, g.done, 0));
    }
    return MCS_OK;
}

// Pass it is steady when every pipeline's first tree was sampled by the
// previous pass: then its launches depend only on the event areas the
// pipelines start in (the area pattern, one bit per pipeline).
bool steady_pass(const mcs_batch *b, uint32_t it, uint64_t *key) {
    if (b->groups.size() > 64) return false;
    uint64_t k = 0;
    for (size_t i = 0; i < b->groups.size(); ++i) {
        const LayeredGroup &g = b->groups[i];
        if (!(g.primed_pos == 0 && g.primed_iter == it)) return false;
        k |= (g.seq & 1u) << i;
    }
    *key = k;
    return true;
}

mcs_status enqueue_pass(mcs_batch *b) {
    if (b->iteration == 0xFFFFFFFFu) return fail(MCS_ERR_STATE, "iteration counter exhausted");
    const uint32_t it = b->iteration + 1;
    uint64_t key = 0;
    if (!b->cfg.use_graph || !steady_pass(b, it, &key)) {
        const mcs_status st = enqueue_pass_launches(b, it);
        if (st != MCS_OK) return st;
        b->iteration = it;
        return MCS_OK;
    }
    int slot = -1;
    for (int i = 0; i < 2 && slot < 0; ++i)
        if (b->pass_exec[i] && b->pass_key[i] == key) slot = i;
    if (slot >= 0) {
        // Replay, and advance the host bookkeeping as the launches would.
        MCS_HIP(hipGraphLaunch(b->pass_exec[slot], b->stream));
        for (LayeredGroup &g : b->groups) {
            g.seq += g.trees.size();
            g.primed_pos = 0;
            g.primed_iter = it + 1;
        }
        b->iteration = it;
        return MCS_OK;
    }
    for (int i = 0; i < 2 && slot < 0; ++i)
        if (!b->pass_exec[i]) slot = i;
    if (slot < 0) return fail(MCS_ERR_STATE, "more than two area patterns in steady passes");
    // Capture this pass (its host bookkeeping advances as for an eager pass),
    // then launch it.
    MCS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const mcs_status st = enqueue_pass_launches(b, it);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (st != MCS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    MCS_HIP(e);
    b->pass_graph[slot] = graph;
    MCS_HIP(hipGraphInstantiate(&b->pass_exec[slot], graph, nullptr, nullptr, 0));
    b->pass_key[slot] = key;
    MCS_HIP(hipGraphLaunch(b->pass_exec[slot], b->stream));
    b->iteration = it;
    return MCS_OK;
}

mcs_status build_batch(const mcs_pdag *pdags, int32_t n_trees, const mcs_sample_shape *shapes,
                       const int32_t *tally_offsets, const int32_t *tally_index,
                       const mcs_batch_config *cfg_in, mcs_batch **out) {
    if (!pdags || !shapes || !out || n_trees <= 0 || (tally_index && !tally_offsets))
        return fail(MCS_ERR_INVALID_ARGUMENT, "null argument or empty batch");
    *out = nullptr;
    auto b = std::make_unique<mcs_batch>();
    mcs_batch_config_default(&b->cfg);
    if (cfg_in) b->cfg = *cfg_in;
    mcs_batch_config &cfg = b->cfg;
    if (cfg.lds_bytes == 0) cfg.lds_bytes = 64u << 10;
    if (cfg.max_tile_words == 0) cfg.max_tile_words = 64;
    if (cfg.item_words == 0) cfg.item_words = 4096;
    if (cfg.workgroup_lanes == 0) cfg.workgroup_lanes = 1024;
    if (cfg.min_fused_tile_words == 0) cfg.min_fused_tile_words = 8;
    if (cfg.layered_streams == 0) cfg.layered_streams = 2;
    if (cfg.layered_streams > 8)
        return fail(MCS_ERR_INVALID_ARGUMENT, "layered_streams must be at most 8");
    if (cfg.max_tile_words > 64 || (cfg.max_tile_words & (cfg.max_tile_words - 1)))
        return fail(MCS_ERR_INVALID_ARGUMENT, "max_tile_words must be a power of two <= 64");
    if (cfg.lds_bytes > mcs::kMaxFusedLds)
        return fail(MCS_ERR_INVALID_ARGUMENT, "lds_bytes exceeds the 160 KiB of a CU");
    if (cfg.workgroup_lanes > 1024 || cfg.workgroup_lanes % 64)
        return fail(MCS_ERR_INVALID_ARGUMENT, "workgroup_lanes must be a multiple of 64 up to 1024");
    b->lanes = cfg.workgroup_lanes;

    // Plans, schedules, tallies, and the route of every tree.
    b->trees.resize(size_t(n_trees));
    for (int32_t t = 0; t < n_trees; ++t) {
        TreeHost &th = b->trees[size_t(t)];
        const mcs_sample_shape &sh = shapes[t];
        th.n_bp = sh.batch_size * sh.bitpacks_per_batch;
        if (sh.batch_size == 0 || sh.bitpacks_per_batch == 0 ||
            th.n_bp / sh.batch_size != sh.bitpacks_per_batch || th.n_bp > (uint64_t(1) << 32))
            return fail(MCS_ERR_INVALID_ARGUMENT,
                        "tree " + std::to_string(t) + ": sample shape must be positive and at most 2^32 words");
        std::string err;
        if (!mcs::build_plan(pdags[t], th.plan, err))
            return fail(MCS_ERR_GRAPH, "tree " + std::to_string(t) + ": " + err);
        th.sch = mcs::build_schedule(th.plan, kBatchMaxFanIn);
        th.n_nodes = th.plan.n_nodes() + th.sch.n_aux;

        std::unordered_map<int32_t, int32_t> node_of_index;
        for (int32_t n = 0; n < th.plan.n_nodes(); ++n) node_of_index[th.plan.node_pdag_index[n]] = n;
        th.tallied.assign(size_t(th.plan.n_nodes()), 0);
        const int32_t t0 = tally_index ? tally_offsets[t] : 0;
        const int32_t t1 = tally_index ? tally_offsets[t + 1] : 0;
        if (t1 < t0) return fail(MCS_ERR_INVALID_ARGUMENT, "tally offsets are not monotone");
        std::vector<int32_t> want;
        for (int32_t i = t0; i < t1; ++i) want.push_back(tally_index[i]);
        if (want.empty()) want.push_back(pdags[t].root);
        for (int32_t idx : want) {
            auto it = node_of_index.find(idx);
            if (it == node_of_index.end())
                return fail(MCS_ERR_GRAPH, "tree " + std::to_string(t) + ": tally index " +
                                               std::to_string(idx) + " is not reachable from the root");
            if (th.tallied[size_t(it->second)]) continue;
            th.tallied[size_t(it->second)] = 1;
            th.tally_nodes.push_back(it->second);
            th.tally_index.push_back(idx);
        }

        uint32_t w = cfg.max_tile_words;
        while (w > 1 && mcs::fused_lds_bytes(uint32_t(th.n_nodes), w) > cfg.lds_bytes) w >>= 1;
        th.fused = w >= cfg.min_fused_tile_words &&
                   mcs::fused_lds_bytes(uint32_t(th.n_nodes), w) <= cfg.lds_bytes &&
                   th.sch.layers.size() <= mcs::kMaxFusedLayers;
        th.tile_words = th.fused ? w : 0;
    }

    MCS_HIP(hipGetDevice(&b->device));
    MCS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    auto keep = [&](void *p) { b->allocations.push_back(p); };
    MCS_HIP(hipMalloc(&b->d_state, 2 * sizeof(uint32_t)));
    keep(b->d_state);
    MCS_HIP(hipMemsetAsync(b->d_state, 0, 2 * sizeof(uint32_t), b->stream));

    // ---- fused part ---------------------------------------------------
    std::vector<mcs::FusedTree> ft;
    std::vector<mcs::EventDesc> ev;
    std::vector<mcs::FusedGate> gates;
    std::vector<uint32_t> layer_off, args;
    std::vector<double> cost_per_tile;
    std::vector<int32_t> fused_ids;
    for (int32_t t = 0; t < n_trees; ++t) {
        TreeHost &th = b->trees[size_t(t)];
        if (!th.fused) continue;
        const mcs::Plan &plan = th.plan;
        const mcs::Schedule &sch = th.sch;
        const int32_t n_ev = plan.n_events();
        const int32_t n_nodes = th.n_nodes;
        const uint32_t w = th.tile_words;
        b->lds_bytes = std::max<uint32_t>(b->lds_bytes, uint32_t(mcs::fused_lds_bytes(uint32_t(n_nodes), w)));

        // LDS node slots: events, then gates in schedule order.
        th.slot_of_node.assign(size_t(n_nodes), -1);
        for (int32_t e = 0; e < n_ev; ++e) th.slot_of_node[size_t(e)] = e;
        int32_t next = n_ev;
        for (const auto &layer : sch.layers)
            for (const mcs::ExecGate &g : layer) th.slot_of_node[size_t(g.out)] = next++;

        mcs::FusedTree d{};
        d.n_bp = th.n_bp;
        d.n_ev = uint32_t(n_ev);
        d.n_nodes = uint32_t(n_nodes);
        d.n_layers = uint32_t(sch.layers.size());
        d.log2_w = uint32_t(__builtin_ctz(w));
        d.ev_begin = uint32_t(ev.size());
        d.gate_begin = uint32_t(gates.size());
        d.layer_begin = uint32_t(layer_off.size());
        d.ones_begin = uint32_t(b->n_ones);
        th.ones_begin = d.ones_begin;
        b->n_ones += uint64_t(n_nodes);

        for (int32_t e = 0; e < n_ev; ++e) {
            mcs::EventDesc x = mcs::make_event_desc(plan.event_pdag_index[e], uint32_t(e), plan.event_p[e]);
            if (th.tallied[size_t(e)]) x.mode |= mcs::kEventTallied;
            ev.push_back(x);
        }
        uint64_t fan = 0;
        uint32_t local = 0;
        for (const auto &layer : sch.layers) {
            layer_off.push_back(local);
            for (const mcs::ExecGate &g : layer) {
                mcs::FusedGate fg{};
                // Slots in the reference's input order (positive, then
                // negated); NOT and NULL read only their last input.
                std::vector<uint32_t> slots;
                for (int32_t i : g.pos_in) slots.push_back(uint32_t(th.slot_of_node[size_t(i)]));
                for (int32_t i : g.neg_in)
                    slots.push_back(uint32_t(th.slot_of_node[size_t(i)]) | mcs::kSlotNegated);
                uint32_t type = g.type;
                if (type == MCS_NOT || type == MCS_NULL) {
                    if (!slots.empty()) slots = {slots.back()};
                    type = type == MCS_NOT ? MCS_NOR : MCS_OR;
                }
                const uint32_t n_in = uint32_t(slots.size());
                fg.n_in = n_in;
                const bool tal = g.out < plan.n_nodes() && th.tallied[size_t(g.out)];
                fg.type_k_flags = type | (uint32_t(uint8_t(g.k)) << 8) |
                                  (tal ? mcs::kGateTallied : 0u);
                if (n_in <= mcs::kInlineInputs) {
                    const bool and_like = type == MCS_AND || type == MCS_NAND;
                    const uint32_t pad = uint32_t(n_nodes) + (and_like ? 1u : 0u);
                    for (uint32_t j = 0; j < mcs::kInlineInputs; ++j)
                        fg.in[j] = j < n_in ? slots[j] : pad;
                } else {
                    fg.in[0] = uint32_t(args.size());
                    args.insert(args.end(), slots.begin(), slots.end());
                }
                gates.push_back(fg);
                fan += n_in;
                ++local;
            }
        }
        layer_off.push_back(local);
        ft.push_back(d);
        // Relative cost of one tile: Philox words dominate.
        cost_per_tile.push_back(double(w) * (double(n_ev) * 40.0 + double(fan) + double(local)));
        fused_ids.push_back(t);
        if (args.size() > 0xFFFFFFFFull || gates.size() > 0xFFFFFFFFull)
            return fail(MCS_ERR_INVALID_ARGUMENT, "batch too large");
    }
    b->n_fused = uint32_t(fused_ids.size());

    if (b->n_fused) {
        // Occupancy: the grid is every workgroup slot LDS and registers allow.
        int per_cu = 0;
        MCS_HIP(mcs::fused_max_workgroups_per_cu(b->lanes, b->lds_bytes, &per_cu));
        if (per_cu <= 0)
            return fail(MCS_ERR_INVALID_ARGUMENT, "LDS and workgroup size leave no workgroup per CU");
        hipDeviceProp_t prop;
        MCS_HIP(hipGetDeviceProperties(&prop, b->device));
        const uint64_t slots = uint64_t(per_cu) * uint64_t(prop.multiProcessorCount);

        // Work items of whole tiles, at most item_words each and small
        // enough that there are four per workgroup slot; largest first.
        uint64_t total_words = 0;
        for (int32_t t : fused_ids) total_words += b->trees[size_t(t)].n_bp;
        const uint64_t item_words = std::max<uint64_t>(
            1, std::min<uint64_t>(cfg.item_words, total_words / std::max<uint64_t>(4 * slots, 1)));
        std::vector<mcs::WorkItem> items;
        std::vector<double> item_cost;
        for (size_t f = 0; f < fused_ids.size(); ++f) {
            const TreeHost &th = b->trees[size_t(fused_ids[f])];
            const uint64_t n_tiles = (th.n_bp + th.tile_words - 1) / th.tile_words;
            const uint64_t per_item = std::max<uint64_t>(1, item_words / th.tile_words);
            for (uint64_t first = 0; first < n_tiles; first += per_item) {
                const uint64_t n = std::min(per_item, n_tiles - first);
                items.push_back({uint32_t(f), uint32_t(first), uint32_t(n), 0});
                item_cost.push_back(cost_per_tile[f] * double(n));
            }
        }
        if (items.size() > 0xFFFFFFFFull) return fail(MCS_ERR_INVALID_ARGUMENT, "too many work items");
        std::vector<size_t> order(items.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return item_cost[x] > item_cost[y]; });
        std::vector<mcs::WorkItem> sorted;
        sorted.reserve(items.size());
        for (size_t i : order) sorted.push_back(items[i]);
        b->n_items = uint32_t(sorted.size());
        b->grid = uint32_t(std::min<uint64_t>(slots, b->n_items));

        mcs::FusedTree *d_trees = nullptr;
        mcs::EventDesc *d_ev = nullptr;
        mcs::FusedGate *d_gates = nullptr;
        uint32_t *d_layer = nullptr, *d_args = nullptr;
        mcs::WorkItem *d_items = nullptr;
        unsigned long long *d_ones = nullptr;
        MCS_HIP(upload(&d_trees, ft, b->stream));
        keep(d_trees);
        MCS_HIP(upload(&d_ev, ev, b->stream));
        keep(d_ev);
        MCS_HIP(upload(&d_gates, gates, b->stream));
        keep(d_gates);
        MCS_HIP(upload(&d_layer, layer_off, b->stream));
        keep(d_layer);
        MCS_HIP(upload(&d_args, args, b->stream));
        keep(d_args);
        MCS_HIP(upload(&d_items, sorted, b->stream));
        keep(d_items);
        MCS_HIP(hipMalloc(&d_ones, std::max<uint64_t>(b->n_ones, 1) * 8));
        keep(d_ones);
        MCS_HIP(hipMemsetAsync(d_ones, 0, std::max<uint64_t>(b->n_ones, 1) * 8, b->stream));
        b->device_bytes += ft.size() * sizeof(mcs::FusedTree) + ev.size() * sizeof(mcs::EventDesc) +
                           gates.size() * sizeof(mcs::FusedGate) + (layer_off.size() + args.size()) * 4 +
                           sorted.size() * sizeof(mcs::WorkItem) + b->n_ones * 8;
        mcs::FusedLaunch &L = b->launch;
        L.trees = d_trees;
        L.events = d_ev;
        L.gates = d_gates;
        L.layer_off = d_layer;
        L.args = d_args;
        L.items = d_items;
        L.ones = d_ones;
        L.state = b->d_state;
        L.n_items = b->n_items;
        L.key0 = cfg.key0;
        L.key1 = cfg.key1;
    }

    // ---- layered pipelines --------------------------------------------
    // Trees of one sample shape are dealt, costliest first, to the least
    // loaded of layered_streams pipelines (cost: Philox words plus gate
    // inputs, per word of the shape).
    // Each pipeline's words arena holds its largest tree, so more pipelines
    // take more memory: the count drops (to 1 at least) until every arena
    // fits in the device's free memory.
    std::map<uint64_t, std::vector<int32_t>> trees_of_shape;
    for (int32_t t = 0; t < n_trees; ++t)
        if (!b->trees[size_t(t)].fused) trees_of_shape[b->trees[size_t(t)].n_bp].push_back(t);
    auto cost = [&](int32_t t) {
        const TreeHost &th = b->trees[size_t(t)];
        return double(th.plan.n_events()) * 40.0 + double(th.n_nodes);
    };
    // Pipeline of every tree of a shape for a given pipeline count.
    auto deal = [&](const std::vector<int32_t> &ids, size_t n_pipes) {
        std::vector<double> load(n_pipes, 0.0);
        std::vector<size_t> pipe(ids.size());
        for (size_t i = 0; i < ids.size(); ++i) {
            pipe[i] = size_t(std::min_element(load.begin(), load.end()) - load.begin());
            load[pipe[i]] += cost(ids[i]);
        }
        return pipe;
    };
    auto arena_bytes = [&](const std::vector<int32_t> &ids, const std::vector<size_t> &pipe,
                           size_t n_pipes, uint64_t n_bp) {
        std::vector<uint64_t> max_ev(n_pipes, 0), max_gate(n_pipes, 0);
        for (size_t i = 0; i < ids.size(); ++i) {
            const TreeHost &th = b->trees[size_t(ids[i])];
            max_ev[pipe[i]] = std::max<uint64_t>(max_ev[pipe[i]], uint64_t(th.plan.n_events()));
            max_gate[pipe[i]] = std::max<uint64_t>(max_gate[pipe[i]], uint64_t(th.n_nodes - th.plan.n_events()));
        }
        const uint64_t cw = pick_chunk_words(n_bp);
        const uint64_t stride = (n_bp + cw - 1) / cw * cw;
        uint64_t total = 0;
        for (size_t k = 0; k < n_pipes; ++k) total += (2 * max_ev[k] + max_gate[k]) * stride * 8;
        return total;
    };
    size_t free_now = 0, total_now = 0;
    if (!trees_of_shape.empty()) MCS_HIP(hipMemGetInfo(&free_now, &total_now));
    size_t n_pipes_max = cfg.layered_streams;
    for (auto &[n_bp, ids] : trees_of_shape)
        std::stable_sort(ids.begin(), ids.end(), [&](int32_t x, int32_t y) { return cost(x) > cost(y); });
    for (; n_pipes_max > 1; --n_pipes_max) {
        uint64_t need = 0;
        for (auto &[n_bp, ids] : trees_of_shape) {
            const size_t n = std::min(n_pipes_max, ids.size());
            need += arena_bytes(ids, deal(ids, n), n, n_bp);
        }
        if (need + (uint64_t(256) << 20) * trees_of_shape.size() * n_pipes_max <= free_now) break;
    }
    for (auto &[n_bp, ids] : trees_of_shape) {
        const size_t n_pipes = std::min(n_pipes_max, ids.size());
        const std::vector<size_t> pipe = deal(ids, n_pipes);
        const size_t first = b->groups.size();
        for (size_t k = 0; k < n_pipes; ++k) {
            b->groups.emplace_back();
            b->groups.back().n_bp = n_bp;
        }
        for (size_t i = 0; i < ids.size(); ++i) {
            b->trees[size_t(ids[i])].group = int32_t(first + pipe[i]);
            b->groups[first + pipe[i]].trees.push_back(ids[i]);
        }
    }
    if (!b->groups.empty())
        MCS_HIP(hipEventCreateWithFlags(&b->began, hipEventDisableTiming));
    for (LayeredGroup &g : b->groups) {
        MCS_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        MCS_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    }
    for (LayeredGroup &g : b->groups) {
        mcs::LaunchGeometry &geo = g.geo;
        geo.chunk_words = pick_chunk_words(g.n_bp);
        geo.n_chunks = uint32_t((g.n_bp + geo.chunk_words - 1) / geo.chunk_words);
        geo.stride = uint64_t(geo.n_chunks) * geo.chunk_words;
        geo.n_bp = g.n_bp;
        uint64_t max_gate_rows = 0;
        for (int32_t t : g.trees) {
            const TreeHost &th = b->trees[size_t(t)];
            g.max_ev = std::max<uint32_t>(g.max_ev, uint32_t(th.plan.n_events()));
            max_gate_rows = std::max<uint64_t>(max_gate_rows, uint64_t(th.n_nodes - th.plan.n_events()));
        }
        const uint64_t rows = 2ull * g.max_ev + max_gate_rows;
        std::vector<mcs::GateDesc> gd;
        std::vector<uint32_t> a0, a1;
        std::vector<mcs::EventDesc> evs;
        uint64_t ones_total = 2ull * g.max_ev;  // leading pad (see enqueue_layered)
        for (int32_t t : g.trees) {
            TreeHost &th = b->trees[size_t(t)];
            const int32_t n_ev = th.plan.n_events();
            const uint32_t gate_rows = uint32_t(th.n_nodes - n_ev);
            LayeredGroup::Tree tr;
            const mcs::RowMap rm{{0u, g.max_ev}, 2u * g.max_ev, 2u * g.max_ev + gate_rows};
            // The fused LDS path keeps th.sch whole; the layered path fuses
            // single-consumer gates into their consumers' tasks.
            mcs::Schedule fsch = th.sch;
            mcs::fuse_single_consumer(fsch, th.plan, kBatchMaxFanIn);
            mcs::build_launch_plan(th.plan, fsch, th.tallied, rm, tr.lp);
            tr.gate_begin = gd.size();
            tr.args_begin = a0.size();
            tr.ev_begin = evs.size();
            tr.n_ev = uint32_t(n_ev);
            gd.insert(gd.end(), tr.lp.gates.begin(), tr.lp.gates.end());
            a0.insert(a0.end(), tr.lp.args[0].begin(), tr.lp.args[0].end());
            a1.insert(a1.end(), tr.lp.args[1].begin(), tr.lp.args[1].end());
            for (int32_t e = 0; e < n_ev; ++e)
                evs.push_back(mcs::make_event_desc(th.plan.event_pdag_index[e], uint32_t(e),
                                                   th.plan.event_p[e]));
            th.group_ones_begin = ones_total;
            ones_total += uint64_t(gate_rows) + uint64_t(n_ev);
            // Grid sizes of every launch of this tree with the largest
            // sampler slice it may carry.
            mcs::PassLaunch pl{};
            for (size_t l = 0; l + 1 < tr.lp.launch_begin.size(); ++l) {
                pl.n_gates = tr.lp.launch_begin[l + 1] - tr.lp.launch_begin[l];
                pl.samp_end = uint64_t(g.max_ev) * geo.n_chunks;
                if (mcs::pass_grid_blocks(pl, geo) * 256ull >= (uint64_t(1) << 32))
                    return fail(MCS_ERR_INVALID_ARGUMENT, "grid too large: reduce words per pass");
            }
            g.t.push_back(std::move(tr));
        }
        if (gd.size() > 0xFFFFFFFFull || a0.size() > 0xFFFFFFFFull)
            return fail(MCS_ERR_INVALID_ARGUMENT, "batch too large");
        size_t free_b = 0, total_b = 0;
        MCS_HIP(hipMemGetInfo(&free_b, &total_b));
        const uint64_t word_bytes = rows * geo.stride * 8;
        if (word_bytes + (uint64_t(256) << 20) > free_b)
            return fail(MCS_ERR_OUT_OF_MEMORY,
                        "layered pipeline needs " + std::to_string(word_bytes >> 20) + " MiB, " +
                            std::to_string(free_b >> 20) + " MiB free: reduce words per pass");
        MCS_HIP(hipMalloc(&g.d_words, word_bytes));
        keep(g.d_words);
        MCS_HIP(upload(&g.d_gates, gd, b->stream));
        keep(g.d_gates);
        MCS_HIP(upload(&g.d_args[0], a0, b->stream));
        keep(g.d_args[0]);
        MCS_HIP(upload(&g.d_args[1], a1, b->stream));
        keep(g.d_args[1]);
        MCS_HIP(upload(&g.d_events, evs, b->stream));
        keep(g.d_events);
        MCS_HIP(hipMalloc(&g.d_ones, ones_total * 8));
        keep(g.d_ones);
        MCS_HIP(hipMemsetAsync(g.d_ones, 0, ones_total * 8, b->stream));
        g.sum_stride = mcs::state_stride_for(geo.n_chunks);
        MCS_HIP(hipMalloc(&g.d_sum, rows * g.sum_stride));
        keep(g.d_sum);
        MCS_HIP(hipMemsetAsync(g.d_sum, mcs::kChunkMixed, rows * g.sum_stride, b->stream));
        b->device_bytes += word_bytes + gd.size() * sizeof(mcs::GateDesc) + (a0.size() + a1.size()) * 4 +
                           evs.size() * sizeof(mcs::EventDesc) + ones_total * 8 + rows * g.sum_stride;
    }
    MCS_HIP(hipStreamSynchronize(b->stream));
    *out = b.release();
    return MCS_OK;
}

}  // namespace

extern "C" {

void mcs_batch_config_default(mcs_batch_config *cfg) {
    if (!cfg) return;
    cfg->lds_bytes = 64u << 10;
    cfg->max_tile_words = 64;
    cfg->item_words = 4096;
    cfg->use_graph = 1;
    cfg->key0 = MCS_REFERENCE_KEY0;
    cfg->key1 = MCS_REFERENCE_KEY1;
    cfg->workgroup_lanes = 1024;
    cfg->min_fused_tile_words = 8;
    cfg->layered_streams = 2;
}

mcs_status mcs_batch_create(const mcs_pdag *pdags, int32_t n_trees, const mcs_sample_shape *shapes,
                            const int32_t *tally_offsets, const int32_t *tally_index,
                            const mcs_batch_config *cfg, mcs_batch **out) {
    return build_batch(pdags, n_trees, shapes, tally_offsets, tally_index, cfg, out);
}

mcs_status mcs_batch_destroy(mcs_batch *batch) {
    delete batch;
    return MCS_OK;
}

mcs_status mcs_batch_pass(mcs_batch *batch, uint32_t n_passes) {
    if (!batch) return fail(MCS_ERR_INVALID_ARGUMENT, "null batch");
    MCS_HIP(hipSetDevice(batch->device));
    for (uint32_t i = 0; i < n_passes; ++i) {
        const mcs_status st = enqueue_pass(batch);
        if (st != MCS_OK) return st;
    }
    return MCS_OK;
}

mcs_status mcs_batch_wait(mcs_batch *batch) {
    if (!batch) return fail(MCS_ERR_INVALID_ARGUMENT, "null batch");
    MCS_HIP(hipSetDevice(batch->device));
    MCS_HIP(hipStreamSynchronize(batch->stream));
    return MCS_OK;
}

mcs_status mcs_batch_collect(mcs_batch *batch, int32_t tree, mcs_tally *out, int32_t n_out) {
    if (!batch || !out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (tree < 0 || tree >= int32_t(batch->trees.size()))
        return fail(MCS_ERR_INVALID_ARGUMENT, "tree out of range");
    const TreeHost &th = batch->trees[size_t(tree)];
    if (n_out != int32_t(th.tally_nodes.size()))
        return fail(MCS_ERR_INVALID_ARGUMENT,
                    "collect needs exactly " + std::to_string(th.tally_nodes.size()) + " records");
    MCS_HIP(hipSetDevice(batch->device));
    const int32_t n_ev = th.plan.n_events();
    std::vector<unsigned long long> ones(size_t(th.n_nodes));
    if (th.fused) {
        MCS_HIP(hipMemcpyAsync(ones.data(), batch->launch.ones + th.ones_begin, ones.size() * 8,
                               hipMemcpyDeviceToHost, batch->stream));
    } else {
        // Layered counters: gates and partials (node - n_ev), then events.
        const LayeredGroup &g = batch->groups[size_t(th.group)];
        MCS_HIP(hipMemcpyAsync(ones.data(), g.d_ones + th.group_ones_begin, ones.size() * 8,
                               hipMemcpyDeviceToHost, batch->stream));
    }
    MCS_HIP(hipStreamSynchronize(batch->stream));
    for (size_t t = 0; t < th.tally_nodes.size(); ++t) {
        mcs_tally &r = out[t];
        const int32_t node = th.tally_nodes[t];
        r.index = th.tally_index[t];
        if (th.fused)
            r.num_one_bits = ones[size_t(th.slot_of_node[size_t(node)])];
        else
            r.num_one_bits = node < n_ev ? ones[size_t(th.n_nodes - n_ev + node)]
                                         : ones[size_t(node - n_ev)];
        r.total_bits = uint64_t(batch->iteration) * th.n_bp * 64ull;
        mcs_tally_point_estimates(&r);
    }
    return MCS_OK;
}

mcs_status mcs_batch_get_info(mcs_batch *batch, mcs_batch_info *out) {
    if (!batch || !out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    out->num_trees = int32_t(batch->trees.size());
    out->num_fused_trees = int32_t(batch->n_fused);
    out->num_layered_trees = int32_t(batch->trees.size()) - int32_t(batch->n_fused);
    uint32_t launches = 0;
    for (const LayeredGroup &g : batch->groups)
        for (const LayeredGroup::Tree &t : g.t) launches += uint32_t(t.lp.launch_begin.size() - 1);
    out->layered_launches = launches;
    out->num_items = batch->n_items;
    out->grid = batch->grid;
    out->lds_bytes = batch->lds_bytes;
    out->workgroup_lanes = batch->lanes;
    out->iteration = batch->iteration;
    out->device_bytes = batch->device_bytes;
    return MCS_OK;
}

mcs_status mcs_batch_tile_words(mcs_batch *batch, int32_t tree, uint32_t *out) {
    if (!batch || !out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (tree < 0 || tree >= int32_t(batch->trees.size()))
        return fail(MCS_ERR_INVALID_ARGUMENT, "tree out of range");
    *out = batch->trees[size_t(tree)].tile_words;
    return MCS_OK;
}

mcs_status mcs_batch_time_passes(mcs_batch *batch, uint32_t n_passes, double *kernel_ms) {
    if (!batch || !kernel_ms) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n_passes == 0) return fail(MCS_ERR_INVALID_ARGUMENT, "n_passes must be positive");
    MCS_HIP(hipSetDevice(batch->device));
    hipEvent_t a = nullptr, z = nullptr;
    MCS_HIP(hipEventCreate(&a));
    hipError_t e = hipEventCreate(&z);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        MCS_HIP(e);
    }
    mcs_status st = MCS_OK;
    e = hipEventRecord(a, batch->stream);
    for (uint32_t i = 0; i < n_passes && e == hipSuccess && st == MCS_OK; ++i) st = enqueue_pass(batch);
    if (e == hipSuccess && st == MCS_OK) e = hipEventRecord(z, batch->stream);
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
