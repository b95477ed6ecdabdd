// Batch engine: many independent trees, one persistent whole-tree-in-LDS
// launch per pass (kernel in fused.hip).
//
// Host side: per tree the reference's plan (event order, layers) and the
// execution schedule, a tile width W chosen so the tree's W words per node
// fit the LDS budget, and work items of whole tiles ordered largest-cost
// first so the persistent workgroups finish together. A pass is a tiny
// kernel that resets the work queue and advances the iteration, then the
// persistent launch; with use_graph the pair is captured once as a hipGraph
// and replayed per pass.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "fused.h"
#include "kernels.h"
#include "layering.h"
#include "mcscram_amd.h"
#include "status.h"

using mcs::fail;

namespace {

constexpr uint32_t kBatchMaxFanIn = 64;

struct TreeHost {
    mcs::Plan plan;
    std::vector<int32_t> slot_of_node;  // plan node id -> LDS node slot
    std::vector<int32_t> tally_nodes;   // plan node ids, first-appearance order
    std::vector<int32_t> tally_index;   // pdag indices
    uint32_t ones_begin = 0;
    uint32_t tile_words = 1;
    uint64_t n_bp = 0;
};

template <typename T>
hipError_t upload(T **dst, const std::vector<T> &src, hipStream_t s) {
    const size_t n = std::max<size_t>(src.size(), 1);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), n * sizeof(T));
    if (e != hipSuccess || src.empty()) return e;
    return hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

}  // namespace

struct mcs_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    mcs_batch_config cfg{};
    std::vector<TreeHost> trees;
    uint32_t iteration = 0;
    uint32_t n_items = 0;
    uint32_t grid = 0;
    uint32_t lds_bytes = 0;
    uint32_t lanes = 0;
    uint64_t n_ones = 0;
    uint64_t device_bytes = 0;
    mcs::FusedLaunch launch{};
    std::vector<void *> allocations;

    ~mcs_batch() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : allocations) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

mcs_status enqueue_pass(mcs_batch *b) {
    if (b->iteration == 0xFFFFFFFFu) return fail(MCS_ERR_STATE, "iteration counter exhausted");
    if (b->graph_exec) {
        MCS_HIP(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
        MCS_HIP(mcs::launch_fused_begin_pass(b->stream, b->launch.state));
        MCS_HIP(mcs::launch_fused(b->stream, b->launch, b->grid, b->lanes, b->lds_bytes));
    }
    ++b->iteration;
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
    if (cfg.max_tile_words > 64 || (cfg.max_tile_words & (cfg.max_tile_words - 1)))
        return fail(MCS_ERR_INVALID_ARGUMENT, "max_tile_words must be a power of two <= 64");
    if (cfg.lds_bytes > mcs::kMaxFusedLds)
        return fail(MCS_ERR_INVALID_ARGUMENT, "lds_bytes exceeds the 160 KiB of a CU");
    if (cfg.workgroup_lanes > 1024 || cfg.workgroup_lanes % 64)
        return fail(MCS_ERR_INVALID_ARGUMENT, "workgroup_lanes must be a multiple of 64 up to 1024");
    b->lanes = cfg.workgroup_lanes;

    std::vector<mcs::FusedTree> ft;
    std::vector<mcs::EventDesc> ev;
    std::vector<mcs::FusedGate> gates;
    std::vector<uint32_t> layer_off, args;
    struct Cost {
        double per_tile;
    };
    std::vector<Cost> cost;
    b->trees.resize(size_t(n_trees));
    std::vector<mcs::Schedule> schedules(static_cast<size_t>(n_trees));
    // Plans first: the launch's LDS budget must fit the largest tree.
    uint64_t budget = cfg.lds_bytes;
    for (int32_t t = 0; t < n_trees; ++t) {
        TreeHost &th = b->trees[size_t(t)];
        std::string err;
        if (!mcs::build_plan(pdags[t], th.plan, err))
            return fail(MCS_ERR_GRAPH, "tree " + std::to_string(t) + ": " + err);
        schedules[size_t(t)] = mcs::build_schedule(th.plan, kBatchMaxFanIn);
        const uint32_t n_nodes = uint32_t(th.plan.n_nodes() + schedules[size_t(t)].n_aux);
        budget = std::max(budget, mcs::fused_lds_bytes(n_nodes, 1));
        if (budget > mcs::kMaxFusedLds)
            return fail(MCS_ERR_INVALID_ARGUMENT,
                        "tree " + std::to_string(t) + " has " + std::to_string(n_nodes) +
                            " nodes: too large for one LDS tile; use mcs_engine_create");
    }
    for (int32_t t = 0; t < n_trees; ++t) {
        TreeHost &th = b->trees[size_t(t)];
        const mcs_sample_shape &sh = shapes[t];
        th.n_bp = sh.batch_size * sh.bitpacks_per_batch;
        if (sh.batch_size == 0 || sh.bitpacks_per_batch == 0 ||
            th.n_bp / sh.batch_size != sh.bitpacks_per_batch || th.n_bp > (uint64_t(1) << 32))
            return fail(MCS_ERR_INVALID_ARGUMENT,
                        "tree " + std::to_string(t) + ": sample shape must be positive and at most 2^32 words");
        const mcs::Plan &plan = th.plan;
        const mcs::Schedule &sch = schedules[size_t(t)];
        const int32_t n_ev = plan.n_events();
        const int32_t n_nodes = plan.n_nodes() + sch.n_aux;
        if (sch.layers.size() > mcs::kMaxFusedLayers)
            return fail(MCS_ERR_INVALID_ARGUMENT,
                        "tree " + std::to_string(t) + " is deeper than " +
                            std::to_string(mcs::kMaxFusedLayers) + " layers; use mcs_engine_create");

        // Tallies.
        std::unordered_map<int32_t, int32_t> node_of_index;
        for (int32_t n = 0; n < plan.n_nodes(); ++n) node_of_index[plan.node_pdag_index[n]] = n;
        std::vector<uint8_t> tallied(size_t(plan.n_nodes()), 0);
        const int32_t t0 = tally_index ? tally_offsets[t] : 0;
        const int32_t t1 = tally_index ? tally_offsets[t + 1] : 0;
        if (t1 < t0) return fail(MCS_ERR_INVALID_ARGUMENT, "tally offsets are not monotone");
        std::vector<int32_t> want(tally_index ? tally_index + t0 : nullptr,
                                  tally_index ? tally_index + t1 : nullptr);
        if (want.empty()) want.push_back(pdags[t].root);
        for (int32_t idx : want) {
            auto it = node_of_index.find(idx);
            if (it == node_of_index.end())
                return fail(MCS_ERR_GRAPH, "tree " + std::to_string(t) + ": tally index " +
                                               std::to_string(idx) + " is not reachable from the root");
            if (tallied[size_t(it->second)]) continue;
            tallied[size_t(it->second)] = 1;
            th.tally_nodes.push_back(it->second);
            th.tally_index.push_back(idx);
        }

        // Tile width: the widest power of two whose tile fits the budget.
        uint32_t w = cfg.max_tile_words;
        while (w > 1 && mcs::fused_lds_bytes(uint32_t(n_nodes), w) > budget) w >>= 1;
        th.tile_words = w;
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
            if (tallied[size_t(e)]) x.mode |= mcs::kEventTallied;
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
                const bool tal = g.out < plan.n_nodes() && tallied[size_t(g.out)];
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
        cost.push_back({double(w) * (double(n_ev) * 40.0 + double(fan) + double(local))});
        if (args.size() > 0xFFFFFFFFull || gates.size() > 0xFFFFFFFFull)
            return fail(MCS_ERR_INVALID_ARGUMENT, "batch too large");
    }

    // Device and occupancy: the grid is every workgroup slot the LDS and
    // register use allow.
    MCS_HIP(hipGetDevice(&b->device));
    MCS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int per_cu = 0;
    MCS_HIP(mcs::fused_max_workgroups_per_cu(b->lanes, b->lds_bytes, &per_cu));
    if (per_cu <= 0)
        return fail(MCS_ERR_INVALID_ARGUMENT, "LDS and workgroup size leave no workgroup per CU");
    hipDeviceProp_t prop;
    MCS_HIP(hipGetDeviceProperties(&prop, b->device));
    const uint64_t slots = uint64_t(per_cu) * uint64_t(prop.multiProcessorCount);

    // Work items of whole tiles, at most item_words each, and small enough
    // that there are four per workgroup slot; largest first.
    uint64_t total_words = 0;
    for (const TreeHost &th : b->trees) total_words += th.n_bp;
    const uint64_t item_words = std::max<uint64_t>(
        1, std::min<uint64_t>(cfg.item_words, total_words / std::max<uint64_t>(4 * slots, 1)));
    std::vector<mcs::WorkItem> items;
    std::vector<double> item_cost;
    for (int32_t t = 0; t < n_trees; ++t) {
        const TreeHost &th = b->trees[size_t(t)];
        const uint64_t n_tiles = (th.n_bp + th.tile_words - 1) / th.tile_words;
        const uint64_t per_item = std::max<uint64_t>(1, item_words / th.tile_words);
        for (uint64_t first = 0; first < n_tiles; first += per_item) {
            const uint64_t n = std::min(per_item, n_tiles - first);
            items.push_back({uint32_t(t), uint32_t(first), uint32_t(n), 0});
            item_cost.push_back(cost[size_t(t)].per_tile * double(n));
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
    uint32_t *d_layer = nullptr, *d_args = nullptr, *d_state = nullptr;
    mcs::WorkItem *d_items = nullptr;
    unsigned long long *d_ones = nullptr;
    auto keep = [&](void *p) { b->allocations.push_back(p); };
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
    MCS_HIP(hipMalloc(&d_state, 2 * sizeof(uint32_t)));
    keep(d_state);
    MCS_HIP(hipMemsetAsync(d_ones, 0, std::max<uint64_t>(b->n_ones, 1) * 8, b->stream));
    MCS_HIP(hipMemsetAsync(d_state, 0, 2 * sizeof(uint32_t), b->stream));
    MCS_HIP(hipStreamSynchronize(b->stream));
    b->device_bytes = ft.size() * sizeof(mcs::FusedTree) + ev.size() * sizeof(mcs::EventDesc) +
                      gates.size() * sizeof(mcs::FusedGate) + (layer_off.size() + args.size()) * 4 +
                      sorted.size() * sizeof(mcs::WorkItem) + b->n_ones * 8 + 8;

    mcs::FusedLaunch &L = b->launch;
    L.trees = d_trees;
    L.events = d_ev;
    L.gates = d_gates;
    L.layer_off = d_layer;
    L.args = d_args;
    L.items = d_items;
    L.ones = d_ones;
    L.state = d_state;
    L.n_items = b->n_items;
    L.key0 = cfg.key0;
    L.key1 = cfg.key1;

    if (cfg.use_graph) {
        MCS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e1 = mcs::launch_fused_begin_pass(b->stream, L.state);
        hipError_t e2 = mcs::launch_fused(b->stream, L, b->grid, b->lanes, b->lds_bytes);
        hipError_t e3 = hipStreamEndCapture(b->stream, &b->graph);
        MCS_HIP(e1);
        MCS_HIP(e2);
        MCS_HIP(e3);
        MCS_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    }
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
    std::vector<unsigned long long> ones(th.slot_of_node.size());
    MCS_HIP(hipMemcpyAsync(ones.data(), batch->launch.ones + th.ones_begin, ones.size() * 8,
                           hipMemcpyDeviceToHost, batch->stream));
    MCS_HIP(hipStreamSynchronize(batch->stream));
    for (size_t t = 0; t < th.tally_nodes.size(); ++t) {
        mcs_tally &r = out[t];
        r.index = th.tally_index[t];
        r.num_one_bits = ones[size_t(th.slot_of_node[size_t(th.tally_nodes[t])])];
        r.total_bits = uint64_t(batch->iteration) * th.n_bp * 64ull;
        mcs_tally_point_estimates(&r);
    }
    return MCS_OK;
}

mcs_status mcs_batch_get_info(mcs_batch *batch, mcs_batch_info *out) {
    if (!batch || !out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    out->num_trees = int32_t(batch->trees.size());
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
