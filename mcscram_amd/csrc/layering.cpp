// Layered topological order and event positions, as the reference's
// layer_manager computes them (see layering.h for the file:line map).
#include "layering.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <unordered_map>

// The reference orders nodes with libstdc++'s std::sort (two unstable sorts
// below). Its order among equal keys fixes every basic event's position, and
// so its Philox counters: built against another C++ library the engine would
// sample other streams, so such a build is refused.
#ifndef __GLIBCXX__
#error "layering.cpp needs libstdc++: its std::sort fixes the reference's event order"
#endif

namespace mcs {
namespace {

struct NodeRef {
    bool is_var;
    int32_t slot;
};

// One element of the sequence the reference sorts (its Node pointers).
struct SortItem {
    int32_t depth;
    bool is_var;
    uint8_t type;
    int32_t slot;
};

}  // namespace

bool build_plan(const mcs_pdag &pd, Plan &plan, std::string &error) {
    if (pd.n_vars < 0 || pd.n_gates <= 0) {
        error = "a PDAG needs at least one gate and a non-negative variable count";
        return false;
    }
    if ((pd.n_vars && (!pd.var_index || !pd.var_p)) || !pd.gate_index || !pd.gate_type ||
        !pd.gate_k || !pd.arg_offsets || (pd.arg_offsets[pd.n_gates] && !pd.args)) {
        error = "null PDAG array";
        return false;
    }
    std::unordered_map<int32_t, NodeRef> by_index;
    by_index.reserve(size_t(pd.n_vars) + size_t(pd.n_gates));
    for (int32_t i = 0; i < pd.n_vars; ++i) {
        const double p = pd.var_p[i];
        if (!(p >= 0.0 && p <= 1.0)) {
            error = "probability of variable " + std::to_string(pd.var_index[i]) + " outside [0, 1]";
            return false;
        }
        if (pd.var_index[i] <= 0 || !by_index.emplace(pd.var_index[i], NodeRef{true, i}).second) {
            error = "duplicate or non-positive variable index " + std::to_string(pd.var_index[i]);
            return false;
        }
    }
    for (int32_t i = 0; i < pd.n_gates; ++i) {
        if (pd.gate_type[i] > MCS_NULL) {
            error = "unknown connective for gate " + std::to_string(pd.gate_index[i]);
            return false;
        }
        if (pd.gate_index[i] <= 0 || !by_index.emplace(pd.gate_index[i], NodeRef{false, i}).second) {
            error = "duplicate or non-positive gate index " + std::to_string(pd.gate_index[i]);
            return false;
        }
        if (pd.arg_offsets[i + 1] < pd.arg_offsets[i]) {
            error = "argument offsets are not monotone";
            return false;
        }
        if (pd.arg_offsets[i + 1] - pd.arg_offsets[i] > 0xFFFF) {
            error = "gate " + std::to_string(pd.gate_index[i]) + " has more than 65535 arguments";
            return false;
        }
    }
    if (pd.arg_offsets[0] != 0) {
        error = "argument offsets must start at 0";
        return false;
    }
    std::vector<std::vector<NodeRef>> args(pd.n_gates);
    for (int32_t i = 0; i < pd.n_gates; ++i) {
        for (int64_t a = pd.arg_offsets[i]; a < pd.arg_offsets[i + 1]; ++a) {
            const int32_t s = pd.args[a];
            auto it = by_index.find(s < 0 ? -s : s);
            if (s == 0 || it == by_index.end()) {
                error = "gate " + std::to_string(pd.gate_index[i]) + " has unknown argument " +
                        std::to_string(s);
                return false;
            }
            args[i].push_back(it->second);
        }
    }
    auto root_it = by_index.find(pd.root);
    if (root_it == by_index.end() || root_it->second.is_var) {
        error = "root " + std::to_string(pd.root) + " is not a gate";
        return false;
    }
    const int32_t root = root_it->second.slot;

    // Depth of every gate reachable from the root (iterative post-order with
    // cycle detection).
    std::vector<int32_t> depth(pd.n_gates, -1);
    std::vector<uint8_t> state(pd.n_gates, 0);  // 0 new, 1 open, 2 done
    {
        std::vector<std::pair<int32_t, size_t>> stack;
        stack.push_back({root, 0});
        state[root] = 1;
        while (!stack.empty()) {
            auto &[g, pos] = stack.back();
            if (pos < args[g].size()) {
                const NodeRef r = args[g][pos++];
                if (r.is_var) continue;
                if (state[r.slot] == 1) {
                    error = "the PDAG has a cycle through gate " + std::to_string(pd.gate_index[r.slot]);
                    return false;
                }
                if (state[r.slot] == 0) {
                    state[r.slot] = 1;
                    stack.push_back({r.slot, 0});
                }
                continue;
            }
            int32_t m = -1;
            for (const NodeRef &r : args[g]) m = std::max(m, r.is_var ? 0 : depth[r.slot]);
            depth[g] = m + 1;
            state[g] = 2;
            stack.pop_back();
        }
    }

    // DFS gather, in the order of the reference's recursion: the gate, its
    // gate arguments recursively, then its not yet visited variables.
    std::vector<SortItem> seq;
    {
        std::vector<uint8_t> gvis(pd.n_gates, 0), vvis(pd.n_vars, 0);
        std::vector<std::pair<int32_t, size_t>> stack;
        gvis[root] = 1;
        seq.push_back({depth[root], false, pd.gate_type[root], root});
        stack.push_back({root, 0});
        while (!stack.empty()) {
            auto &[g, pos] = stack.back();
            bool descended = false;
            while (pos < args[g].size()) {
                const NodeRef r = args[g][pos++];
                if (r.is_var || gvis[r.slot]) continue;
                gvis[r.slot] = 1;
                seq.push_back({depth[r.slot], false, pd.gate_type[r.slot], r.slot});
                stack.push_back({r.slot, 0});
                descended = true;
                break;
            }
            if (descended) continue;
            for (const NodeRef &r : args[g]) {
                if (r.is_var && !vvis[r.slot]) {
                    vvis[r.slot] = 1;
                    seq.push_back({0, true, 0, r.slot});
                }
            }
            stack.pop_back();
        }
    }

    // The reference's two unstable sorts, with its comparators.
    std::sort(seq.begin(), seq.end(),
              [](const SortItem &a, const SortItem &b) { return a.depth < b.depth; });
    std::vector<std::vector<SortItem>> layers(size_t(seq.back().depth) + 1);
    for (const SortItem &n : seq) layers[n.depth].push_back(n);
    for (auto &layer : layers) {
        std::sort(layer.begin(), layer.end(), [](const SortItem &l, const SortItem &r) {
            if (l.is_var && !r.is_var) return true;
            if (!l.is_var && r.is_var) return false;
            if (l.is_var && r.is_var) return false;
            return l.type < r.type;
        });
    }

    plan = Plan{};
    std::vector<int32_t> node_of_var(pd.n_vars, -1), node_of_gate(pd.n_gates, -1);
    for (const SortItem &n : layers[0]) {
        if (!n.is_var) {
            error = "internal: gate at depth 0";
            return false;
        }
        node_of_var[n.slot] = plan.n_events();
        plan.event_pdag_index.push_back(pd.var_index[n.slot]);
        plan.event_p.push_back(pd.var_p[n.slot]);
    }
    const int32_t n_events = plan.n_events();
    int32_t next = n_events;
    plan.layer_begin.push_back(0);
    for (size_t l = 1; l < layers.size(); ++l) {
        for (const SortItem &n : layers[l]) node_of_gate[n.slot] = next++;
        plan.layer_begin.push_back(uint32_t(next - n_events));
    }
    plan.gates.reserve(size_t(next - n_events));
    for (size_t l = 1; l < layers.size(); ++l) {
        for (const SortItem &n : layers[l]) {
            PlannedGate pg;
            pg.pdag_index = pd.gate_index[n.slot];
            pg.type = pd.gate_type[n.slot];
            pg.k = pd.gate_k[n.slot];
            pg.layer = int32_t(l);
            for (int kind = 0; kind < 2; ++kind) {  // variables first, then gates
                for (int64_t a = pd.arg_offsets[n.slot]; a < pd.arg_offsets[n.slot + 1]; ++a) {
                    const NodeRef r = args[n.slot][size_t(a - pd.arg_offsets[n.slot])];
                    if (r.is_var != (kind == 0)) continue;
                    const int32_t id = r.is_var ? node_of_var[r.slot] : node_of_gate[r.slot];
                    (pd.args[a] < 0 ? pg.neg_in : pg.pos_in).push_back(id);
                }
            }
            plan.gates.push_back(std::move(pg));
        }
    }
    plan.node_pdag_index = plan.event_pdag_index;
    for (const PlannedGate &g : plan.gates) plan.node_pdag_index.push_back(g.pdag_index);
    plan.root_node = node_of_gate[root];
    return true;
}

Schedule build_schedule(const Plan &plan, uint32_t max_fan_in) {
    Schedule sch;
    const int32_t n_nodes = plan.n_nodes();
    std::vector<int32_t> depth(size_t(n_nodes), 0);  // events 0; grows with aux nodes
    auto place = [&](ExecGate &&g) {
        int32_t d = 0;
        for (const auto *v : {&g.pos_in, &g.neg_in})
            for (int32_t i : *v) d = std::max(d, depth[size_t(i)]);
        d += 1;
        if (size_t(g.out) >= depth.size()) depth.resize(size_t(g.out) + 1, 0);
        depth[size_t(g.out)] = d;
        if (sch.layers.size() < size_t(d)) sch.layers.resize(size_t(d));
        sch.layers[size_t(d) - 1].push_back(std::move(g));
    };
    const uint32_t cap = std::max<uint32_t>(max_fan_in, 2);
    for (size_t gi = 0; gi < plan.gates.size(); ++gi) {
        const PlannedGate &pg = plan.gates[gi];
        const int32_t node = plan.n_events() + int32_t(gi);
        const size_t n_in = pg.pos_in.size() + pg.neg_in.size();
        uint8_t base = 0xFF;
        if (pg.type == MCS_AND || pg.type == MCS_NAND) base = MCS_AND;
        if (pg.type == MCS_OR || pg.type == MCS_NOR) base = MCS_OR;
        if (pg.type == MCS_XOR) base = MCS_XOR;
        if (base == 0xFF || n_in <= cap) {
            place(ExecGate{node, pg.type, pg.k, pg.pos_in, pg.neg_in, {}});
            continue;
        }
        // Literals (node, negated), shallow first; stable so equal depths keep
        // the plan's order.
        std::vector<std::pair<int32_t, bool>> lits;
        for (int32_t i : pg.pos_in) lits.push_back({i, false});
        for (int32_t i : pg.neg_in) lits.push_back({i, true});
        std::stable_sort(lits.begin(), lits.end(), [&](const auto &a, const auto &b) {
            return depth[size_t(a.first)] < depth[size_t(b.first)];
        });
        while (lits.size() > cap) {
            std::vector<std::pair<int32_t, bool>> next;
            for (size_t b = 0; b < lits.size(); b += cap) {
                ExecGate part{n_nodes + sch.n_aux++, base, 0, {}, {}, {}};
                for (size_t j = b; j < std::min(lits.size(), b + cap); ++j)
                    (lits[j].second ? part.neg_in : part.pos_in).push_back(lits[j].first);
                next.push_back({part.out, false});
                place(std::move(part));
            }
            lits = std::move(next);
        }
        ExecGate top{node, pg.type, pg.k, {}, {}, {}};
        for (const auto &l : lits) (l.second ? top.neg_in : top.pos_in).push_back(l.first);
        place(std::move(top));
    }
    return sch;
}


void fuse_single_consumer(Schedule &sch, const Plan &plan, uint32_t max_rows) {
    const int32_t n_nodes = plan.n_nodes();
    const size_t n_ids = size_t(n_nodes) + size_t(sch.n_aux);
    std::vector<int32_t> n_cons(n_ids, 0), consumer(n_ids, -1);
    std::vector<ExecGate *> by_out(n_ids, nullptr);
    std::vector<uint32_t> rows(n_ids, 0);
    for (auto &layer : sch.layers)
        for (ExecGate &g : layer) {
            by_out[size_t(g.out)] = &g;
            rows[size_t(g.out)] = uint32_t(g.pos_in.size() + g.neg_in.size());
            for (const auto *v : {&g.pos_in, &g.neg_in})
                for (int32_t i : *v) {
                    ++n_cons[size_t(i)];
                    consumer[size_t(i)] = g.out;
                }
        }
    std::vector<uint8_t> removed(n_ids, 0), has_child(n_ids, 0);
    for (auto &layer : sch.layers)
        for (ExecGate &g : layer) {
            const int32_t id = g.out;
            if (id >= n_nodes || id < plan.n_events() || id == plan.root_node) continue;
            if (n_cons[size_t(id)] != 1 || has_child[size_t(id)] || g.type == MCS_ATLEAST) continue;
            ExecGate *c = by_out[size_t(consumer[size_t(id)])];
            if (!c) continue;
            const uint32_t c_in = uint32_t(c->pos_in.size() + c->neg_in.size());
            const bool bitwise = c->type == MCS_AND || c->type == MCS_OR || c->type == MCS_XOR ||
                                 c->type == MCS_NAND || c->type == MCS_NOR;
            const bool small_atleast = c->type == MCS_ATLEAST && c_in <= 15;
            if (!bitwise && !small_atleast) continue;
            if (rows[size_t(c->out)] - 1 + rows[size_t(id)] > max_rows) continue;
            c->inl.push_back(g);
            rows[size_t(c->out)] += rows[size_t(id)] - 1;
            has_child[size_t(c->out)] = 1;
            removed[size_t(id)] = 1;
        }
    // Rebuild the layers: a task's depth is one more than the deepest row it
    // reads, its fused gates' inputs included.
    std::vector<ExecGate> all;
    for (auto &layer : sch.layers)
        for (ExecGate &g : layer)
            if (!removed[size_t(g.out)]) all.push_back(std::move(g));
    std::vector<int32_t> depth(n_ids, 0);
    std::vector<std::vector<ExecGate>> layers;
    for (ExecGate &g : all) {
        int32_t d = 0;
        auto row_depth = [&](int32_t i) { d = std::max(d, depth[size_t(i)]); };
        for (const auto *v : {&g.pos_in, &g.neg_in})
            for (int32_t i : *v) {
                const ExecGate *ch = nullptr;
                for (const ExecGate &c : g.inl)
                    if (c.out == i) ch = &c;
                if (!ch) {
                    row_depth(i);
                    continue;
                }
                for (const auto *w : {&ch->pos_in, &ch->neg_in})
                    for (int32_t j : *w) row_depth(j);
            }
        depth[size_t(g.out)] = d + 1;
        for (const ExecGate &c : g.inl) depth[size_t(c.out)] = d + 1;
        if (layers.size() < size_t(d) + 1) layers.resize(size_t(d) + 1);
        layers[size_t(d)].push_back(std::move(g));
    }
    layers.erase(std::remove_if(layers.begin(), layers.end(),
                                [](const std::vector<ExecGate> &l) { return l.empty(); }),
                 layers.end());
    sch.layers = std::move(layers);
}

}  // namespace mcs
