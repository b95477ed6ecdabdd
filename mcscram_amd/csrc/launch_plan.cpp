// Gate descriptors and launch split for the layered kernel.
#include "launch_plan.h"

#include <algorithm>

#include "mcscram_amd.h"

namespace mcs {

void build_launch_plan(const Plan &plan, const Schedule &sch, const std::vector<uint8_t> &tallied,
                       const RowMap &rows, TreeLaunchPlan &out) {
    const int32_t n_ev = plan.n_events();
    auto arg_row = [&](int32_t node, uint32_t buf) -> uint32_t {
        return node < n_ev ? rows.be_row0[buf] + uint32_t(node)
                           : rows.gate_row0 + uint32_t(node - n_ev);
    };
    out = TreeLaunchPlan{};
    out.launch_begin.push_back(0);
    for (size_t l = 0; l < sch.layers.size(); ++l) {
        double cost = 0;
        auto add_gate = [&](uint32_t out_row, uint32_t type, int32_t k, uint32_t flags,
                            const std::vector<int32_t> &pos, const std::vector<int32_t> &neg) {
            GateDesc d{};
            d.out_row = out_row;
            d.arg_begin = uint32_t(out.args[0].size());
            const uint32_t n_in = uint32_t(pos.size() + neg.size());
            d.counts = n_in | uint32_t(pos.size()) << 16;
            d.type_k_flags = type | (uint32_t(uint8_t(k)) << 8) | flags;
            for (const auto *v : {&pos, &neg})
                for (int32_t i : *v) {
                    out.args[0].push_back(arg_row(i, 0));
                    out.args[1].push_back(arg_row(i, 1));
                }
            out.gates.push_back(d);
            cost += n_in + ((flags & kGateNoStore) ? 0 : 1);
        };
        for (const ExecGate &g : sch.layers[l]) {
            const bool t = g.out < plan.n_nodes() && tallied[size_t(g.out)];
            add_gate(rows.gate_row0 + uint32_t(g.out - n_ev), g.type, g.k, t ? kGateTallied : 0u,
                     g.pos_in, g.neg_in);
        }
        if (l == 0) {
            for (int32_t e = 0; e < n_ev; ++e)
                if (tallied[size_t(e)])
                    add_gate(rows.event_tally_slot0 + uint32_t(e), MCS_NULL, 0,
                             kGateTallied | kGateNoStore, {e}, {});
        }
        out.launch_begin.push_back(uint32_t(out.gates.size()));
        out.launch_cost.push_back(cost);
    }
}

std::vector<uint64_t> split_in_proportion(const std::vector<double> &cost, uint64_t total) {
    double sum = 0;
    for (double c : cost) sum += c;
    std::vector<uint64_t> b{0};
    double acc = 0;
    for (size_t l = 0; l < cost.size(); ++l) {
        acc += cost[l];
        const uint64_t end = (l + 1 == cost.size() || sum <= 0)
                                 ? total
                                 : uint64_t(double(total) * (acc / sum));
        b.push_back(std::max(end, b.back()));
    }
    return b;
}

}  // namespace mcs
