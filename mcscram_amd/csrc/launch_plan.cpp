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
    auto is_tallied = [&](int32_t node) { return node < plan.n_nodes() && tallied[size_t(node)]; };
    out = TreeLaunchPlan{};
    out.launch_begin.push_back(0);
    // Fused gates: descriptors and argument rows collected apart, appended
    // after every task descriptor; inline references are patched then.
    std::vector<GateDesc> kids;
    std::vector<uint32_t> kid_args[2];
    std::vector<std::pair<size_t, uint32_t>> patch;  // (position in args, kid)
    auto desc = [](uint32_t out_row, uint32_t type, int32_t k, uint32_t flags, uint32_t arg_begin,
                   uint32_t n_in, uint32_t n_pos) {
        GateDesc d{};
        d.out_row = out_row;
        d.arg_begin = arg_begin;
        d.counts = n_in | n_pos << 16;
        d.type_k_flags = type | (uint32_t(uint8_t(k)) << 8) | flags;
        return d;
    };
    for (size_t l = 0; l < sch.layers.size(); ++l) {
        double cost = 0;
        auto add_gate = [&](uint32_t out_row, uint32_t type, int32_t k, uint32_t flags,
                            const std::vector<int32_t> &pos, const std::vector<int32_t> &neg,
                            const std::vector<ExecGate> &inl) {
            const uint32_t arg_begin = uint32_t(out.args[0].size());
            const uint32_t n_in = uint32_t(pos.size() + neg.size());
            const size_t kids_before = kids.size();
            for (const auto *v : {&pos, &neg})
                for (int32_t i : *v) {
                    const ExecGate *ch = nullptr;
                    for (const ExecGate &c : inl)
                        if (c.out == i) ch = &c;
                    if (!ch) {
                        out.args[0].push_back(arg_row(i, 0));
                        out.args[1].push_back(arg_row(i, 1));
                        cost += 1;
                        continue;
                    }
                    const uint32_t kid = uint32_t(kids.size());
                    const uint32_t kid_in = uint32_t(ch->pos_in.size() + ch->neg_in.size());
                    kids.push_back(desc(rows.gate_row0 + uint32_t(ch->out - n_ev), ch->type, ch->k,
                                        kGateNoStore | (is_tallied(ch->out) ? kGateTallied : 0u),
                                        uint32_t(kid_args[0].size()), kid_in,
                                        uint32_t(ch->pos_in.size())));
                    for (const auto *w : {&ch->pos_in, &ch->neg_in})
                        for (int32_t j : *w) {
                            kid_args[0].push_back(arg_row(j, 0));
                            kid_args[1].push_back(arg_row(j, 1));
                        }
                    cost += kid_in;
                    patch.push_back({out.args[0].size(), kid});
                    out.args[0].push_back(0);
                    out.args[1].push_back(0);
                }
            if (kids.size() > kids_before) flags |= kGateHasFused;
            out.gates.push_back(desc(out_row, type, k, flags, arg_begin, n_in,
                                     uint32_t(pos.size())));
            cost += (flags & kGateNoStore) ? 0 : 1;
        };
        for (const ExecGate &g : sch.layers[l])
            add_gate(rows.gate_row0 + uint32_t(g.out - n_ev), g.type, g.k,
                     is_tallied(g.out) ? kGateTallied : 0u, g.pos_in, g.neg_in, g.inl);
        if (l == 0) {
            for (int32_t e = 0; e < n_ev; ++e)
                if (tallied[size_t(e)])
                    add_gate(rows.event_tally_slot0 + uint32_t(e), MCS_NULL, 0,
                             kGateTallied | kGateNoStore, {e}, {}, {});
        }
        out.launch_begin.push_back(uint32_t(out.gates.size()));
        out.launch_cost.push_back(cost);
    }
    const uint32_t kid0 = uint32_t(out.gates.size());
    const uint32_t arg0 = uint32_t(out.args[0].size());
    for (GateDesc d : kids) {
        d.arg_begin += arg0;
        out.gates.push_back(d);
    }
    for (uint32_t b = 0; b < 2; ++b)
        out.args[b].insert(out.args[b].end(), kid_args[b].begin(), kid_args[b].end());
    for (const auto &[pos, kid] : patch)
        out.args[0][pos] = out.args[1][pos] = kInlineArg | (kid0 + kid);
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
