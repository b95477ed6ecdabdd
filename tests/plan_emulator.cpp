// CPU emulation of the layered pass over the engine's own launch plan.
//
// TEST INFRASTRUCTURE ONLY. Builds the execution schedule and launch plan
// with the engine's host code (layering.cpp, launch_plan.cpp), then runs the
// launches on the CPU with the gate semantics the kernel implements (gate
// folds, NOT/NULL on the last input, 8-plane ATLEAST count, fused gates
// evaluated inside their consumer's task and tallied there), over basic-event
// words supplied by the caller. Every index the kernel would use is checked:
// rows in range and written by an earlier launch, argument ranges, fused
// references only where the kernel accepts them. The pass runs once from each
// of the two event buffers the engine alternates between (rows 0.. and
// n_events..), and both runs must agree. Output: each gate's words (fused
// gates recomputed as mcs_engine_read_words does) and every node's tally.
//
// Usage: plan_emulator PDAG_FILE EVENT_WORDS_FILE FUSE SPLIT_FAN_IN OUT_FILE
//        plan_emulator --split   (checks the launch's workgroup-to-task map)
//   PDAG_FILE: i32 n_vars, i32 n_gates, i64 n_args, i32 root, then the
//   mcs_pdag arrays in order; EVENT_WORDS_FILE: n_events x 64 uint64 words in
//   event_idx order. OUT_FILE: gate words (plan order, 64 each), then one
//   uint64 tally per plan node, then the plan's PDAG index per node.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.h"
#include "launch_plan.h"
#include "layering.h"

using namespace mcs;

namespace {

constexpr int kWords = 64;

std::vector<char> slurp(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    std::vector<char> b;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + n);
    std::fclose(f);
    return b;
}

struct Emulator {
    const TreeLaunchPlan &lp;
    const std::vector<uint32_t> &args;
    uint64_t n_rows;
    std::vector<uint64_t> words;
    std::vector<uint8_t> written;
    std::vector<uint64_t> ones;

    uint64_t row_word(uint32_t row, int w) const {
        if (row >= n_rows) throw std::runtime_error("row out of range: " + std::to_string(row));
        if (!written[row]) throw std::runtime_error("read of a row no earlier launch wrote: " + std::to_string(row));
        return words[uint64_t(row) * kWords + uint64_t(w)];
    }

    uint64_t eval(const GateDesc &d, int w, bool task) {
        const uint32_t n_in = d.counts & 0xFFFFu, n_pos = d.counts >> 16;
        const uint32_t type = d.type_k_flags & 0xFFu, k8 = (d.type_k_flags >> 8) & 0xFFu;
        if (uint64_t(d.arg_begin) + n_in > args.size()) throw std::runtime_error("arguments out of range");
        auto in = [&](uint32_t i) -> uint64_t {
            const uint32_t a = args[d.arg_begin + i];
            uint64_t v;
            if (a & kInlineArg) {
                if (!task) throw std::runtime_error("fused gate referenced from a fused gate");
                if (type == 4u || type == 7u) throw std::runtime_error("fused gate under NOT/NULL");
                if (type == 2u && n_in > 15u) throw std::runtime_error("fused gate under a wide ATLEAST");
                const uint32_t ref = a & ~kInlineArg;
                if (ref < lp.launch_begin.back() || ref >= lp.gates.size())
                    throw std::runtime_error("fused reference out of range");
                const GateDesc &c = lp.gates[ref];
                if ((c.type_k_flags & 0xFFu) == 2u) throw std::runtime_error("fused ATLEAST gate");
                if (!(c.type_k_flags & kGateNoStore)) throw std::runtime_error("fused gate stores");
                v = eval(c, w, false);
                if (c.type_k_flags & kGateTallied) ones.at(c.out_row) += uint64_t(__builtin_popcountll(v));
            } else {
                v = row_word(a, w);
            }
            return i >= n_pos ? ~v : v;
        };
        uint64_t r;
        if (type == 2u) {
            uint64_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (uint32_t i = 0; i < n_in; ++i) {
                uint64_t v = in(i);
                for (int b = 0; b < 8; ++b) {
                    const uint64_t t = c[b] & v;
                    c[b] ^= v;
                    v = t;
                }
            }
            uint64_t gt = 0, eq = ~0ull;
            for (int b = 7; b >= 0; --b) {
                if ((k8 >> b) & 1u) {
                    eq &= c[b];
                } else {
                    gt |= eq & c[b];
                    eq &= ~c[b];
                }
            }
            r = gt | eq;
        } else if (type == 4u || type == 7u) {
            r = n_in ? in(n_in - 1) : 0;
        } else {
            const bool and_like = type == 0u || type == 5u;
            r = and_like ? ~0ull : 0ull;
            for (uint32_t i = 0; i < n_in; ++i) {
                const uint64_t v = in(i);
                if (and_like) r &= v;
                else if (type == 3u) r ^= v;
                else r |= v;
            }
        }
        if (type == 4u || type == 5u || type == 6u) r = ~r;
        return r;
    }
};

}  // namespace

// Every workgroup of a launch maps to one task, and every gate and sampler
// task of the launch to exactly one workgroup (xcd_share, gates_before); the
// task decode's division by reciprocal is exact.
int check_split() {
    uint64_t state = 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return state;
    };
    std::vector<std::pair<uint64_t, uint64_t>> cases = {{0, 0}, {0, 1}, {1, 0}, {1, 1}, {7, 0},
                                                        {0, 7}, {8, 8}, {9, 1}, {1, 9}, {1000, 3}};
    for (int i = 0; i < 3000; ++i) {
        const uint64_t scale = uint64_t(1) << (next() % 21);
        cases.push_back({next() % scale, next() % scale});
    }
    cases.push_back({(uint64_t(1) << 26) - 5, (uint64_t(1) << 26) + 3});
    for (auto [ng, ns] : cases) {
        std::vector<uint8_t> gseen(ng, 0), sseen(ns, 0);
        uint64_t gsum = 0, ssum = 0;
        for (uint32_t x = 0; x < kXcds; ++x) {
            const XcdShare sh = xcd_share(ng, ns, x);
            const uint64_t frac = interleave_frac(sh.ngx, sh.tx);
            if (frac > (uint64_t(1) << 32)) return 1;
            if (gates_before(sh.tx, sh.ngx, frac) != sh.ngx) return 1;
            for (uint64_t j = 0; j < sh.tx; ++j) {
                const uint64_t b = gates_before(j, sh.ngx, frac);
                const uint64_t u = gates_before(j + 1, sh.ngx, frac);
                if (u > b + 1) return 1;
                if (u > b) {
                    const uint64_t g = sh.g_lo + b;
                    if (g >= ng || gseen[g]++) return 1;
                    ++gsum;
                } else {
                    const uint64_t t = sh.s_lo + (j - b);
                    if (t >= ns || sseen[t]++) return 1;
                    ++ssum;
                }
            }
        }
        if (gsum != ng || ssum != ns) return 1;
    }
    // Task decode: division by a host-computed reciprocal.
    const uint64_t ds[] = {1, 2, 3, 7, 8, 255, 256, 1000, 65535, 65536, 40000, 131072,
                           (uint64_t(1) << 32) - 1, uint64_t(1) << 32, (uint64_t(1) << 63) + 5,
                           ~uint64_t(0)};
    for (int i = 0; i < 200000; ++i) {
        const uint64_t d = i < 16 * 64 ? ds[i % 16] : (next() >> (next() % 64)) | 1u;
        const uint64_t n = i % 64 == 0 ? ~uint64_t(0) - uint64_t(i % 3)
                                       : next() >> (next() % 64);
        if (div_by_reciprocal(n, d, reciprocal64(d)) != n / d) return 1;
        if (mulhi64(n, d) != uint64_t((static_cast<unsigned __int128>(n) * d) >> 64)) return 1;
    }
    std::printf("split ok: %zu cases\n", cases.size());
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 2 && std::string(argv[1]) == "--split") {
        const int rc = check_split();
        if (rc) std::printf("EMULATOR ERROR: workgroup split\n");
        return rc ? 2 : 0;
    }
    if (argc != 6) {
        std::fprintf(stderr, "usage: plan_emulator PDAG EVENTS FUSE SPLIT OUT\n");
        return 64;
    }
    try {
        const std::vector<char> b = slurp(argv[1]);
        const char *p = b.data();
        auto rd = [&](void *dst, size_t n) {
            std::memcpy(dst, p, n);
            p += n;
        };
        int32_t nv, ng, root;
        int64_t na;
        rd(&nv, 4);
        rd(&ng, 4);
        rd(&na, 8);
        rd(&root, 4);
        const size_t snv = size_t(nv), sng = size_t(ng), sna = size_t(na);
        std::vector<int32_t> vi(snv), gi(sng), gk(sng), args(sna);
        std::vector<double> vp(snv);
        std::vector<uint8_t> gt(sng);
        std::vector<int64_t> off(sng + 1);
        rd(vi.data(), 4 * size_t(nv));
        rd(vp.data(), 8 * size_t(nv));
        rd(gi.data(), 4 * size_t(ng));
        rd(gt.data(), size_t(ng));
        rd(gk.data(), 4 * size_t(ng));
        rd(off.data(), 8 * (size_t(ng) + 1));
        rd(args.data(), 4 * size_t(na));
        const mcs_pdag pd{nv, vi.data(), vp.data(), ng, gi.data(), gt.data(), gk.data(),
                          off.data(), args.data(), root};
        Plan plan;
        std::string err;
        if (!build_plan(pd, plan, err)) throw std::runtime_error("plan: " + err);
        Schedule sch = build_schedule(plan, uint32_t(std::atoi(argv[4])));
        if (std::atoi(argv[3])) fuse_single_consumer(sch, plan, 64);
        const int32_t n_ev = plan.n_events();
        const uint32_t gate_row0 = 2u * uint32_t(n_ev);
        std::vector<uint8_t> tallied(size_t(plan.n_nodes()), 1);
        const RowMap rm{{0u, uint32_t(n_ev)}, gate_row0, 0u};
        TreeLaunchPlan lp;
        build_launch_plan(plan, sch, tallied, rm, lp);
        if (lp.args[0].size() != lp.args[1].size()) throw std::runtime_error("argument arrays differ in size");
        const std::vector<char> ev = slurp(argv[2]);
        if (ev.size() != size_t(n_ev) * kWords * 8) throw std::runtime_error("event words file size");
        const uint64_t n_rows = uint64_t(gate_row0) + uint64_t(plan.n_gates()) + uint64_t(sch.n_aux);

        std::vector<uint64_t> out_words, out_ones;
        for (int buf = 0; buf < 2; ++buf) {
            Emulator em{lp, lp.args[buf], n_rows, {}, {}, {}};
            em.words.assign(n_rows * kWords, 0);
            em.written.assign(n_rows, 0);
            em.ones.assign(n_rows, 0);
            const uint64_t be0 = uint64_t(rm.be_row0[buf]);
            std::memcpy(&em.words[be0 * kWords], ev.data(), ev.size());
            for (int32_t e = 0; e < n_ev; ++e) em.written[be0 + uint64_t(e)] = 1;

            for (size_t l = 0; l + 1 < lp.launch_begin.size(); ++l) {
                // Tasks of one launch see only rows of earlier launches.
                std::vector<std::pair<uint32_t, std::vector<uint64_t>>> stores;
                for (uint32_t g = lp.launch_begin[l]; g < lp.launch_begin[l + 1]; ++g) {
                    const GateDesc &d = lp.gates[g];
                    std::vector<uint64_t> o(kWords);
                    for (int w = 0; w < kWords; ++w) o[size_t(w)] = em.eval(d, w, true);
                    if (d.type_k_flags & kGateTallied)
                        for (uint64_t x : o) em.ones.at(d.out_row) += uint64_t(__builtin_popcountll(x));
                    if (!(d.type_k_flags & kGateNoStore)) {
                        if (d.out_row < gate_row0) throw std::runtime_error("store into an event buffer");
                        stores.push_back({d.out_row, o});
                    }
                }
                for (auto &[row, o] : stores) {
                    if (row >= n_rows) throw std::runtime_error("store out of range");
                    std::memcpy(&em.words[uint64_t(row) * kWords], o.data(), kWords * 8);
                    em.written[row] = 1;
                }
            }
            for (size_t g = lp.launch_begin.back(); g < lp.gates.size(); ++g) {
                const GateDesc &d = lp.gates[g];
                if (d.out_row < gate_row0 || d.out_row >= n_rows) throw std::runtime_error("fused row out of range");
                for (int w = 0; w < kWords; ++w)
                    em.words[uint64_t(d.out_row) * kWords + uint64_t(w)] = em.eval(d, w, false);
                em.written[d.out_row] = 1;
            }

            std::vector<uint64_t> words, ones;
            for (int32_t n = n_ev; n < plan.n_nodes(); ++n) {
                const uint64_t row = uint64_t(gate_row0) + uint64_t(n - n_ev);
                if (!em.written[row]) throw std::runtime_error("gate row never written");
                words.insert(words.end(), &em.words[row * kWords], &em.words[row * kWords] + kWords);
            }
            for (int32_t n = 0; n < plan.n_nodes(); ++n)
                ones.push_back(n < n_ev ? em.ones[size_t(n)] : em.ones[size_t(gate_row0) + size_t(n - n_ev)]);
            if (buf == 0) {
                out_words = words;
                out_ones = ones;
            } else if (words != out_words || ones != out_ones) {
                throw std::runtime_error("the two event buffers give different results");
            }
        }

        FILE *f = std::fopen(argv[5], "wb");
        if (!f) throw std::runtime_error("cannot write output");
        std::fwrite(out_words.data(), 8, out_words.size(), f);
        std::fwrite(out_ones.data(), 8, out_ones.size(), f);
        std::fwrite(plan.node_pdag_index.data(), 4, size_t(plan.n_nodes()), f);
        std::fclose(f);
        std::printf("launches %zu fused %zu nodes %d\n", lp.launch_begin.size() - 1,
                    lp.gates.size() - lp.launch_begin.back(), plan.n_nodes());
    } catch (const std::exception &e) {
        std::printf("EMULATOR ERROR: %s\n", e.what());
        return 2;
    }
    return 0;
}
