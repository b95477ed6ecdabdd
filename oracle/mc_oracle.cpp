// CPU oracle for the Monte Carlo fault-tree path.
//
// TEST INFRASTRUCTURE ONLY. This file is a plain, unoptimised restatement of
// the reference's DirectEval Monte Carlo algorithm (mcSCRAM src/mc). Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark may
// load it, and only as the checker. The product library never links it.
//
// What each function restates (reference file:line):
//   orc_philox4x32_10   src/mc/prng/philox128.h:8-54, state128.h:21-26
//   orc_threshold       src/mc/event/node.h:371-376
//   orc_sample_word     src/mc/kernel/basic_event.h:112-142,
//                       src/mc/prng/philox128.h:59-88
//   orc_gate_word       src/mc/kernel/gate.h:262-264, 299-362 (AND/OR/XOR/
//                       NOT/NAND/NOR/NULL), gate.h:506-555 (ATLEAST)
//   orc_layout          src/preprocessor.cc:130-176 (node depth),
//                       src/mc/queue/layer_manager.tpp:33-118 (DFS gather and
//                       the two std::sort calls that fix event_idx)
//   orc_run             src/mc/queue/layer_manager.tpp:202-279 (one pass per
//                       iteration), src/mc/kernel/tally.h:322-360 (tallies),
//                       src/mc/queue/kernel_builder.h:298-426 (input order)
//   orc_point_estimates src/mc/stats/ci_utils.h:167-201
//
// Parity status of each function, and the check that pins it, is stated in
// tests/test_oracle.py. The event_idx permutation (orc_layout) is "parity
// unpinned": it is the reference's DFS gather (layer_manager.tpp:33-60),
// restated from its text, followed by libstdc++'s std::sort with the
// reference's comparators. The sort is pinned by a separate restatement of
// libstdc++'s published introsort (tests/introsort.py); the gather order is
// checked only against other restatements of the same text, and no
// reference binary, vector or fixture holds an event order to compare with.
// Results that depend on it are bit-exact against the restated algorithm.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp). Everything is extern "C" with
// plain pointers so the tests can call it through ctypes.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif
#if defined(__AVX512F__) && defined(__BMI2__)
#include <immintrin.h>
#endif

namespace {

// Connective numbering of the reference PDAG (src/pdag.h:265-275).
enum : uint8_t { kAnd = 0, kOr, kAtleast, kXor, kNot, kNand, kNor, kNull };

// Fixed Philox key of the reference sampler (src/mc/prng/philox128.h:39-42).
constexpr uint32_t kKey0 = 382307844u;
constexpr uint32_t kKey1 = 293830103u;

inline void philox_rounds(uint32_t x[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t pa = uint64_t(0xD2511F53u) * x[0];
        const uint64_t pb = uint64_t(0xCD9E8D57u) * x[2];
        const uint32_t n0 = uint32_t(pb >> 32) ^ x[1] ^ k0;
        const uint32_t n1 = uint32_t(pb);
        const uint32_t n2 = uint32_t(pa >> 32) ^ x[3] ^ k1;
        const uint32_t n3 = uint32_t(pa);
        x[0] = n0; x[1] = n1; x[2] = n2; x[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

struct Graph {
    // node id space: pdag index -> (is_var, slot)
    std::unordered_map<int32_t, std::pair<bool, int32_t>> by_index;
    int32_t n_vars = 0, n_gates = 0;
    const int32_t *var_index = nullptr;
    const double *var_p = nullptr;
    const int32_t *gate_index = nullptr;
    const uint8_t *gate_type = nullptr;
    const int32_t *gate_k = nullptr;
    const int64_t *arg_off = nullptr;
    const int32_t *args = nullptr;
    int32_t root = 0;
};

struct Layout {
    std::vector<int32_t> event_order;  // variable slots in event_idx order
    std::vector<int32_t> gate_order;   // gate slots in evaluation order
    std::vector<int32_t> var_depth, gate_depth;
    std::string error;
};

bool build_graph(Graph &g, std::string &err) {
    for (int32_t i = 0; i < g.n_vars; ++i) {
        if (!g.by_index.emplace(g.var_index[i], std::make_pair(true, i)).second) {
            err = "duplicate variable index";
            return false;
        }
    }
    for (int32_t i = 0; i < g.n_gates; ++i) {
        if (!g.by_index.emplace(g.gate_index[i], std::make_pair(false, i)).second) {
            err = "duplicate gate index";
            return false;
        }
    }
    for (int32_t i = 0; i < g.n_gates; ++i) {
        for (int64_t a = g.arg_off[i]; a < g.arg_off[i + 1]; ++a) {
            if (!g.by_index.count(std::abs(g.args[a]))) {
                err = "unknown argument index";
                return false;
            }
        }
    }
    auto it = g.by_index.find(g.root);
    if (it == g.by_index.end() || it->second.first) {
        err = "root is not a gate";
        return false;
    }
    return true;
}

// One element of the sequence the reference sorts: a node with its depth.
struct SortNode {
    int32_t depth;
    bool is_var;
    uint8_t type;
    int32_t slot;
};

bool layout(const Graph &g, Layout &L) {
    L.var_depth.assign(g.n_vars, -1);
    L.gate_depth.assign(g.n_gates, -1);
    // Node depth: variables 0, gate = 1 + max child depth (preprocessor.cc:130-176).
    std::function<int32_t(int32_t)> gate_depth = [&](int32_t gs) -> int32_t {
        if (L.gate_depth[gs] >= 0) return L.gate_depth[gs];
        int32_t m = -1;
        for (int64_t a = g.arg_off[gs]; a < g.arg_off[gs + 1]; ++a) {
            const auto &ref = g.by_index.at(std::abs(g.args[a]));
            int32_t d;
            if (ref.first) {
                L.var_depth[ref.second] = 0;
                d = 0;
            } else {
                d = gate_depth(ref.second);
            }
            m = std::max(m, d);
        }
        L.gate_depth[gs] = m + 1;
        return m + 1;
    };
    const int32_t root_slot = g.by_index.at(g.root).second;
    gate_depth(root_slot);

    // DFS gather (layer_manager.tpp:33-60): the gate, then its gate arguments
    // recursively in insertion order, then its unvisited variable arguments.
    std::vector<char> gvis(g.n_gates, 0), vvis(g.n_vars, 0);
    std::vector<SortNode> nodes;
    std::function<void(int32_t)> gather = [&](int32_t gs) {
        if (gvis[gs]) return;
        gvis[gs] = 1;
        nodes.push_back({L.gate_depth[gs], false, g.gate_type[gs], gs});
        for (int64_t a = g.arg_off[gs]; a < g.arg_off[gs + 1]; ++a) {
            const auto &ref = g.by_index.at(std::abs(g.args[a]));
            if (!ref.first) gather(ref.second);
        }
        for (int64_t a = g.arg_off[gs]; a < g.arg_off[gs + 1]; ++a) {
            const auto &ref = g.by_index.at(std::abs(g.args[a]));
            if (ref.first && !vvis[ref.second]) {
                vvis[ref.second] = 1;
                nodes.push_back({0, true, 0, ref.second});
            }
        }
    };
    gather(root_slot);

    // First unstable sort: by depth only (layer_manager.tpp:79-82).
    std::sort(nodes.begin(), nodes.end(),
              [](const SortNode &a, const SortNode &b) { return a.depth < b.depth; });
    const int32_t max_layer = nodes.back().depth;
    std::vector<std::vector<SortNode>> layers(max_layer + 1);
    for (const auto &n : nodes) layers[n.depth].push_back(n);
    // Second unstable sort per layer: variables first, gates by connective
    // (layer_manager.tpp:92-115).
    for (auto &layer : layers) {
        std::sort(layer.begin(), layer.end(), [](const SortNode &l, const SortNode &r) {
            if (l.is_var && !r.is_var) return true;
            if (!l.is_var && r.is_var) return false;
            if (l.is_var && r.is_var) return false;
            return l.type < r.type;
        });
    }
    for (const auto &layer : layers) {
        for (const auto &n : layer) {
            if (n.is_var) L.event_order.push_back(n.slot);
            else L.gate_order.push_back(n.slot);
        }
    }
    return true;
}

inline uint64_t threshold_of(double p) {
    constexpr double two_to_32 = 4294967296.0;
    return static_cast<uint64_t>(p * two_to_32);
}

inline uint64_t sample_word_scalar(uint32_t pdag_idx, uint32_t event_idx, uint32_t sample_idx,
                                   uint32_t iteration, uint64_t threshold, uint32_t key0,
                                   uint32_t key1) {
    uint64_t word = 0;
    for (uint32_t gen = 0; gen < 16; ++gen) {
        uint32_t x[4] = {pdag_idx + 1u, event_idx + 1u, sample_idx + 1u,
                         ((iteration + 1u) << 6) + gen};
        philox_rounds(x, key0, key1);
        for (uint32_t j = 0; j < 4; ++j) {
            if (uint64_t(x[j]) < threshold) word |= uint64_t(1) << (4 * gen + j);
        }
    }
    return word;
}

#if defined(__AVX512F__) && defined(__BMI2__)
// The same 16 Philox blocks of a word as the 16 32-bit lanes of one AVX-512
// register per state word: a 32 x 32 -> 64-bit product is two vpmuludq
// (even lanes, then odd lanes shifted down), recombined into hi and lo
// halves; bit 4g + j of the word is lane g of the compare mask of output j.
inline void mul_hilo16(__m512i a, __m512i m, __m512i &hi, __m512i &lo) {
    const __m512i even = _mm512_mul_epu32(a, m);
    const __m512i odd = _mm512_mul_epu32(_mm512_srli_epi64(a, 32), m);
    lo = _mm512_mask_blend_epi32(0xAAAA, even, _mm512_slli_epi64(odd, 32));
    hi = _mm512_mask_blend_epi32(0xAAAA, _mm512_srli_epi64(even, 32), odd);
}

inline uint64_t sample_word_avx512(uint32_t pdag_idx, uint32_t event_idx, uint32_t sample_idx,
                                   uint32_t iteration, uint64_t threshold, uint32_t key0,
                                   uint32_t key1) {
    if (threshold == 0) return 0;
    if (threshold > 0xFFFFFFFFull) return ~uint64_t(0);
    __m512i x0 = _mm512_set1_epi32(int(pdag_idx + 1u));
    __m512i x1 = _mm512_set1_epi32(int(event_idx + 1u));
    __m512i x2 = _mm512_set1_epi32(int(sample_idx + 1u));
    __m512i x3 = _mm512_add_epi32(_mm512_set1_epi32(int((iteration + 1u) << 6)),
                                  _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12,
                                                    13, 14, 15));
    const __m512i m0 = _mm512_set1_epi32(int(0xD2511F53u)), m1 = _mm512_set1_epi32(int(0xCD9E8D57u));
    uint32_t k0 = key0, k1 = key1;
    for (int r = 0; r < 10; ++r) {
        __m512i ha, la, hb, lb;
        mul_hilo16(x0, m0, ha, la);
        mul_hilo16(x2, m1, hb, lb);
        const __m512i n0 = _mm512_xor_si512(_mm512_xor_si512(hb, x1), _mm512_set1_epi32(int(k0)));
        const __m512i n2 = _mm512_xor_si512(_mm512_xor_si512(ha, x3), _mm512_set1_epi32(int(k1)));
        x0 = n0; x1 = lb; x2 = n2; x3 = la;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    const __m512i t = _mm512_set1_epi32(int(uint32_t(threshold)));
    const uint64_t b0 = _mm512_cmplt_epu32_mask(x0, t), b1 = _mm512_cmplt_epu32_mask(x1, t);
    const uint64_t b2 = _mm512_cmplt_epu32_mask(x2, t), b3 = _mm512_cmplt_epu32_mask(x3, t);
    return _pdep_u64(b0, 0x1111111111111111ull) | _pdep_u64(b1, 0x2222222222222222ull) |
           _pdep_u64(b2, 0x4444444444444444ull) | _pdep_u64(b3, 0x8888888888888888ull);
}
#endif

inline uint64_t sample_word(uint32_t pdag_idx, uint32_t event_idx, uint32_t sample_idx,
                            uint32_t iteration, uint64_t threshold, uint32_t key0 = kKey0,
                            uint32_t key1 = kKey1) {
#if defined(__AVX512F__) && defined(__BMI2__)
    return sample_word_avx512(pdag_idx, event_idx, sample_idx, iteration, threshold, key0, key1);
#else
    return sample_word_scalar(pdag_idx, event_idx, sample_idx, iteration, threshold, key0, key1);
#endif
}

// Gate word from its inputs in the reference's order: positive inputs first,
// then negated ones (kernel_builder.h:298-426). `vals` holds n words, the
// first n - n_neg positive.
inline uint64_t gate_word(uint8_t type, int32_t k, const uint64_t *vals, int32_t n, int32_t n_neg) {
    const int32_t neg_off = n - n_neg;
    if (type == kAtleast) {
        // Per bit: 8-bit count of true inputs >= 8-bit k (gate.h:526-541, node.h:676).
        const uint8_t k8 = static_cast<uint8_t>(k);
        uint64_t out = 0;
        for (int bit = 0; bit < 64; ++bit) {
            uint8_t cnt = 0;
            const uint64_t mask = uint64_t(1) << bit;
            for (int32_t i = 0; i < neg_off; ++i) cnt += (vals[i] & mask) ? 1 : 0;
            for (int32_t i = neg_off; i < n; ++i) cnt += (~vals[i] & mask) ? 1 : 0;
            if (cnt >= k8) out |= mask;
        }
        return out;
    }
    uint64_t r = (type == kAnd || type == kNand) ? ~uint64_t(0) : uint64_t(0);
    for (int32_t i = 0; i < n; ++i) {
        const uint64_t v = i < neg_off ? vals[i] : ~vals[i];
        switch (type) {
        case kOr: case kNor: r |= v; break;
        case kAnd: case kNand: r &= v; break;
        case kXor: r ^= v; break;
        case kNot: case kNull: r = v; break;
        default: break;
        }
    }
    if (type == kNand || type == kNor || type == kNot) r = ~r;
    return r;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox_rounds(x, key[0], key[1]);
    std::memcpy(out, x, sizeof(x));
}

uint64_t orc_threshold(double p) { return threshold_of(p); }

uint64_t orc_sample_word(uint32_t pdag_idx, uint32_t event_idx, uint32_t sample_idx,
                         uint32_t iteration, uint64_t threshold) {
    return sample_word(pdag_idx, event_idx, sample_idx, iteration, threshold);
}

uint64_t orc_gate_word(uint8_t type, int32_t k, const uint64_t *vals, int32_t n, int32_t n_neg) {
    return gate_word(type, k, vals, n, n_neg);
}

// Layered order. Outputs: event_idx per variable slot (-1 if unreachable),
// depth per gate slot (-1 if unreachable), and the gate evaluation order.
// Returns the number of reachable gates, or -1 on error.
int32_t orc_layout(int32_t n_vars, const int32_t *var_index, const double *var_p,
                   int32_t n_gates, const int32_t *gate_index, const uint8_t *gate_type,
                   const int32_t *gate_k, const int64_t *arg_off, const int32_t *args,
                   int32_t root, int32_t *event_idx_out, int32_t *gate_depth_out,
                   int32_t *gate_order_out) {
    Graph g;
    g.n_vars = n_vars; g.var_index = var_index; g.var_p = var_p;
    g.n_gates = n_gates; g.gate_index = gate_index; g.gate_type = gate_type;
    g.gate_k = gate_k; g.arg_off = arg_off; g.args = args; g.root = root;
    if (!build_graph(g, g_err)) return -1;
    Layout L;
    layout(g, L);
    for (int32_t i = 0; i < n_vars; ++i) event_idx_out[i] = -1;
    for (size_t e = 0; e < L.event_order.size(); ++e) event_idx_out[L.event_order[e]] = int32_t(e);
    for (int32_t i = 0; i < n_gates; ++i) gate_depth_out[i] = L.gate_depth[i];
    for (size_t i = 0; i < L.gate_order.size(); ++i) gate_order_out[i] = L.gate_order[i];
    return int32_t(L.gate_order.size());
}

// Full Monte Carlo: `n_iters` passes of n_bp words each, starting at 1-based
// iteration `first_iter`. Accumulates popcounts into ones_out for the tallied
// pdag indices. If words_out is non-null it receives every reachable node's
// words of the LAST pass, rows in the order [variables by slot, gates by slot]
// (n_vars + n_gates rows of n_bp words; unreachable rows left untouched).
// Returns 0 on success.
int32_t orc_run_keyed(int32_t n_vars, const int32_t *var_index, const double *var_p,
                      int32_t n_gates, const int32_t *gate_index, const uint8_t *gate_type,
                      const int32_t *gate_k, const int64_t *arg_off, const int32_t *args,
                      int32_t root, uint64_t n_bp, uint32_t first_iter, uint32_t n_iters,
                      int32_t n_tally, const int32_t *tally_index, uint64_t *ones_out,
                      uint64_t *words_out, int32_t n_threads, uint32_t key0, uint32_t key1) {
    Graph g;
    g.n_vars = n_vars; g.var_index = var_index; g.var_p = var_p;
    g.n_gates = n_gates; g.gate_index = gate_index; g.gate_type = gate_type;
    g.gate_k = gate_k; g.arg_off = arg_off; g.args = args; g.root = root;
    if (!build_graph(g, g_err)) return -1;
    for (int32_t i = 0; i < n_vars; ++i) {
        if (!(var_p[i] >= 0.0 && var_p[i] <= 1.0)) {
            g_err = "probability outside [0, 1]";
            return -1;
        }
    }
    Layout L;
    layout(g, L);
    std::vector<int32_t> ev_of_slot(n_vars, -1);
    for (size_t e = 0; e < L.event_order.size(); ++e) ev_of_slot[L.event_order[e]] = int32_t(e);

    // Tally targets as (is_var, slot).
    std::vector<std::pair<bool, int32_t>> tally_ref(n_tally);
    for (int32_t t = 0; t < n_tally; ++t) {
        auto it = g.by_index.find(tally_index[t]);
        if (it == g.by_index.end()) {
            g_err = "unknown tally index";
            return -1;
        }
        const bool reach = it->second.first ? ev_of_slot[it->second.second] >= 0
                                            : L.gate_depth[it->second.second] >= 0;
        if (!reach) {
            g_err = "tally node not reachable from the root";
            return -1;
        }
        tally_ref[t] = it->second;
        ones_out[t] = 0;
    }

    std::vector<uint64_t> thr(n_vars);
    for (int32_t i = 0; i < n_vars; ++i) thr[i] = threshold_of(var_p[i]);

    // Per gate: argument references split as the reference orders them.
    struct Arg { bool is_var; int32_t slot; };
    std::vector<std::vector<Arg>> gin(n_gates);
    std::vector<int32_t> gneg(n_gates, 0);
    for (int32_t gs : L.gate_order) {
        std::vector<Arg> pos, neg;
        for (int pass = 0; pass < 2; ++pass) {  // variables first, then gates
            for (int64_t a = arg_off[gs]; a < arg_off[gs + 1]; ++a) {
                const auto &ref = g.by_index.at(std::abs(args[a]));
                if (ref.first != (pass == 0)) continue;
                (args[a] < 0 ? neg : pos).push_back({ref.first, ref.second});
            }
        }
        gneg[gs] = int32_t(neg.size());
        pos.insert(pos.end(), neg.begin(), neg.end());
        gin[gs] = std::move(pos);
    }

#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    for (uint32_t it = 0; it < n_iters; ++it) {
        const uint32_t iteration = first_iter + it;
        const bool last = (it + 1 == n_iters);
#pragma omp parallel
        {
            std::vector<uint64_t> vword(n_vars, 0), gword(n_gates, 0), vals;
            std::vector<uint64_t> local(n_tally, 0);
#pragma omp for schedule(static)
            for (int64_t s = 0; s < int64_t(n_bp); ++s) {
                for (size_t e = 0; e < L.event_order.size(); ++e) {
                    const int32_t vs = L.event_order[e];
                    vword[vs] = sample_word(uint32_t(var_index[vs]), uint32_t(e), uint32_t(s),
                                            iteration, thr[vs], key0, key1);
                }
                for (int32_t gs : L.gate_order) {
                    const auto &in = gin[gs];
                    vals.resize(in.size());
                    for (size_t i = 0; i < in.size(); ++i)
                        vals[i] = in[i].is_var ? vword[in[i].slot] : gword[in[i].slot];
                    gword[gs] = gate_word(gate_type[gs], gate_k[gs], vals.data(),
                                          int32_t(in.size()), gneg[gs]);
                }
                for (int32_t t = 0; t < n_tally; ++t) {
                    const auto &r = tally_ref[t];
                    local[t] += uint64_t(__builtin_popcountll(r.first ? vword[r.second] : gword[r.second]));
                }
                if (last && words_out) {
                    for (size_t e = 0; e < L.event_order.size(); ++e) {
                        const int32_t vs = L.event_order[e];
                        words_out[uint64_t(vs) * n_bp + uint64_t(s)] = vword[vs];
                    }
                    for (int32_t gs : L.gate_order)
                        words_out[(uint64_t(n_vars) + uint64_t(gs)) * n_bp + uint64_t(s)] = gword[gs];
                }
            }
#pragma omp critical
            for (int32_t t = 0; t < n_tally; ++t) ones_out[t] += local[t];
        }
    }
    return 0;
}

// orc_run_keyed with the reference's fixed key.
int32_t orc_run(int32_t n_vars, const int32_t *var_index, const double *var_p,
                int32_t n_gates, const int32_t *gate_index, const uint8_t *gate_type,
                const int32_t *gate_k, const int64_t *arg_off, const int32_t *args,
                int32_t root, uint64_t n_bp, uint32_t first_iter, uint32_t n_iters,
                int32_t n_tally, const int32_t *tally_index, uint64_t *ones_out,
                uint64_t *words_out, int32_t n_threads) {
    return orc_run_keyed(n_vars, var_index, var_p, n_gates, gate_index, gate_type, gate_k, arg_off,
                         args, root, n_bp, first_iter, n_iters, n_tally, tally_index, ones_out,
                         words_out, n_threads, kKey0, kKey1);
}

// Point estimates from raw counters: out = {mean, std_err, lo95, hi95, lo99, hi99}.
void orc_point_estimates(uint64_t ones, uint64_t total, double out[6]) {
    if (total == 0) {
        for (int i = 0; i < 6; ++i) out[i] = 0.0;
        return;
    }
    const double n = static_cast<double>(total);
    const double p = static_cast<double>(ones) / n;
    const double se = std::sqrt(p * (1.0 - p) / n);
    const double z95 = 1.959963984540054, z99 = 2.5758293035489004;
    out[0] = p;
    out[1] = se;
    out[2] = std::clamp(p - se * z95, 0.0, 1.0);
    out[3] = std::clamp(p + se * z95, 0.0, 1.0);
    out[4] = std::clamp(p - se * z99, 0.0, 1.0);
    out[5] = std::clamp(p + se * z99, 0.0, 1.0);
}

int32_t orc_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
