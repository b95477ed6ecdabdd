// Exact top-event probability by a reduced ordered BDD.
//
// TEST INFRASTRUCTURE ONLY: the exact value the Monte Carlo estimates are
// checked against. The reference computes the same quantity with its BDD
// (src/bdd.{h,cc}; probability at src/probability_analysis.cc:213-258):
// P(f) = p(x) P(f|x=1) + (1 - p(x)) P(f|x=0) at every node. This file is a
// textbook construction (Bryant, IEEE Trans. Comput. C-35(8), 1986): a unique
// table, memoised apply for AND/OR/XOR, ATLEAST k-of-n by the recursion
// T(k, i) = ite(x_i, T(k-1, i+1), T(k, i+1)), and no complement edges.
//
// Variable order: first appearance in a depth-first walk from the root that
// takes a gate's variable arguments first, in their listed order, then
// descends into its gate arguments in their listed order. The order changes
// the BDD's size, never the probability.
//
// Pinned by: brute-force enumeration of small random trees (all 2^n
// assignments), the reference's core known answers (tests/bench_core_tests.cc)
// and the exact probabilities of the Aralia models that the reference's
// benchmark table lists (benchmarks/input.csv).

#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

enum : uint8_t { kAnd = 0, kOr, kAtleast, kXor, kNot, kNand, kNor, kNull };

struct Node {
    int32_t level;  // position in the variable order; terminals: INT32_MAX
    int32_t lo, hi;
};

struct Key3 {
    int64_t a, b, c;
    bool operator==(const Key3 &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct Hash3 {
    size_t operator()(const Key3 &k) const {
        uint64_t h = uint64_t(k.a) * 0x9E3779B97F4A7C15ull;
        h ^= uint64_t(k.b) + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
        h ^= uint64_t(k.c) + 0x85157AF5ull + (h << 6) + (h >> 2);
        return size_t(h ^ (h >> 29));
    }
};

class Bdd {
  public:
    explicit Bdd(size_t limit) : limit_(limit) {
        nodes_.push_back({INT32_MAX, 0, 0});  // 0 = false
        nodes_.push_back({INT32_MAX, 1, 1});  // 1 = true
    }
    bool overflow() const { return overflow_; }
    size_t size() const { return nodes_.size(); }

    int32_t var(int32_t level) { return make(level, 0, 1); }

    int32_t make(int32_t level, int32_t lo, int32_t hi) {
        if (lo == hi) return lo;
        const Key3 k{level, lo, hi};
        auto it = unique_.find(k);
        if (it != unique_.end()) return it->second;
        if (nodes_.size() >= limit_) {
            overflow_ = true;
            return 0;
        }
        nodes_.push_back({level, lo, hi});
        const int32_t id = int32_t(nodes_.size() - 1);
        unique_.emplace(k, id);
        return id;
    }

    // op: 0 AND, 1 OR, 2 XOR
    int32_t apply(int op, int32_t f, int32_t g) {
        if (overflow_) return 0;
        if (op == 0) {
            if (f == 0 || g == 0) return 0;
            if (f == 1) return g;
            if (g == 1) return f;
            if (f == g) return f;
        } else if (op == 1) {
            if (f == 1 || g == 1) return 1;
            if (f == 0) return g;
            if (g == 0) return f;
            if (f == g) return f;
        } else {
            if (f == 0) return g;
            if (g == 0) return f;
            if (f == g) return 0;
        }
        if (f > g) std::swap(f, g);  // all three operations commute
        const Key3 k{op, f, g};
        auto it = cache_.find(k);
        if (it != cache_.end()) return it->second;
        const Node nf = nodes_[f], ng = nodes_[g];
        const int32_t lv = nf.level < ng.level ? nf.level : ng.level;
        const int32_t f0 = nf.level == lv ? nf.lo : f, f1 = nf.level == lv ? nf.hi : f;
        const int32_t g0 = ng.level == lv ? ng.lo : g, g1 = ng.level == lv ? ng.hi : g;
        const int32_t lo = apply(op, f0, g0);
        const int32_t hi = apply(op, f1, g1);
        const int32_t r = make(lv, lo, hi);
        cache_.emplace(k, r);
        return r;
    }

    int32_t negate(int32_t f) { return apply(2, f, 1); }

    int32_t ite(int32_t c, int32_t t, int32_t e) {
        return apply(1, apply(0, c, t), apply(0, negate(c), e));
    }

    double probability(int32_t f, const std::vector<double> &p_of_level) {
        std::unordered_map<int32_t, double> memo;
        return prob(f, p_of_level, memo);
    }

  private:
    double prob(int32_t f, const std::vector<double> &p, std::unordered_map<int32_t, double> &memo) {
        if (f <= 1) return double(f);
        auto it = memo.find(f);
        if (it != memo.end()) return it->second;
        const Node n = nodes_[f];
        const double q = p[size_t(n.level)];
        const double r = q * prob(n.hi, p, memo) + (1.0 - q) * prob(n.lo, p, memo);
        memo.emplace(f, r);
        return r;
    }

    size_t limit_;
    bool overflow_ = false;
    std::vector<Node> nodes_;
    std::unordered_map<Key3, int32_t, Hash3> unique_;
    std::unordered_map<Key3, int32_t, Hash3> cache_;
};

std::string g_bdd_err;

}  // namespace

extern "C" {

const char *orc_bdd_last_error() { return g_bdd_err.c_str(); }

// Exact P(root). Returns the probability, or -1 when the BDD exceeds
// max_nodes, or -2 on a malformed PDAG (message in orc_bdd_last_error).
// out_nodes (nullable) receives the number of BDD nodes built.
double orc_bdd_probability(int32_t n_vars, const int32_t *var_index, const double *var_p,
                           int32_t n_gates, const int32_t *gate_index, const uint8_t *gate_type,
                           const int32_t *gate_k, const int64_t *arg_off, const int32_t *args,
                           int32_t root, uint64_t max_nodes, uint64_t *out_nodes) {
    std::unordered_map<int32_t, int32_t> var_slot, gate_slot;
    for (int32_t i = 0; i < n_vars; ++i) var_slot[var_index[i]] = i;
    for (int32_t i = 0; i < n_gates; ++i) gate_slot[gate_index[i]] = i;
    if (!gate_slot.count(root)) {
        g_bdd_err = "root is not a gate";
        return -2.0;
    }
    // Variable order and gate post-order from one depth-first walk.
    std::vector<int32_t> level_of_var(size_t(n_vars), -1);
    std::vector<double> p_of_level;
    std::vector<int32_t> post;
    std::vector<uint8_t> state(size_t(n_gates), 0);  // 0 new, 1 open, 2 done
    struct Frame {
        int32_t slot;
        int64_t next;
        int pass;  // 0: variables, 1: gates
    };
    std::vector<Frame> stack{{gate_slot[root], arg_off[gate_slot[root]], 0}};
    state[size_t(gate_slot[root])] = 1;
    while (!stack.empty()) {
        Frame &fr = stack.back();
        const int64_t end = arg_off[fr.slot + 1];
        if (fr.next >= end) {
            if (fr.pass == 0) {
                fr.pass = 1;
                fr.next = arg_off[fr.slot];
                continue;
            }
            state[size_t(fr.slot)] = 2;
            post.push_back(fr.slot);
            stack.pop_back();
            continue;
        }
        const int32_t a = args[fr.next++];
        const int32_t idx = a < 0 ? -a : a;
        if (fr.pass == 0) {
            auto v = var_slot.find(idx);
            if (v != var_slot.end() && level_of_var[size_t(v->second)] < 0) {
                level_of_var[size_t(v->second)] = int32_t(p_of_level.size());
                p_of_level.push_back(var_p[v->second]);
            }
            if (v == var_slot.end() && !gate_slot.count(idx)) {
                g_bdd_err = "unknown argument index " + std::to_string(idx);
                return -2.0;
            }
        } else {
            auto g = gate_slot.find(idx);
            if (g == gate_slot.end()) continue;
            if (state[size_t(g->second)] == 1) {
                g_bdd_err = "cycle through gate " + std::to_string(idx);
                return -2.0;
            }
            if (state[size_t(g->second)] == 0) {
                state[size_t(g->second)] = 1;
                stack.push_back({g->second, arg_off[g->second], 0});
            }
        }
    }

    Bdd bdd{static_cast<size_t>(max_nodes)};
    std::vector<int32_t> gate_bdd(size_t(n_gates), 0);
    auto arg_bdd = [&](int32_t a) {
        const int32_t idx = a < 0 ? -a : a;
        auto v = var_slot.find(idx);
        const int32_t f = v != var_slot.end() ? bdd.var(level_of_var[size_t(v->second)])
                                              : gate_bdd[size_t(gate_slot[idx])];
        return a < 0 ? bdd.negate(f) : f;
    };
    for (int32_t s : post) {
        const int64_t b = arg_off[s], e = arg_off[s + 1];
        const uint8_t t = gate_type[s];
        int32_t r = 0;
        if (t == kAtleast) {
            const int32_t n = int32_t(e - b), k = gate_k[s];
            if (k <= 0) {
                r = 1;
            } else if (k > n) {
                r = 0;
            } else {
                // row[j] = T(j, i): at least j of args i..n-1.
                std::vector<int32_t> row(size_t(k) + 1, 0);
                row[0] = 1;
                for (int32_t i = n - 1; i >= 0; --i) {
                    const int32_t x = arg_bdd(args[b + i]);
                    for (int32_t j = k; j >= 1; --j) row[size_t(j)] = bdd.ite(x, row[size_t(j - 1)], row[size_t(j)]);
                }
                r = row[size_t(k)];
            }
        } else if (t == kNot || t == kNull) {
            r = e > b ? arg_bdd(args[e - 1]) : 0;
            if (t == kNot) r = bdd.negate(r);
        } else {
            const bool and_like = t == kAnd || t == kNand;
            const int op = and_like ? 0 : (t == kXor ? 2 : 1);
            r = and_like ? 1 : 0;
            for (int64_t i = b; i < e; ++i) r = bdd.apply(op, r, arg_bdd(args[i]));
            if (t == kNand || t == kNor) r = bdd.negate(r);
        }
        gate_bdd[size_t(s)] = r;
        if (bdd.overflow()) break;
    }
    if (out_nodes) *out_nodes = bdd.size();
    if (bdd.overflow()) {
        g_bdd_err = "BDD node limit exceeded";
        return -1.0;
    }
    return bdd.probability(gate_bdd[size_t(gate_slot[root])], p_of_level);
}

}  // extern "C"
