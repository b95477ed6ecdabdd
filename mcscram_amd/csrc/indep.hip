// Independence approximation of the top event for many parameter samples:
// the closed-form baseline whose residual the uncertainty-analysis
// surrogate learns. No counterpart in the reference's Monte Carlo path.
//
// Each gate's probability with its inputs treated as independent: AND the
// product, OR one minus the product of the complements, XOR by the parity
// recursion p <- p (1 - q) + (1 - p) q, ATLEAST the Poisson-binomial tail
// P(count >= k), NOT/NULL the input the reference's NOT/NULL reads (the
// last in its input order), NAND/NOR/NOT complemented; a
// complemented input is 1 - q. Exact for trees without shared events, an
// approximation (the top event's logit to within a few tenths on the
// workloads here) otherwise. FP64 throughout, so probabilities near 1 keep
// their complement.
//
// Layout: one thread per sample; node probabilities [node][sample] in HBM,
// so every load and store of a gate step is coalesced across a wave. The
// gates run in a depth-first post-order from the root (every input before
// its gate), the same order for every thread; descriptors are wave-uniform
// scalar loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

#include "mcscram_amd.h"
#include "status.h"

using mcs::fail;

namespace {

// The ATLEAST tail is a dynamic programme over count states held in
// registers: min(k, n - k + 1) + 1 of them (counting true inputs up to k,
// or false ones up to n - k), at most kMaxStates.
constexpr int kMaxStates = 17;

struct IndepGate {       // 16 bytes
    uint32_t type;       // connective
    uint32_t k;          // ATLEAST threshold
    uint32_t arg_begin;  // into the argument array
    uint32_t n_in;
};

typedef const __attribute__((address_space(4))) uint32_t const_u32;

__device__ __forceinline__ uint32_t sload(const uint32_t *p, uint64_t i) {
    return ((const_u32 *)p)[i];
}

// p[i][s] = P[s][i]: the caller's sample-major probabilities, node-major.
__global__ void to_node_major(const float *__restrict__ P, int64_t n, int32_t n_vars,
                              double *__restrict__ val, int64_t stride) {
    __shared__ float tile[32][33];
    const int64_t s0 = int64_t(blockIdx.x) * 32, i0 = int64_t(blockIdx.y) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int r = ty; r < 32; r += 8) {
        const int64_t s = s0 + r, i = i0 + tx;
        tile[r][tx] = (s < n && i < n_vars) ? P[s * n_vars + i] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = i0 + r, s = s0 + tx;
        if (i < n_vars && s < n) val[i * stride + s] = double(tile[tx][r]);
    }
}

// Argument word: node id << 1 | complemented.
__device__ __forceinline__ double arg_p(const double *__restrict__ val, int64_t stride, int64_t s,
                                        uint32_t a) {
    const double q = val[int64_t(a >> 1) * stride + s];
    return (a & 1u) ? 1.0 - q : q;
}

__global__ __launch_bounds__(256) void indep_eval(const IndepGate *__restrict__ gates,
                                                  const uint32_t *__restrict__ args,
                                                  int32_t n_gates, int32_t first_gate_node,
                                                  double *__restrict__ val, int64_t stride,
                                                  int64_t n, uint32_t root_node,
                                                  double *__restrict__ logit) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t *gd = reinterpret_cast<const uint32_t *>(gates);
    for (int32_t g = 0; g < n_gates; ++g) {
        const uint32_t type = sload(gd, 4u * g), k = sload(gd, 4u * g + 1);
        const uint32_t b = sload(gd, 4u * g + 2), m = sload(gd, 4u * g + 3);
        double v = 0.0;
        if (type == MCS_AND || type == MCS_NAND) {
            v = 1.0;
            for (uint32_t i = 0; i < m; ++i) v *= arg_p(val, stride, s, sload(args, b + i));
        } else if (type == MCS_OR || type == MCS_NOR) {
            double c = 1.0;
            for (uint32_t i = 0; i < m; ++i) c *= 1.0 - arg_p(val, stride, s, sload(args, b + i));
            v = 1.0 - c;
        } else if (type == MCS_XOR) {
            for (uint32_t i = 0; i < m; ++i) {
                const double q = arg_p(val, stride, s, sload(args, b + i));
                v = v * (1.0 - q) + (1.0 - v) * q;
            }
        } else if (type == MCS_NOT || type == MCS_NULL) {
            v = m ? arg_p(val, stride, s, sload(args, b + m - 1)) : 0.0;
        } else {  // ATLEAST: the host checked min(k, m - k + 1) + 1 <= kMaxStates
            // Count true inputs up to k (tail = the state "k or more"), or,
            // when that is shorter, false inputs up to m - k (the gate holds
            // while at most m - k are false).
            const bool by_false = k > m - k + 1;
            const uint32_t cap = by_false ? m - k + 1 : k;  // tail state index
            double dp[kMaxStates];
#pragma unroll
            for (int j = 0; j < kMaxStates; ++j) dp[j] = j == 0 ? 1.0 : 0.0;
            for (uint32_t i = 0; i < m; ++i) {
                double q = arg_p(val, stride, s, sload(args, b + i));
                if (by_false) q = 1.0 - q;
#pragma unroll
                for (int j = kMaxStates - 1; j >= 1; --j) {
                    if (uint32_t(j) > cap) continue;
                    dp[j] = (uint32_t(j) == cap ? dp[j] : dp[j] * (1.0 - q)) + dp[j - 1] * q;
                }
                dp[0] *= 1.0 - q;
            }
            double tail = 0.0;
#pragma unroll
            for (int j = 0; j < kMaxStates; ++j)
                if (uint32_t(j) == cap) tail = dp[j];
            v = k == 0 ? 1.0 : (k > m ? 0.0 : (by_false ? 1.0 - tail : tail));
        }
        if (type == MCS_NOT || type == MCS_NAND || type == MCS_NOR) v = 1.0 - v;
        val[int64_t(first_gate_node + g) * stride + s] = v;
    }
    const double p = val[int64_t(root_node) * stride + s];
    const double lo = 1e-300, hi = 1.0 - 0x1p-52;
    const double pc = p < lo ? lo : (p > hi ? hi : p);
    logit[s] = std::log(pc) - std::log1p(-pc);
}

}  // namespace

extern "C" mcs_status mcs_independence_logits(const mcs_pdag *pdag, const float *p_samples,
                                              int64_t n_samples, double *logit_out) {
    if (!pdag || !p_samples || !logit_out || n_samples < 0 || pdag->n_vars < 0 ||
        pdag->n_gates <= 0 || !pdag->arg_offsets || !pdag->args || !pdag->gate_index ||
        !pdag->gate_type || !pdag->gate_k || (pdag->n_vars > 0 && !pdag->var_index))
        return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n_samples == 0) return MCS_OK;
    const int32_t nv = pdag->n_vars;
    // Nodes: variables 0 .. nv-1 in the caller's column order, then the
    // gates reachable from the root in post-order.
    std::unordered_map<int32_t, int32_t> var_slot, gate_slot;
    for (int32_t i = 0; i < nv; ++i)
        if (!var_slot.emplace(pdag->var_index[i], i).second)
            return fail(MCS_ERR_GRAPH, "duplicate variable index");
    for (int32_t g = 0; g < pdag->n_gates; ++g)
        if (var_slot.count(pdag->gate_index[g]) || !gate_slot.emplace(pdag->gate_index[g], g).second)
            return fail(MCS_ERR_GRAPH, "duplicate gate index");
    auto root_it = gate_slot.find(pdag->root);
    if (root_it == gate_slot.end()) return fail(MCS_ERR_GRAPH, "the root is not a gate");
    std::vector<int32_t> order;             // gate slots, post-order
    std::vector<uint8_t> state(size_t(pdag->n_gates), 0);  // 1 on stack, 2 done
    std::vector<std::pair<int32_t, int64_t>> stack{{root_it->second, pdag->arg_offsets[root_it->second]}};
    state[size_t(root_it->second)] = 1;
    while (!stack.empty()) {
        auto &[g, next] = stack.back();
        if (next == pdag->arg_offsets[g + 1]) {
            state[size_t(g)] = 2;
            order.push_back(g);
            stack.pop_back();
            continue;
        }
        const int32_t a = std::abs(pdag->args[next++]);
        if (var_slot.count(a)) continue;
        auto it = gate_slot.find(a);
        if (it == gate_slot.end()) return fail(MCS_ERR_GRAPH, "argument " + std::to_string(a) + " is not a node");
        if (state[size_t(it->second)] == 1) return fail(MCS_ERR_GRAPH, "cycle through gate " + std::to_string(a));
        if (state[size_t(it->second)] == 0) {
            state[size_t(it->second)] = 1;
            stack.push_back({it->second, pdag->arg_offsets[it->second]});
        }
    }
    std::vector<int32_t> node_of_gate(size_t(pdag->n_gates), -1);
    for (size_t t = 0; t < order.size(); ++t) node_of_gate[size_t(order[t])] = nv + int32_t(t);
    std::vector<IndepGate> gates;
    std::vector<uint32_t> args;
    for (int32_t g : order) {
        IndepGate d{};
        d.type = pdag->gate_type[g];
        d.k = uint32_t(std::max(pdag->gate_k[g], 0));
        d.arg_begin = uint32_t(args.size());
        d.n_in = uint32_t(pdag->arg_offsets[g + 1] - pdag->arg_offsets[g]);
        if (d.type > MCS_NULL) return fail(MCS_ERR_GRAPH, "unknown connective");
        if (d.type == MCS_ATLEAST && d.k >= 1 && d.k <= d.n_in &&
            std::min(d.k, d.n_in - d.k + 1) + 1 > uint32_t(kMaxStates))
            return fail(MCS_ERR_INVALID_ARGUMENT,
                        "ATLEAST gate " + std::to_string(pdag->gate_index[g]) +
                            " needs more count states than the baseline keeps (" +
                            std::to_string(kMaxStates) + ")");
        // NOT and NULL take their last input in the reference's order:
        // positive variables, positive gates, complemented variables,
        // complemented gates (src/mc/queue/kernel_builder.h:298-426,
        // src/mc/kernel/gate.h:321-351); only that input is kept.
        int64_t j0 = pdag->arg_offsets[g], j1 = pdag->arg_offsets[g + 1];
        if ((d.type == MCS_NOT || d.type == MCS_NULL) && j1 > j0) {
            int64_t best = j0;
            int best_rank = -1;
            for (int64_t j = j0; j < j1; ++j) {
                const int32_t a = pdag->args[j];
                const int rank = (a < 0 ? 2 : 0) + (var_slot.count(std::abs(a)) ? 0 : 1);
                if (rank >= best_rank) {
                    best_rank = rank;
                    best = j;
                }
            }
            j0 = best;
            j1 = best + 1;
            d.n_in = 1;
        }
        for (int64_t j = j0; j < j1; ++j) {
            const int32_t a = pdag->args[j];
            const int32_t abs_a = std::abs(a);
            auto v = var_slot.find(abs_a);
            const int32_t node = v != var_slot.end() ? v->second : node_of_gate[size_t(gate_slot.at(abs_a))];
            args.push_back(uint32_t(node) << 1 | (a < 0 ? 1u : 0u));
        }
        gates.push_back(d);
    }
    const int64_t n_nodes = int64_t(nv) + int64_t(order.size());
    const uint32_t root_node = uint32_t(node_of_gate[size_t(root_it->second)]);

    // Samples in chunks whose node matrix stays within 4 GiB.
    const int64_t chunk = std::max<int64_t>(256, std::min<int64_t>(
                                                     n_samples, (int64_t(4) << 30) / (8 * n_nodes)));
    IndepGate *d_gates = nullptr;
    uint32_t *d_args = nullptr;
    float *d_p = nullptr;
    double *d_val = nullptr, *d_logit = nullptr;
    hipStream_t stream = nullptr;
    auto cleanup = [&]() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *q : {static_cast<void *>(d_gates), static_cast<void *>(d_args),
                        static_cast<void *>(d_p), static_cast<void *>(d_val), static_cast<void *>(d_logit)})
            if (q) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    };
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d_gates, gates.size() * sizeof(IndepGate));
    if (e == hipSuccess) e = hipMalloc(&d_args, std::max<size_t>(args.size(), 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&d_p, size_t(chunk) * size_t(std::max(nv, 1)) * 4);
    if (e == hipSuccess) e = hipMalloc(&d_val, size_t(n_nodes) * size_t(chunk) * 8);
    if (e == hipSuccess) e = hipMalloc(&d_logit, size_t(chunk) * 8);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_gates, gates.data(), gates.size() * sizeof(IndepGate), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && !args.empty())
        e = hipMemcpyAsync(d_args, args.data(), args.size() * 4, hipMemcpyHostToDevice, stream);
    for (int64_t a = 0; a < n_samples && e == hipSuccess; a += chunk) {
        const int64_t n = std::min(chunk, n_samples - a);
        if (nv > 0) {
            e = hipMemcpyAsync(d_p, p_samples + a * nv, size_t(n) * size_t(nv) * 4, hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(to_node_major, dim3(uint32_t((n + 31) / 32), uint32_t((nv + 31) / 32)), dim3(256),
                               0, stream, d_p, n, nv, d_val, chunk);
            e = hipGetLastError();
            if (e != hipSuccess) break;
        }
        hipLaunchKernelGGL(indep_eval, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream, d_gates, d_args,
                           int32_t(gates.size()), nv, d_val, chunk, n, root_node, d_logit);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(logit_out + a, d_logit, size_t(n) * 8, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
    cleanup();
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? MCS_ERR_OUT_OF_MEMORY : MCS_ERR_DEVICE,
                    std::string("independence baseline: ") + hipGetErrorString(e));
    return MCS_OK;
}
