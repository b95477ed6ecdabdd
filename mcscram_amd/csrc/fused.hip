// Persistent whole-tree-in-LDS kernel for batches of small trees (CDNA4).
//
// Workgroups of 64 to 1024 lanes pull work items (tree, tiles) from a queue in HBM.
// For each tile of W words a workgroup samples every basic event of the tree
// into LDS, then evaluates the gate layers from LDS with a workgroup barrier
// between layers. Lane l of a step owns word (l mod W) of node (l / W), so
// the W words of one node are W neighbouring lanes of one wave: their
// popcounts are summed with lane shuffles and added to the node's LDS
// counter by one lane. Counters are flushed to HBM with one 64-bit atomic
// per node per item.
//
// Gate semantics are those of the layered kernel (src/mc/kernel/gate.h:
// 262-264, 299-362; ATLEAST 506-555 with the 8-bit count); the sampler is
// the same Bernoulli word (src/mc/kernel/basic_event.h:112-142).
#include <hip/hip_runtime.h>

#include "device_ops.h"
#include "fused.h"

namespace mcs {
namespace {

constexpr uint32_t kFusedMaxLanes = 1024;
static_assert(sizeof(FusedGate) == 32, "gate descriptors are loaded as two 16-byte words");

// Sum of v over the aligned group of w lanes (w a power of two <= 64).
__device__ __forceinline__ uint32_t group_sum(uint32_t v, uint32_t w) {
    for (uint32_t off = 1; off < w; off <<= 1) v += __shfl_xor(v, int(off), 64);
    return v;
}

__device__ __forceinline__ uint64_t lds_word(const uint64_t *lds, uint32_t node, uint32_t log2_w,
                                             uint32_t w) {
    return lds[(node << log2_w) + w];
}

// A gate descriptor held in registers: two 16-byte words,
// a = {n_in, type_k_flags, in0, in1}, b = {in2, in3, in4, in5}.
struct GateRegs {
    uint4 a, b;
};

__device__ __forceinline__ GateRegs load_gate(const FusedGate *g) {
    const uint4 *p = reinterpret_cast<const uint4 *>(g);
    return {p[0], p[1]};
}

// The word of an input slot (node | negated << 31), complemented if negated.
__device__ __forceinline__ uint64_t slot_word(const uint64_t *lds, uint32_t slot, uint32_t log2_w,
                                              uint32_t w) {
    const uint64_t x = lds[((slot & 0x7FFFFFFFu) << log2_w) + w];
    const uint64_t neg = uint64_t(int64_t(int32_t(slot)) >> 31);  // all ones if negated
    return x ^ neg;
}

// Carry-save count of the inputs, then count >= k8 on the P planes; k8 at
// or above 2^P can never be met.
template <int P, typename Each>
__device__ __forceinline__ uint64_t atleast_planes(Each each, uint32_t k8) {
    uint64_t c[P];
#pragma unroll
    for (int b = 0; b < P; ++b) c[b] = 0;
    each([&](uint64_t v) {
#pragma unroll
        for (int b = 0; b < P; ++b) {
            const uint64_t t = c[b] & v;
            c[b] ^= v;
            v = t;
        }
    });
    if (P < 8 && k8 >= (1u << P)) return 0;
    return atleast_compare<uint64_t, P>(c, k8);
}

// Gate word. Descriptors are normalised on the host: NOT and NULL arrive as
// NOR and OR of their last input, and the unused inline slots point at an
// all-zeros (OR, XOR, NOR, ATLEAST) or all-ones (AND, NAND) LDS word, so the
// six inline inputs fold without conditions.
__device__ __forceinline__ uint64_t eval_gate_word(const GateRegs &g,
                                                   const uint32_t *__restrict__ args,
                                                   const uint64_t *lds, uint32_t log2_w,
                                                   uint32_t w) {
    const uint32_t n_in = g.a.x;
    const uint32_t type = g.a.y & 0xFFu;
    const uint32_t k8 = (g.a.y >> 8) & 0xFFu;
    uint64_t r;
    if (n_in <= kInlineInputs) {
        const uint32_t slot[kInlineInputs] = {g.a.z, g.a.w, g.b.x, g.b.y, g.b.z, g.b.w};
        uint64_t v[kInlineInputs];
#pragma unroll
        for (uint32_t j = 0; j < kInlineInputs; ++j) v[j] = slot_word(lds, slot[j], log2_w, w);
        if (type == 2u) {  // at most six inputs: three planes
            r = atleast_planes<3>(
                [&](auto &&acc) {
#pragma unroll
                    for (uint32_t j = 0; j < kInlineInputs; ++j) acc(v[j]);
                },
                k8);
        } else if (type == 0u || type == 5u) {
            r = v[0] & v[1] & v[2] & v[3] & v[4] & v[5];
        } else if (type == 3u) {
            r = v[0] ^ v[1] ^ v[2] ^ v[3] ^ v[4] ^ v[5];
        } else {
            r = v[0] | v[1] | v[2] | v[3] | v[4] | v[5];
        }
    } else {
        const uint32_t *a = args + g.a.z;
        auto each = [&](auto &&acc) {
            for (uint32_t i = 0; i < n_in; ++i) acc(slot_word(lds, a[i], log2_w, w));
        };
        if (type == 2u) {
            r = atleast_planes<8>(each, k8);  // the reference's 8-bit count
        } else if (type == 0u || type == 5u) {
            r = ~0ull;
            each([&](uint64_t v) { r &= v; });
        } else if (type == 3u) {
            r = 0;
            each([&](uint64_t v) { r ^= v; });
        } else {
            r = 0;
            each([&](uint64_t v) { r |= v; });
        }
    }
    if (type == 5u || type == 6u) r = ~r;  // NAND, NOR (and NOT, sent as NOR)
    return r;
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations, then for every wave, and leaves global loads in flight (the
// kernel stores nothing to global memory inside a tile).
__device__ __forceinline__ void lds_barrier() {
    __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

}  // namespace

__global__ void fused_begin_pass(uint32_t *state) {
    if (threadIdx.x == 0) {
        state[0] = 0;
        state[1] += 1;
    }
}

__global__ __launch_bounds__(1024) void fused_trees(FusedLaunch L) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_layer_off[kMaxFusedLayers + 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t lanes = blockDim.x;  // a multiple of 64
    const uint32_t iteration = L.state[1];
    const uint32_t c3base = (iteration + 1u) << 6;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(L.state, 1u);
        __syncthreads();
        // Wave-uniform: everything derived from it is loaded by the scalar unit.
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= L.n_items) break;
        const WorkItem wi = L.items[item];
        const FusedTree T = L.trees[wi.tree];
        const uint32_t log2_w = T.log2_w;
        const uint32_t W = 1u << log2_w;
        // LDS: W words per node, then the all-zeros and all-ones words that
        // pad unused gate inputs, then one counter per node.
        uint32_t *cnt = reinterpret_cast<uint32_t *>(lds + (uint64_t(T.n_nodes + 2u) << log2_w));
        for (uint32_t n = tid; n < T.n_nodes; n += lanes) cnt[n] = 0;
        for (uint32_t i = tid; i < 2u * W; i += lanes)
            lds[(uint64_t(T.n_nodes) << log2_w) + i] = i < W ? 0ull : ~0ull;
        __syncthreads();
        const EventDesc *events = L.events + T.ev_begin;
        const FusedGate *gates = L.gates + T.gate_begin;
        // Layer offsets of the tree, staged once per item.
        for (uint32_t l = tid; l <= T.n_layers; l += lanes) s_layer_off[l] = L.layer_off[T.layer_begin + l];
        __syncthreads();
        const uint32_t *layer_off = s_layer_off;
        for (uint32_t tile = 0; tile < wi.n_tiles; ++tile) {
            const uint64_t base = uint64_t(wi.first_tile + tile) << log2_w;
            // Basic events of the tile; the next item's descriptor is loaded
            // while the current word is computed. Descriptor fields:
            // {ctr0, ctr1, threshold, mode}.
            const uint32_t n_ev_items = T.n_ev << log2_w;
            const uint4 *ev4 = reinterpret_cast<const uint4 *>(events);
            // Prefetch indices are clamped to the last event: every load is
            // in bounds and unconditional.
            const uint32_t last_ev = T.n_ev ? T.n_ev - 1u : 0u;
            uint4 d = ev4[min(tid >> log2_w, last_ev)];
            for (uint32_t idx = tid; idx < n_ev_items; idx += lanes) {
                const uint32_t e = idx >> log2_w, w = idx & (W - 1u);
                const uint4 d_next = ev4[min((idx + lanes) >> log2_w, last_ev)];
                const uint64_t s = base + w;
                const uint32_t mode = d.w & 3u;
                uint64_t word = 0;
                if (s < T.n_bp) {
                    if (mode == kEventDraw)
                        word = bernoulli_word(d.x, d.y, uint32_t(s) + 1u, c3base, L.key0, L.key1,
                                              d.z);
                    else if (mode == kEventAlways)
                        word = ~0ull;
                }
                lds[idx] = word;
                if (d.w & kEventTallied) {
                    const uint32_t pc = group_sum(uint32_t(__popcll(word)), W);
                    if (w == 0 && pc) atomicAdd(cnt + e, pc);
                }
                d = d_next;
            }
            // Gate layers, each from LDS only. The descriptor of a lane's
            // first gate of layer l + 1 is loaded before layer l's barrier.
            uint32_t g_lo = layer_off[0];
            uint32_t g_hi = T.n_layers ? layer_off[1] : g_lo;
            // Prefetch indices are clamped to the tree's last gate.
            const uint32_t last_gate = layer_off[T.n_layers] - 1u;
            GateRegs gd_next = load_gate(gates + min(g_lo + (tid >> log2_w), last_gate));
            lds_barrier();
            for (uint32_t l = 0; l < T.n_layers; ++l) {
                const uint32_t g0 = g_lo;
                const uint32_t n_items = (g_hi - g0) << log2_w;
                GateRegs gd = gd_next;
                g_lo = g_hi;
                g_hi = l + 2 <= T.n_layers ? layer_off[l + 2] : g_hi;
                gd_next = load_gate(gates + min(g_lo + (tid >> log2_w), last_gate));
                for (uint32_t idx = tid; idx < n_items; idx += lanes) {
                    const uint32_t g = g0 + (idx >> log2_w), w = idx & (W - 1u);
                    if (idx != tid) gd = load_gate(gates + g);
                    const uint64_t r = eval_gate_word(gd, L.args, lds, log2_w, w);
                    const uint32_t node = T.n_ev + g;
                    lds[(node << log2_w) + w] = r;
                    if (gd.a.y & kGateTallied) {
                        const uint32_t pc =
                            group_sum(base + w < T.n_bp ? uint32_t(__popcll(r)) : 0u, W);
                        if (w == 0 && pc) atomicAdd(cnt + node, pc);
                    }
                }
                lds_barrier();
            }
        }
        for (uint32_t n = tid; n < T.n_nodes; n += lanes) {
            const uint32_t c = cnt[n];
            if (c) atomicAdd(L.ones + T.ones_begin + n, (unsigned long long)c);
        }
        __syncthreads();
    }
}

hipError_t fused_max_workgroups_per_cu(uint32_t lanes, uint32_t lds_bytes, int *out) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, fused_trees, int(lanes),
                                                        size_t(lds_bytes));
}

hipError_t launch_fused_begin_pass(hipStream_t stream, uint32_t *state) {
    hipLaunchKernelGGL(fused_begin_pass, dim3(1), dim3(64), 0, stream, state);
    return hipGetLastError();
}

hipError_t launch_fused(hipStream_t stream, const FusedLaunch &l, uint32_t grid, uint32_t lanes,
                        uint32_t lds_bytes) {
    if (grid == 0) return hipSuccess;
    if (lanes == 0 || lanes > kFusedMaxLanes || lanes % 64u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fused_trees, dim3(grid), dim3(lanes), lds_bytes, stream, l);
    return hipGetLastError();
}

}  // namespace mcs
