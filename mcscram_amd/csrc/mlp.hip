// Surrogate network for uncertainty analysis: an MLP n_in -> H -> H -> H -> 1
// with GELU, BF16 weights and activations, FP32 accumulation, MSE loss,
// Adam on FP32 master weights. Forward and backward products run on the
// matrix cores (v_mfma_f32_16x16x32_bf16) through one strided GEMM kernel;
// the rest are elementwise and reduction kernels.
//
// Shapes: batch rows B (a multiple of 64), input width n_in padded with
// zero columns to Dp (a multiple of 64), hidden width H (a multiple of 64).
// Parameters (FP32 master order): W1 [n_in x H], b1 [H], W2 [H x H], b2,
// W3 [H x H], b3, w4 [H], b4 [1]; x @ W is the layer product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "mcscram_amd.h"
#include "status.h"

using mcs::fail;

namespace mcs {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return uint16_t((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__host__ __device__ __forceinline__ float from_bf16(uint16_t h) {
    const uint32_t u = uint32_t(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// GELU(x) = x * Phi(x) and its derivative Phi(x) + x * phi(x).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

struct GemmArgs {
    const uint16_t *A;  // element (m, k) at A[m * a_sm + k * a_sk]
    int64_t a_sm, a_sk;
    const uint16_t *B;  // element (k, n) at B[k * b_sk + n * b_sn]
    int64_t b_sk, b_sn;
    float *C;           // element (m, n) at C[m * c_sm + n]
    int64_t c_sm;
    int32_t M, N, K;    // M, N multiples of 64, K a multiple of 32
    // Split-K: blockIdx.z covers k-steps [z * steps, (z + 1) * steps) and,
    // with more than one split, writes its 64 x 64 partial to
    // partial[z][m][n] (M x N floats each) for gemm_reduce; steps = all
    // k-steps and partial unused otherwise.
    int32_t steps;
    float *partial;
};

constexpr int kTileM = 64, kTileN = 64, kTileK = 32, kPadK = kTileK + 8;

// 8 consecutive k of one row of A (or one column of B) as one 16-byte word:
// a single load when the k stride is 1, eight 2-byte loads otherwise.
__device__ __forceinline__ uint4 load8(const uint16_t *p, int64_t stride) {
    if (stride == 1) return *reinterpret_cast<const uint4 *>(p);
    uint16_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[int64_t(j) * stride];
    uint4 r;
    __builtin_memcpy(&r, v, 16);
    return r;
}

// C = A x B. Workgroup: 64 x 64 outputs, four waves of 32 x 32, each a 2 x 2
// grid of 16 x 16 MFMA tiles; K in steps of 32 through LDS. A is staged as
// [m][k] and B transposed as [n][k], so both operand fragments (8 BF16 of
// one row and consecutive k) are single 16-byte LDS reads. The next step's
// operands are loaded into registers while the matrix cores work on the
// current one.
__global__ __launch_bounds__(256) void gemm_bf16(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kTileM][kPadK];
    __shared__ __attribute__((aligned(16))) uint16_t Bt[kTileN][kPadK];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * kTileM, n0 = blockIdx.x * kTileN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int r = lane & 15, q = lane >> 4;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Each thread stages 8 consecutive k of one row of A and of one column of B.
    const int sm = tid >> 2, sk = (tid & 3) * 8;
    const int k_begin = int(blockIdx.z) * g.steps * kTileK;
    const int k_end = min(g.K, k_begin + g.steps * kTileK);
    const uint16_t *pa = g.A + int64_t(m0 + sm) * g.a_sm + int64_t(k_begin + sk) * g.a_sk;
    const uint16_t *pb = g.B + int64_t(k_begin + sk) * g.b_sk + int64_t(n0 + sm) * g.b_sn;
    const int64_t a_step = int64_t(kTileK) * g.a_sk, b_step = int64_t(kTileK) * g.b_sk;
    uint4 ra = load8(pa, g.a_sk), rb = load8(pb, g.b_sk);
    for (int k0 = k_begin; k0 < k_end; k0 += kTileK) {
        *reinterpret_cast<uint4 *>(&As[sm][sk]) = ra;
        *reinterpret_cast<uint4 *>(&Bt[sm][sk]) = rb;
        __syncthreads();
        if (k0 + kTileK < k_end) {
            pa += a_step;
            pb += b_step;
            ra = load8(pa, g.a_sk);
            rb = load8(pb, g.b_sk);
        }
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + i * 16 + r][q * 8]);
#pragma unroll
        for (int j = 0; j < 2; ++j) bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bt[wn + j * 16 + r][q * 8]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // Accumulator layout of a 16 x 16 tile: lane holds column (lane % 16),
    // rows 4 * (lane / 16) + 0..3.
    float *out = g.C;
    int64_t out_sm = g.c_sm;
    if (gridDim.z > 1) {
        out = g.partial + int64_t(blockIdx.z) * g.M * g.N;
        out_sm = g.N;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + i * 16 + q * 4 + v, col = n0 + wn + j * 16 + r;
                out[int64_t(row) * out_sm + col] = acc[i][j][v];
            }
}

// C = the sum of the split-K partials, in split order (deterministic).
__global__ void gemm_reduce(const float *__restrict__ partial, int32_t splits, int32_t M, int32_t N,
                            float *__restrict__ C, int64_t c_sm) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t mn = int64_t(M) * N;
    if (i >= mn) return;
    float s = 0.f;
    for (int32_t z = 0; z < splits; ++z) s += partial[int64_t(z) * mn + i];
    const int64_t row = i / N, col = i - row * N;
    C[row * c_sm + col] = s;
}

// Rows of x (FP32, n_in wide) gathered by index into BF16 rows Dp wide,
// zero-padded; y gathered alongside.
__global__ void gather_rows(const uint16_t *__restrict__ x, int32_t dp, const float *__restrict__ y,
                            const int32_t *__restrict__ idx, int32_t rows, uint16_t *__restrict__ xb,
                            float *__restrict__ yb) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = int64_t(rows) * dp;
    if (i < n) {
        const int64_t r = i / dp, c = i - r * dp;
        xb[i] = x[int64_t(idx[r]) * dp + c];
    }
    if (i < rows) yb[i] = y[idx[i]];
}

// FP32 [rows x n_in] -> BF16 [rows x dp] (zero columns beyond n_in). With
// `map` ([scale | shift | cap], n_in each) column c becomes
// min(cap_c, exp(scale_c x + shift_c)) before rounding.
__global__ void to_bf16_padded(const float *__restrict__ x, int64_t rows, int32_t n_in, int32_t dp,
                               const float *__restrict__ map, uint16_t *__restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows * dp) return;
    const int64_t r = i / dp, c = i - r * dp;
    if (c >= n_in) {
        out[i] = 0;
        return;
    }
    float v = x[r * n_in + c];
    if (map) v = fminf(map[2 * n_in + c], expf(fmaf(map[c], v, map[n_in + c])));
    out[i] = to_bf16(v);
}

// z = c + b (kept for the backward pass), a = BF16(GELU(z)).
__global__ void bias_gelu(float *__restrict__ z, const float *__restrict__ b, int32_t h, int64_t n,
                          uint16_t *__restrict__ a) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = z[i] + b[i % h];
    z[i] = v;
    a[i] = to_bf16(gelu(v));
}

// Output layer and loss: yhat[r] = sum_j a[r][j] * w4[j] + b4 (FP32 over
// BF16 operands); d[r] = 2 (yhat - y) / B; loss[r] = (yhat - y)^2.
__global__ void out_layer(const uint16_t *__restrict__ a, const uint16_t *__restrict__ w4,
                          const float *__restrict__ b4, int32_t h, int32_t rows,
                          const float *__restrict__ y, float *__restrict__ yhat,
                          float *__restrict__ d, float *__restrict__ sq) {
    const int r = blockIdx.x;  // one wave per row
    const int lane = threadIdx.x;
    float s = 0.f;
    for (int j = lane; j < h; j += 64) s += from_bf16(a[int64_t(r) * h + j]) * from_bf16(w4[j]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        const float v = s + b4[0];
        yhat[r] = v;
        if (y) {
            const float e = v - y[r];
            d[r] = 2.f * e / float(rows);
            sq[r] = e * e;
        }
    }
}

// Backward through the output layer and the last GELU:
// dz3[r][j] = BF16(d[r] * w4[j] * GELU'(z3[r][j])).
__global__ void out_back(const float *__restrict__ d, const uint16_t *__restrict__ w4,
                         const float *__restrict__ z, int32_t h, int64_t n,
                         uint16_t *__restrict__ dz) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = i / h;
    dz[i] = to_bf16(d[r] * from_bf16(w4[i % h]) * gelu_grad(z[i]));
}

// dz = BF16(da * GELU'(z)).
__global__ void gelu_back(const float *__restrict__ da, const float *__restrict__ z, int64_t n,
                          uint16_t *__restrict__ dz) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dz[i] = to_bf16(da[i] * gelu_grad(z[i]));
}

// Column sums over the rows of a BF16 [rows x h] matrix, each row weighted
// by w[r] when w is given: out[j] = sum_r w[r] m[r][j] (bias gradients, and
// the output layer's weight gradient). A block of 16 waves covers 64
// columns, one per lane; wave k takes rows k, k + 16, ..., so every load of
// a wave is one coalesced 128-byte row segment, and the 16 partial sums
// meet in LDS in wave order: the result does not depend on scheduling.
constexpr int kColWaves = 16;

__global__ __launch_bounds__(64 * kColWaves) void col_sums(const uint16_t *__restrict__ m,
                                                           const float *__restrict__ w,
                                                           int32_t rows, int32_t h,
                                                           float *__restrict__ out) {
    __shared__ float part[kColWaves][64];
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    const int j = int(blockIdx.x) * 64 + lane;
    float s = 0.f;
    if (j < h) {
        for (int r = wave; r < rows; r += kColWaves) {
            const float v = from_bf16(m[int64_t(r) * h + j]);
            s = w ? fmaf(v, w[r], s) : s + v;
        }
    }
    part[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && j < h) {
        float t = 0.f;
        for (int k = 0; k < kColWaves; ++k) t += part[k][lane];
        out[j] = t;
    }
}

// Adam on FP32 master weights (bias-corrected), with decoupled weight decay
// wd (AdamW: w -= lr * wd * w beside the Adam step; 0 is plain Adam),
// refreshing the BF16 copy.
__global__ void adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                     float *__restrict__ v, uint16_t *__restrict__ pb, int64_t n, float lr,
                     float b1, float b2, float eps, float c1, float c2, float wd) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float w = p[i] - lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
    p[i] = w;
    if (pb) pb[i] = to_bf16(w);
}

__global__ void sum_vec(const float *__restrict__ x, int32_t n, float *__restrict__ out) {
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 64) s += x[i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) out[0] = s;
}

__global__ void cvt_bf16(const float *__restrict__ x, int64_t n, uint16_t *__restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = to_bf16(x[i]);
}

uint32_t blocks_for(int64_t n, int t = 256) { return uint32_t((n + t - 1) / t); }

}  // namespace

// Split-K workspace: enough for 512 workgroups' 64 x 64 partials.
constexpr int64_t kSplitWorkspaceFloats = int64_t(512) * kTileM * kTileN;

// With fewer output tiles than twice the CUs, K is split so that about 512
// workgroups run, each over at least four k-steps, when a workspace of
// kSplitWorkspaceFloats floats is given.
hipError_t launch_gemm_bf16(hipStream_t s, GemmArgs g, float *workspace = nullptr) {
    if (g.M % kTileM || g.N % kTileN || g.K % kTileK || g.M <= 0 || g.N <= 0 || g.K <= 0)
        return hipErrorInvalidValue;
    const int64_t tiles = int64_t(g.M / kTileM) * (g.N / kTileN);
    const int32_t ksteps = g.K / kTileK;
    int32_t splits = 1;
    if (workspace && tiles < 512) {
        splits = int32_t(std::min<int64_t>(512 / tiles, ksteps / 4));
        splits = std::max(splits, 1);
        if (int64_t(splits) * g.M * g.N > kSplitWorkspaceFloats) splits = 1;
    }
    g.steps = (ksteps + splits - 1) / splits;
    splits = (ksteps + g.steps - 1) / g.steps;
    g.partial = workspace;
    hipLaunchKernelGGL(gemm_bf16, dim3(g.N / kTileN, g.M / kTileM, splits), dim3(256), 0, s, g);
    if (splits > 1) {
        const int64_t mn = int64_t(g.M) * g.N;
        hipLaunchKernelGGL(gemm_reduce, dim3(uint32_t((mn + 255) / 256)), dim3(256), 0, s,
                           static_cast<const float *>(workspace), splits, g.M, g.N, g.C, g.c_sm);
    }
    return hipGetLastError();
}

}  // namespace mcs

// ---------------------------------------------------------------------------
// Host side.

struct mcs_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t n_in = 0, dp = 0, h = 0;
    int64_t n_params = 0;
    // FP32 master parameters, Adam moments and gradients share one layout:
    // [W1 (dp x h) | b1 | W2 | b2 | W3 | b3 | w4 | b4] (W1 rows >= n_in stay 0).
    int64_t off_w1 = 0, off_b1 = 0, off_w2 = 0, off_b2 = 0, off_w3 = 0, off_b3 = 0, off_w4 = 0,
            off_b4 = 0, n_dev = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *pb = nullptr;  // BF16 copy (same layout)
    float *map = nullptr;    // input map [scale | shift | cap] or null
    int64_t adam_t = 0;
    float weight_decay = 0.f;  // decoupled (AdamW); 0: plain Adam
    // Workspace for a batch of `cap` rows.
    int32_t cap = 0;
    uint16_t *xb = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *dz = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *da = nullptr, *yb = nullptr, *yhat = nullptr,
          *d = nullptr, *sq = nullptr, *scal = nullptr;
    int32_t *idx = nullptr;
    float *splitws = nullptr;  // split-K partials (mcs::kSplitWorkspaceFloats)
    std::vector<void *> ws;

    ~mcs_mlp() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        for (void *x : {static_cast<void *>(p), static_cast<void *>(g), static_cast<void *>(m),
                        static_cast<void *>(v), static_cast<void *>(pb), static_cast<void *>(map)})
            if (x) (void)hipFree(x);
        for (void *x : ws) (void)hipFree(x);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

mcs_status ensure_workspace(mcs_mlp *n, int32_t rows) {
    if (rows <= n->cap) return MCS_OK;
    for (void *x : n->ws) (void)hipFree(x);
    n->ws.clear();
    auto alloc = [&](auto **ptr, size_t bytes) -> hipError_t {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, bytes);
        if (e == hipSuccess) n->ws.push_back(q);
        *ptr = static_cast<std::remove_reference_t<decltype(**ptr)> *>(q);
        return e;
    };
    const size_t R = size_t(rows), H = size_t(n->h);
    MCS_HIP(alloc(&n->xb, R * size_t(n->dp) * 2));
    MCS_HIP(alloc(&n->a1, R * H * 2));
    MCS_HIP(alloc(&n->a2, R * H * 2));
    MCS_HIP(alloc(&n->a3, R * H * 2));
    MCS_HIP(alloc(&n->dz, R * H * 2));
    MCS_HIP(alloc(&n->z1, R * H * 4));
    MCS_HIP(alloc(&n->z2, R * H * 4));
    MCS_HIP(alloc(&n->z3, R * H * 4));
    MCS_HIP(alloc(&n->da, R * std::max<size_t>(H, 1) * 4));
    MCS_HIP(alloc(&n->yb, R * 4));
    MCS_HIP(alloc(&n->yhat, R * 4));
    MCS_HIP(alloc(&n->d, R * 4));
    MCS_HIP(alloc(&n->sq, R * 4));
    MCS_HIP(alloc(&n->scal, 64 * 4));
    MCS_HIP(alloc(&n->idx, R * 4));
    MCS_HIP(alloc(&n->splitws, size_t(mcs::kSplitWorkspaceFloats) * 4));
    n->cap = rows;
    return MCS_OK;
}

mcs::GemmArgs gemm(const uint16_t *A, int64_t a_sm, int64_t a_sk, const uint16_t *B, int64_t b_sk,
                   int64_t b_sn, float *C, int64_t c_sm, int32_t M, int32_t N, int32_t K) {
    return mcs::GemmArgs{A, a_sm, a_sk, B, b_sk, b_sn, C, c_sm, M, N, K};
}

// Forward over the rows staged in n->xb (and targets in n->yb if y).
mcs_status forward(mcs_mlp *n, int32_t rows, bool with_loss) {
    hipStream_t s = n->stream;
    const int32_t H = n->h;
    const int64_t RH = int64_t(rows) * H;
    const uint16_t *W1 = n->pb + n->off_w1, *W2 = n->pb + n->off_w2, *W3 = n->pb + n->off_w3;
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->xb, n->dp, 1, W1, H, 1, n->z1, H, rows, H, n->dp), n->splitws));
    hipLaunchKernelGGL(mcs::bias_gelu, dim3(mcs::blocks_for(RH)), dim3(256), 0, s, n->z1,
                       n->p + n->off_b1, H, RH, n->a1);
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->a1, H, 1, W2, H, 1, n->z2, H, rows, H, H), n->splitws));
    hipLaunchKernelGGL(mcs::bias_gelu, dim3(mcs::blocks_for(RH)), dim3(256), 0, s, n->z2,
                       n->p + n->off_b2, H, RH, n->a2);
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->a2, H, 1, W3, H, 1, n->z3, H, rows, H, H), n->splitws));
    hipLaunchKernelGGL(mcs::bias_gelu, dim3(mcs::blocks_for(RH)), dim3(256), 0, s, n->z3,
                       n->p + n->off_b3, H, RH, n->a3);
    hipLaunchKernelGGL(mcs::out_layer, dim3(rows), dim3(64), 0, s, n->a3, n->pb + n->off_w4,
                       n->p + n->off_b4, H, rows, with_loss ? n->yb : nullptr, n->yhat, n->d, n->sq);
    MCS_HIP(hipGetLastError());
    return MCS_OK;
}

// Backward into n->g (gradients of the mean squared error over `rows`).
mcs_status backward(mcs_mlp *n, int32_t rows) {
    hipStream_t s = n->stream;
    const int32_t H = n->h;
    const int64_t RH = int64_t(rows) * H;
    float *g = n->g;
    const uint16_t *W2 = n->pb + n->off_w2, *W3 = n->pb + n->off_w3;
    // Output layer: gw4[j] = sum_r a3[r][j] d[r], gb4 = sum_r d[r].
    hipLaunchKernelGGL(mcs::col_sums, dim3(mcs::blocks_for(H, 64)), dim3(64 * mcs::kColWaves), 0, s,
                       n->a3, n->d, rows, H, g + n->off_w4);
    hipLaunchKernelGGL(mcs::sum_vec, dim3(1), dim3(64), 0, s, n->d, rows, g + n->off_b4);
    hipLaunchKernelGGL(mcs::out_back, dim3(mcs::blocks_for(RH)), dim3(256), 0, s, n->d,
                       n->pb + n->off_w4, n->z3, H, RH, n->dz);
    // Layer 3: gW3 = a2^T dz3, gb3, da2 = dz3 W3^T.
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->a2, 1, H, n->dz, H, 1, g + n->off_w3, H, H, H, rows), n->splitws));
    hipLaunchKernelGGL(mcs::col_sums, dim3(mcs::blocks_for(H, 64)), dim3(64 * mcs::kColWaves), 0, s,
                       n->dz, nullptr, rows, H, g + n->off_b3);
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->dz, H, 1, W3, 1, H, n->da, H, rows, H, H), n->splitws));
    hipLaunchKernelGGL(mcs::gelu_back, dim3(mcs::blocks_for(RH)), dim3(256), 0, s, n->da, n->z2, RH,
                       n->dz);
    // Layer 2.
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->a1, 1, H, n->dz, H, 1, g + n->off_w2, H, H, H, rows), n->splitws));
    hipLaunchKernelGGL(mcs::col_sums, dim3(mcs::blocks_for(H, 64)), dim3(64 * mcs::kColWaves), 0, s,
                       n->dz, nullptr, rows, H, g + n->off_b2);
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->dz, H, 1, W2, 1, H, n->da, H, rows, H, H), n->splitws));
    hipLaunchKernelGGL(mcs::gelu_back, dim3(mcs::blocks_for(RH)), dim3(256), 0, s, n->da, n->z1, RH,
                       n->dz);
    // Layer 1: gW1 = x^T dz1 (dp x h), gb1.
    MCS_HIP(mcs::launch_gemm_bf16(s, gemm(n->xb, 1, n->dp, n->dz, H, 1, g + n->off_w1, H, n->dp, H, rows), n->splitws));
    hipLaunchKernelGGL(mcs::col_sums, dim3(mcs::blocks_for(H, 64)), dim3(64 * mcs::kColWaves), 0, s,
                       n->dz, nullptr, rows, H, g + n->off_b1);
    MCS_HIP(hipGetLastError());
    return MCS_OK;
}

mcs_status adam_step(mcs_mlp *n, float lr) {
    ++n->adam_t;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, float(n->adam_t)), c2 = 1.f - std::pow(b2, float(n->adam_t));
    hipLaunchKernelGGL(mcs::adam, dim3(mcs::blocks_for(n->n_dev)), dim3(256), 0, n->stream, n->p,
                       n->g, n->m, n->v, n->pb, n->n_dev, lr, b1, b2, eps, c1, c2,
                       n->weight_decay);
    MCS_HIP(hipGetLastError());
    return MCS_OK;
}

// Stage host rows [first, first + rows) of x (FP32, n_in wide) and y.
mcs_status stage_rows(mcs_mlp *n, const float *x, const float *y, int64_t first, int32_t rows) {
    float *tmp = nullptr;
    const size_t bytes = size_t(rows) * size_t(n->n_in) * 4;
    MCS_HIP(hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpyAsync(tmp, x + first * n->n_in, bytes, hipMemcpyHostToDevice, n->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mcs::to_bf16_padded, dim3(mcs::blocks_for(int64_t(rows) * n->dp)), dim3(256),
                           0, n->stream, tmp, int64_t(rows), n->n_in, n->dp, n->map, n->xb);
        e = hipGetLastError();
    }
    if (e == hipSuccess && y)
        e = hipMemcpyAsync(n->yb, y + first, size_t(rows) * 4, hipMemcpyHostToDevice, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    (void)hipFree(tmp);
    MCS_HIP(e);
    return MCS_OK;
}

// Rows whose count is not a multiple of 64 are padded with copies of row 0
// for the forward pass (predictions of the pad are dropped).
int32_t padded(int64_t rows) { return int32_t((rows + 63) / 64 * 64); }

}  // namespace

extern "C" {

mcs_status mcs_mlp_create(int32_t n_in, int32_t hidden, uint64_t seed, mcs_mlp **out) {
    if (!out || n_in <= 0 || hidden <= 0 || hidden % 64)
        return fail(MCS_ERR_INVALID_ARGUMENT, "n_in > 0 and hidden a positive multiple of 64");
    *out = nullptr;
    auto n = std::make_unique<mcs_mlp>();
    n->n_in = n_in;
    n->dp = (n_in + 63) / 64 * 64;
    n->h = hidden;
    const int64_t D = n->dp, H = hidden;
    n->off_w1 = 0;
    n->off_b1 = n->off_w1 + D * H;
    n->off_w2 = n->off_b1 + H;
    n->off_b2 = n->off_w2 + H * H;
    n->off_w3 = n->off_b2 + H;
    n->off_b3 = n->off_w3 + H * H;
    n->off_w4 = n->off_b3 + H;
    n->off_b4 = n->off_w4 + H;
    n->n_dev = n->off_b4 + 1;
    n->n_params = int64_t(n_in) * H + 3 * H + 2 * H * H + H + 1;
    // He-uniform initialisation from SplitMix64 (weights); zero biases.
    std::vector<float> p(size_t(n->n_dev), 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto uni = [&]() { return double(next() >> 11) * (1.0 / 9007199254740992.0); };
    auto init = [&](int64_t off, int64_t rows, int64_t cols, int64_t fan_in) {
        const double a = std::sqrt(6.0 / double(fan_in));
        for (int64_t r = 0; r < rows; ++r)
            for (int64_t c = 0; c < cols; ++c) p[size_t(off + r * cols + c)] = float((2.0 * uni() - 1.0) * a);
    };
    init(n->off_w1, n_in, H, n_in);  // rows >= n_in stay zero
    init(n->off_w2, H, H, H);
    init(n->off_w3, H, H, H);
    init(n->off_w4, 1, H, H);
    std::vector<uint16_t> pb(p.size());
    for (size_t i = 0; i < p.size(); ++i) pb[i] = mcs::to_bf16(p[i]);
    MCS_HIP(hipGetDevice(&n->device));
    MCS_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    const size_t bytes = size_t(n->n_dev) * 4;
    MCS_HIP(hipMalloc(&n->p, bytes));
    MCS_HIP(hipMalloc(&n->g, bytes));
    MCS_HIP(hipMalloc(&n->m, bytes));
    MCS_HIP(hipMalloc(&n->v, bytes));
    MCS_HIP(hipMalloc(&n->pb, size_t(n->n_dev) * 2));
    MCS_HIP(hipMemcpy(n->p, p.data(), bytes, hipMemcpyHostToDevice));
    MCS_HIP(hipMemcpy(n->pb, pb.data(), pb.size() * 2, hipMemcpyHostToDevice));
    MCS_HIP(hipMemset(n->g, 0, bytes));
    MCS_HIP(hipMemset(n->m, 0, bytes));
    MCS_HIP(hipMemset(n->v, 0, bytes));
    *out = n.release();
    return MCS_OK;
}

mcs_status mcs_mlp_destroy(mcs_mlp *net) {
    delete net;
    return MCS_OK;
}

mcs_status mcs_mlp_num_params(mcs_mlp *net, int64_t *n) {
    if (!net || !n) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    *n = net->n_params;
    return MCS_OK;
}

// Host order: W1 [n_in x h] | b1 | W2 | b2 | W3 | b3 | w4 | b4.
static mcs_status copy_params(mcs_mlp *net, float *host, int64_t n, bool to_host, const float *src_dev) {
    if (n != net->n_params) return fail(MCS_ERR_INVALID_ARGUMENT, "parameter count mismatch");
    std::vector<float> all(size_t(net->n_dev));
    const float *src = src_dev ? src_dev : net->p;
    if (to_host) {
        MCS_HIP(hipStreamSynchronize(net->stream));
        MCS_HIP(hipMemcpy(all.data(), src, all.size() * 4, hipMemcpyDeviceToHost));
    }
    const int64_t w1 = int64_t(net->n_in) * net->h;
    const int64_t rest = net->n_dev - net->off_b1;
    if (to_host) {
        std::memcpy(host, all.data(), size_t(w1) * 4);
        std::memcpy(host + w1, all.data() + net->off_b1, size_t(rest) * 4);
        return MCS_OK;
    }
    std::fill(all.begin(), all.end(), 0.f);
    std::memcpy(all.data(), host, size_t(w1) * 4);
    std::memcpy(all.data() + net->off_b1, host + w1, size_t(rest) * 4);
    std::vector<uint16_t> pb(all.size());
    for (size_t i = 0; i < all.size(); ++i) pb[i] = mcs::to_bf16(all[i]);
    MCS_HIP(hipStreamSynchronize(net->stream));
    MCS_HIP(hipMemcpy(net->p, all.data(), all.size() * 4, hipMemcpyHostToDevice));
    MCS_HIP(hipMemcpy(net->pb, pb.data(), pb.size() * 2, hipMemcpyHostToDevice));
    return MCS_OK;
}

mcs_status mcs_mlp_get_params(mcs_mlp *net, float *out, int64_t n) {
    if (!net || !out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    return copy_params(net, out, n, true, nullptr);
}

mcs_status mcs_mlp_set_params(mcs_mlp *net, const float *in, int64_t n) {
    if (!net || !in) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    net->adam_t = 0;
    MCS_HIP(hipMemset(net->m, 0, size_t(net->n_dev) * 4));
    MCS_HIP(hipMemset(net->v, 0, size_t(net->n_dev) * 4));
    return copy_params(net, const_cast<float *>(in), n, false, nullptr);
}

mcs_status mcs_mlp_set_weight_decay(mcs_mlp *net, float weight_decay) {
    if (!net) return fail(MCS_ERR_INVALID_ARGUMENT, "null network");
    if (!(weight_decay >= 0.f) || weight_decay > 1.f)
        return fail(MCS_ERR_INVALID_ARGUMENT, "weight decay must be in [0, 1]");
    net->weight_decay = weight_decay;
    return MCS_OK;
}

mcs_status mcs_mlp_set_input_map(mcs_mlp *net, const float *scale, const float *shift,
                                 const float *cap) {
    if (!net) return fail(MCS_ERR_INVALID_ARGUMENT, "null network");
    const int nulls = !scale + !shift + !cap;
    if (nulls != 0 && nulls != 3)
        return fail(MCS_ERR_INVALID_ARGUMENT, "input map: give scale, shift and cap, or none");
    MCS_HIP(hipSetDevice(net->device));
    MCS_HIP(hipStreamSynchronize(net->stream));
    if (nulls == 3) {
        if (net->map) MCS_HIP(hipFree(net->map));
        net->map = nullptr;
        return MCS_OK;
    }
    const size_t n = size_t(net->n_in);
    std::vector<float> h(3 * n);
    for (size_t c = 0; c < n; ++c) {
        if (!std::isfinite(scale[c]) || !std::isfinite(shift[c]) || std::isnan(cap[c]))
            return fail(MCS_ERR_INVALID_ARGUMENT, "input map: non-finite scale or shift, or NaN cap");
        h[c] = scale[c];
        h[n + c] = shift[c];
        h[2 * n + c] = cap[c];
    }
    if (!net->map) MCS_HIP(hipMalloc(&net->map, h.size() * 4));
    MCS_HIP(hipMemcpy(net->map, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    return MCS_OK;
}

mcs_status mcs_mlp_predict(mcs_mlp *net, const float *x, int64_t n_rows, float *y_out) {
    if (!net || !x || !y_out || n_rows < 0) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    MCS_HIP(hipSetDevice(net->device));
    const int32_t chunk = 4096;
    mcs_status st = ensure_workspace(net, chunk);
    if (st != MCS_OK) return st;
    std::vector<float> pad_x;
    for (int64_t first = 0; first < n_rows; first += chunk) {
        const int32_t rows = int32_t(std::min<int64_t>(chunk, n_rows - first));
        const int32_t prow = padded(rows);
        const float *src = x;
        int64_t f = first;
        if (prow != rows) {  // pad the last chunk with copies of its first row
            pad_x.assign(size_t(prow) * size_t(net->n_in), 0.f);
            std::memcpy(pad_x.data(), x + first * net->n_in, size_t(rows) * size_t(net->n_in) * 4);
            for (int32_t r = rows; r < prow; ++r)
                std::memcpy(pad_x.data() + size_t(r) * size_t(net->n_in), x + first * net->n_in,
                            size_t(net->n_in) * 4);
            src = pad_x.data();
            f = 0;
        }
        st = stage_rows(net, src, nullptr, f, prow);
        if (st != MCS_OK) return st;
        st = forward(net, prow, false);
        if (st != MCS_OK) return st;
        MCS_HIP(hipMemcpyAsync(y_out + first, net->yhat, size_t(rows) * 4, hipMemcpyDeviceToHost,
                               net->stream));
        MCS_HIP(hipStreamSynchronize(net->stream));
    }
    return MCS_OK;
}

mcs_status mcs_mlp_loss_grad(mcs_mlp *net, const float *x, const float *y, int64_t n_rows,
                             float *loss_out, float *grad_out) {
    if (!net || !x || !y || !loss_out) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (n_rows <= 0 || n_rows % 64 || n_rows > (int64_t(1) << 20))
        return fail(MCS_ERR_INVALID_ARGUMENT, "rows must be a positive multiple of 64");
    MCS_HIP(hipSetDevice(net->device));
    const int32_t rows = int32_t(n_rows);
    mcs_status st = ensure_workspace(net, rows);
    if (st == MCS_OK) st = stage_rows(net, x, y, 0, rows);
    if (st == MCS_OK) st = forward(net, rows, true);
    if (st == MCS_OK) st = backward(net, rows);
    if (st != MCS_OK) return st;
    hipLaunchKernelGGL(mcs::sum_vec, dim3(1), dim3(64), 0, net->stream, net->sq, rows, net->scal);
    MCS_HIP(hipGetLastError());
    float s = 0;
    MCS_HIP(hipMemcpyAsync(&s, net->scal, 4, hipMemcpyDeviceToHost, net->stream));
    MCS_HIP(hipStreamSynchronize(net->stream));
    *loss_out = s / float(rows);
    if (grad_out) return copy_params(net, grad_out, net->n_params, true, net->g);
    return MCS_OK;
}

mcs_status mcs_mlp_fit(mcs_mlp *net, const float *x, const float *y, int64_t n_rows, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float *epoch_loss) {
    if (!net || !x || !y) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    if (batch <= 0 || batch % 64 || n_rows < batch || epochs <= 0)
        return fail(MCS_ERR_INVALID_ARGUMENT, "batch a positive multiple of 64, rows >= batch");
    MCS_HIP(hipSetDevice(net->device));
    mcs_status st = ensure_workspace(net, batch);
    if (st != MCS_OK) return st;
    // The whole training set, BF16 and padded, stays on the device.
    uint16_t *xall = nullptr;
    float *yall = nullptr, *tmp = nullptr, *losses = nullptr;
    const int64_t steps = n_rows / batch;
    auto cleanup = [&]() {
        for (void *q : {static_cast<void *>(xall), static_cast<void *>(yall), static_cast<void *>(tmp),
                        static_cast<void *>(losses)})
            if (q) (void)hipFree(q);
    };
    hipError_t e = hipMalloc(&xall, size_t(n_rows) * size_t(net->dp) * 2);
    if (e == hipSuccess) e = hipMalloc(&yall, size_t(n_rows) * 4);
    if (e == hipSuccess) e = hipMalloc(&losses, size_t(steps) * 4);
    const int64_t chunk = 8192;
    if (e == hipSuccess) e = hipMalloc(&tmp, size_t(chunk) * size_t(net->n_in) * 4);
    for (int64_t f = 0; f < n_rows && e == hipSuccess; f += chunk) {
        const int64_t r = std::min(chunk, n_rows - f);
        e = hipMemcpyAsync(tmp, x + f * net->n_in, size_t(r) * size_t(net->n_in) * 4,
                           hipMemcpyHostToDevice, net->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(mcs::to_bf16_padded, dim3(mcs::blocks_for(r * net->dp)), dim3(256), 0,
                               net->stream, tmp, r, net->n_in, net->dp, net->map, xall + f * net->dp);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(yall, y, size_t(n_rows) * 4, hipMemcpyHostToDevice, net->stream);
    if (e != hipSuccess) {
        cleanup();
        MCS_HIP(e);
    }
    uint64_t st_rng = seed;
    auto next = [&]() {
        st_rng += 0x9E3779B97F4A7C15ull;
        uint64_t z = st_rng;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    std::vector<int32_t> perm(static_cast<size_t>(n_rows));
    std::vector<float> host_losses(static_cast<size_t>(steps));
    const double pi = 3.14159265358979323846;
    for (int32_t ep = 0; ep < epochs && st == MCS_OK; ++ep) {
        const float lr_ep = float(double(lr) * 0.5 * (1.0 + std::cos(pi * double(ep) / double(epochs))));
        for (int64_t i = 0; i < n_rows; ++i) perm[size_t(i)] = int32_t(i);
        for (int64_t i = n_rows - 1; i > 0; --i) {  // Fisher-Yates
            const int64_t j = int64_t(next() % uint64_t(i + 1));
            std::swap(perm[size_t(i)], perm[size_t(j)]);
        }
        for (int64_t sidx = 0; sidx < steps && st == MCS_OK; ++sidx) {
            e = hipMemcpyAsync(net->idx, perm.data() + sidx * batch, size_t(batch) * 4,
                               hipMemcpyHostToDevice, net->stream);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(mcs::gather_rows, dim3(mcs::blocks_for(int64_t(batch) * net->dp)),
                               dim3(256), 0, net->stream, xall, net->dp, yall, net->idx, batch, net->xb,
                               net->yb);
            st = forward(net, batch, true);
            if (st == MCS_OK) st = backward(net, batch);
            if (st == MCS_OK) st = adam_step(net, lr_ep);
            if (st == MCS_OK) {
                hipLaunchKernelGGL(mcs::sum_vec, dim3(1), dim3(64), 0, net->stream, net->sq, batch,
                                   losses + sidx);
                e = hipGetLastError();
                if (e != hipSuccess) break;
            }
            // The host vector of indices is reused next step: wait for the copy.
            if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
            if (e != hipSuccess) break;
        }
        if (st == MCS_OK && e == hipSuccess) {
            e = hipMemcpy(host_losses.data(), losses, size_t(steps) * 4, hipMemcpyDeviceToHost);
            double s = 0;
            for (float l : host_losses) s += l;
            if (epoch_loss) epoch_loss[ep] = float(s / double(steps * batch));
        }
        if (e != hipSuccess) break;
    }
    cleanup();
    if (st != MCS_OK) return st;
    MCS_HIP(e);
    return MCS_OK;
}

// Test hook: C = A x B for BF16 inputs given as FP32 host matrices (rounded
// to BF16), row-major A [M x K] and B [K x N]; transposes select which
// operand is read transposed from a row-major buffer of the swapped shape.
mcs_status mcs_mlp_gemm_test(const float *a, const float *b, int32_t M, int32_t N, int32_t K,
                             int32_t trans_a, int32_t trans_b, float *c) {
    if (!a || !b || !c) return fail(MCS_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<uint16_t> ha(size_t(M) * K), hb(size_t(K) * N);
    for (size_t i = 0; i < ha.size(); ++i) ha[i] = mcs::to_bf16(a[i]);
    for (size_t i = 0; i < hb.size(); ++i) hb[i] = mcs::to_bf16(b[i]);
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *dw = nullptr;
    MCS_HIP(hipMalloc(&da, ha.size() * 2));
    hipError_t e = hipMalloc(&db, hb.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&dc, size_t(M) * N * 4);
    if (e == hipSuccess) e = hipMemcpy(da, ha.data(), ha.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, hb.data(), hb.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        // a is [M x K] row-major, or [K x M] row-major when trans_a.
        const int64_t a_sm = trans_a ? 1 : K, a_sk = trans_a ? M : 1;
        const int64_t b_sk = trans_b ? 1 : N, b_sn = trans_b ? K : 1;
        e = hipMalloc(&dw, size_t(mcs::kSplitWorkspaceFloats) * 4);
        if (e == hipSuccess)
            e = mcs::launch_gemm_bf16(nullptr, gemm(da, a_sm, a_sk, db, b_sk, b_sn, dc, N, M, N, K),
                                      dw);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(c, dc, size_t(M) * N * 4, hipMemcpyDeviceToHost);
    for (void *q : {static_cast<void *>(da), static_cast<void *>(db), static_cast<void *>(dc),
                    static_cast<void *>(dw)})
        if (q) (void)hipFree(q);
    MCS_HIP(e);
    return MCS_OK;
}

}  // extern "C"
