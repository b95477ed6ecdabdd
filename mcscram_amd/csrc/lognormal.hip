// Lognormal parameter samples of the basic events, drawn on the GPU, for the
// uncertainty-analysis outer loop. The reference draws each deviate on the
// host from std::mt19937 (src/expression/random_deviate.cc:91-108, called per
// sample from src/uncertainty_analysis.h:154-169); its Monte Carlo path has
// no device-side parameter sampling, so this stream is this engine's own and
// is not meant to reproduce the reference's deviates.
//
// Row r, variable i (row-major [rows][n_vars] outputs):
//   one Philox4x32-10 block per pair of variables (2j, 2j + 1):
//     key = (seed mod 2^32, seed >> 32),
//     counter = (j, r mod 2^32, r >> 32, kLognormalTag);
//   u1 = (top 53 bits of (y0, y1) + 1/2) / 2^53 in (0, 1),
//   u2 = top 53 bits of (y2, y3) / 2^53 in [0, 1);
//   Box-Muller: rad = sqrt(-2 ln u1), z_{2j} = rad cos(2 pi u2),
//   z_{2j+1} = rad sin(2 pi u2);
//   p = min(1, exp(mu_i + sigma_i z_i)), in FP64; z is stored as FP32 (the
//   surrogate network's input).
// The tag keeps these counters apart from the Monte Carlo sampler's, whose
// word 3 is ((iteration + 1) << 6) + generation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "device_ops.h"
#include "mcscram_amd.h"
#include "status.h"

using mcs::fail;

namespace {

constexpr uint32_t kLognormalTag = 0x4C4E524Du;

__global__ __launch_bounds__(256) void lognormal_rows(const double *__restrict__ mu,
                                                      const double *__restrict__ sigma,
                                                      int32_t n_vars, uint32_t key0, uint32_t key1,
                                                      uint64_t first_row, int64_t rows,
                                                      float *__restrict__ z_out,
                                                      double *__restrict__ p_out) {
    const int32_t pairs = (n_vars + 1) >> 1;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= rows * int64_t(pairs)) return;
    const int64_t r = t / pairs;
    const int32_t j = int32_t(t - r * pairs);
    const uint64_t row = first_row + uint64_t(r);
    uint32_t x0 = uint32_t(j), x1 = uint32_t(row), x2 = uint32_t(row >> 32), x3 = kLognormalTag;
    uint32_t k0 = key0, k1 = key1;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        mcs::philox_round(x0, x1, x2, x3, k0, k1);
        k0 += mcs::kPhiloxW0;
        k1 += mcs::kPhiloxW1;
    }
    const double two53 = 9007199254740992.0;
    const uint64_t b1 = ((uint64_t(x0) << 32) | x1) >> 11;
    const uint64_t b2 = ((uint64_t(x2) << 32) | x3) >> 11;
    const double u1 = (double(b1) + 0.5) / two53;
    const double u2 = double(b2) / two53;
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincos(2.0 * M_PI * u2, &s, &c);
    const int32_t i = 2 * j;
    const int64_t o = r * int64_t(n_vars) + i;
    const double za = rad * c, zb = rad * s;
    if (z_out) z_out[o] = float(za);
    if (p_out) p_out[o] = fmin(1.0, exp(mu[i] + sigma[i] * za));
    if (i + 1 < n_vars) {
        if (z_out) z_out[o + 1] = float(zb);
        if (p_out) p_out[o + 1] = fmin(1.0, exp(mu[i + 1] + sigma[i + 1] * zb));
    }
}

}  // namespace

extern "C" mcs_status mcs_lognormal_samples(const double *mu, const double *sigma, int32_t n_vars,
                                            uint64_t seed, int64_t first_row, int64_t n_rows,
                                            float *z_out, double *p_out) {
    if (n_vars < 0 || n_rows < 0 || first_row < 0)
        return fail(MCS_ERR_INVALID_ARGUMENT, "lognormal samples: negative size or row");
    if (n_vars == 0 || n_rows == 0 || (!z_out && !p_out)) return MCS_OK;
    if (!mu || !sigma) return fail(MCS_ERR_INVALID_ARGUMENT, "lognormal samples: null mu or sigma");
    for (int32_t i = 0; i < n_vars; ++i)
        if (!std::isfinite(mu[i]) || !std::isfinite(sigma[i]) || sigma[i] < 0.0)
            return fail(MCS_ERR_INVALID_ARGUMENT,
                        "lognormal samples: variable " + std::to_string(i) +
                            " needs a finite mu and a finite sigma >= 0");
    const int64_t pairs = (int64_t(n_vars) + 1) / 2;
    // Rows in chunks whose outputs stay within 1 GiB of device memory.
    const int64_t row_bytes = int64_t(n_vars) * ((z_out ? 4 : 0) + (p_out ? 8 : 0));
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n_rows, (int64_t(1) << 30) / row_bytes));
    double *d_mu = nullptr, *d_sigma = nullptr, *d_p = nullptr;
    float *d_z = nullptr;
    hipStream_t stream = nullptr;
    auto cleanup = [&]() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *q : {static_cast<void *>(d_mu), static_cast<void *>(d_sigma), static_cast<void *>(d_p),
                        static_cast<void *>(d_z)})
            if (q) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    };
    const size_t nv = size_t(n_vars);
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d_mu, nv * 8);
    if (e == hipSuccess) e = hipMalloc(&d_sigma, nv * 8);
    if (e == hipSuccess && z_out) e = hipMalloc(&d_z, size_t(chunk) * nv * 4);
    if (e == hipSuccess && p_out) e = hipMalloc(&d_p, size_t(chunk) * nv * 8);
    if (e == hipSuccess) e = hipMemcpyAsync(d_mu, mu, nv * 8, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sigma, sigma, nv * 8, hipMemcpyHostToDevice, stream);
    for (int64_t a = 0; a < n_rows && e == hipSuccess; a += chunk) {
        const int64_t n = std::min(chunk, n_rows - a);
        const int64_t threads = n * pairs;
        hipLaunchKernelGGL(lognormal_rows, dim3(uint32_t((threads + 255) / 256)), dim3(256), 0, stream,
                           d_mu, d_sigma, n_vars, uint32_t(seed), uint32_t(seed >> 32),
                           uint64_t(first_row + a), n, d_z, d_p);
        e = hipGetLastError();
        if (e == hipSuccess && z_out)
            e = hipMemcpyAsync(z_out + a * n_vars, d_z, size_t(n) * nv * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess && p_out)
            e = hipMemcpyAsync(p_out + a * n_vars, d_p, size_t(n) * nv * 8, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
    cleanup();
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? MCS_ERR_OUT_OF_MEMORY : MCS_ERR_DEVICE,
                    std::string("lognormal samples: ") + hipGetErrorString(e));
    return MCS_OK;
}
