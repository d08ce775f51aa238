// selftest.hip -- GPU check of the range-specific correctly rounded helpers
// of the Dykstra projection, of the simplex projection's small-constant
// quotients and of the histograms' bin fraction against the compiler's IEEE
// sqrtf and division.
#include "sct_internal.hpp"
#include "potts_math.hpp"

#include <algorithm>

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Operands: x, b uniform over the float bit patterns of [1, 2^64); a uniform
// over the patterns of [2^-23, b); for the histograms' fraction, b uniform
// over the patterns of [2^-32, 2^32] and |a| over those of [b 2^-40, b], of
// either sign.  Counts results that differ in any bit.
__global__ __launch_bounds__(256) void cr_math_kernel(uint64_t n, uint64_t seed, unsigned long long *bad)
{
    constexpr uint32_t lo = 0x3f800000u, hi = 0x5f800000u, a_lo = 0x34000000u;
    unsigned long long bs = 0, bd = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t h1 = splitmix64(seed ^ (2 * i)), h2 = splitmix64(seed ^ (2 * i + 1));
        const float x = __uint_as_float(lo + uint32_t(h1 % (hi - lo)));
        const uint32_t b_bits = lo + uint32_t((h1 >> 32) % (hi - lo));
        const float b = __uint_as_float(b_bits);
        const float a = __uint_as_float(a_lo + uint32_t(h2 % (b_bits - a_lo)));
        bs += __float_as_uint(sct::sqrt_cr_ge1(x)) != __float_as_uint(sqrtf(x));
        bd += __float_as_uint(sct::div_cr_normal(a, b)) != __float_as_uint(a / b);
        // the histograms' fraction (x - start) / (end - start) on its short path
        const uint64_t h3 = splitmix64(~seed ^ (2 * i)), h4 = splitmix64(~seed ^ (2 * i + 1));
        constexpr uint32_t hb_lo = 0x2f800000u, hb_hi = 0x4f800000u;   // 2^-32, 2^32
        const uint32_t hb = hb_lo + uint32_t(h3 % (hb_hi - hb_lo + 1));
        const uint32_t ha_lo = hb - (40u << 23);                       // b 2^-40, a normal float
        const float hbf = __uint_as_float(hb);
        const float haf = __uint_as_float((ha_lo + uint32_t(h4 % (hb - ha_lo + 1))) | uint32_t(h3 >> 63) << 31);
        bd += __float_as_uint(sct::div_cr_normal(haf, hbf)) != __float_as_uint(haf / hbf);
        // the small-constant quotients on any finite operand other than -0
        const float y = __uint_as_float(uint32_t(h2 >> 32));
        if (isfinite(y) && __float_as_uint(y) != 0x80000000u) {
            bd += __float_as_uint(sct::div_small<3>(y)) != __float_as_uint(y / 3.0f);
            bd += __float_as_uint(sct::div_small<5>(y)) != __float_as_uint(y / 5.0f);
            bd += __float_as_uint(sct::div_small<7>(y)) != __float_as_uint(y / 7.0f);
        }
    }
    if (bs) atomicAdd(&bad[0], bs);
    if (bd) atomicAdd(&bad[1], bd);
}

}  // namespace

extern "C" int sct_selftest_cr_math(sct_ctx *ctx, uint64_t n, uint64_t seed, uint64_t *bad_sqrt, uint64_t *bad_div)
{
    if (!ctx || !bad_sqrt || !bad_div) return SCT_ERR_INVALID;
    unsigned long long *d = nullptr;
    int rc = sct_alloc(ctx, (void **)&d, 2 * sizeof(unsigned long long));
    if (rc) return rc;
    unsigned long long h[2] = {0, 0};
    if (hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = SCT_ERR_HIP;
    if (!rc && n) {
        const uint32_t grid = uint32_t(std::min<uint64_t>((n + 255) / 256, uint64_t(ctx->cu_count) * 16));
        cr_math_kernel<<<grid, 256, 0, ctx->stream>>>(n, seed, d);
        if (hipGetLastError() != hipSuccess) rc = SCT_ERR_HIP;
    }
    if (!rc && hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) rc = SCT_ERR_HIP;
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = SCT_ERR_HIP;
    sct_free(ctx, d);
    if (rc) return sct_fail(ctx, rc, "cr_math self-test failed to run");
    *bad_sqrt = h[0];
    *bad_div = h[1];
    return SCT_OK;
}

namespace {

// The solver's rounding of a float to the stored Xi / V-bar element.
template <typename St>
__global__ __launch_bounds__(256) void store_round_kernel(const float *__restrict__ x, uint64_t n,
                                                          uint16_t *__restrict__ out)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        out[i] = __builtin_bit_cast(uint16_t, St::st(x[i]));
}

}  // namespace

extern "C" int sct_selftest_store_round(sct_ctx *ctx, int storage, const float *x, uint64_t n, uint16_t *out)
{
    if (!ctx || (n && (!x || !out))) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (storage != SCT_STORAGE_BF16 && storage != SCT_STORAGE_F16)
        return sct_fail(ctx, SCT_ERR_INVALID, "storage must be SCT_STORAGE_BF16 or SCT_STORAGE_F16");
    if (!n) return SCT_OK;
    float *dx = nullptr;
    uint16_t *dy = nullptr;
    int rc = sct_alloc(ctx, (void **)&dx, n * sizeof(float));
    if (!rc) rc = sct_alloc(ctx, (void **)&dy, n * sizeof(uint16_t));
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(dx, x, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess) {
        const uint32_t grid = uint32_t(std::min<uint64_t>((n + 255) / 256, uint64_t(ctx->cu_count) * 16));
        if (storage == SCT_STORAGE_BF16)
            store_round_kernel<sct::StorageBF16><<<grid, 256, 0, ctx->stream>>>(dx, n, dy);
        else
            store_round_kernel<sct::StorageF16><<<grid, 256, 0, ctx->stream>>>(dx, n, dy);
        e = hipGetLastError();
    }
    if (!rc && e == hipSuccess) e = hipMemcpyAsync(out, dy, n * sizeof(uint16_t), hipMemcpyDeviceToHost, ctx->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (!rc && e != hipSuccess) {
        (void)hipStreamSynchronize(ctx->stream);
        rc = sct_hip_fail(ctx, e, "store_round self-test");
    }
    sct_free(ctx, dx);
    sct_free(ctx, dy);
    return rc;
}
