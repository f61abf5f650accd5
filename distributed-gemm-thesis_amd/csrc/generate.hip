// Seeded, counter-based matrix generation (SplitMix64, Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014).
// Element i takes output i of the generator seeded with `seed`, so the whole
// matrix is produced in parallel and equals the sequential stream.
#include "dg_internal.hpp"

namespace dg {

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

// The output function applied to the generator state z = seed + (i+1) gamma.
__device__ __forceinline__ uint64_t splitmix64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Second, independent stream for the exponent of the wide distribution.
constexpr uint64_t kExpSeedXor = 0xD1B54A32D192ED03ull;

// Element value from the states of the value stream (z) and, for the wide
// distribution, of the exponent stream (ze).
__device__ __forceinline__ double value_of(uint64_t zs, int dist, uint64_t zes) {
    const uint64_t z = splitmix64_mix(zs);
    const double u01 = (double)(z >> 11) * 0x1.0p-53;  // exact: 53-bit integer times 2^-53
    if (dist == DG_DIST_UNIFORM_01) return u01;
    double v = 2.0 * u01 - 1.0;  // exact in [-1, 1)
    if (dist == DG_DIST_WIDE) {
        const uint64_t ze = splitmix64_mix(zes);
        const int e = (int)((ze >> 32) % 41ull) - 20;
        v = ldexp(v, e);  // exact power-of-two scaling
    }
    return v;
}

__device__ __forceinline__ double generated_value(uint64_t seed, int dist, int64_t i) {
    const uint64_t step = ((uint64_t)i + 1) * kGamma;
    return value_of(seed + step, dist, (seed ^ kExpSeedXor) + step);
}

// One element per thread and pass (two per thread measured slower); the
// output function costs about 10 32-bit multiplies per element, which the
// 8-byte stores still hide.
__global__ void generate_kernel(uint64_t seed, int dist, double* __restrict__ out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = generated_value(seed, dist, i);
}

// FP32 form: the same FP64 values rounded to nearest (ties to even).
__global__ void generate_f32_kernel(uint64_t seed, int dist, float* __restrict__ out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)generated_value(seed, dist, i);
}

}  // namespace

}  // namespace dg

template <typename T>
static int dg_generate_common(uint64_t seed, int dist, T* ptr, int64_t n, void* stream, const char* name) {
    using namespace dg;
    if (n < 0) return fail(DG_ERR_INVALID, "%s: negative size", name);
    if (dist != DG_DIST_UNIFORM_PM1 && dist != DG_DIST_WIDE && dist != DG_DIST_UNIFORM_01)
        return fail(DG_ERR_INVALID, "%s: unknown distribution %d", name, dist);
    if (n == 0) return DG_OK;
    if (!ptr) return fail(DG_ERR_INVALID, "%s: null pointer", name);
    int rc = require_device();
    if (rc != DG_OK) return rc;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if constexpr (sizeof(T) == 8) {
        hipLaunchKernelGGL(generate_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), seed, dist, ptr, n);
    } else {
        hipLaunchKernelGGL(generate_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), seed, dist, ptr, n);
    }
    DG_LAUNCH_CHECK();
    return DG_OK;
}

extern "C" int dg_generate(uint64_t seed, int dist, double* ptr, int64_t n, void* stream) {
    return dg_generate_common(seed, dist, ptr, n, stream, "dg_generate");
}

extern "C" int dg_generate_f32(uint64_t seed, int dist, float* ptr, int64_t n, void* stream) {
    return dg_generate_common(seed, dist, ptr, n, stream, "dg_generate_f32");
}
