// Seeded, counter-based matrix generation (SplitMix64, Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014).
// Element i takes output i of the generator seeded with `seed`, so the whole
// matrix is produced in parallel and equals the sequential stream.
#include "dg_internal.hpp"

namespace dg {

namespace {

__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Second, independent stream for the exponent of the wide distribution.
constexpr uint64_t kExpSeedXor = 0xD1B54A32D192ED03ull;

__device__ __forceinline__ double generated_value(uint64_t seed, int dist, int64_t i) {
    const uint64_t z = splitmix64_at(seed, (uint64_t)i);
    const double u01 = (double)(z >> 11) * 0x1.0p-53;  // exact: 53-bit integer times 2^-53
    if (dist == DG_DIST_UNIFORM_01) return u01;
    double v = 2.0 * u01 - 1.0;  // exact in [-1, 1)
    if (dist == DG_DIST_WIDE) {
        const uint64_t ze = splitmix64_at(seed ^ kExpSeedXor, (uint64_t)i);
        const int e = (int)((ze >> 32) % 41ull) - 20;
        v = ldexp(v, e);  // exact power-of-two scaling
    }
    return v;
}

typedef double gdv2 __attribute__((ext_vector_type(2)));

// Two consecutive elements per thread and pass, one 16-byte store (the
// buffer is 16-byte aligned: checked by the host); an odd last element is
// written by thread 0.
__global__ void generate_kernel(uint64_t seed, int dist, double* __restrict__ out, int64_t n) {
    const int64_t np = n / 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x)
        *reinterpret_cast<gdv2*>(out + 2 * i) = gdv2{generated_value(seed, dist, 2 * i), generated_value(seed, dist, 2 * i + 1)};
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) out[n - 1] = generated_value(seed, dist, n - 1);
}

// Unaligned buffers: one element per thread and pass.
__global__ void generate_unaligned_kernel(uint64_t seed, int dist, double* __restrict__ out, int64_t n) {
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
        if (reinterpret_cast<uintptr_t>(ptr) & 15)
            hipLaunchKernelGGL(generate_unaligned_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), seed,
                               dist, ptr, n);
        else
            hipLaunchKernelGGL(generate_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), seed, dist,
                               ptr, n);
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
