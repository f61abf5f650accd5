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

// E = 16/sizeof(T) consecutive elements per thread, written as one 16-byte
// vector, the grid covering the whole matrix (one store per thread keeps
// the resident workgroups on a compact window of the output).  The first
// `peel` elements (up to a 16-byte boundary) and the tail elements go
// singly.  Without 16-byte vectors (VEC false) one element per thread.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void generate_kernel(uint64_t seed, int dist, T* __restrict__ out, int64_t n,
                                                       int peel) {
    typedef __attribute__((address_space(1))) T gT;
    typedef T vT __attribute__((ext_vector_type(16 / sizeof(T))));
    typedef __attribute__((address_space(1))) vT gV;
    constexpr int E = VEC ? 16 / sizeof(T) : 1;
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nv = (n - peel) / E;
    for (int64_t v = gid; v < nv; v += stride) {
        const int64_t i0 = peel + v * E;
        if constexpr (VEC) {
            vT w;
#pragma unroll
            for (int e = 0; e < E; ++e) w[e] = (T)generated_value(seed, dist, i0 + e);
            *(gV*)(out + i0) = w;
        } else {
            ((gT*)out)[i0] = (T)generated_value(seed, dist, i0);
        }
    }
    if (gid < peel) ((gT*)out)[gid] = (T)generated_value(seed, dist, gid);
    const int64_t t0 = peel + nv * E;
    if (gid < n - t0) ((gT*)out)[t0 + gid] = (T)generated_value(seed, dist, t0 + gid);
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
    // The FP32 form rounds the FP64 value to nearest (ties to even).
    const uintptr_t addr = reinterpret_cast<uintptr_t>(ptr);
    const bool vec = (addr % sizeof(T)) == 0;
    int peel = vec ? (int)(((16 - (addr & 15)) & 15) / sizeof(T)) : 0;
    if (peel > n) peel = (int)n;
    const int64_t per = vec ? 16 / sizeof(T) : 1;
    int64_t blocks = ((n - peel) / per + 255) / 256;
    if (blocks > (1 << 30)) blocks = 1 << 30;  // beyond that the kernel loops
    if (blocks < 1) blocks = 1;
    if (vec)
        hipLaunchKernelGGL((generate_kernel<T, true>), dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), seed,
                           dist, ptr, n, peel);
    else
        hipLaunchKernelGGL((generate_kernel<T, false>), dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), seed,
                           dist, ptr, n, 0);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

extern "C" int dg_generate(uint64_t seed, int dist, double* ptr, int64_t n, void* stream) {
    return dg_generate_common(seed, dist, ptr, n, stream, "dg_generate");
}

extern "C" int dg_generate_f32(uint64_t seed, int dist, float* ptr, int64_t n, void* stream) {
    return dg_generate_common(seed, dist, ptr, n, stream, "dg_generate_f32");
}
