// Bit-exact layout kernels: 2D-block and block-cyclic pack/unpack, strided
// block copy and the 2.5D layer reduction.  All are HBM-bound copies; a
// workgroup moves whole column runs of four columns at once, so that both
// the strided global side and the contiguous packed side are accessed in
// 1 KB wavefront runs of 16-byte vectors where alignment allows.
#include "dg_internal.hpp"

namespace dg {

namespace {

constexpr int kThreads = 256;
constexpr int kColsPerBlock = 4;
constexpr int kMaxParts = 256;

typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
typedef double dv2 __attribute__((ext_vector_type(2)));

// Copies runs of `len` elements, both sides contiguous, for nc <= 4
// columns at once: every load of a thread's pass over the columns is issued
// before the first store.  When every run starts 16-byte aligned on both
// sides, the runs move in 16-byte vectors (E elements each) and threads
// 0..len%E-1 copy the tail elements of each run; otherwise element by element.
template <typename T>
__device__ __forceinline__ void copy_cols(const T* const (&src)[kColsPerBlock], T* const (&dst)[kColsPerBlock],
                                          int nc, int64_t len) {
    typedef __attribute__((address_space(1))) const T gcT;
    typedef __attribute__((address_space(1))) T gT;
    typedef T vT __attribute__((ext_vector_type(16 / sizeof(T))));
    typedef __attribute__((address_space(1))) const vT gcV;
    typedef __attribute__((address_space(1))) vT gV;
    constexpr int E = 16 / sizeof(T);
    constexpr int U = 2;  // vectors per column and pass
    bool vec = true;
#pragma unroll
    for (int i = 0; i < kColsPerBlock; ++i)
        if (i < nc) vec = vec && !((reinterpret_cast<uintptr_t>(src[i]) | reinterpret_cast<uintptr_t>(dst[i])) & 15);
    const int64_t tid = threadIdx.x;
    if (vec) {
        const int64_t nv = len / E;
        for (int64_t x0 = 0; x0 < nv; x0 += U * kThreads) {
            vT v[kColsPerBlock][U];
#pragma unroll
            for (int i = 0; i < kColsPerBlock; ++i)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t x = x0 + u * kThreads + tid;
                    if (i < nc && x < nv) v[i][u] = ((gcV*)src[i])[x];
                }
#pragma unroll
            for (int i = 0; i < kColsPerBlock; ++i)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t x = x0 + u * kThreads + tid;
                    if (i < nc && x < nv) ((gV*)dst[i])[x] = v[i][u];
                }
        }
        const int64_t t0 = nv * E;
        if (tid < len - t0) {
#pragma unroll
            for (int i = 0; i < kColsPerBlock; ++i)
                if (i < nc) ((gT*)dst[i])[t0 + tid] = ((gcT*)src[i])[t0 + tid];
        }
    } else {
        for (int64_t x0 = 0; x0 < len; x0 += U * kThreads) {
            T v[kColsPerBlock][U];
#pragma unroll
            for (int i = 0; i < kColsPerBlock; ++i)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t x = x0 + u * kThreads + tid;
                    if (i < nc && x < len) v[i][u] = ((gcT*)src[i])[x];
                }
#pragma unroll
            for (int i = 0; i < kColsPerBlock; ++i)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t x = x0 + u * kThreads + tid;
                    if (i < nc && x < len) ((gT*)dst[i])[x] = v[i][u];
                }
        }
    }
}

struct Bands {
    int64_t off[kMaxParts + 1];
    int n;
};

// Row band `band` (rows [off[b], off[b+1])) of every column: the packed
// index of global (r, c) in band b is off[b]*cols + (r - off[b]) + c*rows_b
// (tiles of one band are contiguous in rank order and share ld = rows_b).
template <bool PACK>
__global__ void block2d_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t ldg,
                               int64_t cols, Bands bands) {
    const int b = blockIdx.x;
    const int64_t r0 = bands.off[b];
    const int64_t rb = bands.off[b + 1] - r0;
    if (rb <= 0) return;
    const int64_t band_base = r0 * cols;
    for (int64_t c = (int64_t)blockIdx.y * kColsPerBlock; c < cols; c += (int64_t)gridDim.y * kColsPerBlock) {
        const int nc = (int)(c + kColsPerBlock < cols ? kColsPerBlock : cols - c);
        const double* s[kColsPerBlock];
        double* d[kColsPerBlock];
#pragma unroll
        for (int i = 0; i < kColsPerBlock; ++i) {
            const int64_t cc = c + (i < nc ? i : 0);
            const int64_t g = r0 + cc * ldg;
            const int64_t p = band_base + cc * rb;
            s[i] = PACK ? src + g : src + p;
            d[i] = PACK ? dst + p : dst + g;
        }
        copy_cols(s, d, nc, rb);
    }
}

struct CyclicGeom {
    int64_t rows, cols, mb, nb;
    int P, Q;
    int64_t lrows[kMaxParts];    // numroc(rows, mb, pr, 0, P)
    int64_t rowpre[kMaxParts];   // prefix of lrows
};

// Row block `ib` (rows [ib*mb, ...)) of column c belongs to process row
// ib % P at local row (ib / P)*mb; column c to process column (c/nb) % Q at
// local column ((c/nb)/Q)*nb + c % nb.  Local matrix of rank pr*Q + pc
// starts at rowpre[pr]*cols + lrows[pr]*colpre[pc] in the packed buffer.
template <bool PACK>
__global__ void cyclic_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t ldg,
                              const CyclicGeom* __restrict__ geo_p, const int64_t* __restrict__ colpre) {
    const CyclicGeom& geo = *geo_p;
    const int64_t ib = blockIdx.x;
    const int64_t row0 = ib * geo.mb;
    const int64_t len = (geo.rows - row0) < geo.mb ? (geo.rows - row0) : geo.mb;
    const int pr = (int)(ib % geo.P);
    const int64_t lrow0 = (ib / geo.P) * geo.mb;
    const int64_t lrows = geo.lrows[pr];
    for (int64_t c = (int64_t)blockIdx.y * kColsPerBlock; c < geo.cols; c += (int64_t)gridDim.y * kColsPerBlock) {
        const int nc = (int)(c + kColsPerBlock < geo.cols ? kColsPerBlock : geo.cols - c);
        const double* s[kColsPerBlock];
        double* d[kColsPerBlock];
#pragma unroll
        for (int i = 0; i < kColsPerBlock; ++i) {
            const int64_t cc = c + (i < nc ? i : 0);
            const int64_t jb = cc / geo.nb;
            const int pc = (int)(jb % geo.Q);
            const int64_t lcol = (jb / geo.Q) * geo.nb + (cc - jb * geo.nb);
            const int64_t base = geo.rowpre[pr] * geo.cols + lrows * colpre[pc];
            const int64_t p = base + lrow0 + lcol * lrows;
            const int64_t g = row0 + cc * ldg;
            s[i] = PACK ? src + g : src + p;
            d[i] = PACK ? dst + p : dst + g;
        }
        copy_cols(s, d, nc, len);
    }
}

template <typename T>
__global__ void copy_block_kernel(int64_t rows, int64_t cols, const T* __restrict__ src, int64_t lds,
                                  T* __restrict__ dst, int64_t ldd) {
    for (int64_t c = (int64_t)blockIdx.x * kColsPerBlock; c < cols; c += (int64_t)gridDim.x * kColsPerBlock) {
        const int nc = (int)(c + kColsPerBlock < cols ? kColsPerBlock : cols - c);
        const T* s[kColsPerBlock];
        T* d[kColsPerBlock];
#pragma unroll
        for (int i = 0; i < kColsPerBlock; ++i) {
            const int64_t cc = c + (i < nc ? i : 0);
            s[i] = src + cc * lds;
            d[i] = dst + cc * ldd;
        }
        copy_cols(s, d, nc, rows);
    }
}

// C[x] += partial[0*n + x] + partial[1*n + x] + ... in layer order.  The
// 128-bit path needs n even so that every layer starts 16-byte aligned.
template <bool VEC>
__global__ void layer_reduce_kernel(double* __restrict__ C, const double* __restrict__ partial, int64_t n,
                                    int nlayers) {
    gdouble* c = (gdouble*)C;
    gcdouble* p = (gcdouble*)partial;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (VEC) {
        typedef __attribute__((address_space(1))) dv2 gdv2;
        typedef __attribute__((address_space(1))) const dv2 gcdv2;
        // two 16-byte vectors per thread and pass, `stride` apart: every
        // load of a pass is issued before the adds
        const int64_t nv = n / 2;
        for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < nv; x += 2 * stride) {
            const bool two = x + stride < nv;
            dv2 acc0 = *(gdv2*)(c + 2 * x);
            dv2 acc1 = two ? *(gdv2*)(c + 2 * (x + stride)) : dv2{0.0, 0.0};
            for (int l = 0; l < nlayers; ++l) {
                const dv2 v0 = *(gcdv2*)(p + (int64_t)l * n + 2 * x);
                const dv2 v1 = two ? *(gcdv2*)(p + (int64_t)l * n + 2 * (x + stride)) : dv2{0.0, 0.0};
                acc0.x += v0.x;
                acc0.y += v0.y;
                acc1.x += v1.x;
                acc1.y += v1.y;
            }
            *(gdv2*)(c + 2 * x) = acc0;
            if (two) *(gdv2*)(c + 2 * (x + stride)) = acc1;
        }
    } else {
        for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += stride) {
            double acc = c[x];
            for (int l = 0; l < nlayers; ++l) acc += p[(int64_t)l * n + x];
            c[x] = acc;
        }
    }
}

unsigned col_blocks(int64_t cols, int64_t runs) {
    // Enough blocks to fill the GPU several times over, within grid.y limits.
    int64_t want = (cols + kColsPerBlock - 1) / kColsPerBlock;
    int64_t cap = 65535;
    if (runs > 0 && runs * want > (1 << 20)) want = ((1 << 20) + runs - 1) / runs;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (unsigned)want;
}

int check_offsets(const char* who, const int64_t* off, int parts, int64_t total) {
    if (!off) return fail(DG_ERR_INVALID, "%s: null offsets", who);
    if (parts < 1 || parts > kMaxParts) return fail(DG_ERR_INVALID, "%s: %d parts (1..%d)", who, parts, kMaxParts);
    if (off[0] != 0 || off[parts] != total)
        return fail(DG_ERR_INVALID, "%s: offsets must run from 0 to %lld", who, (long long)total);
    for (int i = 0; i < parts; ++i)
        if (off[i + 1] < off[i]) return fail(DG_ERR_INVALID, "%s: offsets decrease at %d", who, i);
    return DG_OK;
}

template <bool PACK>
int block2d(const char* who, const double* src, double* dst, int64_t ldg, int64_t rows, int64_t cols,
            const int64_t* row_off, int nrb, const int64_t* col_off, int ncb, void* stream) {
    if (rows < 0 || cols < 0) return fail(DG_ERR_INVALID, "%s: negative size", who);
    if (ldg < (rows > 1 ? rows : 1)) return fail(DG_ERR_INVALID, "%s: ldg %lld < rows %lld", who, (long long)ldg, (long long)rows);
    int rc = check_offsets(who, row_off, nrb, rows);
    if (rc) return rc;
    rc = check_offsets(who, col_off, ncb, cols);
    if (rc) return rc;
    if (rows == 0 || cols == 0) return DG_OK;
    if (!src || !dst) return fail(DG_ERR_INVALID, "%s: null pointer", who);
    rc = require_device();
    if (rc) return rc;
    Bands b;
    b.n = nrb;
    for (int i = 0; i <= nrb; ++i) b.off[i] = row_off[i];
    dim3 grid((unsigned)nrb, col_blocks(cols, nrb));
    hipLaunchKernelGGL(block2d_kernel<PACK>, grid, dim3(kThreads), 0, as_stream(stream), src, dst, ldg, cols, b);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

template <bool PACK>
int cyclic(const char* who, const double* src, double* dst, int64_t ldg, int64_t rows, int64_t cols, int64_t mb,
           int64_t nb, int P, int Q, void* stream) {
    if (rows < 0 || cols < 0) return fail(DG_ERR_INVALID, "%s: negative size", who);
    if (mb <= 0 || nb <= 0) return fail(DG_ERR_INVALID, "%s: block size must be positive", who);
    if (P < 1 || Q < 1 || P > kMaxParts || Q > kMaxParts)
        return fail(DG_ERR_INVALID, "%s: process grid %dx%d (1..%d per side)", who, P, Q, kMaxParts);
    if (ldg < (rows > 1 ? rows : 1)) return fail(DG_ERR_INVALID, "%s: ldg %lld < rows %lld", who, (long long)ldg, (long long)rows);
    if (rows == 0 || cols == 0) return DG_OK;
    if (!src || !dst) return fail(DG_ERR_INVALID, "%s: null pointer", who);
    int rc = require_device();
    if (rc) return rc;
    CyclicGeom g{};
    g.rows = rows; g.cols = cols; g.mb = mb; g.nb = nb; g.P = P; g.Q = Q;
    int64_t acc = 0;
    for (int pr = 0; pr < P; ++pr) {
        g.lrows[pr] = dg_numroc(rows, mb, pr, 0, P);
        g.rowpre[pr] = acc;
        acc += g.lrows[pr];
    }
    std::vector<int64_t> colpre(Q + 1);
    acc = 0;
    for (int pc = 0; pc < Q; ++pc) { colpre[pc] = acc; acc += dg_numroc(cols, nb, pc, 0, Q); }
    colpre[Q] = acc;
    TableUseScope uses{as_stream(stream)};  // after the launch below returns
    const CyclicGeom* dg_ = static_cast<const CyclicGeom*>(cached_table(&g, sizeof(g), as_stream(stream)));
    if (!dg_) return DG_ERR_ALLOC;
    const int64_t* dcp = static_cast<const int64_t*>(cached_table(colpre.data(), colpre.size() * sizeof(int64_t), as_stream(stream)));
    if (!dcp) return DG_ERR_ALLOC;
    const int64_t row_blocks = (rows + mb - 1) / mb;
    if (row_blocks > INT32_MAX) return fail(DG_ERR_INVALID, "%s: too many row blocks", who);
    dim3 grid((unsigned)row_blocks, col_blocks(cols, row_blocks));
    hipLaunchKernelGGL(cyclic_kernel<PACK>, grid, dim3(kThreads), 0, as_stream(stream), src, dst, ldg, dg_, dcp);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

}  // namespace

}  // namespace dg

extern "C" {

int dg_block_splits(int64_t n, int parts, int64_t* offsets) {
    using namespace dg;
    if (n < 0 || parts < 1 || !offsets) return fail(DG_ERR_INVALID, "dg_block_splits: n=%lld parts=%d", (long long)n, parts);
    const int64_t base = n / parts;
    for (int i = 0; i < parts; ++i) offsets[i] = (int64_t)i * base;
    offsets[parts] = n;
    return DG_OK;
}

int dg_virtual_grid(int size, int* drow, int* dcol) {
    using namespace dg;
    if (size < 1 || !drow || !dcol) return fail(DG_ERR_INVALID, "dg_virtual_grid: size %d", size);
    int px = 1;
    while ((int64_t)(px + 1) * (px + 1) <= size) ++px;  // floor(sqrt(size)) in integers
    int d = px;
    while (d > 1 && size % d != 0) --d;
    *dcol = d;
    *drow = size / d;
    return DG_OK;
}

int64_t dg_numroc(int64_t n, int64_t nb, int iproc, int isrcproc, int nprocs) {
    if (n <= 0 || nb <= 0 || nprocs <= 0) return 0;
    const int64_t dist = ((int64_t)nprocs + iproc - isrcproc) % nprocs;
    const int64_t nblocks = n / nb;
    int64_t r = (nblocks / nprocs) * nb;
    const int64_t extra = nblocks % nprocs;
    if (dist < extra) r += nb;
    else if (dist == extra) r += n % nb;
    return r;
}

int64_t dg_block_cyclic_size(int64_t rows, int64_t cols, int64_t mb, int64_t nb, int P, int Q) {
    int64_t total = 0;
    for (int pr = 0; pr < P; ++pr)
        for (int pc = 0; pc < Q; ++pc) total += dg_numroc(rows, mb, pr, 0, P) * dg_numroc(cols, nb, pc, 0, Q);
    return total;
}

int dg_pack_block2d(const double* G, int64_t ldg, int64_t rows, int64_t cols, const int64_t* row_off, int nrb,
                    const int64_t* col_off, int ncb, double* packed, void* stream) {
    return dg::block2d<true>("dg_pack_block2d", G, packed, ldg, rows, cols, row_off, nrb, col_off, ncb, stream);
}

int dg_unpack_block2d(const double* packed, const int64_t* row_off, int nrb, const int64_t* col_off, int ncb,
                      double* G, int64_t ldg, int64_t rows, int64_t cols, void* stream) {
    return dg::block2d<false>("dg_unpack_block2d", packed, G, ldg, rows, cols, row_off, nrb, col_off, ncb, stream);
}

int dg_pack_block_cyclic(const double* G, int64_t ldg, int64_t rows, int64_t cols, int64_t mb, int64_t nb, int P,
                         int Q, double* packed, void* stream) {
    return dg::cyclic<true>("dg_pack_block_cyclic", G, packed, ldg, rows, cols, mb, nb, P, Q, stream);
}

int dg_unpack_block_cyclic(const double* packed, int64_t mb, int64_t nb, int P, int Q, double* G, int64_t ldg,
                           int64_t rows, int64_t cols, void* stream) {
    return dg::cyclic<false>("dg_unpack_block_cyclic", packed, G, ldg, rows, cols, mb, nb, P, Q, stream);
}

}  // extern "C"

namespace dg {
namespace {
template <typename T>
int copy_block(const char* name, int64_t rows, int64_t cols, const T* src, int64_t lds, T* dst, int64_t ldd,
               void* stream) {
    if (rows < 0 || cols < 0) return fail(DG_ERR_INVALID, "%s: negative size", name);
    if (lds < rows || ldd < rows) return fail(DG_ERR_INVALID, "%s: leading dimension < rows", name);
    if (rows == 0 || cols == 0) return DG_OK;
    if (!src || !dst) return fail(DG_ERR_INVALID, "%s: null pointer", name);
    int rc = require_device();
    if (rc) return rc;
    hipLaunchKernelGGL(copy_block_kernel<T>, dim3(col_blocks(cols, 1)), dim3(kThreads), 0, as_stream(stream), rows,
                       cols, src, lds, dst, ldd);
    DG_LAUNCH_CHECK();
    return DG_OK;
}
}  // namespace
}  // namespace dg

extern "C" {

int dg_copy_block(int64_t rows, int64_t cols, const double* src, int64_t lds, double* dst, int64_t ldd,
                  void* stream) {
    return dg::copy_block("dg_copy_block", rows, cols, src, lds, dst, ldd, stream);
}

int dg_copy_block_f32(int64_t rows, int64_t cols, const float* src, int64_t lds, float* dst, int64_t ldd,
                      void* stream) {
    return dg::copy_block("dg_copy_block_f32", rows, cols, src, lds, dst, ldd, stream);
}

int dg_layer_reduce(double* C, const double* partial, int64_t n, int nlayers, void* stream) {
    using namespace dg;
    if (n < 0 || nlayers < 0) return fail(DG_ERR_INVALID, "dg_layer_reduce: negative size");
    if (n == 0 || nlayers == 0) return DG_OK;
    if (!C || !partial) return fail(DG_ERR_INVALID, "dg_layer_reduce: null pointer");
    if ((reinterpret_cast<uintptr_t>(C) | reinterpret_cast<uintptr_t>(partial)) & 15)
        return fail(DG_ERR_INVALID, "dg_layer_reduce: buffers must be 16-byte aligned");
    int rc = require_device();
    if (rc) return rc;
    const bool vec = (n % 2) == 0;
    int64_t blocks = ((vec ? n / 2 : n) + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (vec)
        hipLaunchKernelGGL(layer_reduce_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), C,
                           partial, n, nlayers);
    else
        hipLaunchKernelGGL(layer_reduce_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), C,
                           partial, n, nlayers);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

}  // extern "C"
