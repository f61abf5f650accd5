// Bit-exact layout kernels: 2D-block and block-cyclic pack/unpack, strided
// block copy and the 2.5D layer reduction.  All are HBM-bound copies.  A
// workgroup (one wave) moves one chunk of one column run (or, for short
// runs, the same chunk of several columns), one 16-byte vector per thread,
// and the grid walks row chunks fastest: every thread has exactly one load
// in flight and the workgroups resident at any moment cover a compact
// window of both buffers, which is what keeps the HBM queues full.
// Measured on MI355X for a 16384^2 column copy: 6.3 TB/s in this shape with
// 256-thread workgroups, against 5.5-5.8 TB/s for four columns or two
// vectors per thread and 4.4-4.9 TB/s for grid-stride loops; nontemporal
// accesses and one-wave workgroups then took the kernels from 73-78 % to
// 78-83 % of the 8 TB/s peak (a plain 2 GiB copy: 79 -> 83.5 %).
#include "dg_internal.hpp"

#include <algorithm>

namespace dg {

namespace {

constexpr int kThreads = 64;  // one wave per workgroup
constexpr int kMaxParts = 256;

typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
typedef double dv2 __attribute__((ext_vector_type(2)));

// Every byte these kernels move is read once and written once, so loads and
// stores carry the nontemporal hint.
template <typename G>
__device__ __forceinline__ void move_nt(const G* src, G* dst) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(src), dst);
}

// Copies one chunk of `len` <= lanes*E elements (E = 16/sizeof(T)) with the
// calling thread's `lane` of `lanes`.  When source and destination share
// their offset within 16 bytes, the first elements up to a 16-byte boundary
// are copied singly, then one 16-byte vector per lane, then the tail
// elements; otherwise element by element.
template <typename T>
__device__ __forceinline__ void copy_chunk(const T* src, T* dst, int64_t len, int lane, int lanes) {
    typedef __attribute__((address_space(1))) const T gcT;
    typedef __attribute__((address_space(1))) T gT;
    typedef T vT __attribute__((ext_vector_type(16 / sizeof(T))));
    typedef __attribute__((address_space(1))) const vT gcV;
    typedef __attribute__((address_space(1))) vT gV;
    constexpr int E = 16 / sizeof(T);
    const uintptr_t as = reinterpret_cast<uintptr_t>(src), ad = reinterpret_cast<uintptr_t>(dst);
    if (!((as ^ ad) & 15)) {
        int64_t p = (int64_t)(((16 - (as & 15)) & 15) / sizeof(T));
        if (p > len) p = len;
        const int64_t nv = (len - p) / E;
        const int64_t t0 = p + nv * E;
        if (lane < nv) move_nt((gcV*)(src + p) + lane, (gV*)(dst + p) + lane);
        if (lane < p) move_nt((gcT*)src + lane, (gT*)dst + lane);
        if (lane < len - t0) move_nt((gcT*)src + t0 + lane, (gT*)dst + t0 + lane);
    } else {
        for (int64_t x = lane; x < len; x += lanes) move_nt((gcT*)src + x, (gT*)dst + x);
    }
}

// Workgroup shape for runs of `run` elements: `lanes` threads (a power of
// two, 16..kThreads) per column chunk of lanes*E elements, kThreads/lanes
// columns per workgroup.
struct RunShape {
    int lanes_log2;
    int64_t chunk;    // elements per column chunk
    int64_t chunks;   // chunks per run
    int cols_per_wg;
};

RunShape run_shape(int64_t run, int E) {
    RunShape r{};
    const int64_t vecs = (run + E - 1) / E;
    int l2 = 4;
    while ((1 << l2) < vecs && (1 << l2) < kThreads) ++l2;
    r.lanes_log2 = l2;
    r.chunk = (int64_t)(1 << l2) * E;
    r.chunks = run > 0 ? (run + r.chunk - 1) / r.chunk : 1;
    r.cols_per_wg = kThreads >> l2;
    return r;
}

unsigned col_groups(int64_t cols, int cols_per_wg) {
    int64_t g = (cols + cols_per_wg - 1) / cols_per_wg;
    if (g > 65535) g = 65535;  // the kernels loop over the remaining column groups
    return (unsigned)(g < 1 ? 1 : g);
}

struct Bands {
    int64_t off[kMaxParts + 1];
    int32_t cpre[kMaxParts + 1];  // first chunk of each band (prefix of its chunk counts)
    int n;
};

// Row band b (rows [off[b], off[b+1])) of every column: the packed index of
// global (r, c) in band b is off[b]*cols + (r - off[b]) + c*rows_b (tiles of
// one band are contiguous in rank order and share ld = rows_b).  The grid
// walks the source in memory order: packing reads global columns top to
// bottom (blockIdx.x = chunk of the column across the bands, blockIdx.y =
// column group); unpacking reads each band's packed region front to back
// (blockIdx.x = chunk of the band's column run, blockIdx.y = column group,
// blockIdx.z = band).
template <bool PACK>
__global__ __launch_bounds__(kThreads) void block2d_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                           int64_t ldg, int64_t cols, Bands bands, int lanes_log2) {
    const int x = blockIdx.x;
    int b;
    int64_t q0;  // first row of the chunk within its band
    const int64_t chunk = (int64_t)2 << lanes_log2;
    if (PACK) {
        int lo = 0, hi = bands.n - 1;  // band b with cpre[b] <= x < cpre[b+1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bands.cpre[mid] <= x) lo = mid; else hi = mid - 1;
        }
        b = lo;
        q0 = (int64_t)(x - bands.cpre[b]) * chunk;
    } else {
        b = blockIdx.z;
        q0 = (int64_t)x * chunk;
    }
    const int64_t r0 = bands.off[b];
    const int64_t rb = bands.off[b + 1] - r0;
    const int64_t len = rb - q0 < chunk ? rb - q0 : chunk;
    if (len <= 0) return;
    const int lane = threadIdx.x & ((1 << lanes_log2) - 1);
    const int cpw = kThreads >> lanes_log2;
    const int64_t band_base = r0 * cols;
    for (int64_t c = (int64_t)blockIdx.y * cpw + (threadIdx.x >> lanes_log2); c < cols; c += (int64_t)gridDim.y * cpw) {
        const int64_t g = r0 + q0 + c * ldg;
        const int64_t p = band_base + q0 + c * rb;
        copy_chunk(PACK ? src + g : src + p, PACK ? dst + p : dst + g, len, lane, 1 << lanes_log2);
    }
}

// numroc(n, b, p, 0, procs) = base*b + (b if p < extra, rem if p == extra,
// else 0) with full = n / b, rem = n % b, base = full / procs, extra = full
// % procs, and its prefix over the processes before p.
struct Numroc {
    uint32_t b, rem, base, extra;
};

Numroc make_numroc(uint32_t n, uint32_t b, uint32_t procs) {
    const uint32_t full = n / b;
    return Numroc{b, n - full * b, full / procs, full % procs};
}

__device__ __forceinline__ uint32_t local_extent(const Numroc& r, uint32_t p) {
    return r.base * r.b + (p < r.extra ? r.b : (p == r.extra ? r.rem : 0u));
}
__device__ __forceinline__ uint32_t local_prefix(const Numroc& r, uint32_t p) {
    return p * r.base * r.b + (p < r.extra ? p : r.extra) * r.b + (p > r.extra ? r.rem : 0u);
}

// Block-cyclic geometry (rows, cols, mb, nb < 2^31; source process 0), all
// in kernel arguments, so that a thread's first load waits on no table read.
struct Cyclic {
    int64_t ldg;
    uint32_t rows, cols, mb, nb, P, Q;
    Numroc nr, nc;            // row and column distributions
    FastDiv dP, dQ, dnb, dchunks;
};

// Row block ib (rows [ib*mb, ...)) of column c belongs to process row
// ib % P at local row (ib / P)*mb; column c to process column (c/nb) % Q at
// local column ((c/nb)/Q)*nb + c % nb.  The local matrix of rank (pr, pc)
// (column-major, ld = its local rows) starts at rowpre(pr)*cols +
// lrows(pr)*colpre(pc) in the packed buffer.  Packing walks the global
// columns top to bottom (blockIdx.x = row block * chunks + chunk,
// blockIdx.y = column group); unpacking walks every rank's local matrix in
// memory order (blockIdx.x = local row block * chunks + chunk, blockIdx.y =
// local column group, blockIdx.z = rank pr*Q + pc, looping past 65535).
template <bool PACK>
__global__ __launch_bounds__(kThreads) void cyclic_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                          Cyclic g, int lanes_log2) {
    const uint32_t chunk = 2u << lanes_log2;
    const uint32_t blk = fdiv(blockIdx.x, g.dchunks);
    const uint32_t q0 = (blockIdx.x - blk * g.dchunks.d) * chunk;  // first row of the chunk within its row block
    const int lane = threadIdx.x & ((1 << lanes_log2) - 1);
    const uint32_t cpw = kThreads >> lanes_log2;
    if (PACK) {
        const uint32_t row0 = blk * g.mb + q0;
        if (row0 >= g.rows || q0 >= g.mb) return;
        uint32_t len = g.mb - q0 < chunk ? g.mb - q0 : chunk;
        if (g.rows - row0 < len) len = g.rows - row0;
        const uint32_t bq = fdiv(blk, g.dP), pr = blk - bq * g.P;
        const uint32_t lrows = local_extent(g.nr, pr);
        const int64_t rbase = (int64_t)local_prefix(g.nr, pr) * g.cols + bq * g.mb + q0;
        for (uint32_t c = blockIdx.y * cpw + (threadIdx.x >> lanes_log2); c < g.cols; c += gridDim.y * cpw) {
            const uint32_t jb = fdiv(c, g.dnb);
            const uint32_t jq = fdiv(jb, g.dQ), pc = jb - jq * g.Q;
            const uint32_t lcol = jq * g.nb + (c - jb * g.nb);
            const int64_t p = rbase + (int64_t)lrows * (local_prefix(g.nc, pc) + lcol);
            copy_chunk(src + row0 + c * g.ldg, dst + p, len, lane, 1 << lanes_log2);
        }
    } else {
        for (uint32_t rank = blockIdx.z; rank < g.P * g.Q; rank += gridDim.z) {
            const uint32_t pr = fdiv(rank, g.dQ), pc = rank - pr * g.Q;
            const uint32_t lrows = local_extent(g.nr, pr);
            const uint32_t lr0 = blk * g.mb + q0;  // first local row of the chunk
            if (lr0 >= lrows || q0 >= g.mb) continue;
            uint32_t len = g.mb - q0 < chunk ? g.mb - q0 : chunk;
            if (lrows - lr0 < len) len = lrows - lr0;
            const uint32_t lcols = local_extent(g.nc, pc);
            const int64_t base = (int64_t)local_prefix(g.nr, pr) * g.cols + (int64_t)lrows * local_prefix(g.nc, pc) + lr0;
            const int64_t row0 = (int64_t)(blk * g.P + pr) * g.mb + q0;
            for (uint32_t lc = blockIdx.y * cpw + (threadIdx.x >> lanes_log2); lc < lcols; lc += gridDim.y * cpw) {
                const uint32_t lb = fdiv(lc, g.dnb);
                const int64_t c = (int64_t)(lb * g.Q + pc) * g.nb + (lc - lb * g.nb);
                copy_chunk(src + base + (int64_t)lc * lrows, dst + row0 + c * g.ldg, len, lane, 1 << lanes_log2);
            }
        }
    }
}

// blockIdx.x = chunk of a column, blockIdx.y = column group.
template <typename T>
__global__ __launch_bounds__(kThreads) void copy_block_kernel(int64_t rows, int64_t cols, const T* __restrict__ src,
                                                              int64_t lds, T* __restrict__ dst, int64_t ldd,
                                                              int lanes_log2) {
    constexpr int E = 16 / sizeof(T);
    const int64_t chunk = (int64_t)E << lanes_log2;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t len = rows - r0 < chunk ? rows - r0 : chunk;
    if (len <= 0) return;
    const int lane = threadIdx.x & ((1 << lanes_log2) - 1);
    const int cpw = kThreads >> lanes_log2;
    for (int64_t c = (int64_t)blockIdx.y * cpw + (threadIdx.x >> lanes_log2); c < cols; c += (int64_t)gridDim.y * cpw)
        copy_chunk(src + r0 + c * lds, dst + r0 + c * ldd, len, lane, 1 << lanes_log2);
}

// C[x] += partial[0*n + x] + partial[1*n + x] + ... in layer order.  The
// 128-bit path needs n even so that every layer starts 16-byte aligned; one
// vector (or element) per thread, the grid covering n.
template <bool VEC>
__global__ __launch_bounds__(kThreads) void layer_reduce_kernel(double* __restrict__ C,
                                                                const double* __restrict__ partial, int64_t n,
                                                                int nlayers) {
    gdouble* c = (gdouble*)C;
    gcdouble* p = (gcdouble*)partial;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (VEC) {
        typedef __attribute__((address_space(1))) dv2 gdv2;
        typedef __attribute__((address_space(1))) const dv2 gcdv2;
        const int64_t nv = n / 2;
        for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < nv; x += stride) {
            dv2 acc = __builtin_nontemporal_load((gdv2*)(c + 2 * x));
            for (int l = 0; l < nlayers; ++l) {
                const dv2 v = __builtin_nontemporal_load((gcdv2*)(p + (int64_t)l * n + 2 * x));
                acc.x += v.x;
                acc.y += v.y;
            }
            __builtin_nontemporal_store(acc, (gdv2*)(c + 2 * x));
        }
    } else {
        for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += stride) {
            double acc = __builtin_nontemporal_load(c + x);
            for (int l = 0; l < nlayers; ++l) acc += __builtin_nontemporal_load(p + (int64_t)l * n + x);
            __builtin_nontemporal_store(acc, c + x);
        }
    }
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
    int64_t tallest = 0;
    for (int i = 0; i <= nrb; ++i) b.off[i] = row_off[i];
    for (int i = 0; i < nrb; ++i) tallest = std::max<int64_t>(tallest, row_off[i + 1] - row_off[i]);
    const RunShape rs = run_shape(tallest, 2);
    int64_t chunks = 0;
    for (int i = 0; i < nrb; ++i) {
        b.cpre[i] = (int32_t)chunks;
        chunks += (row_off[i + 1] - row_off[i] + rs.chunk - 1) / rs.chunk;
        if (chunks > INT32_MAX) return fail(DG_ERR_INVALID, "%s: too many row chunks", who);
    }
    b.cpre[nrb] = (int32_t)chunks;
    // packing: chunks of whole global columns; unpacking: chunks of the tallest band, one z per band
    dim3 grid(PACK ? (unsigned)chunks : (unsigned)((tallest + rs.chunk - 1) / rs.chunk),
              col_groups(cols, rs.cols_per_wg), PACK ? 1u : (unsigned)nrb);
    hipLaunchKernelGGL(block2d_kernel<PACK>, grid, dim3(kThreads), 0, as_stream(stream), src, dst, ldg, cols, b,
                       rs.lanes_log2);
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
    if (rows > INT32_MAX || cols > INT32_MAX || mb > INT32_MAX || nb > INT32_MAX)
        return fail(DG_ERR_UNSUPPORTED, "%s: rows, cols and block sizes must be below 2^31", who);
    Cyclic g{};
    g.ldg = ldg;
    g.rows = (uint32_t)rows; g.cols = (uint32_t)cols; g.mb = (uint32_t)mb; g.nb = (uint32_t)nb;
    g.P = (uint32_t)P; g.Q = (uint32_t)Q;
    g.nr = make_numroc(g.rows, g.mb, g.P);
    g.nc = make_numroc(g.cols, g.nb, g.Q);
    const RunShape rs = run_shape(mb < rows ? mb : rows, 2);
    g.dP = make_fastdiv(g.P); g.dQ = make_fastdiv(g.Q); g.dnb = make_fastdiv(g.nb);
    g.dchunks = make_fastdiv((uint32_t)rs.chunks);
    // packing: every global row block; unpacking: process row 0's local row
    // blocks (the most) and process column 0's local columns (the most)
    const int64_t row_blocks = PACK ? (rows + mb - 1) / mb : (dg_numroc(rows, mb, 0, 0, P) + mb - 1) / mb;
    const int64_t ncols = PACK ? cols : dg_numroc(cols, nb, 0, 0, Q);
    if (row_blocks > INT32_MAX / rs.chunks) return fail(DG_ERR_INVALID, "%s: too many row chunks", who);
    const int ranks = P * Q;
    dim3 grid((unsigned)(row_blocks * rs.chunks), col_groups(ncols, rs.cols_per_wg),
              PACK ? 1u : (unsigned)(ranks < 65535 ? ranks : 65535));
    hipLaunchKernelGGL(cyclic_kernel<PACK>, grid, dim3(kThreads), 0, as_stream(stream), src, dst, g, rs.lanes_log2);
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
    const RunShape rs = run_shape(rows, 16 / (int)sizeof(T));
    if (rs.chunks > INT32_MAX) return fail(DG_ERR_INVALID, "%s: too many rows", name);
    hipLaunchKernelGGL(copy_block_kernel<T>, dim3((unsigned)rs.chunks, col_groups(cols, rs.cols_per_wg)),
                       dim3(kThreads), 0, as_stream(stream), rows, cols, src, lds, dst, ldd, rs.lanes_log2);
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
    int64_t blocks = ((vec ? n / 2 : n) + kThreads - 1) / kThreads;
    if (blocks > (1 << 30)) blocks = 1 << 30;  // beyond that the kernel loops
    if (vec)
        hipLaunchKernelGGL(layer_reduce_kernel<true>, dim3((unsigned)blocks), dim3(kThreads), 0, as_stream(stream), C,
                           partial, n, nlayers);
    else
        hipLaunchKernelGGL(layer_reduce_kernel<false>, dim3((unsigned)blocks), dim3(kThreads), 0, as_stream(stream),
                           C, partial, n, nlayers);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

}  // extern "C"
