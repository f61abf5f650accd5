// Internal helpers shared by the HIP sources of libdgamd: error state,
// HIP call checking, device checks and the device table cache.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dgamd.h"

namespace dg {

// Sets the calling thread's last-error message and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Returns DG_OK when the current device is a gfx950 GPU.
int require_device();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Normalises 'n','N','t','T','c','C' to 'N' or 'T'; returns 0 on a bad op.
inline char norm_op(char op) {
    switch (op) {
        case 'n': case 'N': return 'N';
        case 't': case 'T': case 'c': case 'C': return 'T';
        default: return 0;
    }
}

// Device-resident copy of a small host table (problem / segment descriptors).
// Identical tables (same bytes) share one device buffer, so repeated calls
// with the same arguments upload nothing.  Returns nullptr and sets the
// error on failure.  Uploading while `stream` is being captured into a HIP
// graph is refused: tables must be built before capture; a table looked up
// during a capture stays resident for the life of the process (the graph
// holds its address).  Beyond 256 tables per device, a new table reuses the
// buffer of the least recently used one whose last use has completed on the
// device, found by event queries (nothing waits, nothing is freed).
const void* cached_table(const void* host, size_t bytes, hipStream_t stream);

// Marks the kernels that read this thread's tables looked up since the last
// mark as enqueued on `stream`: records the completion event that a later
// reuse of those buffers waits for.  TableUseScope calls it when a call that
// looks up tables returns, after its launches.
void table_uses_enqueued(hipStream_t stream);
struct TableUseScope {
    hipStream_t stream;
    ~TableUseScope() { table_uses_enqueued(stream); }
};

#define DG_HIP(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return ::dg::fail(DG_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__,   \
                              #call, hipGetErrorString(e_));                    \
    } while (0)

#define DG_LAUNCH_CHECK()                                                       \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess)                                                   \
            return ::dg::fail(DG_ERR_HIP, "%s:%d kernel launch: %s", __FILE__,  \
                              __LINE__, hipGetErrorString(e_));                 \
    } while (0)

// ---------------------------------------------------------------- GEMM core
// One K-segment of an output tile's reduction: op(A) rows x klen starting at
// `a` (leading dimension lda of the stored matrix) times op(B) klen x cols at
// `b`.  A problem walks its segments in order; Cannon rotations and SUMMA
// panels are segments whose pointers point into other ranks' tiles.
struct Seg {
    const double* a;
    const double* b;
    int64_t lda;
    int64_t ldb;
    int64_t klen;
};

struct Prob {
    int64_t m, n;
    double* c;
    int64_t ldc;
    double alpha, beta;
    int32_t seg0, nseg;    // segment range in the segment table
    int32_t tile0[2];      // first global tile id of this problem, per tile shape
    int32_t tiles_m[2], tiles_n[2];
    // direct-to-LDS kernel: bit 0 (bit 1) set when the row-contiguous A (B)
    // starts 8 bytes past a 16-byte boundary; its loads then start one row
    // earlier and the tile grid runs over rows + 1
    int32_t shift;
};

// Division of 32-bit unsigned integers by a launch constant d >= 1 with
// one high multiply, a subtract, an add and two shifts (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994, figure 4.1): l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// t = mulhi(m, n), n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0).
struct FastDiv {
    uint32_t d, m, s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{};
    uint32_t l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;
    f.d = d;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 1 ? l - 1 : 0;
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    const uint32_t t = __umulhi(f.m, n);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Tile scheduling of the FP64 GEMM launch.
enum {
    DG_GEMM_MODE_ROUND_ROBIN = 0,  // tile = workgroup id (tiles dealt round-robin over XCDs)
    DG_GEMM_MODE_XCD = 1,          // each XCD gets one contiguous tile range (L2 reuse)
    DG_GEMM_MODE_PERSISTENT = 2,   // CU-count workgroups pull tiles from an atomic counter
};

// Output tile shapes of the FP64 GEMM.
enum {
    DG_GEMM_TILE_128x128 = 0,  // 4 waves, two workgroups per CU
    DG_GEMM_TILE_256x128 = 1,  // 8 waves, one workgroup per CU
    DG_GEMM_TILE_128x128W8 = 2,  // 8 waves of 64x32, two workgroups per CU
    DG_GEMM_TILE_DEFAULT = DG_GEMM_TILE_128x128,
};

// Launches the FP64 MFMA GEMM over `probs` (host vector; uploaded through
// the table cache).  All problems share opA/opB.
int launch_gemm_f64(char opA, char opB, const std::vector<Prob>& probs,
                    const std::vector<Seg>& segs, hipStream_t stream, int mode);

// A batch split by kernel.  Each problem is routed on its own properties,
// so it sums in the same order alone or inside any batch: the direct-to-LDS
// kernel takes it when every segment has K > 0 and even leading dimensions,
// its k-contiguous operand (A with opA = T, B with opB = N) has 16-byte
// aligned bases and even K, and its row-contiguous operand (A with opA = N,
// B with opB = T) has 8-byte aligned bases with the same offset within 16
// bytes in every segment; the register-staged kernel takes the others.
// DG_GEMM_NO_DMA routes everything to the register-staged kernel.
// Direct-to-LDS problems whose tiles are all interior (no row shift, m and n
// multiples of 128, whole k-steps, an even step count) go to `dma`, which
// the A-in-registers kernel runs; the others go to `dmae` (the edge-handling
// kernel).  With split_edges, a direct-to-LDS problem with edges and at
// least 8 interior tiles is cut into sub-problems: an interior block on
// `dma`, and strips on `dmae` around it, placed so that the tile count does
// not grow (a row-shifted problem's interior starts at its row 127).  The
// sub-problems' segments are appended to `segs` (which starts as a copy of
// the input segments); every element still sums in the same order.  The
// strips past the interior that are narrower than a tile go instead to
// `nar[0]` (columns past it, 128x32 tiles) and `nar[1]` (rows past it, 64x128
// tiles), which the register-staged kernel runs in the same k order
// (DG_GEMM_NO_NARROW keeps them on `dmae`).
struct GemmBatch {
    std::vector<Prob> dma, dmae, gen;  // tile ranges assigned per list
    int tiles_dma = 0;                 // 128x128 tiles of `dma`
    int tiles_dmae = 0;                // 128x128 tiles of `dmae`
    int tiles_gen[2] = {0, 0};         // tiles of `gen` per tile shape
    bool vec = true;                   // every `gen` segment allows 128-bit loads
    std::vector<Prob> nar[2];          // narrow strips: 128x32 tiles, 64x128 tiles
    int tiles_nar[2] = {0, 0};
    bool vec_nar[2] = {true, true};
    std::vector<Seg> segs;             // the segment table all lists index
};
int prepare_gemm(char opA, char opB, const std::vector<Prob>& probs_in, const std::vector<Seg>& segs,
                 GemmBatch& out, bool split_edges = true);

// The same batch with its tables resident on the device.
struct GemmBatchDev {
    const Prob* dma = nullptr;
    int ndma = 0, tiles_dma = 0;
    const Prob* dmae = nullptr;
    int ndmae = 0, tiles_dmae = 0;
    const Prob* gen = nullptr;
    int ngen = 0, tiles_gen[2] = {0, 0};
    bool vec = true;
    const Prob* nar[2] = {nullptr, nullptr};
    int nnar[2] = {0, 0}, tiles_nar[2] = {0, 0};
    bool vec_nar[2] = {true, true};
    const Seg* segs = nullptr;
};
// Device half.  The persistent mode needs a device int `counter`, which it
// zeroes with a kernel before each launch (graph safe).  `dma` runs the
// A-in-registers kernel (the lean direct-to-LDS kernel under
// DG_GEMM_NO_AREG), `dmae` the edge-handling direct-to-LDS kernel, `gen` the
// register-staged kernel of tile shape cfg; one launch per non-empty list.
int launch_gemm_f64_dev(char opA, char opB, int mode, int cfg, const GemmBatchDev& b, int* counter,
                        hipStream_t stream);
// Tile shape used by the one-shot entry points (DG_GEMM_TILE env override).
int default_tile_cfg();

// C[x] = beta*C[x] over an m x n column-major matrix (beta == 0 writes 0).
int launch_scale(int64_t m, int64_t n, double beta, double* C, int64_t ldc, hipStream_t stream);

// Smallest CRT moduli count (2..16) whose scaling reaches q = 53 bits for
// inner dimension k (16 when none does).
int crt_moduli_for_fp64(int64_t k);

// Ozaki scheme II in two parts.  A prepared operand holds the residues and
// line exponents of `rows` lines of length k (crt_operand_bytes bytes);
// crt_prepare builds one from lines x[i*rs + k*ks] with a scan buffer of
// crt_scan_bytes bytes; crt_multiply forms C = alpha op(A) op(B) + beta C
// from prepared op(A) rows and op(B) columns with a digit buffer of
// crt_digit_bytes bytes.  Buffers 256-byte aligned.
int64_t crt_operand_bytes(int64_t rows, int64_t k, int nmod);
int64_t crt_scan_bytes(int64_t rows, int64_t k);
int64_t crt_digit_bytes(int64_t m, int64_t n, int nmod);
int crt_prepare(const double* x, int64_t rows, int64_t k, int64_t rs, int64_t ks, int nmod, void* operand,
                void* scan, hipStream_t st);
int crt_multiply(int64_t m, int64_t n, int64_t k, int nmod, double alpha, const void* a_op, const void* b_op,
                 double beta, double* C, int64_t ldc, void* digits, hipStream_t st);

}  // namespace dg
