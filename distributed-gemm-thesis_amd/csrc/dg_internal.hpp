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
// graph is refused: tables must be built before capture.
const void* cached_table(const void* host, size_t bytes, hipStream_t stream);

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
    int32_t pad_;
};

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

// Host half of a launch: drops empty problems, assigns tile ranges and tells
// whether every segment allows 128-bit loads (`vec`) and whether every tile
// is interior for the direct-to-LDS kernel (`dma`: m and n multiples of 128,
// every segment's K a multiple of 8, and `vec`).
int prepare_gemm(const std::vector<Prob>& probs_in, const std::vector<Seg>& segs, std::vector<Prob>& probs,
                 bool* vec, int tiles[2], bool* dma = nullptr);
// Device half: tables already resident.  The persistent mode needs a device
// int `counter`, which it zeroes with a kernel before the launch (graph safe).
// With `dma` (every tile interior, see prepare_gemm) the 128x128 non-persistent
// launches use the direct-to-LDS kernel.
int launch_gemm_f64_dev(char opA, char opB, bool vec, int mode, int cfg, const Prob* dp, int np, const Seg* ds,
                        const int tiles[2], int* counter, hipStream_t stream, bool dma = false);
// Tile shape used by the one-shot entry points (DG_GEMM_TILE env override).
int default_tile_cfg();

// C[x] = beta*C[x] over an m x n column-major matrix (beta == 0 writes 0).
int launch_scale(int64_t m, int64_t n, double beta, double* C, int64_t ldc, hipStream_t stream);

// Smallest CRT moduli count (2..16) whose scaling reaches q = 53 bits for
// inner dimension k (16 when none does).
int crt_moduli_for_fp64(int64_t k);

}  // namespace dg
