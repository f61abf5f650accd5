// FP32 GEMM on the CDNA4 matrix cores (v_mfma_f32_16x16x4_f32).
//
// The FP64 engine's direct-to-LDS pipeline at the same byte rate: a k-step
// is 16 floats (64 bytes per row), so a 128x128 tile moves exactly the bytes
// of an FP64 k-step of 8 and does twice the flops, which the matrix cores
// deliver at twice the FP64 rate.
//   * 128x128 tiles, 4 waves of 64x64 (4x4 MFMA tiles of 16x16), two
//     workgroups per CU, a 3-stage LDS ring of 16 KB stages filled by
//     global_load_lds_dwordx4 (no register staging).
//   * k-contiguous operand (A for opA = T, B for opB = N): LDS [row][k],
//     64-byte rows, 16-byte chunk c of row r at position c ^ ((r >> 1) & 3);
//     a lane reads its four k with one 128-bit read.
//   * row-contiguous operand (A for opA = N, B for opB = T): LDS [k][row]
//     without padding, two k-rows per 1 KB load.  Lane lr of an MFMA block i
//     stands for row 4 lr + i, so a lane's rows of all four blocks are one
//     16-byte chunk of a k-row: one 128-bit read per k sub-step (16 lanes read
//     256 consecutive bytes) instead of four 32-bit ones.
// Lane l of an MFMA supplies k = 4*(l/16) + s at sub-step s (s = 0..3) for
// both operands, so each product term pairs the right A and B elements
// whatever order the matrix core sums in.  Result layout of
// v_mfma_f32_16x16x4_f32: lane l, register v holds D(4*(l/16) + v, l%16);
// with either row mapping each lane stores runs of four consecutive rows of
// one C column (16 bytes).
// The barrier runs one step ahead, as in the FP64 kernel: after 48 of the 64
// MFMAs of step k each wave waits for its own loads of step k+2, one barrier
// publishes them and frees the stage of step k (read during step k-1), which
// then receives step k+3 (+1.4 % over a 4-stage ring with the barrier after
// the MFMAs, at 16384^3).
//
// Interior problems with a k-contiguous operand (opA = T, or opA = opB = N)
// run sgemm_dma_xr_kernel instead, which keeps that operand in registers and
// only the other in LDS (+2.3 % at NN 16384^3; described at the kernel), and
// uneven ones are cut into such an interior block and edge strips
// (run_sgemm_dma), with the same bits as the whole problem on this kernel.
//
// Edges: rows past the end of a k-contiguous operand read its last row;
// row chunks past the end of a row-contiguous operand read the last 16-byte
// aligned chunk, which holds a valid element and so cannot cross a page;
// the products of those rows land only in unstored rows / columns of C.
// k past K reads a zero page (K must be a multiple of 4 when an operand is
// k-contiguous, so a 16-byte chunk never straddles K), and an odd step count
// gets one zero step.  Operands that are not 16-byte aligned, leading
// dimensions that are not multiples of 4, or K % 4 != 0 with a k-contiguous
// operand run sgemm_generic_kernel (LDS-tiled FMA) instead.
#include "dg_internal.hpp"

namespace dg {

namespace {

typedef float fv4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void gvoid_t;

namespace sk {
constexpr int BM = 128, BN = 128, BK = 16, STG = 3;
constexpr int PART = 128 * BK * 4;  // 8 KB per operand per stage
constexpr int SSZ = 2 * PART;
constexpr size_t LDS_BYTES = (size_t)STG * SSZ;  // 48 KB
constexpr int GROUP_M = 8;

__host__ __device__ constexpr int swz(int r) { return (r >> 1) & 3; }

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt field");
    // gfx9 s_waitcnt: vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt left at their maxima
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Zeros for k past K and for the padding step (never written): 16 bytes per lane.
__device__ __align__(256) float g_sgemm_zero[256];

struct Cursor {
    const float* p;
    int64_t second, step;  // offset of the wave's second load, advance per k-step (floats)
};

// Cursor of this lane at k = 0.  `r0` is the tile's first row of the
// operand, `rows` its row count.
template <bool KC, bool EDGE>
__device__ __forceinline__ Cursor make_cursor(const float* base, int64_t ld, int64_t r0, int64_t rows, int wave,
                                              int lane) {
    if constexpr (KC) {
        // rows 32w .. 32w+31 in two loads of 16 rows; lane -> row lane/4, slot lane%4
        const int row = lane >> 2;
        const int chunk = (lane & 3) ^ swz(row);
        int64_t ra = r0 + 32 * wave + row, rb = ra + 16;
        if (EDGE) {
            ra = min(ra, rows - 1);
            rb = min(rb, rows - 1);
        }
        return Cursor{base + ra * ld + 4 * chunk, (rb - ra) * ld, BK};
    } else {
        // k-rows 4w + h (first load) and 4w + 2 + h (second), h = lane / 32;
        // the lane moves rows 4 (lane % 32) .. + 3 of its k-row
        const int h = lane >> 5;
        const int c = lane & 31;
        int64_t r = r0 + 4 * c;
        if (EDGE) r = min(r, (rows - 1) & ~(int64_t)3);
        return Cursor{base + r + (int64_t)(4 * wave + h) * ld, 2 * ld, BK * ld};
    }
}

__device__ __forceinline__ Cursor zero_cursor(int lane) { return Cursor{g_sgemm_zero + 4 * lane, 0, 0}; }

// One k-step of one operand into `dst`; with tail != 0 only the first `tail`
// k of the step are real and the lanes past them load zeros.
template <bool KC>
__device__ __forceinline__ void issue_operand(char* dst, Cursor& c, int wave, int lane, int tail) {
    char* d0 = KC ? dst + (32 * wave) * 64 : dst + (2 * wave) * 1024;
    char* d1 = KC ? d0 + 16 * 64 : d0 + 1024;
    const float* p0 = c.p;
    const float* p1 = c.p + c.second;
    if (tail) {
        const float* z = g_sgemm_zero + 4 * lane;
        if constexpr (KC) {
            const int k4 = 4 * ((lane & 3) ^ swz(lane >> 2));
            if (k4 >= tail) { p0 = z; p1 = z; }
        } else {
            const int h = lane >> 5;
            if (4 * wave + h >= tail) p0 = z;
            if (4 * wave + 2 + h >= tail) p1 = z;
        }
    }
    __builtin_amdgcn_global_load_lds((gvoid_t*)p0, (lds_void_t*)d0, 16, 0, 0);
    __builtin_amdgcn_global_load_lds((gvoid_t*)p1, (lds_void_t*)d1, 16, 0, 0);
    c.p += c.step;
}

// f[i][s] = X(row, k = 4 lg + s) of one stage, row = w0 + 16i + lr for a
// k-contiguous operand and w0 + 4lr + i for a row-contiguous one (whose four
// rows of a lane are then one 16-byte chunk of each k-row: 16 lanes read 256
// consecutive bytes).
template <bool KC>
__device__ __forceinline__ void read_operand(float (&f)[4][4], const char* part, int w0, int lr, int lg) {
    if constexpr (KC) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const fv4 v = *reinterpret_cast<const fv4*>(part + (w0 + 16 * i + lr) * 64 + 16 * (lg ^ swz(lr)));
#pragma unroll
            for (int s = 0; s < 4; ++s) f[i][s] = v[s];
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const fv4 v = *reinterpret_cast<const fv4*>(part + (4 * lg + s) * 512 + (w0 + 4 * lr) * 4);
#pragma unroll
            for (int i = 0; i < 4; ++i) f[i][s] = v[i];
        }
    }
}
}  // namespace sk

struct SgemmArgs {
    const float* a;
    const float* b;
    float* c;
    int64_t lda, ldb, ldc;
    int64_t m, n, k;
    float alpha, beta;
    int tiles_m, tiles_n;
    int cvec;  // C 16-byte aligned with ldc % 4 == 0: vector stores
};

__device__ __forceinline__ int xcd_remap(int bid, int total) {
    constexpr int X = 8;
    const int q = total / X, r = total % X;
    const int x = bid % X, slot = bid / X;
    return x * q + min(x, r) + slot;
}

// EDGE = false: m, n multiples of 128, K a multiple of 16 with an even step count.
template <bool AK, bool BKC, bool EDGE>
__global__ __launch_bounds__(256, 2) void sgemm_dma_kernel(SgemmArgs P) {
    using namespace sk;
    extern __shared__ __attribute__((aligned(1024))) char lds[];
    const int total = P.tiles_m * P.tiles_n;
    const int tile = xcd_remap(blockIdx.x, total);
    const int per_group = GROUP_M * P.tiles_n;
    const int group = tile / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(P.tiles_m - first_m, GROUP_M);
    const int in_group = tile - group * per_group;
    const int64_t m0 = (int64_t)(first_m + in_group % gsize) * BM;
    const int64_t n0 = (int64_t)(in_group / gsize) * BN;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lr = lane & 15, lg = lane >> 4;

    int64_t left = (P.k + BK - 1) / BK;
    const int tail = (int)(P.k % BK);
    Cursor ca = make_cursor<AK, EDGE>(P.a, P.lda, m0, P.m, wave, lane);
    Cursor cb = make_cursor<BKC, EDGE>(P.b, P.ldb, n0, P.n, wave, lane);
    int ist = 0;
    auto issue_next = [&]() {
        char* st = lds + ist * SSZ;
        const int t = (EDGE && left == 1) ? tail : 0;
        issue_operand<AK>(st, ca, wave, lane, t);
        issue_operand<BKC>(st + PART, cb, wave, lane, t);
        ist = ist + 1 == STG ? 0 : ist + 1;
        if (--left == 0) {
            left = INT64_MAX;
            ca = zero_cursor(lane);
            cb = zero_cursor(lane);
        }
    };
    int64_t nk = (P.k + BK - 1) / BK;
    nk += nk & 1;

    fv4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fv4{0.f, 0.f, 0.f, 0.f};

    static_assert(STG == 3, "the waits below assume a 3-stage ring");
    for (int64_t s = 0; s < STG && s < nk; ++s) issue_next();
    // steps 0 and 1 published (4 loads per wave and step; step 2 may be in flight)
    if (nk >= 3) wait_vmcnt<4>(); else wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    int rst = 1;  // stage of step kt + 1
    float fa0[4][4], fb0[4][4], fa1[4][4], fb1[4][4];
    read_operand<AK>(fa0, lds, wm, lr, lg);
    read_operand<BKC>(fb0, lds + PART, wn, lr, lg);
    // LDS read instructions of each operand per step and the MFMAs interleaved
    // after each; MFMAs f = 16s + 4i + j in order, the barrier after f = 47
    constexpr int NRA = 4, NRB = 4;  // one 128-bit read per block (k-contiguous) or k sub-step (row-contiguous)
    constexpr int MA = 8, MB = 3;
    int64_t kt = 0;
#define SGEMM_MFMAS(CA, CB, F0, F1)                                                                 \
    _Pragma("unroll") for (int f_ = F0; f_ < F1; ++f_) {                                            \
        const int s_ = f_ >> 4, i_ = (f_ >> 2) & 3, j_ = f_ & 3;                                    \
        acc[i_][j_] = __builtin_amdgcn_mfma_f32_16x16x4f32(CA[i_][s_], CB[j_][s_], acc[i_][j_], 0, 0, 0); \
    }
#define SGEMM_STEP(CA, CB, NA, NB)                                                                  \
    {                                                                                               \
        const char* nst = lds + rst * SSZ; /* harmless past the last step */                        \
        read_operand<AK>(NA, nst, wm, lr, lg);                                                      \
        SGEMM_MFMAS(CA, CB, 0, 48)                                                                  \
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                          \
        _Pragma("unroll") for (int g_ = 0; g_ < NRA; ++g_) {                                        \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                      \
            __builtin_amdgcn_sched_group_barrier(0x008, MA, 0);                                     \
        }                                                                                           \
        rst = rst + 1 == STG ? 0 : rst + 1;                                                         \
        if (kt + 2 < nk) {                                                                          \
            /* LDS reads retire in order: once at most 4 of the >= 4 reads just issued are pending, \
               the stage of step kt (read during step kt-1) is no longer being read */             \
            __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4) | (4 << 8));                  \
            wait_vmcnt<0>(); /* step kt+2, the only loads in flight */                              \
            __builtin_amdgcn_s_barrier();                                                           \
            if (kt + STG < nk) issue_next();                                                        \
        }                                                                                           \
        read_operand<BKC>(NB, nst + PART, wn, lr, lg);                                              \
        SGEMM_MFMAS(CA, CB, 48, 64)                                                                 \
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                          \
        _Pragma("unroll") for (int g_ = 0; g_ < NRB; ++g_) {                                        \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                      \
            __builtin_amdgcn_sched_group_barrier(0x008, MB, 0);                                     \
        }                                                                                           \
        ++kt;                                                                                       \
    }
    while (kt < nk) {
        SGEMM_STEP(fa0, fb0, fa1, fb1);
        SGEMM_STEP(fa1, fb1, fa0, fb0);
    }
#undef SGEMM_STEP
#undef SGEMM_MFMAS

    // acc[i][j][v] = sum for C(row(i, 4lg + v), col(j, lr)): row(i, r) = wm + 16i + r
    // for a k-contiguous A and wm + 4r + i for a row-contiguous one, col(j, c)
    // likewise for B; each lane stores runs of four consecutive rows
    const float alpha = P.alpha, beta = P.beta;
    const bool interior = !EDGE || (m0 + BM <= P.m && n0 + BN <= P.n);
    auto store4 = [&](int64_t m, int64_t n, fv4 acc4) {
        if (interior && P.cvec) {
            __attribute__((address_space(1))) fv4* cp = (__attribute__((address_space(1))) fv4*)(P.c + m + n * P.ldc);
            fv4 val = alpha * acc4;
            if (beta != 0.f) {
                const fv4 old = *cp;
#pragma unroll
                for (int v = 0; v < 4; ++v) val[v] = fmaf(beta, old[v], val[v]);
            }
            *cp = val;
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (m + v >= P.m || n >= P.n) continue;
                __attribute__((address_space(1))) float* cp =
                    (__attribute__((address_space(1))) float*)(P.c + (m + v) + n * P.ldc);
                float val = alpha * acc4[v];
                if (beta != 0.f) val = fmaf(beta, *cp, val);
                *cp = val;
            }
        }
    };
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t n = n0 + (BKC ? wn + 16 * j + lr : wn + 4 * lr + j);
        if constexpr (AK) {
#pragma unroll
            for (int i = 0; i < 4; ++i) store4(m0 + wm + 16 * i + 4 * lg, n, acc[i][j]);
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v)
                store4(m0 + wm + 16 * lg + 4 * v, n, fv4{acc[0][j][v], acc[1][j][v], acc[2][j][v], acc[3][j][v]});
        }
    }
}

// Interior problems (m, n multiples of 128, K a multiple of 16 with an even
// step count) with one k-contiguous operand X straight into registers and the
// other, Y, through the LDS ring (8 KB stages): each wave owns 32 rows of X
// (rows of C for X = A, columns for X = B) against all 128 rows of Y,
// loads them from global memory into a 3-step register ring (one 16-byte
// load per lane and 16-row block: the lane's four k of a step) and reads
// only Y from LDS.  A stays the MFMA's first operand and k = 4 lg + s at
// sub-step s, sub-steps in order within each step, so every element is
// summed in the order of sgemm_dma_kernel: the two agree bit for bit.
// X = A for opA = T (any opB), X = B for opA = N, opB = N; both
// row-contiguous (opA = N, opB = T) stays on sgemm_dma_kernel.  With X in
// registers only Y crosses LDS, and the next step's Y fragments are all read
// before the barrier (a row-contiguous Y lost 5 % when half of them were
// read after it).  Loads in issue order: Y(k+3) at the barrier of step k,
// X(k+3) after its MFMAs, so the barrier's wait for Y(k+2) leaves the loads
// of X(k+2) in flight.
// XB: X is B.  YKC: Y is k-contiguous.
template <bool XB, bool YKC>
__global__ __launch_bounds__(256, 2) void sgemm_dma_xr_kernel(SgemmArgs P) {
    using namespace sk;
    typedef __attribute__((address_space(1))) const fv4 gfv4;
    extern __shared__ __attribute__((aligned(1024))) char lds[];
    const int total = P.tiles_m * P.tiles_n;
    const int tile = xcd_remap(blockIdx.x, total);
    const int per_group = GROUP_M * P.tiles_n;
    const int group = tile / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(P.tiles_m - first_m, GROUP_M);
    const int in_group = tile - group * per_group;
    const int64_t m0 = (int64_t)(first_m + in_group % gsize) * BM;
    const int64_t n0 = (int64_t)(in_group / gsize) * BN;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 15, lg = lane >> 4;
    const int64_t x0 = (XB ? n0 : m0) + 32 * wave;  // the wave's first row of X
    const int64_t y0 = XB ? m0 : n0;
    const int64_t ldx = XB ? P.ldb : P.lda;

    // this lane's X at the next step to load: row x0 + 16i + lr, k = 4lg .. 4lg + 3
    const float* xp = (XB ? P.b : P.a) + (x0 + lr) * ldx + 4 * lg;
    const int64_t x_blk = 16 * ldx;
    // rx[slot][i] of the slot's step: X(row x0 + 16i + lr, k = 4lg + s) in component s
    fv4 rx[3][2];
    auto load_x = [&](fv4 (&r)[2]) {
        r[0] = *(gfv4*)xp;
        r[1] = *(gfv4*)(xp + x_blk);
        xp += BK;
    };
    Cursor cy = make_cursor<YKC, false>(XB ? P.a : P.b, XB ? P.lda : P.ldb, y0, XB ? P.m : P.n, wave, lane);
    int ist = 0;
    auto issue_y = [&]() {
        issue_operand<YKC>(lds + ist * PART, cy, wave, lane, 0);
        ist = ist + 1 == STG ? 0 : ist + 1;
    };
    const int64_t nk = P.k / BK;  // even, >= 2

    fv4 acc[2][8];  // acc[i][j]: X block i against Y block j = 4h + jj
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = fv4{0.f, 0.f, 0.f, 0.f};

    static_assert(STG == 3, "the waits below assume a 3-stage ring");
    issue_y();
    load_x(rx[0]);
    issue_y();
    load_x(rx[1]);
    if (nk > 2) {
        issue_y();
        load_x(rx[2]);
        wait_vmcnt<6>();  // Y of steps 0 and 1 landed; X(1), Y(2), X(2) may be in flight
    } else {
        wait_vmcnt<0>();
    }
    __builtin_amdgcn_s_barrier();
    int rst = 1;  // stage of step kt + 1
    float fy0[2][4][4], fy1[2][4][4];  // f[h][jj][s]: Y block 4h + jj, k = 4lg + s
    read_operand<YKC>(fy0[0], lds, 0, lr, lg);
    read_operand<YKC>(fy0[1], lds, 64, lr, lg);
    int64_t kt = 0;
    // MFMAs f = 32h + 8s + 4i + jj in order, the barrier after f = 47
#define XR_MFMAS(SLOT, CY, F0, F1)                                                                  \
    _Pragma("unroll") for (int f_ = F0; f_ < F1; ++f_) {                                            \
        const int h_ = f_ >> 5, s_ = (f_ >> 3) & 3, i_ = (f_ >> 2) & 1, jj_ = f_ & 3;               \
        const float xv_ = rx[SLOT][i_][s_], yv_ = CY[h_][jj_][s_];                                  \
        acc[i_][4 * h_ + jj_] = XB ? __builtin_amdgcn_mfma_f32_16x16x4f32(yv_, xv_, acc[i_][4 * h_ + jj_], 0, 0, 0) \
                                   : __builtin_amdgcn_mfma_f32_16x16x4f32(xv_, yv_, acc[i_][4 * h_ + jj_], 0, 0, 0); \
    }
// FULL: kt + 3 < nk is known (the barrier, the issue of Y(kt+3) and the load of X(kt+3) all run)
#define XR_STEP(FULL, SLOT, CY, NY)                                                                 \
    {                                                                                               \
        const char* nst = lds + rst * PART; /* harmless past the last step */                       \
        /* row-contiguous Y: the previous step's fragment reads (all before its barrier) are done;  \
           saying so spares the compiler's own waits, which cannot count across the LDS-direct      \
           loads.  Not for a k-contiguous Y: there the wait makes the compiler sink MFMAs of        \
           earlier steps past the barriers until it spills */                                       \
        if constexpr (!YKC) __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4));             \
        read_operand<YKC>(NY[0], nst, 0, lr, lg);                                                   \
        read_operand<YKC>(NY[1], nst, 64, lr, lg);                                                  \
        XR_MFMAS(SLOT, CY, 0, 48)                                                                   \
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                          \
        _Pragma("unroll") for (int g_ = 0; g_ < 8; ++g_) {                                          \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                      \
            __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);                                      \
        }                                                                                           \
        rst = rst + 1 == STG ? 0 : rst + 1;                                                         \
        if (FULL || kt + 2 < nk) {                                                                  \
            /* the stage of step kt was read during step kt-1: once at most the 8 reads just        \
               issued are pending, none of it is still being read */                                \
            __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4) | (8 << 8));                  \
            wait_vmcnt<2>(); /* Y(kt+2); X(kt+2), issued after it, may be in flight */              \
            __builtin_amdgcn_s_barrier();                                                           \
            if (FULL || kt + STG < nk) issue_y();                                                   \
        }                                                                                           \
        XR_MFMAS(SLOT, CY, 48, 64)                                                                  \
        if (FULL || kt + STG < nk) load_x(rx[SLOT]);                                                \
        ++kt;                                                                                       \
    }
    // X slot = step % 3, Y fragments alternate: six steps per turn.  The
    // steady turns have no conditions (so the waits the compiler derives
    // across the loop stay partial); the last 2 to 8 steps, nk even, follow.
    while (kt + 9 <= nk) {
        XR_STEP(true, 0, fy0, fy1);
        XR_STEP(true, 1, fy1, fy0);
        XR_STEP(true, 2, fy0, fy1);
        XR_STEP(true, 0, fy1, fy0);
        XR_STEP(true, 1, fy0, fy1);
        XR_STEP(true, 2, fy1, fy0);
    }
    while (true) {
        XR_STEP(false, 0, fy0, fy1);
        XR_STEP(false, 1, fy1, fy0);
        if (kt >= nk) break;
        XR_STEP(false, 2, fy0, fy1);
        XR_STEP(false, 0, fy1, fy0);
        if (kt >= nk) break;
        XR_STEP(false, 1, fy0, fy1);
        XR_STEP(false, 2, fy1, fy0);
        if (kt >= nk) break;
    }
#undef XR_STEP
#undef XR_MFMAS

    // Result v of block (i, j = 4h + jj) in a lane: the MFMA's row 4lg + v
    // and column lr.  X = A: C(x0 + 16i + 4lg + v, n0 + yrow(h, jj, lr));
    // X = B: C(m0 + yrow(h, jj, 4lg + v), x0 + 16i + lr), with yrow(h, jj, r)
    // = 64h + 16jj + r for a k-contiguous Y and 64h + 4r + jj for a
    // row-contiguous one.  Each lane stores runs of four consecutive rows.
    const float alpha = P.alpha, beta = P.beta;
    auto store4 = [&](int64_t m, int64_t n, fv4 acc4) {
        float* c = P.c + m + n * P.ldc;
        fv4 val = alpha * acc4;
        if (P.cvec) {
            __attribute__((address_space(1))) fv4* cp = (__attribute__((address_space(1))) fv4*)c;
            if (beta != 0.f) {
                const fv4 old = *cp;
#pragma unroll
                for (int v = 0; v < 4; ++v) val[v] = fmaf(beta, old[v], val[v]);
            }
            *cp = val;
        } else {
            __attribute__((address_space(1))) float* cp = (__attribute__((address_space(1))) float*)c;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float x = val[v];
                if (beta != 0.f) x = fmaf(beta, cp[v], x);
                cp[v] = x;
            }
        }
    };
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if constexpr (!XB) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj)
                    store4(x0 + 16 * i + 4 * lg, n0 + (YKC ? 64 * h + 16 * jj + lr : 64 * h + 4 * lr + jj),
                           acc[i][4 * h + jj]);
            } else if constexpr (YKC) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) store4(m0 + 64 * h + 16 * jj + 4 * lg, x0 + 16 * i + lr, acc[i][4 * h + jj]);
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    store4(m0 + 64 * h + 16 * lg + 4 * v, x0 + 16 * i + lr,
                           fv4{acc[i][4 * h][v], acc[i][4 * h + 1][v], acc[i][4 * h + 2][v], acc[i][4 * h + 3][v]});
            }
        }
    }
}

// Any layout: 64x64 tiles, k-steps of 16 staged through LDS with guarded
// loads (zeros past the edges), each thread a 4x4 block of C in FP32 FMAs.
__global__ __launch_bounds__(256) void sgemm_generic_kernel(int ta, int tb, SgemmArgs P) {
    __shared__ float As[16][64 + 1];
    __shared__ float Bs[16][64 + 1];
    const int t = threadIdx.x;
    const int tx = t & 15, ty = t >> 4;
    const int64_t m0 = (int64_t)blockIdx.x * 64, n0 = (int64_t)blockIdx.y * 64;
    float acc[4][4] = {};
    for (int64_t k0 = 0; k0 < P.k; k0 += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            // A: element (row, kk) of op(A); stored (row, kk) at row + kk*lda (N) or kk + row*lda (T)
            int r, kk;
            if (!ta) { r = e & 63; kk = e >> 6; } else { kk = e & 15; r = e >> 4; }
            int64_t gr = m0 + r, gk = k0 + kk;
            As[kk][r] = (gr < P.m && gk < P.k) ? (ta ? P.a[gk + gr * P.lda] : P.a[gr + gk * P.lda]) : 0.f;
            // B: element (kk, col) of op(B); stored at kk + col*ldb (N) or col + kk*ldb (T)
            if (!tb) { kk = e & 15; r = e >> 4; } else { r = e & 63; kk = e >> 6; }
            gr = n0 + r; gk = k0 + kk;
            Bs[kk][r] = (gr < P.n && gk < P.k) ? (tb ? P.b[gr + gk * P.ldb] : P.b[gk + gr * P.ldb]) : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t m = m0 + tx + 16 * i;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t n = n0 + ty + 16 * j;
            if (m >= P.m || n >= P.n) continue;
            float* cp = P.c + m + n * P.ldc;
            float val = P.alpha * acc[i][j];
            if (P.beta != 0.f) val = fmaf(P.beta, *cp, val);
            *cp = val;
        }
    }
}

__global__ void sgemm_scale_kernel(int64_t m, int64_t n, float beta, float* __restrict__ C, int64_t ldc) {
    const int64_t total = m * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % m, j = e / m;
        float* p = C + i + j * ldc;
        *p = beta == 0.f ? 0.f : beta * *p;
    }
}

template <bool AK, bool BKC>
void launch_sgemm_dma(bool edge, const SgemmArgs& a, int tiles, hipStream_t st) {
    // DG_GEMM_NO_AREG keeps both operands in LDS (same results; for comparisons)
    static const bool no_areg = getenv("DG_GEMM_NO_AREG") != nullptr;
    constexpr size_t xr_lds = (size_t)sk::STG * sk::PART;
    if constexpr (AK || BKC) {
        if (!edge && !no_areg) {
            // the k-contiguous operand into registers: A if it is, else B (opA = N, opB = N)
            hipLaunchKernelGGL((sgemm_dma_xr_kernel<!AK, AK ? BKC : false>), dim3(tiles), dim3(256), xr_lds, st, a);
            return;
        }
    }
    if (edge)
        hipLaunchKernelGGL((sgemm_dma_kernel<AK, BKC, true>), dim3(tiles), dim3(256), sk::LDS_BYTES, st, a);
    else
        hipLaunchKernelGGL((sgemm_dma_kernel<AK, BKC, false>), dim3(tiles), dim3(256), sk::LDS_BYTES, st, a);
}

// One direct-to-LDS problem.  With a k-contiguous operand, whole even
// k-steps and at least 8 whole interior tiles, an uneven problem is cut into
// its interior block (register-operand kernel) and the strips after it (rows
// past the interior, then the columns past it beside the interior rows),
// which run the edge kernel: the cut adds no tiles, and since both kernels
// sum every element in the same order it changes no result bit.
template <bool AK, bool BKC>
void run_sgemm_dma(const SgemmArgs& a, hipStream_t st) {
    auto one = [&](SgemmArgs p) {
        p.tiles_m = (int)((p.m + sk::BM - 1) / sk::BM);
        p.tiles_n = (int)((p.n + sk::BN - 1) / sk::BN);
        const int64_t steps = (p.k + sk::BK - 1) / sk::BK;
        const bool edge = p.m % sk::BM || p.n % sk::BN || p.k % sk::BK || (steps & 1);
        launch_sgemm_dma<AK, BKC>(edge, p, p.tiles_m * p.tiles_n, st);
    };
    static const bool no_areg = getenv("DG_GEMM_NO_AREG") != nullptr;
    const int64_t mi = a.m / sk::BM * sk::BM, ni = a.n / sk::BN * sk::BN;
    const bool k_ok = a.k % sk::BK == 0 && !((a.k / sk::BK) & 1);
    if (!(AK || BKC) || no_areg || !k_ok || (mi == a.m && ni == a.n) || mi * ni < 8 * sk::BM * sk::BN) {
        one(a);
        return;
    }
    SgemmArgs in = a;
    in.m = mi;
    in.n = ni;
    one(in);
    if (mi < a.m) {
        SgemmArgs s = a;  // rows mi .. m of every column
        s.a = AK ? a.a + mi * a.lda : a.a + mi;
        s.c = a.c + mi;
        s.m = a.m - mi;
        one(s);
    }
    if (ni < a.n) {
        SgemmArgs s = a;  // columns ni .. n of rows 0 .. mi
        s.b = BKC ? a.b + ni * a.ldb : a.b + ni;
        s.c = a.c + ni * a.ldc;
        s.m = mi;
        s.n = a.n - ni;
        one(s);
    }
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

}  // namespace

}  // namespace dg

extern "C" int dg_sgemm(char opA, char opB, int64_t m, int64_t n, int64_t k, float alpha, const float* A,
                        int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc,
                        void* stream) {
    using namespace dg;
    const char oa = norm_op(opA), ob = norm_op(opB);
    if (!oa) return fail(DG_ERR_INVALID, "dg_sgemm: invalid opA '%c'", opA);
    if (!ob) return fail(DG_ERR_INVALID, "dg_sgemm: invalid opB '%c'", opB);
    if (m < 0 || n < 0 || k < 0) return fail(DG_ERR_INVALID, "dg_sgemm: negative dimension");
    const int64_t rowsA = oa == 'N' ? m : k;
    const int64_t rowsB = ob == 'N' ? k : n;
    if (lda < (rowsA > 1 ? rowsA : 1)) return fail(DG_ERR_INVALID, "dg_sgemm: lda %lld < %lld", (long long)lda, (long long)rowsA);
    if (ldb < (rowsB > 1 ? rowsB : 1)) return fail(DG_ERR_INVALID, "dg_sgemm: ldb %lld < %lld", (long long)ldb, (long long)rowsB);
    if (ldc < (m > 1 ? m : 1)) return fail(DG_ERR_INVALID, "dg_sgemm: ldc %lld < %lld", (long long)ldc, (long long)m);
    if (m == 0 || n == 0) return DG_OK;
    int rc = require_device();
    if (rc != DG_OK) return rc;
    hipStream_t st = as_stream(stream);
    if (!C) return fail(DG_ERR_INVALID, "dg_sgemm: null C");
    if (k == 0 || alpha == 0.f) {
        int64_t blocks = (m * n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(sgemm_scale_kernel, dim3((unsigned)blocks), dim3(256), 0, st, m, n, beta, C, ldc);
        DG_LAUNCH_CHECK();
        return DG_OK;
    }
    if (!A || !B) return fail(DG_ERR_INVALID, "dg_sgemm: null matrix pointer");
    const bool ak = oa == 'T', bkc = ob == 'N';  // operands stored contiguous along k
    const int64_t tm = (m + sk::BM - 1) / sk::BM, tn = (n + sk::BN - 1) / sk::BN;
    if (tm * tn > INT32_MAX || (m + 63) / 64 > INT32_MAX || (n + 63) / 64 > 65535)
        return fail(DG_ERR_INVALID, "dg_sgemm: %lld x %lld output is too large", (long long)m, (long long)n);
    SgemmArgs a{A, B, C, lda, ldb, ldc, m, n, k, alpha, beta, (int)tm, (int)tn, 0};
    a.cvec = aligned16(C) && ldc % 4 == 0;
    const bool dma = aligned16(A) && aligned16(B) && lda % 4 == 0 && ldb % 4 == 0 && (!(ak || bkc) || k % 4 == 0);
    if (dma) {
        if (ak && bkc) run_sgemm_dma<true, true>(a, st);
        else if (ak) run_sgemm_dma<true, false>(a, st);
        else if (bkc) run_sgemm_dma<false, true>(a, st);
        else run_sgemm_dma<false, false>(a, st);
    } else {
        hipLaunchKernelGGL(sgemm_generic_kernel, dim3((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64)), dim3(256),
                           0, st, (int)ak, (int)(ob == 'T'), a);
    }
    DG_LAUNCH_CHECK();
    return DG_OK;
}
