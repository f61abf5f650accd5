// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64), grouped over
// problems and segmented along K.
//
// One launch computes every 128x128 output tile of every problem in the
// problem table.  A problem is C = alpha * sum_s opA(seg s) * opB(seg s)
// + beta*C, where segment s supplies its own A/B base pointers and leading
// dimensions.  A plain DGEMM is one problem with one segment; a virtual-grid
// Cannon rank is one problem whose segments are the tiles the rotations
// would have brought in; a SUMMA rank is one problem whose segments are the
// owners' panels.
//
// Three kernels share the problem/segment tables and the tile order.  Batches
// whose tiles are all interior run gemm_f64_dma_ar_kernel (A straight into
// registers, B direct to LDS); the others that take direct loads run
// gemm_f64_dma_kernel (both operands direct to LDS); both are described below.  The rest run gemm_f64_kernel:
// Tiling (256 threads = 4 waves of 64 lanes):
//   block tile 128 x 128, k-step 16, LDS double buffered (2 blocks per CU);
//   each wave owns a 64 x 64 sub-tile = 4 x 4 MFMA tiles of 16 x 16.
// LDS holds both operands k-contiguous, [row][k] with a padded stride of 18
// doubles (144 B), so a lane reads its 4 consecutive k values with two
// 128-bit LDS reads and 16 lanes of one read touch 64 distinct banks.
// Lane l of an MFMA supplies k = 4*(l/16) + kk for sub-step kk; A and B use
// the same k assignment, so the product is exact whatever k order the
// matrix core sums in.
// The MFMA is issued as D = B^T A^T, so each output register holds 16
// consecutive rows of one C column: the epilogue stores are 128-byte runs.
// Result layout of v_mfma_f64_16x16x4_f64: lane l, register v holds
// D(l/16 + 4v, l%16) (checked on the device by the exact-integer tests).
#include <string>

#include "dg_internal.hpp"

namespace dg {

namespace {

typedef double dv2 __attribute__((ext_vector_type(2)));
typedef double dv4 __attribute__((ext_vector_type(4)));

// Global-address-space views: loads through them compile to global_load_*
// (counted by vmcnt only), not flat_load_*, whose lgkmcnt share would make
// every LDS wait also wait for the prefetch of the next k-step.
typedef __attribute__((address_space(1))) const double gdouble;
typedef __attribute__((address_space(1))) const dv2 gdv2;
__device__ __forceinline__ double gld(const double* p) { return *(gdouble*)p; }
__device__ __forceinline__ dv2 gld2(const double* p) { return *(gdv2*)p; }

constexpr int BK = 16;
constexpr int LDK = BK + 2;  // padded LDS row stride (doubles)
constexpr int GROUP_M = 8;   // tile rows per raster group

// Tile shapes: BM x BN output tiles computed by NTH threads, each wave owning
// a WM x WN sub-tile (WM/16 x WN/16 MFMA tiles of 16 x 16); OCC workgroups
// per CU.
template <int BM_, int BN_, int NTH_, int OCC_, int WM_ = 64, int WN_ = 64>
struct TileCfg {
    static constexpr int BM = BM_, BN = BN_, NTH = NTH_, OCC = OCC_, WM = WM_, WN = WN_;
    static constexpr int WAVES_N = BN / WN;
    static constexpr int FM = WM / 16, FN = WN / 16;
    static constexpr int PA = BM * BK / (2 * NTH);  // 128-bit loads per thread per k-step (A)
    static constexpr int PB = BN * BK / (2 * NTH);  // (B)
    static_assert((BM / WM) * (BN / WN) * 64 == NTH, "one wave per WM x WN sub-tile");
};
using Cfg128 = TileCfg<128, 128, 256, 2>;  // 73.7 KB LDS, two workgroups per CU
using Cfg256 = TileCfg<256, 128, 512, 1>;  // 110.6 KB LDS, 8 waves, half the B traffic per flop
using Cfg128w8 = TileCfg<128, 128, 512, 2, 64, 32>;  // 8 waves of 64 x 32: four waves per SIMD
// Strips narrower than a 128x128 tile, cut off uneven direct-to-LDS problems:
// columns past the interior (fewer than 128) and rows past it.
using CfgN32 = TileCfg<128, 32, 256, 2, 32, 32>;  // 4 waves of 32 x 32 stacked in M
using CfgM64 = TileCfg<64, 128, 256, 2, 32, 64>;  // 2 x 2 waves of 32 x 64 (2 ms per step faster than
                                                  // 32 x 128 tiles on the 20000 batch's 41-row strips)
constexpr int64_t kNarrowMax = 64;  // widest strip (rows or columns) put on narrow tiles

// Block id -> tile id so that each XCD (blocks are dealt round-robin to the
// 8 XCDs) receives one contiguous range of tiles, which the grouped raster
// below turns into a compact patch of C whose A/B panels share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int total) {
    constexpr int X = 8;
    const int q = total / X, r = total % X;
    const int x = bid % X, slot = bid / X;
    return x * q + min(x, r) + slot;
}

// Loads one k-step of an operand stored "k-contiguous" (element (row, k) at
// base[k + row*ld]) into registers: ROWS rows x 16 k; thread t takes k pair
// 2*(t%8) of rows t/8 + (NTH/8)*q.
template <int ROWS, int NTH, bool EDGE, bool VEC>
__device__ __forceinline__ void load_kc(dv2 (&r)[ROWS * BK / (2 * NTH)], const double* __restrict__ base, int64_t ld,
                                        int64_t row0, int64_t rows, int64_t k0, int64_t klen) {
    const int t = threadIdx.x;
    const int kp = (t & 7) * 2;
    const int rr = t >> 3;
#pragma unroll
    for (int q = 0; q < ROWS * BK / (2 * NTH); ++q) {
        int64_t row = row0 + rr + (NTH / 8) * q;
        int64_t k = k0 + kp;
        if (!EDGE) {
            const double* p = base + k + row * ld;
            if (VEC) {
                r[q] = gld2(p);
            } else {
                r[q].x = gld(p);
                r[q].y = gld(p + 1);
            }
        } else {
            row = row < rows ? row : rows - 1;  // clamped rows feed outputs that are never stored
            const double* p = base + row * ld;
            r[q].x = (k < klen) ? gld(p + k) : 0.0;
            r[q].y = (k + 1 < klen) ? gld(p + k + 1) : 0.0;
        }
    }
}

// Loads one k-step of an operand stored "row-contiguous" (element (row, k)
// at base[row + k*ld]): thread t takes rows 2*(t%(ROWS/2)), +1 of k column
// t/(ROWS/2) + (NTH/(ROWS/2))*q.
template <int ROWS, int NTH, bool EDGE, bool VEC>
__device__ __forceinline__ void load_rc(dv2 (&r)[ROWS * BK / (2 * NTH)], const double* __restrict__ base, int64_t ld,
                                        int64_t row0, int64_t rows, int64_t k0, int64_t klen) {
    constexpr int HP = ROWS / 2;
    const int t = threadIdx.x;
    const int rp = (t % HP) * 2;
    const int kc = t / HP;
#pragma unroll
    for (int q = 0; q < ROWS * BK / (2 * NTH); ++q) {
        int64_t row = row0 + rp;
        int64_t k = k0 + kc + (NTH / HP) * q;
        if (!EDGE) {
            const double* p = base + row + k * ld;
            if (VEC) {
                r[q] = gld2(p);
            } else {
                r[q].x = gld(p);
                r[q].y = gld(p + 1);
            }
        } else {
            if (k < klen) {
                const double* p = base + k * ld;
                int64_t ra = row < rows ? row : rows - 1;
                int64_t rb = row + 1 < rows ? row + 1 : rows - 1;
                r[q].x = gld(p + ra);
                r[q].y = gld(p + rb);
            } else {
                r[q].x = 0.0;
                r[q].y = 0.0;
            }
        }
    }
}

template <int ROWS, int NTH>
__device__ __forceinline__ void store_kc(double* __restrict__ s, const dv2 (&r)[ROWS * BK / (2 * NTH)]) {
    const int t = threadIdx.x;
    const int kp = (t & 7) * 2;
    const int rr = t >> 3;
#pragma unroll
    for (int q = 0; q < ROWS * BK / (2 * NTH); ++q) *reinterpret_cast<dv2*>(&s[(rr + (NTH / 8) * q) * LDK + kp]) = r[q];
}

template <int ROWS, int NTH>
__device__ __forceinline__ void store_rc(double* __restrict__ s, const dv2 (&r)[ROWS * BK / (2 * NTH)]) {
    constexpr int HP = ROWS / 2;
    const int t = threadIdx.x;
    const int rp = (t % HP) * 2;
    const int kc = t / HP;
#pragma unroll
    for (int q = 0; q < ROWS * BK / (2 * NTH); ++q) {
        s[rp * LDK + kc + (NTH / HP) * q] = r[q].x;
        s[(rp + 1) * LDK + kc + (NTH / HP) * q] = r[q].y;
    }
}

// op(A) is k-contiguous in memory when A is transposed; op(B) when B is not.
template <class C, bool TA, bool EDGE, bool VEC>
__device__ __forceinline__ void load_a(dv2 (&r)[C::PA], const Seg& s, int64_t m0, int64_t M, int64_t k0) {
    if (TA) load_kc<C::BM, C::NTH, EDGE, VEC>(r, s.a, s.lda, m0, M, k0, s.klen);
    else    load_rc<C::BM, C::NTH, EDGE, VEC>(r, s.a, s.lda, m0, M, k0, s.klen);
}
template <class C, bool TB, bool EDGE, bool VEC>
__device__ __forceinline__ void load_b(dv2 (&r)[C::PB], const Seg& s, int64_t n0, int64_t N, int64_t k0) {
    if (TB) load_rc<C::BN, C::NTH, EDGE, VEC>(r, s.b, s.ldb, n0, N, k0, s.klen);
    else    load_kc<C::BN, C::NTH, EDGE, VEC>(r, s.b, s.ldb, n0, N, k0, s.klen);
}
template <class C, bool TA>
__device__ __forceinline__ void store_a(double* s, const dv2 (&r)[C::PA]) {
    if (TA) store_kc<C::BM, C::NTH>(s, r); else store_rc<C::BM, C::NTH>(s, r);
}
template <class C, bool TB>
__device__ __forceinline__ void store_b(double* s, const dv2 (&r)[C::PB]) {
    if (TB) store_rc<C::BN, C::NTH>(s, r); else store_kc<C::BN, C::NTH>(s, r);
}

// Computes one output tile of the grouped problem set.
template <class C, bool TA, bool TB, bool VEC>
__device__ __forceinline__ void gemm_tile(const Prob* __restrict__ probs, int nprob, const Seg* __restrict__ segs,
                                          int tile, double (&As)[2][C::BM * LDK], double (&Bs)[2][C::BN * LDK]) {
    constexpr int BM = C::BM, BN = C::BN;
    constexpr int ci = BM == 256 ? 1 : 0;  // which tile range of the problem table

    // Problem owning this tile: last problem with tile0 <= tile.
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (probs[mid].tile0[ci] <= tile) lo = mid; else hi = mid - 1;
    }
    const Prob P = probs[lo];
    const int local = tile - P.tile0[ci];

    // Grouped raster: GROUP_M tile rows sweep the tile columns together.
    const int per_group = GROUP_M * P.tiles_n[ci];
    const int group = local / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(P.tiles_m[ci] - first_m, GROUP_M);
    const int in_group = local - group * per_group;
    const int tm = first_m + in_group % gsize;
    const int tn = in_group / gsize;

    const int64_t m0 = (int64_t)tm * BM;
    const int64_t n0 = (int64_t)tn * BN;
    const bool interior = (m0 + BM <= P.m) && (n0 + BN <= P.n);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = (wave / C::WAVES_N) * C::WM;
    const int wn = (wave % C::WAVES_N) * C::WN;
    const int lr = lane & 15;
    const int lg = lane >> 4;

    constexpr int FM = C::FM, FN = C::FN;
    dv4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = dv4{0.0, 0.0, 0.0, 0.0};

    int s = P.seg0;
    const int s_end = P.seg0 + P.nseg;
    Seg cur = segs[s];
    int64_t kk = 0;

    dv2 ra[C::PA], rb[C::PB];
    if (interior && BK <= cur.klen) {
        load_a<C, TA, false, VEC>(ra, cur, m0, P.m, 0);
        load_b<C, TB, false, VEC>(rb, cur, n0, P.n, 0);
    } else {
        load_a<C, TA, true, false>(ra, cur, m0, P.m, 0);
        load_b<C, TB, true, false>(rb, cur, n0, P.n, 0);
    }
    int buf = 0;
    store_a<C, TA>(As[0], ra);
    store_b<C, TB>(Bs[0], rb);
    __syncthreads();

    while (true) {
        // Next k-step position (possibly in the next segment).
        int64_t nk = kk + BK;
        int ns = s;
        Seg nxt = cur;
        if (nk >= cur.klen) {
            ++ns;
            nk = 0;
            if (ns < s_end) nxt = segs[ns];
        }
        const bool has_next = ns < s_end;
        if (has_next) {
            if (interior && nk + BK <= nxt.klen) {
                load_a<C, TA, false, VEC>(ra, nxt, m0, P.m, nk);
                load_b<C, TB, false, VEC>(rb, nxt, n0, P.n, nk);
            } else {
                load_a<C, TA, true, false>(ra, nxt, m0, P.m, nk);
                load_b<C, TB, true, false>(rb, nxt, n0, P.n, nk);
            }
        }

        const double* as = As[buf];
        const double* bs = Bs[buf];
        // Each half h of the k-step is one 8-deep step of the direct-to-LDS
        // kernels, summed in their order: lane group lg supplies k = 8h + 2lg
        // + q to MFMA q (one 128-bit LDS read per half), so every element
        // takes the same sequence of MFMAs in all three kernels and a
        // problem cut between them gives the same bits (but for the sign of
        // an exactly zero sum when K leaves this kernel a zero-padded half).
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            dv2 af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
                af[i] = *reinterpret_cast<const dv2*>(&as[(wm + 16 * i + lr) * LDK + 8 * h + 2 * lg]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                bf[j] = *reinterpret_cast<const dv2*>(&bs[(wn + 16 * j + lr) * LDK + 8 * h + 2 * lg]);
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j][q], af[i][q], acc[i][j], 0, 0, 0);
        }

        if (!has_next) break;
        buf ^= 1;
        store_a<C, TA>(As[buf], ra);
        store_b<C, TB>(Bs[buf], rb);
        __syncthreads();
        s = ns;
        kk = nk;
        cur = nxt;
    }

    // Epilogue.  Register v of lane group lg holds MFMA result row lg + 4v
    // (the f64 16x16x4 result layout), i.e.
    // acc[i][j][v] = sum for C(m0 + wm + 16i + lr, n0 + wn + 16j + lg + 4v).
    const double alpha = P.alpha, beta = P.beta;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int64_t m = m0 + wm + 16 * i + lr;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t n = n0 + wn + 16 * j + lg + 4 * v;
                if (interior || (m < P.m && n < P.n)) {
                    __attribute__((address_space(1))) double* cp =
                        (__attribute__((address_space(1))) double*)(P.c + m + n * P.ldc);
                    double val = alpha * acc[i][j][v];
                    if (beta != 0.0) val = fma(beta, *cp, val);
                    *cp = val;
                }
            }
        }
    }
}

// MODE 0: one workgroup per tile, tile = blockIdx.x (round-robin over XCDs);
// MODE 1: one workgroup per tile, XCD-contiguous tile ranges (xcd_remap);
// MODE 2: persistent: every workgroup takes tiles from `counter` (zeroed
//         before the launch) until the tile space is exhausted.
template <class C, bool TA, bool TB, bool VEC, int MODE>
__global__ __launch_bounds__(C::NTH, C::OCC) void gemm_f64_kernel(const Prob* __restrict__ probs, int nprob,
                                                                  const Seg* __restrict__ segs, int total_tiles,
                                                                  int* __restrict__ counter) {
    __shared__ double As[2][C::BM * LDK];
    __shared__ double Bs[2][C::BN * LDK];
    if (MODE == 0) {
        gemm_tile<C, TA, TB, VEC>(probs, nprob, segs, blockIdx.x, As, Bs);
    } else if (MODE == 1) {
        gemm_tile<C, TA, TB, VEC>(probs, nprob, segs, xcd_remap(blockIdx.x, total_tiles), As, Bs);
    } else {
        __shared__ int next;
        while (true) {
            __syncthreads();  // previous tile's LDS reads and `next` reads are complete
            if (threadIdx.x == 0) next = atomicAdd(counter, 1);
            __syncthreads();
            const int tile = next;
            if (tile >= total_tiles) break;
            gemm_tile<C, TA, TB, VEC>(probs, nprob, segs, tile, As, Bs);
        }
    }
}

// ---------------------------------------------------------------------------
// Direct-to-LDS form of the same GEMM for batches whose tiles are all
// interior (m, n multiples of 128, every segment's K a multiple of 8 and an
// even number of k-steps per problem, 16-byte aligned operands with even
// leading dimensions).  128x128 tiles, 4 waves of
// 64x64, two workgroups per CU, k-steps of 8 in a 3-stage LDS ring filled by
// global_load_lds_dwordx4 (no register staging).  Each operand is staged in
// the layout its memory order gives:
//  * contiguous along the tile's rows (A for opA = N, B for opB = T): one load
//    of a wave moves one k-row of 128 rows (1 KB); LDS [k][row] with a
//    1088-byte row stride, so the 64-bit fragment reads of four k-rows that
//    one MFMA needs fall in different bank halves;
//  * contiguous along k (A for opA = T, B for opB = N): one load moves 16
//    rows x 8 k (64 B each); LDS [row][k] with 16-byte chunk c of row r at
//    position c ^ ((r >> 1) & 3), the placement at which a wave's 128-bit
//    reads of 16 rows x 4 chunks run at the full LDS rate.
// Lane l of an MFMA supplies k = 2*(l/16) + s for sub-step s of a k-step:
// A and B use the same k assignment.  The fragments of step k+1 are read while
// the MFMAs of step k run, and the barrier runs one step ahead: after 24 of
// the 32 MFMAs of step k each wave waits for its own loads of step k+2, one
// barrier publishes them and frees the stage of step k (read completely during
// step k-1), which then receives step k+3.  Measured on a 16384^3 DGEMM: three
// stages beat four (fewer loads in flight) and a barrier inside the MFMA run
// beats one after it, together +2.2 %; five stages (unpadded, swizzled) lose 2 %.
namespace dmak {
constexpr int DBM = 128, DBN = 128, DBK = 8, DSTG = 3;
constexpr int RS = 1088;                   // k-row stride of a row-contiguous operand (bytes)
constexpr int RPART = DBK * RS;             // 8704 bytes
constexpr int KPART = 128 * DBK * 8;        // 8192 bytes
template <bool KC> constexpr int part() { return KC ? KPART : RPART; }
template <bool AK, bool BKC> constexpr int stage() { return part<AK>() + part<BKC>(); }
constexpr size_t lds_bytes(bool ak, bool bkc) { return (size_t)DSTG * ((ak ? KPART : RPART) + (bkc ? KPART : RPART)); }

__host__ __device__ constexpr int swz(int r) { return (r >> 1) & 3; }

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt field");
    // gfx9 s_waitcnt: vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt left at their maxima
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void gvoid_t;

// Zeros for the loads of k past a segment's end and of the padding step
// that makes a tile's step count even (never written).
__device__ __align__(256) double g_dma_zero[128];

// Per-lane source of one operand's loads within a segment: `p` points at
// this lane's first element of the current k-step, `second` is the offset of
// the wave's second load and `step` the advance per k-step (all in doubles).
struct OperandCursor {
    const double* p;
    int64_t second, step;
};

// Cursor at k = 0 of a segment: `base` is the segment's operand pointer,
// `r0` the tile's first row, `ld` the leading dimension, `rows` the
// operand's row count.
//  * k-contiguous operand: rows past the end are read from the last row
//    (their products land in rows or columns of C that are never stored).
//  * row-contiguous operand: a lane loads the 16-byte aligned row pair
//    (2i, 2i+1) of a grid that starts `shift` (0 or 1) rows before `base`;
//    rows past the end read the last pair.  Every such pair holds at least
//    one element of the operand, and a 16-byte aligned access never crosses
//    a page, so the load stays inside mapped memory; the row before the
//    first (shift 1) and the row after the last (odd row count) only feed
//    rows or columns of C that are never stored.
// EDGE = false: the tile is interior and shift is 0 (no clamping).
template <bool KC, bool EDGE>
__device__ __forceinline__ OperandCursor operand_cursor(const double* base, int64_t ld, int64_t r0, int64_t rows,
                                                        int shift, int wave, int lane) {
    if constexpr (KC) {
        // rows 32w .. 32w+31 in two loads of 16 rows; lane -> row lane/4, position lane%4
        const int row = lane >> 2;
        const int chunk = (lane & 3) ^ swz(row);
        const int64_t ra = EDGE ? min(r0 + 32 * wave + row, rows - 1) : r0 + 32 * wave + row;
        const int64_t rb = EDGE ? min(r0 + 32 * wave + row + 16, rows - 1) : ra + 16;
        return OperandCursor{base + ra * ld + 2 * chunk, (rb - ra) * ld, DBK};
    } else {
        // k-rows 2w, 2w+1 of 128 rows; lane -> rows 2*lane, 2*lane+1 of the pair grid
        if constexpr (!EDGE) return OperandCursor{base + r0 + 2 * lane + (int64_t)(2 * wave) * ld, ld, DBK * ld};
        const int64_t last = (rows + shift - 1) & ~(int64_t)1;  // first row of the last pair
        const int64_t r = min(r0 + 2 * lane, last);
        return OperandCursor{base - shift + r + (int64_t)(2 * wave) * ld, ld, DBK * ld};
    }
}

// Cursor past the last segment: every load reads zeros.
__device__ __forceinline__ OperandCursor zero_cursor(int lane) { return OperandCursor{g_dma_zero + 2 * lane, 0, 0}; }

// One k-step of one operand into `dst` (its part of a stage), then advance.
// With tail != 0 the step is a segment's last and only its first `tail`
// (even for k-contiguous operands) k are real: the lanes past them load zeros.
template <bool KC>
__device__ __forceinline__ void issue_operand(char* dst, OperandCursor& c, int wave, int lane, int tail) {
    char* d0 = KC ? dst + (32 * wave) * 64 : dst + (2 * wave) * RS;
    char* d1 = KC ? d0 + 16 * 64 : d0 + RS;
    const double* p0 = c.p;
    const double* p1 = c.p + c.second;
    if (tail) {
        const double* z = g_dma_zero + 2 * lane;
        if constexpr (KC) {
            // this lane's k pair is 2 * chunk, 2 * chunk + 1 of the step
            const int k2 = 2 * ((lane & 3) ^ swz(lane >> 2));
            if (k2 >= tail) { p0 = z; p1 = z; }
        } else {
            // load 0 carries k-row 2w, load 1 k-row 2w + 1
            if (2 * wave >= tail) p0 = z;
            if (2 * wave + 1 >= tail) p1 = z;
        }
    }
    __builtin_amdgcn_global_load_lds((gvoid_t*)p0, (lds_void_t*)d0, 16, 0, 0);
    __builtin_amdgcn_global_load_lds((gvoid_t*)p1, (lds_void_t*)d1, 16, 0, 0);
    c.p += c.step;
}

// Fragments of one operand for one k-step: f[i][s] = X(row w0 + 16i + lr, k = 2lg + s).
template <bool KC>
__device__ __forceinline__ void read_operand(double (&f)[4][2], const char* part, int w0, int lr, int lg) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if constexpr (KC) {
            const dv2 v = *reinterpret_cast<const dv2*>(part + (w0 + 16 * i + lr) * 64 + 16 * (lg ^ swz(lr)));
            f[i][0] = v.x;
            f[i][1] = v.y;
        } else {
#pragma unroll
            for (int s = 0; s < 2; ++s)
                f[i][s] = *reinterpret_cast<const double*>(part + (2 * lg + s) * RS + (w0 + 16 * i + lr) * 8);
        }
    }
}
}  // namespace dmak

// EDGE = false: every tile interior, no row shift, every segment's K a
// multiple of 8 and an even step count (the same arithmetic without the
// edge paths).
template <bool AK, bool BKC, bool EDGE>
__device__ __forceinline__ void dma_tile(const Prob* __restrict__ probs, int nprob, const Seg* __restrict__ segs,
                                         int tile, char* lds) {
    using dmak::DBM; using dmak::DBN; using dmak::DBK; using dmak::DSTG;
    using dmak::read_operand; using dmak::wait_vmcnt;
    constexpr int SSZ = dmak::stage<AK, BKC>(), APART = dmak::part<AK>();
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (probs[mid].tile0[0] <= tile) lo = mid; else hi = mid - 1;
    }
    const Prob P = probs[lo];
    const int local = tile - P.tile0[0];
    const int per_group = GROUP_M * P.tiles_n[0];
    const int group = local / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(P.tiles_m[0] - first_m, GROUP_M);
    const int in_group = local - group * per_group;
    const int64_t m0 = (int64_t)(first_m + in_group % gsize) * DBM;
    const int64_t n0 = (int64_t)(in_group / gsize) * DBN;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lr = lane & 15, lg = lane >> 4;

    // issue cursor over the problem's segments, one k-step (8) at a time; a
    // segment's last step may be partial, and past the last segment the
    // cursors read zeros (the padding step of an odd step count)
    const int sa = EDGE ? P.shift & 1 : 0, sb = EDGE ? (P.shift >> 1) & 1 : 0;  // row shifts of a row-contiguous A / B
    int iseg = P.seg0;
    Seg is = segs[iseg];
    int64_t left = (is.klen + DBK - 1) / DBK;  // k-steps left in the current segment
    int tail = (int)(is.klen % DBK);
    dmak::OperandCursor ca = dmak::operand_cursor<AK, EDGE>(is.a, is.lda, m0, P.m, sa, wave, lane);
    dmak::OperandCursor cb = dmak::operand_cursor<BKC, EDGE>(is.b, is.ldb, n0, P.n, sb, wave, lane);
    int ist = 0;
    auto issue_next = [&]() {
        char* st = lds + ist * SSZ;
        if (EDGE && left == 1 && tail) {
            dmak::issue_operand<AK>(st, ca, wave, lane, tail);
            dmak::issue_operand<BKC>(st + APART, cb, wave, lane, tail);
        } else {
            dmak::issue_operand<AK>(st, ca, wave, lane, 0);
            dmak::issue_operand<BKC>(st + APART, cb, wave, lane, 0);
        }
        ist = ist + 1 == DSTG ? 0 : ist + 1;
        if (--left == 0) {
            if (++iseg < P.seg0 + P.nseg) {
                is = segs[iseg];
                left = (is.klen + DBK - 1) / DBK;
                tail = (int)(is.klen % DBK);
                ca = dmak::operand_cursor<AK, EDGE>(is.a, is.lda, m0, P.m, sa, wave, lane);
                cb = dmak::operand_cursor<BKC, EDGE>(is.b, is.ldb, n0, P.n, sb, wave, lane);
            } else {
                left = INT64_MAX;
                tail = 0;
                ca = dmak::zero_cursor(lane);
                cb = dmak::zero_cursor(lane);
            }
        }
    };
    int64_t nk = 0;  // k-steps over all segments, rounded up to even (>= 2)
    for (int sg = P.seg0; sg < P.seg0 + P.nseg; ++sg) nk += (segs[sg].klen + DBK - 1) / DBK;
    nk += nk & 1;

    dv4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = dv4{0.0, 0.0, 0.0, 0.0};

    static_assert(DSTG == 3, "the waits below assume a 3-stage ring");
    for (int64_t s = 0; s < DSTG && s < nk; ++s) issue_next();
    // steps 0 and 1 published (4 loads per wave and step; step 2 may be in flight)
    if (nk >= 3) wait_vmcnt<4>(); else wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    int rst = 1;  // stage of step kt + 1
    double fa0[4][2], fb0[4][2], fa1[4][2], fb1[4][2];
    read_operand<AK>(fa0, lds, wm, lr, lg);
    read_operand<BKC>(fb0, lds + APART, wn, lr, lg);
    int64_t kt = 0;
    // one step on (CA, CB) while the next step's fragments are read into (NA, NB):
    // MFMAs f = 16q + 4i + j in order, the barrier after f = 23
#define DMA_MFMAS(CA, CB, F0, F1)                                                                   \
    _Pragma("unroll") for (int f_ = F0; f_ < F1; ++f_) {                                            \
        const int q_ = f_ >> 4, i_ = (f_ >> 2) & 3, j_ = f_ & 3;                                    \
        acc[i_][j_] = __builtin_amdgcn_mfma_f64_16x16x4f64(CB[j_][q_], CA[i_][q_], acc[i_][j_], 0, 0, 0); \
    }
// FULL: kt + 3 < nk is known (the barrier and the issue of step kt+3 both run)
#define DMA_STEP(FULL, CA, CB, NA, NB)                                                                    \
    {                                                                                               \
        const char* nst = lds + rst * SSZ; /* harmless past the last step */                        \
        read_operand<AK>(NA, nst, wm, lr, lg);                                                      \
        DMA_MFMAS(CA, CB, 0, 24)                                                                    \
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                          \
        _Pragma("unroll") for (int g_ = 0; g_ < 4; ++g_) {                                          \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                      \
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                      \
        }                                                                                           \
        __builtin_amdgcn_sched_group_barrier(0x008, 14, 0);                                         \
        rst = rst + 1 == DSTG ? 0 : rst + 1;                                                        \
        if (FULL || kt + 2 < nk) {                                                                          \
            /* the stage of step kt was read during step kt-1; LDS reads retire in order, so once  \
               at most the 4 reads just issued are pending, none of it is still being read */       \
            __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4) | (4 << 8));                  \
            wait_vmcnt<0>(); /* step kt+2, the only loads in flight */                              \
            __builtin_amdgcn_s_barrier();                                                           \
            if (FULL || kt + DSTG < nk) issue_next();                                                        \
        }                                                                                           \
        read_operand<BKC>(NB, nst + APART, wn, lr, lg);                                             \
        DMA_MFMAS(CA, CB, 24, 32)                                                                   \
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                          \
        _Pragma("unroll") for (int g_ = 0; g_ < 3; ++g_) {                                          \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                      \
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                      \
        }                                                                                           \
        ++kt;                                                                                       \
    }
    // Steady turns of twelve steps without conditions, then the last steps
    // two at a time.  The compiler drains every load at a loop head (it
    // cannot count across the direct-to-LDS loads there), which with a
    // two-step loop was a wait for step kt+2's loads every other step:
    // twelve-step turns run 5.5 % faster at 16384^3.
    while (kt + 15 <= nk) {
        DMA_STEP(true, fa0, fb0, fa1, fb1);
        DMA_STEP(true, fa1, fb1, fa0, fb0);
        DMA_STEP(true, fa0, fb0, fa1, fb1);
        DMA_STEP(true, fa1, fb1, fa0, fb0);
        DMA_STEP(true, fa0, fb0, fa1, fb1);
        DMA_STEP(true, fa1, fb1, fa0, fb0);
        DMA_STEP(true, fa0, fb0, fa1, fb1);
        DMA_STEP(true, fa1, fb1, fa0, fb0);
        DMA_STEP(true, fa0, fb0, fa1, fb1);
        DMA_STEP(true, fa1, fb1, fa0, fb0);
        DMA_STEP(true, fa0, fb0, fa1, fb1);
        DMA_STEP(true, fa1, fb1, fa0, fb0);
    }
    while (kt < nk) {
        DMA_STEP(false, fa0, fb0, fa1, fb1);
        DMA_STEP(false, fa1, fb1, fa0, fb0);
    }
#undef DMA_STEP
#undef DMA_MFMAS

    // acc[i][j][v] = sum for C(m0 + wm + 16i + lr - sa, n0 + wn + 16j + lg + 4v - sb)
    const double alpha = P.alpha, beta = P.beta;
    const bool interior = !EDGE || (sa == 0 && sb == 0 && m0 + DBM <= P.m && n0 + DBN <= P.n);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t m = m0 + wm + 16 * i + lr - sa;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t n = n0 + wn + 16 * j + lg + 4 * v - sb;
                if (!interior && (m < 0 || m >= P.m || n < 0 || n >= P.n)) continue;
                __attribute__((address_space(1))) double* cp =
                    (__attribute__((address_space(1))) double*)(P.c + m + n * P.ldc);
                double val = alpha * acc[i][j][v];
                if (beta != 0.0) val = fma(beta, *cp, val);
                *cp = val;
            }
        }
    }
}

// MODE as gemm_f64_kernel: 0 tile = blockIdx.x, 1 XCD-contiguous ranges,
// 2 persistent (tiles taken from `counter`, zeroed before the launch).
template <bool AK, bool BKC, bool EDGE, int MODE>
__global__ __launch_bounds__(256, 2) void gemm_f64_dma_kernel(const Prob* __restrict__ probs, int nprob,
                                                               const Seg* __restrict__ segs, int total_tiles,
                                                               int* __restrict__ counter) {
    extern __shared__ __attribute__((aligned(1024))) char lds[];
    if (MODE == 0) {
        dma_tile<AK, BKC, EDGE>(probs, nprob, segs, blockIdx.x, lds);
    } else if (MODE == 1) {
        dma_tile<AK, BKC, EDGE>(probs, nprob, segs, xcd_remap(blockIdx.x, total_tiles), lds);
    } else {
        __shared__ int next;
        while (true) {
            __syncthreads();  // the previous tile's LDS reads and `next` reads are complete
            if (threadIdx.x == 0) next = atomicAdd(counter, 1);
            __syncthreads();
            const int tile = next;
            if (tile >= total_tiles) break;
            dma_tile<AK, BKC, EDGE>(probs, nprob, segs, tile, lds);
        }
    }
}

// A in registers (interior batches): each wave owns 32 rows of the 128x128
// tile and all 128 columns (2 x 8 MFMA tiles of 16x16), so no two waves use
// the same element of A, and A needs neither LDS nor a barrier.  Lane
// (lr, lg) holds rows 2lr, 2lr+1 of its wave's 32 at k = 2lg + q (q = 0, 1):
// MFMA tile i = 0 takes the even rows, i = 1 the odd ones.  Two 16-byte
// global loads per k-step: with A row-contiguous (opA = N) one per k (the
// row pair), with A k-contiguous (opA = T) one per row (the k pair).  A runs three k-steps ahead in a ring of registers
// (loaded after the last MFMA that reads the slot); B goes through the
// 3-stage direct-to-LDS ring of dma_tile, whose fragments of step k+1 are read
// in two halves around the barrier.  Loads in issue order: B(k+3) at the
// barrier of step k, A(k+3) after its MFMAs, so the barrier's wait for B(k+2)
// leaves the two loads of A(k+2) in flight.
template <bool AK, bool BKC>
__device__ __forceinline__ void dma_tile_ar(const Prob* __restrict__ probs, int nprob, const Seg* __restrict__ segs,
                                            int tile, char* lds) {
    using dmak::DBM; using dmak::DBN; using dmak::DBK; using dmak::DSTG;
    using dmak::read_operand; using dmak::wait_vmcnt;
    constexpr int SSZ = dmak::part<BKC>();
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (probs[mid].tile0[0] <= tile) lo = mid; else hi = mid - 1;
    }
    const Prob P = probs[lo];
    const int local = tile - P.tile0[0];
    const int per_group = GROUP_M * P.tiles_n[0];
    const int group = local / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(P.tiles_m[0] - first_m, GROUP_M);
    const int in_group = local - group * per_group;
    const int64_t m0 = (int64_t)(first_m + in_group % gsize) * DBM;
    const int64_t n0 = (int64_t)(in_group / gsize) * DBN;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 15, lg = lane >> 4;
    const int64_t arow = m0 + 32 * wave + 2 * lr;

    // issue cursors: every segment a whole number of k-steps, none past the end
    int iseg = P.seg0;
    Seg is = segs[iseg];
    int64_t left = is.klen / DBK;
    // A at (arow, k = 2lg) of the next step to issue; the second load is one ld away
    auto a_origin = [&](const Seg& g) { return AK ? g.a + arow * g.lda + 2 * lg : g.a + arow + (int64_t)(2 * lg) * g.lda; };
    const double* ap = a_origin(is);
    int64_t alda = is.lda;
    dmak::OperandCursor cb = dmak::operand_cursor<BKC, false>(is.b, is.ldb, n0, P.n, 0, wave, lane);
    const double* apend = ap;  // A of the step issued last
    int64_t apld = alda;
    int ist = 0;
    auto issue_next = [&]() {
        dmak::issue_operand<BKC>(lds + ist * SSZ, cb, wave, lane, 0);
        ist = ist + 1 == DSTG ? 0 : ist + 1;
        apend = ap;
        apld = alda;
        ap += AK ? DBK : DBK * alda;
        if (--left == 0 && ++iseg < P.seg0 + P.nseg) {
            is = segs[iseg];
            left = is.klen / DBK;
            ap = a_origin(is);
            alda = is.lda;
            cb = dmak::operand_cursor<BKC, false>(is.b, is.ldb, n0, P.n, 0, wave, lane);
        }
    };
    // ra[slot][h] of the slot's step: opA = N, h = q: A(rows 2lr, 2lr + 1; k = 2lg + q);
    // opA = T, h = i: A(row 2lr + i; k = 2lg, 2lg + 1)
    dv2 ra[3][2];
    auto load_a = [&](dv2 (&r)[2]) {
        r[0] = gld2(apend);
        r[1] = gld2(apend + apld);
    };
    int64_t nk = 0;  // even (an interior batch)
    for (int sg = P.seg0; sg < P.seg0 + P.nseg; ++sg) nk += segs[sg].klen / DBK;

    dv4 acc[2][8];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = dv4{0.0, 0.0, 0.0, 0.0};

    static_assert(DSTG == 3, "the waits below assume a 3-stage ring");
    issue_next();
    load_a(ra[0]);
    issue_next();
    load_a(ra[1]);
    if (nk > 2) {
        issue_next();
        load_a(ra[2]);
        wait_vmcnt<6>();  // B of steps 0 and 1 landed; A(1), B(2), A(2) may be in flight
    } else {
        wait_vmcnt<0>();
    }
    __builtin_amdgcn_s_barrier();
    int rst = 1;  // stage of step kt + 1
    double fb0[2][4][2], fb1[2][4][2];  // f[h][jj][q] = B(row 64h + 16jj + lr, k = 2lg + q)
    read_operand<BKC>(fb0[0], lds, 0, lr, lg);
    read_operand<BKC>(fb0[1], lds, 64, lr, lg);
    int64_t kt = 0;
    // MFMAs f = 16q + 8i + j in order, the barrier after f = 23
#define AR_MFMAS(SLOT, CB, F0, F1)                                                                  \
    _Pragma("unroll") for (int f_ = F0; f_ < F1; ++f_) {                                            \
        const int q_ = f_ >> 4, i_ = (f_ >> 3) & 1, j_ = f_ & 7;                                    \
        acc[i_][j_] = __builtin_amdgcn_mfma_f64_16x16x4f64(CB[j_ >> 2][j_ & 3][q_],                \
                                                           AK ? ra[SLOT][i_][q_] : ra[SLOT][q_][i_], acc[i_][j_], 0, 0, 0); \
    }
// FULL: kt + 3 < nk is known (the barrier, the issue of B(kt+3) and the load of A(kt+3) all run)
#define AR_STEP(FULL, SLOT, CB, NB)                                                                 \
    {                                                                                               \
        const char* nst = lds + rst * SSZ; /* harmless past the last step */                        \
        read_operand<BKC>(NB[0], nst, 0, lr, lg);                                                   \
        AR_MFMAS(SLOT, CB, 0, 24)                                                                   \
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                          \
        _Pragma("unroll") for (int g_ = 0; g_ < 4; ++g_) {                                          \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                      \
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                      \
        }                                                                                           \
        __builtin_amdgcn_sched_group_barrier(0x008, 14, 0);                                         \
        rst = rst + 1 == DSTG ? 0 : rst + 1;                                                        \
        if (FULL || kt + 2 < nk) {                                                                  \
            /* the stage of step kt was read during step kt-1: once at most the 4 reads just        \
               issued are pending, none of it is still being read */                                \
            __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4) | (4 << 8));                  \
            wait_vmcnt<2>(); /* B(kt+2); A(kt+2), issued after it, may be in flight */              \
            __builtin_amdgcn_s_barrier();                                                           \
            if (FULL || kt + DSTG < nk) issue_next();                                                \
        }                                                                                           \
        read_operand<BKC>(NB[1], nst, 64, lr, lg);                                                  \
        AR_MFMAS(SLOT, CB, 24, 32)                                                                  \
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                          \
        _Pragma("unroll") for (int g_ = 0; g_ < 3; ++g_) {                                          \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                      \
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);                                      \
        }                                                                                           \
        if (FULL || kt + DSTG < nk) load_a(ra[SLOT]);                                               \
        ++kt;                                                                                       \
    }
    // A slot = step % 3, B fragments alternate: six steps make a cycle, and a
    // steady turn runs four of them.  The steady turns have no conditions (so
    // the waits the compiler derives inside them stay partial); the compiler
    // drains every load at the loop head, once per turn (at 16384^3 twelve-step
    // turns beat six-step ones by 0.3 % on NN and 2 % on TN, 24-step turns
    // twelve-step ones by 0.2 %); the last 2 to 26 steps, nk even, follow.
    while (kt + 27 <= nk) {
        AR_STEP(true, 0, fb0, fb1);
        AR_STEP(true, 1, fb1, fb0);
        AR_STEP(true, 2, fb0, fb1);
        AR_STEP(true, 0, fb1, fb0);
        AR_STEP(true, 1, fb0, fb1);
        AR_STEP(true, 2, fb1, fb0);
        AR_STEP(true, 0, fb0, fb1);
        AR_STEP(true, 1, fb1, fb0);
        AR_STEP(true, 2, fb0, fb1);
        AR_STEP(true, 0, fb1, fb0);
        AR_STEP(true, 1, fb0, fb1);
        AR_STEP(true, 2, fb1, fb0);
        AR_STEP(true, 0, fb0, fb1);
        AR_STEP(true, 1, fb1, fb0);
        AR_STEP(true, 2, fb0, fb1);
        AR_STEP(true, 0, fb1, fb0);
        AR_STEP(true, 1, fb0, fb1);
        AR_STEP(true, 2, fb1, fb0);
        AR_STEP(true, 0, fb0, fb1);
        AR_STEP(true, 1, fb1, fb0);
        AR_STEP(true, 2, fb0, fb1);
        AR_STEP(true, 0, fb1, fb0);
        AR_STEP(true, 1, fb0, fb1);
        AR_STEP(true, 2, fb1, fb0);
    }
    while (true) {
        AR_STEP(false, 0, fb0, fb1);
        AR_STEP(false, 1, fb1, fb0);
        if (kt >= nk) break;
        AR_STEP(false, 2, fb0, fb1);
        AR_STEP(false, 0, fb1, fb0);
        if (kt >= nk) break;
        AR_STEP(false, 1, fb0, fb1);
        AR_STEP(false, 2, fb1, fb0);
        if (kt >= nk) break;
    }
#undef AR_STEP
#undef AR_MFMAS

    // acc[i][j][v] = sum for C(arow + i, n0 + 16j + lg + 4v)
    const double alpha = P.alpha, beta = P.beta;
    typedef __attribute__((address_space(1))) double gd;
    typedef __attribute__((address_space(1))) dv2 gd2;
    const bool vec = ((reinterpret_cast<uintptr_t>(P.c) & 15) == 0) && !(P.ldc & 1);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t n = n0 + 16 * j + lg + 4 * v;
            double* cp = P.c + arow + n * P.ldc;
            dv2 val = dv2{alpha * acc[0][j][v], alpha * acc[1][j][v]};
            if (vec) {
                if (beta != 0.0) {
                    const dv2 c = *(gd2*)cp;
                    val = dv2{fma(beta, c.x, val.x), fma(beta, c.y, val.y)};
                }
                *(gd2*)cp = val;
            } else {
                if (beta != 0.0) val = dv2{fma(beta, ((gd*)cp)[0], val.x), fma(beta, ((gd*)cp)[1], val.y)};
                ((gd*)cp)[0] = val.x;
                ((gd*)cp)[1] = val.y;
            }
        }
    }
}

template <bool AK, bool BKC, int MODE>
__global__ __launch_bounds__(256, 2) void gemm_f64_dma_ar_kernel(const Prob* __restrict__ probs, int nprob,
                                                                  const Seg* __restrict__ segs, int total_tiles,
                                                                  int* __restrict__ counter) {
    extern __shared__ __attribute__((aligned(1024))) char lds[];
    if (MODE == 0) {
        dma_tile_ar<AK, BKC>(probs, nprob, segs, blockIdx.x, lds);
    } else if (MODE == 1) {
        dma_tile_ar<AK, BKC>(probs, nprob, segs, xcd_remap(blockIdx.x, total_tiles), lds);
    } else {
        __shared__ int next;
        while (true) {
            __syncthreads();  // the previous tile's LDS reads and `next` reads are complete
            if (threadIdx.x == 0) next = atomicAdd(counter, 1);
            __syncthreads();
            const int tile = next;
            if (tile >= total_tiles) break;
            dma_tile_ar<AK, BKC>(probs, nprob, segs, tile, lds);
        }
    }
}

template <bool AK, bool BKC>
void launch_dma_ar(int mode, const Prob* dp, int np, const Seg* ds, int tiles, int grid, int* counter, hipStream_t st) {
    const size_t lb = (size_t)dmak::DSTG * dmak::part<BKC>();
    if (mode == DG_GEMM_MODE_ROUND_ROBIN)
        hipLaunchKernelGGL((gemm_f64_dma_ar_kernel<AK, BKC, 0>), dim3(tiles), dim3(256), lb, st, dp, np, ds, tiles, counter);
    else if (mode == DG_GEMM_MODE_XCD)
        hipLaunchKernelGGL((gemm_f64_dma_ar_kernel<AK, BKC, 1>), dim3(tiles), dim3(256), lb, st, dp, np, ds, tiles, counter);
    else
        hipLaunchKernelGGL((gemm_f64_dma_ar_kernel<AK, BKC, 2>), dim3(grid), dim3(256), lb, st, dp, np, ds, tiles, counter);
}

template <bool AK, bool BKC, bool EDGE>
void launch_dma_e(int mode, const Prob* dp, int np, const Seg* ds, int tiles, int grid, int* counter, hipStream_t st) {
    const size_t lb = dmak::lds_bytes(AK, BKC);
    if (mode == DG_GEMM_MODE_ROUND_ROBIN)
        hipLaunchKernelGGL((gemm_f64_dma_kernel<AK, BKC, EDGE, 0>), dim3(tiles), dim3(256), lb, st, dp, np, ds, tiles, counter);
    else if (mode == DG_GEMM_MODE_XCD)
        hipLaunchKernelGGL((gemm_f64_dma_kernel<AK, BKC, EDGE, 1>), dim3(tiles), dim3(256), lb, st, dp, np, ds, tiles, counter);
    else
        hipLaunchKernelGGL((gemm_f64_dma_kernel<AK, BKC, EDGE, 2>), dim3(grid), dim3(256), lb, st, dp, np, ds, tiles, counter);
}
template <bool AK, bool BKC>
void launch_dma(bool edge, int mode, const Prob* dp, int np, const Seg* ds, int tiles, int grid, int* counter,
                hipStream_t st) {
    static const bool no_areg = getenv("DG_GEMM_NO_AREG") != nullptr;
    if (!edge && !no_areg) return launch_dma_ar<AK, BKC>(mode, dp, np, ds, tiles, grid, counter, st);
    if (edge) launch_dma_e<AK, BKC, true>(mode, dp, np, ds, tiles, grid, counter, st);
    else launch_dma_e<AK, BKC, false>(mode, dp, np, ds, tiles, grid, counter, st);
}

__global__ void zero_counter_kernel(int* counter) { *counter = 0; }

__global__ void scale_kernel(int64_t m, int64_t n, double beta, double* __restrict__ C, int64_t ldc) {
    const int64_t total = m * n;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < total;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = x / m, row = x - col * m;
        double* p = C + row + col * ldc;
        *p = (beta == 0.0) ? 0.0 : beta * *p;
    }
}

template <class C, bool TA, bool TB, bool VEC>
void launch_mode(int mode, const Prob* dp, int np, const Seg* ds, int tiles, int grid, int* counter,
                 hipStream_t st) {
    if (mode == DG_GEMM_MODE_ROUND_ROBIN)
        hipLaunchKernelGGL((gemm_f64_kernel<C, TA, TB, VEC, 0>), dim3(tiles), dim3(C::NTH), 0, st, dp, np, ds, tiles,
                           counter);
    else if (mode == DG_GEMM_MODE_XCD)
        hipLaunchKernelGGL((gemm_f64_kernel<C, TA, TB, VEC, 1>), dim3(tiles), dim3(C::NTH), 0, st, dp, np, ds, tiles,
                           counter);
    else
        hipLaunchKernelGGL((gemm_f64_kernel<C, TA, TB, VEC, 2>), dim3(grid), dim3(C::NTH), 0, st, dp, np, ds, tiles,
                           counter);
}

template <class C, bool TA, bool TB>
void launch_kind(bool vec, int mode, const Prob* dp, int np, const Seg* ds, int tiles, int grid, int* counter,
                 hipStream_t st) {
    if (vec) launch_mode<C, TA, TB, true>(mode, dp, np, ds, tiles, grid, counter, st);
    else     launch_mode<C, TA, TB, false>(mode, dp, np, ds, tiles, grid, counter, st);
}

template <class C>
void launch_cfg(char opA, char opB, bool vec, int mode, const Prob* dp, int np, const Seg* ds, int tiles, int grid,
                int* counter, hipStream_t st) {
    const bool ta = opA == 'T', tb = opB == 'T';
    if (!ta && !tb) launch_kind<C, false, false>(vec, mode, dp, np, ds, tiles, grid, counter, st);
    else if (!ta && tb) launch_kind<C, false, true>(vec, mode, dp, np, ds, tiles, grid, counter, st);
    else if (ta && !tb) launch_kind<C, true, false>(vec, mode, dp, np, ds, tiles, grid, counter, st);
    else launch_kind<C, true, true>(vec, mode, dp, np, ds, tiles, grid, counter, st);
}

}  // namespace

int launch_scale(int64_t m, int64_t n, double beta, double* C, int64_t ldc, hipStream_t stream) {
    if (m <= 0 || n <= 0) return DG_OK;
    int64_t total = m * n;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, m, n, beta, C, ldc);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

int prepare_gemm(char opA, char opB, const std::vector<Prob>& probs_in, const std::vector<Seg>& segs,
                 GemmBatch& out, bool split_edges) {
    out = GemmBatch{};
    out.segs = segs;
    const bool ak = opA == 'T', bkc = opB != 'T';  // k-contiguous operands
    static const bool no_dma = getenv("DG_GEMM_NO_DMA") != nullptr;
    int64_t tdma = 0, tdmae = 0, tgen[2] = {0, 0};
    const int bm[2] = {Cfg128::BM, Cfg256::BM}, bn[2] = {Cfg128::BN, Cfg256::BN};
    constexpr int64_t TM = dmak::DBM, TN = dmak::DBN;
    // a direct-to-LDS problem (or sub-problem) onto its list
    auto add_dma = [&](Prob p, bool interior) {
        const int sa = p.shift & 1, sb = (p.shift >> 1) & 1;
        p.tiles_m[0] = p.tiles_m[1] = (int32_t)((p.m + sa + TM - 1) / TM);
        p.tiles_n[0] = p.tiles_n[1] = (int32_t)((p.n + sb + TN - 1) / TN);
        int64_t& t = interior ? tdma : tdmae;
        p.tile0[0] = p.tile0[1] = (int32_t)t;
        t += (int64_t)p.tiles_m[0] * p.tiles_n[0];
        (interior ? out.dma : out.dmae).push_back(p);
    };
    // rows [r0, r0 + nr) x columns [c0, c0 + nc) of p as a problem of its own,
    // its segments re-based and appended; its shifts follow from the new bases
    auto sub = [&](const Prob& p, int64_t r0, int64_t nr, int64_t c0, int64_t nc) {
        Prob q = p;
        q.m = nr;
        q.n = nc;
        q.c = p.c + r0 + c0 * p.ldc;
        q.seg0 = (int32_t)out.segs.size();
        for (int s = p.seg0; s < p.seg0 + p.nseg; ++s) {
            Seg g = segs[s];
            g.a += ak ? r0 * g.lda : r0;
            g.b += bkc ? c0 * g.ldb : c0;
            out.segs.push_back(g);
        }
        const Seg& g0 = out.segs[q.seg0];
        const int sa = !ak && (reinterpret_cast<uintptr_t>(g0.a) & 15) == 8;
        const int sb = !bkc && (reinterpret_cast<uintptr_t>(g0.b) & 15) == 8;
        q.shift = sa | (sb << 1);
        return q;
    };
    // a strip narrower than a tile onto narrow list w (0: 128x32 tiles, 1: 64x128)
    static const bool no_narrow = getenv("DG_GEMM_NO_NARROW") != nullptr;
    const int nbm[2] = {CfgN32::BM, CfgM64::BM}, nbn[2] = {CfgN32::BN, CfgM64::BN};
    int64_t tnar[2] = {0, 0};
    auto add_nar = [&](Prob q, int w) {
        for (int s = q.seg0; s < q.seg0 + q.nseg; ++s) {
            const Seg& g = out.segs[s];
            if ((reinterpret_cast<uintptr_t>(g.a) & 15) || (reinterpret_cast<uintptr_t>(g.b) & 15) || (g.lda & 1) ||
                (g.ldb & 1))
                out.vec_nar[w] = false;
        }
        q.tiles_m[0] = q.tiles_m[1] = (int32_t)((q.m + nbm[w] - 1) / nbm[w]);
        q.tiles_n[0] = q.tiles_n[1] = (int32_t)((q.n + nbn[w] - 1) / nbn[w]);
        q.tile0[0] = q.tile0[1] = (int32_t)tnar[w];
        tnar[w] += (int64_t)q.tiles_m[0] * q.tiles_n[0];
        out.nar[w].push_back(q);
    };
    // a strip of `width` rows (columns) fills 128-wide tiles to width / 128,
    // a 64-row (32-column) tile to width / 64 (width / (32 ceil(width / 32))),
    // and the narrow tiles pay up to this width
    auto narrow = [&](int64_t width) { return !no_narrow && width <= kNarrowMax; };
    for (const Prob& p0 : probs_in) {
        if (p0.m <= 0 || p0.n <= 0 || p0.nseg <= 0) continue;
        bool v = true, d = !no_dma;
        int off_a = -1, off_b = -1;  // byte offset within 16 of a row-contiguous A / B
        for (int s = p0.seg0; s < p0.seg0 + p0.nseg; ++s) {
            const Seg& g = segs[s];
            const int oa = (int)(reinterpret_cast<uintptr_t>(g.a) & 15), ob = (int)(reinterpret_cast<uintptr_t>(g.b) & 15);
            if (oa || ob || (g.lda & 1) || (g.ldb & 1)) v = false;
            if ((g.lda & 1) || (g.ldb & 1) || g.klen <= 0) d = false;
            if (ak) { if (oa || (g.klen & 1)) d = false; }   // k pairs, 16-byte aligned
            else { if ((oa & 7) || (off_a >= 0 && oa != off_a)) d = false; off_a = oa; }
            if (bkc) { if (ob || (g.klen & 1)) d = false; }
            else { if ((ob & 7) || (off_b >= 0 && ob != off_b)) d = false; off_b = ob; }
        }
        Prob p = p0;
        if (!d) {
            out.vec = out.vec && v;
            for (int c = 0; c < 2; ++c) {
                p.tiles_m[c] = (int32_t)((p.m + bm[c] - 1) / bm[c]);
                p.tiles_n[c] = (int32_t)((p.n + bn[c] - 1) / bn[c]);
                p.tile0[c] = (int32_t)tgen[c];
                tgen[c] += (int64_t)p.tiles_m[c] * p.tiles_n[c];
            }
            out.gen.push_back(p);
            continue;
        }
        const int sa = off_a == 8, sb = off_b == 8;
        p.shift = sa | (sb << 1);
        int64_t steps = 0;
        bool whole = true;  // every segment a whole number of k-steps
        for (int s = p.seg0; s < p.seg0 + p.nseg; ++s) {
            steps += (segs[s].klen + dmak::DBK - 1) / dmak::DBK;
            if (segs[s].klen % dmak::DBK) whole = false;
        }
        const bool k_ok = whole && !(steps & 1);
        if (!p.shift && p.m % TM == 0 && p.n % TN == 0 && k_ok) {
            add_dma(p, true);
            continue;
        }
        // A shifted row-contiguous operand is aligned from row (column) 127
        // on: the strip before it is one tile row (column) of the shifted
        // tiling, so cutting there adds no tiles.
        const int64_t r0 = sa ? TM - 1 : 0, c0 = sb ? TN - 1 : 0;
        const int64_t mi = p.m > r0 ? (p.m - r0) / TM * TM : 0, ni = p.n > c0 ? (p.n - c0) / TN * TN : 0;
        if (!split_edges || !k_ok || mi * ni < 8 * TM * TN) {
            add_dma(p, false);
            continue;
        }
        add_dma(sub(p, r0, mi, c0, ni), true);                                       // interior
        if (r0 > 0) add_dma(sub(p, 0, r0, 0, p.n), false);                           // rows before it
        if (r0 + mi < p.m) {                                                         // rows after it
            const Prob q = sub(p, r0 + mi, p.m - r0 - mi, 0, p.n);
            if (narrow(q.m)) add_nar(q, 1); else add_dma(q, false);
        }
        if (c0 > 0) add_dma(sub(p, r0, mi, 0, c0), false);                           // columns before it
        if (c0 + ni < p.n) {                                                         // columns after it
            const Prob q = sub(p, r0, mi, c0 + ni, p.n - c0 - ni);
            if (narrow(q.n)) add_nar(q, 0); else add_dma(q, false);
        }
    }
    if (tdma > INT32_MAX || tdmae > INT32_MAX || tgen[0] > INT32_MAX || tgen[1] > INT32_MAX ||
        tnar[0] > INT32_MAX || tnar[1] > INT32_MAX || out.segs.size() > (size_t)INT32_MAX)
        return fail(DG_ERR_INVALID, "gemm: output tiles exceed one launch");
    out.tiles_nar[0] = (int)tnar[0];
    out.tiles_nar[1] = (int)tnar[1];
    out.tiles_dma = (int)tdma;
    out.tiles_dmae = (int)tdmae;
    out.tiles_gen[0] = (int)tgen[0];
    out.tiles_gen[1] = (int)tgen[1];
    return DG_OK;
}

int persistent_grid(int cfg) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return cus * (cfg == DG_GEMM_TILE_256x128 ? Cfg256::OCC : cfg == DG_GEMM_TILE_128x128W8 ? Cfg128w8::OCC : Cfg128::OCC);
}

int launch_gemm_f64_dev(char opA, char opB, int mode, int cfg, const GemmBatchDev& b, int* counter,
                        hipStream_t stream) {
    if (mode == DG_GEMM_MODE_PERSISTENT && !counter)
        return fail(DG_ERR_INVALID, "gemm: persistent launch without a work counter");
    for (int e = 0; e < 2; ++e) {  // interior list, then edge list
        const bool edge = e == 1;
        const int t0 = edge ? b.tiles_dmae : b.tiles_dma;
        if (t0 <= 0) continue;
        const Prob* dp = edge ? b.dmae : b.dma;
        const int np = edge ? b.ndmae : b.ndma;
        int grid = 0;
        if (mode == DG_GEMM_MODE_PERSISTENT) {
            grid = persistent_grid(DG_GEMM_TILE_128x128);
            if (grid > t0) grid = t0;
            hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
        }
        const bool ak = opA == 'T', bkc = opB != 'T';
        if (ak && bkc) launch_dma<true, true>(edge, mode, dp, np, b.segs, t0, grid, counter, stream);
        else if (ak) launch_dma<true, false>(edge, mode, dp, np, b.segs, t0, grid, counter, stream);
        else if (bkc) launch_dma<false, true>(edge, mode, dp, np, b.segs, t0, grid, counter, stream);
        else launch_dma<false, false>(edge, mode, dp, np, b.segs, t0, grid, counter, stream);
        DG_LAUNCH_CHECK();
    }
    for (int w = 0; w < 2; ++w) {  // narrow strips: 128x32 tiles, then 64x128 tiles
        const int tw = b.tiles_nar[w];
        if (tw <= 0) continue;
        int grid = 0;
        if (mode == DG_GEMM_MODE_PERSISTENT) {
            grid = persistent_grid(DG_GEMM_TILE_128x128);  // two workgroups per CU, as CfgN32 / CfgM64
            if (grid > tw) grid = tw;
            hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
        }
        if (w == 0) launch_cfg<CfgN32>(opA, opB, b.vec_nar[0], mode, b.nar[0], b.nnar[0], b.segs, tw, grid, counter, stream);
        else launch_cfg<CfgM64>(opA, opB, b.vec_nar[1], mode, b.nar[1], b.nnar[1], b.segs, tw, grid, counter, stream);
        DG_LAUNCH_CHECK();
    }
    const int t = b.tiles_gen[cfg == DG_GEMM_TILE_256x128 ? 1 : 0];
    if (t <= 0) return DG_OK;
    int grid = 0;
    if (mode == DG_GEMM_MODE_PERSISTENT) {
        grid = persistent_grid(cfg);
        if (grid > t) grid = t;
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
    }
    if (cfg == DG_GEMM_TILE_256x128) launch_cfg<Cfg256>(opA, opB, b.vec, mode, b.gen, b.ngen, b.segs, t, grid, counter, stream);
    else if (cfg == DG_GEMM_TILE_128x128W8) launch_cfg<Cfg128w8>(opA, opB, b.vec, mode, b.gen, b.ngen, b.segs, t, grid, counter, stream);
    else launch_cfg<Cfg128>(opA, opB, b.vec, mode, b.gen, b.ngen, b.segs, t, grid, counter, stream);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

int default_tile_cfg() {
    static const int cfg = [] {
        const char* e = getenv("DG_GEMM_TILE");
        if (e && std::string(e) == "256x128") return (int)DG_GEMM_TILE_256x128;
        if (e && std::string(e) == "128x128") return (int)DG_GEMM_TILE_128x128;
        if (e && std::string(e) == "128x128w8") return (int)DG_GEMM_TILE_128x128W8;
        return (int)DG_GEMM_TILE_DEFAULT;
    }();
    return cfg;
}

int launch_gemm_f64(char opA, char opB, const std::vector<Prob>& probs_in, const std::vector<Seg>& segs,
                    hipStream_t stream, int mode) {
    GemmBatch hb;
    int rc = prepare_gemm(opA, opB, probs_in, segs, hb);
    if (rc) return rc;
    if (hb.dma.empty() && hb.dmae.empty() && hb.gen.empty() && hb.nar[0].empty() && hb.nar[1].empty()) return DG_OK;
    GemmBatchDev b;
    TableUseScope uses{stream};  // after the launch below returns
    auto table = [&](const auto& v, auto** dst) -> bool {
        using T = typename std::decay_t<decltype(v)>::value_type;
        *dst = nullptr;
        if (v.empty()) return true;
        *dst = static_cast<const T*>(cached_table(v.data(), v.size() * sizeof(T), stream));
        return *dst != nullptr;
    };
    if (!table(hb.dma, &b.dma) || !table(hb.dmae, &b.dmae) || !table(hb.gen, &b.gen) || !table(hb.segs, &b.segs) ||
        !table(hb.nar[0], &b.nar[0]) || !table(hb.nar[1], &b.nar[1]))
        return DG_ERR_ALLOC;
    for (int w = 0; w < 2; ++w) {
        b.nnar[w] = (int)hb.nar[w].size();
        b.tiles_nar[w] = hb.tiles_nar[w];
        b.vec_nar[w] = hb.vec_nar[w];
    }
    b.ndma = (int)hb.dma.size();
    b.tiles_dma = hb.tiles_dma;
    b.ndmae = (int)hb.dmae.size();
    b.tiles_dmae = hb.tiles_dmae;
    b.ngen = (int)hb.gen.size();
    b.tiles_gen[0] = hb.tiles_gen[0];
    b.tiles_gen[1] = hb.tiles_gen[1];
    b.vec = hb.vec;
    if (mode == DG_GEMM_MODE_PERSISTENT) mode = DG_GEMM_MODE_XCD;  // needs a counter: plans only
    return launch_gemm_f64_dev(opA, opB, mode, default_tile_cfg(), b, nullptr, stream);
}

}  // namespace dg

extern "C" int dg_dgemm(char opA, char opB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                        int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                        void* stream) {
    using namespace dg;
    const char oa = norm_op(opA), ob = norm_op(opB);
    if (!oa) return fail(DG_ERR_INVALID, "dg_dgemm: invalid opA '%c'", opA);
    if (!ob) return fail(DG_ERR_INVALID, "dg_dgemm: invalid opB '%c'", opB);
    if (m < 0 || n < 0 || k < 0) return fail(DG_ERR_INVALID, "dg_dgemm: negative dimension");
    const int64_t rowsA = oa == 'N' ? m : k;
    const int64_t rowsB = ob == 'N' ? k : n;
    if (lda < (rowsA > 1 ? rowsA : 1)) return fail(DG_ERR_INVALID, "dg_dgemm: lda %lld < %lld", (long long)lda, (long long)rowsA);
    if (ldb < (rowsB > 1 ? rowsB : 1)) return fail(DG_ERR_INVALID, "dg_dgemm: ldb %lld < %lld", (long long)ldb, (long long)rowsB);
    if (ldc < (m > 1 ? m : 1)) return fail(DG_ERR_INVALID, "dg_dgemm: ldc %lld < %lld", (long long)ldc, (long long)m);
    if (m == 0 || n == 0) return DG_OK;
    int rc = require_device();
    if (rc != DG_OK) return rc;
    hipStream_t st = as_stream(stream);
    if (k == 0 || alpha == 0.0) return launch_scale(m, n, beta, C, ldc, st);
    if (!A || !B || !C) return fail(DG_ERR_INVALID, "dg_dgemm: null matrix pointer");
    std::vector<Seg> segs{Seg{A, B, lda, ldb, k}};
    Prob p{};
    p.m = m; p.n = n; p.c = C; p.ldc = ldc; p.alpha = alpha; p.beta = beta; p.seg0 = 0; p.nseg = 1;
    std::vector<Prob> probs{p};
    return launch_gemm_f64(oa, ob, probs, segs, st, DG_GEMM_MODE_XCD);
}
