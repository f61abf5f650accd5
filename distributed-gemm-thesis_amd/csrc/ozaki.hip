// Ozaki split DGEMM on the low-precision matrix cores.
//
// Ozaki, Ogita, Oishi, Rump, "Error-free transformations of matrix
// multiplication by using fast routines of matrix multiplication and its
// applications", Numer. Algorithms 59 (2012); Ootomo, Ozaki, Yokota, "DGEMM on
// integer matrix multiplication unit", IJHPCA (2024).
//
//   1. e_i = frexp exponent of max_k |op(A)_ik| (max < 2^e_i), f_j likewise
//      for the columns of op(B).
//   2. r_0 = x * 2^-e, d_t = trunc(r_{t-1} * 2^b), r_t = r_{t-1} * 2^b - d_t
//      for t = 1..s, all exact in FP64: INT8 digits carry b = 7 bits
//      (|d| <= 127), BF16 digits b = 5 bits (|d| <= 31, exact in bf16).
//   3. For level l = 2..s+1 and t = 1..l-1 the exact integer product
//      A_t B_{l-t} is accumulated on the matrix cores (INT8 -> INT32,
//      BF16 -> FP32 below 2^24, both exact); the 32-bit accumulator is added
//      into an FP64 partial sum W as ldexp(acc, -b*l) whenever it could
//      overflow next (the host computes the flush points) and at the end of
//      each level.
//   4. C = alpha * ldexp(W, e_i + f_j) + beta * C, without fused multiply-add
//      so that the CPU oracle reproduces every bit.
//
// Slices are stored k-contiguous, [t][row][Kp] with rows padded to 256 and K
// padded to 128 bytes with zeros, so the product kernels have no edge cases
// until the final store.  Each run of products between two flushes (a
// group) runs as one long-K product through the scheme II pipeline
// (oz_plane_kernel: 128-byte k-steps, described with crt_gemm_kernel) into a
// 32-bit plane per group (INT32 sums, or the FP32 sums of BF16 slices), and
// oz_combine_kernel adds the planes into the FP64 partial sums W (workspace)
// in schedule order, with the operations a flush inside the product loop
// would make.
#include "dg_internal.hpp"

#include <type_traits>

// Every FP64 operation of the recombination is rounded on its own (no fused
// multiply-add), matching the CPU oracle bit for bit.
#pragma clang fp contract(off)

namespace dg {

namespace {

typedef int iv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));
typedef double dv4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

constexpr int PAD = 256;     // row padding of the slices and residues (a multiple of BM and BN)
constexpr int KSB = 128;     // K padding granule of the slices and residues, bytes
// Product kernels: BM x BN output tiles of 8 waves of WM x WN (FM x FN
// MFMA tiles of 16x16), one workgroup per CU.
constexpr int BM = 256, BN = 256;
constexpr int WM = 128, WN = 64;
constexpr int OCC = 1;
constexpr int NW = (BM / WM) * (BN / WN);  // waves per workgroup
constexpr int CNT = NW * 64;               // threads per workgroup
constexpr int FM = WM / 16, FN = WN / 16;  // 16x16 MFMA tiles per wave
static_assert(PAD % BM == 0 && PAD % BN == 0, "slice rows are padded to whole tiles");
constexpr int NT = 256;
constexpr int GROUP_M = 4;

__device__ __forceinline__ int xcd_remap(int bid, int total) {
    constexpr int X = 8;
    const int q = total / X, r = total % X;
    const int x = bid % X, slot = bid / X;
    return x * q + min(x, r) + slot;
}

// One product of the level schedule: slices (t, u) over k range [k0, k0+klen)
// of the padded K, then optionally flush the 32-bit accumulator with weight
// 2^-(bits*level).
struct OzSeg {
    int32_t t, u;       // 0-based slice indices
    int32_t k0, klen;   // in elements, multiples of the k-step
    int32_t level;      // t + u + 2 (1-based level)
    int32_t flush;      // 0: keep accumulating; 1: add into W; 2: first flush (W = x);
                        // 3: last flush (C = alpha*ldexp(W + x, e+f) + beta*C); 4: first and last
};

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt field");
    // gfx9 s_waitcnt: vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt left at their maxima
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void gvoid_t;

// A pointer the caller knows to be equal across the wave, in scalar
// registers (a uniform base plus a 32-bit lane offset lets a load take the
// scalar-base address form instead of a 64-bit address per lane).
__device__ __forceinline__ const int8_t* wave_uniform(const int8_t* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<const int8_t*>(((uint64_t)hi << 32) | lo);
}

// Output tile of workgroup slot `bid` among `tiles`: XCD-aware, GROUP_M tile
// rows sweep the tile columns together.  With 256x256 tiles and 32 resident
// workgroups per XCD, GROUP_M = 4 gives each XCD a 4 x 8 block of tiles (12
// distinct operand slices per k-step, as for 8 x 4) while the A panels all
// XCDs revisit stay at half the size, which keeps them in the last-level
// cache (measured on the 16384^3 CRT product: 4.8% faster than 8 x 4).
__device__ __forceinline__ void tile_origin(int bid, int tiles, int tiles_m, int tiles_n, int64_t& m0, int64_t& n0) {
    const int tile = xcd_remap(bid, tiles);
    const int per_group = GROUP_M * tiles_n;
    const int group = tile / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int in_group = tile - group * per_group;
    m0 = (int64_t)(first_m + in_group % gsize) * BM;
    n0 = (int64_t)(in_group / gsize) * BN;
}

// ------------------------------------------------------------ Ozaki scheme II
// Ozaki, Uchino, Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique"
// (2025).  A' = rint(op(A)_ik * 2^sigma_i) (rounded to nearest, which
// halves the RMS scaling error of truncation) and B' = rint(op(B)_kj *
// 2^tau_j) are integers; the shifts follow the paper's Cauchy-Schwarz bound
// (its fast mode): every scaled line has 2-norm <= 2^Q before rounding, with
// Q the largest such that 2^(2Q) + ceil(sqrt(k)) 2^Q + k <= P/4
// (P = prod p_t), so |V| = |A'B'| <= P/4.  The norm bound is an exact integer
// sum (crt_bound_term), so the shifts do not depend on summation order; a
// line whose norm is far below sqrt(k) times its largest entry (wide-ranging
// data) gets more bits than the per-entry bound would give it, up to 53.
// V is computed modulo N pairwise coprime moduli p_t <= 256 (one INT8 GEMM
// each, residues balanced into [-128, 127]) and rebuilt exactly by the
// Chinese remainder theorem, V = sum_t x_t (P/p_t) mod P, balanced.
// Shifts are stored as e' = Q - sigma (e'_i, f'_j), and
// C = alpha * ldexp(double(V), e'_i + f'_j - 2Q) + beta*C with double(V)
// correctly rounded.

constexpr int kMaxModuli = 16;
// Pairwise coprime: 2^8, 3*5*17, 11*23, 13*19, 7*31 and primes.
__constant__ int c_moduli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193};


// Direct CRT: V = (sum_t x_t * M_t) mod P with M_t = P / p_t and
// x_t = r_t * (M_t^-1 mod p_t) mod p_t, then the balanced representative.
struct CrtTables {
    int32_t n;                          // moduli in use
    int32_t cinv[kMaxModuli];           // M_t^-1 mod p_t
    double mt[kMaxModuli][3];           // M_t in 40-bit limbs (as doubles), least significant first
    double Pl[3];                       // P = prod p_t in 40-bit limbs (as doubles), least significant first
    double invPd;                       // 1 / fl(P) rounded to double
};

// p_t as a constant expression, so that x % crt_modulus(t) in an unrolled
// loop compiles to a multiply-shift.
__host__ __device__ constexpr int crt_modulus(int t) {
    return t == 0 ? 256 : t == 1 ? 255 : t == 2 ? 253 : t == 3 ? 251 : t == 4 ? 247 : t == 5 ? 241 : t == 6 ? 239
         : t == 7 ? 233 : t == 8 ? 229 : t == 9 ? 227 : t == 10 ? 223 : t == 11 ? 217 : t == 12 ? 211
         : t == 13 ? 199 : t == 14 ? 197 : 193;
}

// Balanced residues of A' = rint(op(X)(row, k) * 2^(q - e_row)) for every
// modulus, [t][row][Kp] int8 (zeros in the padding).  A workgroup stages a
// 16-row x 128-k tile of op(X) through LDS (loaded along whichever direction
// is contiguous in memory; 16.5 KB, so eight workgroups share a CU and the
// load and arithmetic phases of different workgroups overlap); each thread
// then turns 8 consecutive k of one row into one 8-byte store per modulus,
// so a wave writes whole 128-byte lines.  Residues in [-floor(p/2),
// ceil(p/2) - 1] from an exact split of A' (below), five FP64 operations
// per element and modulus.
constexpr int RT_R = 16, RT_K = 128, RT_T = 256 / RT_R, RT_KPT = RT_K / RT_T;  // threads per row, k per thread
// Per-modulus constants of the residue pass, passed as kernel arguments so
// that they sit in scalar registers (a VOP3 FMA takes them directly; gfx9
// VOP3 encodings take no 64-bit literals).
struct ResidueConsts {
    double c24[kMaxModuli];   // 2^24 mod p
    double inv_p[kMaxModuli]; // fl(1/p)
    double pd[kMaxModuli];    // p
};

// Tile order of the residue pass (blockIdx.x -> row tile, k tile): lines
// contiguous in k walk each tile row along k; strided lines go in groups of
// RT_KG k tiles, row tiles within a group, so that the workgroups in flight
// read long runs of each column of op(X) and still write 2 KB runs of each
// residue row.
constexpr uint32_t RT_KG = 16;

template <int NMOD>
__global__ __launch_bounds__(256) void crt_residue_kernel(const double* __restrict__ x, int64_t rows, int64_t K,
                                                          int64_t rs, int64_t ks, const int32_t* __restrict__ e, int q,
                                                          int8_t* __restrict__ out, int64_t Kp, int64_t slice_bytes,
                                                          ResidueConsts rc, uint32_t nkt, uint32_t nrt) {
    __shared__ double tile[RT_R][RT_K + 1];
    uint32_t kt, rt;
    if (ks == 1) {
        rt = blockIdx.x / nkt;
        kt = blockIdx.x - rt * nkt;
    } else {
        const uint32_t grp = blockIdx.x / (RT_KG * nrt);
        const uint32_t first = grp * RT_KG;
        const uint32_t gk = nkt - first < RT_KG ? nkt - first : RT_KG;
        const uint32_t in = blockIdx.x - grp * RT_KG * nrt;
        rt = in / gk;
        kt = first + (in - rt * gk);
    }
    const int64_t r0 = (int64_t)rt * RT_R;
    const int64_t k0 = (int64_t)kt * RT_K;
    const int t = threadIdx.x;
    // Eight loads per thread, all issued before the first wait: addresses
    // outside op(X) are clamped onto its last row / k (in bounds) and the
    // loaded value replaced by zero.  op(X) is read once here and the
    // residues written once, so both go nontemporal.
    if (ks == 1) {
        const int64_t k = k0 + (t & 127);
        const int64_t kc = k < K ? k : K - 1;
        double v[RT_R / 2];
#pragma unroll
        for (int i = 0; i < RT_R / 2; ++i) {
            const int64_t r = r0 + (t >> 7) + 2 * i;
            v[i] = __builtin_nontemporal_load(x + (r < rows ? r : rows - 1) * rs + kc);
        }
#pragma unroll
        for (int i = 0; i < RT_R / 2; ++i) {
            const int rr = (t >> 7) + 2 * i;
            tile[rr][t & 127] = (r0 + rr < rows && k < K) ? v[i] : 0.0;
        }
    } else {
        const int64_t r = r0 + (t % RT_R);
        const int64_t rc = r < rows ? r : rows - 1;
        double v[RT_K * RT_R / 256];
#pragma unroll
        for (int i = 0; i < RT_K * RT_R / 256; ++i) {
            const int64_t k = k0 + t / RT_R + (256 / RT_R) * i;
            v[i] = __builtin_nontemporal_load(x + rc * rs + (k < K ? k : K - 1) * ks);
        }
#pragma unroll
        for (int i = 0; i < RT_K * RT_R / 256; ++i) {
            const int kk = t / RT_R + (256 / RT_R) * i;
            tile[t % RT_R][kk] = (r < rows && k0 + kk < K) ? v[i] : 0.0;
        }
    }
    __syncthreads();
    const int rr = t / RT_T;
    const int kq = (t % RT_T) * RT_KPT;
    const int64_t r = r0 + rr;
    // a = ah 2^24 + al exactly (al in [0, 2^24)), so for each modulus
    // y = ah (2^24 mod p) + al = a (mod p) with |y| < 2^38, whose quotient
    // estimate fma(y, fl(1/p), 1/2) is within 2^-22 of y/p + 1/2 (never closer
    // than 1/(2p) to an integer for odd p; exact for p = 256, where y = al):
    // r = y - p floor(y/p + 1/2) is the balanced residue, computed exactly.
    double ah[RT_KPT], al[RT_KPT];
    const int sh = r < rows ? q - e[r] : 0;
#pragma unroll
    for (int c = 0; c < RT_KPT; ++c) {
        const double a = r < rows ? rint(ldexp(tile[rr][kq + c], sh)) : 0.0;
        ah[c] = floor(a * 0x1p-24);
        al[c] = fma(-ah[c], 0x1p24, a);
    }
    int8_t* dst = out + r * Kp + k0 + kq;
#pragma unroll
    for (int m = 0; m < NMOD; ++m) {
        static_assert(RT_KPT == 8, "one 8-byte store per modulus");
        uint32_t v[RT_KPT];
#pragma unroll
        for (int c = 0; c < RT_KPT; ++c) {
            const double y = m == 0 ? al[c] : fma(ah[c], rc.c24[m], al[c]);
            const double qd = floor(fma(y, rc.inv_p[m], 0.5));
            v[c] = (uint32_t)(int)fma(-qd, rc.pd[m], y);
        }
        // bytes 0 of four words into one word: two selections of two bytes, then one of the two halves
        const uint32_t w0 = __builtin_amdgcn_perm(__builtin_amdgcn_perm(v[3], v[2], 0x0c0c0400u),
                                                  __builtin_amdgcn_perm(v[1], v[0], 0x0c0c0400u), 0x05040100u);
        const uint32_t w1 = __builtin_amdgcn_perm(__builtin_amdgcn_perm(v[7], v[6], 0x0c0c0400u),
                                                  __builtin_amdgcn_perm(v[5], v[4], 0x0c0c0400u), 0x05040100u);
        __builtin_nontemporal_store(u32x2{w0, w1}, reinterpret_cast<u32x2*>(dst + m * slice_bytes));
    }
}

struct CrtArgs {
    const int8_t* a;     // residues of A', [N][Mp][Kp]
    const int8_t* b;     // residues of B', [N][Np][Kp]
    uint8_t* res;        // CRT digits x_t of V, [N][Np][Mp] (column-major per modulus, planes res_plane apart)
    int64_t Mp, Np, Kp, res_plane;
    int64_t a_slice, b_slice;
    int32_t tiles_m, tiles_n, nmod;
    int32_t cinv[kMaxModuli];  // (P / p_t)^-1 mod p_t
};

// The scheme II product pipeline moves 128-byte k-steps: one
// global_load_lds_dwordx4 of a wave covers 8 rows x 128 bytes, so every L2
// request is a whole 128-byte line (64-byte half lines cost the same request
// slot and halve the operand stream).  A step's A rows and B rows each fill
// a 32 KB slot; five slots (160 KB) form a ring in which A_0 B_0 A_1 B_1 ...
// take consecutive slots.  Within a row, 16-byte chunk c sits at position
// c ^ (row % 8): the fragment reads (lane l: row l % 16, chunk 4h + l / 16 of
// half-step h) then run at the full LDS rate.  A step is two MFMA half-steps
// of 64 bytes; the barrier sits between them: it publishes step s+1 (A_{s+1}
// loaded two steps earlier, B_{s+1} one step earlier) and releases the slots
// of step s, whose fragments are all in registers by then, to B_{s+2} and
// A_{s+3}.  The second half-step's MFMAs run while the fragments of the
// first half of step s+1 are read, and the first half-step's while those of
// the second half are read.
constexpr int CK2 = 128;                  // bytes per scheme II k-step
constexpr int CSLOT = BM * CK2;           // one operand's rows of one step
constexpr int NSLOT = 5;
constexpr size_t kCrtLds = (size_t)NSLOT * CSLOT;
static_assert(BM == BN && BM / NW == 32, "scheme II loader: 32 rows per wave and operand");

__device__ __forceinline__ int next_slot(int x) { return x + 1 == NSLOT ? 0 : x + 1; }

// This wave's 32 rows of one operand for k-step s into `slot`; g (uniform
// across the wave) points at the wave's first row, loff is the lane's
// offset to its row and chunk, rows `ld` bytes apart.  A uniform base and a
// 32-bit lane offset let every load take the scalar-base address form.
__device__ __forceinline__ void crt_issue(int8_t* lds, int slot, int wave, const int8_t* g, uint32_t loff, int64_t s,
                                          int64_t ld) {
    int8_t* d = lds + slot * CSLOT + wave * 4096;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        __builtin_amdgcn_global_load_lds((gvoid_t*)(wave_uniform(g + s * CK2 + (int64_t)(8 * i) * ld) + loff),
                                         (lds_void_t*)(d + 1024 * i), 16, 0, 0);
}

// One 64-byte MFMA of the pipeline: 16x16 INT8 products summed in INT32
// (iv4 accumulators) or BF16 products summed in FP32 (fv4 accumulators, for
// the BF16 slices of scheme I).  Either way a lane's A and B operands are 16
// bytes of one row, so the LDS layout is the same.
__device__ __forceinline__ void mfma64(iv4& acc, const iv4& a, const iv4& b) {
    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ void mfma64(fv4& acc, const iv4& a, const iv4& b) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), acc, 0, 0, 0);
}

// One MFMA half-step on fa / cb, reading the fragments of the following
// half-step meanwhile: fa[i] right after its last use, B into nb.
template <class ACC>
__device__ __forceinline__ void crt_half(ACC (&acc)[FM][FN], iv4 (&fa)[FM], const iv4 (&cb)[FN], iv4 (&nb)[FN],
                                         const int8_t* nas, const int8_t* nbs) {
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) mfma64(acc[i][j], fa[i], cb[j]);
        if (i == 0) {
#pragma unroll
            for (int j = 0; j < FN; ++j) nb[j] = *reinterpret_cast<const iv4*>(nbs + 2048 * j);
        }
        fa[i] = *reinterpret_cast<const iv4*>(nas + 2048 * i);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, FN, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, FN + 1, 0);
#pragma unroll
    for (int g = 1; g < FM; ++g) {
        __builtin_amdgcn_sched_group_barrier(0x008, FN, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    }
}

// crt_half that also issues this wave's loads of B step sb (into slot_b)
// and A step sa (into slot_a), one global_load_lds after each group of FN
// MFMAs:
// issued back to back, the eight 1 KB loads of every wave queue at the
// vector memory unit while the matrix cores wait.
template <class ACC>
__device__ __forceinline__ void crt_half_ld(ACC (&acc)[FM][FN], iv4 (&fa)[FM], const iv4 (&cb)[FN], iv4 (&nb)[FN],
                                            const int8_t* nas, const int8_t* nbs, int8_t* lds, int slot_b,
                                            int slot_a, int wave, const int8_t* gb, const int8_t* ga, uint32_t loff,
                                            int64_t sb, int64_t sa, int64_t ld) {
    static_assert(FM == 8, "one load per MFMA group: four of B, then four of A");
    int8_t* db = lds + slot_b * CSLOT + wave * 4096;
    int8_t* da = lds + slot_a * CSLOT + wave * 4096;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) mfma64(acc[i][j], fa[i], cb[j]);
        if (i == 0) {
#pragma unroll
            for (int j = 0; j < FN; ++j) nb[j] = *reinterpret_cast<const iv4*>(nbs + 2048 * j);
        }
        fa[i] = *reinterpret_cast<const iv4*>(nas + 2048 * i);
        if (i < 4)
            __builtin_amdgcn_global_load_lds((gvoid_t*)(wave_uniform(gb + sb * CK2 + (int64_t)(8 * i) * ld) + loff),
                                             (lds_void_t*)(db + 1024 * i), 16, 0, 0);
        else
            __builtin_amdgcn_global_load_lds(
                (gvoid_t*)(wave_uniform(ga + sa * CK2 + (int64_t)(8 * (i - 4)) * ld) + loff),
                (lds_void_t*)(da + 1024 * (i - 4)), 16, 0, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, FN, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, FN + 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
#pragma unroll
    for (int g = 1; g < FM; ++g) {
        __builtin_amdgcn_sched_group_barrier(0x008, FN, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
}

// One INT8 GEMM per modulus in one launch: workgroup b computes tile
// b % tiles of modulus b / tiles and stores the CRT digit
// x_t = ((A'_t B'_t) mod p_t) * ((P / p_t)^-1 mod p_t) mod p_t as a byte.
__global__ __launch_bounds__(CNT, OCC) void crt_gemm_kernel(CrtArgs g) {
    extern __shared__ __attribute__((aligned(1024))) int8_t lds[];
    const int tiles = g.tiles_m * g.tiles_n;
    const int tmod = blockIdx.x / tiles;
    int64_t m0, n0;
    tile_origin(blockIdx.x - tmod * tiles, tiles, g.tiles_m, g.tiles_n, m0, n0);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave / (BN / WN)) * WM;
    const int wn = (wave % (BN / WN)) * WN;
    const int lr = lane & 15;
    const int lg = lane >> 4;

    // loader: lane -> row lane / 8 of an 8-row group, position lane % 8
    // holding global chunk (lane % 8) ^ (row % 8)
    const int lrow = lane >> 3;
    const int lchunk = (lane & 7) ^ lrow;
    const int8_t* ga = g.a + tmod * g.a_slice + (m0 + 32 * wave) * g.Kp;
    const int8_t* gb = g.b + tmod * g.b_slice + (n0 + 32 * wave) * g.Kp;
    const uint32_t loff = (uint32_t)(lrow * g.Kp + 16 * lchunk);
    const int roff0 = lr * CK2 + 16 * (lg ^ (lr & 7));
    const int roff1 = lr * CK2 + 16 * ((4 + lg) ^ (lr & 7));

    iv4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = iv4{0, 0, 0, 0};

    const int64_t nk = g.Kp / CK2;  // >= 1
    // Every step issues the same loads, so that every wait is the same: a
    // load of a step past the end reloads the last step into a slot that is
    // free and never read again (its data unused).
    const auto clamp = [nk](int64_t s) { return s < nk ? s : nk - 1; };
    // prologue: A_0 B_0 A_1 B_1 A_2 into slots 0..4
    crt_issue(lds, 0, wave, ga, loff, 0, g.Kp);
    crt_issue(lds, 1, wave, gb, loff, 0, g.Kp);
    crt_issue(lds, 2, wave, ga, loff, clamp(1), g.Kp);
    crt_issue(lds, 3, wave, gb, loff, clamp(1), g.Kp);
    crt_issue(lds, 4, wave, ga, loff, clamp(2), g.Kp);
    int isl = 0;
    // step 0 landed: at most A_1 B_1 A_2 (4 loads each) pending
    wait_vmcnt<12>();
    __builtin_amdgcn_s_barrier();
    int csl = 0;  // slot of A_s; B_s in the next one
    iv4 fa[FM], cb[FN], nb[FN];
    {
        const int8_t* as = lds + wm * CK2 + roff0;
        const int8_t* bs = lds + CSLOT + wn * CK2 + roff0;
#pragma unroll
        for (int j = 0; j < FN; ++j) cb[j] = *reinterpret_cast<const iv4*>(bs + 2048 * j);
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = *reinterpret_cast<const iv4*>(as + 2048 * i);
    }
    for (int64_t s = 0; s + 1 < nk; ++s) {
        const int bsl = next_slot(csl), asl1 = next_slot(bsl), bsl1 = next_slot(asl1);
        // first half of step s; fragments of its second half read meanwhile
        crt_half(acc, fa, cb, nb, lds + csl * CSLOT + wm * CK2 + roff1, lds + bsl * CSLOT + wn * CK2 + roff1);
        // every fragment of step s is in registers once this wave's LDS reads
        // are done; A_{s+1} and B_{s+1} landed with at most A_{s+2} pending
        __builtin_amdgcn_s_waitcnt((0xF) | (0x3 << 14) | (0x7 << 4));  // lgkmcnt(0)
        wait_vmcnt<4>();
        __builtin_amdgcn_s_barrier();
        // second half of step s, fragments of the first half of step s+1 read
        // meanwhile, B_{s+2} and A_{s+3} issued among its MFMAs into the slots
        // of step s
        const int sb = isl, sa = next_slot(isl);
        isl = next_slot(sa);
        crt_half_ld(acc, fa, nb, cb, lds + asl1 * CSLOT + wm * CK2 + roff0, lds + bsl1 * CSLOT + wn * CK2 + roff0,
                    lds, sb, sa, wave, gb, ga, loff, clamp(s + 2), clamp(s + 3), g.Kp);
        csl = asl1;
    }
    {
        // last step: no loads; the second half's reads of a following step
        // are in bounds and unused
        const int bsl = next_slot(csl), asl1 = next_slot(bsl), bsl1 = next_slot(asl1);
        crt_half(acc, fa, cb, nb, lds + csl * CSLOT + wm * CK2 + roff1, lds + bsl * CSLOT + wn * CK2 + roff1);
        crt_half(acc, fa, nb, cb, lds + asl1 * CSLOT + wm * CK2 + roff0, lds + bsl1 * CSLOT + wn * CK2 + roff0);
    }
    wait_vmcnt<0>();  // the reloads past the end land before the workgroup's LDS is released
    // acc[i][j][v] = V_t(m0 + wm + 16i + 4lg + v, n0 + wn + 16j + lr): four
    // consecutive rows of one column, stored as one 32-bit word of digits.
    // x = acc * cinv is exact (|x| < 2^39) and floor(fl(x fl(1/p) + 2^-10)) is
    // floor(x/p) exactly (the estimate errs by < 2^-20 and the fractional
    // part of x/p lies in [0, 1 - 1/p]), so x - p floor(x/p) is the digit.
    const int p = c_moduli[tmod];
    const double pd = p, inv_p = 1.0 / p, cinv = g.cinv[tmod];
    uint8_t* out = g.res + (int64_t)tmod * g.res_plane;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int64_t n = n0 + wn + 16 * j + lr;
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int64_t m = m0 + wm + 16 * i + 4 * lg;
            uint32_t w = 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double x = (double)acc[i][j][v] * cinv;
                const double qd = floor(fma(x, inv_p, 0x1p-10));
                w |= (uint32_t)(int)fma(-qd, pd, x) << (8 * v);
            }
            *reinterpret_cast<uint32_t*>(out + n * g.Mp + m) = w;
        }
    }
}

// ---------------------------------------------- scheme I on the scheme II pipeline
// The products of scheme I between two flushes (a group: slices t0 + p of A
// against slices u0 - p of B for p < count, one level, one k range) add up
// to one long-K product, so a group's k-steps run back to back through the
// scheme II pipeline above.  Workgroup b computes tile b % tiles of the
// launch's group b / tiles and stores its 32-bit sums (INT32 for INT8
// slices, FP32 for BF16 slices) as a plane, [Np][Mp] column-major;
// oz_combine_kernel then adds the planes into FP64 in group order.
struct OzGroup {
    int32_t t0, u0;   // 0-based slices of the first product
    int32_t count;    // products
    int32_t level;    // t + u + 2 (t, u 1-based) of every product
    int32_t kb;       // byte offset of the k range within a slice row
    FastDiv nks;      // 128-byte k-steps per product (nks.d), as a divisor
};

struct OzPlaneArgs {
    const int8_t* a;             // A slices, [s][Mp] rows of rb bytes
    const int8_t* b;             // B slices, [s][Np] rows of rb bytes
    int32_t* planes;             // one [Np][Mp] plane per group of the launch
    const OzGroup* groups;
    int32_t g0;                  // first group of the launch
    int32_t tiles_m, tiles_n;
    int64_t Mp, rb;
    int64_t a_slice, b_slice;    // bytes per slice
    int64_t plane;               // elements per plane
};

// This wave's rows of k-step s of a group (clamped to its last step, for
// the pipeline's reloads past the end): product p = s / nks, step s % nks
// within it, the product's slice `stride` bytes from the previous one's.
// Branch-free scalar arithmetic, so the loop body stays one basic block.
__device__ __forceinline__ const int8_t* group_step(const int8_t* first, int64_t stride, uint32_t s, uint32_t last,
                                                    const FastDiv& nks) {
    const uint32_t sc = s < last ? s : last;
    const uint32_t p = fdiv(sc, nks);
    return first + (int64_t)p * stride + (int64_t)(sc - p * nks.d) * CK2;
}

template <bool BF16>
__global__ __launch_bounds__(CNT, OCC) void oz_plane_kernel(OzPlaneArgs g) {
    typedef typename std::conditional<BF16, fv4, iv4>::type acc_t;
    extern __shared__ __attribute__((aligned(1024))) int8_t lds[];
    const int tiles = g.tiles_m * g.tiles_n;
    const int gi = blockIdx.x / tiles;
    int64_t m0, n0;
    tile_origin(blockIdx.x - gi * tiles, tiles, g.tiles_m, g.tiles_n, m0, n0);
    const OzGroup gr = g.groups[g.g0 + gi];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave / (BN / WN)) * WM;
    const int wn = (wave % (BN / WN)) * WN;
    const int lr = lane & 15;
    const int lg = lane >> 4;
    const int lrow = lane >> 3;
    const int lchunk = (lane & 7) ^ lrow;
    const uint32_t loff = (uint32_t)(lrow * g.rb + 16 * lchunk);
    const int roff0 = lr * CK2 + 16 * (lg ^ (lr & 7));
    const int roff1 = lr * CK2 + 16 * ((4 + lg) ^ (lr & 7));

    const uint32_t last = (uint32_t)gr.count * gr.nks.d - 1;  // the group's last k-step
    const int8_t* ga = g.a + gr.t0 * g.a_slice + (m0 + 32 * wave) * g.rb + gr.kb;
    const int8_t* gb = g.b + gr.u0 * g.b_slice + (n0 + 32 * wave) * g.rb + gr.kb;
    const auto a_at = [&](uint32_t s) { return group_step(ga, g.a_slice, s, last, gr.nks); };
    const auto b_at = [&](uint32_t s) { return group_step(gb, -g.b_slice, s, last, gr.nks); };

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // prologue: A_0 B_0 A_1 B_1 A_2 into slots 0..4, as in crt_gemm_kernel
    crt_issue(lds, 0, wave, a_at(0), loff, 0, g.rb);
    crt_issue(lds, 1, wave, b_at(0), loff, 0, g.rb);
    crt_issue(lds, 2, wave, a_at(1), loff, 0, g.rb);
    crt_issue(lds, 3, wave, b_at(1), loff, 0, g.rb);
    crt_issue(lds, 4, wave, a_at(2), loff, 0, g.rb);
    int isl = 0;
    wait_vmcnt<12>();
    __builtin_amdgcn_s_barrier();
    int csl = 0;
    iv4 fa[FM], fb[FN], nb[FN];
    {
        const int8_t* as = lds + wm * CK2 + roff0;
        const int8_t* bs = lds + CSLOT + wn * CK2 + roff0;
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = *reinterpret_cast<const iv4*>(bs + 2048 * j);
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = *reinterpret_cast<const iv4*>(as + 2048 * i);
    }
    for (uint32_t s = 0; s < last; ++s) {
        const int bsl = next_slot(csl), asl1 = next_slot(bsl), bsl1 = next_slot(asl1);
        crt_half(acc, fa, fb, nb, lds + csl * CSLOT + wm * CK2 + roff1, lds + bsl * CSLOT + wn * CK2 + roff1);
        __builtin_amdgcn_s_waitcnt((0xF) | (0x3 << 14) | (0x7 << 4));  // lgkmcnt(0)
        wait_vmcnt<4>();
        __builtin_amdgcn_s_barrier();
        const int sb = isl, sa = next_slot(isl);
        isl = next_slot(sa);
        crt_half_ld(acc, fa, nb, fb, lds + asl1 * CSLOT + wm * CK2 + roff0, lds + bsl1 * CSLOT + wn * CK2 + roff0,
                    lds, sb, sa, wave, b_at(s + 2), a_at(s + 3), loff, 0, 0, g.rb);
        csl = asl1;
    }
    {
        const int bsl = next_slot(csl), asl1 = next_slot(bsl), bsl1 = next_slot(asl1);
        crt_half(acc, fa, fb, nb, lds + csl * CSLOT + wm * CK2 + roff1, lds + bsl * CSLOT + wn * CK2 + roff1);
        crt_half(acc, fa, nb, fb, lds + asl1 * CSLOT + wm * CK2 + roff0, lds + bsl1 * CSLOT + wn * CK2 + roff0);
    }
    wait_vmcnt<0>();
    // acc[i][j][v] = sum at (m0 + wm + 16i + 4lg + v, n0 + wn + 16j + lr): one 16-byte store
    int32_t* out = g.planes + (int64_t)gi * g.plane;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int64_t n = n0 + wn + 16 * j + lr;
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int64_t m = m0 + wm + 16 * i + 4 * lg;
            __builtin_nontemporal_store(__builtin_bit_cast(iv4, acc[i][j]), reinterpret_cast<iv4*>(out + n * g.Mp + m));
        }
    }
}

// The FP64 recombination of scheme I from the planes of groups [gb, ge) of
// ng, for one thread's 4 rows of one column of the padded Mp x Np extent:
// each group's x = sum * 2^-(bits*level) (exact) is added in group order,
// W = x for the first group of all and W + x after (W carried in HBM from
// one launch to the next); after the last group C = alpha*ldexp(W, e_i + f_j)
// + beta*C.  Every FP64 operation is rounded on its own, in this order.
template <bool BF16>
__global__ __launch_bounds__(256) void oz_combine_kernel(const int32_t* __restrict__ planes, int64_t plane,
                                                         const OzGroup* __restrict__ groups, int gb, int ge, int ng,
                                                         int bits, double* __restrict__ W, int64_t Mp, int64_t m,
                                                         int64_t n, const int32_t* __restrict__ ea,
                                                         const int32_t* __restrict__ eb, double alpha, double beta,
                                                         double* __restrict__ C, int64_t ldc, FastDiv dqm,
                                                         uint32_t total) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;  // one thread per 4 rows
    if (x >= total) return;
    const uint32_t j = fdiv(x, dqm);
    const int64_t i0 = (int64_t)(x - j * dqm.d) * 4;
    const int64_t e0 = (int64_t)j * Mp + i0;
    double cur[4] = {0.0, 0.0, 0.0, 0.0};
    if (gb > 0) {
        const dv4 w = __builtin_nontemporal_load(reinterpret_cast<const dv4*>(W + e0));
#pragma unroll
        for (int e = 0; e < 4; ++e) cur[e] = w[e];
    }
    for (int gg = gb; gg < ge; ++gg) {
        const double wgt = ldexp(1.0, -bits * groups[gg].level);
        const iv4 v = __builtin_nontemporal_load(reinterpret_cast<const iv4*>(planes + (int64_t)(gg - gb) * plane + e0));
        // the FP32 sums of BF16 slices: the whole vector is bit-cast (a bit cast
        // of one vector element, v[e], read the wrong element with this compiler)
        const fv4 vf = __builtin_bit_cast(fv4, v);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double xv = (BF16 ? (double)vf[e] : (double)v[e]) * wgt;
            cur[e] = gg == 0 ? xv : __dadd_rn(cur[e], xv);
        }
    }
    if (ge < ng) {
        __builtin_nontemporal_store(dv4{cur[0], cur[1], cur[2], cur[3]}, reinterpret_cast<dv4*>(W + e0));
        return;
    }
    if ((int64_t)j >= n) return;
    const int jf = eb[j];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t i = i0 + e;
        if (i >= m) break;
        double* cp = C + i + (int64_t)j * ldc;
        double val = __dmul_rn(alpha, ldexp(cur[e], ea[i] + jf));
        if (beta != 0.0) val = __dadd_rn(val, __dmul_rn(beta, *cp));
        *cp = val;
    }
}

// The balanced CRT value V of one element, correctly rounded to double, from
// the limb sums L_l = sum_t x_t * limb_l(M_t) over 40-bit limbs (exact
// doubles below 2^52), S = L0 + L1 2^40 + L2 2^80 = V + jP with j in [0, 16].
// Since |V| <= P/4 (the scaling bound), S/P lies within 1/4 of the integer j
// and j = rint(fl(S) fl(1/P)) exactly (the estimate errs by ~2^-47).  Then
// in FP64, every step exact: D_l = L_l - j P_l (|D_l| < 2^52), and two carries
// bring D0, D1 into [0, 2^40), so V = D2 2^80 + D1 2^40 + D0 with D2 signed.
// The three limbs split into 32-bit words that make V in 128-bit two's
// complement, whose magnitude is rounded to double once: the top 64 bits
// with a sticky bit for everything below, converted with one rounding.
__host__ __device__ inline double crt_value(double L0, double L1, double L2, const CrtTables& tb) {
    const double Sd = fma(L2, 0x1p80, fma(L1, 0x1p40, L0));
    const double jd = rint(Sd * tb.invPd);
    double D0 = fma(-jd, tb.Pl[0], L0);
    double D1 = fma(-jd, tb.Pl[1], L1);
    double D2 = fma(-jd, tb.Pl[2], L2);
    double c = floor(D0 * 0x1p-40);
    D0 = fma(-c, 0x1p40, D0);
    D1 = D1 + c;
    c = floor(D1 * 0x1p-40);
    D1 = fma(-c, 0x1p40, D1);
    D2 = D2 + c;
    // 32-bit halves: D0, D1 in [0, 2^40) (high halves below 2^8), D2 signed below 2^44
    const double h0 = floor(D0 * 0x1p-32), h1 = floor(D1 * 0x1p-32), h2 = floor(D2 * 0x1p-32);
    const uint32_t lo0 = (uint32_t)fma(-h0, 0x1p32, D0), lo1 = (uint32_t)fma(-h1, 0x1p32, D1);
    const uint32_t lo2 = (uint32_t)fma(-h2, 0x1p32, D2);
    const uint32_t hi0 = (uint32_t)h0, hi1 = (uint32_t)h1, hi2 = (uint32_t)(int32_t)h2;
    const uint32_t w0 = lo0;
    const uint32_t w1 = hi0 | (lo1 << 8);
    const uint32_t w2 = (lo1 >> 24) | (hi1 << 8) | (lo2 << 16);
    const uint32_t w3 = (lo2 >> 16) | (hi2 << 16);
    uint64_t hi = ((uint64_t)w3 << 32) | w2;
    uint64_t lo = ((uint64_t)w1 << 32) | w0;
    const bool neg = (int64_t)hi < 0;
    // magnitude
    const uint64_t nlo = ~lo + 1;
    const uint64_t nhi = ~hi + (nlo == 0 ? 1u : 0u);
    lo = neg ? nlo : lo;
    hi = neg ? nhi : hi;
    const int z = hi ? __builtin_clzll(hi) : 64;
    const uint64_t top = z == 64 ? lo : (z ? (hi << z) | (lo >> (64 - z)) : hi);
    const uint64_t rest = z == 64 ? 0 : (z ? (lo << z) : lo);
    const double v = ldexp((double)(top | (rest != 0 ? 1ull : 0ull)), z == 64 ? 0 : 64 - z);
    return neg ? -v : v;
}

// Four consecutive rows per thread (one 32-bit load of CRT digits per modulus).
// The limb sums L_l = sum_t x_t * limb_l(M_t) over 40-bit limbs of M_t are
// exact FP64 FMAs (x_t < 256, limbs < 2^40, at most 16 terms: below 2^52);
// crt_value reduces S = sum_t x_t M_t < 16 P modulo P to the balanced V and
// rounds it to double once.
// C = alpha * ldexp(V, e_i + f_j - 2q) + beta*C, four rows per 32-byte access
// when C allows it.
// NMOD (the moduli count) is a template argument so that every digit load
// of a thread is issued before the first is used.
// Thread x (x < ceil(m/4) n < 2^32) takes rows 4 (x % qm) .. + 3 of column
// x / qm, qm = ceil(m/4), divided by multiply-shift.
template <int NMOD>
__global__ __launch_bounds__(256) void crt_reconstruct_kernel(const uint8_t* __restrict__ res, CrtTables tb, int64_t m,
                                                              int64_t n, int64_t Mp, int64_t plane,
                                                              const int32_t* __restrict__ ea,
                                                              const int32_t* __restrict__ eb, int q, double alpha,
                                                              double beta, double* __restrict__ C, int64_t ldc,
                                                              FastDiv dqm) {
    const uint32_t total = dqm.d * (uint32_t)n;
    {
        const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;  // one thread per 4 rows
        if (x >= total) return;
        const uint32_t jq = fdiv(x, dqm);
        const int64_t j = jq, i0 = (int64_t)(x - jq * dqm.d) * 4;
        const uint8_t* rp = res + j * Mp + i0;
        uint32_t w[NMOD];
#pragma unroll
        for (int t = 0; t < NMOD; ++t)  // digits and C are touched once: nontemporal accesses
            w[t] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(rp + (int64_t)t * plane));
        double L[4][3];
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int l = 0; l < 3; ++l) L[e][l] = 0.0;
#pragma unroll
        for (int t = 0; t < NMOD; ++t) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double xd = (double)((w[t] >> (8 * e)) & 255u);
#pragma unroll
                for (int l = 0; l < 3; ++l) L[e][l] = fma(xd, tb.mt[t][l], L[e][l]);
            }
        }
        const int jf = eb[j];
        double val[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t i = i0 + e < m ? i0 + e : m - 1;  // rows past m are computed but never stored
            val[e] = __dmul_rn(alpha, ldexp(crt_value(L[e][0], L[e][1], L[e][2], tb), ea[i] + jf - 2 * q));
        }
        double* cp = C + i0 + j * ldc;
        if (i0 + 3 < m && (reinterpret_cast<uintptr_t>(cp) & 31) == 0) {
            dv4* c4 = reinterpret_cast<dv4*>(cp);
            if (beta != 0.0) {
                const dv4 old = __builtin_nontemporal_load(c4);
#pragma unroll
                for (int e = 0; e < 4; ++e) val[e] = __dadd_rn(val[e], __dmul_rn(beta, old[e]));
            }
            __builtin_nontemporal_store(dv4{val[0], val[1], val[2], val[3]}, c4);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (i0 + e >= m) break;
                double v = val[e];
                if (beta != 0.0) v = __dadd_rn(v, __dmul_rn(beta, cp[e]));
                cp[e] = v;
            }
        }
    }
}

typedef double pdv2 __attribute__((ext_vector_type(2)));

// Rows contiguous in k whose every row starts 16-byte aligned: the scans read
// them in 16-byte pairs, 16 lanes per row (host and device take the same test).
__host__ __device__ inline bool pair_rows(const double* x, int64_t rs) {
    return !(rs & 1) && !(reinterpret_cast<uintptr_t>(x) & 15);
}

// Row maxima of op(X): partial max |x| over a k-chunk for every row.
// Element (row, k) of op(X) is at x[row*rs + k*ks].
__global__ void oz_partial_max_kernel(const double* __restrict__ x, int64_t rows, int64_t K, int64_t rs, int64_t ks,
                                      int64_t kchunk, double* __restrict__ part) {
    // blockIdx.y: k-chunk.  Rows strided in memory (ks != 1): a thread per
    // row, consecutive lanes on consecutive rows, four loads in flight.
    // Rows contiguous in k (ks == 1): a wave per row, lanes along k, then a
    // wave reduction.
    const int64_t k0 = blockIdx.y * kchunk;
    const int64_t k1 = min(K, k0 + kchunk);
    if (ks != 1) {
        for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
            const double* xr = x + r * rs;
            double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
            int64_t k = k0;
            for (; k + 4 <= k1; k += 4) {
                m0 = fmax(m0, fabs(xr[k * ks]));
                m1 = fmax(m1, fabs(xr[(k + 1) * ks]));
                m2 = fmax(m2, fabs(xr[(k + 2) * ks]));
                m3 = fmax(m3, fabs(xr[(k + 3) * ks]));
            }
            for (; k < k1; ++k) m0 = fmax(m0, fabs(xr[k * ks]));
            part[blockIdx.y * rows + r] = fmax(fmax(m0, m1), fmax(m2, m3));
        }
    } else if (pair_rows(x, rs)) {
        // 16 lanes per row, k pairs 32 apart: a load of the wave reads 256
        // contiguous bytes of each of its four rows
        const int l16 = threadIdx.x & 15;
        const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 4);
        for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 4) + (threadIdx.x >> 4); r < rows; r += groups) {
            const double* xr = x + r * rs;
            double mx = 0.0;
            for (int64_t k = k0 + 2 * l16; k < k1; k += 32) {
                if (k + 1 < k1) {
                    const pdv2 v = *reinterpret_cast<const pdv2*>(xr + k);
                    mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
                } else {
                    mx = fmax(mx, fabs(xr[k]));
                }
            }
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
            if (l16 == 0) part[blockIdx.y * rows + r] = mx;
        }
    } else {
        const int lane = threadIdx.x & 63;
        const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
        for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); r < rows; r += waves) {
            const double* xr = x + r * rs;
            double mx = 0.0;
            for (int64_t k = k0 + lane; k < k1; k += 64) mx = fmax(mx, fabs(xr[k]));
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
            if (lane == 0) part[blockIdx.y * rows + r] = mx;
        }
    }
}

// Reduces the chunk maxima and writes the frexp exponent (0 for zero rows,
// and for the padding rows beyond `rows`).
__global__ void oz_exponent_kernel(const double* __restrict__ part, int64_t rows, int64_t rows_p, int nchunk,
                                   int32_t* __restrict__ e) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows_p; r += (int64_t)gridDim.x * blockDim.x) {
        int ex = 0;
        if (r < rows) {
            double mx = 0.0;
            for (int c = 0; c < nchunk; ++c) mx = fmax(mx, part[(int64_t)c * rows + r]);
            if (mx != 0.0) frexp(mx, &ex);
        }
        e[r] = ex;
    }
}

// Ozaki scheme II scaling (Cauchy-Schwarz bound).  Entry bound: |x| = m 2^ex
// (m in [0.5, 1)) counts ceil(256 m)^2 2^(16 - 2d), d = e - ex >= 0 for the
// exponent e of the chunk's maximum, rounded up to an integer (at least 1)
// when 2d > 16: an exact integer upper bound of 2^32 (x 2^-e)^2 whose sums do
// not depend on the summation order.  Computed in FP64 without branches:
// M^2 <= 2^16 scaled by 2^-min(2d - 16, 64) and rounded up is the integer
// exactly (a zero entry gives m = 0, hence 0), and chunk sums of at most
// 64 terms of at most 2^32 stay exact.
__device__ __forceinline__ double crt_bound_term(double x, int e) {
    int ex;
    const double m = frexp(fabs(x), &ex);
    const double M = ceil(m * 256.0);
    return ceil(ldexp(M * M, -min(2 * (e - ex) - 16, 64)));
}

// The scheme II line scan, one read of op(X) per call.  Each line is cut
// into k-chunks of CRT_KC entries; chunk c gets its own max exponent e_c
// (max |x| < 2^e_c) and the exact integer bound S_c = sum crt_bound_term(x,
// e_c) (at most CRT_KC * 2^32 = 2^38), stored as one word
// ((e_c + 2048) << 48) | S_c, or 0 for an all-zero chunk.  crt_shift_kernel
// rescales the chunk bounds to the line's max exponent.  Rows strided in
// memory (ks != 1): a 64-line x 64-k tile through LDS, loaded along the
// lines.  Lines contiguous in k: 16 lanes per line with 16-byte pair loads
// when every line starts 16-byte aligned, else a wave per line.  Every term
// and chunk sum is an integer below 2^39, so the sums are exact in any order.
constexpr int CRT_KC = 64;

__device__ __forceinline__ uint64_t crt_chunk_word(double amax, uint64_t sum) {
    if (amax == 0.0) return 0;
    int ec;
    frexp(amax, &ec);
    return ((uint64_t)(ec + 2048) << 48) | sum;
}

__device__ __forceinline__ int crt_chunk_exp(double amax) {
    int ec = 0;
    if (amax != 0.0) frexp(amax, &ec);
    return ec;
}

// Lines contiguous in k: 16 lanes per line with 16-byte pair loads when every
// line starts 16-byte aligned, else a wave per line.
__device__ __forceinline__ void scan_contiguous(const double* __restrict__ x, int64_t rows, int64_t K, int64_t rs,
                                                int64_t tile, int64_t chunk, uint64_t* __restrict__ part) {
    const int64_t k0 = chunk * CRT_KC;
    if (pair_rows(x, rs)) {
        // 16 lanes per line, two 16-byte pairs each (k0 + 2 l16 and 32 further)
        const int l16 = threadIdx.x & 15;
        const int64_t r = tile * (int64_t)(blockDim.x >> 4) + (threadIdx.x >> 4);
        if (r >= rows) return;  // whole 16-lane groups leave together
        const double* xr = x + r * rs;
        double v[4];
        if (k0 + CRT_KC <= K) {  // the whole chunk inside the line: both loads in flight together
            const pdv2 w0 = __builtin_nontemporal_load(reinterpret_cast<const pdv2*>(xr + k0 + 2 * l16));
            const pdv2 w1 = __builtin_nontemporal_load(reinterpret_cast<const pdv2*>(xr + k0 + 32 + 2 * l16));
            v[0] = w0.x; v[1] = w0.y; v[2] = w1.x; v[3] = w1.y;
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t k = k0 + 32 * h + 2 * l16;
                if (k + 1 < K) {
                    const pdv2 w = *reinterpret_cast<const pdv2*>(xr + k);
                    v[2 * h] = w.x;
                    v[2 * h + 1] = w.y;
                } else {
                    v[2 * h] = k < K ? xr[k] : 0.0;
                    v[2 * h + 1] = 0.0;
                }
            }
        }
        double mx = fmax(fmax(fabs(v[0]), fabs(v[1])), fmax(fabs(v[2]), fabs(v[3])));
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        const int ec = crt_chunk_exp(mx);
        double sum = (crt_bound_term(v[0], ec) + crt_bound_term(v[1], ec)) +
                     (crt_bound_term(v[2], ec) + crt_bound_term(v[3], ec));
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if (l16 == 0) part[chunk * rows + r] = crt_chunk_word(mx, (uint64_t)sum);
    } else {
        const int lane = threadIdx.x & 63;
        const int64_t r = tile * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
        if (r >= rows) return;
        const double* xr = x + r * rs;
        const double v = k0 + lane < K ? xr[k0 + lane] : 0.0;
        double mx = fabs(v);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        const int ec = crt_chunk_exp(mx);
        double sum = crt_bound_term(v, ec);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if (lane == 0) part[chunk * rows + r] = crt_chunk_word(mx, (uint64_t)sum);
    }
}

// Lines strided in memory (ks != 1): a thread per line, consecutive lanes on
// consecutive lines (a wave's load reads 512 contiguous bytes when rs == 1),
// every load of the chunk in flight at once.  Each entry is kept as its
// exponent ex and mantissa bound M = ceil(256 m) (|x| = m 2^ex, M in
// [128, 256], 0 for a zero entry) packed as ((ex + 2048) << 9) | M, so the
// largest packed word of the chunk carries the exponent e of its largest
// entry, and the terms of crt_bound_term follow from (M, ex).  blockIdx.x =
// line tile, blockIdx.y = chunk (the grid walks down the columns).
__global__ __launch_bounds__(256) void crt_scan_lines_kernel(const double* __restrict__ x, int64_t rows, int64_t K,
                                                             int64_t rs, int64_t ks, uint64_t* __restrict__ part) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    const int64_t k0 = (int64_t)blockIdx.y * CRT_KC;
    const int64_t kn = K - k0 < CRT_KC ? K - k0 : CRT_KC;
    const double* xr = x + r * rs + k0 * ks;
    // every load unconditional (a clamped, in-bounds address past the end
    // of the line), so that all of them issue before the first wait
    double raw[CRT_KC];
#pragma unroll
    for (int i = 0; i < CRT_KC; ++i) raw[i] = __builtin_nontemporal_load(xr + (i < kn ? i : kn - 1) * ks);
    uint32_t wd[CRT_KC];
    uint32_t top = 0;
#pragma unroll
    for (int i = 0; i < CRT_KC; ++i) {
        const double v = i < kn ? raw[i] : 0.0;
        int ex;
        const double m = frexp(fabs(v), &ex);
        const uint32_t M = (uint32_t)ceil(m * 256.0);
        wd[i] = M ? ((uint32_t)(ex + 2048) << 9) | M : 0u;
        top = top > wd[i] ? top : wd[i];
    }
    uint64_t word = 0;
    if (top) {
        const int ec = (int)(top >> 9) - 2048;
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < CRT_KC; ++i) {
            const double M = (double)(wd[i] & 511u);
            const int ex = (int)(wd[i] >> 9) - 2048;
            sum += wd[i] ? ceil(ldexp(M * M, -min(2 * (ec - ex) - 16, 64))) : 0.0;
        }
        word = ((uint64_t)(ec + 2048) << 48) | (uint64_t)sum;
    }
    part[(int64_t)blockIdx.y * rows + r] = word;
}

// Lines contiguous in k: blockIdx.x = line tile * nchunk + chunk, so the grid
// walks along the lines (consecutive workgroups read consecutive chunks of
// the same lines).
__global__ __launch_bounds__(256) void crt_scan_rows_kernel(const double* __restrict__ x, int64_t rows, int64_t K,
                                                            int64_t rs, uint32_t nchunk, uint64_t* __restrict__ part) {
    const uint32_t tile = blockIdx.x / nchunk;
    scan_contiguous(x, rows, K, rs, tile, blockIdx.x - tile * nchunk, part);
}

// Line shifts: e = the largest chunk exponent, S = sum over the nonzero
// chunks of ceil(S_c / 4^(e - e_c)), sigma the largest shift with
// 2^(2(sigma + e)) S <= 2^(2Q + 32) (sigma = Q + 16 - e - ceil(log4 S)),
// capped at 53 - e; an all-zero line keeps sigma = 0.  Stored as
// e' = Q - sigma, the form the residue and reconstruction kernels take.
// A workgroup takes 64 lines; its four waves split the chunks (lanes on
// consecutive lines, so every load is coalesced) and combine through LDS.
__global__ __launch_bounds__(256) void crt_shift_kernel(const uint64_t* __restrict__ part, int64_t rows, int64_t rows_p,
                                                        int nchunk, int Q, int32_t* __restrict__ e) {
    __shared__ uint64_t red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * 64 + lane;
    const bool valid = r < rows;
    uint64_t top = 0;
    if (valid)
        for (int c = w; c < nchunk; c += 4) top = max(top, part[(int64_t)c * rows + r] >> 48);
    red[w][lane] = top;
    __syncthreads();
    top = max(max(red[0][lane], red[1][lane]), max(red[2][lane], red[3][lane]));
    __syncthreads();
    uint64_t S = 0;
    if (valid && top != 0)
        for (int c = w; c < nchunk; c += 4) {
            const uint64_t wd = part[(int64_t)c * rows + r];
            if (wd == 0) continue;
            const uint64_t sc = wd & ((1ull << 48) - 1);
            const int d2 = 2 * (int)(top - (wd >> 48));
            S += d2 >= 62 ? 1 : (sc + (1ull << d2) - 1) >> d2;
        }
    red[w][lane] = S;
    __syncthreads();
    if (w != 0 || r >= rows_p) return;
    S = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    int sig = 0;
    if (valid && top != 0) {
        const int bl = S > 1 ? 64 - __builtin_clzll(S - 1) : 0;  // bit length of S - 1
        const int j = (bl + 1) >> 1;                              // ceil(log4 S)
        const int er = (int)top - 2048;
        sig = min(53 - er, Q + 16 - er - j);
    }
    e[r] = Q - sig;
}

// Writes the s digits of op(X)(row, k) into slices[t][row][Kp] (zeros in the
// padding).  A 64x64 tile goes through LDS so that both the read of op(X)
// and the k-contiguous slice writes are coalesced whatever op(X)'s layout.
template <bool BF16>
__global__ void oz_slice_kernel(const double* __restrict__ x, int64_t rows, int64_t K, int64_t rs, int64_t ks,
                                const int32_t* __restrict__ e, int s, int bits, int8_t* __restrict__ out,
                                int64_t Kp, int64_t slice_bytes) {
    __shared__ double tile[64][65];
    const int64_t r0 = (int64_t)blockIdx.y * 64;
    const int64_t k0 = (int64_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;  // 0..3
    // load: along the contiguous direction of op(X)
    if (ks == 1) {
        for (int rr = ty; rr < 64; rr += 4) {
            const int64_t r = r0 + rr, k = k0 + tx;
            tile[rr][tx] = (r < rows && k < K) ? x[r * rs + k] : 0.0;
        }
    } else {
        for (int kk = ty; kk < 64; kk += 4) {
            const int64_t r = r0 + tx, k = k0 + kk;
            tile[tx][kk] = (r < rows && k < K) ? x[r * rs + k * ks] : 0.0;
        }
    }
    __syncthreads();
    const double scale = ldexp(1.0, bits);
    // thread: row t/4 of the tile, 16 consecutive k -> 16 (INT8) or 32 (BF16) bytes per slice
    const int rr = threadIdx.x >> 2;
    const int kq = (threadIdx.x & 3) * 16;
    const int64_t r = r0 + rr;
    double v[16];
    const int sh = r < rows ? -e[r] : 0;
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = r < rows ? ldexp(tile[rr][kq + c], sh) : 0.0;
    int8_t* dst = out + (r * Kp + k0 + kq) * (BF16 ? 2 : 1);
    for (int t = 0; t < s; ++t) {
        uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            v[c] = v[c] * scale;
            const double d = trunc(v[c]);
            v[c] = v[c] - d;
            if (BF16) {
                const float f = (float)d;  // |d| <= 31: exact, low mantissa bits zero
                w[c >> 1] |= (__float_as_uint(f) >> 16) << (16 * (c & 1));
            } else {
                w[c >> 2] |= (uint32_t)(uint8_t)(int8_t)(int)d << (8 * (c & 3));
            }
        }
        iv4* o = reinterpret_cast<iv4*>(dst + t * slice_bytes);
        o[0] = iv4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
        if (BF16) o[1] = iv4{(int)w[4], (int)w[5], (int)w[6], (int)w[7]};
    }
}

constexpr int64_t kMaxChunk = 256;  // k-chunk of the max reduction

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace

// Prepared operand of the CRT form (one side of a product): the balanced
// residues of its scaled lines, [nmod][rows_p][Kp] int8 with rows_p = rows
// rounded up to the tile height and Kp = k rounded up to the k-step (zeros
// in the padding), then the lines' exponents e' = Q - sigma (rows_p int32).
int64_t crt_operand_bytes(int64_t rows, int64_t k, int nmod) {
    const int64_t rp = round_up(rows > 0 ? rows : 1, PAD), kp = round_up(k > 0 ? k : 1, KSB);
    return round_up((int64_t)nmod * rp * kp, 256) + rp * 4;
}

int64_t crt_scan_bytes(int64_t rows, int64_t k) {
    const int64_t nchunk = k > 0 ? (k + CRT_KC - 1) / CRT_KC : 1;
    return nchunk * (rows > 0 ? rows : 1) * 8;
}

int64_t crt_digit_bytes(int64_t m, int64_t n, int nmod) {
    return (int64_t)nmod * round_up(m > 0 ? m : 1, PAD) * round_up(n > 0 ? n : 1, PAD);
}

namespace {

// Scheme I groups whose planes one launch holds (the combine reads them all
// per element): 8 planes of 4 bytes per output element.
constexpr int kPlaneBatch = 8;

struct Plan {
    int64_t Mp, Np, Kp, esz, a_slice, b_slice, res_plane, nchunk_a, nchunk_b, kchunk_max;
    int64_t plane_batch;  // scheme I: groups per product launch
    size_t off_a, off_b, off_ea, off_eb, off_part, off_w, off_planes, total;
};

// Host side of the CRT form: q and the reconstruction tables for n moduli.
int crt_setup(int64_t k, int nmod, CrtTables* tb, int* q_out) {
    unsigned __int128 P = 1;
    for (int t = 0; t < nmod; ++t) P *= (unsigned)crt_modulus(t);
    // largest Q with 2^(2Q) + ceil(sqrt(k)) 2^Q + k <= P/4: lines of 2-norm
    // <= 2^Q before rounding keep |A'B'| <= P/4 (Cauchy-Schwarz)
    const uint64_t kk = (uint64_t)(k > 0 ? k : 0);
    uint64_t rk = (uint64_t)sqrt((double)kk);
    while (rk * rk > kk) --rk;
    while ((rk + 1) * (rk + 1) <= kk) ++rk;
    if (rk * rk < kk) ++rk;  // ceil(sqrt(k))
    int q = 62;
    while (q >= 1) {
        const unsigned __int128 need = ((unsigned __int128)1 << (2 * q)) + ((unsigned __int128)rk << q) + kk;
        if (need <= (P >> 2)) break;
        --q;
    }
    if (q < 1) return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: %d moduli leave no bits for k=%lld", nmod, (long long)k);
    *q_out = q;
    tb->n = nmod;
    for (int t = 0; t < nmod; ++t) {
        const int p = crt_modulus(t);
        const unsigned __int128 M = P / (unsigned)p;
        for (int l = 0; l < 3; ++l) tb->mt[t][l] = (double)(uint64_t)((M >> (40 * l)) & (((unsigned __int128)1 << 40) - 1));
        // inverse of M mod p by extended Euclid (M and p are coprime)
        int64_t a = (int64_t)(M % (unsigned)p), b = p, x0 = 1, x1 = 0;
        while (b) {
            const int64_t qq = a / b, tmp = a - qq * b;
            a = b; b = tmp;
            const int64_t tx = x0 - qq * x1;
            x0 = x1; x1 = tx;
        }
        tb->cinv[t] = (int32_t)(((x0 % p) + p) % p);
    }
    for (int l = 0; l < 3; ++l) tb->Pl[l] = (double)(uint64_t)((P >> (40 * l)) & (l < 2 ? (((unsigned __int128)1 << 40) - 1) : ~(unsigned __int128)0));
    tb->invPd = 1.0 / (double)P;
    return DG_OK;
}

}  // namespace

int crt_moduli_for_fp64(int64_t k) {
    if (k >= ((int64_t)1 << 19)) return kMaxModuli;
    unsigned __int128 P = 1;
    const unsigned __int128 need = (unsigned __int128)(k > 0 ? k : 1) << (2 * 53 + 2);
    for (int t = 0; t < kMaxModuli; ++t) {
        P *= (unsigned)crt_modulus(t);
        if (t >= 1 && need <= P) return t + 1;
    }
    return kMaxModuli;
}

namespace {

// Workspace of one CRT call: the prepared A and B operands (crt_operand_bytes
// each), the line-scan words, then the CRT digits of V.
Plan make_crt_plan(int64_t m, int64_t n, int64_t k, int nmod) {
    Plan p{};
    p.esz = 1;
    p.Mp = round_up(m > 0 ? m : 1, PAD);
    p.Np = round_up(n > 0 ? n : 1, PAD);
    p.Kp = round_up(k > 0 ? k : 1, KSB);
    p.a_slice = p.Mp * p.Kp;
    p.b_slice = p.Np * p.Kp;
    p.res_plane = p.Mp * p.Np;
    p.kchunk_max = CRT_KC;  // the line scan's k-chunk (k < 2^17 keeps grid.y below 2^11)
    p.nchunk_a = (k + CRT_KC - 1) / CRT_KC;
    if (p.nchunk_a < 1) p.nchunk_a = 1;
    p.nchunk_b = p.nchunk_a;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = (off + bytes + 255) / 256 * 256; return o; };
    p.off_a = take((size_t)crt_operand_bytes(m, k, nmod));
    p.off_b = take((size_t)crt_operand_bytes(n, k, nmod));
    p.off_part = take((size_t)crt_scan_bytes(m > n ? m : n, k));
    p.off_w = take((size_t)crt_digit_bytes(m, n, nmod));
    p.total = off;
    return p;
}

std::vector<OzSeg> make_schedule(int s, int bits, int64_t Kp, int64_t K, bool bf16, int* err);

Plan make_plan(int64_t m, int64_t n, int64_t k, int s, bool bf16) {
    Plan p{};
    p.esz = bf16 ? 2 : 1;
    p.Mp = round_up(m > 0 ? m : 1, PAD);
    p.Np = round_up(n > 0 ? n : 1, PAD);
    p.Kp = round_up(k > 0 ? k : 1, KSB / p.esz);
    p.a_slice = p.Mp * p.Kp * p.esz;
    p.b_slice = p.Np * p.Kp * p.esz;
    p.kchunk_max = kMaxChunk;
    while ((k + p.kchunk_max - 1) / p.kchunk_max > 4096) p.kchunk_max *= 2;  // keep grid.y small
    p.nchunk_a = (k + p.kchunk_max - 1) / p.kchunk_max;
    if (p.nchunk_a < 1) p.nchunk_a = 1;
    p.nchunk_b = p.nchunk_a;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = (off + bytes + 255) / 256 * 256; return o; };
    p.off_a = take((size_t)s * p.a_slice);
    p.off_b = take((size_t)s * p.b_slice);
    p.off_ea = take((size_t)p.Mp * 4);
    p.off_eb = take((size_t)p.Np * 4);
    p.off_part = take((size_t)p.nchunk_a * (size_t)(m > n ? m : n) * 8);
    p.off_w = take((size_t)p.Mp * (size_t)p.Np * 8);
    {  // group planes
        int err = 0;
        int64_t groups = 0;
        for (const OzSeg& g : make_schedule(s, bf16 ? 5 : 7, p.Kp, k, bf16, &err)) groups += g.flush != 0;
        p.plane_batch = groups < kPlaneBatch ? (groups > 0 ? groups : 1) : kPlaneBatch;
        p.off_planes = take((size_t)p.plane_batch * (size_t)p.Mp * (size_t)p.Np * 4);
    }
    p.total = off;
    return p;
}

// Product schedule in the oracle's order: level, k-chunk, t.
std::vector<OzSeg> make_schedule(int s, int bits, int64_t Kp, int64_t K, bool bf16, int* err) {
    const int64_t dmax = (1 << bits) - 1;
    const int64_t kstep = KSB / (bf16 ? 2 : 1);
    int64_t kchunk = Kp;
    int64_t flush;
    if (bf16) {
        // fp32 sums stay exact below 2^24: chunk K so that one product fits
        const int64_t cap = ((int64_t)1 << 24) / (dmax * dmax);
        if (kchunk > cap) kchunk = cap / kstep * kstep;
        const int64_t keff = K < kchunk ? K : kchunk;
        flush = ((int64_t)1 << 24) / (dmax * dmax * (keff > 0 ? keff : 1));
    } else {
        flush = (int64_t)INT32_MAX / (dmax * dmax * (K > 0 ? K : 1));
    }
    *err = (flush < 1 || kchunk < kstep) ? 1 : 0;
    std::vector<OzSeg> segs;
    for (int l = 2; l <= s + 1; ++l)
        for (int64_t k0 = 0; k0 < Kp; k0 += kchunk) {
            const int64_t kl = (Kp - k0) < kchunk ? (Kp - k0) : kchunk;
            int cnt = 0;
            for (int t = 1; t <= l - 1; ++t) {
                OzSeg g{};
                g.t = t - 1;
                g.u = l - t - 1;
                g.k0 = (int32_t)k0;
                g.klen = (int32_t)kl;
                g.level = l;
                ++cnt;
                g.flush = (cnt == flush || t == l - 1) ? 1 : 0;
                if (g.flush) cnt = 0;
                segs.push_back(g);
            }
        }
    // Mark the first and the last flush (W is written, not read, by the first;
    // the last one writes C).
    bool seen = false;
    for (auto& g : segs)
        if (g.flush) {
            if (!seen) { g.flush = 2; seen = true; }
        }
    for (auto it = segs.rbegin(); it != segs.rend(); ++it)
        if (it->flush) { it->flush = it->flush == 2 ? 4 : 3; break; }
    return segs;
}

// The schedule's products cut at its flushes: each group is consecutive
// products of one level and one k range with t rising and u falling by one.
// Returns false if the schedule does not have that form.
bool make_groups(const std::vector<OzSeg>& sched, int64_t esz, std::vector<OzGroup>& out) {
    out.clear();
    for (size_t i = 0; i < sched.size();) {
        size_t j = i;
        while (j < sched.size() && !sched[j].flush) ++j;
        if (j == sched.size()) return false;
        const OzSeg& f = sched[i];
        for (size_t q = i; q <= j; ++q) {
            const OzSeg& x = sched[q];
            if (x.level != f.level || x.k0 != f.k0 || x.klen != f.klen || x.t != f.t + (int)(q - i) ||
                x.u != f.u - (int)(q - i))
                return false;
        }
        if ((f.klen * esz) % CK2 != 0 || (f.k0 * esz) % CK2 != 0 || f.klen <= 0) return false;
        out.push_back(OzGroup{f.t, f.u, (int32_t)(j - i + 1), f.level, (int32_t)(f.k0 * esz),
                              make_fastdiv((uint32_t)(f.klen * esz / CK2))});
        if ((int64_t)out.back().count * out.back().nks.d > INT32_MAX) return false;
        i = j + 1;
    }
    return !out.empty();
}

}  // namespace

// Scheme II, one operand: line scan, shifts and residues into `operand`
// (crt_operand_bytes(rows, k, nmod) bytes, 256-byte aligned) for lines
// x[i*rs + k*ks], using `scan` (crt_scan_bytes(rows, k) bytes).  The result
// depends only on the operand, k and nmod, so a prepared operand serves
// every product that reads it.
int crt_prepare(const double* x, int64_t rows, int64_t k, int64_t rs, int64_t ks, int nmod, void* operand,
                void* scan, hipStream_t st) {
    CrtTables tables{};
    int q = 0;
    int rc = crt_setup(k, nmod, &tables, &q);
    if (rc) return rc;
    const int64_t rp = round_up(rows, PAD), kp = round_up(k, KSB);
    const int nchunk = (int)((k + CRT_KC - 1) / CRT_KC);
    int8_t* slices = static_cast<int8_t*>(operand);
    int32_t* e = reinterpret_cast<int32_t*>(static_cast<char*>(operand) + round_up((int64_t)nmod * rp * kp, 256));
    uint64_t* bp = static_cast<uint64_t*>(scan);
    // Line exponents and Cauchy-Schwarz shifts: one scan (chunk exponents
    // and chunk bounds), then the per-line combination.  Strided lines: a
    // thread per line; contiguous lines: 16 lanes per line when the lines
    // pair up in 16-byte loads, else a wave per line.
    if (ks != 1) {
        hipLaunchKernelGGL(crt_scan_lines_kernel, dim3((unsigned)((rows + 255) / 256), (unsigned)nchunk), dim3(NT), 0,
                           st, x, rows, k, rs, ks, bp);
    } else {
        const int64_t per = pair_rows(x, rs) ? NT / 16 : NT / 64;
        const int64_t blocks = (rows + per - 1) / per * nchunk;
        if (blocks > INT32_MAX) return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: too many lines for one scan");
        hipLaunchKernelGGL(crt_scan_rows_kernel, dim3((unsigned)blocks), dim3(NT), 0, st, x, rows, k, rs,
                           (uint32_t)nchunk, bp);
    }
    hipLaunchKernelGGL(crt_shift_kernel, dim3((unsigned)((rp + 63) / 64)), dim3(NT), 0, st, bp, rows, rp, nchunk, q, e);
    const uint32_t nkt = (uint32_t)(kp / RT_K), nrt = (uint32_t)(rp / RT_R);
    if ((int64_t)nkt * nrt > INT32_MAX) return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: too many residue tiles");
    decltype(&crt_residue_kernel<16>) rsk = crt_residue_kernel<16>;
    switch (nmod) {
#define DG_CRT_CASE(N) case N: rsk = crt_residue_kernel<N>; break;
        DG_CRT_CASE(2) DG_CRT_CASE(3) DG_CRT_CASE(4) DG_CRT_CASE(5) DG_CRT_CASE(6) DG_CRT_CASE(7)
        DG_CRT_CASE(8) DG_CRT_CASE(9) DG_CRT_CASE(10) DG_CRT_CASE(11) DG_CRT_CASE(12) DG_CRT_CASE(13)
        DG_CRT_CASE(14) DG_CRT_CASE(15)
#undef DG_CRT_CASE
        default: break;
    }
    ResidueConsts rcs{};
    for (int t = 0; t < kMaxModuli; ++t) {
        const int pm = crt_modulus(t);
        rcs.c24[t] = (double)((1 << 24) % pm);
        rcs.inv_p[t] = 1.0 / (double)pm;
        rcs.pd[t] = (double)pm;
    }
    hipLaunchKernelGGL(rsk, dim3(nkt * nrt), dim3(256), 0, st, x, rows, k, rs, ks, e, q, slices, kp, rp * kp, rcs, nkt,
                       nrt);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

// Scheme II, the product of two prepared operands (m lines of op(A), n lines
// of op(B)^T, both over k with nmod moduli): the nmod INT8 GEMMs into the CRT
// digits (`digits`, crt_digit_bytes(m, n, nmod) bytes), then
// C = alpha * V * 2^(e'_i + f'_j - 2Q) + beta * C.
int crt_multiply(int64_t m, int64_t n, int64_t k, int nmod, double alpha, const void* a_op, const void* b_op,
                 double beta, double* C, int64_t ldc, void* digits, hipStream_t st) {
    CrtTables tables{};
    int q = 0;
    int rc = crt_setup(k, nmod, &tables, &q);
    if (rc) return rc;
    const int64_t mp = round_up(m, PAD), np = round_up(n, PAD), kp = round_up(k, KSB);
    CrtArgs cg{};
    cg.a = static_cast<const int8_t*>(a_op);
    cg.b = static_cast<const int8_t*>(b_op);
    cg.res = static_cast<uint8_t*>(digits);
    cg.Mp = mp; cg.Np = np; cg.Kp = kp; cg.a_slice = mp * kp; cg.b_slice = np * kp;
    cg.res_plane = mp * np;
    cg.tiles_m = (int32_t)(mp / BM); cg.tiles_n = (int32_t)(np / BN); cg.nmod = nmod;
    for (int t = 0; t < kMaxModuli; ++t) cg.cinv[t] = tables.cinv[t];
    const int32_t* ea = reinterpret_cast<const int32_t*>(static_cast<const char*>(a_op) + round_up((int64_t)nmod * mp * kp, 256));
    const int32_t* eb = reinterpret_cast<const int32_t*>(static_cast<const char*>(b_op) + round_up((int64_t)nmod * np * kp, 256));
    const int64_t grid = (int64_t)cg.tiles_m * cg.tiles_n * nmod;
    if (grid > INT32_MAX) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: too many tiles");
    const int64_t quads = ((m + 3) / 4) * n;
    if (quads > (int64_t)UINT32_MAX - 255)
        return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: too many output elements for one reconstruction");
    hipLaunchKernelGGL(crt_gemm_kernel, dim3((unsigned)grid), dim3(CNT), kCrtLds, st, cg);
    DG_LAUNCH_CHECK();
    decltype(&crt_reconstruct_kernel<16>) rk = crt_reconstruct_kernel<16>;
    switch (nmod) {
#define DG_CRT_CASE(N) case N: rk = crt_reconstruct_kernel<N>; break;
        DG_CRT_CASE(2) DG_CRT_CASE(3) DG_CRT_CASE(4) DG_CRT_CASE(5) DG_CRT_CASE(6) DG_CRT_CASE(7)
        DG_CRT_CASE(8) DG_CRT_CASE(9) DG_CRT_CASE(10) DG_CRT_CASE(11) DG_CRT_CASE(12) DG_CRT_CASE(13)
        DG_CRT_CASE(14) DG_CRT_CASE(15)
#undef DG_CRT_CASE
        default: break;
    }
    hipLaunchKernelGGL(rk, dim3((unsigned)((quads + 255) / 256)), dim3(256), 0, st, cg.res, tables, m, n, mp,
                       cg.res_plane, ea, eb, q, alpha, beta, C, ldc, make_fastdiv((uint32_t)((m + 3) / 4)));
    DG_LAUNCH_CHECK();
    return DG_OK;
}

}  // namespace dg

extern "C" {

int64_t dg_ozaki_workspace(int64_t m, int64_t n, int64_t k, int slices, int kind) {
    if (m < 0 || n < 0 || k < 0 || slices < 1 || slices > 16) return -1;
    if (kind == DG_OZAKI_INT8_CRT) {
        if (slices < 2) return -1;
        return (int64_t)dg::make_crt_plan(m, n, k, slices).total;
    }
    if (kind != DG_OZAKI_INT8 && kind != DG_OZAKI_BF16) return -1;
    return (int64_t)dg::make_plan(m, n, k, slices, kind == DG_OZAKI_BF16).total;
}

int dg_ozaki_dgemm(char opA, char opB, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                   const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int slices, int kind,
                   void* workspace, int64_t workspace_bytes, double* err_bound, void* stream) {
    using namespace dg;
    const char oa = norm_op(opA), ob = norm_op(opB);
    if (!oa || !ob) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: invalid op '%c%c'", opA, opB);
    if (kind != DG_OZAKI_INT8 && kind != DG_OZAKI_BF16 && kind != DG_OZAKI_INT8_CRT)
        return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: kind %d", kind);
    if (slices < 1 || slices > 16) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: slices %d (1..16)", slices);
    const bool crt = kind == DG_OZAKI_INT8_CRT;
    if (crt && slices < 2) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: the CRT form needs 2..16 moduli");
    if (m < 0 || n < 0 || k < 0) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: negative dimension");
    const int64_t rowsA = oa == 'N' ? m : k, rowsB = ob == 'N' ? k : n;
    if (lda < (rowsA > 1 ? rowsA : 1) || ldb < (rowsB > 1 ? rowsB : 1) || ldc < (m > 1 ? m : 1))
        return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: leading dimension too small");
    const bool bf16 = kind == DG_OZAKI_BF16;
    const int bits = bf16 ? 5 : 7;
    CrtTables tables{};
    int q = 0;
    if (crt) {
        if (k > 131071) return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: CRT form needs k < 131072 (exact int32 sums)");
        int rc0 = crt_setup(k, slices, &tables, &q);
        if (rc0) return rc0;
        // every line's shift reaches sigma + e >= qe = min(53, Q - ceil(log4 k)) (its
        // bound sum is at most k 2^32), so |A - 2^-sigma A'| <= 2^(e - qe - 1) per entry
        // and the product error is below k * 2^(1-qe) * 2^(e_i+f_j)
        int l4 = 0;
        while (((int64_t)1 << (2 * l4)) < k) ++l4;
        const int qe = q - l4 < 53 ? q - l4 : 53;
        if (err_bound) *err_bound = (double)k * ldexp(1.0, 1 - qe);
    } else if (err_bound) {
        *err_bound = (double)k * (slices + 3) * ldexp(1.0, -bits * slices);
    }
    if (m == 0 || n == 0) return DG_OK;
    int rc = require_device();
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    if (k == 0 || alpha == 0.0) return launch_scale(m, n, beta, C, ldc, st);
    if (!A || !B || !C) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: null pointer");
    if (k > INT32_MAX / 2) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: k too large");
    const Plan p = crt ? make_crt_plan(m, n, k, slices) : make_plan(m, n, k, slices, bf16);
    if (p.Mp / 32 > 65535 || p.Np / 32 > 65535 || p.Kp / 64 > INT32_MAX)
        return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: matrix too large for one slicing launch");
    if (!workspace || workspace_bytes < (int64_t)p.total)
        return fail(DG_ERR_WORKSPACE, "dg_ozaki_dgemm: workspace %lld < %zu bytes", (long long)workspace_bytes, p.total);
    if (reinterpret_cast<uintptr_t>(workspace) & 255)
        return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: workspace must be 256-byte aligned");
    int serr = 0;
    std::vector<OzSeg> sched;
    if (!crt) {
        sched = make_schedule(slices, bits, p.Kp, k, bf16, &serr);
        if (serr)
            return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: k=%lld too large for exact 32-bit slice sums", (long long)k);
    }

    char* ws = static_cast<char*>(workspace);
    int8_t* sa = reinterpret_cast<int8_t*>(ws + p.off_a);
    int8_t* sb = reinterpret_cast<int8_t*>(ws + p.off_b);
    int32_t* ea = reinterpret_cast<int32_t*>(ws + p.off_ea);
    int32_t* eb = reinterpret_cast<int32_t*>(ws + p.off_eb);
    double* part = reinterpret_cast<double*>(ws + p.off_part);

    // op(A)(i, k) at A[i*rs + k*ks]; op(B) as the rows of its transpose: B^T(j, k).
    const int64_t a_rs = oa == 'N' ? 1 : lda, a_ks = oa == 'N' ? lda : 1;
    const int64_t b_rs = ob == 'N' ? ldb : 1, b_ks = ob == 'N' ? 1 : ldb;
    const int nchunk = (int)p.nchunk_a;
    const unsigned gy = (unsigned)nchunk;

    if (crt) {
        // op(A)'s rows and op(B)'s columns prepared in the workspace, then their product
        rc = crt_prepare(A, m, k, a_rs, a_ks, slices, ws + p.off_a, ws + p.off_part, st);
        if (!rc) rc = crt_prepare(B, n, k, b_rs, b_ks, slices, ws + p.off_b, ws + p.off_part, st);
        if (!rc) rc = crt_multiply(m, n, k, slices, alpha, ws + p.off_a, ws + p.off_b, beta, C, ldc, ws + p.off_w, st);
        return rc;
    }
    {
        // Row exponents of op(A), column exponents of op(B).
        // ks == 1: 16 lanes per row when the rows pair up (NT/16 rows per block), else
        // a wave per row (NT/64); otherwise a thread per row
        auto pm_grid = [](int64_t rows, int64_t ks, const double* x, int64_t rs) {
            const int64_t per = ks == 1 ? (pair_rows(x, rs) ? NT / 16 : NT / 64) : NT;
            const int64_t g = (rows + per - 1) / per;
            return (unsigned)(g < 65535 ? g : 65535);
        };
        unsigned gx = pm_grid(m, a_ks, A, a_rs);
        hipLaunchKernelGGL(oz_partial_max_kernel, dim3(gx, gy), dim3(NT), 0, st, A, m, k, a_rs, a_ks, p.kchunk_max, part);
        DG_LAUNCH_CHECK();
        hipLaunchKernelGGL(oz_exponent_kernel, dim3((unsigned)((p.Mp + NT - 1) / NT)), dim3(NT), 0, st, part, m, p.Mp,
                           nchunk, ea);
        DG_LAUNCH_CHECK();
        gx = pm_grid(n, b_ks, B, b_rs);
        hipLaunchKernelGGL(oz_partial_max_kernel, dim3(gx, gy), dim3(NT), 0, st, B, n, k, b_rs, b_ks, p.kchunk_max, part);
        DG_LAUNCH_CHECK();
        hipLaunchKernelGGL(oz_exponent_kernel, dim3((unsigned)((p.Np + NT - 1) / NT)), dim3(NT), 0, st, part, n, p.Np,
                           nchunk, eb);
        DG_LAUNCH_CHECK();
    }
    // Slices (the padding is written as zeros by the same kernels).
    {
        dim3 ga((unsigned)(p.Kp / 64), (unsigned)(p.Mp / 64));
        dim3 gb((unsigned)(p.Kp / 64), (unsigned)(p.Np / 64));
        if (bf16) {
            hipLaunchKernelGGL(oz_slice_kernel<true>, ga, dim3(NT), 0, st, A, m, k, a_rs, a_ks, ea, slices, bits, sa,
                               p.Kp, p.a_slice);
            hipLaunchKernelGGL(oz_slice_kernel<true>, gb, dim3(NT), 0, st, B, n, k, b_rs, b_ks, eb, slices, bits, sb,
                               p.Kp, p.b_slice);
        } else {
            hipLaunchKernelGGL(oz_slice_kernel<false>, ga, dim3(NT), 0, st, A, m, k, a_rs, a_ks, ea, slices, bits, sa,
                               p.Kp, p.a_slice);
            hipLaunchKernelGGL(oz_slice_kernel<false>, gb, dim3(NT), 0, st, B, n, k, b_rs, b_ks, eb, slices, bits, sb,
                               p.Kp, p.b_slice);
        }
        DG_LAUNCH_CHECK();
    }
    // INT8 products: each launch computes up to plane_batch groups into
    // planes, then the combine adds them into W (or, after the last group, C).
    std::vector<OzGroup> groups;
    if (!make_groups(sched, p.esz, groups))
        return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: internal error: product schedule without group form");
    TableUseScope uses{st};  // after the launches below return
    const OzGroup* dgroups = static_cast<const OzGroup*>(cached_table(groups.data(), groups.size() * sizeof(OzGroup), st));
    if (!dgroups) return DG_ERR_ALLOC;
    const int ng = (int)groups.size();
    const int64_t tiles = (p.Mp / BM) * (p.Np / BN);
    if (tiles * p.plane_batch > INT32_MAX || (p.Mp / 4) * p.Np > UINT32_MAX)
        return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: too many tiles");
    OzPlaneArgs g{};
    g.a = sa; g.b = sb; g.planes = reinterpret_cast<int32_t*>(ws + p.off_planes); g.groups = dgroups;
    g.tiles_m = (int32_t)(p.Mp / BM); g.tiles_n = (int32_t)(p.Np / BN);
    g.Mp = p.Mp; g.rb = p.Kp * p.esz; g.a_slice = p.a_slice; g.b_slice = p.b_slice; g.plane = p.Mp * p.Np;
    double* W = reinterpret_cast<double*>(ws + p.off_w);
    const uint32_t total = (uint32_t)((p.Mp / 4) * p.Np);
    const FastDiv dqm = make_fastdiv((uint32_t)(p.Mp / 4));
    for (int g0 = 0; g0 < ng; g0 += (int)p.plane_batch) {
        const int cnt = ng - g0 < (int)p.plane_batch ? ng - g0 : (int)p.plane_batch;
        g.g0 = g0;
        const dim3 grid((unsigned)(tiles * cnt));
        const dim3 cgrid((total + 255) / 256);
        if (bf16) {
            hipLaunchKernelGGL(oz_plane_kernel<true>, grid, dim3(CNT), kCrtLds, st, g);
            hipLaunchKernelGGL(oz_combine_kernel<true>, cgrid, dim3(256), 0, st, g.planes, g.plane, dgroups, g0,
                               g0 + cnt, ng, bits, W, p.Mp, m, n, ea, eb, alpha, beta, C, ldc, dqm, total);
        } else {
            hipLaunchKernelGGL(oz_plane_kernel<false>, grid, dim3(CNT), kCrtLds, st, g);
            hipLaunchKernelGGL(oz_combine_kernel<false>, cgrid, dim3(256), 0, st, g.planes, g.plane, dgroups, g0,
                               g0 + cnt, ng, bits, W, p.Mp, m, n, ea, eb, alpha, beta, C, ldc, dqm, total);
        }
        DG_LAUNCH_CHECK();
    }
    return DG_OK;
}

}  // extern "C"
