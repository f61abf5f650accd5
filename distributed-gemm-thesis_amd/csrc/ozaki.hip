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
//      into an FP64 register accumulator as ldexp(acc, -b*l) whenever it
//      could overflow next (the host computes the flush points) and at the
//      end of each level.
//   4. C = alpha * ldexp(acc, e_i + f_j) + beta * C, without fused multiply-add
//      so that the CPU oracle reproduces every bit.
//
// Slices are stored k-contiguous, [t][row][Kp] with rows padded to 128 and K
// padded to one 128-byte k-step with zeros, so the GEMM kernel has no edge
// cases until the final store.  The GEMM kernel uses 128x128 output tiles,
// 4 waves of 64x64, 128-byte k-steps double buffered in LDS with a 144-byte
// row stride (conflict-free 128-bit fragment reads).
#include "dg_internal.hpp"

// Every FP64 operation of the recombination is rounded on its own (no fused
// multiply-add), matching the CPU oracle bit for bit.
#pragma clang fp contract(off)

namespace dg {

namespace {

typedef int iv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));
typedef double dv4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int OT = 128;          // output tile (rows and columns)
constexpr int KSB = 128;         // k-step in bytes
constexpr int LDB = KSB + 16;    // LDS row stride in bytes
constexpr int NT = 256;
constexpr int GROUP_M = 8;

typedef __attribute__((address_space(1))) const iv4 giv4;

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
    int32_t flush;      // add the accumulator into FP64 after this product
};

struct OzArgs {
    const int8_t* a;      // A slices, [s][Mp][Kp] elements
    const int8_t* b;      // B slices, [s][Np][Kp]
    const int32_t* ea;    // row exponents (Mp)
    const int32_t* eb;    // column exponents (Np)
    const OzSeg* segs;
    int32_t nseg;
    int32_t bits;
    int64_t m, n, Mp, Np, Kp;   // Kp in elements
    int64_t a_slice, b_slice;   // bytes per slice
    double alpha, beta;
    double* c;
    int64_t ldc;
    int32_t tiles_m, tiles_n;
};

template <bool BF16>
__device__ __forceinline__ void mma(iv4& acc_i, fv4& acc_f, const iv4& x, const iv4& y) {
    if (BF16) {
        acc_f = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, x), __builtin_bit_cast(bf8, y), acc_f,
                                                        0, 0, 0);
    } else {
        acc_i = __builtin_amdgcn_mfma_i32_16x16x64_i8(x, y, acc_i, 0, 0, 0);
    }
}

// Loads one 128-row x 128-byte k-step: thread t takes 16-byte chunk t%8 of
// rows t/8 + 32*q.
__device__ __forceinline__ void oz_load(iv4 (&r)[4], const int8_t* __restrict__ base, int64_t row_bytes,
                                        int64_t row0, int64_t kbyte0) {
    const int t = threadIdx.x;
    const int ch = t & 7;
    const int rr = t >> 3;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int8_t* p = base + (row0 + rr + 32 * q) * row_bytes + kbyte0 + ch * 16;
        r[q] = *(giv4*)p;
    }
}

__device__ __forceinline__ void oz_store(int8_t* __restrict__ s, const iv4 (&r)[4]) {
    const int t = threadIdx.x;
    const int ch = t & 7;
    const int rr = t >> 3;
#pragma unroll
    for (int q = 0; q < 4; ++q) *reinterpret_cast<iv4*>(s + (rr + 32 * q) * LDB + ch * 16) = r[q];
}

template <bool BF16>
__global__ __launch_bounds__(NT, 1) void oz_gemm_kernel(OzArgs g) {
    __shared__ __attribute__((aligned(16))) int8_t As[2][OT * LDB];
    __shared__ __attribute__((aligned(16))) int8_t Bs[2][OT * LDB];

    const int total = g.tiles_m * g.tiles_n;
    const int tile = xcd_remap(blockIdx.x, total);
    const int per_group = GROUP_M * g.tiles_n;
    const int group = tile / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(g.tiles_m - first_m, GROUP_M);
    const int in_group = tile - group * per_group;
    const int64_t m0 = (int64_t)(first_m + in_group % gsize) * OT;
    const int64_t n0 = (int64_t)(in_group / gsize) * OT;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 64;
    const int wn = (wave & 1) * 64;
    const int lr = lane & 15;
    const int lg = lane >> 4;

    const int esz = BF16 ? 2 : 1;
    const int64_t row_bytes = g.Kp * esz;

    dv4 accd[4][4];
    iv4 acci[4][4];
    fv4 accf[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            accd[i][j] = dv4{0.0, 0.0, 0.0, 0.0};
            acci[i][j] = iv4{0, 0, 0, 0};
            accf[i][j] = fv4{0.f, 0.f, 0.f, 0.f};
        }

    int sidx = 0;
    OzSeg cur = g.segs[0];
    int64_t kb = 0;  // byte offset within the current segment
    const int64_t seg_bytes0 = (int64_t)cur.klen * esz;
    (void)seg_bytes0;

    iv4 ra[4], rb[4];
    oz_load(ra, g.a + cur.t * g.a_slice, row_bytes, m0, (int64_t)cur.k0 * esz);
    oz_load(rb, g.b + cur.u * g.b_slice, row_bytes, n0, (int64_t)cur.k0 * esz);
    int buf = 0;
    oz_store(As[0], ra);
    oz_store(Bs[0], rb);
    __syncthreads();

    while (true) {
        int64_t nkb = kb + KSB;
        int ns = sidx;
        OzSeg nxt = cur;
        if (nkb >= (int64_t)cur.klen * esz) {
            ++ns;
            nkb = 0;
            if (ns < g.nseg) nxt = g.segs[ns];
        }
        const bool has_next = ns < g.nseg;
        if (has_next) {
            oz_load(ra, g.a + nxt.t * g.a_slice, row_bytes, m0, (int64_t)nxt.k0 * esz + nkb);
            oz_load(rb, g.b + nxt.u * g.b_slice, row_bytes, n0, (int64_t)nxt.k0 * esz + nkb);
        }

        const int8_t* as = As[buf];
        const int8_t* bs = Bs[buf];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            iv4 af[4], bfr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *reinterpret_cast<const iv4*>(as + (wm + 16 * i + lr) * LDB + h * 64 + lg * 16);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bfr[j] = *reinterpret_cast<const iv4*>(bs + (wn + 16 * j + lr) * LDB + h * 64 + lg * 16);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) mma<BF16>(acci[i][j], accf[i][j], bfr[j], af[i]);
        }

        // End of the current product: flush the 32-bit sums into FP64.
        const bool seg_end = (kb + KSB >= (int64_t)cur.klen * esz);
        if (seg_end && cur.flush) {
            const double w = ldexp(1.0, -g.bits * cur.level);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double x = BF16 ? (double)accf[i][j][v] : (double)acci[i][j][v];
                        accd[i][j][v] = __dadd_rn(accd[i][j][v], x * w);  // x*w exact (power of two)
                    }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acci[i][j] = iv4{0, 0, 0, 0};
                    accf[i][j] = fv4{0.f, 0.f, 0.f, 0.f};
                }
        }

        if (!has_next) break;
        buf ^= 1;
        oz_store(As[buf], ra);
        oz_store(Bs[buf], rb);
        __syncthreads();
        sidx = ns;
        kb = nkb;
        cur = nxt;
    }

    // Epilogue.  16x16 i32/f32 result layout: lane l, register v holds
    // D(4*(l/16) + v, l%16); D = B^T A^T, so acc[i][j][v] is
    // C(m0 + wm + 16i + lr, n0 + wn + 16j + 4lg + v).
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t m = m0 + wm + 16 * i + lr;
        if (m >= g.m) continue;
        const int ea = g.ea[m];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t n = n0 + wn + 16 * j + 4 * lg + v;
                if (n < g.n) {
                    double* cp = g.c + m + n * g.ldc;
                    double val = __dmul_rn(g.alpha, ldexp(accd[i][j][v], ea + g.eb[n]));
                    if (g.beta != 0.0) val = __dadd_rn(val, __dmul_rn(g.beta, *cp));
                    *cp = val;
                }
            }
        }
    }
}

// Row maxima of op(X): partial max |x| over a k-chunk for every row.
// Element (row, k) of op(X) is at x[row*rs + k*ks].
__global__ void oz_partial_max_kernel(const double* __restrict__ x, int64_t rows, int64_t K, int64_t rs, int64_t ks,
                                      int64_t kchunk, double* __restrict__ part) {
    // blockIdx.y: k-chunk; threads stride over rows, or over k when rows are
    // contiguous in k (ks == 1), with a block reduction.
    const int64_t k0 = blockIdx.y * kchunk;
    const int64_t k1 = min(K, k0 + kchunk);
    if (ks != 1) {
        for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
            double mx = 0.0;
            for (int64_t k = k0; k < k1; ++k) mx = fmax(mx, fabs(x[r * rs + k * ks]));
            part[blockIdx.y * rows + r] = mx;
        }
    } else {
        __shared__ double red[NT];
        for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
            double mx = 0.0;
            for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) mx = fmax(mx, fabs(x[r * rs + k]));
            red[threadIdx.x] = mx;
            __syncthreads();
            for (int w = NT / 2; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
                __syncthreads();
            }
            if (threadIdx.x == 0) part[blockIdx.y * rows + r] = red[0];
            __syncthreads();
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
    for (int rr = ty; rr < 64; rr += 4) {
        const int64_t r = r0 + rr, k = k0 + tx;
        double v = (r < rows) ? ldexp(tile[rr][tx], -e[r]) : 0.0;
        for (int t = 0; t < s; ++t) {
            v = v * scale;
            const double d = trunc(v);
            v = v - d;
            if (BF16) {
                const float f = (float)d;  // |d| <= 31: exact, low mantissa bits zero
                const uint16_t hb = (uint16_t)(__float_as_uint(f) >> 16);
                reinterpret_cast<uint16_t*>(out + t * slice_bytes)[r * Kp + k] = hb;
            } else {
                out[t * slice_bytes + r * Kp + k] = (int8_t)(int)d;
            }
        }
    }
}

struct Plan {
    int64_t Mp, Np, Kp, esz, a_slice, b_slice, nchunk_a, nchunk_b, kchunk_max;
    size_t off_a, off_b, off_ea, off_eb, off_part, total;
};

constexpr int64_t kMaxChunk = 1024;  // k-chunk of the max reduction

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

Plan make_plan(int64_t m, int64_t n, int64_t k, int s, bool bf16) {
    Plan p{};
    p.esz = bf16 ? 2 : 1;
    p.Mp = round_up(m > 0 ? m : 1, OT);
    p.Np = round_up(n > 0 ? n : 1, OT);
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
    return segs;
}

}  // namespace

}  // namespace dg

extern "C" {

int64_t dg_ozaki_workspace(int64_t m, int64_t n, int64_t k, int slices, int kind) {
    if (m < 0 || n < 0 || k < 0 || slices < 1 || slices > 16 || (kind != DG_OZAKI_INT8 && kind != DG_OZAKI_BF16))
        return -1;
    return (int64_t)dg::make_plan(m, n, k, slices, kind == DG_OZAKI_BF16).total;
}

int dg_ozaki_dgemm(char opA, char opB, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                   const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int slices, int kind,
                   void* workspace, int64_t workspace_bytes, double* err_bound, void* stream) {
    using namespace dg;
    const char oa = norm_op(opA), ob = norm_op(opB);
    if (!oa || !ob) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: invalid op '%c%c'", opA, opB);
    if (kind != DG_OZAKI_INT8 && kind != DG_OZAKI_BF16) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: kind %d", kind);
    if (slices < 1 || slices > 16) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: slices %d (1..16)", slices);
    if (m < 0 || n < 0 || k < 0) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: negative dimension");
    const int64_t rowsA = oa == 'N' ? m : k, rowsB = ob == 'N' ? k : n;
    if (lda < (rowsA > 1 ? rowsA : 1) || ldb < (rowsB > 1 ? rowsB : 1) || ldc < (m > 1 ? m : 1))
        return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: leading dimension too small");
    const bool bf16 = kind == DG_OZAKI_BF16;
    const int bits = bf16 ? 5 : 7;
    if (err_bound) *err_bound = (double)k * (slices + 3) * ldexp(1.0, -bits * slices);
    if (m == 0 || n == 0) return DG_OK;
    int rc = require_device();
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    if (k == 0 || alpha == 0.0) return launch_scale(m, n, beta, C, ldc, st);
    if (!A || !B || !C) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: null pointer");
    if (k > INT32_MAX / 2) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: k too large");
    const Plan p = make_plan(m, n, k, slices, bf16);
    if (p.Mp / 64 > 65535 || p.Np / 64 > 65535 || p.Kp / 64 > INT32_MAX)
        return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: matrix too large for one slicing launch");
    if (!workspace || workspace_bytes < (int64_t)p.total)
        return fail(DG_ERR_WORKSPACE, "dg_ozaki_dgemm: workspace %lld < %zu bytes", (long long)workspace_bytes, p.total);
    if (reinterpret_cast<uintptr_t>(workspace) & 255)
        return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: workspace must be 256-byte aligned");
    int serr = 0;
    std::vector<OzSeg> sched = make_schedule(slices, bits, p.Kp, k, bf16, &serr);
    if (serr) return fail(DG_ERR_UNSUPPORTED, "dg_ozaki_dgemm: k=%lld too large for exact 32-bit slice sums", (long long)k);

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

    // Row exponents of op(A), column exponents of op(B).
    {
        unsigned gx = a_ks == 1 ? (unsigned)(m < 65535 ? m : 65535) : (unsigned)((m + NT - 1) / NT < 65535 ? (m + NT - 1) / NT : 65535);
        hipLaunchKernelGGL(oz_partial_max_kernel, dim3(gx, gy), dim3(NT), 0, st, A, m, k, a_rs, a_ks, p.kchunk_max, part);
        DG_LAUNCH_CHECK();
        hipLaunchKernelGGL(oz_exponent_kernel, dim3((unsigned)((p.Mp + NT - 1) / NT)), dim3(NT), 0, st, part, m, p.Mp,
                           nchunk, ea);
        DG_LAUNCH_CHECK();
        gx = b_ks == 1 ? (unsigned)(n < 65535 ? n : 65535) : (unsigned)((n + NT - 1) / NT < 65535 ? (n + NT - 1) / NT : 65535);
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
    const OzSeg* dsched = static_cast<const OzSeg*>(cached_table(sched.data(), sched.size() * sizeof(OzSeg), st));
    if (!dsched) return DG_ERR_ALLOC;
    OzArgs g{};
    g.a = sa; g.b = sb; g.ea = ea; g.eb = eb; g.segs = dsched; g.nseg = (int32_t)sched.size(); g.bits = bits;
    g.m = m; g.n = n; g.Mp = p.Mp; g.Np = p.Np; g.Kp = p.Kp; g.a_slice = p.a_slice; g.b_slice = p.b_slice;
    g.alpha = alpha; g.beta = beta; g.c = C; g.ldc = ldc;
    g.tiles_m = (int32_t)(p.Mp / OT);
    g.tiles_n = (int32_t)(p.Np / OT);
    const int64_t tiles = (int64_t)g.tiles_m * g.tiles_n;
    if (tiles > INT32_MAX) return fail(DG_ERR_INVALID, "dg_ozaki_dgemm: too many tiles");
    if (bf16) hipLaunchKernelGGL(oz_gemm_kernel<true>, dim3((unsigned)tiles), dim3(NT), 0, st, g);
    else      hipLaunchKernelGGL(oz_gemm_kernel<false>, dim3((unsigned)tiles), dim3(NT), 0, st, g);
    DG_LAUNCH_CHECK();
    return DG_OK;
}

}  // extern "C"
