/*
 * dgamd.h — C ABI of the MI355X (gfx950) per-rank GEMM engine.
 *
 * Every entry point takes plain pointers and 64-bit sizes, returns a
 * dg_status, never throws, and is stream-ordered on the HIP stream passed as
 * `stream` (NULL = the default stream).  Matrices are column-major in the BLAS
 * convention; device pointers are owned by the caller.  On failure the
 * message of the last error on the calling thread is available from
 * dg_last_error().
 *
 * HIP graph capture: the GEMM, SUMMA, Cannon, pack and Ozaki entry points
 * enqueue only kernel launches.  Problem and segment descriptors live in a
 * device-resident cache keyed by their contents, so one call made before
 * capture with the same shapes and pointers makes them resident; a call that
 * would have to upload a new table during capture fails with DG_ERR_INVALID
 * instead.
 *
 * Reference interfaces that each entry point replaces are cited as
 * path:line inside bvrettos/Distributed-GEMM-Thesis.
 */
#ifndef DGAMD_H
#define DGAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dg_status {
    DG_OK = 0,
    DG_ERR_INVALID = 1,     /* bad dimension, leading dimension, grid or op   */
    DG_ERR_UNSUPPORTED = 2, /* valid request this build does not implement    */
    DG_ERR_HIP = 3,         /* a HIP runtime call failed                      */
    DG_ERR_NO_DEVICE = 4,   /* no usable gfx950 device                        */
    DG_ERR_ALLOC = 5,       /* device or pinned allocation failed             */
    DG_ERR_SCHEDULE = 6,    /* Cannon rotation schedule does not cover K      */
    DG_ERR_WORKSPACE = 7    /* workspace argument too small                   */
} dg_status;

/* Message describing the last failure on this thread ("" if none). */
const char* dg_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int dg_version(void);
/* Non-zero when a device is present and its architecture is gfx950. */
int dg_device_ok(void);

/* ------------------------------------------------------------------ GEMM -- */

/* C = alpha*op(A)*op(B) + beta*C on the FP64 matrix cores
 * (v_mfma_f64_16x16x4_f64).  op is 'N'/'n', 'T'/'t' or 'C'/'c' (real data:
 * 'C' == 'T').  beta == 0 does not read C (NaN in C does not propagate).
 * Replaces: cublasDgemm(handle, opA, opB, m, n, k, &alpha, A, lda, B, ldb,
 *           &beta, C, ldc) at src/25DCannon.cpp:324-325 and 356-357,
 *           src/25DCannonNCCL.cpp:165-166, 202-203, 260-261, 291-292;
 *           op parsing as charToCublasTransOp, src/Utilities/TransposeEmblems.cpp:49-65. */
int dg_dgemm(char opA, char opB, int64_t m, int64_t n, int64_t k,
             double alpha, const double* A, int64_t lda,
             const double* B, int64_t ldb,
             double beta, double* C, int64_t ldc, void* stream);

/* C = alpha*op(A)*op(B) + beta*C in FP32 on the FP32 matrix cores
 * (v_mfma_f32_16x16x4_f32; FP32 products and sums), same conventions as
 * dg_dgemm (column-major, 'N'/'T'/'C', beta == 0 does not read C).
 * 16-byte aligned A and B with lda, ldb multiples of 4 (and K a multiple of
 * 4 when an operand is stored contiguous along k) take the direct-to-LDS
 * kernel; other layouts a slower LDS-tiled FMA kernel.
 * Replaces: the FP32 instantiations of the reference's GEMM entry points,
 *           slateGEMM<float> (src/Wrappers/Slate/slateWrapped.cpp:334-335)
 *           and cosmaPreDistributedOptimalGemm<float>
 *           (src/Wrappers/Cosma/cosmaWrapped.cpp:430-431), one rank's
 *           local product. */
int dg_sgemm(char opA, char opB, int64_t m, int64_t n, int64_t k,
             float alpha, const float* A, int64_t lda,
             const float* B, int64_t ldb,
             float beta, float* C, int64_t ldc, void* stream);

/* ------------------------------------------------------ 2D block layout -- */

/* Split `n` into `parts` contiguous ranges, the last range absorbing the
 * remainder: offsets[0..parts], offsets[i] = i*(n/parts), offsets[parts] = n.
 * Replaces the tile extents of Summa25Decomposer (localM = M/dRow,
 * src/25DCannon.cpp:44-46) and BlockSequentialDecomposer
 * (src/Decompositions/2DBlockSequentialDecomposition.cpp:11-25, 77-89). */
int dg_block_splits(int64_t n, int parts, int64_t* offsets);

/* Most-square process grid for `size` devices: *dcol = largest divisor of
 * size that is <= floor(sqrt(size)), *drow = size / *dcol.
 * Replaces: BlockSequentialDecomposer::calculateVirtualDeviceGrid,
 * src/Decompositions/2DBlockSequentialDecomposition.cpp:27-54 (and defines the
 * calculateProcessGrid that src/Decompositions/pblasDecomposition.cpp:31 calls
 * but the reference never defines). */
int dg_virtual_grid(int size, int* drow, int* dcol);

/* Pack a global column-major matrix G (rows x cols, leading dim ldg) into
 * per-rank tiles.  Tile (i,j) covers rows [row_off[i], row_off[i+1]) and
 * columns [col_off[j], col_off[j+1]); tiles are stored contiguously in rank
 * order r = i*ncb + j, each column-major with leading dimension = its row
 * count.  row_off / col_off are HOST arrays of nrb+1 / ncb+1 offsets
 * (nrb, ncb <= 256).  Bit-exact copy.
 * Replaces: Summa25Decomposer::decompose2D, src/25DCannon.cpp:87-149;
 *           BlockSequentialDecomposer::deliverMatrix,
 *           src/Decompositions/2DBlockSequentialDecomposition.cpp:56-162;
 *           copyBlock/transferBlock, src/Transfers/sharedTransfers.cpp:29-63,
 *           src/Transfers/mpiTransfers.cpp:16-35. */
int dg_pack_block2d(const double* G, int64_t ldg, int64_t rows, int64_t cols,
                    const int64_t* row_off, int nrb, const int64_t* col_off, int ncb,
                    double* packed, void* stream);

/* Inverse of dg_pack_block2d: scatter the packed tiles back into G.
 * Replaces: Summa25Decomposer::reorderMatrix, src/25DCannon.cpp:231-284;
 *           BlockSequentialDecomposer::gatherResult,
 *           src/Decompositions/2DBlockSequentialDecomposition.cpp:164-253;
 *           receiveBlock, src/Transfers/mpiTransfers.cpp:37-57. */
int dg_unpack_block2d(const double* packed,
                      const int64_t* row_off, int nrb, const int64_t* col_off, int ncb,
                      double* G, int64_t ldg, int64_t rows, int64_t cols, void* stream);

/* Strided block copy: dst[r + c*ldd] = src[r + c*lds] for r < rows, c < cols
 * (device to device).  Replaces: copyBlock, src/Transfers/sharedTransfers.cpp:29-63
 * and memcpy2D, :65-81. */
int dg_copy_block(int64_t rows, int64_t cols, const double* src, int64_t lds,
                  double* dst, int64_t ldd, void* stream);

/* FP32 form of dg_copy_block.
 * Replaces: copyBlock<float>, src/Transfers/sharedTransfers.cpp:88. */
int dg_copy_block_f32(int64_t rows, int64_t cols, const float* src, int64_t lds,
                      float* dst, int64_t ldd, void* stream);

/* Host to device: `height` columns of `width` bytes, column j at
 * host + j*host_pitch, to dev + j*dev_pitch (pitches in bytes, >= width).
 * Pinned host memory is copied by DMA directly.  Pageable host memory goes
 * through a process-wide ring of three 32 MiB pinned buffers (allocated at
 * the first such copy, kept until the process ends): up to eight host
 * threads fill one buffer while the DMA engine moves the previous one, and
 * such copies run one at a time per process.  The copy follows earlier work
 * on `stream` and is complete when the call returns; it cannot be captured
 * into a graph.
 * Replaces: the host-to-device branches of copy, copyBlock and memcpy2D,
 *           src/Transfers/sharedTransfers.cpp:22-81, as decompose2D uses
 *           them on host matrices, src/25DCannon.cpp:87-149. */
int dg_copy2d_h2d(void* dev, int64_t dev_pitch, const void* host, int64_t host_pitch,
                  int64_t width, int64_t height, void* stream);

/* Device to host, the same way; returns once `host` holds the data.
 * Replaces: the device-to-host branches of the same functions, as
 *           reorderMatrix gathers C into host memory, src/25DCannon.cpp:231-284. */
int dg_copy2d_d2h(void* host, int64_t host_pitch, const void* dev, int64_t dev_pitch,
                  int64_t width, int64_t height, void* stream);

/* ------------------------------------------------- block-cyclic layout -- */

/* ScaLAPACK NUMROC: rows (or columns) of an n-long dimension split in nb
 * blocks that process iproc owns among nprocs, the first block on isrcproc.
 * Replaces: numroc, src/Decompositions/pblasDecomposition.cpp:3-23. */
int64_t dg_numroc(int64_t n, int64_t nb, int iproc, int isrcproc, int nprocs);

/* Elements of the packed block-cyclic buffer: sum over the P x Q process
 * grid of numroc(rows, mb, pr, 0, P) * numroc(cols, nb, pc, 0, Q). */
int64_t dg_block_cyclic_size(int64_t rows, int64_t cols, int64_t mb, int64_t nb, int P, int Q);

/* Pack a global column-major matrix into the local matrices of a P x Q
 * row-major process grid (rank = pr*Q + pc, as procRow = rank/dCol), block
 * size mb x nb, first block on process (0,0).  Local matrices are stored
 * contiguously in rank order, each column-major with leading dimension
 * numroc(rows, mb, pr, 0, P).  Bit-exact.
 * Replaces: pblasDecomposer::scatterMatrix,
 *           src/Decompositions/pblasDecomposition.cpp:73-148. */
int dg_pack_block_cyclic(const double* G, int64_t ldg, int64_t rows, int64_t cols,
                         int64_t mb, int64_t nb, int P, int Q, double* packed, void* stream);

/* Inverse of dg_pack_block_cyclic.
 * Replaces: pblasDecomposer::gatherMatrix,
 *           src/Decompositions/pblasDecomposition.cpp:150-221. */
int dg_unpack_block_cyclic(const double* packed, int64_t mb, int64_t nb, int P, int Q,
                           double* G, int64_t ldg, int64_t rows, int64_t cols, void* stream);

/* ------------------------------------------------ virtual-grid executors -- */

#define DG_SCHED_AUTO 0      /* reference rotation where it covers K, else standard */
#define DG_SCHED_REFERENCE 1 /* the reference's rotation offsets r(r+1)/2          */
#define DG_SCHED_STANDARD 2  /* textbook Cannon: shift by one per step             */

/* Elements of workspace dg_cannon25d needs: (c-1) * M * N. */
int64_t dg_cannon25d_workspace(int64_t M, int64_t N, int p, int c);

/* 2D (c = 1) / 2.5D (c > 1) Cannon on a virtual p x p x c grid, one GPU.
 * A_packed, B_packed, C_packed hold the tiles of the STORED matrices in the
 * dg_pack_block2d layout over a p x p grid (uniform splits: p divides each
 * stored dimension).  Layer k, step r, rank (i,j) multiplies op(A)_{i,m} by
 * op(B)_{m,j} with m = (i + j + off[k][r]) mod p; the shifts of the reference
 * are these index rotations, nothing is moved.  Layer 0 starts from beta*C,
 * layers k > 0 from 0 into `workspace` ((c-1)*M*N doubles), and the layers
 * are summed into C_packed in HBM.  `schedule` is one of DG_SCHED_*.
 * Replaces: Summa25Decomposer::multiply, src/25DCannon.cpp:286-375 (skew
 *           303-315, rotations 331-366, beta policy 322/352, reduceResult
 *           203-228, broadcastToStack 174-194) and serializedMultiplyNCCL,
 *           src/25DCannonNCCL.cpp:216-304. */
int dg_cannon25d(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha,
                 const double* A_packed, const double* B_packed, double beta,
                 double* C_packed, int p, int c, int schedule,
                 double* workspace, int64_t workspace_elems, void* stream);

/* Rank (i,j,k)'s m index at step r under `schedule` (host helper for tests
 * and callers that print the schedule); -1 if the arguments are invalid.
 * Replaces the target arithmetic of src/25DCannon.cpp:303-310, 333-344 and
 * truncatedDivisionRemainer, src/25DCannonNCCL.cpp:6-9. */
int dg_cannon_m_index(int p, int c, int schedule, int i, int j, int k, int r);

/* SUMMA on a virtual P x Q grid over block-cyclic local matrices
 * (dg_pack_block_cyclic layout): C = alpha op(A) op(B) + beta C.  C is M x N
 * with blocks mb x nb.  The stored A is M x K with blocks mb x kb (opA = 'N')
 * or K x M with blocks kb x mb (opA = 'T'); the stored B is K x N with blocks
 * kb x nb (opB = 'N') or N x K with blocks nb x kb (opB = 'T'); all three are
 * distributed over the same P x Q grid.  Panel t of width kb is read in place
 * from its owners (for 'N' operands process column t mod Q of A and process
 * row t mod P of B): the panel broadcasts are index rotations.
 * New relative to the reference (include/summa.hpp:1-28 is empty; SUMMA
 * exists there only inside cublasMpGemm, whose TransA/TransB arguments this
 * follows: src/Wrappers/cuBLASMp/cuBLASMpWrapped.cpp:198-223, 565-574, 703). */
int dg_summa(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A_packed, const double* B_packed, double beta, double* C_packed,
             int64_t mb, int64_t nb, int64_t kb, int P, int Q, void* stream);

/* C[x] += sum_{l < nlayers} partial[l*n + x] for x < n (HBM reduction of the
 * 2.5D layers, summed in layer order).
 * Replaces: Summa25Decomposer::reduceResult (MPI_Reduce SUM),
 *           src/25DCannon.cpp:203-228; reduceResultNCCL, src/25DCannonNCCL.cpp:306-320. */
int dg_layer_reduce(double* C, const double* partial, int64_t n, int nlayers, void* stream);

/* ------------------------------------------------------- grouped batch -- */

/* One GEMM of a grouped batch: C = alpha*op(A)*op(B) + beta*C. */
typedef struct dg_gemm_desc {
    char opA, opB;     /* 'N'/'T'/'C' as in dg_dgemm */
    char reserved[6];
    int64_t m, n, k;
    double alpha, beta;
    const double* A;
    int64_t lda;
    const double* B;
    int64_t ldb;
    double* C;
    int64_t ldc;
} dg_gemm_desc;

/* Every GEMM of `descs` in one launch per distinct (opA, opB) pair: the
 * tiles of all problems are dealt from one tile space, largest K first.
 * Output ranges must not overlap, and no problem's output may be another
 * problem's input: the problems run in no specified order.
 * Replaces the per-rank local GEMMs the reference runs one rank at a time
 * on BlockSequentialDecomposer tiles (PARALiADgemm per rank,
 * src/distributedParalia.cpp:57-64, 159-167) and the per-step cublasDgemm
 * calls of src/25DCannon.cpp:324-357. */
int dg_grouped_dgemm(const dg_gemm_desc* descs, int count, void* stream);

/* dg_grouped_dgemm that allocates nothing: the batch's descriptor tables go
 * into the library's static device scratch array (256 KiB per GPU, held
 * until the process ends) instead of the table cache, written by kernels
 * whose arguments carry them, so the call also records cleanly into a HIP
 * graph.  A batch whose tables exceed the array (about 1900 single-segment
 * problems) fails with DG_ERR_WORKSPACE before any launch.  Calls that
 * overlap on one GPU share the array and race.  No counterpart in the
 * reference: for callers that cannot allocate device memory (graph capture,
 * tight serving loops). */
int dg_grouped_dgemm_noalloc(const dg_gemm_desc* descs, int count, void* stream);

/* A grouped batch prepared once and executed many times: descriptor tables
 * are uploaded at creation, the tile shape / scheduling variant is chosen by
 * the autotuner (timed on first use, cached per shape class), and
 * dg_grouped_plan_graph_launch captures the launch into a HIP graph on its
 * first call and replays it afterwards.  The operand pointers are fixed at
 * creation.  New relative to the reference (no batching exists there). */
typedef struct dg_grouped_plan dg_grouped_plan;

#define DG_PLAN_AUTOTUNE 0       /* time the variants on first execution     */
#define DG_PLAN_NONPERSISTENT 1  /* one workgroup per output tile           */
#define DG_PLAN_PERSISTENT 2     /* one workgroup per CU slot, tiles taken from
                                    an atomic work counter, largest K first */

int dg_grouped_plan_create(const dg_gemm_desc* descs, int count, int mode, dg_grouped_plan** plan);
int dg_grouped_plan_execute(dg_grouped_plan* plan, void* stream);
int dg_grouped_plan_graph_launch(dg_grouped_plan* plan, void* stream);
/* Variant the plan runs (DG_PLAN_NONPERSISTENT or DG_PLAN_PERSISTENT; the
 * autotuner decides on the first execution when created with AUTOTUNE). */
int dg_grouped_plan_variant(const dg_grouped_plan* plan);
int dg_grouped_plan_destroy(dg_grouped_plan* plan);

/* ---------------------------------------------------- Ozaki split DGEMM -- */

#define DG_OZAKI_INT8 0      /* 7-bit digits in int8, v_mfma_i32_16x16x64_i8, exact INT32 sums  */
#define DG_OZAKI_BF16 1      /* 5-bit digits in bf16, v_mfma_f32_16x16x32_bf16, exact FP32 sums */
#define DG_OZAKI_INT8_CRT 2  /* Ozaki scheme II: `slices` = moduli (2..16), one INT8 GEMM each,
                                exact CRT reconstruction (Ozaki, Uchino, Imamura 2025)       */

/* Bytes of device workspace dg_ozaki_dgemm needs (-1 on bad arguments). */
int64_t dg_ozaki_workspace(int64_t m, int64_t n, int64_t k, int slices, int kind);

/* C = alpha*op(A)*op(B) + beta*C emulated with the Ozaki split: every row of
 * op(A) and column of op(B) is scaled by a power of two and cut into
 * `slices` digits (INT8 or BF16 kind); the s(s+1)/2 slice products with
 * t + u <= s + 1 run on the low-precision matrix cores with exact 32-bit
 * accumulation and are recombined in FP64 on the GPU.  `workspace` is
 * caller-owned device memory of dg_ozaki_workspace() bytes, 256-byte
 * aligned.  If err_bound is not NULL it receives k*(s+3)*2^(-b*s) (b = 7 for
 * INT8, 5 for BF16): |C - op(A)op(B)|_ij <= err_bound * 2^(e_i + f_j) for the
 * product part, e_i / f_j being the row / column scales.  With
 * DG_OZAKI_INT8_CRT, `slices` is the number of moduli N; every row of op(A)
 * and column of op(B) is scaled by a power of two to integers (at most 53
 * bits) whose lines have 2-norm <= 2^Q, Q the largest with
 * 2^(2Q) + ceil(sqrt(k))*2^Q + k <= P/4 (P the moduli's product; the
 * Cauchy-Schwarz scaling of Ozaki scheme II), their integer product is formed
 * exactly from N INT8 GEMMs and rebuilt by the Chinese remainder theorem, and
 * err_bound = k*2^(1-q) with q = min(53, Q - ceil(log4 k)) (every line gets
 * at least q bits; lines whose norm is far below sqrt(k) times their largest
 * entry get more).
 * New relative to the reference (no split DGEMM exists there); it replaces
 * the same cublasDgemm call sites as dg_dgemm (src/25DCannon.cpp:324-325). */
int dg_ozaki_dgemm(char opA, char opB, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb,
                   double beta, double* C, int64_t ldc, int slices, int kind,
                   void* workspace, int64_t workspace_bytes, double* err_bound, void* stream);

/* ------------------------------------------- grouped batch, Ozaki form -- */

#define DG_FORM_FP64 0        /* FP64 MFMA for every problem                           */
#define DG_FORM_OZAKI_INT8 1  /* per problem, Ozaki scheme II (DG_OZAKI_INT8_CRT) where it
                                 is measured faster than the FP64 MFMA path, else FP64  */

/* A grouped plan in the given arithmetic form.  DG_FORM_FP64 is
 * dg_grouped_plan_create.  With DG_FORM_OZAKI_INT8, `moduli` is the CRT
 * moduli count (2..16), or 0 to take per problem the smallest count whose
 * scaling reaches q = 53 bits (the FP64-level bound k*2^-52, 16 when none
 * does).  On the first execution the autotuner times each problem of at
 * least 2^32 flops both ways on its own operands (outputs restored
 * afterwards; choices cached per shape and moduli count) and keeps the
 * faster; smaller problems stay FP64.  The FP64 problems still run as one
 * grouped launch; the Ozaki problems follow on the same stream, sharing one
 * workspace the plan owns; dg_grouped_plan_graph_launch captures all of it
 * in one HIP graph.  New relative to the reference. */
int dg_grouped_plan_create_form(const dg_gemm_desc* descs, int count, int mode, int form, int moduli,
                                dg_grouped_plan** plan);
/* Form chosen for problem `index` (DG_FORM_*; -1 before the first execution
 * or on a bad index); *moduli receives its moduli count (0 for FP64). */
int dg_grouped_plan_problem_form(const dg_grouped_plan* plan, int index, int* moduli);

/* ------------------------------------------------------------ generation -- */

#define DG_DIST_UNIFORM_PM1 0  /* u in [-1, 1): 2*(x>>11)*2^-53 - 1                 */
#define DG_DIST_WIDE 1         /* u * 2^e, e uniform integer in [-20, 20]          */
#define DG_DIST_UNIFORM_01 2   /* (x>>11)*2^-53 in [0, 1)                          */

/* Fill ptr[0..n) from SplitMix64 (Steele, Lea, Flood 2014): element i uses
 * output i of the generator seeded with `seed` (counter based, so any
 * element can be produced independently).  Bit-identical to the CPU oracle.
 * Replaces: MatrixInit / generateMatrixGPU (cuRAND, clock seeded),
 *           src/Utilities/cmatrix.cpp:104-169, 239-297. */
int dg_generate(uint64_t seed, int dist, double* ptr, int64_t n, void* stream);

/* FP32 form of dg_generate: the same FP64 values rounded to nearest, ties to
 * even.  Replaces: generateMatrixGPU<float> / MatrixInit<float>,
 * src/Utilities/cmatrix.cpp:301-309. */
int dg_generate_f32(uint64_t seed, int dist, float* ptr, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DGAMD_H */
