/*
 * oracle.h — CPU restatement of the reference's per-rank GEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker; the product (libdgamd.so) never calls it.
 *
 * Each function restates one reference function of
 * bvrettos/Distributed-GEMM-Thesis (cited path:line) in plain C, following
 * the reference's own loop structure where it has one, so that the GPU
 * kernels (which use closed-form index maps) are checked against an
 * independent derivation.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/Decompositions/pblasDecomposition.cpp:3-23 */
int64_t or_numroc(int64_t n, int64_t nb, int iproc, int isrproc, int nprocs);

/* src/Decompositions/2DBlockSequentialDecomposition.cpp:27-54 */
void or_virtual_grid(int size, int* drow, int* dcol);

/* tile extents: src/25DCannon.cpp:44-46 and
 * src/Decompositions/2DBlockSequentialDecomposition.cpp:11-25, 118-127 */
void or_block_splits(int64_t n, int parts, int64_t* offsets);

/* src/Transfers/sharedTransfers.cpp:29-63 (column-major branch) */
void or_copy_block(int64_t rows, int64_t cols, const double* src, int64_t lds, double* dst, int64_t ldd);

/* decompose2D src/25DCannon.cpp:87-149 / deliverMatrix
 * src/Decompositions/2DBlockSequentialDecomposition.cpp:56-162: tile (i,j)
 * goes to rank i*ncb + j; rank buffers are concatenated in rank order. */
void or_pack_block2d(const double* G, int64_t ldg, int64_t rows, int64_t cols, const int64_t* row_off, int nrb,
                     const int64_t* col_off, int ncb, double* packed);
/* reorderMatrix src/25DCannon.cpp:231-284 / gatherResult :164-253 */
void or_unpack_block2d(const double* packed, const int64_t* row_off, int nrb, const int64_t* col_off, int ncb,
                       double* G, int64_t ldg, int64_t rows, int64_t cols);

/* scatterMatrix src/Decompositions/pblasDecomposition.cpp:73-148 (its
 * running rowOffset/columnOffset bookkeeping, kept per rank). */
void or_pack_block_cyclic(const double* G, int64_t ldg, int64_t rows, int64_t cols, int64_t mb, int64_t nb, int P,
                          int Q, double* packed);
/* BlockSequentialDecomposer::deliverMatrix (2DBlockSequentialDecomposition.cpp:56-162):
 * every rank's A, B, C tiles in rank order, column-major (col_major != 0) or
 * row-major, each with the reference's local leading dimension. */
void or_block_seq_tiles(int col_major, int64_t M, int64_t N, int64_t K, int dRow, int dCol, const double* gA,
                        const double* gB, const double* gC, double* tA, double* tB, double* tC);
/* gatherMatrix src/Decompositions/pblasDecomposition.cpp:150-221 */
void or_unpack_block_cyclic(const double* packed, int64_t mb, int64_t nb, int P, int Q, double* G, int64_t ldg,
                            int64_t rows, int64_t cols);

/* Simulates the skew and rotation messages of Summa25Decomposer::multiply
 * (src/25DCannon.cpp:301-366) on a p x p x c grid and writes, for every
 * rank (i,j,k) and step r, the k-tile index m of the A tile it holds:
 * out[((k*p + i)*p + j)*(p/c) + r].  Also records the B tile's row index in
 * outB (same layout).  Returns 0. */
int or_cannon_simulate(int p, int c, int* outA, int* outB);

/* Textbook DGEMM with long double accumulation; op 'N' or 'T'. beta == 0
 * does not read C (cublasDgemm/cblas semantics). */
void or_dgemm(char opA, char opB, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc);

/* Whole 2.5D Cannon on the CPU: packed tiles as dg_cannon25d, the rank's
 * tile products in the simulated rotation order (src/25DCannon.cpp:322-366),
 * beta on layer 0 only, then the layer sum (reduceResult, :203-228).
 * `sched_std` != 0 uses shift-by-one rotations instead of the reference's. */
void or_cannon25d(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha, const double* Ap,
                  const double* Bp, double beta, double* Cp, int p, int c, int sched_std);

/* SUMMA on block-cyclic local matrices: C_loc = beta*C_loc + alpha * sum over
 * panels t of A_loc(pr, t mod Q)[:, panel] * B_loc(t mod P, pc)[panel, :]. */
void or_summa(int64_t M, int64_t N, int64_t K, double alpha, const double* Ap, const double* Bp, double beta,
              double* Cp, int64_t mb, int64_t nb, int64_t kb, int P, int Q);
/* The same with opA / opB = 'T': the stored A is K x M (blocks kb x mb),
 * the stored B is N x K (blocks nb x kb); TransA/TransB of cublasMpGemm,
 * src/Wrappers/cuBLASMp/cuBLASMpWrapped.cpp:565-574, 703. */
void or_summa_op(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha, const double* Ap,
                 const double* Bp, double beta, double* Cp, int64_t mb, int64_t nb, int64_t kb, int P, int Q);

/* C[x] += sum_l partial[l*n + x], layer order (MPI_Reduce SUM,
 * src/25DCannon.cpp:203-228). */
void or_layer_reduce(double* C, const double* partial, int64_t n, int nlayers);

/* src/Utilities/Validation.cpp:96-104, 85-94, 57-83 (messages omitted). */
double or_Derror(double a, double b);
long or_Dvec_diff(const double* a, const double* b, int64_t size, double eps);
short or_Dtest_equality(const double* C_comp, const double* C, int64_t size);

/* SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): sequential generator. */
uint64_t or_splitmix64_next(uint64_t* state);
/* out[i] from output i of the sequential stream seeded with `seed`;
 * dist 0: 2*(x>>11)*2^-53 - 1; 1: that times 2^e with e = ((x'>>32) mod 41) - 20
 * where x' is output i of the stream seeded with seed ^ 0xD1B54A32D192ED03;
 * 2: (x>>11)*2^-53. */
void or_generate(uint64_t seed, int dist, double* out, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
