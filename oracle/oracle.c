/*
 * oracle.c — CPU restatement of the reference's per-rank GEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker; never by the product.
 * Functions follow the cited reference code (bvrettos/Distributed-GEMM-Thesis)
 * loop by loop where it has loops; see oracle.h for the citations.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

int64_t or_numroc(int64_t n, int64_t nb, int iproc, int isrproc, int nprocs) {
    /* pblasDecomposition.cpp:3-23 */
    int64_t extraBlocks, myDistance, numBlocks, r;
    myDistance = (nprocs + iproc - isrproc) % nprocs;
    numBlocks = n / nb;
    r = (numBlocks / nprocs) * nb;
    extraBlocks = numBlocks % nprocs;
    if (myDistance < extraBlocks) r += nb;
    else if (myDistance == extraBlocks) r += n % nb;
    return r;
}

void or_virtual_grid(int size, int* drow, int* dcol) {
    /* 2DBlockSequentialDecomposition.cpp:27-54 */
    int Px = (int)sqrt((double)size);
    int Py = Px;
    int counter;
    if (Px == 0) { Py = size; Px = 1; }
    for (counter = Px; counter > 0; --counter)
        if (size % counter == 0) break;
    if (counter == 0) { Px = size; Py = 1; }
    else { Px = counter; Py = size / counter; }
    *drow = Py;
    *dcol = Px;
}

void or_block_splits(int64_t n, int parts, int64_t* off) {
    /* tile i starts at i*(n/parts); the last tile absorbs n % parts
       (2DBlockSequentialDecomposition.cpp:118-121) */
    int i;
    for (i = 0; i < parts; ++i) off[i] = (int64_t)i * (n / parts);
    off[parts] = n;
}

void or_copy_block(int64_t rows, int64_t cols, const double* src, int64_t lds, double* dst, int64_t ldd) {
    /* sharedTransfers.cpp:59-90, column by column */
    int64_t c;
    if (rows * cols <= 0) return;
    for (c = 0; c < cols; ++c) memcpy(dst + ldd * c, src + lds * c, sizeof(double) * (size_t)rows);
}

static int64_t rank_base_2d(const int64_t* row_off, const int64_t* col_off, int ncb, int i, int j) {
    /* sum of the sizes of all tiles of ranks before i*ncb + j */
    int64_t base = 0;
    int ii, jj;
    for (ii = 0; ii <= i; ++ii)
        for (jj = 0; jj < ncb; ++jj) {
            if (ii == i && jj == j) return base;
            base += (row_off[ii + 1] - row_off[ii]) * (col_off[jj + 1] - col_off[jj]);
        }
    return base;
}

void or_pack_block2d(const double* G, int64_t ldg, int64_t rows, int64_t cols, const int64_t* row_off, int nrb,
                     const int64_t* col_off, int ncb, double* packed) {
    /* decompose2D 25DCannon.cpp:113-128: copyBlock(localM, localK, &A[offset], aTile, M, localM) per (i,j) */
    int i, j;
    (void)rows; (void)cols;
    for (i = 0; i < nrb; ++i)
        for (j = 0; j < ncb; ++j) {
            int64_t tm = row_off[i + 1] - row_off[i], tn = col_off[j + 1] - col_off[j];
            or_copy_block(tm, tn, G + row_off[i] + col_off[j] * ldg, ldg,
                          packed + rank_base_2d(row_off, col_off, ncb, i, j), tm);
        }
}

/* copyBlock with colMajor = false (sharedTransfers.cpp:29-63): rows x cols
 * of a row-major source (ld = lds) into a row-major destination (ld = ldd),
 * row by row. */
static void or_copy_block_rows(int64_t rows, int64_t cols, const double* src, int64_t lds, double* dst, int64_t ldd) {
    int64_t r;
    if (rows * cols <= 0) return;
    for (r = 0; r < rows; ++r) memcpy(dst + ldd * r, src + lds * r, sizeof(double) * (size_t)cols);
}

/* BlockSequentialDecomposer::deliverMatrix, 2DBlockSequentialDecomposition.cpp:56-162:
 * every rank's tiles A_r (tileM x K), B_r (K x tileN), C_r (tileM x tileN), in
 * rank order r = i*dCol + j, each tile in the storage order of the globals:
 * column-major (lld = tile rows; offsets :100-108) or row-major (lld = tile
 * columns; offsets :89-99).  Tiles of the last grid row / column absorb the
 * remainder (:79-88).  Offsets are 64-bit here (the reference's are int). */
void or_block_seq_tiles(int col_major, int64_t M, int64_t N, int64_t K, int dRow, int dCol, const double* gA,
                        const double* gB, const double* gC, double* tA, double* tB, double* tC) {
    int i, j;
    int64_t pa = 0, pb = 0, pc = 0;
    for (i = 0; i < dRow; ++i) {
        int64_t tileM = M / dRow + (i + 1 == dRow ? M % dRow : 0);
        for (j = 0; j < dCol; ++j) {
            int64_t tileN = N / dCol + (j + 1 == dCol ? N % dCol : 0);
            if (col_major) {
                or_copy_block(tileM, K, gA + (M / dRow) * i, M, tA + pa, tileM);
                or_copy_block(K, tileN, gB + (N / dCol) * K * j, K, tB + pb, K);
                or_copy_block(tileM, tileN, gC + (N / dCol) * M * j + (M / dRow) * i, M, tC + pc, tileM);
            } else {
                or_copy_block_rows(tileM, K, gA + (M / dRow) * K * i, K, tA + pa, K);
                or_copy_block_rows(K, tileN, gB + (N / dCol) * j, N, tB + pb, tileN);
                or_copy_block_rows(tileM, tileN, gC + (M / dRow) * N * i + (N / dCol) * j, N, tC + pc, tileN);
            }
            pa += tileM * K;
            pb += K * tileN;
            pc += tileM * tileN;
        }
    }
}

void or_unpack_block2d(const double* packed, const int64_t* row_off, int nrb, const int64_t* col_off, int ncb,
                       double* G, int64_t ldg, int64_t rows, int64_t cols) {
    /* reorderMatrix 25DCannon.cpp:254-271: copyBlock(localM, localN, cTile, &C[offset], localM, M) */
    int i, j;
    (void)rows; (void)cols;
    for (i = 0; i < nrb; ++i)
        for (j = 0; j < ncb; ++j) {
            int64_t tm = row_off[i + 1] - row_off[i], tn = col_off[j + 1] - col_off[j];
            or_copy_block(tm, tn, packed + rank_base_2d(row_off, col_off, ncb, i, j), tm,
                          G + row_off[i] + col_off[j] * ldg, ldg);
        }
}

/* Shared walk of scatterMatrix/gatherMatrix (pblasDecomposition.cpp:134-170
 * and 211-245): row blocks cycle the receiving process row, column blocks
 * the process column; every rank keeps its running rowOffset/columnOffset,
 * advanced modulo its local extent. */
static void cyclic_walk(const double* G_in, double* G_out, int64_t ldg, int64_t M, int64_t N, int64_t Mb, int64_t Nb,
                        int P, int Q, const double* pk_in, double* pk_out) {
    int nr = P * Q, r;
    int64_t* localRows = (int64_t*)calloc((size_t)P, sizeof(int64_t));
    int64_t* localCols = (int64_t*)calloc((size_t)Q, sizeof(int64_t));
    int64_t* base = (int64_t*)calloc((size_t)nr, sizeof(int64_t));
    int64_t* rowOffset = (int64_t*)calloc((size_t)nr, sizeof(int64_t));
    int64_t* columnOffset = (int64_t*)calloc((size_t)nr, sizeof(int64_t));
    int64_t acc = 0, row, col;
    int pr, pc, rpr, rpc;
    for (pr = 0; pr < P; ++pr) localRows[pr] = or_numroc(M, Mb, pr, 0, P);
    for (pc = 0; pc < Q; ++pc) localCols[pc] = or_numroc(N, Nb, pc, 0, Q);
    for (r = 0; r < nr; ++r) { base[r] = acc; acc += localRows[r / Q] * localCols[r % Q]; }
    rpr = 0;
    for (row = 0; row < M; row += Mb, rpr = (rpr + 1) % P) {
        int64_t rowsToSend = Mb;
        rpc = 0;
        if (M - row < Mb) rowsToSend = M - row;
        for (col = 0; col < N; col += Nb, rpc = (rpc + 1) % Q) {
            int64_t columnsToSend = Nb;
            int64_t lr;
            double* loc;
            r = rpr * Q + rpc;
            if (N - col < Nb) columnsToSend = N - col;
            lr = localRows[rpr];
            if (G_in) {
                loc = pk_out + base[r] + columnOffset[r] * lr + rowOffset[r];
                or_copy_block(rowsToSend, columnsToSend, G_in + col * ldg + row, ldg, loc, lr);
            } else {
                or_copy_block(rowsToSend, columnsToSend, pk_in + base[r] + columnOffset[r] * lr + rowOffset[r], lr,
                              G_out + col * ldg + row, ldg);
            }
            columnOffset[r] = (columnOffset[r] + columnsToSend) % localCols[rpc];
        }
        for (r = 0; r < nr; ++r)
            if (r / Q == rpr) rowOffset[r] = (rowOffset[r] + rowsToSend) % localRows[rpr];
    }
    free(localRows); free(localCols); free(base); free(rowOffset); free(columnOffset);
}

void or_pack_block_cyclic(const double* G, int64_t ldg, int64_t rows, int64_t cols, int64_t mb, int64_t nb, int P,
                          int Q, double* packed) {
    cyclic_walk(G, NULL, ldg, rows, cols, mb, nb, P, Q, NULL, packed);
}

void or_unpack_block_cyclic(const double* packed, int64_t mb, int64_t nb, int P, int Q, double* G, int64_t ldg,
                            int64_t rows, int64_t cols) {
    cyclic_walk(NULL, G, ldg, rows, cols, mb, nb, P, Q, packed, NULL);
}

static int tmod(int a, int b) { return ((a % b) + b) % b; } /* 25DCannonNCCL.cpp:6-9 */

int or_cannon_simulate(int p, int c, int* outA, int* outB) {
    /* 25DCannon.cpp:37-42 (coordinates, pc3 = p/c), 301-315 (skew), 331-366 (rotations) */
    int pc3 = p / c, k, i, j, rot;
    size_t n = (size_t)c * p * p;
    int* hA = (int*)malloc(n * sizeof(int));
    int* hB = (int*)malloc(n * sizeof(int));
    int* nA = (int*)malloc(n * sizeof(int));
    int* nB = (int*)malloc(n * sizeof(int));
#define IX(k_, i_, j_) (((size_t)(k_) * p + (i_)) * p + (j_))
    /* after broadcastToStack every layer holds layer 0's A_ij, B_ij */
    for (k = 0; k < c; ++k)
        for (i = 0; i < p; ++i)
            for (j = 0; j < p; ++j) { hA[IX(k, i, j)] = j; hB[IX(k, i, j)] = i; }
    /* skew: A_ij to column s = (j - i + k*pc3) mod p; B_ij to row s' = (i - j + k*pc3) mod p */
    for (k = 0; k < c; ++k)
        for (i = 0; i < p; ++i)
            for (j = 0; j < p; ++j) {
                nA[IX(k, i, tmod(j - i + k * pc3, p))] = hA[IX(k, i, j)];
                nB[IX(k, tmod(i - j + k * pc3, p), j)] = hB[IX(k, i, j)];
            }
    memcpy(hA, nA, n * sizeof(int));
    memcpy(hB, nB, n * sizeof(int));
    for (rot = 0; rot < pc3; ++rot) {
        if (rot > 0) {
            /* rotation `rot`: A to column (j + rot) % p, B to row (i + rot) % p (333-344, 364-365) */
            for (k = 0; k < c; ++k)
                for (i = 0; i < p; ++i)
                    for (j = 0; j < p; ++j) {
                        nA[IX(k, i, (j + rot) % p)] = hA[IX(k, i, j)];
                        nB[IX(k, (i + rot) % p, j)] = hB[IX(k, i, j)];
                    }
            memcpy(hA, nA, n * sizeof(int));
            memcpy(hB, nB, n * sizeof(int));
        }
        for (k = 0; k < c; ++k)
            for (i = 0; i < p; ++i)
                for (j = 0; j < p; ++j) {
                    outA[IX(k, i, j) * pc3 + rot] = hA[IX(k, i, j)];
                    outB[IX(k, i, j) * pc3 + rot] = hB[IX(k, i, j)];
                }
    }
#undef IX
    free(hA); free(hB); free(nA); free(nB);
    return 0;
}

void or_dgemm(char opA, char opB, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
    int64_t i, j, l;
    int ta = (opA == 'T' || opA == 't' || opA == 'C' || opA == 'c');
    int tb = (opB == 'T' || opB == 't' || opB == 'C' || opB == 'c');
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) {
            long double s = 0.0L;
            double v;
            for (l = 0; l < k; ++l) {
                double a = ta ? A[l + i * lda] : A[i + l * lda];
                double b = tb ? B[j + l * ldb] : B[l + j * ldb];
                s += (long double)a * (long double)b;
            }
            v = (double)((long double)alpha * s);
            if (beta != 0.0) v = (double)((long double)v + (long double)beta * (long double)C[i + j * ldc]);
            C[i + j * ldc] = v;
        }
}

void or_cannon25d(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha, const double* Ap,
                  const double* Bp, double beta, double* Cp, int p, int c, int sched_std) {
    int steps = p / c, k, i, j, r;
    int64_t mt = M / p, nt = N / p, kt = K / p;
    int64_t tA = mt * kt, tB = kt * nt, tC = mt * nt;
    int ta = (opA == 'T' || opA == 't' || opA == 'C' || opA == 'c');
    int tb = (opB == 'T' || opB == 't' || opB == 'C' || opB == 'c');
    size_t nsim = (size_t)c * p * p * steps;
    int* mA = (int*)malloc(nsim * sizeof(int));
    int* mB = (int*)malloc(nsim * sizeof(int));
    double* part = (double*)calloc((size_t)(c > 1 ? c - 1 : 1) * (size_t)(M * N), sizeof(double));
    or_cannon_simulate(p, c, mA, mB);
    for (k = 0; k < c; ++k)
        for (i = 0; i < p; ++i)
            for (j = 0; j < p; ++j) {
                double* Ct = k == 0 ? Cp + ((int64_t)i * p + j) * tC : part + (int64_t)(k - 1) * M * N + ((int64_t)i * p + j) * tC;
                for (r = 0; r < steps; ++r) {
                    size_t s = (((size_t)k * p + i) * p + j) * steps + r;
                    int m = sched_std ? tmod(i + j - k * steps - r, p) : mA[s];
                    const double* At = ta ? Ap + ((int64_t)m * p + i) * tA : Ap + ((int64_t)i * p + m) * tA;
                    const double* Bt = tb ? Bp + ((int64_t)j * p + m) * tB : Bp + ((int64_t)m * p + j) * tB;
                    /* first product: beta on layer 0, 0 elsewhere (25DCannon.cpp:322); later: 1 (352) */
                    double be = r == 0 ? (k == 0 ? beta : 0.0) : 1.0;
                    or_dgemm(opA, opB, mt, nt, kt, alpha, At, ta ? kt : mt, Bt, tb ? nt : kt, be, Ct, mt);
                }
            }
    if (c > 1) or_layer_reduce(Cp, part, M * N, c - 1);
    free(mA); free(mB); free(part);
}

void or_summa(int64_t M, int64_t N, int64_t K, double alpha, const double* Ap, const double* Bp, double beta,
              double* Cp, int64_t mb, int64_t nb, int64_t kb, int P, int Q) {
    /* Local matrices in rank order; base of rank r = sum of earlier ranks' sizes. */
    int pr, pc, r;
    int64_t t, npanels = (K + kb - 1) / kb;
    int nr = P * Q;
    int64_t *bA = (int64_t*)calloc((size_t)nr, sizeof(int64_t)), *bB = (int64_t*)calloc((size_t)nr, sizeof(int64_t)),
            *bC = (int64_t*)calloc((size_t)nr, sizeof(int64_t));
    int64_t aA = 0, aB = 0, aC = 0;
    for (r = 0; r < nr; ++r) {
        int rp = r / Q, cp = r % Q;
        bA[r] = aA; aA += or_numroc(M, mb, rp, 0, P) * or_numroc(K, kb, cp, 0, Q);
        bB[r] = aB; aB += or_numroc(K, kb, rp, 0, P) * or_numroc(N, nb, cp, 0, Q);
        bC[r] = aC; aC += or_numroc(M, mb, rp, 0, P) * or_numroc(N, nb, cp, 0, Q);
    }
    for (pr = 0; pr < P; ++pr)
        for (pc = 0; pc < Q; ++pc) {
            int64_t lm = or_numroc(M, mb, pr, 0, P), ln = or_numroc(N, nb, pc, 0, Q);
            double* Cl = Cp + bC[pr * Q + pc];
            if (lm == 0 || ln == 0) continue;
            if (npanels == 0) {
                int64_t x;
                for (x = 0; x < lm * ln; ++x) Cl[x] = beta == 0.0 ? 0.0 : beta * Cl[x];
            }
            for (t = 0; t < npanels; ++t) {
                int qa = (int)(t % Q), pb = (int)(t % P);
                int64_t w = (K - t * kb) < kb ? (K - t * kb) : kb;
                int64_t ldb = or_numroc(K, kb, pb, 0, P);
                const double* Al = Ap + bA[pr * Q + qa] + (t / Q) * kb * lm;
                const double* Bl = Bp + bB[pb * Q + pc] + (t / P) * kb;
                or_dgemm('N', 'N', lm, ln, w, alpha, Al, lm, Bl, ldb, t == 0 ? beta : 1.0, Cl, lm);
            }
        }
    free(bA); free(bB); free(bC);
}

/* SUMMA with either operand transposed, restated on the global matrices:
 * unpack the stored operands (A is M x K or K x M, B is K x N or N x K, on
 * the same P x Q grid with the blocks of the matching dimensions), add the
 * panels of width kb in order 0, 1, ... (beta on the first), pack C.  The
 * result per element is the same panel-ordered sum as or_summa's. */
void or_summa_op(char opA, char opB, int64_t M, int64_t N, int64_t K, double alpha, const double* Ap,
                 const double* Bp, double beta, double* Cp, int64_t mb, int64_t nb, int64_t kb, int P, int Q) {
    int ta = opA == 'T' || opA == 't' || opA == 'C' || opA == 'c';
    int tb = opB == 'T' || opB == 't' || opB == 'C' || opB == 'c';
    int64_t rA = ta ? K : M, cA = ta ? M : K, rB = tb ? N : K, cB = tb ? K : N;
    int64_t t, npanels = (K + kb - 1) / kb, x;
    double* Ag = (double*)malloc((size_t)(rA * cA > 0 ? rA * cA : 1) * sizeof(double));
    double* Bg = (double*)malloc((size_t)(rB * cB > 0 ? rB * cB : 1) * sizeof(double));
    double* Cg = (double*)malloc((size_t)(M * N > 0 ? M * N : 1) * sizeof(double));
    if (M == 0 || N == 0) { free(Ag); free(Bg); free(Cg); return; }
    or_unpack_block_cyclic(Ap, ta ? kb : mb, ta ? mb : kb, P, Q, Ag, rA > 0 ? rA : 1, rA, cA);
    or_unpack_block_cyclic(Bp, tb ? nb : kb, tb ? kb : nb, P, Q, Bg, rB > 0 ? rB : 1, rB, cB);
    or_unpack_block_cyclic(Cp, mb, nb, P, Q, Cg, M, M, N);
    if (npanels == 0)
        for (x = 0; x < M * N; ++x) Cg[x] = beta == 0.0 ? 0.0 : beta * Cg[x];
    for (t = 0; t < npanels; ++t) {
        int64_t k0 = t * kb, w = (K - k0) < kb ? (K - k0) : kb;
        const double* Ak = ta ? Ag + k0 : Ag + k0 * M;   /* op(A)(:, k0:k0+w) */
        const double* Bk = tb ? Bg + k0 * N : Bg + k0;   /* op(B)(k0:k0+w, :) */
        or_dgemm(ta ? 'T' : 'N', tb ? 'T' : 'N', M, N, w, alpha, Ak, ta ? K : M, Bk, tb ? N : K,
                 t == 0 ? beta : 1.0, Cg, M);
    }
    or_pack_block_cyclic(Cg, M, M, N, mb, nb, P, Q, Cp);
    free(Ag); free(Bg); free(Cg);
}

void or_layer_reduce(double* C, const double* partial, int64_t n, int nlayers) {
    int64_t x;
    int l;
    for (x = 0; x < n; ++x) {
        double acc = C[x];
        for (l = 0; l < nlayers; ++l) acc += partial[(int64_t)l * n + x];
        C[x] = acc;
    }
}

/* Validation.cpp:96-104 */
double or_Derror(double a, double b) {
    double d = a - b;
    if (d < 0) d = -d;
    if (a == 0) return d;
    return d / a;
}

/* Validation.cpp:85-94 */
long or_Dvec_diff(const double* a, const double* b, int64_t size, double eps) {
    long failed = 0;
    int64_t i;
    for (i = 0; i < size; i++)
        if (or_Derror(a[i], b[i]) > eps) failed++;
    return failed;
}

/* Validation.cpp:57-83 (return value only) */
short or_Dtest_equality(const double* C_comp, const double* C, int64_t size) {
    long acc = 8, failed;
    double eps = 1e-8;
    failed = or_Dvec_diff(C_comp, C, size, eps);
    while (eps > DBL_MIN && !failed && acc < 30) {
        eps *= 0.1;
        acc++;
        failed = or_Dvec_diff(C_comp, C, size, eps);
    }
    if (8 == acc) return 0;
    return (short)acc;
}

uint64_t or_splitmix64_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_generate(uint64_t seed, int dist, double* out, int64_t n) {
    uint64_t s = seed, se = seed ^ 0xD1B54A32D192ED03ull;
    int64_t i;
    for (i = 0; i < n; ++i) {
        uint64_t z = or_splitmix64_next(&s);
        double u = (double)(z >> 11) * 0x1.0p-53;
        if (dist == 2) {
            out[i] = u;
        } else {
            double v = 2.0 * u - 1.0;
            if (dist == 1) {
                uint64_t ze = or_splitmix64_next(&se);
                int e = (int)((ze >> 32) % 41ull) - 20;
                v = ldexp(v, e);
            }
            out[i] = v;
        }
    }
}
