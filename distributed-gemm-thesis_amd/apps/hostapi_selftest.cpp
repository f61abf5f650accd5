// Self-test of the C++ host API on one GPU: every drop-in class and driver
// on small problems, validated with Dtest_equality and the relative
// Frobenius error against a host triple-loop DGEMM.  Exit status 0 on
// success; prints one line per check.
//
//   hostapi_selftest          run every check
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include <cstdlib>
#include <unistd.h>
#include <sstream>
#include <fstream>

#include <hip/hip_runtime.h>

#include "25DCannon.hpp"
#include "MatrixUtilities.hpp"
#include "cmatrix.h"
#include "2DBlockSequentialDecomposition.hpp"
#include "gridLayout.hpp"
#include "pblasDecomposition.hpp"

namespace {

int failures = 0;

void report(const char* name, bool ok, double err) {
    printf("%-48s %s (rel err %.3e)\n", name, ok ? "ok" : "FAILED", err);
    if (!ok) ++failures;
}

// C = alpha*op(A)*op(B) + beta*C, column-major, long double sums.
void host_gemm(char ta, char tb, long long m, long long n, long long k, double alpha, const double* A, long long lda,
               const double* B, long long ldb, double beta, double* C, long long ldc) {
    for (long long j = 0; j < n; ++j)
        for (long long i = 0; i < m; ++i) {
            long double s = 0;
            for (long long l = 0; l < k; ++l) {
                const double a = ta == 'N' ? A[i + l * lda] : A[l + i * lda];
                const double b = tb == 'N' ? B[l + j * ldb] : B[j + l * ldb];
                s += (long double)a * b;
            }
            const double v = (double)(alpha * s);
            C[i + j * ldc] = beta == 0.0 ? v : v + beta * C[i + j * ldc];
        }
}

void test_cannon(char ta, char tb, long long M, long long N, long long K, int p, int c) {
    const long long ra = ta == 'N' ? M : K, ca = ta == 'N' ? K : M;
    const long long rb = tb == 'N' ? K : N, cb = tb == 'N' ? N : K;
    std::vector<double> A(ra * ca), B(rb * cb), C(M * N), ref;
    setMatrixSeed(11);
    MatrixInit(A.data(), ra * ca, -1);
    MatrixInit(B.data(), rb * cb, -1);
    MatrixInit(C.data(), M * N, -1);
    ref = C;
    host_gemm(ta, tb, M, N, K, 1.25, A.data(), ra, B.data(), rb, 0.5, ref.data(), M);
    Summa25Decomposer dec(M, N, K, p, p, c, false);
    dec.setOperands(ta, tb);
    dec.decompose2D(0, A.data(), B.data(), C.data(), dec.localA, dec.localB, dec.localC);
    dec.multiply(ta, tb, A.data(), B.data(), C.data(), 1.25, 0.5);
    std::vector<double> out(M * N, 0.0);
    dec.reorderMatrix(0, out.data(), dec.localC);
    const double err = relativeFrobeniusError(ref.data(), out.data(), M * N);
    char name[96];
    snprintf(name, sizeof(name), "Summa25Decomposer %c%c %lldx%lldx%lld on %dx%dx%d", ta, tb, M, N, K, p, p, c);
    report(name, err <= 1e-13, err);
}

// The reference's digit-count check on a GPU result whose entries are bounded
// away from zero (A, B in [0.5, 1], beta = 0: every C entry is a sum of K
// products in [0.25, 1]), against the long-double host product: the FP64
// result agrees to at least 13 digits everywhere.
void test_validation_digits() {
    const long long M = 256, N = 192, K = 320;
    std::vector<double> A(M * K), B(K * N), C(M * N, 0.0), ref(M * N, 0.0);
    setMatrixSeed(13);
    MatrixInit(A.data(), M * K, -1);
    MatrixInit(B.data(), K * N, -1);
    for (double& x : A) x = 0.75 + 0.25 * x;  // [-1, 1) -> [0.5, 1)
    for (double& x : B) x = 0.75 + 0.25 * x;
    host_gemm('N', 'N', M, N, K, 1.0, A.data(), M, B.data(), K, 0.0, ref.data(), M);
    Summa25Decomposer dec(M, N, K, 2, 2, 1, false);
    dec.decompose2D(0, A.data(), B.data(), C.data(), dec.localA, dec.localB, dec.localC);
    dec.multiply('N', 'N', A.data(), B.data(), C.data(), 1.0, 0.0);
    std::vector<double> out(M * N, 0.0);
    dec.reorderMatrix(0, out.data(), dec.localC);
    const short digits = Dtest_equality(ref.data(), out.data(), M * N);
    const double err = relativeFrobeniusError(ref.data(), out.data(), M * N);
    printf("Dtest_equality digits on the 2x2x1 Cannon result: %d\n", (int)digits);
    report("Dtest_equality >= 13 digits on a GPU result", digits >= 13 && err <= 1e-14, err);
}

void test_full_offload_driver() {
    const long long n = 256;
    std::vector<double> A(n * n), B(n * n), C(n * n), ref;
    setMatrixSeed(5);
    MatrixInit(A.data(), n * n, -1);
    MatrixInit(B.data(), n * n, -1);
    MatrixInit(C.data(), n * n, -1);
    ref = C;
    host_gemm('N', 'N', n, n, n, 1.0, A.data(), n, B.data(), n, 0.0, ref.data(), n);
    fullOffloadSumma25Dgemm('N', 'N', n, n, n, 1.0, A.data(), n, B.data(), n, 0.0, C.data(), n, 2, 2, 1, 1, false, true);
    const double err = relativeFrobeniusError(ref.data(), C.data(), n * n);
    report("fullOffloadSumma25Dgemm 256^3 on 2x2x1", err <= 1e-13, err);
}

void test_pblas() {
    const long long M = 203, N = 117;
    pblasDecomposer dec(M, N, 16, 8, 6);  // 6 ranks -> 3 x 2 grid
    std::vector<double> G(M * N), back(M * N, -1.0);
    setMatrixSeed(3);
    MatrixInit(G.data(), M * N, -1);
    double* packed = nullptr;
    HIP_CHECK(hipMalloc(&packed, sizeof(double) * dec.packedSize()));
    dec.scatterMatrix(0, G.data(), packed);
    dec.gatherMatrix(0, back.data(), packed);
    // element (I, J) of rank (pr, pc)'s local matrix
    std::vector<double> pk(dec.packedSize());
    HIP_CHECK(hipMemcpy(pk.data(), packed, sizeof(double) * pk.size(), hipMemcpyDeviceToHost));
    bool ok = memcmp(G.data(), back.data(), sizeof(double) * M * N) == 0 && dec.dRow == 3 && dec.dCol == 2;
    for (long long J = 0; J < N && ok; J += 7)
        for (long long I = 0; I < M && ok; I += 5) {
            const int pr = (int)((I / 16) % 3), pc = (int)((J / 8) % 2), r = pr * 2 + pc;
            const long long li = (I / 16 / 3) * 16 + I % 16, lj = (J / 8 / 2) * 8 + J % 8;
            ok = pk[dec.localOffset(r) + li + lj * dec.localRowsOf(r)] == G[I + J * M];
        }
    HIP_CHECK(hipFree(packed));
    report("pblasDecomposer scatter/gather 203x117, 16x8 on 3x2", ok, 0.0);
}

// Row-major globals (ld = columns), as the reference's colMajor = false
// branch: C = alpha A B + beta C with every piece row-major.
void test_block_sequential_row_major() {
    const long long M = 230, N = 170, K = 70;
    BlockSequentialDecomposer dec(M, N, K, 6, false);
    std::vector<double> A(M * K), B(K * N), C(M * N), ref;
    setMatrixSeed(17);
    MatrixInit(A.data(), M * K, -1);
    MatrixInit(B.data(), K * N, -1);
    MatrixInit(C.data(), M * N, -1);
    ref = C;
    // row-major C = A B is column-major C^T = B^T A^T
    host_gemm('N', 'N', N, M, K, 0.5, B.data(), N, A.data(), K, -1.0, ref.data(), N);
    dec.deliverMatrix(A.data(), B.data(), C.data());
    // rank (1, 1)'s C tile is row-major with ld = its columns
    std::vector<double> tile(dec.tileRows(1) * dec.tileColumns(1));
    HIP_CHECK(hipMemcpy(tile.data(), dec.blockC(1, 1), sizeof(double) * tile.size(), hipMemcpyDeviceToHost));
    bool ok = dec.dRow == 3 && dec.dCol == 2;
    const long long r0 = M / 3, c0 = N / 2, tn = dec.tileColumns(1);
    for (long long r = 0; r < dec.tileRows(1) && ok; ++r)
        for (long long c = 0; c < tn && ok; ++c) ok = tile[r * tn + c] == C[(r0 + r) * N + c0 + c];
    dec.multiply(0.5, -1.0);
    std::vector<double> out(M * N);
    dec.gatherResult(0, out.data());
    const double err = relativeFrobeniusError(ref.data(), out.data(), M * N);
    report("BlockSequentialDecomposer row-major 230x170x70 on 3x2", ok && err <= 1e-13, err);
}

void test_block_sequential() {
    const long long M = 300, N = 260, K = 90;
    BlockSequentialDecomposer dec(M, N, K, 35);
    std::vector<double> A(M * K), B(K * N), C(M * N), ref;
    setMatrixSeed(7);
    MatrixInit(A.data(), M * K, -1);
    MatrixInit(B.data(), K * N, -1);
    MatrixInit(C.data(), M * N, -1);
    ref = C;
    host_gemm('N', 'N', M, N, K, 1.0, A.data(), M, B.data(), K, 2.0, ref.data(), M);
    dec.deliverMatrix(A.data(), B.data(), C.data());
    dec.multiply(1.0, 2.0);
    std::vector<double> out(M * N);
    dec.gatherResult(0, out.data());
    const double err = relativeFrobeniusError(ref.data(), out.data(), M * N);
    report("BlockSequentialDecomposer 300x260x90 on 7x5", err <= 1e-13 && dec.dRow == 7 && dec.dCol == 5 &&
                                                              dec.tileRows(6) == 300 - 6 * 42, err);
}

void test_grid_layout() {
    GridLayout g(std::vector<int64_t>{0, 3, 10, 11}, std::vector<int64_t>{0, 5, 6});
    std::vector<double> X(11 * 6), Y(11 * 6, 0.0);
    for (size_t i = 0; i < X.size(); ++i) X[i] = (double)i;
    g.scatterMatrix(X.data(), 11);
    std::vector<double> blk(g.block(1, 0).rows * g.block(1, 0).columns);
    HIP_CHECK(hipMemcpy(blk.data(), g.block(1, 0).data, sizeof(double) * blk.size(), hipMemcpyDeviceToHost));
    bool ok = blk[0] == X[3] && blk[blk.size() - 1] == X[9 + 4 * 11] && g.owners[5] == 5;
    g.gatherMatrix(Y.data(), 11);
    ok = ok && memcmp(X.data(), Y.data(), sizeof(double) * X.size()) == 0;
    report("GridLayout uneven 3x2 round trip", ok, 0.0);
}

void test_validation_and_errors() {
    std::vector<double> ref = {1.0, 2.0, -3.0, 0.0}, x = {1.0, 2.0 * (1 + 3e-11), -3.0 * 1.5, 0.0};
    const short d = Dtest_equality(ref.data(), x.data(), 4);
    bool ok = d == 11;
    dgSetErrorHandler(dgThrowingErrorHandler);
    try {
        Summa25Decomposer bad(64, 64, 64, 2, 4, 1, false);
        ok = false;
    } catch (const DgHostError& e) {
        ok = ok && e.code == DG_ABORT_GRID;
    }
    try {
        Summa25Decomposer bad(64, 64, 64, 2, 2, 1, false);
        bad.multiplyNCCL('N', 'N', nullptr, nullptr, nullptr, 1.0, 0.0);
        ok = false;
    } catch (const DgHostError& e) {
        ok = ok && e.code == DG_ABORT_WRONG_VARIANT;
    }
    dgSetErrorHandler(nullptr);
    ok = ok && truncatedDivisionRemainer(-3, 4) == 1 && charToTransOp('c') == 'T' && getSlurmNumNodes() >= 1;
    report("Dtest_equality, error codes, helpers", ok, 0.0);
}

// MatrixInit on the GPU (FP64 and FP32) equals the host generation bit for bit.
void test_generation() {
    const long long n = 10007;
    std::vector<double> hd(n), gd(n);
    std::vector<float> hf(n), gf(n);
    setMatrixSeed(99);
    MatrixInit(hd.data(), n, -1);
    setMatrixSeed(99);
    MatrixInit(hf.data(), n, -1);
    double* dd = nullptr;
    float* df = nullptr;
    bool ok = hipMalloc(&dd, n * sizeof(double)) == hipSuccess && hipMalloc(&df, n * sizeof(float)) == hipSuccess;
    if (ok) {
        setMatrixSeed(99);
        MatrixInit(dd, n, 0);
        setMatrixSeed(99);
        MatrixInit(df, n, 0);
        ok = hipMemcpy(gd.data(), dd, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(gf.data(), df, n * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
    }
    ok = ok && memcmp(hd.data(), gd.data(), n * sizeof(double)) == 0 && memcmp(hf.data(), gf.data(), n * sizeof(float)) == 0;
    for (long long i = 0; ok && i < n; ++i) ok = hf[i] == (float)hd[i];
    if (dd) (void)hipFree(dd);
    if (df) (void)hipFree(df);
    report("MatrixInit<double/float> GPU == host", ok, 0.0);
}

// writeMatrix: one line per column of a column-major array (per row of a
// row-major one), values space-separated, then an empty line; appends.
void test_write_matrix() {
    double m[6] = {1, 2, 3, 4, 5, 6};  // 2 x 3, column-major: columns (1 2) (3 4) (5 6)
    float f[4] = {0.5f, 1.5f, 2.5f, 3.5f};
    char path[] = "/tmp/dg_writematrix_XXXXXX";
    const int fd = mkstemp(path);
    bool ok = fd >= 0;
    if (ok) {
        close(fd);
        std::string name(path);
        writeMatrix(m, 2, 3, name);
        writeMatrix(f, 2, 2, name, MatrixLayout::RowMajor);
        std::ifstream in(name);
        std::stringstream ss;
        ss << in.rdbuf();
        ok = ss.str() == "1 2 \n3 4 \n5 6 \n\n0.5 1.5 \n2.5 3.5 \n\n";
        unlink(path);
    }
    report("writeMatrix column- and row-major text", ok, 0.0);
}

}  // namespace

int main() {
    test_cannon('N', 'N', 256, 256, 256, 2, 1);
    test_cannon('N', 'N', 256, 128, 384, 4, 2);
    test_cannon('N', 'T', 256, 128, 384, 4, 2);
    test_cannon('T', 'N', 192, 192, 96, 4, 4);
    test_full_offload_driver();
    test_pblas();
    test_block_sequential();
    test_block_sequential_row_major();
    test_validation_digits();
    test_grid_layout();
    test_validation_and_errors();
    test_generation();
    test_write_matrix();
    printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
