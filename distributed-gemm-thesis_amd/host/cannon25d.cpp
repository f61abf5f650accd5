// Summa25Decomposer and the 2.5D Cannon drivers of the C++ host API (the
// reference's src/25DCannon.cpp and src/25DCannonNCCL.cpp) over the virtual
// grid executor dg_cannon25d.
#include <cmath>
#include <string>
#include <vector>

#include "25DCannon.hpp"

int truncatedDivisionRemainer(int a, int b) { return ((a % b) + b) % b; }

Summa25Decomposer::Summa25Decomposer(long long M_, long long N_, long long K_, int dRow_, int dCol_, int c_,
                                     bool overlap)
    : M(M_), N(N_), K(K_), dRow(dRow_), dCol(dCol_), c(c_), rank(0), broadcastComplete(false),
      schedule(DG_SCHED_AUTO), localA(nullptr), localB(nullptr), localC(nullptr), workspace(nullptr),
      communicationTime(0), executionTime(0), communicationComputationOverlap(overlap), stream(nullptr) {
    // src/25DCannon.cpp:7-23: square process grid, p*p*c ranks.  Rectangular
    // matrices are accepted here as long as p divides every dimension.
    if (dRow != dCol) {
        printf("2.5D Summa currently supports only Square Process Grids (dRow = dCol)\n");
        dgHostError(DG_ABORT_GRID, "Summa25Decomposer: dRow %d != dCol %d", dRow, dCol);
        return;
    }
    if (c < 1 || dRow < 1 || dRow % c != 0) {
        printf("Process Dimensions not matching with size, change communicator size\n");
        dgHostError(DG_ABORT_GRID, "Summa25Decomposer: c=%d must divide p=%d", c, dRow);
        return;
    }
    if (M % dRow || N % dRow || K % dRow) {
        dgHostError(DG_ABORT_NON_SQUARE_MATRIX, "Summa25Decomposer: p=%d must divide M=%lld, N=%lld, K=%lld", dRow, M,
                    N, K);
        return;
    }
    size = dRow * dCol * c;
    processStack = rank / (dRow * dCol);
    processRow = (rank / dCol) % dRow;
    processColumn = (rank % dCol) % dRow;
    pc3 = dRow / c;
    pc = dRow;
    localM = M / dRow;
    localN = N / dCol;
    localK = K / dRow;
    llda = localM;
    lldb = localK;
    lldc = localM;
    HIP_CHECK(hipMalloc(&localA, sizeof(double) * (size_t)(M * K > 0 ? M * K : 1)));
    HIP_CHECK(hipMalloc(&localB, sizeof(double) * (size_t)(K * N > 0 ? K * N : 1)));
    HIP_CHECK(hipMalloc(&localC, sizeof(double) * (size_t)(M * N > 0 ? M * N : 1)));
    const long long ws = dg_cannon25d_workspace(M, N, dRow, c);
    if (ws > 0) HIP_CHECK(hipMalloc(&workspace, sizeof(double) * (size_t)ws));
}

Summa25Decomposer::~Summa25Decomposer() {
    (void)hipFree(localA);
    (void)hipFree(localB);
    (void)hipFree(localC);
    if (workspace) (void)hipFree(workspace);
}

void Summa25Decomposer::setOperands(char TransA, char TransB) {
    opA_ = charToTransOp(TransA);
    opB_ = charToTransOp(TransB);
}

double* Summa25Decomposer::tileA(int i, int j) const { return localA + ((long long)i * dCol + j) * (M / dRow) * (K / dRow); }
double* Summa25Decomposer::tileB(int i, int j) const { return localB + ((long long)i * dCol + j) * (K / dRow) * (N / dRow); }
double* Summa25Decomposer::tileC(int i, int j) const { return localC + ((long long)i * dCol + j) * localM * localN; }

void Summa25Decomposer::multiply(char TransA, char TransB, double* A, double* B, double* C, double alpha, double beta) {
    (void)A; (void)B; (void)C;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, stream));
    const long long ws = dg_cannon25d_workspace(M, N, dRow, c);
    DG_CHECK(dg_cannon25d(charToTransOp(TransA), charToTransOp(TransB), M, N, K, alpha, localA, localB, beta, localC,
                          dRow, c, schedule, workspace, ws, stream));
    HIP_CHECK(hipEventRecord(e1, stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    executionTime += ms / 1e3;
    broadcastComplete = true;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void Summa25Decomposer::multiplyNCCL(char TransA, char TransB, double* A, double* B, double* C, double alpha,
                                     double beta) {
    // src/25DCannonNCCL.cpp:111-114: the NCCL variant needs the overlap setup.
    if (!communicationComputationOverlap) {
        printf("Call multiply, not multiply NNCL\n");
        dgHostError(DG_ABORT_WRONG_VARIANT, "multiplyNCCL on a decomposer built without overlap");
        return;
    }
    multiply(TransA, TransB, A, B, C, alpha, beta);
}

void Summa25Decomposer::serializedMultiplyNCCL(char TransA, char TransB, double* A, double* B, double* C, double alpha,
                                               double beta) {
    multiplyNCCL(TransA, TransB, A, B, C, alpha, beta);
}

void Summa25Decomposer::decompose2D(int senderRank, double* A, double* B, double* C, double* aTile, double* bTile,
                                    double* cTile) {
    if (senderRank < 0 || senderRank >= size) dgHostError(DG_ABORT_GRID, "decompose2D: sender %d", senderRank);
    const int p = dRow;
    auto pack = [&](double* G, long long rows, long long cols, double* out) {
        if (!G || !out) return;
        std::vector<int64_t> ro(p + 1), co(p + 1);
        DG_CHECK(dg_block_splits(rows, p, ro.data()));
        DG_CHECK(dg_block_splits(cols, p, co.data()));
        double* d = G;
        const bool host = !isDevicePointer(G);
        if (host) {
            HIP_CHECK(hipMalloc(&d, sizeof(double) * (size_t)(rows * cols > 0 ? rows * cols : 1)));
            const int64_t w = (int64_t)sizeof(double) * rows;
            DG_CHECK(dg_copy2d_h2d(d, w, G, w, w, cols, stream));
        }
        DG_CHECK(dg_pack_block2d(d, rows, rows, cols, ro.data(), p, co.data(), p, out, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        if (host) HIP_CHECK(hipFree(d));
    };
    const long long ra = opA_ == 'N' ? M : K, ca = opA_ == 'N' ? K : M;
    const long long rb = opB_ == 'N' ? K : N, cb = opB_ == 'N' ? N : K;
    pack(A, ra, ca, aTile);
    pack(B, rb, cb, bTile);
    pack(C, M, N, cTile);
    broadcastComplete = false;
}

void Summa25Decomposer::reorderMatrix(int gatherRank, double* C, double* cTile) {
    if (gatherRank < 0 || gatherRank >= size) dgHostError(DG_ABORT_GRID, "reorderMatrix: rank %d", gatherRank);
    const int p = dRow;
    std::vector<int64_t> ro(p + 1), co(p + 1);
    DG_CHECK(dg_block_splits(M, p, ro.data()));
    DG_CHECK(dg_block_splits(N, p, co.data()));
    double* d = C;
    const bool host = !isDevicePointer(C);
    if (host) HIP_CHECK(hipMalloc(&d, sizeof(double) * (size_t)(M * N > 0 ? M * N : 1)));
    DG_CHECK(dg_unpack_block2d(cTile, ro.data(), p, co.data(), p, d, M, M, N, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (host) {
        const int64_t w = (int64_t)sizeof(double) * M;
        DG_CHECK(dg_copy2d_d2h(C, w, d, w, w, N, stream));
        HIP_CHECK(hipFree(d));
    }
}

void Summa25Decomposer::resetMetricCounters() {
    communicationTime = 0;
    executionTime = 0;
}

namespace {

enum Mode { PRE_DISTRIBUTED, FULL_OFFLOAD };

void run_driver(Mode mode, const char* backend, char TransA, char TransB, long long M, long long N, long long K,
                double alpha, double* A, double* B, double beta, double* C, int dRow, int dCol, int c,
                int numberOfRuns, bool logging, bool gatherResults) {
    const int numberOfNodes = getSlurmNumNodes();
    FILE* logfile = nullptr;
    if (logging) {
        std::string filename = std::string("DGEMM_execution_logs-") +
                               (mode == FULL_OFFLOAD ? "FullOffload_GEMM-" : "PreDistributed_GEMM-") +
                               getMachineName() + "-2.5DCANNON-" + backend + ".csv";
        logfile = createLogCsv(filename, kCannonCsvHeader);
    }
    const double t0 = csecond();
    Summa25Decomposer decomposer(M, N, K, dRow, dCol, c, false);
    decomposer.setOperands(TransA, TransB);
    if (mode == FULL_OFFLOAD) {
        decomposer.decompose2D(0, A, B, C, decomposer.localA, decomposer.localB, decomposer.localC);
    } else {
        // Pre-distributed: every rank generates its own tiles on the device
        // (src/25DCannon.cpp:489-496).
        MatrixInit(decomposer.localA, M * K, 0);
        MatrixInit(decomposer.localB, K * N, 0);
        MatrixInit(decomposer.localC, M * N, 0);
    }
    const double t1 = csecond();
    for (int run = 0; run < numberOfRuns; ++run) {
        const double b0 = csecond();
        decomposer.multiply(TransA, TransB, A, B, C, alpha, beta);
        const double b1 = csecond();
        if (gatherResults && C) decomposer.reorderMatrix(0, C, decomposer.localC);
        const double b2 = csecond();
        if (logging) {
            char line[512];
            snprintf(line, sizeof(line), "%s,%lld,%lld,%lld,%d,%d,%d,%d,%d,%lf,%lf,%lf,%lf,%s,%lf,%lf\n",
                     "2.5D CANNON cuBLAS", M, N, K, numberOfNodes, dRow * dCol * c, dRow, dCol, c, t1 - t0, b1 - b0,
                     b2 - b1, calculateGflops(M, N, K, b1 - b0), mode == FULL_OFFLOAD ? "host" : "devices",
                     decomposer.communicationTime, decomposer.executionTime);
            writeLineToFile(logfile, line);
            decomposer.resetMetricCounters();
        }
    }
    if (logfile) fclose(logfile);
}

}  // namespace

void preDistributedSumma25Dgemm(char TransA, char TransB, long long M, long long N, long long K, double alpha,
                                double* A, long long lda, double* B, long long ldb, double beta, double* C,
                                long long ldc, int dRow, int dCol, int c, int numberOfRuns, bool logging,
                                bool gatherResults) {
    (void)lda; (void)ldb; (void)ldc;
    run_driver(PRE_DISTRIBUTED, "cuBLAS", TransA, TransB, M, N, K, alpha, A, B, beta, C, dRow, dCol, c, numberOfRuns,
               logging, gatherResults);
}

void fullOffloadSumma25Dgemm(char TransA, char TransB, long long M, long long N, long long K, double alpha, double* A,
                             long long lda, double* B, long long ldb, double beta, double* C, long long ldc, int dRow,
                             int dCol, int c, int numberOfRuns, bool logging, bool gatherResults) {
    (void)lda; (void)ldb; (void)ldc;
    run_driver(FULL_OFFLOAD, "cuBLAS", TransA, TransB, M, N, K, alpha, A, B, beta, C, dRow, dCol, c, numberOfRuns,
               logging, gatherResults);
}

void preDistributedCannon25NCCL(char TransA, char TransB, long long M, long long N, long long K, double alpha,
                                double* A, long long lda, double* B, long long ldb, double beta, double* C,
                                long long ldc, int dRow, int dCol, int c, int numberOfRuns, bool logging,
                                bool gatherResults) {
    (void)lda; (void)ldb; (void)ldc;
    run_driver(PRE_DISTRIBUTED, "cuBLAS-NCCL", TransA, TransB, M, N, K, alpha, A, B, beta, C, dRow, dCol, c,
               numberOfRuns, logging, gatherResults);
}
