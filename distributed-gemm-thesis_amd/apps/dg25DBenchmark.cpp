// Command-line driver of the 2.5D Cannon hot path on one GPU, with the
// argument list of the reference's cublas25DBenchmark
// (src/Benchmarks/cuBLAS25Benchmark.cpp:12-46):
//
//   dg25DBenchmark '{TransA}{TransB}' M N K alpha beta dRow dCol c numberOfRuns {host|devices} [fullOffload]
//
// dRow x dCol x c is the virtual process grid.  "devices" generates every
// rank's tiles on the GPU (pre-distributed); a 13th argument selects the
// full-offload path (global host matrices decomposed onto the grid and C
// gathered back).  Each run appends a CSV line like the reference.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "25DCannon.hpp"

int main(int argc, char* argv[]) {
    if (argc < 12) {
        fprintf(stderr,
                "Usage: ./dg25DBenchmark '{TransA}{TransB}' M N K alpha beta dRow dCol C numberOfRuns "
                "dataLocation={host,devices} fullOffload{default=false}\n");
        return 1;
    }
    const char transA = argv[1][0];
    const char transB = argv[1][1] ? argv[1][1] : 'N';
    const long long M = atoll(argv[2]), N = atoll(argv[3]), K = atoll(argv[4]);
    const double alpha = atof(argv[5]), beta = atof(argv[6]);
    const int dRow = atoi(argv[7]), dCol = atoi(argv[8]), c = atoi(argv[9]);
    const int runs = atoi(argv[10]);
    const std::string loc = argv[11];
    const bool fullOffload = argc == 13;
    if (loc != "host" && loc != "devices") {
        printf("Initial Data location not recognized, check usage\n");
        return 2;
    }
    std::vector<double> C((size_t)(M * N));
    if (fullOffload) {
        printf("Running Full Offload 2.5D Cannon Run (%d runs) with M=%lld, N=%lld, K=%lld on a %dx%dx%d virtual grid\n",
               runs, M, N, K, dRow, dCol, c);
        std::vector<double> A((size_t)(M * K)), B((size_t)(K * N));
        MatrixInit(A.data(), M * K, loc == "host" ? -1 : 0);
        MatrixInit(B.data(), K * N, loc == "host" ? -1 : 0);
        MatrixInit(C.data(), M * N, loc == "host" ? -1 : 0);
        fullOffloadSumma25Dgemm(transA, transB, M, N, K, alpha, A.data(), M, B.data(), K, beta, C.data(), M, dRow, dCol,
                                c, runs, true, true);
    } else {
        printf("Running Pre Distributed 2.5D Cannon Run (%d runs) with M=%lld, N=%lld, K=%lld on a %dx%dx%d virtual grid\n",
               runs, M, N, K, dRow, dCol, c);
        preDistributedSumma25Dgemm(transA, transB, M, N, K, alpha, nullptr, M, nullptr, K, beta, C.data(), M, dRow, dCol,
                                   c, runs, true, false);
    }
    return 0;
}
