// Uneven 2D block decomposition on a virtual device grid (drop-in for the
// reference's include/2DBlockSequentialDecomposition.hpp and
// src/Decompositions/2DBlockSequentialDecomposition.cpp).  Rank (i,j) gets
// the row panel A_i (tileM x K), the column panel B_j (K x tileN) and the
// block C_ij; the last grid row / column absorbs the remainder.  All ranks'
// pieces live in packed device buffers (panels of A in rank-row order,
// panels of B in rank-column order, C blocks in rank order), and
// multiply() runs every rank's full-K GEMM in one grouped launch.
#ifndef BLOCK_SEQUENTIAL_DECOMPOSITION_HPP
#define BLOCK_SEQUENTIAL_DECOMPOSITION_HPP

#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

class BlockSequentialDecomposer {
  public:
    long long M, N, K;
    long long localM, localN, localK;  // extents of `rank`'s pieces
    int dRow, dCol, numberOfDevices;
    bool colMajor;
    int rank, size;
    int processRow, processColumn;

    // Device buffers (owned): every rank's A row panel, B column panel, C block.
    double *localA = nullptr, *localB = nullptr, *localC = nullptr;

    // `size` virtual ranks; the scalar fields describe `rank`.  colMajor =
    // false takes row-major global matrices (ld = columns) and keeps every
    // piece row-major with ld = its columns, as the reference's branch.
    BlockSequentialDecomposer(long long M, long long N, long long K, int size, bool colMajor = true, int rank = 0);
    ~BlockSequentialDecomposer();
    BlockSequentialDecomposer(const BlockSequentialDecomposer&) = delete;
    BlockSequentialDecomposer& operator=(const BlockSequentialDecomposer&) = delete;

    // dRow x dCol = most-square factorisation of `size` (:27-54).
    void calculateVirtualDeviceGrid();

    // Global A (M x K), B (K x N), C (M x N), host or device, column-major
    // (ld M, K, M) or row-major (ld K, N, N) -> every rank's pieces
    // (deliverMatrix :56-162).  The localX arguments
    // receive the owned packed buffers (may be nullptr).
    void deliverMatrix(double* globalA, double* globalB, double* globalC, double** localA_out = nullptr,
                       double** localB_out = nullptr, double** localC_out = nullptr);
    // Every rank's C block -> global C (gatherResult :164-253).
    void gatherResult(int gatherRank, double* C, double* localC_in = nullptr);

    // C_ij = alpha * A_i * B_j + beta * C_ij for every rank, one grouped launch.
    void multiply(double alpha, double beta, hipStream_t stream = nullptr);

    // Extents and packed offsets of rank r's pieces.
    long long tileRows(int i) const { return rowOff_[i + 1] - rowOff_[i]; }
    long long tileColumns(int j) const { return colOff_[j + 1] - colOff_[j]; }
    double* panelA(int i) const;
    double* panelB(int j) const;
    double* blockC(int i, int j) const;

  private:
    std::vector<int64_t> rowOff_, colOff_;
};

#endif
