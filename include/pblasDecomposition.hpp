// ScaLAPACK block-cyclic decomposition on a virtual process grid (drop-in
// for the reference's include/pblasDecomposition.hpp and
// src/Decompositions/pblasDecomposition.cpp).  The MPI communicator becomes
// the number of virtual ranks: one GPU holds every rank's local matrix, all
// of them in one packed device buffer in rank order (dg_pack_block_cyclic).
#ifndef PBLAS_DECOMPOSITION_HPP
#define PBLAS_DECOMPOSITION_HPP

#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

// ScaLAPACK NUMROC (pblasDecomposition.cpp:3-23).
int numroc(int n, int nb, int iproc, int isrproc, int nprocs);
// Most-square process grid for `size` ranks (the reference calls
// calculateProcessGrid, pblasDecomposition.cpp:31, without defining it;
// this is the rule of calculateVirtualDeviceGrid,
// 2DBlockSequentialDecomposition.cpp:27-54).
void calculateProcessGrid(int* dRow, int* dCol, int size);

class pblasDecomposer {
  public:
    long long M, N;              // matrix dimensions
    long long Mb, Nb;            // block sizes
    int dRow, dCol;              // process grid (row-major ranks: rank = procRow*dCol + procColumn)
    int rank, size;              // rank: the "current" virtual rank the scalar fields below describe
    int procRow, procColumn;
    int myBlocks;
    long long localRows, localColumns;
    int gridRows, gridColumns;   // blocks per dimension

    // `size` virtual ranks on the most-square grid, fields for `rank`.
    pblasDecomposer(long long M, long long N, long long Mb, long long Nb, int size, int rank = 0);
    // Explicit dRow x dCol grid.
    pblasDecomposer(long long M, long long N, long long Mb, long long Nb, int dRow, int dCol, int rank);
    ~pblasDecomposer();

    // Doubles of the packed buffer that holds every rank's local matrix.
    long long packedSize() const;
    // Offset of rank r's local matrix in the packed buffer, its rows and columns.
    long long localOffset(int r) const;
    long long localRowsOf(int r) const;
    long long localColumnsOf(int r) const;

    // Global column-major matrix (ld = M, host or device) -> every rank's
    // local matrix in `localMatrix` (packed device buffer of packedSize()),
    // and back (scatterMatrix :73-148, gatherMatrix :150-221).  senderRank /
    // receiverRank name the rank that would hold the global matrix; on one
    // GPU all ranks are local, so it only has to be a valid rank.
    void scatterMatrix(int senderRank, double* globalMatrix, double* localMatrix, hipStream_t stream = nullptr);
    void gatherMatrix(int receiverRank, double* globalMatrix, double* localMatrix, hipStream_t stream = nullptr);

  private:
    std::vector<long long> rowsOf_, colsOf_, offset_;
    void init(int rank);
};

#endif
