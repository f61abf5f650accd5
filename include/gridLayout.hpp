// Uneven 2D grid layout with owners (the working form of the reference's
// include/gridLayout.hpp:6-27 stub: Block{data, ld, rows, columns} and
// GridLayout{rowBlocks, columnBlocks, rowSplits, columnSplits, owners,
// localBlocks; scatterMatrix}).  Blocks of one matrix live in one packed
// device buffer in owner (rank) order, each column-major with ld = its rows
// (the dg_pack_block2d layout).
#ifndef GRID_LAYOUT_HPP
#define GRID_LAYOUT_HPP

#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

class Block {
  public:
    double* data = nullptr;   // device pointer into the packed buffer
    long long ld = 0;
    long long rows = 0;
    long long columns = 0;
    long long rowOffset = 0;      // position in the global matrix
    long long columnOffset = 0;
    int owner = 0;
    void* getDataPointer() { return data; }
};

class GridLayout {
  public:
    long long M = 0, N = 0;
    int rowBlocks = 0, columnBlocks = 0;
    std::vector<int64_t> rowSplits;       // rowBlocks + 1 offsets
    std::vector<int64_t> columnSplits;    // columnBlocks + 1 offsets
    std::vector<int> owners;              // owners[i*columnBlocks + j] = rank of block (i,j) = i*columnBlocks + j
    std::vector<Block> localBlocks;       // every block (one GPU holds all virtual ranks)
    double* packed = nullptr;             // device buffer of M*N doubles (owned)

    // Even split with the remainder on the last block row / column
    // (BlockSequentialDecomposer, 2DBlockSequentialDecomposition.cpp:11-25).
    GridLayout(long long M, long long N, int rowBlocks, int columnBlocks);
    // Arbitrary splits: rowSplits from 0 to M, columnSplits from 0 to N.
    GridLayout(const std::vector<int64_t>& rowSplits, const std::vector<int64_t>& columnSplits);
    ~GridLayout();
    GridLayout(const GridLayout&) = delete;
    GridLayout& operator=(const GridLayout&) = delete;

    // Global column-major matrix (host or device, leading dimension ld) ->
    // the blocks, and back.  Bit-exact.
    void scatterMatrix(const double* matrix, long long ld, hipStream_t stream = nullptr);
    void gatherMatrix(double* matrix, long long ld, hipStream_t stream = nullptr);

    Block& block(int i, int j) { return localBlocks[(size_t)i * columnBlocks + j]; }

  private:
    void build();
};

#endif
