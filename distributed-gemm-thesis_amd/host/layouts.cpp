// Grid layouts of the C++ host API: GridLayout (uneven 2D blocks with
// owners), pblasDecomposer (block-cyclic) and BlockSequentialDecomposer
// (uneven row/column panels), all over packed device buffers.
#include <cmath>
#include <vector>

#include "2DBlockSequentialDecomposition.hpp"
#include "errorHandling.hpp"
#include "gridLayout.hpp"
#include "pblasDecomposition.hpp"
#include "transfers.hpp"

namespace {

// Runs `fn(devptr)` on a device view of a (possibly host) column-major
// matrix of ld x cols doubles; copies in before and out after as asked.
template <typename F>
void with_device_matrix(double* p, long long ld, long long cols, bool copy_in, bool copy_out, F fn) {
    if (isDevicePointer(p)) {
        fn(p);
        return;
    }
    double* d = nullptr;
    const int64_t w = (int64_t)sizeof(double) * ld;
    HIP_CHECK(hipMalloc(&d, (size_t)w * (size_t)(cols > 0 ? cols : 1)));
    if (copy_in) DG_CHECK(dg_copy2d_h2d(d, w, p, w, w, cols, nullptr));
    fn(d);
    HIP_CHECK(hipDeviceSynchronize());
    if (copy_out) DG_CHECK(dg_copy2d_d2h(p, w, d, w, w, cols, nullptr));
    HIP_CHECK(hipFree(d));
}

}  // namespace

// ------------------------------------------------------------------ GridLayout

GridLayout::GridLayout(long long M_, long long N_, int rb, int cb) : M(M_), N(N_), rowBlocks(rb), columnBlocks(cb) {
    if (rb < 1 || cb < 1 || M_ < 0 || N_ < 0) dgHostError(DG_ABORT_GRID, "GridLayout: bad grid %dx%d", rb, cb);
    rowSplits.resize(rb + 1);
    columnSplits.resize(cb + 1);
    DG_CHECK(dg_block_splits(M_, rb, rowSplits.data()));
    DG_CHECK(dg_block_splits(N_, cb, columnSplits.data()));
    build();
}

GridLayout::GridLayout(const std::vector<int64_t>& rs, const std::vector<int64_t>& cs)
    : rowSplits(rs), columnSplits(cs) {
    if (rs.size() < 2 || cs.size() < 2) dgHostError(DG_ABORT_GRID, "GridLayout: need at least one block per side");
    rowBlocks = (int)rs.size() - 1;
    columnBlocks = (int)cs.size() - 1;
    M = rs.back();
    N = cs.back();
    build();
}

void GridLayout::build() {
    HIP_CHECK(hipMalloc(&packed, sizeof(double) * (size_t)(M * N > 0 ? M * N : 1)));
    owners.resize((size_t)rowBlocks * columnBlocks);
    localBlocks.resize((size_t)rowBlocks * columnBlocks);
    long long off = 0;
    for (int i = 0; i < rowBlocks; ++i)
        for (int j = 0; j < columnBlocks; ++j) {
            Block& b = localBlocks[(size_t)i * columnBlocks + j];
            b.rows = rowSplits[i + 1] - rowSplits[i];
            b.columns = columnSplits[j + 1] - columnSplits[j];
            b.ld = b.rows > 0 ? b.rows : 1;
            b.rowOffset = rowSplits[i];
            b.columnOffset = columnSplits[j];
            b.owner = i * columnBlocks + j;
            b.data = packed + off;
            owners[(size_t)i * columnBlocks + j] = b.owner;
            off += b.rows * b.columns;
        }
}

GridLayout::~GridLayout() {
    if (packed) (void)hipFree(packed);
}

void GridLayout::scatterMatrix(const double* matrix, long long ld, hipStream_t stream) {
    with_device_matrix(const_cast<double*>(matrix), ld, N, true, false, [&](double* d) {
        DG_CHECK(dg_pack_block2d(d, ld, M, N, rowSplits.data(), rowBlocks, columnSplits.data(), columnBlocks, packed,
                                 stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    });
}

void GridLayout::gatherMatrix(double* matrix, long long ld, hipStream_t stream) {
    with_device_matrix(matrix, ld, N, true, true, [&](double* d) {
        DG_CHECK(dg_unpack_block2d(packed, rowSplits.data(), rowBlocks, columnSplits.data(), columnBlocks, d, ld, M, N,
                                   stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    });
}

// ------------------------------------------------------------------ block-cyclic

int numroc(int n, int nb, int iproc, int isrproc, int nprocs) {
    return (int)dg_numroc(n, nb, iproc, isrproc, nprocs);
}

void calculateProcessGrid(int* dRow, int* dCol, int size) { DG_CHECK(dg_virtual_grid(size, dRow, dCol)); }

pblasDecomposer::pblasDecomposer(long long M_, long long N_, long long Mb_, long long Nb_, int size_, int rank_)
    : M(M_), N(N_), Mb(Mb_), Nb(Nb_), rank(0), size(size_) {
    if (size_ < 1) dgHostError(DG_ABORT_GRID, "pblasDecomposer: size %d", size_);
    calculateProcessGrid(&dRow, &dCol, size_);
    init(rank_);
}

pblasDecomposer::pblasDecomposer(long long M_, long long N_, long long Mb_, long long Nb_, int dRow_, int dCol_,
                                 int rank_)
    : M(M_), N(N_), Mb(Mb_), Nb(Nb_), dRow(dRow_), dCol(dCol_), rank(0), size(dRow_ * dCol_) {
    if (dRow_ < 1 || dCol_ < 1) dgHostError(DG_ABORT_GRID, "pblasDecomposer: grid %dx%d", dRow_, dCol_);
    init(rank_);
}

void pblasDecomposer::init(int rank_) {
    if (Mb <= 0 || Nb <= 0) dgHostError(DG_ABORT_GRID, "pblasDecomposer: block sizes must be positive");
    if (rank_ < 0 || rank_ >= size) dgHostError(DG_ABORT_GRID, "pblasDecomposer: rank %d of %d", rank_, size);
    rank = rank_;
    procRow = rank / dCol;
    procColumn = rank % dCol;
    gridRows = (int)((M + Mb - 1) / Mb);
    gridColumns = (int)((N + Nb - 1) / Nb);
    rowsOf_.resize(size);
    colsOf_.resize(size);
    offset_.resize(size + 1);
    long long acc = 0;
    for (int r = 0; r < size; ++r) {
        rowsOf_[r] = dg_numroc(M, Mb, r / dCol, 0, dRow);
        colsOf_[r] = dg_numroc(N, Nb, r % dCol, 0, dCol);
        offset_[r] = acc;
        acc += rowsOf_[r] * colsOf_[r];
    }
    offset_[size] = acc;
    localRows = rowsOf_[rank];
    localColumns = colsOf_[rank];
    myBlocks = 0;
    for (int i = 0; i < gridRows; i++)
        if (i % dRow == procRow)
            for (int j = 0; j < gridColumns; j++)
                if (j % dCol == procColumn) myBlocks++;
}

pblasDecomposer::~pblasDecomposer() {}

long long pblasDecomposer::packedSize() const { return offset_[size]; }
long long pblasDecomposer::localOffset(int r) const { return offset_[r]; }
long long pblasDecomposer::localRowsOf(int r) const { return rowsOf_[r]; }
long long pblasDecomposer::localColumnsOf(int r) const { return colsOf_[r]; }

void pblasDecomposer::scatterMatrix(int senderRank, double* globalMatrix, double* localMatrix, hipStream_t stream) {
    if (senderRank < 0 || senderRank >= size) dgHostError(DG_ABORT_GRID, "scatterMatrix: sender %d", senderRank);
    with_device_matrix(globalMatrix, M, N, true, false, [&](double* d) {
        DG_CHECK(dg_pack_block_cyclic(d, M, M, N, Mb, Nb, dRow, dCol, localMatrix, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    });
}

void pblasDecomposer::gatherMatrix(int receiverRank, double* globalMatrix, double* localMatrix, hipStream_t stream) {
    if (receiverRank < 0 || receiverRank >= size) dgHostError(DG_ABORT_GRID, "gatherMatrix: receiver %d", receiverRank);
    with_device_matrix(globalMatrix, M, N, false, true, [&](double* d) {
        DG_CHECK(dg_unpack_block_cyclic(localMatrix, Mb, Nb, dRow, dCol, d, M, M, N, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    });
}

// ------------------------------------------------------------------ uneven 2D panels

BlockSequentialDecomposer::BlockSequentialDecomposer(long long M_, long long N_, long long K_, int size_,
                                                     bool colMajor_, int rank_)
    : M(M_), N(N_), K(K_), numberOfDevices(size_), colMajor(colMajor_), rank(rank_), size(size_) {
    if (size_ < 1 || rank_ < 0 || rank_ >= size_) dgHostError(DG_ABORT_GRID, "BlockSequentialDecomposer: rank %d of %d", rank_, size_);
    calculateVirtualDeviceGrid();
    rowOff_.resize(dRow + 1);
    colOff_.resize(dCol + 1);
    DG_CHECK(dg_block_splits(M, dRow, rowOff_.data()));
    DG_CHECK(dg_block_splits(N, dCol, colOff_.data()));
    processRow = rank / dCol;
    processColumn = rank % dCol;
    localM = tileRows(processRow);
    localN = tileColumns(processColumn);
    localK = K;
    HIP_CHECK(hipMalloc(&localA, sizeof(double) * (size_t)(M * K > 0 ? M * K : 1)));
    HIP_CHECK(hipMalloc(&localB, sizeof(double) * (size_t)(K * N > 0 ? K * N : 1)));
    HIP_CHECK(hipMalloc(&localC, sizeof(double) * (size_t)(M * N > 0 ? M * N : 1)));
}

BlockSequentialDecomposer::~BlockSequentialDecomposer() {
    (void)hipFree(localA);
    (void)hipFree(localB);
    (void)hipFree(localC);
}

void BlockSequentialDecomposer::calculateVirtualDeviceGrid() { DG_CHECK(dg_virtual_grid(numberOfDevices, &dRow, &dCol)); }

// Row panel i of A (rows rowOff_[i].., all K columns) at rowOff_[i]*K; column
// panel j of B (all K rows, columns colOff_[j]..) at colOff_[j]*K.
// Column-major (ld of a piece = its rows): block (i,j) of C at
// rowOff_[i]*N + rows_i*colOff_[j] -- the dg_pack_block2d layouts of A over
// a dRow x 1 grid, B over 1 x dCol and C over dRow x dCol.
// Row-major (ld of a piece = its columns, as the reference's colMajor = false
// branch, src/Decompositions/2DBlockSequentialDecomposition.cpp:89-100, 190-200):
// a row-major matrix is the column-major transpose, so A^T (K x M) packs over
// 1 x dRow, B^T (N x K) over dCol x 1 and C^T (N x M) over dCol x dRow; block
// (i,j) of C is block (j,i) of C^T, at colOff_[j]*M + columns_j*rowOff_[i].
double* BlockSequentialDecomposer::panelA(int i) const { return localA + rowOff_[i] * K; }
double* BlockSequentialDecomposer::panelB(int j) const { return localB + colOff_[j] * K; }
double* BlockSequentialDecomposer::blockC(int i, int j) const {
    if (!colMajor) return localC + colOff_[j] * M + tileColumns(j) * rowOff_[i];
    return localC + rowOff_[i] * N + tileRows(i) * colOff_[j];
}

void BlockSequentialDecomposer::deliverMatrix(double* gA, double* gB, double* gC, double** la, double** lb,
                                              double** lc) {
    const int64_t kOff[2] = {0, K};
    if (colMajor) {
        with_device_matrix(gA, M, K, true, false, [&](double* d) {
            DG_CHECK(dg_pack_block2d(d, M, M, K, rowOff_.data(), dRow, kOff, 1, localA, nullptr));
        });
        with_device_matrix(gB, K, N, true, false, [&](double* d) {
            DG_CHECK(dg_pack_block2d(d, K, K, N, kOff, 1, colOff_.data(), dCol, localB, nullptr));
        });
        if (gC) {
            with_device_matrix(gC, M, N, true, false, [&](double* d) {
                DG_CHECK(dg_pack_block2d(d, M, M, N, rowOff_.data(), dRow, colOff_.data(), dCol, localC, nullptr));
            });
        }
    } else {
        with_device_matrix(gA, K, M, true, false, [&](double* d) {
            DG_CHECK(dg_pack_block2d(d, K, K, M, kOff, 1, rowOff_.data(), dRow, localA, nullptr));
        });
        with_device_matrix(gB, N, K, true, false, [&](double* d) {
            DG_CHECK(dg_pack_block2d(d, N, N, K, colOff_.data(), dCol, kOff, 1, localB, nullptr));
        });
        if (gC) {
            with_device_matrix(gC, N, M, true, false, [&](double* d) {
                DG_CHECK(dg_pack_block2d(d, N, N, M, colOff_.data(), dCol, rowOff_.data(), dRow, localC, nullptr));
            });
        }
    }
    HIP_CHECK(hipDeviceSynchronize());
    if (la) *la = localA;
    if (lb) *lb = localB;
    if (lc) *lc = localC;
}

void BlockSequentialDecomposer::gatherResult(int gatherRank, double* C, double* lc) {
    if (gatherRank < 0 || gatherRank >= size) dgHostError(DG_ABORT_GRID, "gatherResult: rank %d", gatherRank);
    double* src = lc ? lc : localC;
    if (colMajor) {
        with_device_matrix(C, M, N, false, true, [&](double* d) {
            DG_CHECK(dg_unpack_block2d(src, rowOff_.data(), dRow, colOff_.data(), dCol, d, M, M, N, nullptr));
            HIP_CHECK(hipDeviceSynchronize());
        });
    } else {
        with_device_matrix(C, N, M, false, true, [&](double* d) {
            DG_CHECK(dg_unpack_block2d(src, colOff_.data(), dCol, rowOff_.data(), dRow, d, N, N, M, nullptr));
            HIP_CHECK(hipDeviceSynchronize());
        });
    }
}

void BlockSequentialDecomposer::multiply(double alpha, double beta, hipStream_t stream) {
    std::vector<dg_gemm_desc> descs;
    descs.reserve((size_t)dRow * dCol);
    for (int i = 0; i < dRow; ++i)
        for (int j = 0; j < dCol; ++j) {
            dg_gemm_desc d{};
            d.opA = 'N';
            d.opB = 'N';
            d.k = K;
            d.alpha = alpha;
            d.beta = beta;
            d.C = blockC(i, j);
            if (colMajor) {
                d.m = tileRows(i);
                d.n = tileColumns(j);
                d.A = panelA(i);
                d.lda = d.m > 0 ? d.m : 1;
                d.B = panelB(j);
                d.ldb = K > 0 ? K : 1;
            } else {
                // row-major C_ij = A_i B_j is column-major C_ij^T = B_j^T A_i^T
                d.m = tileColumns(j);
                d.n = tileRows(i);
                d.A = panelB(j);
                d.lda = d.m > 0 ? d.m : 1;
                d.B = panelA(i);
                d.ldb = K > 0 ? K : 1;
            }
            d.ldc = d.m > 0 ? d.m : 1;
            descs.push_back(d);
        }
    DG_CHECK(dg_grouped_dgemm(descs.data(), (int)descs.size(), stream));
    HIP_CHECK(hipStreamSynchronize(stream));
}
