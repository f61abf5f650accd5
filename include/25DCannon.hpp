// 2D / 2.5D Cannon on a virtual p x p x c process grid, one GPU (drop-in for
// the reference's include/25DCannon.hpp, src/25DCannon.cpp and
// src/25DCannonNCCL.cpp).
//
// One object drives every virtual rank.  The ranks' tiles of layer 0 live in
// packed device buffers (tile (i,j) of the stored matrix at rank i*p + j,
// dg_pack_block2d layout); the skew and rotations are index rotations of
// dg_cannon25d, the stack broadcast is free (all layers read the same tiles)
// and the c - 1 extra layers are partial C buffers summed in HBM.
//
// Generalises the reference: M, N, K may differ (p must divide each) and
// 'T' operands multiply the transposed tiles.  Like the reference, the grid
// must be square (dRow == dCol, else exit -2) and c must divide p.  A and B
// are never modified, so repeated multiply() calls see the same operands.
#ifndef CANNON_HPP
#define CANNON_HPP

#include <hip/hip_runtime.h>

#include "cmatrix.h"
#include "errorHandling.hpp"
#include "generalUtilities.hpp"
#include "logging.hpp"
#include "transfers.hpp"
#include "validation.hpp"

// ((a % b) + b) % b (src/25DCannonNCCL.cpp:6-9).
int truncatedDivisionRemainer(int a, int b);

class Summa25Decomposer {
  public:
    long long M, N, K;
    int dRow, dCol, c;   // dRow = dCol = p
    int rank, size;      // size = p*p*c virtual ranks; rank 0 drives them all
    int processRow, processColumn, processStack;  // (i,j,k) of `rank`
    int pc3, pc;         // p/c rotation steps per layer, p
    bool broadcastComplete;
    int schedule;        // DG_SCHED_* used by multiply()

    // Layer-0 tiles of every rank (device, owned): A and B tiles of the stored
    // matrices, C tiles; workspace holds the (c-1) partial layers.
    double *localA, *localB, *localC;
    double* workspace;
    long long localM, localN, localK;
    long long llda, lldb, lldc;
    double communicationTime, executionTime;  // communicationTime stays 0: nothing moves
    bool communicationComputationOverlap;
    hipStream_t stream;

    Summa25Decomposer(long long M, long long N, long long K, int dRow, int dCol, int c,
                      bool communicationComputationOverlap);
    ~Summa25Decomposer();
    Summa25Decomposer(const Summa25Decomposer&) = delete;
    Summa25Decomposer& operator=(const Summa25Decomposer&) = delete;

    // C_ij = alpha * sum_m op(A)_im op(B)_mj + beta * C_ij on every rank
    // (multiply, src/25DCannon.cpp:286-375).  The A, B, C arguments are
    // ignored, as in the reference: the operands are the packed tiles.
    void multiply(char TransA, char TransB, double* A, double* B, double* C, double alpha, double beta);
    // The NCCL variants of the reference (src/25DCannonNCCL.cpp:109-304)
    // compute the same product; here they are the same call.
    void multiplyNCCL(char TransA, char TransB, double* A, double* B, double* C, double alpha, double beta);
    void serializedMultiplyNCCL(char TransA, char TransB, double* A, double* B, double* C, double alpha,
                                double beta);

    // Global A (rows x cols of the stored matrix, ld = rows), B, C (host or
    // device) -> the packed tiles aTile/bTile/cTile (decompose2D :87-149).
    // The stored shapes follow the last TransA/TransB given to setOperands
    // (default 'N','N': A is M x K, B is K x N).
    void decompose2D(int senderRank, double* A, double* B, double* C, double* aTile, double* bTile, double* cTile);
    // Packed C tiles -> global C (ld = M, host or device) (reorderMatrix :231-284).
    void reorderMatrix(int gatherRank, double* C, double* cTile);
    void resetMetricCounters();

    // Stored-operand layout used by decompose2D ('T': A stored K x M, B stored N x K).
    void setOperands(char TransA, char TransB);

    // Device pointer of rank (i,j)'s layer-0 tile.
    double* tileA(int i, int j) const;
    double* tileB(int i, int j) const;
    double* tileC(int i, int j) const;

  private:
    char opA_ = 'N', opB_ = 'N';
};

// Drivers of the reference (src/25DCannon.cpp:392-538,
// src/25DCannonNCCL.cpp:322-400).  Each builds a decomposer, initialises or
// distributes the operands, runs numberOfRuns multiplies, appends one CSV
// line per run when `logging` (file DGEMM_execution_logs-<mode>-<machine>-
// 2.5DCANNON-<backend>.csv, header kCannonCsvHeader) and gathers C into the
// caller's C (ld = M) when `gatherResults`.
void preDistributedSumma25Dgemm(char TransA, char TransB, long long M, long long N, long long K, double alpha,
                                double* A, long long lda, double* B, long long ldb, double beta, double* C,
                                long long ldc, int dRow, int dCol, int c, int numberOfRuns, bool logging,
                                bool gatherResults);

void fullOffloadSumma25Dgemm(char TransA, char TransB, long long M, long long N, long long K, double alpha,
                             double* A, long long lda, double* B, long long ldb, double beta, double* C,
                             long long ldc, int dRow, int dCol, int c, int numberOfRuns, bool logging,
                             bool gatherResults);

void preDistributedCannon25NCCL(char TransA, char TransB, long long M, long long N, long long K, double alpha,
                                double* A, long long lda, double* B, long long ldb, double beta, double* C,
                                long long ldc, int dRow, int dCol, int c, int numberOfRuns, bool logging,
                                bool gatherResults);

#endif
