// General utilities of the C++ host API (drop-in for the reference's
// include/generalUtilities.hpp and src/Utilities/{GeneralUtilities,TransposeEmblems}.cpp).
#ifndef GENERAL_UTILITIES_HPP
#define GENERAL_UTILITIES_HPP

#include "errorHandling.hpp"

// Nodes of the SLURM job, 1 when SLURM_JOB_NUM_NODES is unset (the
// reference's atoi(getenv(...)) crashes there, GeneralUtilities.cpp:3-6).
int getSlurmNumNodes();

// GFLOP/s = 2*M*N*K*1e-9 / seconds (GeneralUtilities.cpp:8-11).
double calculateGflops(const long long M, const long long N, const long long K, const double executionTime);

// 'n'/'N' -> 'N', 't'/'T'/'c'/'C' -> 'T'; anything else warns and returns 'N'
// (charToCublasTransOp / charToCblasTransOp, TransposeEmblems.cpp:17-65).
char charToTransOp(char operation);
// 'N' -> 'n', 'T' -> 't' (cblasTransOpToChar / cublasTransOpToChar).
char transOpToChar(char operation);

// Free and total device memory in bytes (GeneralUtilities.cpp:13-35).
void getGPUMemoryInfo(long long* freeMemory, long long* maxMemory, const int deviceID);

// Wall-clock seconds with microsecond resolution (GeneralUtilities.cpp:37-53).
double csecond(void);

#endif
