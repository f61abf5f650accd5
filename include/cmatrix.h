// Matrix generation, printing and copying of the C++ host API (drop-in for
// the reference's include/cmatrix.h and src/Utilities/cmatrix.cpp; the
// templated include/MatrixUtilities.hpp names are provided too).
//
// Generation is seeded and reproducible (the reference seeds cuRAND and
// rand() from the wall clock): element i of a matrix is output i of
// SplitMix64 seeded with the current seed, mapped to uniform[-1, 1) by
// default (dg_generate in dgamd.h).  Device memory is filled on the GPU, host
// memory by the same generator on the CPU, bit-identically.
#ifndef CMATRIX_H
#define CMATRIX_H

#include <cstdint>
#include <string>

#include "errorHandling.hpp"
#include "generalUtilities.hpp"

// Seed and distribution (DG_DIST_*) used by MatrixInit / generateMatrix*;
// every call advances the seed by one so consecutive matrices differ.
void setMatrixSeed(uint64_t seed, int distribution = DG_DIST_UNIFORM_PM1);
uint64_t currentMatrixSeed();

/* Print related functions (cmatrix.cpp:3-23) */
void printMatrix(double* array, const long long rows, const long long columns, int rank);
void printMatrixColumnMajor(double* array, const long long rows, const long long columns, int rank);
void printInLine(double* array, const long long rows, const long long columns, int rank);

/* Write matrix to file functions (cmatrix.cpp:25-69) */
void writeMatrixToFile(double* matrix, const int rows, const int columns, const std::string& filename);
void writeMatrixToFileColumnMajor(double* matrix, const long long rows, const long long columns,
                                  const std::string& filename);

/* Copy and generate matrices functions (cmatrix.cpp:72-237) */
template <typename T>
T* copyMatrix(T* matrix, const long long rows, const long long columns);

template <typename T>
void generateMatrix(T* array, const long long rows, const long long columns);
template <typename T>
void generateMatrix(T* array, const long long size);

template <typename T>
void generateMatrixGPU(T* array, const long long rows, const long long columns, const int deviceID);
template <typename T>
void generateMatrixGPU(T* array, const long long size, const int deviceID);

// loc == -1: generate on the CPU; loc in [0, devices): on that GPU (device or
// host memory destination); anything else prints the usage and exits -2
// (cmatrix.cpp:239-297).
template <typename T>
void MatrixInit(T* matrix, const long long rows, const long long columns, int loc);
template <typename T>
void MatrixInit(T* matrix, const long long size, int loc);

#endif
