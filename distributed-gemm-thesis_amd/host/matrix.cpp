// Matrix generation, printing and block copies of the C++ host API (the
// reference's src/Utilities/cmatrix.cpp, src/Utilities/MatrixUtilities.cpp and
// src/Transfers/sharedTransfers.cpp).
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <mutex>
#include <type_traits>
#include <vector>

#include "MatrixUtilities.hpp"
#include "cmatrix.h"
#include "transfers.hpp"

// ------------------------------------------------------------------ seeds

namespace {
std::mutex g_seed_mu;
uint64_t g_seed = 1;
int g_dist = DG_DIST_UNIFORM_PM1;

// The next (seed, distribution) pair; every matrix gets its own seed.
void take_seed(uint64_t* seed, int* dist) {
    std::lock_guard<std::mutex> lock(g_seed_mu);
    *seed = g_seed++;
    *dist = g_dist;
}

uint64_t splitmix64_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// CPU form of dg_generate: identical values.
void generate_host(uint64_t seed, int dist, double* out, long long n) {
    uint64_t s = seed, se = seed ^ 0xD1B54A32D192ED03ull;
    for (long long i = 0; i < n; ++i) {
        const uint64_t z = splitmix64_next(&s);
        const double u = (double)(z >> 11) * 0x1.0p-53;
        double v;
        if (dist == DG_DIST_UNIFORM_01) {
            v = u;
        } else {
            v = 2.0 * u - 1.0;
            if (dist == DG_DIST_WIDE) {
                const uint64_t ze = splitmix64_next(&se);
                v = ldexp(v, (int)((ze >> 32) % 41ull) - 20);
            }
        }
        out[i] = v;
    }
}
}  // namespace

void setMatrixSeed(uint64_t seed, int distribution) {
    std::lock_guard<std::mutex> lock(g_seed_mu);
    g_seed = seed;
    g_dist = distribution;
}

uint64_t currentMatrixSeed() {
    std::lock_guard<std::mutex> lock(g_seed_mu);
    return g_seed;
}

// ------------------------------------------------------------------ printing

void printMatrix(double* array, const long long rows, const long long columns, int rank) {
    std::cout << "RANK: " << rank << std::endl;
    for (long long i = 0; i < rows; i++) {
        for (long long j = 0; j < columns; j++) std::cout << array[i * columns + j] << " ";
        std::cout << std::endl;
    }
    std::cout << std::endl;
}

void printMatrixColumnMajor(double* array, const long long rows, const long long columns, int rank) {
    printf("RANK: %d\n", rank);
    for (long long i = 0; i < rows; i++) {
        for (long long j = 0; j < columns; j++) printf("%lf ", array[j * rows + i]);
        printf("\n");
    }
}

void printInLine(double* array, const long long rows, const long long columns, int rank) {
    std::cout << "RANK: " << rank << std::endl;
    for (long long i = 0; i < rows * columns; i++) std::cout << array[i] << " ";
    std::cout << std::endl;
}

void writeMatrixToFile(double* array, const int rows, const int columns, const std::string& filename) {
    std::ofstream out(filename, std::ofstream::out | std::ofstream::app);
    if (!out.is_open()) {
        std::cerr << "Error opening file: " << filename << std::endl;
        return;
    }
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < columns; j++) out << array[(long long)i * columns + j] << " ";
        out << std::endl;
    }
    out << std::endl;
}

void writeMatrixToFileColumnMajor(double* array, const long long rows, const long long columns,
                                  const std::string& filename) {
    std::ofstream out(filename, std::ofstream::out | std::ofstream::app);
    if (!out.is_open()) {
        std::cerr << "Error opening file: " << filename << std::endl;
        return;
    }
    for (long long i = 0; i < rows; i++) {
        for (long long j = 0; j < columns; j++) out << array[j * rows + i] << " ";
        out << std::endl;
    }
    out << std::endl;
}

// ------------------------------------------------------------------ memory kinds

bool isDevicePointer(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory reports an error: clear it
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

hipMemcpyKind findMemcpyKind(const void* src, const void* dst) {
    const bool sd = isDevicePointer(src), dd = isDevicePointer(dst);
    if (sd && dd) return hipMemcpyDeviceToDevice;
    if (!sd && dd) return hipMemcpyHostToDevice;
    if (sd && !dd) return hipMemcpyDeviceToHost;
    return hipMemcpyHostToHost;
}

template <typename T>
void copy(size_t size, const T* src, T* dst, hipMemcpyKind dir) {
    static_assert(std::is_trivially_copyable<T>::value, "elements must be trivially copyable");
    const int64_t bytes = (int64_t)(sizeof(T) * size);
    if (dir == hipMemcpyHostToDevice) DG_CHECK(dg_copy2d_h2d(dst, bytes, src, bytes, bytes, 1, nullptr));
    else if (dir == hipMemcpyDeviceToHost) DG_CHECK(dg_copy2d_d2h(dst, bytes, src, bytes, bytes, 1, nullptr));
    else HIP_CHECK(hipMemcpy(dst, src, sizeof(T) * size, dir));
}

void memcpy2D(void* src, long long lds, void* dst, long long ldd, long long rows, long long columns,
              size_t elementSize, bool columnMajor) {
    if (!src || !dst) {
        printf("On memcpy2D: Pointers not initialized correctly\n");
        return;
    }
    if (!columnMajor) std::swap(rows, columns);
    if (rows <= 0 || columns <= 0) return;
    const int64_t w = rows * (int64_t)elementSize, sp = lds * (int64_t)elementSize, dp = ldd * (int64_t)elementSize;
    const hipMemcpyKind dir = findMemcpyKind(src, dst);
    if (dir == hipMemcpyHostToDevice) DG_CHECK(dg_copy2d_h2d(dst, dp, src, sp, w, columns, nullptr));
    else if (dir == hipMemcpyDeviceToHost) DG_CHECK(dg_copy2d_d2h(dst, dp, src, sp, w, columns, nullptr));
    else HIP_CHECK(hipMemcpy2D(dst, dp, src, sp, w, columns, dir));
}

template <typename T>
void copyBlock(long long rows, long long columns, T* src, T* dst, const long long lds, const long long ldd,
               bool columnMajor) {
    if (src == NULL || dst == NULL) {
        printf("On copyBlock: Pointers not initialized correctly\n");
        return;
    }
    if (rows * columns <= 0) return;
    if (!columnMajor) std::swap(rows, columns);
    const hipMemcpyKind dir = findMemcpyKind(src, dst);
    if (dir == hipMemcpyDeviceToDevice) {
        if constexpr (std::is_same<T, double>::value) {
            DG_CHECK(dg_copy_block(rows, columns, src, lds, dst, ldd, nullptr));
        } else {
            DG_CHECK(dg_copy_block_f32(rows, columns, src, lds, dst, ldd, nullptr));
        }
        HIP_CHECK(hipDeviceSynchronize());
        return;
    }
    if (rows == lds && rows == ldd) {
        copy((size_t)(rows * columns), src, dst, dir);
        return;
    }
    memcpy2D(src, lds, dst, ldd, rows, columns, sizeof(T), true);
}

template void copy<double>(size_t, const double*, double*, hipMemcpyKind);
template void copy<float>(size_t, const float*, float*, hipMemcpyKind);
template void copyBlock<double>(long long, long long, double*, double*, const long long, const long long, bool);
template void copyBlock<float>(long long, long long, float*, float*, const long long, const long long, bool);

// ------------------------------------------------------------------ generation

template <typename T>
T* copyMatrix(T* matrix, const long long rows, const long long columns) {
    T* out = (T*)malloc(sizeof(T) * rows * columns);
    for (long long i = 0; i < rows * columns; i++) out[i] = matrix[i];
    return out;
}

template <typename T>
void generateMatrix(T* array, const long long size) {
    uint64_t seed;
    int dist;
    take_seed(&seed, &dist);
    if (std::is_same<T, double>::value) {
        generate_host(seed, dist, reinterpret_cast<double*>(array), size);
    } else {
        std::vector<double> tmp((size_t)size);
        generate_host(seed, dist, tmp.data(), size);
        for (long long i = 0; i < size; ++i) array[i] = (T)tmp[(size_t)i];
    }
}

template <typename T>
void generateMatrix(T* array, const long long rows, const long long columns) {
    generateMatrix(array, rows * columns);
}

template <typename T>
void generateMatrixGPU(T* array, const long long size, const int deviceID) {
    int prev = 0;
    HIP_CHECK(hipGetDevice(&prev));
    HIP_CHECK(hipSetDevice(deviceID));
    uint64_t seed;
    int dist;
    take_seed(&seed, &dist);
    const bool dev = isDevicePointer(array);
    T* target = array;
    if (!dev) HIP_CHECK(hipMalloc(&target, sizeof(T) * (size_t)(size > 0 ? size : 1)));
    if constexpr (std::is_same<T, double>::value) {
        DG_CHECK(dg_generate(seed, dist, target, size, nullptr));
    } else {
        DG_CHECK(dg_generate_f32(seed, dist, target, size, nullptr));
    }
    if (target != array) {
        const int64_t bytes = (int64_t)(sizeof(T) * size);
        DG_CHECK(dg_copy2d_d2h(array, bytes, target, bytes, bytes, 1, nullptr));
        HIP_CHECK(hipFree(target));
    }
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipSetDevice(prev));
}

template <typename T>
void generateMatrixGPU(T* array, const long long rows, const long long columns, const int deviceID) {
    generateMatrixGPU(array, rows * columns, deviceID);
}

template <typename T>
void MatrixInit(T* matrix, const long long size, int loc) {
    if (!matrix) {
        printf("Matrix not initialized in the right way\n");
        dgHostError(-1, "MatrixInit: null matrix");
        return;
    }
    int maxGPUs = 0;
    if (hipGetDeviceCount(&maxGPUs) != hipSuccess) maxGPUs = 0;
    if (loc < -1 || loc >= maxGPUs) {
        printf("Location == -1 (CPU), Location {0,...gpuSize-1} == DeviceID \n");
        dgHostError(-2, "MatrixInit: location %d outside -1..%d", loc, maxGPUs - 1);
        return;
    }
    if (loc == -1) generateMatrix(matrix, size);
    else generateMatrixGPU(matrix, size, loc);
}

template <typename T>
void MatrixInit(T* matrix, const long long rows, const long long columns, int loc) {
    MatrixInit(matrix, rows * columns, loc);
}

template double* copyMatrix<double>(double*, const long long, const long long);
template float* copyMatrix<float>(float*, const long long, const long long);
template void generateMatrix<double>(double*, const long long, const long long);
template void generateMatrix<float>(float*, const long long, const long long);
template void generateMatrix<double>(double*, const long long);
template void generateMatrix<float>(float*, const long long);
template void generateMatrixGPU<double>(double*, const long long, const long long, const int);
template void generateMatrixGPU<float>(float*, const long long, const long long, const int);
template void generateMatrixGPU<double>(double*, const long long, const int);
template void generateMatrixGPU<float>(float*, const long long, const int);
template void MatrixInit<double>(double*, const long long, const long long, int);
template void MatrixInit<float>(float*, const long long, const long long, int);
template void MatrixInit<double>(double*, const long long, int);
template void MatrixInit<float>(float*, const long long, int);

// ------------------------------------------------------------------ templated I/O (MatrixUtilities.hpp)

template <typename scalar_t>
void printMatrix(scalar_t* array, long long rows, long long columns, int rank, MatrixLayout layout) {
    if (!array) {
        dgHostError(-1, "printMatrix: null array");
        return;
    }
    printf("Rank: %d\n", rank);
    if (layout == MatrixLayout::ColumnMajor) std::swap(rows, columns);
    for (long long i = 0; i < rows; ++i) {
        for (long long j = 0; j < columns; ++j) printf("%lf ", (double)array[i * columns + j]);
        printf("\n");
    }
    printf("\n");
}

template <typename scalar_t>
void writeMatrix(scalar_t* array, long long rows, long long columns, std::string& filename, MatrixLayout layout) {
    std::ofstream out(filename, std::ofstream::out | std::ofstream::app);
    if (!out.is_open()) {
        std::cerr << "Error opening file: " << filename << std::endl;
        return;
    }
    if (layout == MatrixLayout::ColumnMajor) std::swap(rows, columns);
    for (long long i = 0; i < rows; ++i) {
        for (long long j = 0; j < columns; ++j) out << array[i * columns + j] << " ";
        out << std::endl;
    }
    out << std::endl;
}

template void printMatrix<float>(float*, long long, long long, int, MatrixLayout);
template void printMatrix<double>(double*, long long, long long, int, MatrixLayout);
template void writeMatrix<float>(float*, long long, long long, std::string&, MatrixLayout);
template void writeMatrix<double>(double*, long long, long long, std::string&, MatrixLayout);
