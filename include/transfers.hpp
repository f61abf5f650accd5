// Block copies of the C++ host API (drop-in for the reference's
// include/transfers.hpp and src/Transfers/sharedTransfers.cpp).  The MPI
// point-to-point pair transferBlock/receiveBlock has no counterpart: on one
// GPU every virtual rank's tile is addressed directly.
#ifndef TRANSFERS_HPP
#define TRANSFERS_HPP

#include <cstddef>

#include <hip/hip_runtime.h>

// Direction of a copy between the two pointers (findMemcpyKind,
// sharedTransfers.cpp:3-20; unregistered host memory counts as host).
hipMemcpyKind findMemcpyKind(const void* sourcePointer, const void* destinationPointer);

template <typename T>
void copy(size_t size, const T* sourcePointer, T* destinationPointer, hipMemcpyKind copyDirection);

// Copies a rows x columns block between leading dimensions ldSource and
// ldDestination (column-major; columnMajor = false swaps the roles as in
// sharedTransfers.cpp:72-73).  Device-to-device copies run the dg_copy_block
// kernel; copies that involve host memory use one strided hipMemcpy2D.
template <typename T>
void copyBlock(long long rows, long long columns, T* sourcePointer, T* destinationPointer,
               const long long ldSource, const long long ldDestination, bool columnMajor = true);

// hipMemcpy2D of a strided block (memcpy2D, sharedTransfers.cpp:65-81).
void memcpy2D(void* sourcePointer, long long ldSource, void* destinationPointer, long long ldDestination,
              long long rows, long long columns, size_t elementSize, bool columnMajor = true);

// True when the pointer is device memory of the current device.
bool isDevicePointer(const void* p);

#endif
