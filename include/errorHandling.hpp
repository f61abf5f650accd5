// Error handling of the C++ host API (drop-in for the reference's
// include/errorHandling.hpp:5-56, whose *_CHECK macros print file:line and
// exit(EXIT_FAILURE), and whose drivers MPI_Abort with codes -1/-2/-5/-6).
//
// Every failure goes through dgHostError(code, message).  The default
// handler prints the message and calls exit(code) like the reference; tests
// and embedding applications can install a handler that throws instead.
#ifndef ERROR_HANDLING_HPP
#define ERROR_HANDLING_HPP

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime.h>

#include "dgamd.h"

// Exit codes of the reference's MPI_Abort calls (src/25DCannon.cpp:7-23,
// src/25DCannonNCCL.cpp:45-47, 111-114) and MatrixInit (src/Utilities/cmatrix.cpp:250-254).
enum DgAbortCode {
    DG_ABORT_NON_SQUARE_MATRIX = -1,
    DG_ABORT_GRID = -2,
    DG_ABORT_NO_MEMORY = -5,
    DG_ABORT_WRONG_VARIANT = -6,
};

class DgHostError : public std::runtime_error {
  public:
    DgHostError(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

typedef void (*DgErrorHandler)(int code, const char* message);

// Installs a handler (nullptr restores print-and-exit); returns the old one.
DgErrorHandler dgSetErrorHandler(DgErrorHandler handler);
// Handler that throws DgHostError, for tests and embedding applications.
void dgThrowingErrorHandler(int code, const char* message);
// Reports a failure through the installed handler (does not return when the
// handler exits or throws).
void dgHostError(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIP_CHECK(call)                                                                         \
    do {                                                                                        \
        hipError_t status_ = (call);                                                            \
        if (status_ != hipSuccess)                                                              \
            dgHostError(EXIT_FAILURE, "HIP error at %s:%d : %s", __FILE__, __LINE__,            \
                        hipGetErrorString(status_));                                            \
    } while (0)

#define DG_CHECK(call)                                                                          \
    do {                                                                                        \
        int status_ = (call);                                                                   \
        if (status_ != DG_OK)                                                                   \
            dgHostError(EXIT_FAILURE, "dgamd error at %s:%d : status %d: %s", __FILE__, __LINE__, \
                        status_, dg_last_error());                                              \
    } while (0)

#endif
