// Error handling, general utilities, validation and CSV logging of the C++
// host API (the reference's src/Utilities/{errorHandling,GeneralUtilities,
// TransposeEmblems,Validation,Logging}.cpp).
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <limits.h>
#include <mutex>

#include "errorHandling.hpp"
#include "generalUtilities.hpp"
#include "logging.hpp"
#include "validation.hpp"

// ------------------------------------------------------------------ errors

namespace {
std::mutex g_handler_mu;
DgErrorHandler g_handler = nullptr;
}  // namespace

DgErrorHandler dgSetErrorHandler(DgErrorHandler handler) {
    std::lock_guard<std::mutex> lock(g_handler_mu);
    DgErrorHandler old = g_handler;
    g_handler = handler;
    return old;
}

void dgThrowingErrorHandler(int code, const char* message) { throw DgHostError(code, message); }

void dgHostError(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    DgErrorHandler h;
    {
        std::lock_guard<std::mutex> lock(g_handler_mu);
        h = g_handler;
    }
    if (h) {
        h(code, buf);
        return;
    }
    fprintf(stderr, "%s\n", buf);
    exit(code);
}

// ------------------------------------------------------------------ general utilities

int getSlurmNumNodes() {
    const char* v = getenv("SLURM_JOB_NUM_NODES");
    if (!v || !*v) return 1;
    int n = atoi(v);
    return n > 0 ? n : 1;
}

double calculateGflops(const long long M, const long long N, const long long K, const double executionTime) {
    return (2.0 * (double)M * (double)N * (double)K * 1e-9) / executionTime;
}

char charToTransOp(char operation) {
    switch (operation) {
        case 'n': case 'N': return 'N';
        case 't': case 'T': case 'c': case 'C': return 'T';
        default:
            std::cerr << "Error: Invalid Transpose Operation Character: " << operation << ". Returning 'N' Transpose"
                      << std::endl;
            return 'N';
    }
}

char transOpToChar(char operation) { return charToTransOp(operation) == 'T' ? 't' : 'n'; }

void getGPUMemoryInfo(long long* freeMemory, long long* maxMemory, const int deviceID) {
    int prev = 0;
    HIP_CHECK(hipGetDevice(&prev));
    HIP_CHECK(hipSetDevice(deviceID));
    size_t fr = 0, mx = 0;
    HIP_CHECK(hipMemGetInfo(&fr, &mx));
    *freeMemory = (long long)fr;
    *maxMemory = (long long)mx;
    HIP_CHECK(hipSetDevice(prev));
}

// Seconds on the monotonic clock (callers take differences of two readings).
double csecond(void) {
    const auto t = std::chrono::steady_clock::now().time_since_epoch();
    return std::chrono::duration<double>(t).count();
}

// ------------------------------------------------------------------ validation

template <typename T>
T abs(T x) {
    return x < 0 ? -x : x;
}

template <typename T>
T Terror(T a, T b) {
    if (a == 0) return abs<T>(a - b);
    return abs<T>(a - b) / a;
}

template <typename T>
long int vectorDifference(T* a, T* b, long long size, double eps) {
    long int failed = 0;
    for (long long i = 0; i < size; i++)
        if (Terror(a[i], b[i]) > eps) failed++;
    return failed;
}

namespace {
// The digit-count loop shared by testEquality and Dtest_equality.
template <typename T, typename Err>
short digits(T* ref, T* x, long long size, Err err) {
    long int acc = 8, failed;
    double eps = 1e-8;
    auto count = [&](double e) {
        long int f = 0;
        for (long long i = 0; i < size; i++)
            if (err(ref[i], x[i]) > e) f++;
        return f;
    };
    failed = count(eps);
    while (eps > DBL_MIN && !failed && acc < 30) {
        eps *= 0.1;
        acc++;
        failed = count(eps);
    }
    if (8 == acc) {
        fprintf(stderr, "Test failed %ld times\n", failed);
        int ctr = 0;
        long long itt = 0;
        while (ctr < 10 && itt < size) {
            if (err(ref[itt], x[itt]) > eps) {
                fprintf(stderr, "Baseline vs Tested(adr = %p, itt = %lld): %.15lf vs %.15lf\n", (void*)&x[itt], itt,
                        (double)ref[itt], (double)x[itt]);
                ctr++;
            }
            itt++;
        }
        return 0;
    }
    fprintf(stderr, "Test passed(Accuracy= %ld digits, %ld/%lld breaking for %ld)\n\n", acc, failed, size, acc + 1);
    return (short)acc;
}
}  // namespace

template <typename T>
short testEquality(T* referenceArray, T* array, long long size) {
    return digits(referenceArray, array, size, [](T a, T b) { return (double)Terror<T>(a, b); });
}

double dabs(double x) { return x < 0 ? -x : x; }

double Derror(double a, double b) {
    if (a == 0) return dabs(a - b);
    return dabs(a - b) / a;
}

long int Dvec_diff(double* a, double* b, long long size, double eps) {
    long int failed = 0;
    for (long long i = 0; i < size; i++)
        if (Derror(a[i], b[i]) > eps) failed++;
    return failed;
}

short Dtest_equality(double* C_comp, double* C, long long size) { return digits(C_comp, C, size, Derror); }

double relativeFrobeniusError(const double* C_ref, const double* C, long long size) {
    long double num = 0.0L, den = 0.0L;
    for (long long i = 0; i < size; ++i) {
        const long double d = (long double)C[i] - (long double)C_ref[i];
        num += d * d;
        den += (long double)C_ref[i] * (long double)C_ref[i];
    }
    if (den == 0.0L) return num == 0.0L ? 0.0 : (double)sqrtl(num);
    return (double)sqrtl(num / den);
}

template double abs<double>(double);
template float abs<float>(float);
template double Terror<double>(double, double);
template float Terror<float>(float, float);
template long int vectorDifference<double>(double*, double*, long long, double);
template long int vectorDifference<float>(float*, float*, long long, double);
template short testEquality<double>(double*, double*, long long);
template short testEquality<float>(float*, float*, long long);

// ------------------------------------------------------------------ logging

const char* const kCannonCsvHeader =
    "Algo,M,N,K,TotalNodes,TotalGPUs,dRow,dCol,StackSize,DecompositionTime,ExecutionTime,GatherTime,GFlops,"
    "DataLocation,CommunicationTime,GemmCallTime";

// Opens `filename` for appending and writes `header` unless the file's first
// line already is it.  The two console messages are the reference's
// (src/Utilities/Logging.cpp), kept as drop-in output.
FILE* createLogCsv(std::string filename, std::string header) {
    bool header_present = false, file_has_line = false;
    {
        std::ifstream in(filename);
        std::string first;
        if (in && std::getline(in, first)) {
            file_has_line = true;
            header_present = first == header && !in.eof();  // the line ended with a newline
        }
    }
    FILE* logfile = fopen(filename.c_str(), "a");
    if (logfile == NULL) dgHostError(EXIT_FAILURE, "cannot open log file %s", filename.c_str());
    if (header_present) {
        std::cout << "Header already placed, skip it" << std::endl;
        return logfile;
    }
    if (!file_has_line) std::cout << "Header not placed" << std::endl;
    fprintf(logfile, "%s\n", header.c_str());
    return logfile;
}

void writeLineToFile(FILE* logfile, char* dataLine) { fputs(dataLine, logfile); }

const char* getMachineName() {
    static char name[HOST_NAME_MAX + 1] = "";
    if (!name[0]) {
        if (gethostname(name, sizeof(name)) != 0) strcpy(name, "unknown");
        name[HOST_NAME_MAX] = '\0';
    }
    return name;
}
