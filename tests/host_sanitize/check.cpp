// CPU-only checks of the host API's utilities (validation, timing, CSV
// logging, op helpers), built with AddressSanitizer and UndefinedBehavior-
// Sanitizer by tests/test_host_sanitize.py.  No GPU call is made.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "generalUtilities.hpp"
#include "logging.hpp"
#include "validation.hpp"

static int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) { std::fprintf(stderr, "FAILED: %s (line %d)\n", #c, __LINE__); ++failures; } \
    } while (0)

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    // op helpers and GFLOP count
    CHECK(charToTransOp('n') == 'N' && charToTransOp('T') == 'T');
    CHECK(transOpToChar('N') == 'n' && transOpToChar('t') == 't');
    CHECK(std::fabs(calculateGflops(1000, 1000, 1000, 2.0) - 1.0) < 1e-12);
    // timer: monotone, in seconds
    const double t0 = csecond(), t1 = csecond();
    CHECK(t1 >= t0 && t0 > 0.0);
    // validation on data bounded away from zero: equal arrays give full digits,
    // a perturbation at 1e-10 relative gives about 10
    const long long n = 4097;
    std::vector<double> a(n), b(n);
    for (long long i = 0; i < n; ++i) a[i] = b[i] = 1.0 + (double)(i % 97) / 97.0;
    CHECK(Dtest_equality(a.data(), b.data(), n) >= 15);
    CHECK(Dvec_diff(a.data(), b.data(), n, 1e-300) == 0);
    for (long long i = 0; i < n; i += 3) b[i] *= 1.0 + 1e-10;
    const short d = Dtest_equality(a.data(), b.data(), n);
    CHECK(d >= 9 && d <= 11);
    CHECK(Dvec_diff(a.data(), b.data(), n, 1e-12) == (n + 2) / 3);
    const double rel = relativeFrobeniusError(a.data(), b.data(), n);
    CHECK(rel > 1e-11 && rel < 1e-10);
    std::vector<float> fa(n, 2.0f), fb(n, 2.0f);
    CHECK(testEquality<float>(fa.data(), fb.data(), n) >= 7);
    // CSV log: the header once, lines appended
    const std::string path = dir + "/sanitize_log.csv";
    std::remove(path.c_str());
    FILE* f = createLogCsv(path, "A,B,C");
    char line[] = "1,2,3\n";
    writeLineToFile(f, line);
    std::fclose(f);
    f = createLogCsv(path, "A,B,C");
    writeLineToFile(f, line);
    std::fclose(f);
    FILE* r = std::fopen(path.c_str(), "r");
    CHECK(r != nullptr);
    std::string text;
    char buf[256];
    while (r && std::fgets(buf, sizeof(buf), r)) text += buf;
    if (r) std::fclose(r);
    CHECK(text == "A,B,C\n1,2,3\n1,2,3\n");
    CHECK(std::strlen(getMachineName()) > 0);
    std::printf("host utilities: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
