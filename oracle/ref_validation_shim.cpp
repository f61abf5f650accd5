// C entry points over the reference's own validation code, compiled from
// src/Utilities/Validation.cpp where it lies in the reference tree (see
// oracle/Makefile).  Test infrastructure only: it lets the oracle's
// restatement of Dtest_equality be checked against the original.
#include "validation.hpp"

extern "C" short ref_Dtest_equality(double* C_comp, double* C, long long size) {
    return Dtest_equality(C_comp, C, size);
}

extern "C" long ref_Dvec_diff(double* a, double* b, long long size, double eps) {
    return Dvec_diff(a, b, size, eps);
}
