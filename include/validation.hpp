// Validation of the C++ host API (drop-in for the reference's
// include/validation.hpp and src/Utilities/Validation.cpp).
#ifndef VALIDATION_HPP
#define VALIDATION_HPP

#include <cfloat>
#include <cstdio>

template <typename T>
T abs(T x);

template <typename T>
T Terror(T a, T b);

template <typename T>
long int vectorDifference(T* a, T* b, long long size, double eps);

// Digits of agreement of `array` with `referenceArray` (templated form of
// Dtest_equality; the reference declares but never instantiates it).
template <typename T>
short testEquality(T* referenceArray, T* array, long long size);

// Counts entries whose relative error |a-b|/a exceeds eps, starting at 1e-8
// and dividing eps by 10 while nothing fails (up to 30 digits); returns the
// first digit count with failures, 0 if entries fail already at 1e-8.  The
// error divides by the signed reference value, so negative references never
// fail; a zero reference uses the absolute error (Validation.cpp:57-104).
short Dtest_equality(double* C_comp, double* C, long long size);
long int Dvec_diff(double* a, double* b, long long size, double eps);
double Derror(double a, double b);
double dabs(double x);

// ||C - C_ref||_F / ||C_ref||_F over `size` entries (0 if both are zero):
// the acceptance check the reference's digit count cannot give near zeros.
double relativeFrobeniusError(const double* C_ref, const double* C, long long size);

#endif
