// Templated matrix I/O and generation of the C++ host API: drop-in for the
// reference's include/MatrixUtilities.hpp (src/Utilities/MatrixUtilities.cpp).
//
// printMatrix / writeMatrix print the host array line by line; for a
// column-major array (the default layout) each printed line is one column,
// as in the reference.  The generation templates are those of cmatrix.h
// (seeded SplitMix64, bit-identical on CPU and GPU).
#ifndef MATRIX_UTILITIES_HPP
#define MATRIX_UTILITIES_HPP

#include <string>

#include "Enums.hpp"
#include "cmatrix.h"

/* Matrix I/O (MatrixUtilities.hpp:27-31; host arrays) */
template <typename scalar_t>
void printMatrix(scalar_t* array, long long rows, long long columns, int rank = 0,
                 MatrixLayout layout = MatrixLayout::ColumnMajor);

// Appends to `filename`: one line per row (RowMajor) or per column
// (ColumnMajor), values separated by spaces, then an empty line.
template <typename scalar_t>
void writeMatrix(scalar_t* array, long long rows, long long columns, std::string& filename,
                 MatrixLayout layout = MatrixLayout::ColumnMajor);

#endif
