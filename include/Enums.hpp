// Enumerations of the C++ host API.  The reference's MatrixUtilities.hpp
// includes <Enums.hpp> for MatrixLayout, which its tree does not contain;
// this header supplies the type with the two values that code uses.
#ifndef DG_ENUMS_HPP
#define DG_ENUMS_HPP

enum class MatrixLayout { RowMajor, ColumnMajor };

#endif
