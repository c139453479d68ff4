/*
 * Sparse-matrix containers shared by every entry point of libspai_mi355x.
 *
 * Field order, names and types are identical to the reference's containers
 * (include/common/dataType.h:24-42) so that code written against the
 * reference links against this library unchanged: 0-based indices, int
 * pointers and indices, double values.  `mPtr` has nCol+1 (CSC) or nRow+1
 * (CSR) entries; `n` is the order of a square matrix.
 */
#ifndef SPAI_TYPES_H
#define SPAI_TYPES_H

#ifndef MATRIX_DATATYPE_H
typedef struct {
    int *mIndex;
    double *mData;
    int *mPtr;
    int nRow;
    int nCol;
    int n;
    int nonzeroes;
} CSR_Matrix;

typedef struct {
    int *mIndex;
    double *mData;
    int *mPtr;
    int nRow;
    int nCol;
    int n;
    int nonzeroes;
} CSC_Matrix;
#endif

#endif
