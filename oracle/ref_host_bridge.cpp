// Exposes the reference's host-only helpers with C linkage so the tests can
// check the oracle and the MatrixMarket reader against them.  Compiled by
// oracle/Makefile (target `ref`) against the reference headers where they lie
// (-I<reference>/include/common); the output goes to oracle/_ref/ and is
// never shipped with the library.
//
//   include/common/init.h:43-94      computeN1MAX, computeN2MAX
//   include/common/assemble.h:34-75  CSC2CSR
//   include/common/read.h:1296-1321  readMatrixToCSC
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace std;

#include "dataType.h"
#include "init.h"
#include "assemble.h"
#include "read.h"

extern "C" {

int ref_compute_n1max(int n, int *colptr, int *rowidx, double *val) {
    CSC_Matrix A;
    A.n = A.nRow = A.nCol = n;
    A.mPtr = colptr;
    A.mIndex = rowidx;
    A.mData = val;
    A.nonzeroes = colptr[n];
    return computeN1MAX(&A, &A);  // pattern(M) = pattern(A)
}

int ref_compute_n2max(int n, int *colptr) {
    CSC_Matrix A;
    A.n = A.nRow = A.nCol = n;
    A.mPtr = colptr;
    A.mIndex = nullptr;
    A.mData = nullptr;
    A.nonzeroes = colptr[n];
    return computeN2MAX(&A);
}

// CSR of a CSC matrix into caller buffers (rowptr n+1, colidx/val nnz).
int ref_csc2csr(int n, int *colptr, int *rowidx, double *val, int *rowptr, int *colidx, double *csrval) {
    CSC_Matrix A;
    A.n = A.nRow = A.nCol = n;
    A.mPtr = colptr;
    A.mIndex = rowidx;
    A.mData = val;
    A.nonzeroes = colptr[n];
    CSR_Matrix R;
    CSC2CSR(&A, &R, (unsigned)n);
    std::memcpy(rowptr, R.mPtr, sizeof(int) * (size_t)(n + 1));
    std::memcpy(colidx, R.mIndex, sizeof(int) * (size_t)R.nonzeroes);
    std::memcpy(csrval, R.mData, sizeof(double) * (size_t)R.nonzeroes);
    free(R.mPtr);
    free(R.mIndex);
    free(R.mData);
    return 0;
}

// Reads a MatrixMarket file; returns a handle whose arrays are fetched with
// ref_mtx_get and released with ref_mtx_free.
void *ref_read_mtx(const char *path, int *n, int *nnz) {
    CSC_Matrix *A = (CSC_Matrix *)calloc(1, sizeof(CSC_Matrix));
    std::vector<char> p(path, path + strlen(path) + 1);
    readMatrixToCSC(p.data(), A);
    *n = A->nCol;
    *nnz = A->nonzeroes;
    return A;
}

void ref_mtx_get(void *h, int *colptr, int *rowidx, double *val) {
    CSC_Matrix *A = (CSC_Matrix *)h;
    std::memcpy(colptr, A->mPtr, sizeof(int) * (size_t)(A->nCol + 1));
    std::memcpy(rowidx, A->mIndex, sizeof(int) * (size_t)A->nonzeroes);
    std::memcpy(val, A->mData, sizeof(double) * (size_t)A->nonzeroes);
}

void ref_mtx_free(void *h) {
    CSC_Matrix *A = (CSC_Matrix *)h;
    free(A->mPtr);
    free(A->mIndex);
    free(A->mData);
    free(A);
}
}
