/*
 * Drop-in C++ API with the reference's declarations
 * (include/common/spai_multi_dcu_api.h:1-24, C++ linkage), implemented on
 * one MI355X by libspai_mi355x.
 *
 * Semantics at one device (dcu_units = 1), the only configuration built:
 *   - initMultiDCU(requestedDCUs): any value selects the current HIP device
 *     (the reference uses all devices for requestedDCUs <= 0,
 *     src/spai_multi_dcu.cpp:1309-1313; here there is exactly one).
 *   - RunStaticSPAI_MultiDCU_MPI / RunDynamicSPAI_MultiDCU_MPI
 *     (src/spai_multi_dcu.cpp:1867-1875): CSC_A is the caller's host CSC
 *     (read only).  The functions build M with pattern(M) = pattern(A) for
 *     columns [globalColStart, globalColEnd) and FILL devCSC_M_global with
 *     device CSC arrays (hipMalloc'd on the context's device, mData in
 *     double; columns outside the range are zero), which the caller frees
 *     with hipFree.  The reference leaves M empty when its chunked branch is
 *     not taken (src/spai_multi_dcu.cpp:1670-1674, 1844) and only times the
 *     sizing phases (:1569, :1846-1849); here the full M is always returned
 *     and the returned float is the device time of the whole build in ms.
 *     Static uses balanced_columns scheduling, dynamic balanced_nnz.
 *   - Errors print a message and exit(EXIT_FAILURE), as the reference does
 *     (src/spai_multi_dcu.cpp:37-43); the C ABI in spai_mi355x.h returns
 *     status codes instead.
 *   - cublas2_pbicgstabv2 (include/bicgstab/bicgstab_solver.h:33): device
 *     CSR A and M, device b and x (x = x0 on entry), absolute TOL on
 *     ||r||_2, MAX_ITER <= 0 means n; returns the reference's iteration
 *     count.  Unlike the reference, b is not overwritten with A x.
 *   - cuCSC2CSR (include/common/cuFormatConversion.h:453): device arrays.
 */
#ifndef SPAI_MULTI_DCU_API_H
#define SPAI_MULTI_DCU_API_H

#include "spai_types.h"

struct MultiDCU_Context;

MultiDCU_Context *initMultiDCU(int requestedDCUs = -1);

void cleanupMultiDCU(MultiDCU_Context *ctx);

float RunStaticSPAI_MultiDCU_MPI(MultiDCU_Context *ctx, const CSC_Matrix *CSC_A, CSC_Matrix *devCSC_M_global,
                                 int globalColStart, int globalColEnd);

float RunDynamicSPAI_MultiDCU_MPI(MultiDCU_Context *ctx, const CSC_Matrix *CSC_A, CSC_Matrix *devCSC_M_global,
                                  int globalColStart, int globalColEnd);

int cublas2_pbicgstabv2(CSR_Matrix *devA, CSR_Matrix *devM, double *dev_b, double *dev_x, double TOL, int MAX_ITER);

void cuCSC2CSR(int m, int n, int nnz, double *cscData, int *cscIndex, int *cscPtr, double *csrData, int *csrIndex,
               int *csrPtr);

#endif
