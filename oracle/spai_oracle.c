/*
 * CPU oracle for the static-SPAI build and the SPAI-preconditioned BiCGSTAB
 * solve.  TEST INFRASTRUCTURE ONLY: nothing in the product library links,
 * loads or calls this file.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker (or as the
 * timed CPU baseline), never as the thing measured for the GPU.
 *
 * It is a from-scratch restatement, in plain C, of the reference algorithm:
 *
 *   SPAI (one least-squares problem per column k, pattern(M) = pattern(A)):
 *     J_k  = row indices of A(:,k) in storage order
 *            (src/spai_single_dcu.cpp:263-283, computeJ_SSPAIv10)
 *     I_k  = sorted ascending union of the row patterns of A(:,J_k)
 *            (src/spai_single_dcu.cpp:510-615, computeIShared_Symbol_SpMMv1)
 *     Ahat = dense A(I_k,J_k), row-major, first match in each column
 *            (src/spai_single_dcu.cpp:933-971, ComputeTildeACSR_SSPAIv10)
 *     modified Gram-Schmidt QR of Ahat in the reference's operation order
 *            (src/spai_single_dcu.cpp:1057-1121, QR_RShared_SSPAIv10)
 *     E_k  = position of k in I_k or -1
 *            (src/spai_single_dcu.cpp:973-992, ComputeTildeE_SSPAIv10)
 *     x    = R^-1 (row E_k of Q), x = 0 when E_k = -1
 *            (src/spai_single_dcu.cpp:1016-1055, Sol_SSPAIv10)
 *     M(J_k[j],k) = x[j] at A's storage position
 *            (src/spai_single_dcu.cpp:1123-1140, 2611-2655, modifyData4_SSPAIv10)
 *   A second QR mode (Householder, LAPACK dgeqr2 convention) is provided to
 *   cross-check the MGS restatement; the product's GPU kernels use
 *   Householder.
 *
 *   BiCGSTAB, right preconditioned, restating include/bicgstab/bicgstab_solver.h
 *   lines 106-232 (recurrences, iteration counting, absolute stop ||r|| < TOL),
 *   with x0 and TOL as parameters instead of the reference's fixed b = x0 = 1,
 *   TOL = 1e-7 (src/spai_multi_dcu.cpp:1920-1934).  b is never overwritten
 *   (the reference overwrites it with A*x at exit, bicgstab_solver.h:237-246).
 *   A switch (oracle_set_shadow_fp32) takes r* = fl32(r0), as the GPU's
 *   mixed-precision solve stores it.
 *
 *   CSC -> CSR with rows ordered by column (include/common/cuFormatConversion.h
 *   lines 453-585; include/common/assemble.h:34-75).
 *
 *   Adjoint of the linear solve (no reference counterpart): A^T lambda = g
 *   solved with BiCGSTAB preconditioned by M^T, dL/dA_ij = -lambda_i x_j on
 *   pattern(A), dL/db = lambda.
 *
 * Pinning (tests/test_oracle.py): numpy lstsq on dense A(:,J_k) per column,
 * closed forms (diagonal A, dense pattern = inverse, square tiles), MGS vs
 * Householder, scipy.sparse.linalg.bicgstab iterate-by-iterate, scipy CSR
 * transpose, and the reference's own host helpers compiled into oracle/_ref
 * (computeN1MAX / computeN2MAX from include/common/init.h, CSC2CSR from
 * include/common/assemble.h, readMatrixToCSC from include/common/read.h).
 * The reductions inside hipSPARSE/hipBLAS (the reference's third-party
 * arithmetic) are not restated bit-for-bit: parity unpinned at that level,
 * covered by the +-1 iteration tolerance.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_OK 0
#define ORACLE_ERR_ARG -1
#define ORACLE_ERR_NOMEM -3

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* I_k: sorted unique union of the rows of the columns listed in J.
 * Returns |I_k|; I must hold sum of the column counts. */
static int build_row_set(const int *colptr, const int *rowidx, const int *J, int n2, int *I) {
    int t = 0;
    for (int j = 0; j < n2; ++j) {
        int c = J[j];
        for (int p = colptr[c]; p < colptr[c + 1]; ++p) I[t++] = rowidx[p];
    }
    if (t == 0) return 0;
    qsort(I, (size_t)t, sizeof(int), cmp_int);
    int u = 1;
    for (int i = 1; i < t; ++i)
        if (I[i] != I[u - 1]) I[u++] = I[i];
    return u;
}

/* Modified Gram-Schmidt in the order of QR_RShared_SSPAIv10
 * (src/spai_single_dcu.cpp:1078-1117).  Q (n1 x n2, row-major, stride n2)
 * is overwritten, R (n2 x n2 row-major) receives the upper triangle. */
static void mgs_qr(double *Q, double *R, double *Rs, int n1, int n2) {
    for (int i = 0; i < n2; ++i) {
        for (int j = i; j < n2; ++j) {
            double t = 0.0;
            for (int k = 0; k < n1; ++k) t += Q[k * n2 + i] * Q[k * n2 + j];
            Rs[j - i] = t;
        }
        double rii = sqrt(Rs[0]);
        for (int k = 0; k < n1; ++k) Q[k * n2 + i] /= rii;
        for (int j = i; j < n2; ++j) {
            Rs[j - i] /= rii;
            R[i * n2 + j] = Rs[j - i];
        }
        for (int j = i + 1; j < n2; ++j)
            for (int k = 0; k < n1; ++k) Q[k * n2 + j] -= Rs[j - i] * Q[k * n2 + i];
    }
}

/* Sol_SSPAIv10 (src/spai_single_dcu.cpp:1016-1055): x = row E of Q, then
 * back substitution with R. */
static void mgs_solve(const double *Q, const double *R, int n2, int E, double *x) {
    for (int i = 0; i < n2; ++i) x[i] = (E < 0) ? 0.0 : Q[E * n2 + i];
    for (int i = n2 - 1; i >= 0; --i) {
        x[i] /= R[i * n2 + i];
        for (int j = 0; j < i; ++j) x[j] -= R[j * n2 + i] * x[i];
    }
}

/* Householder QR (dgeqr2 convention: H = I - tau v v^T, v[j] = 1,
 * beta = -sign(alpha) * norm) of the n1 x n2 row-major tile, then
 * x = R^-1 (Q^T e_E)[0:n2].  Used as the cross-check mode. */
static void householder_solve(double *A, int n1, int n2, int E, double *tau, double *y, double *x) {
    int kmax = n2 < n1 ? n2 : n1;
    for (int j = 0; j < kmax; ++j) {
        double alpha = A[j * n2 + j];
        double sigma = 0.0;
        for (int i = j + 1; i < n1; ++i) sigma += A[i * n2 + j] * A[i * n2 + j];
        if (sigma == 0.0) {
            tau[j] = 0.0;
            continue;
        }
        double nrm = sqrt(alpha * alpha + sigma);
        double beta = alpha >= 0.0 ? -nrm : nrm;
        tau[j] = (beta - alpha) / beta;
        double scal = 1.0 / (alpha - beta);
        for (int i = j + 1; i < n1; ++i) A[i * n2 + j] *= scal;
        A[j * n2 + j] = beta;
        for (int c = j + 1; c < n2; ++c) {
            double w = A[j * n2 + c];
            for (int i = j + 1; i < n1; ++i) w += A[i * n2 + j] * A[i * n2 + c];
            w *= tau[j];
            A[j * n2 + c] -= w;
            for (int i = j + 1; i < n1; ++i) A[i * n2 + c] -= w * A[i * n2 + j];
        }
    }
    for (int i = 0; i < n1; ++i) y[i] = (i == E) ? 1.0 : 0.0;
    for (int j = 0; j < kmax; ++j) {
        if (tau[j] == 0.0) continue;
        double w = y[j];
        for (int i = j + 1; i < n1; ++i) w += A[i * n2 + j] * y[i];
        w *= tau[j];
        y[j] -= w;
        for (int i = j + 1; i < n1; ++i) y[i] -= w * A[i * n2 + j];
    }
    for (int i = n2 - 1; i >= 0; --i) {
        double t = (i < n1) ? y[i] : 0.0;
        for (int c = i + 1; c < n2; ++c) t -= A[i * n2 + c] * x[c];
        x[i] = (E < 0) ? 0.0 : t / A[i * n2 + i];
    }
}

/* Per-column statistics used by the planner cross-checks: n2_k = |J_k|,
 * t_k = sum of column counts gathered, n1_k = |I_k|, e_k = position of k. */
int oracle_column_sets(int n, const int *colptr, const int *rowidx, int *n1, int *n2, int *t, int *e) {
    if (n < 0 || !colptr || !rowidx) return ORACLE_ERR_ARG;
    int fail = 0;
#pragma omp parallel
    {
        int cap = 0;
        int *I = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int k = 0; k < n; ++k) {
            int c0 = colptr[k], m = colptr[k + 1] - c0;
            int tot = 0;
            for (int j = 0; j < m; ++j) {
                int c = rowidx[c0 + j];
                tot += colptr[c + 1] - colptr[c];
            }
            if (tot > cap) {
                free(I);
                cap = tot * 2 + 16;
                I = (int *)malloc(sizeof(int) * (size_t)cap);
                if (!I) { fail = 1; cap = 0; continue; }
            }
            int u = (m > 0) ? build_row_set(colptr, rowidx, rowidx + c0, m, I) : 0;
            int pos = -1;
            for (int i = 0; i < u; ++i)
                if (I[i] == k) { pos = i; break; }
            if (n1) n1[k] = u;
            if (n2) n2[k] = m;
            if (t) t[k] = tot;
            if (e) e[k] = pos;
        }
        free(I);
    }
    return fail ? ORACLE_ERR_NOMEM : ORACLE_OK;
}

/* Static SPAI with pattern(M) = pattern(A).
 *   qr_mode 0: modified Gram-Schmidt in the reference's order.
 *   qr_mode 1: Householder (cross-check).
 * mval receives M's values in A's CSC storage order (M.mPtr = A.mPtr,
 * M.mIndex = A.mIndex, src/spai_single_dcu.cpp:2617-2624).  Columns outside
 * [col_begin, col_end) are left untouched. */
int oracle_spai_pattern(int n, const int *colptr, const int *rowidx, const double *val, const int *pcolptr,
                        const int *prowidx, int qr_mode, int col_begin, int col_end, double *mval);

int oracle_spai(int n, const int *colptr, const int *rowidx, const double *val, int qr_mode,
                int col_begin, int col_end, double *mval) {
    return oracle_spai_pattern(n, colptr, rowidx, val, colptr, rowidx, qr_mode, col_begin, col_end, mval);
}

/* As oracle_spai with M's pattern given separately (CSC pcolptr/prowidx):
 * J_k = prowidx[pcolptr[k] .. pcolptr[k+1]), the least squares still over
 * A's columns, mval in the pattern's order.  The reference's alternative
 * patterns (GetSparsityPattern, cuFormatConversion.h:598-974: A^T, E + A,
 * E + A + A^T) are built by the caller. */
int oracle_spai_pattern(int n, const int *colptr, const int *rowidx, const double *val, const int *pcolptr,
                        const int *prowidx, int qr_mode, int col_begin, int col_end, double *mval) {
    if (n < 0 || !colptr || !rowidx || !val || !pcolptr || !prowidx || !mval) return ORACLE_ERR_ARG;
    if (col_begin < 0) col_begin = 0;
    if (col_end > n) col_end = n;
    int fail = 0;
#pragma omp parallel
    {
        int capI = 0, capT = 0, capN = 0;
        int *I = NULL;
        double *tile = NULL, *R = NULL, *Rs = NULL, *x = NULL, *y = NULL, *tau = NULL;
#pragma omp for schedule(dynamic, 32)
        for (int k = col_begin; k < col_end; ++k) {
            int c0 = pcolptr[k], n2 = pcolptr[k + 1] - c0;
            if (n2 == 0) continue;
            const int *J = prowidx + c0;
            int tot = 0;
            for (int j = 0; j < n2; ++j) tot += colptr[J[j] + 1] - colptr[J[j]];
            if (tot + 1 > capI) {
                free(I);
                capI = 2 * tot + 16;
                I = (int *)malloc(sizeof(int) * (size_t)capI);
            }
            if (n2 > capN) {
                free(R); free(Rs); free(x); free(tau);
                capN = 2 * n2 + 8;
                R = (double *)malloc(sizeof(double) * (size_t)capN * (size_t)capN);
                Rs = (double *)malloc(sizeof(double) * (size_t)capN);
                x = (double *)malloc(sizeof(double) * (size_t)capN);
                tau = (double *)malloc(sizeof(double) * (size_t)capN);
            }
            if (!I || !R || !Rs || !x || !tau) { fail = 1; continue; }
            int n1 = build_row_set(colptr, rowidx, J, n2, I);
            long need = (long)n1 * n2;
            if (need > capT || n1 > capT) {
                free(tile); free(y);
                capT = (int)(need > n1 ? need : n1) * 2 + 16;
                tile = (double *)malloc(sizeof(double) * (size_t)capT);
                y = (double *)malloc(sizeof(double) * (size_t)capT);
                if (!tile || !y) { fail = 1; capT = 0; continue; }
            }
            /* Ahat(i,j) = A(I[i], J[j]): first match in column J[j]
             * (ComputeTildeACSR_SSPAIv10 searches linearly). */
            for (int i = 0; i < n1; ++i) {
                for (int j = 0; j < n2; ++j) {
                    int c = J[j];
                    double v = 0.0;
                    for (int p = colptr[c]; p < colptr[c + 1]; ++p)
                        if (rowidx[p] == I[i]) { v = val[p]; break; }
                    tile[i * n2 + j] = v;
                }
            }
            int E = -1;
            for (int i = 0; i < n1; ++i)
                if (I[i] == k) { E = i; break; }
            if (qr_mode == 0) {
                mgs_qr(tile, R, Rs, n1, n2);
                mgs_solve(tile, R, n2, E, x);
            } else {
                householder_solve(tile, n1, n2, E, tau, y, x);
            }
            for (int j = 0; j < n2; ++j) mval[c0 + j] = x[j];
        }
        free(I); free(tile); free(R); free(Rs); free(x); free(y); free(tau);
    }
    return fail ? ORACLE_ERR_NOMEM : ORACLE_OK;
}

/* CSC -> CSR, rows ordered by column index (stable counting sort over the
 * columns in order).  Also returns perm[csc_pos] = csr_pos when perm != 0. */
int oracle_csc2csr(int nrow, int ncol, const int *colptr, const int *rowidx, const double *val,
                   int *rowptr, int *colidx, double *csrval, int *perm) {
    if (nrow < 0 || ncol < 0 || !colptr || !rowidx || !rowptr) return ORACLE_ERR_ARG;
    int nnz = colptr[ncol];
    for (int i = 0; i <= nrow; ++i) rowptr[i] = 0;
    for (int p = 0; p < nnz; ++p) {
        if (rowidx[p] < 0 || rowidx[p] >= nrow) return ORACLE_ERR_ARG;
        rowptr[rowidx[p] + 1]++;
    }
    for (int i = 0; i < nrow; ++i) rowptr[i + 1] += rowptr[i];
    int *cur = (int *)malloc(sizeof(int) * (size_t)(nrow > 0 ? nrow : 1));
    if (!cur) return ORACLE_ERR_NOMEM;
    memcpy(cur, rowptr, sizeof(int) * (size_t)nrow);
    for (int c = 0; c < ncol; ++c) {
        for (int p = colptr[c]; p < colptr[c + 1]; ++p) {
            int q = cur[rowidx[p]]++;
            if (colidx) colidx[q] = c;
            if (csrval && val) csrval[q] = val[p];
            if (perm) perm[p] = q;
        }
    }
    free(cur);
    return ORACLE_OK;
}

/* y = A x for CSR A (row-parallel, each row summed in storage order). */
static void spmv(int n, const int *rp, const int *ci, const double *v, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) t += v[p] * x[ci[p]];
        y[i] = t;
    }
}

void oracle_spmv(int n, const int *rp, const int *ci, const double *v, const double *x, double *y) {
    spmv(n, rp, ci, v, x, y);
}

static double dot(int n, const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* rho' of the BiCGSTAB loop: 0 (default) (r*, r) after the r update, as
 * bicgstab_solver.h:196-199; 1 by linearity, (r*, s) - omega (r*, A z), the
 * rearrangement the GPU solver uses so that beta is known before its next
 * M-apply.  Mode 1 only serves to compare the two roundings' iteration-count
 * distributions in the tests. */
static int g_rho_mode = 0;
void oracle_set_rho_mode(int mode) { g_rho_mode = mode; }

/* Shadow residual of the mixed-precision solve: 0 (default) r* = r0 as the
 * reference (bicgstab_solver.h:117); 1 r* = r0 rounded to FP32, every inner
 * product with r* (rho and the alpha denominator) reading the rounded value.
 * This is how the GPU solver stores r* next to an FP32 M; every other vector
 * stays FP64. */
static int g_shadow32 = 0;
void oracle_set_shadow_fp32(int on) { g_shadow32 = on; }

/* Right-preconditioned BiCGSTAB, restating bicgstab_solver.h:106-232.
 *   x      : in = x0, out = solution
 *   tol    : absolute threshold on the recursive residual ||r||_2 (strict <)
 *   max_iter: as the reference (<= 0 means n); returns the reference's
 *             iter_times (iterations executed on convergence, max_iter when
 *             the loop runs out after max_iter - 1 iterations).
 *   info[0] = final recursive ||r||_2, info[1] = ||r0||_2,
 *   info[2] = 1 if the stop test fired, else 0.
 * M = NULL means no preconditioner (y = p). */
int oracle_bicgstab_history(int n, const int *arp, const int *aci, const double *av, const int *mrp,
                            const int *mci, const double *mv, const double *b, double *x, double tol, int max_iter,
                            double *info, double *hist);

int oracle_bicgstab(int n, const int *arp, const int *aci, const double *av, const int *mrp, const int *mci,
                    const double *mv, const double *b, double *x, double tol, int max_iter, double *info) {
    return oracle_bicgstab_history(n, arp, aci, av, mrp, mci, mv, b, x, tol, max_iter, info, NULL);
}

/* As oracle_bicgstab; hist (max_iter doubles, may be NULL) receives ||r_k||
 * after iteration k = 1, 2, ... */
int oracle_bicgstab_history(int n, const int *arp, const int *aci, const double *av, const int *mrp,
                            const int *mci, const double *mv, const double *b, double *x, double tol, int max_iter,
                            double *info, double *hist) {
    if (n <= 0 || !arp || !aci || !av || !b || !x) return ORACLE_ERR_ARG;
    size_t sz = sizeof(double) * (size_t)n;
    double *r = (double *)malloc(sz), *rs = (double *)malloc(sz), *p = (double *)malloc(sz);
    double *y = (double *)malloc(sz), *ay = (double *)malloc(sz), *s = (double *)malloc(sz);
    double *z = (double *)malloc(sz), *az = (double *)malloc(sz);
    if (!r || !rs || !p || !y || !ay || !s || !z || !az) {
        free(r); free(rs); free(p); free(y); free(ay); free(s); free(z); free(az);
        return ORACLE_ERR_NOMEM;
    }
    spmv(n, arp, aci, av, x, r);
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    memcpy(rs, r, sz);
    memcpy(p, r, sz);
    if (g_shadow32)
        for (int i = 0; i < n; ++i) rs[i] = (double)(float)rs[i];
    double err0 = sqrt(dot(n, r, r));
    double rho = dot(n, rs, r);
    double err = err0;
    int converged = 0;
    int it = 1;
    if (max_iter <= 0) max_iter = n;
    while (it < max_iter) {
        if (mrp) spmv(n, mrp, mci, mv, p, y); else memcpy(y, p, sz);
        double rho_old = rho;
        spmv(n, arp, aci, av, y, ay);
        double alpha = rho_old / dot(n, rs, ay);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) s[i] = r[i] + (-alpha) * ay[i];
        if (mrp) spmv(n, mrp, mci, mv, s, z); else memcpy(z, s, sz);
        spmv(n, arp, aci, av, z, az);
        double ts = dot(n, az, s), tt = dot(n, az, az);
        double w = ts / tt;
        double rs_s = 0.0, rs_t = 0.0;
        if (g_rho_mode == 1) {
            rs_s = dot(n, rs, s);
            rs_t = dot(n, rs, az);
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * y[i] + w * z[i];
            r[i] = s[i] + (-w) * az[i];
        }
        rho = g_rho_mode == 1 ? rs_s - w * rs_t : dot(n, rs, r);
        double beta = (rho / rho_old) * (alpha / w);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) p[i] = (beta * p[i] + r[i]) + (-beta * w) * ay[i]; /* scal, axpy, axpy */
        err = sqrt(dot(n, r, r));
        if (hist) hist[it - 1] = err;
        if (err < tol) { converged = 1; break; }
        it++;
    }
    if (info) { info[0] = err; info[1] = err0; info[2] = converged; }
    free(r); free(rs); free(p); free(y); free(ay); free(s); free(z); free(az);
    return it;
}

/* ||b - A x||_2 */
double oracle_true_residual(int n, const int *arp, const int *aci, const double *av, const double *b,
                            const double *x) {
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    if (!t) return -1.0;
    spmv(n, arp, aci, av, x, t);
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (b[i] - t[i]) * (b[i] - t[i]);
    free(t);
    return sqrt(s);
}

/* Gradient of L = 0.5 ||x - x_t||^2 with A x = b, on A's CSC pattern.
 * Inputs: CSC A (colptr,rowidx,val), CSC M (same pattern, values mval), the
 * converged forward solution x, target xt.  Solves A^T lambda = x - xt with
 * BiCGSTAB right-preconditioned by M^T: CSR(A^T) is A's CSC arrays and
 * CSR(M^T) is M's CSC arrays.  Outputs dA (CSC order, nnz) and db (n).
 * info as in oracle_bicgstab plus info[3] = iterations. */
int oracle_adjoint(int n, const int *colptr, const int *rowidx, const double *val, const double *mval,
                   const double *x, const double *xt, double tol, int max_iter, double *dA, double *db,
                   double *info) {
    if (n <= 0 || !colptr || !rowidx || !val || !x || !xt || !dA || !db) return ORACLE_ERR_ARG;
    double *g = (double *)malloc(sizeof(double) * (size_t)n);
    if (!g) return ORACLE_ERR_NOMEM;
    for (int i = 0; i < n; ++i) { g[i] = x[i] - xt[i]; db[i] = 0.0; }
    double inf[3] = {0, 0, 0};
    int it = oracle_bicgstab(n, colptr, rowidx, val, mval ? colptr : NULL, rowidx, mval, g, db, tol, max_iter,
                             inf);
    free(g);
    if (it < 0) return it;
    /* dL/dA_ij = -lambda_i x_j for each stored (i, j): CSC column j, row i. */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int p = colptr[j]; p < colptr[j + 1]; ++p) dA[p] = -db[rowidx[p]] * x[j];
    if (info) { info[0] = inf[0]; info[1] = inf[1]; info[2] = inf[2]; info[3] = it; }
    return it;
}

/* Threads used by the parallel loops (changes the reduction order of the
 * dot products, used to measure how sensitive iteration counts are to it). */
void oracle_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
