/*
 * libspai_mi355x: static SPAI preconditioner construction and the
 * SPAI-preconditioned BiCGSTAB solve on one AMD Instinct MI355X (gfx950).
 *
 * C ABI: plain pointers and sizes, int status codes (0 = success, negative =
 * error, message via spai_last_error()), no exit() on error.  Every entry
 * point is synchronous with respect to the host: when it returns, its device
 * work is complete.  Calls on one spai_context must not overlap (a context
 * owns one HIP stream and its workspaces); distinct contexts are independent.
 *
 * Reference interfaces these entry points replace (file:line in the
 * reference tree):
 *   - spai_system_build            StaticSPAIv20 (src/spai_single_dcu.cpp:1224)
 *                                  and StaticSPAIv20_MultiDCU
 *                                  (src/spai_multi_dcu.cpp:1952) at one device
 *   - spai_system_solve            cublas2_pbicgstabv2
 *                                  (include/bicgstab/bicgstab_solver.h:33)
 *   - spai_csc_to_csr_device       cuCSC2CSR
 *                                  (include/common/cuFormatConversion.h:453)
 *   - spai_problem_read_mtx        readMatrixToCSC (include/common/read.h:1296)
 *   - the C++ drop-ins (initMultiDCU, RunStaticSPAI_MultiDCU_MPI, ...) are
 *     declared in spai_multi_dcu_api.h.
 */
#ifndef SPAI_MI355X_H
#define SPAI_MI355X_H

#include <stdint.h>

#include "spai_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status ---- */
#define SPAI_OK 0
#define SPAI_ERR_INVALID (-1)     /* bad argument */
#define SPAI_ERR_HIP (-2)         /* HIP runtime error */
#define SPAI_ERR_NOMEM (-3)       /* device or host allocation failed */
#define SPAI_ERR_UNSUPPORTED (-4) /* e.g. a column larger than every tile tier */
#define SPAI_ERR_BREAKDOWN (-5)   /* BiCGSTAB produced a non-finite scalar */
#define SPAI_ERR_NOT_BUILT (-6)   /* solve before spai_system_build */

/* precision of the SPAI build and of the stored M */
#define SPAI_PRECISION_FP64 0     /* FP64 Householder QR, FP64 M */
#define SPAI_PRECISION_MIXED 1    /* FP32 QR + one FP64 refinement sweep, FP32 M */

/* column scheduling (the reference's --partition values) */
#define SPAI_SCHEDULE_BALANCED_COLUMNS 0 /* static contiguous column blocks */
#define SPAI_SCHEDULE_BALANCED_NNZ 1     /* size-tiered atomic work queues */

/* per-system solver outcome */
#define SPAI_SOLVE_RUNNING 0
#define SPAI_SOLVE_CONVERGED 1
#define SPAI_SOLVE_MAX_ITER 2
#define SPAI_SOLVE_BREAKDOWN 3
#define SPAI_SOLVE_CONVERGED_AT_START 4

const char *spai_status_string(int status);
/* Message of the last failing call on this thread. */
const char *spai_last_error(void);
/* Library version string.  0.2: spai_build_report gained tile_shape[3];
 * 0.3: phase_ms[4] (the struct grew; rebuild callers against this header). */
const char *spai_version(void);

/* ------------------------------------------------------------ context --- */
typedef struct spai_context spai_context;

/* Creates a context on HIP device `device` with its own stream. */
int spai_context_create(int device, spai_context **out);
int spai_context_destroy(spai_context *ctx);
/* Name of the device (e.g. "AMD Instinct MI355X") and its CU count. */
int spai_context_device_info(spai_context *ctx, char *name, int name_len, int *cus);

/* ------------------------------------------------------------- system --- */
/* One square sparse matrix, or a batch of independent square matrices
 * stored block-diagonally, resident on the context's device with everything
 * the build and the solves need: CSC and CSR of A, the CSC->CSR permutation,
 * row tiles, and (after spai_system_build) M in CSC and CSR order. */
typedef struct spai_system spai_system;

/* From a host CSC (0-based, square; rows inside a column in any order). */
int spai_system_create(spai_context *ctx, const CSC_Matrix *host_A, spai_system **out);
/* From `nsys` host CSC matrices; system s occupies rows/columns
 * [row_offset[s], row_offset[s+1]) of the block-diagonal batch. */
int spai_system_create_batched(spai_context *ctx, int nsys, const CSC_Matrix *host_systems, spai_system **out);
/* From a device CSC (pointers on the context's device); copied. */
int spai_system_create_device(spai_context *ctx, const CSC_Matrix *dev_A, spai_system **out);
int spai_system_destroy(spai_system *sys);

typedef struct {
    int n;                   /* total order */
    int nnz;
    int nsys;
    int m_precision;         /* -1 before a build */
    float setup_ms;          /* upload + CSC->CSR + tiles (device time) */
} spai_system_info;
int spai_system_get_info(const spai_system *sys, spai_system_info *info);
/* Row offsets of the systems (nsys + 1 ints). */
int spai_system_row_offsets(const spai_system *sys, int *row_offsets);

/* Replaces A's values (same pattern, CSC order, host or device pointer). */
int spai_system_set_values(spai_system *sys, const double *values, int values_on_device);

/* ------------------------------------------------------------- build ---- */
typedef struct {
    float ms;                /* device time: planning + column kernels + M in CSC/CSR */
    int tier_columns[5];     /* columns per size tier (three register buckets, LDS wave, global workgroup) */
    int max_gathered;        /* max_k sum_{j in J_k} |A(:,j)| */
    int max_n2;              /* max_k |J_k| (the reference's n2max) */
    int empty_columns;
    long long sum_gathered;  /* entries of A gathered over all columns */
    int static_tier;         /* tier used by balanced_columns, -1 for balanced_nnz */
    int tile_shape[3];       /* candidate group shape of each register tier (see spai_system_autotune) */
    float phase_ms[4];       /* device ms of the phases: planner, register tiers, LDS tier, global tier
                                (the reference times its SPAI phases with events,
                                src/spai_single_dcu.cpp:2153-2190) */
} spai_build_report;

/* Sparsity pattern of M.  SPAI_PATTERN_A (the default) is pattern(A), as
 * the reference's SPAI uses; the others are the reference's
 * GetSparsityPattern flags (include/common/cuFormatConversion.h:598), which
 * its drivers never call: 0 = A^T, 1 = E + A, 2 = E + A + A^T (E the
 * identity).  E + A and E + A + A^T give every column a row set that holds
 * its own index whenever A's row k is non-empty, so M has no zero column
 * where A lacks a diagonal entry. */
#define SPAI_PATTERN_A (-1)
#define SPAI_PATTERN_AT 0
#define SPAI_PATTERN_I_PLUS_A 1
#define SPAI_PATTERN_I_PLUS_A_PLUS_AT 2
/* Sets M's pattern for the next builds (M must be built again); no
 * counterpart in the reference's API. */
int spai_system_set_pattern(spai_system *sys, int pattern);
/* M's pattern as CSC: *nnz entries, colptr (n + 1) and rowidx (nnz) host
 * arrays (any of the three may be NULL). */
int spai_system_m_pattern(const spai_system *sys, int *nnz, int *colptr, int *rowidx);

/* Builds M on its pattern (pattern(A) unless spai_system_set_pattern chose
 * another) for columns [col_begin, col_end) (col_end < 0 means n); other
 * columns of M are zero. */
int spai_system_build(spai_system *sys, int precision, int schedule, int col_begin, int col_end,
                      spai_build_report *report);

/* ------------------------------------------- tile-shape autotuner ---- */
/* The three register tiers are size buckets of the columns' least squares
 * (largest T_k = sum_{j in J_k} |A(:,j)| of 32, 64, 128 and largest |J_k| of
 * 8, 8, 16).  Each has candidate group shapes: lanes per column, gathered
 * entries per lane, largest |J_k| and tile rows per lane (|I_k| above
 * lanes x rows hands the column to the next bucket).  Candidate 0 is the
 * default.  The shape orders the group sums differently, so M moves in its
 * last bits between shapes.  No counterpart in the reference (whose kernel
 * shapes are fixed, src/spai_single_dcu.cpp:1245-1252). */
#define SPAI_TUNE_TIERS 3
#define SPAI_TUNE_MAX_CANDIDATES 8
typedef struct {
    int lanes;
    int entries_per_lane;
    int max_cols;
    int rows_per_lane;
} spai_tile_shape;
/* Number of candidates of register tier `tier` (0..2); 0 for other tiers. */
int spai_tile_shape_count(int tier);
int spai_tile_shape_get(int tier, int candidate, spai_tile_shape *out);
typedef struct {
    int shape[SPAI_TUNE_TIERS];      /* chosen candidate per tier */
    int columns[SPAI_TUNE_TIERS];    /* columns solved in each tier under the default shapes */
    float ms[SPAI_TUNE_TIERS][SPAI_TUNE_MAX_CANDIDATES]; /* tier kernels' device ms, best of reps; -1: not timed */
    float tune_ms;                   /* host wall time of the whole search */
} spai_tuning;
/* Times every candidate of each populated tier on this system's own
 * columns (a coordinate search from the smallest tier; best of `reps`
 * builds after one untimed one) and keeps the fastest for the system's later
 * builds.  M is left built with the chosen shapes.  `out` may be NULL. */
int spai_system_autotune(spai_system *sys, int precision, int schedule, int reps, spai_tuning *out);
/* Pins the shapes (one candidate index per tier) for the next builds, e.g.
 * a table an earlier autotune returned, so that builds are reproducible. */
int spai_system_set_tile_shapes(spai_system *sys, const int *shape);
int spai_system_get_tile_shapes(const spai_system *sys, int *shape);

/* M's values in the CSC storage order of its pattern (A's, by default),
 * widened to double. */
int spai_system_get_m(spai_system *sys, double *values, int values_on_device);
/* SPAI quality: ||A M - I||_F of the built M (as stored: FP32 in mixed
 * mode), and optionally ||A m_k - e_k||_2 per column (n doubles, may be
 * NULL).  The SPAI least squares minimise each column's term.  Computed
 * from the product below (a column of A M is A(:, J_k) m_k on I_k).  No
 * counterpart in the reference's drivers. */
int spai_system_quality(spai_system *sys, double *frobenius, double *column_residuals);

/* C = A M in CSC (the reference's SpMMWithoutSparsity,
 * include/common/SpMM_lib.h:15, which its drivers never call): column k is
 * A(:, J_k) m_k on the row set I_k, rows ascending, duplicates summed in
 * storage order.  Call with colptr = rowidx = values = NULL to get *nnz;
 * then pass host arrays of n + 1, nnz and nnz entries. */
int spai_system_product(spai_system *sys, long long *nnz, int *colptr, int *rowidx, double *values);

/* |I_k| per column from the last build (n ints), for diagnostics. */
int spai_system_get_row_set_sizes(spai_system *sys, int *n1);

/* ------------------------------------------------------------- solve ---- */
typedef struct {
    double rel_tol;          /* stop when ||r||_2 < rel_tol * ||b||_2 (per system) */
    double abs_tol;          /* > 0: stop when ||r||_2 < abs_tol instead (the reference's TOL) */
    int max_iter;            /* reference semantics: max_iter - 1 iterations at most */
    int use_graph;           /* 1: iterations replayed from a hipGraph */
    int graph_iters;         /* iterations per graph launch */
    int check_every;         /* iterations between host convergence checks without a graph */
    int precondition;        /* 0: no preconditioner (M = I) */
    int record_history;      /* 1: keep ||r_k||_2 per iteration (spai_system_residual_history) */
    int true_residual_restart; /* 1: after the recursive residual converges, restart (at most 4
                                  times) any system whose ||b - A x||_2 is still >= TOL; 0: the
                                  reference's stop rule alone */
} spai_solve_options;

void spai_solve_options_default(spai_solve_options *opt);

typedef struct {
    float ms;                /* device time of the solve incl. the true residual */
    int iterations_max;
    long long iterations_sum;
    int host_checks;
    int all_converged;
} spai_solve_report;

typedef struct {
    int status;              /* SPAI_SOLVE_* */
    int iterations;          /* the reference's iteration count */
    double residual;         /* final recursive ||r||_2 */
    double residual0;        /* ||r_0||_2 */
    double true_residual;    /* ||b - A x||_2 */
    double tol;              /* absolute threshold used */
    int restarts;            /* true-residual restarts taken */
} spai_system_outcome;

/* Solves A x = b for every system of the batch.  x holds x0 on entry.
 * `on_device` = 1: b and x are device pointers (n doubles each). */
int spai_system_solve(spai_system *sys, const double *b, double *x, int on_device, const spai_solve_options *opt,
                      spai_solve_report *report);
/* Allocation-free variant of spai_system_solve.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc/hipFree or other allocating routine:
 * the solver's work vectors, per-system states, counters and reduction
 * partials live in a static 64 MiB device array, one per device, looked up
 * when the device's first context is created and held until the process
 * ends.  b and x are device pointers (x holds x0 on entry).  Limit: 72 bytes
 * per row, 132 per system and 36 per reduction tree node (about 1 node per
 * 450 rows), plus under 1 KiB of alignment, must fit in 64 MiB: about
 * 930,000 rows for one system; larger systems return SPAI_ERR_UNSUPPORTED.  opt->use_graph and opt->record_history must be 0
 * (SPAI_ERR_INVALID otherwise); x updates are never deferred, so results
 * equal spai_system_solve's bit for bit.  The report's ms is host time.
 * Every call on one device shares that device's array: calls that overlap
 * on one device (from any context or thread) race. */
int spai_system_solve_noalloc(spai_system *sys, const double *b, double *x, const spai_solve_options *opt,
                              spai_solve_report *report);
/* Per-system outcome of the last solve (nsys entries). */
int spai_system_outcomes(spai_system *sys, spai_system_outcome *out);
/* ||r_k||_2 after iterations k = 1.. of system s in the last solve or adjoint
 * run with record_history = 1; returns the number of values written (>= 0)
 * or a negative status. */
int spai_system_residual_history(spai_system *sys, int s, double *out, int maxlen);

/* Adjoint of the solve for L(x): given the converged x and g = dL/dx, solves
 * A^T lambda = g with BiCGSTAB right-preconditioned by M^T (CSC(M) used as
 * CSR(M^T)), then dL/db = lambda and dL/dA_ij = -lambda_i x_j on A's CSC
 * pattern (nnz doubles in CSC order). */
int spai_system_adjoint(spai_system *sys, const double *x, const double *g, double *dA, double *db, int on_device,
                        const spai_solve_options *opt, spai_solve_report *report);

/* Average device time (ms) of one launch of a solver kernel repeated `reps`
 * times on the context's stream: which = 0 bare y = A x, 5 bare y = M x,
 * 1..4 the iteration kernels K1..K4 (see spai_system_time_iteration) on the
 * vectors and scalars of the last solve. */
int spai_system_time_spmv(spai_system *sys, int which, int reps, float *ms);

/* Average device time (ms) of each iteration kernel over `reps` consecutive
 * iterations from x0 = 0, b = 1, every system active (no threshold is
 * reached), with HIP events between the kernels on the context's stream: the
 * per-kernel times of the real iteration sequence.  K1 (x/r update and
 * y = M p') runs fused or split in two (a streaming update, then y = M p');
 * variant 0 times the fused form, 1 the split form, -1 the form the solver
 * uses when every unit is active.  ms[5]: K1 or its streaming half, the
 * M-apply half (0 when fused), K2 (a = A y, alpha), K3 (z = M s), K4
 * (t = A z, omega, beta).  Overwrites the solver's work vectors; M must be
 * built. */
int spai_system_time_iteration(spai_system *sys, int reps, int variant, float *ms);

/* Row layout chosen for the solver's operators: SPAI_LAYOUT_SELL64 (slices
 * of 64 rows, column-major, when padding adds at most 20%; 32-bit column
 * indices), SPAI_LAYOUT_SELL64_OFF16 (the same with 16-bit column - row
 * offsets, when every entry of A's pattern has |column - row| < 2^15),
 * SPAI_LAYOUT_SELL64_DIAG (one column - row offset per column block of a
 * slice, shared by its 64 rows, when a slice's distinct offsets add at most
 * 10% to the row-length padding: stencil matrices) or SPAI_LAYOUT_ROW_TILES;
 * SPAI_LAYOUT_ROW_TILES_BLOCKED is row tiles whose products split by column
 * blocks (one system of at least 2^19 rows: each block's slice of the
 * vector stays in the L2 of the XCDs it runs on; SPAI_MI355X_COLUMN_BLOCKS=0
 * at creation turns them off).  `entries` = stored entries per operator (nnz
 * plus padding, or nnz). */
#define SPAI_LAYOUT_ROW_TILES 0
#define SPAI_LAYOUT_SELL64 1
#define SPAI_LAYOUT_SELL64_OFF16 2
#define SPAI_LAYOUT_SELL64_DIAG 3
#define SPAI_LAYOUT_ROW_TILES_BLOCKED 4
int spai_system_layout(const spai_system *sys, int *kind, long long *entries);

/* Average device time (ms) of a plain copy kernel that reads and writes
 * `bytes` bytes each (the calibration kernel for hardware byte counters and
 * the practical streaming bandwidth). */
int spai_calibration_copy(spai_context *ctx, long long bytes, int reps, float *ms);

/* ------------------------------------------------- format conversion ---- */
/* Device CSC -> CSR with rows ordered by column (drop-in for cuCSC2CSR);
 * all pointers on the current device, caller-allocated. */
int spai_csc_to_csr_device(int m, int n, int nnz, const double *cscData, const int *cscIndex, const int *cscPtr,
                           double *csrData, int *csrIndex, int *csrPtr);

/* -------------------------------------------------- host problems ------- */
typedef struct spai_problem spai_problem;

/* Seeded synthetic problem: config 1..5 (random 4096, conv-diff 512^2,
 * power-law 1e6, batch of 64, 64^3 GRF diffusion). */
int spai_problem_generate(int config, uint64_t seed, spai_problem **out);
/* Power-law generator at a chosen size (tests). */
int spai_problem_generate_power_law(int n, uint64_t seed, spai_problem **out);
/* Batch generator with a chosen system count (tests). */
int spai_problem_generate_batched(int nsys, uint64_t seed, spai_problem **out);
/* Lognormal-GRF 3D diffusion (the 64^3 generator) on an m^3 grid (tests). */
int spai_problem_generate_grf(int m, uint64_t seed, spai_problem **out);
/* 2D conv-diff on an m x m grid / 3D diffusion with given cell kappa. */
int spai_problem_convection_diffusion(int m, double peclet, double vx, double vy, spai_problem **out);
/* MatrixMarket file with the reference's reader semantics; b = ones. */
int spai_problem_read_mtx(const char *path, spai_problem **out);
int spai_problem_free(spai_problem *p);
int spai_problem_num_systems(const spai_problem *p);
/* Host CSC of system s; pointers stay owned by the problem. */
int spai_problem_system(const spai_problem *p, int s, CSC_Matrix *out);
/* Right-hand side / target of system s (NULL when absent). */
const double *spai_problem_rhs(const spai_problem *p, int s);
const double *spai_problem_target(const spai_problem *p, int s);
const char *spai_problem_name(const spai_problem *p);
const char *spai_problem_notes(const spai_problem *p);

#ifdef __cplusplus
}
#endif

#endif
