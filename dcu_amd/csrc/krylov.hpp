// Device-side sparse formats, row tiles and the fused BiCGSTAB solver.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.hpp"

namespace spai {

// Per-system solver state kept in device memory; the fused kernels' last
// workgroup of each system updates it, the host only reads it.
struct alignas(64) SysState {
    double rho;      // (r*, r)
    double alpha;
    double omega;
    double beta;
    double rr;       // ||r||^2 (recursive residual)
    double tol;      // absolute threshold on ||r||_2
    double err0;     // ||r_0||_2
    double true_rr;  // ||b - A x||^2 after the solve
    int active;      // 1 while iterating
    int iters;       // the reference's iteration count once stopped
    int status;      // SolveStatus
    int max_iter;
    int restarts;    // restarts from the true residual (true_residual_restart)
};

enum SolveStatus { kRunning = 0, kConverged = 1, kMaxIter = 2, kBreakdown = 3, kConvergedAtStart = 4 };

// Row tiles: contiguous row ranges, none crossing a system boundary, each
// holding about `target_nnz` nonzeros (at most one LDS staging buffer plus one
// row).  Kernels without a reduction run one workgroup per tile; kernels that
// reduce run one workgroup per group of `group_size` consecutive tiles of one
// system, so that each workgroup publishes one partial for several tiles.
struct TileMap {
    int ntiles = 0;
    int target_nnz = 0;
    DeviceArray<int> row;        // ntiles + 1 row boundaries
    DeviceArray<int> sys;        // system of each tile
    DeviceArray<int> sys_tiles;  // nsys + 1 tile offsets per system
    std::vector<int> h_sys_tiles;
    int ngroups = 0;
    int group_size = 0;
    DeviceArray<int> group_tile;  // ngroups + 1: first tile of each group
    DeviceArray<int> group_sys;   // system of each group
    DeviceArray<int> sys_groups;  // nsys + 1 group offsets per system
    // buckets of consecutive groups of one system (two-level reduction count)
    int nbuckets = 0;
    DeviceArray<int> group_bucket;  // bucket of each group
    DeviceArray<int> bucket_group;  // nbuckets + 1: first group of each bucket
    DeviceArray<int> sys_buckets;   // nsys + 1: first bucket of each system
};

// Device CSC -> CSR (rows ordered by column) plus perm[csc_pos] = csr_pos.
// Deterministic: counts by atomics, rows sorted by column afterwards.
void csc_to_csr_device(int n, int nnz, const int *colptr, const int *rowidx, const double *cscval, int *rowptr,
                       int *colidx, double *csrval, int *perm, hipStream_t st);

// Builds tiles over rows of a CSR-like pointer array `ptr` (n+1 entries, on
// device) for systems with row offsets h_row_off (nsys+1, host) and nonzero
// offsets h_nnz_off (nsys+1, host).
void build_tiles(TileMap &tm, const int *ptr, const std::vector<int> &h_row_off,
                 const std::vector<long long> &h_nnz_off, hipStream_t st);

struct SolveOptions {
    double rel_tol = 1e-8;   // stop when ||r|| < rel_tol * ||b|| (per system)
    double abs_tol = 0.0;    // > 0: absolute threshold instead (the reference's TOL)
    int max_iter = 1000;     // reference semantics: at most max_iter - 1 iterations
    int use_graph = 0;       // capture iterations in a hipGraph
    int graph_iters = 8;     // iterations per graph launch (rounded up to even)
    int check_every = 1;     // iterations between host convergence checks (no graph)
    int record_history = 0;  // keep ||r_k|| per iteration and system
    // After the recursive residual converges, check ||b - A x||; a system
    // whose true residual is still >= TOL restarts from its current x with
    // r = b - A x (at most 4 times).  Off: the reference's stop rule only.
    int true_residual_restart = 0;
};

struct SolveResult {
    float ms = 0.0f;         // device time of the whole solve incl. true residual
    int iterations_max = 0;  // max over systems of the reported iteration count
    long long iterations_sum = 0;
    int host_checks = 0;     // host round trips for convergence
    int all_converged = 0;
};

// Matrix operand of the solver: CSR arrays (or CSC arrays for a transposed
// operator) with values in double (A) and double or float (M).
struct OperatorView {
    const int *ptr = nullptr;
    const int *idx = nullptr;
    const void *val = nullptr;
    int fp32 = 0;
};

struct Ctx;  // device-side view of the solver state (krylov.hip)

class BiCGStab {
public:
    BiCGStab();
    ~BiCGStab();
    // Solves every system of the batch: A x = b with right preconditioner M,
    // x holds x0 on entry.  Rows of system s are [row_off[s], row_off[s+1]).
    void solve(const OperatorView &A, const OperatorView &M, const TileMap &tiles, int n, int nsys, const double *b,
               double *x, const SolveOptions &opt, hipStream_t st, SolveResult *res);
    // Per-system results of the last solve (host copies).
    const std::vector<SysState> &states() const { return h_states_; }
    // ||r_k||_2 after iteration k = 1.. of system s in the last solve that
    // recorded a history; returns the number of values written.
    int history(int s, double *out, int maxlen) const;
    // Average device time of one A*vector tile launch over `reps` launches on
    // the solver's stream (roofline measurement of the dominant kernel).
    float time_kernel(int which, const OperatorView &A, const OperatorView &M, const TileMap &tiles, int n, int nsys,
                      int reps, hipStream_t st);

private:
    void ensure(int n, int nsys, int ntiles, int nbuckets);
    Ctx make_ctx(const TileMap &tm, int nsys);
    void enqueue_iteration(const OperatorView &A, const OperatorView &M, const TileMap &tiles, int parity,
                           hipStream_t st);
    int n_ = 0, nsys_ = 0, ntiles_ = 0, nbuckets_ = -1;
    DeviceArray<double> vec_;       // x-independent work vectors
    DeviceArray<double> partials_, bpartials_;
    DeviceArray<SysState> state_;
    DeviceArray<int> counter_;      // per system, + n_active at [nsys]
    std::vector<SysState> h_states_;
    double *r_ = nullptr, *rs_ = nullptr, *p_[2] = {nullptr, nullptr}, *y_ = nullptr, *a_ = nullptr, *s_ = nullptr,
           *z_ = nullptr, *t_ = nullptr;
    const double *b_ = nullptr;
    double *x_ = nullptr;
    // graph cache
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    const void *graph_key_[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int graph_iters_ = 0;
    void drop_graph();
    DeviceArray<double> hist_;
    double *hist_ptr_ = nullptr;
    int hist_stride_ = 0;
    std::vector<double> h_hist_;
};

// dL/dA on A's CSC pattern: dA[p] = -lambda[rowidx[p]] * x[col(p)].
void adjoint_scatter(int n, const int *colptr, const int *rowidx, const double *lambda, const double *x, double *dA,
                     hipStream_t st);

// Elementwise helpers used by the C-ABI.
void convert_f32_to_f64(const float *in, double *out, long long n, hipStream_t st);

}  // namespace spai
