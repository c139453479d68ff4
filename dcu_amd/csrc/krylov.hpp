// Device-side sparse formats, row tiles and the fused BiCGSTAB solver.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.hpp"

namespace spai {

// Per-system solver state kept in device memory; one workgroup of each
// system (its first unit's) updates it, the host only reads it.  No kernel
// reads a field that the same kernel writes.
struct alignas(64) SysState {
    double rho;      // (r*, r)
    double alpha;
    double omega;
    double beta;
    double rr;       // ||r||^2 (recursive residual)
    double tol;      // absolute threshold on ||r||_2
    double err0;     // ||r_0||_2
    double true_rr;  // ||b - A x||^2 after the solve
    double alpha_prev, omega_prev;  // alpha, omega of the iteration whose x update is deferred
    double rho_next;  // rho' formed by the last K1, the rho of the next K3 (K2 moves it to rho)
    int active;      // 1 while iterating
    int iters;       // the reference's iteration count once stopped
    int status;      // SolveStatus
    int max_iter;
    int restarts;    // restarts from the true residual (true_residual_restart)
    int xdefer;      // 1: stopped with x missing the update held in the odd buffers
    int iters_base;  // iterations before the last (re)start
    int k1_count;    // K1 launches since the last (re)start
    int k2_count;    // K2 launches since the last (re)start
    int k1_defer;    // the last K1 left its x update for the next one
};
static_assert(sizeof(SysState) == 128, "two 64-byte lines per system");

enum SolveStatus { kRunning = 0, kConverged = 1, kMaxIter = 2, kBreakdown = 3, kConvergedAtStart = 4 };

// Row layout of the solver's operators.  Every kernel runs one workgroup per
// unit; a unit's rows are contiguous and belong to one system.
//
//  * SELL-64 (kLayoutSell): rows in slices of 64 (one wavefront, lane = row),
//    each slice stored column-major with the width of its longest row; short
//    rows are padded with (own row, 0.0).  A unit is up to kSlicesPerUnit
//    slices of one system.  Chosen when the padding adds at most 20% to the nonzeros (stencil
//    and other near-uniform rows).  A and M share the column indices (M has
//    A's pattern); values are stored per operator in SELL order.
//  * Row tiles (kLayoutTiles): contiguous row ranges holding about
//    `target_nnz` nonzeros (at most one LDS staging buffer, a longer row gets
//    a tile of its own), over a CSR (or, for a transposed operator, CSC)
//    pointer array.  Used for irregular rows and for operators given as CSR.
//
// The iteration kernels publish one partial sum per group of units of a
// system (a unit alone when the system has at most kMaxGroups units), and the
// next kernel that needs the total sums the system's group partials itself.
// The set-up reductions combine per-unit partials up a tree per system (each
// node counts at most kFanIn arrivals).
// `active` lists the units of systems still iterating (rebuilt on the device
// from the system states); launches that use it exit at once past its count.
enum LayoutKind { kLayoutTiles = 0, kLayoutSell = 1 };
#ifndef SPAI_MAX_GROUPS
#define SPAI_MAX_GROUPS 512
#endif
constexpr int kMaxGroups = SPAI_MAX_GROUPS;  // partial-sum groups per system
static_assert(kMaxGroups % 256 == 0 && kMaxGroups <= 1024, "kMaxGroups: a multiple of 256, at most 1024");
// units per group of a system with `units` units
inline __host__ __device__ int group_units(int units) { return units <= kMaxGroups ? 1 : (units + kMaxGroups - 1) / kMaxGroups; }
// A unit is 8 slices (512 rows), one per wavefront of its workgroup: with
// few units active (the end of a batched solve) eight wavefronts per unit hide
// more latency than four wavefronts of two slices each.
#ifndef SPAI_SELL_SLICES_PER_WAVE
#define SPAI_SELL_SLICES_PER_WAVE 1
#endif
constexpr int kSliceRows = 64;
constexpr int kSellSlicesPerWave = SPAI_SELL_SLICES_PER_WAVE;
#ifndef SPAI_SELL_WAVES
#define SPAI_SELL_WAVES 8
#endif
constexpr int kSellWaves = SPAI_SELL_WAVES;  // wavefronts per SELL workgroup
constexpr int kSlicesPerUnit = kSellWaves * kSellSlicesPerWave;

struct TileMap {
    // distinct for every build_layout call in the process: captured solver
    // graphs are keyed by it, so a graph is never replayed on the device
    // arrays of another (or a rebuilt) layout that happen to reuse addresses
    unsigned long long generation = 0;
    int kind = kLayoutTiles;
    int target_nnz = 0;
    int nunits = 0;
    DeviceArray<int> unit_row;   // nunits + 1 row boundaries
    DeviceArray<int> unit_sys;   // system of each unit
    DeviceArray<int> sys_units;  // nsys + 1 unit offsets per system
    DeviceArray<int> sys_row;    // nsys + 1 row offsets per system
    std::vector<int> h_sys_units;
    // SELL-64 only
    int nslices = 0;
    long long sell_len = 0;       // padded entries (a multiple of 64)
    DeviceArray<int> unit_slice;  // nunits + 1: first slice of each unit
    DeviceArray<int> slice_row;   // nslices + 1 row boundaries
    DeviceArray<int> slice_blk;   // nslices + 1 offsets in blocks of 64 entries
    DeviceArray<int> sell_col;    // sell_len column indices
    DeviceArray<short> sell_off;  // sell_len column - row offsets; empty when one does not fit 16 bits
    DeviceArray<int> sell_diag;   // diagonal slices: one column - row offset per block (sell_col, sell_off empty)
    DeviceArray<int> sell_pos;    // nnz: CSR position -> SELL position
    // reduction tree per system (level 0 = its units, then groups of up to
    // kFanIn nodes per level): nodes of system s start at sys_nodes[s]
    int nnodes = 0;
    DeviceArray<int> sys_nodes;  // nsys + 1
    // partial-sum groups of the iteration kernels: system s with U units has
    // ceil(U / G) groups of G = ceil(U / kMaxGroups) consecutive units,
    // starting at sys_groups[s]
    int ngroups = 0;
    DeviceArray<int> sys_groups;  // nsys + 1
    // active-unit list: [0, nunits) unit ids in XCD-blocked order;
    // compaction scratch per system
    DeviceArray<int> active;
    DeviceArray<int> sys_base;
    // entries of an operator's value array in this layout
    long long value_len(long long nnz) const { return kind == kLayoutSell ? sell_len : nnz; }
    // Column blocks (row tiles of one system whose vectors exceed an XCD's
    // L2): the products of the split iteration are formed per block of
    // n / cblocks columns, on the XCDs the block is assigned to, as partial
    // row sums that the consuming kernel adds.  The entries are stored block
    // by block (row-major inside a block, CSR order inside a row).
    int cblocks = 0;                  // 0: none
    int cb_units_max = 0;             // most tiles of any block
    long long cb_nnz = 0;             // entries
    DeviceArray<int> cb_ptr;          // cblocks * n + 1: entries of (block c, row i) start at cb_ptr[c * n + i]
    DeviceArray<int> cb_idx;          // nnz column indices, block by block
    DeviceArray<int> cb_pos;          // nnz: CSR position -> block position
    DeviceArray<int> cb_tile;         // 4 per tile: first row, end row, first entry, end entry (block numbering)
    int cb_toff[9] = {};              // cblocks + 1: block c's tiles are [cb_toff[c], cb_toff[c + 1])
};

// Device CSC -> CSR of an m x n matrix (m rows: rowptr holds m + 1 entries;
// n columns: colptr holds n + 1), rows ordered by column, plus
// perm[csc_pos] = csr_pos.  Deterministic: counts by atomics, rows sorted by
// column afterwards.
void csc_to_csr_device(int m, int n, int nnz, const int *colptr, const int *rowidx, const double *cscval,
                       int *rowptr, int *colidx, double *csrval, int *perm, hipStream_t st);

// Builds the row layout over a CSR-like pointer/index pair (`ptr` n+1 and
// `idx` entries, on the device) for systems with row offsets h_row_off
// (nsys+1, host) and nonzero offsets h_nnz_off (nsys+1, host).  SELL-64 is
// considered only when allow_sell is set (the operators then share `idx`'s
// pattern); otherwise, or when its padding is too large, row tiles.  ptr2
// (optional) is a second operator's row pointers over the same rows (M with
// its own pattern): tiles are cut so that neither operator overflows one.
// allow_blocks: a row-tile layout of one system with at least
// kColBlockMinRows rows gets column blocks (unless SPAI_MI355X_COLUMN_BLOCKS=0).
void build_layout(TileMap &tm, const int *ptr, const int *idx, const std::vector<int> &h_row_off,
                  const std::vector<long long> &h_nnz_off, bool allow_sell, hipStream_t st,
                  const int *ptr2 = nullptr, bool allow_blocks = false);

// Values of a CSR-ordered array in the layout's order: out[sell_pos[p]] =
// in[p] (SELL, padding 0) or a copy (tiles).  (MT_OUT, MT_IN): (double,
// double), (float, double) or (float, float).
template <class MT_OUT, class MT_IN = double>
void layout_values(const TileMap &tm, long long nnz, const MT_IN *in, MT_OUT *out, hipStream_t st);

// perm_out[p] = position of CSC entry p in the layout's value order, from
// perm_csr[p] = its CSR position.
void layout_perm(const TileMap &tm, int nnz, const int *perm_csr, int *perm_out, hipStream_t st);

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

// Matrix operand of the solver in the layout's order: CSR arrays (or CSC
// arrays for a transposed operator) for row tiles, the shared SELL column
// indices (ptr unused) for SELL-64; values in double (A) and double or float
// (M).
struct OperatorView {
    const int *ptr = nullptr;
    const int *idx = nullptr;
    const void *val = nullptr;
    int fp32 = 0;
    const int *blk = nullptr;  // SELL-64: the operator's own slice offsets when its pattern is not A's
    const short *off = nullptr;  // SELL-64: 16-bit column - row offsets (used instead of idx when set)
    const int *diag = nullptr;   // SELL-64 diagonal slices: one offset per block (used instead of idx/off)
};

// SELL-64 arrays of an operator whose pattern differs from A's (an
// alternative SPAI pattern for M), on the layout's slices.
struct SellOperator {
    long long len = 0;            // padded entries
    DeviceArray<int> slice_blk;   // nslices + 1 offsets in blocks of 64 entries
    DeviceArray<int> col;         // len column indices
    DeviceArray<short> off;       // len column - row offsets; empty when one does not fit 16 bits
    DeviceArray<int> pos;         // nnz: CSR position -> SELL position
};

// Builds an operator's SELL-64 arrays on the layout's slices from its CSR
// pattern (rowptr n+1, colidx nnz, on the device).
void build_sell_operator(const TileMap &tm, const int *rowptr, const int *colidx, int nnz, SellOperator &op,
                         hipStream_t st);

// perm_out[p] = SELL position of the entry whose CSR position is perm_csr[p].
void compose_positions(int nnz, const int *perm_csr, const int *sell_pos, int *perm_out, hipStream_t st);

struct Ctx;  // device-side view of the solver state (krylov.hip)

class BiCGStab {
public:
    BiCGStab();
    ~BiCGStab();
    // Solves every system of the batch: A x = b with right preconditioner M,
    // x holds x0 on entry.  Rows of system s are [row_off[s], row_off[s+1]).
    // states_out (host, nsys entries, nullable) receives the per-system
    // states; a solver bound to a workspace requires it.
    void solve(const OperatorView &A, const OperatorView &M, const TileMap &tiles, int n, int nsys, const double *b,
               double *x, const SolveOptions &opt, hipStream_t st, SolveResult *res, SysState *states_out = nullptr);
    // Binds the solver to caller-owned device memory (256-byte aligned)
    // instead of allocations of its own: the work vectors, states, counters
    // and reduction partials of every later solve are carved from
    // [base, base + bytes), and a solve that does not fit is refused.  Such
    // a solver allocates nothing: it never captures graphs, records no
    // history, creates no events and never defers x updates (batches
    // included; the results are bit-identical to the deferring form).
    void bind_workspace(void *base, size_t bytes);
    // Device bytes a bound workspace needs for a solve of n rows and nsys
    // systems on the layout tm.
    static size_t workspace_bytes(int n, int nsys, const TileMap &tm);
    // Drops the captured graphs of the layout `generation` (a system that
    // is destroyed or laid out again).
    void forget_layout(unsigned long long generation);
    // Per-system results of the last solve (host copies).
    const std::vector<SysState> &states() const { return h_states_; }
    // ||r_k||_2 after iteration k = 1.. of system s in the last solve that
    // recorded a history; returns the number of values written.
    int history(int s, double *out, int maxlen) const;
    // Average device time of one launch of a single kernel repeated `reps`
    // times on the solver's stream (which: 0 bare A*x, 5 bare M*x, 1..4
    // K1..K4), on the solver's own vectors and states seeded as a solve from
    // x0 = 0, b = 1 starts, every system active.
    float time_kernel(int which, const OperatorView &A, const OperatorView &M, const TileMap &tiles, int n, int nsys,
                      int reps, hipStream_t st);
    // Average device time of each iteration kernel over `reps` consecutive
    // iterations from x0 = 0, b = 1 with every system active (no threshold is
    // reached), HIP events between the kernels on the solver's stream.
    // variant: 0 fused K1, 1 split K1, -1 the solver's choice for every unit
    // active.  ms: K1 (or its streaming half), its M-apply half (0 when
    // fused), K2, K3, K4.
    void time_iteration(const OperatorView &A, const OperatorView &M, const TileMap &tiles, int n, int nsys,
                        int reps, int variant, hipStream_t st, float ms[5]);

private:
    void ensure(int n, int nsys, const TileMap &tm);
    // b = 1 and x = 0 in tvec_, the states and work vectors of a solve from
    // them (OpInit); returns x
    double *seed_timing_state(const OperatorView &A, const TileMap &tm, int n, int nsys, int max_iter,
                              hipStream_t st);
    void compact(const TileMap &tm, int nsys, hipStream_t st, int *pub = nullptr);
    Ctx make_ctx(const TileMap &tm, int nsys);
    void enqueue_iteration(const OperatorView &A, const OperatorView &M, const TileMap &tiles, int parity, int grid,
                           bool split, hipStream_t st);
    void enqueue_k1(const OperatorView &A, const OperatorView &M, const TileMap &tiles, int parity, const int *list,
                    int grid, bool split, hipStream_t st);
    void enqueue_k2(const OperatorView &A, const TileMap &tiles, int parity, const int *list, int grid,
                    hipStream_t st);
    void enqueue_k3(const OperatorView &M, const TileMap &tiles, int parity, const int *list, int grid, bool split,
                    hipStream_t st);
    void enqueue_k4(const OperatorView &A, const TileMap &tiles, int parity, const int *list, int grid,
                    hipStream_t st);
    void enqueue_init(const OperatorView &A, const TileMap &tiles, const double *x, const double *b, double rel_tol,
                      double abs_tol, int restart, hipStream_t st);
    // y = M p' of the split K1
    void enqueue_mapply_p(const OperatorView &M, const TileMap &tiles, int parity, const int *list, int grid,
                          hipStream_t st);
    // Column-blocked products for this solve (a one-system row-tile layout
    // with column blocks, M on A's pattern, no bound workspace): the
    // operators' values in the blocks' order, refreshed here, before any
    // launch of the solve.
    void prepare_blocks(const OperatorView &A, const OperatorView &M, const TileMap &tiles, int n, int nsys,
                        hipStream_t st);
    bool cb_on_ = false;
    DeviceArray<double> cb_a_, cb_m64_, cb_part_;
    DeviceArray<float> cb_m32_;
    int n_ = 0, nsys_ = 0, nunits_ = -1, nnodes_ = -1, ngroups_ = -1;
    DeviceArray<double> vec_;       // x-independent work vectors
    DeviceArray<double> partials_;  // tree partials, then the iteration partials of the groups and the units
    DeviceArray<SysState> state_;
    DeviceArray<int> counter_;      // [nsys] active systems, [nsys + 1] active units, [nsys + 2 + node] tree
                                    // nodes, then one arrival counter per group
    DeviceArray<double> tvec_;      // b = 1 and x = 0 of time_iteration
    // what the kernels use: the arrays above, or a bound workspace
    SysState *st_ = nullptr;
    int *cnt_ = nullptr;
    double *part_ = nullptr, *gpart_ = nullptr, *upart_ = nullptr;
    void *ws_base_ = nullptr;
    size_t ws_bytes_ = 0;
    std::vector<SysState> h_states_;
    // y and z by iteration parity: x takes two iterations' updates in every
    // other K1, so the previous iteration's y and z must survive one more
    double *r_ = nullptr, *rs_ = nullptr, *p_[2] = {nullptr, nullptr}, *y_[2] = {nullptr, nullptr}, *a_ = nullptr,
           *s_ = nullptr, *z_[2] = {nullptr, nullptr}, *t_ = nullptr;
    const double *b_ = nullptr;
    double *x_ = nullptr;
    // graph cache
    // replayable iteration graphs of the current operands, one per grid rung
    // and K1 form, kept across solves
    struct GraphEntry {
        int iters, grid, compact, split;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs_;
    static constexpr int kGraphKey = 18;  // operands and layout the graphs were captured with
    static constexpr int kGraphKeyGeneration = 12;  // the slot holding the layout's generation
    const void *graph_key_[kGraphKey] = {};
    hipEvent_t check_ev_[2] = {nullptr, nullptr};  // host checks of the graph loop
    int *h_pub_ = nullptr, *d_pub_ = nullptr;      // counts published by the graphs (host-coherent)
    SysState *h_states_pin_ = nullptr;               // pinned staging of the states read back after a solve
    int h_states_cap_ = 0;
    // iterations the last graph-replayed solve on the current graph operands
    // launched before its systems stopped, with its stop rule (0: unknown);
    // the next such solve shortens its final graphs to match
    int pred_launches_ = 0, pred_max_iter_ = 0;
    double pred_rel_ = 0.0, pred_abs_ = 0.0;
    void drop_graph();
    // Consecutive unit launches walk the units in alternating directions, so
    // that a kernel starts on the rows the one before it touched last (still
    // in the memory-side cache).  Captured graphs keep the directions they
    // were captured with.
    int sweep_ = 0;
    bool flip() { return (sweep_++ & 1) != 0; }
    // A batch updates x in every other K1 (fewer bytes where bandwidth
    // bounds the launches); a single system, latency-bound, in every K1.
    bool defer_x_ = false;
    // r* stored in FP32 (an FP32 M); every other vector is FP64
    bool rs32_ = false;
    int x_mode(int parity) const { return defer_x_ ? (parity == 0 ? 1 : 2) : 0; }
    // without deferral one y and one z serve every iteration (read for x,
    // then overwritten, row by row, as before the double buffers)
    double *ybuf(int parity) const { return y_[defer_x_ ? parity : 0]; }
    double *zbuf(int parity) const { return z_[defer_x_ ? parity : 0]; }
    // the second buffers, for OpInit to clear (null when not deferring)
    double *y1() const { return defer_x_ ? y_[1] : nullptr; }
    double *z1() const { return defer_x_ ? z_[1] : nullptr; }
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
