// Sparse formats, row tiles and the fused right-preconditioned BiCGSTAB.
//
// The recurrences are those of include/bicgstab/bicgstab_solver.h:106-232
// (one iteration: y = M p, a = A y, alpha = rho / (r*, a), s = r - alpha a,
// z = M s, t = A z, omega = (t, s) / (t, t), x += alpha y + omega z,
// r = s - omega t, rho' = (r*, r), beta = (rho'/rho)(alpha/omega),
// p = r + beta (p - omega a), stop on ||r||_2 < TOL).  The reference issues
// 18 hipSPARSE/hipBLAS calls and 5 blocking scalar reads per iteration; here
// one iteration is four kernels and every scalar stays in device memory:
//
//   K1  finishes the previous iteration: x += alpha y + omega z and
//       r = s - omega t per row, ||r||^2 reduced, the stop test and the
//       iteration count by the last unit; then y = M p' with
//       p' = beta p + r - beta omega a formed while gathering, p' stored
//   K2  a = A y,   (r*, a) reduced, alpha = rho / (r*, a) by the last unit
//   K3  z = M s    with s = r - alpha a formed while gathering, s stored
//   K4  t = A z,   (t, s), (t, t), (r*, s), (r*, t) reduced; omega,
//       rho' = (r*, s) - omega (r*, t) = (r*, r), beta by the last unit
//
// rho' is the reference's (r*, r) expanded by linearity (r = s - omega t),
// so that beta is known before K1 needs it; ||r|| for the stop test is
// reduced from r itself, as the reference computes it.  A system stops in
// the K1 after its last iteration, with x and r updated.
//
// Every kernel runs one workgroup per unit of the row layout (krylov.hpp):
// four SELL-64 slices, lane = row, the slice's column indices and values
// streamed with nontemporal coalesced loads and each row summed in storage
// order in registers; or a row tile whose nonzeros are staged in LDS as
// products and summed per row in storage order.  Reductions are per-unit
// partials summed in a fixed order by the last unit of each system, so
// results are reproducible run to run.  All systems of a batch share the
// launches; launches inside the iteration loop walk a list of the units of
// systems still iterating, so converged systems cost nothing once the list
// has been rebuilt.
#include "krylov.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace spai {

// Two kinds of reduction.  The iteration kernels (K1, K2, K4) publish one
// partial per group of units and leave the total to the kernel that needs it
// next: each of its workgroups sums the system's group partials in a fixed
// order, and the system's first unit commits the scalars formed from them to
// the system's state.  So no iteration kernel waits for the last of its
// workgroups.  A group of one unit (systems of at most kMaxGroups units)
// stores its partial directly; larger groups count arrivals and the last one
// sums the group's unit partials.
// The set-up reductions (r0 and the true residual) run up a tree per system:
// level 0 holds one partial per unit, each further level one partial per
// group of up to kFanIn consecutive nodes of the level below, and the root's
// value goes to the op's finish step.  Same-address atomics serialise at the
// memory side, so a counter that sees few arrivals (kFanIn) finishes sooner
// than one shared by many units.
struct Ctx {
    SysState *st;
    const int *sys_units;  // nsys + 1: first unit of each system
    const int *sys_nodes;  // nsys + 1: first tree node of each system
    int *counter;          // [0, nsys) unused, [nsys] active systems, [nsys + 1] active units,
                           // [nsys + 2 + node] node arrivals
    double *partials;      // kMaxRed per tree node
    int nsys;
    double *hist;          // optional ||r_k|| history, hist_stride entries per system
    int hist_stride;
    const int *sys_groups;  // nsys + 1: first partial-sum group of each system
    double *gpart;          // kSlots x ngroups group partials (slot-major)
    double *upart;          // kSlots x nunits unit partials of groups of several units (slot-major)
    int *gcount;            // ngroups arrival counters
    int ngroups, nunits;
};

namespace {

constexpr int kThreads = 256;
constexpr int kLdsProducts = 2048;  // products staged per tile (16 KB)
#ifndef SPAI_REDUCE_FANIN
#define SPAI_REDUCE_FANIN 8
#endif
constexpr int kFanIn = SPAI_REDUCE_FANIN;  // children per reduction tree node
// K1 splits into a streaming update and a one-vector M-apply while at least
// this many units are active (bandwidth-bound); below it the fused K1 saves a
// launch (latency-bound).  SELL-64 units (8 wavefronts of one slice) keep the
// fused K1 fed up to a larger count than row tiles.
#ifndef SPAI_SPLIT_UNITS
#define SPAI_SPLIT_UNITS 4096
#endif
#ifndef SPAI_SPLIT_UNITS_TILES
#define SPAI_SPLIT_UNITS_TILES 1024
#endif
constexpr int kSplitUnits = SPAI_SPLIT_UNITS;
// Column blocks of large one-system row-tile layouts: blocks (dividing the 8
// XCDs; each block's slice of an FP64 vector is n / kColBlocks * 8 bytes) and
// the smallest system that gets them (its FP64 vector, 4 MiB, fills an XCD's
// L2).
#ifndef SPAI_COLUMN_BLOCKS
#define SPAI_COLUMN_BLOCKS 4
#endif
constexpr int kColBlocks = SPAI_COLUMN_BLOCKS;
static_assert(kColBlocks >= 1 && 8 % kColBlocks == 0, "column blocks divide the 8 XCDs");
constexpr int kColBlockMinRows = 1 << 19;
// Rows per tile of a column block (a block holds about 1 / kColBlocks of a
// row's entries, so its tiles take more rows for the same products).
#ifndef SPAI_PART_TILE_ROWS
#define SPAI_PART_TILE_ROWS 512
#endif
constexpr int kRowsPerPartTile = SPAI_PART_TILE_ROWS;
constexpr int kSplitUnitsTiles = SPAI_SPLIT_UNITS_TILES;
#ifndef SPAI_COMPACT_EVERY
#define SPAI_COMPACT_EVERY 8
#endif
constexpr int kCompactEvery = SPAI_COMPACT_EVERY;  // iterations between active-list rebuilds inside a graph
constexpr int kMaxRed = 4;          // most values any kernel reduces (partials per tree node)
// Slots of the iteration partials: K1's (r, r), K2's (r*, a), K4's (t, s),
// (t, t), (r*, s), (r*, t).
enum { kSlotRR = 0, kSlotRsA = 1, kSlotOmega = 2, kSlots = 6 };
constexpr int kTakePer = kMaxGroups / kThreads;  // group partials per thread a consuming workgroup loads

// ------------------------------------------------------------ transpose ----
__global__ void k_fill_int(int *p, int v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_count_rows(const int *rowidx, int nnz, int *cnt) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) atomicAdd(&cnt[rowidx[p]], 1);
}

// Exclusive scan, three passes: per-block scan, scan of block sums (one
// block, sequential carry), add offsets.
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(const int *in, int *out, int n, int *bsum) {
    __shared__ int sh[kScanBlock];
    int i = blockIdx.x * kScanBlock + threadIdx.x;
    int v = i < n ? in[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
        int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = sh[kScanBlock - 1];
}

__global__ __launch_bounds__(kScanBlock) void k_scan_sums(int *bsum, int nb, int *total) {
    __shared__ int sh[kScanBlock];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += kScanBlock) {
        int i = base + threadIdx.x;
        int v = i < nb ? bsum[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kScanBlock; o <<= 1) {
            int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += sh[kScanBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(int *out, int n, const int *bsum, const int *total, int *last) {
    int i = blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) out[i] += bsum[blockIdx.x];
    if (i == 0 && last) *last = *total;
}

// out[0..n) = exclusive scan of in[0..n), out[n] = total
void exclusive_scan(const int *in, int *out, int n, DeviceArray<int> &tmp, hipStream_t st) {
    int nb = std::max(1, ceil_div(n, kScanBlock));
    tmp.reserve((size_t)nb + 1);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, st, in, out, n, tmp.get());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, st, tmp.get(), nb, tmp.get() + nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, st, out, n, tmp.get(), tmp.get() + nb, out + n);
    SPAI_CHECK_LAUNCH();
}

// One thread per column: append (column, csc position) into each row's slots.
__global__ void k_fill_rows(const int *colptr, const int *rowidx, int n, const int *rowptr, int *cursor, int *colidx,
                            int *src) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int p = colptr[c]; p < colptr[c + 1]; ++p) {
        int r = rowidx[p];
        int q = rowptr[r] + atomicAdd(&cursor[r], 1);
        colidx[q] = c;
        src[q] = p;
    }
}

// One thread per row: order the row by column (insertion sort up to 32
// entries, shell sort beyond), then write perm and the CSR values.
__global__ void k_sort_rows(int n, const int *rowptr, int *colidx, int *src, const double *cscval, double *csrval,
                            int *perm) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int b = rowptr[r], e = rowptr[r + 1], len = e - b;
    int *ci = colidx + b;
    int *sp = src + b;
    if (len <= 32) {
        for (int i = 1; i < len; ++i) {
            int kc = ci[i], ks = sp[i], j = i - 1;
            while (j >= 0 && ci[j] > kc) {
                ci[j + 1] = ci[j];
                sp[j + 1] = sp[j];
                --j;
            }
            ci[j + 1] = kc;
            sp[j + 1] = ks;
        }
    } else {
        int gap = 1;
        while (gap < len / 3) gap = 3 * gap + 1;
        for (; gap > 0; gap /= 3) {
            for (int i = gap; i < len; ++i) {
                int kc = ci[i], ks = sp[i], j = i;
                while (j >= gap && ci[j - gap] > kc) {
                    ci[j] = ci[j - gap];
                    sp[j] = sp[j - gap];
                    j -= gap;
                }
                ci[j] = kc;
                sp[j] = ks;
            }
        }
    }
    for (int q = b; q < e; ++q) {
        perm[src[q]] = q;
        if (csrval) csrval[q] = cscval[src[q]];
    }
}

// ----------------------------------------------------------- reductions ----
template <int NRED>
__device__ __forceinline__ void block_reduce(double (&v)[NRED], double *sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        if (lane == 0) sh[q * 4 + wave] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NRED; ++q) v[q] = (sh[q * 4 + 0] + sh[q * 4 + 1]) + (sh[q * 4 + 2] + sh[q * 4 + 3]);
    }
    __syncthreads();
}

// Cross-workgroup reductions (reduce_system below) publish partials and
// count arrivals per tree node; the last arrival combines them.
//
// No __threadfence(): on gfx950 an agent-scope fence writes back and
// invalidates the whole L2 (buffer_wbl2 / buffer_inv), which, issued by
// thousands of workgroups, evicts the vectors the others are gathering.
// Instead partials are agent-scope atomic stores (coherent across the XCDs'
// L2s), the thread waits until they are acknowledged (vmcnt(0)) before the
// relaxed counter increment, and the last workgroup reads them back with
// agent-scope atomic loads.
#ifndef SPAI_WAVE_SUM_DPP
#define SPAI_WAVE_SUM_DPP 1
#endif
template <int NRED>
__device__ __forceinline__ void wave_sum(double (&v)[NRED]) {
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
        if constexpr (SPAI_WAVE_SUM_DPP != 0) {
            v[q] = group_allsum<64>(v[q]);  // DPP inside rows of 16 lanes, then 2 permutes
        } else {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        }
    }
}

// Lane 0 publishes the partials, waits until they are acknowledged, then
// counts; returns (to the whole wavefront) whether this was the last of
// `expect` arrivals, in which case the counter is reset for the next launch.
template <int NRED>
__device__ __forceinline__ bool publish_and_count(double *dst, const double (&v)[NRED], int *counter, int expect,
                                                  size_t stride = 1) {
    int last = 0;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < NRED; ++q)
            __hip_atomic_store(&dst[q * stride], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const int prev = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = prev == expect - 1;
        if (last) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __shfl(last, 0, 64) != 0;
}

// Sum of src[t * kMaxRed + q] over t in [first, last), by one wavefront, in
// a fixed order (lane-strided partial sums, then a butterfly).
template <int NRED>
__device__ __forceinline__ void wave_sum_range(const double *src, int first, int last, double (&acc)[NRED]) {
#pragma unroll
    for (int q = 0; q < NRED; ++q) acc[q] = 0.0;
    for (int t = first + (int)(threadIdx.x & 63); t < last; t += 64) {
#pragma unroll
        for (int q = 0; q < NRED; ++q) acc[q] += load_coherent(&src[(size_t)t * kMaxRed + q]);
    }
    wave_sum<NRED>(acc);
}

// Per-unit partials up the system's reduction tree, then op.finish.  After
// the workgroup's partials are combined in LDS only wavefront 0 stays: the
// others retire at once instead of waiting on the publish round trips.  The
// last arrival at each node sums that node's children in a fixed order, so
// the result does not depend on arrival order.  Node numbering follows from
// the system's unit count alone: level l starts after the levels below it.
// W = wavefronts of the workgroup (sh holds NRED * W doubles)
template <int NRED, int W, class Op>
__device__ __forceinline__ void reduce_system(const Ctx &c, int unit, int s, double (&v)[NRED], Op &op, double *sh) {
    static_assert(NRED >= 1 && NRED <= kMaxRed, "partials hold at most kMaxRed values per tree node");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wave_sum<NRED>(v);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NRED; ++q) sh[q * W + wave] = v[q];
    }
    __syncthreads();
    if (wave != 0) return;
    double val[NRED];
#pragma unroll
    for (int q = 0; q < NRED; ++q) {
        val[q] = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w) val[q] += sh[q * W + w];
    }
    const int u0 = c.sys_units[s];
    int n = c.sys_units[s + 1] - u0;  // nodes in the current level
    int idx = unit - u0;              // this node's index in its level
    int node = c.sys_nodes[s];        // first node of the current level
    while (n > 1) {
        const int parent = idx / kFanIn;
        const int c0 = parent * kFanIn, c1 = min(c0 + kFanIn, n);
        const int next = node + n;  // first node of the parent level
        if (!publish_and_count<NRED>(&c.partials[(size_t)(node + idx) * kMaxRed], val,
                                     &c.counter[c.nsys + 2 + next + parent], c1 - c0))
            return;
        wave_sum_range<NRED>(c.partials + (size_t)node * kMaxRed, c0, c1, val);
        node = next;
        idx = parent;
        n = (n + kFanIn - 1) / kFanIn;
    }
    if (lane == 0) op.finish(c, s, val);
}

// ------------------------------------------------ iteration reductions ----
// What an op does with the iteration partials (members it may declare):
//   kTree     reduces up the tree (reduce_system) and finishes there
//   kPublish  leading values of its NRED sums published in slots
//             [kSlot, kSlot + kPublish) for the next kernels
//   kTake     consumes slots [kTakeSlot, kTakeSlot + kTake): every workgroup
//             sums the system's group partials before its rows (need_take()
//             and take(v))
//   kFold     the committing workgroup folds values preloaded by preload()
//             into its sums (the last of its NRED values) before they are
//             reduced, so that the loads overlap the rows
//   kCommit   the system's first unit calls finish(c, s, sums) once its
//             workgroup's sums are reduced
#define SPAI_OP_TRAIT(Name, Member, Type, Default)                                                  \
    template <class Op, class = void>                                                               \
    struct Name : std::integral_constant<Type, Default> {};                                         \
    template <class Op>                                                                             \
    struct Name<Op, std::void_t<decltype(Op::Member)>> : std::integral_constant<Type, Op::Member> {};
SPAI_OP_TRAIT(OpTree, kTree, bool, false)
SPAI_OP_TRAIT(OpPublish, kPublish, int, 0)
SPAI_OP_TRAIT(OpTake, kTake, int, 0)
SPAI_OP_TRAIT(OpFold, kFold, bool, false)
SPAI_OP_TRAIT(OpCommit, kCommit, bool, false)
#undef SPAI_OP_TRAIT

// The partial-sum group of `unit` (a unit of system s).
struct UnitGroup {
    int g;      // group index
    int first;  // first unit of the group
    int size;   // units in the group
};
__device__ __forceinline__ UnitGroup unit_group(const Ctx &c, int s, int unit) {
    const int u0 = c.sys_units[s], nu = c.sys_units[s + 1] - u0;
    const int G = group_units(nu);
    const int local = (unit - u0) / G;
    return {c.sys_groups[s] + local, u0 + local * G, min(G, nu - local * G)};
}

// This thread's share of system s's group partials of slots [slot, slot +
// NT): groups g0 + t, g0 + t + blockDim.x, ... (PER of them: a system has
// at most kMaxGroups groups).  Plain loads issued at once; summed where
// part() is used.
template <int NT, int PER = kTakePer>
struct GroupShare {
    double v[NT][PER];
    __device__ __forceinline__ void load(const Ctx &c, int s, int slot) {
        const int g0 = c.sys_groups[s], g1 = c.sys_groups[s + 1];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int g = g0 + (int)threadIdx.x + k * (int)blockDim.x;
#pragma unroll
            for (int q = 0; q < NT; ++q) v[q][k] = g < g1 ? c.gpart[(size_t)(slot + q) * c.ngroups + g] : 0.0;
        }
    }
    __device__ __forceinline__ double part(int q) const {
        double a = v[q][0];
#pragma unroll
        for (int k = 1; k < PER; ++k) a += v[q][k];
        return a;
    }
};

// Sum over the workgroup (W wavefronts) that every thread receives: wave
// sums, then the wavefronts' sums in wavefront order.  Every workgroup of W
// wavefronts forms the same value from the same inputs.
template <int NT, int W>
__device__ __forceinline__ void block_allsum(double (&v)[NT], double *shk) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wave_sum<NT>(v);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NT; ++q) shk[q * W + wave] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        double t = shk[q * W];
#pragma unroll
        for (int w = 1; w < W; ++w) t += shk[q * W + w];
        v[q] = t;
    }
}

// The scalars an op forms from the previous kernels' partials (every
// workgroup alike), in two steps: take_issue requests the group partials
// (and the committer's preloads) first thing, before the matrix stream, so
// that waiting for them does not wait for the stream; take_finish sums them
// once the op has read the system's state.
template <class Op, int W>
using TakeShare = GroupShare<(OpTake<Op>::value > 0 ? OpTake<Op>::value : 1), (kMaxGroups + W * 64 - 1) / (W * 64)>;

template <int W, class Op>
__device__ __forceinline__ void take_issue(const Ctx &c, int s, int unit, Op &op, TakeShare<Op, W> &gs) {
    if constexpr (OpTake<Op>::value > 0) gs.load(c, s, Op::kTakeSlot);
    if constexpr (OpFold<Op>::value) {
        if (unit == c.sys_units[s]) op.preload(c, s);
    }
}

template <int W, class Op>
__device__ __forceinline__ void take_finish(Op &op, const TakeShare<Op, W> &gs, double *shk) {
    constexpr int NT = OpTake<Op>::value;
    if constexpr (NT > 0) {
        if (op.need_take()) {  // uniform over the workgroup
            double v[NT];
#pragma unroll
            for (int q = 0; q < NT; ++q) v[q] = gs.part(q);
            block_allsum<NT, W>(v, shk);
            op.take(v);
        }
    }
}

// Sum of src[q * stride + t] over t in [first, last) by one wavefront, in a
// fixed order.
template <int NP>
__device__ __forceinline__ void wave_sum_units(const double *src, size_t stride, int first, int last,
                                               double (&acc)[NP]) {
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = 0.0;
    for (int t = first + (int)(threadIdx.x & 63); t < last; t += 64) {
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[q] += load_coherent(&src[(size_t)q * stride + t]);
    }
    wave_sum<NP>(acc);
}

// After the rows.  Tree ops: reduce_system.  Otherwise the workgroup's NRED
// sums are reduced in the block (wavefront 0 keeps them), the first
// kPublish go to the unit's group, and the system's first unit commits.
template <int W, class Op>
__device__ __forceinline__ void finish_unit(const Ctx &c, int unit, int s, double (&v)[Op::NRED > 0 ? Op::NRED : 1],
                                            Op &op, double *sh) {
    constexpr int NRED = Op::NRED;
    if constexpr (OpTree<Op>::value) {
        reduce_system<NRED, W>(c, unit, s, v, op, sh);
    } else if constexpr (NRED > 0) {
        constexpr int NP = OpPublish<Op>::value;
        static_assert(NP <= NRED, "published values are the leading sums");
        if constexpr (OpFold<Op>::value) op.fold(v);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        wave_sum<NRED>(v);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < NRED; ++q) sh[q * W + wave] = v[q];
        }
        __syncthreads();
        if (wave != 0) return;
        double val[NRED];
#pragma unroll
        for (int q = 0; q < NRED; ++q) {
            val[q] = 0.0;
#pragma unroll
            for (int w = 0; w < W; ++w) val[q] += sh[q * W + w];
        }
        if constexpr (NP > 0) {
            const UnitGroup ug = unit_group(c, s, unit);
            if (ug.size == 1) {
                if (lane == 0) {
#pragma unroll
                    for (int q = 0; q < NP; ++q) c.gpart[(size_t)(Op::kSlot + q) * c.ngroups + ug.g] = val[q];
                }
            } else {
                double pub[NP];
#pragma unroll
                for (int q = 0; q < NP; ++q) pub[q] = val[q];
                if (publish_and_count<NP>(&c.upart[(size_t)Op::kSlot * c.nunits + unit], pub, &c.gcount[ug.g],
                                          ug.size, (size_t)c.nunits)) {
                    double tot[NP];
                    wave_sum_units<NP>(c.upart + (size_t)Op::kSlot * c.nunits, (size_t)c.nunits, ug.first,
                                       ug.first + ug.size, tot);
                    if (lane == 0) {
#pragma unroll
                        for (int q = 0; q < NP; ++q) c.gpart[(size_t)(Op::kSlot + q) * c.ngroups + ug.g] = tot[q];
                    }
                }
            }
        }
        if constexpr (OpCommit<Op>::value) {
            if (lane == 0 && unit == c.sys_units[s]) op.finish(c, s, val);
        }
    } else if constexpr (OpCommit<Op>::value) {
        if (threadIdx.x == 0 && unit == c.sys_units[s]) op.finish(c, s, v);
    }
}

// ------------------------------------------------------- unit kernels ----
struct UnitView {
    const int *row;        // nunits + 1 row boundaries
    const int *sys;        // system of each unit
    const int *slice;      // nunits + 1 first slice of each unit (SELL)
    const int *slice_row;  // nslices + 1 (SELL)
    const int *slice_blk;  // nslices + 1, offsets in blocks of 64 entries (SELL)
    const int *list;          // active units in XCD-blocked order, nullptr: every unit (grid = nunits)
    const int *count;         // number of entries of the list
    const int *sys_row;       // nsys + 1 row boundaries of the systems
    int reverse;              // each XCD walks its range of units from the end
    const double *psum;       // column-block partial sums (k_tiles with PS > 0): block c of row i at psum[c * pstride + i]
    int pstride;
};

// The workgroup's unit, or -1 past the end of the active list.
// XCD-blocked order.  Workgroup b is dispatched to XCD b % 8; position L of
// a run of G units goes to workgroup xcd_workgroup(L, G) so that XCD x runs
// the contiguous positions [start_x, start_x + G/8 (+1 for x < G % 8)), and
// neighbouring rows (the gathers' halo) stay in one XCD's L2.
__device__ __forceinline__ int xcd_position(int b, int G) {
    const int x = b & 7, k = b >> 3, base = G >> 3, rem = G & 7;
    return x * base + min(x, rem) + k;
}
__device__ __forceinline__ int xcd_workgroup(int L, int G) {
    const int base = G >> 3, rem = G & 7;
    int x, k;
    if (L < rem * (base + 1)) {
        x = L / (base + 1);
        k = L - x * (base + 1);
    } else {  // base > 0 here, since L < G
        const int L2 = L - rem * (base + 1);
        x = rem + L2 / base;
        k = L2 - (x - rem) * base;
    }
    return 8 * k + x;
}

// Every unit kernel runs in XCD-blocked order, so the gathers' halo rows stay
// in one XCD's L2 (a batched step measured 3.7% faster than in layout
// order).  The compaction stores the active list in that order; without a
// list the workgroup index is mapped.
__device__ __forceinline__ int unit_of(const UnitView &uv) {
    const int b = (int)blockIdx.x;
    if (!uv.list) {
        if (!uv.reverse) return xcd_position(b, (int)gridDim.x);
        const int G = (int)gridDim.x, x = b & 7, k = b >> 3, base = G >> 3, rem = G & 7;
        const int nx = base + (x < rem ? 1 : 0);
        return x * base + min(x, rem) + (nx - 1 - k);
    }
    const int cnt = *uv.count;
    if (b >= cnt) return -1;
    int bb = b;
    if (uv.reverse) {
        const int x = b & 7, nx = (cnt - x + 7) >> 3;
        bb = 8 * (nx - 1 - (b >> 3)) + x;
    }
    return uv.list[bb];  // the grid never exceeds the list's capacity
}

template <class T>
__device__ __forceinline__ T stream_load(const T *p) {
    return __builtin_nontemporal_load(p);
}

constexpr int kRowsPerTile = 512;  // row cap of a tile (host-built)
constexpr int kRowsPerThread = kRowsPerTile / kThreads;
// Rows of a tile with more entries than this are summed by a whole wavefront
// (lane-strided partial sums, then a wave sum); shorter ones by the thread
// that owns the row, in storage order.
#ifndef SPAI_TILE_SHORT_ROW
#define SPAI_TILE_SHORT_ROW 16
#endif
constexpr int kTileShortRow = SPAI_TILE_SHORT_ROW;

// Row tiles.  For every row i of the unit's tile: y_i = sum_p val[p] *
// op.gather(idx[p]) (stored as products in LDS, then summed per row), then
// op.row(i, y_i, pre_i, acc) where pre_i = op.pre(i, acc) is the row's own
// work that does not need the product (operand loads, and outputs that can be
// stored at once), done before the product so that it overlaps with it and
// holds few registers across the gathers.  One tile needs two dependent
// global round trips: the index/value/row-pointer/operand loads, then the
// gathers.  A row of at most kTileShortRow entries is summed in storage order
// by the thread that owns it; a longer one (power-law rows reach 1000) by one
// wavefront: lane l sums entries l, l + 64, ... in order, then a fixed DPP
// tree adds the lanes, so that no thread walks a long row alone while the
// workgroup waits.  Every order is fixed: results are reproducible.
// HAS_MATRIX = false skips the product (streaming ops); with PS > 0 the
// product of row i is instead the sum of PS column-block partials
// uv.psum[c * uv.pstride + i], added in block order (see k_tiles_part).
template <bool HAS_MATRIX, class MT, class Op, int PS = 0>
__global__ __launch_bounds__(kThreads) void k_tiles(const int *__restrict__ ptr, const int *__restrict__ idx,
                                                    const MT *__restrict__ val, UnitView uv, Ctx c, Op op) {
    constexpr int NRED = Op::NRED;
    __shared__ double prod[HAS_MATRIX ? kLdsProducts : 1];
    __shared__ int rp[HAS_MATRIX ? kRowsPerTile + 1 : 1];
    __shared__ double rowsum[HAS_MATRIX ? kRowsPerTile : 1];  // sums of the long rows
    __shared__ double sh[(NRED > 0 ? NRED : 1) * 4];
    __shared__ double shk[(OpTake<Op>::value > 0 ? OpTake<Op>::value : 1) * 4];
    const int unit = unit_of(uv);
    if (unit < 0) return;
    const int s = uv.sys[unit];
    if (!op.enabled(c.st[s])) return;
    TakeShare<Op, kThreads / 64> gs;
    take_issue<kThreads / 64>(c, s, unit, op, gs);
    op.load(c.st[s]);
    const int tid = threadIdx.x;
    double acc[NRED > 0 ? NRED : 1];
#pragma unroll
    for (int q = 0; q < (NRED > 0 ? NRED : 1); ++q) acc[q] = 0.0;
    const int r0 = uv.row[unit], r1 = uv.row[unit + 1];
    const int nrows = r1 - r0;
    if (!HAS_MATRIX) {
        take_finish<kThreads / 64>(op, gs, shk);
        for (int i = r0 + tid; i < r1; i += kThreads) {
            const typename Op::Pre pr = op.pre(i, acc);
            double prod = 0.0;
            if constexpr (PS > 0) {
                prod = uv.psum[i];
#pragma unroll
                for (int q = 1; q < PS; ++q) prod += uv.psum[(size_t)q * uv.pstride + i];
            }
            op.row(i, prod, pr, acc);
        }
    } else {
        const int p0 = ptr[r0];
        const int cnt = ptr[r1] - p0;
        if (cnt <= kLdsProducts && nrows <= kRowsPerTile) {
            for (int k = tid; k <= nrows; k += kThreads) rp[k] = ptr[r0 + k] - p0;
            constexpr int U = kLdsProducts / kThreads;
            int jj[U];
            MT vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = u * kThreads + tid;
                jj[u] = q < cnt ? stream_load(&idx[p0 + q]) : 0;
                vv[u] = q < cnt ? stream_load(&val[p0 + q]) : (MT)0;
            }
            // the scalars (pre and the gathers use them) while the tile streams in
            take_finish<kThreads / 64>(op, gs, shk);
            typename Op::Pre pre[kRowsPerThread];
#pragma unroll
            for (int k = 0; k < kRowsPerThread; ++k) {
                const int i = r0 + tid + k * kThreads;
                if (i < r1) pre[k] = op.pre(i, acc);
            }
            double gv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = u * kThreads + tid;
                gv[u] = q < cnt ? op.gather(jj[u]) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = u * kThreads + tid;
                if (q < cnt) prod[q] = (double)vv[u] * gv[u];
            }
            __syncthreads();
            // long rows: wavefront w takes the rows of [64 w, 64 w + 64),
            // [64 w + 256, ...) that have more than kTileShortRow entries
            {
                const int lane = tid & 63;
                for (int base = tid - lane; base < nrows; base += kThreads) {
                    const int li = base + lane;
                    const bool lng = li < nrows && rp[li + 1] - rp[li] > kTileShortRow;
                    unsigned long long m = __ballot(lng);
                    while (m) {
                        const int j = __ffsll((long long)m) - 1;
                        m &= m - 1;
                        const int b = rp[base + j], e = rp[base + j + 1];
                        double part = 0.0;
                        for (int q = b + lane; q < e; q += 64) part += prod[q];
                        part = group_allsum<64>(part);
                        if (lane == 0) rowsum[base + j] = part;
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kRowsPerThread; ++k) {
                const int li = tid + k * kThreads;
                if (li < nrows) {
                    const int b = rp[li], e = rp[li + 1];
                    double sum = 0.0;
                    if (e - b > kTileShortRow) {
                        sum = rowsum[li];
                    } else {
                        for (int q = b; q < e; ++q) sum += prod[q];
                    }
                    op.row(r0 + li, sum, pre[k], acc);
                }
            }
        } else {
            // a row longer than the LDS buffer (it has a tile of its own):
            // the whole workgroup reduces it
            take_finish<kThreads / 64>(op, gs, shk);
            for (int i = r0; i < r1; ++i) {
                const int b = ptr[i], e = ptr[i + 1];
                double part[1] = {0.0};
                for (int q = b + tid; q < e; q += kThreads) part[0] += (double)val[q] * op.gather(idx[q]);
                block_reduce<1>(part, sh);
                if (tid == 0) op.row(i, part[0], op.pre(i, acc), acc);
            }
        }
    }
    finish_unit<kThreads / 64>(c, unit, s, acc, op, sh);
}

// Column-blocked products (row tiles of one large system, TileMap::cblocks =
// C): workgroup b runs on XCD b % 8, which serves column block (b % 8) % C.
// A tile forms, for each of its rows, the sum of the products over the
// block's columns (the gathered values lie in the block's 1/C of the vector,
// which stays in the XCD's L2) and stores it as part[c * n + i]; every row of
// every block is written (0.0 for a row with no entry in the block).  The row
// sums follow k_tiles: storage order per row, long rows by a wavefront in a
// fixed tree.  Only ops whose gather reads one stored vector use it (no
// state, no sums).
//
// A workgroup's first load is its tile's descriptor (rows and entries,
// host-built), then the row pointers, column indices and values together,
// then the gathers: three dependent round trips instead of five (block
// offsets, tile bounds, row pointers, entries, gathers; power-law products
// 6% faster).  With kPartPipe the launch is persistent: the block's
// workgroups take its tiles in a fixed stride and request the next tile's
// operands behind the current tile's gathers.
struct PartTiles {
    const int4 *tile;     // per tile: first row, end row, first entry, end entry
    int off[kColBlocks + 1];  // block c's tiles are [off[c], off[c + 1])
};
#ifndef SPAI_PART_WG_PER_CU
#define SPAI_PART_WG_PER_CU 4
#endif
constexpr int kPartWgPerCu = SPAI_PART_WG_PER_CU;  // persistent workgroups per CU (32 CUs per XCD)
// Products staged per tile of a column block.  A block's rows hold few
// entries each (a quarter of a row's), so its tiles are cut by rows
// (kRowsPerPartTile) long before they fill the buffer; 1024 keeps fewer idle
// slots in registers and measured the same (power-law: 278 against 275 ms
// for 1000 iterations).
#ifndef SPAI_PART_PRODUCTS
#define SPAI_PART_PRODUCTS 2048
#endif
constexpr int kPartProducts = SPAI_PART_PRODUCTS;
static_assert(kPartProducts % kThreads == 0 && kPartProducts <= kLdsProducts, "part tile products");
// 1: persistent workgroups that request the next tile's operands behind the
// current tile's gathers; 0: one workgroup per tile.  On the power-law
// matrix (1000 iterations) one workgroup per tile took 275 ms, the
// persistent form 278-294 ms (4, 6 or 8 workgroups per CU): the products
// are bound by their gathers, not by the dependent round trips.
#ifndef SPAI_PART_PIPE
#define SPAI_PART_PIPE 0
#endif
constexpr bool kPartPipe = SPAI_PART_PIPE != 0;
constexpr int kPartRp = (kRowsPerPartTile + 1 + kThreads - 1) / kThreads;  // row pointers per thread

template <class MT>
struct PartOperands {
    static constexpr int U = kPartProducts / kThreads;
    int j[U];
    MT v[U];
    int rp[kPartRp];
};

// Requests a tile's operands (nothing for a tile beyond the staging buffer).
template <class MT>
__device__ __forceinline__ void part_load(const int4 d, const int *__restrict__ bp, const int *__restrict__ idx,
                                          const MT *__restrict__ val, int tid, PartOperands<MT> &o) {
    constexpr int U = PartOperands<MT>::U;
    const int nrows = d.y - d.x, cnt = d.w - d.z;
    const bool fits = cnt <= kPartProducts && nrows <= kRowsPerPartTile;
#pragma unroll
    for (int k = 0; k < kPartRp; ++k) {
        const int kk = k * kThreads + tid;
        o.rp[k] = fits && kk <= nrows ? bp[d.x + kk] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int q = u * kThreads + tid;
        const bool in = fits && q < cnt;
        o.j[u] = in ? stream_load(&idx[d.z + q]) : 0;
        o.v[u] = in ? stream_load(&val[d.z + q]) : (MT)0;
    }
}

template <class MT, class Op, int C>
__global__ __launch_bounds__(kThreads) void k_tiles_part(const int *__restrict__ ptr, const int *__restrict__ idx,
                                                         const MT *__restrict__ val, PartTiles pt, int n,
                                                         double *__restrict__ part, Ctx c, Op op) {
    constexpr int U = PartOperands<MT>::U;
    __shared__ double prod[kPartProducts];  // a long row's sum replaces its first product
    __shared__ int rp[kRowsPerPartTile + 1];
    __shared__ double sh[4];
    const int b = (int)blockIdx.x, x = b & 7, blk = x % C;
    constexpr int per = 8 / C;                           // XCDs per block
    const int w = (b >> 3) * per + x / C;                // the workgroup among its block's
    const int nw = ((int)gridDim.x >> 3) * per;          // workgroups per block (the grid is a multiple of 8)
    int t0 = 0, t1 = 0;
#pragma unroll
    for (int k = 0; k < C; ++k)
        if (k == blk) {
            t0 = pt.off[k];
            t1 = pt.off[k + 1];
        }
    int t = t0 + w;
    if (t >= t1) return;
    if (!op.enabled(c.st[0])) return;
    const int tid = threadIdx.x;
    const int *bp = ptr + (size_t)blk * n;  // bp[i]: first entry of row i in the block; bp[n] ends the block
    double *out = part + (size_t)blk * n;
    int4 d = pt.tile[t];
    PartOperands<MT> cur;
    part_load(d, bp, idx, val, tid, cur);
    for (;;) {
        const int tn = t + nw;
        const bool more = kPartPipe && tn < t1;
        int4 dn = d;
        if (more) dn = pt.tile[tn];
        const int r0 = d.x, nrows = d.y - d.x, p0 = d.z, cnt = d.w - d.z;
        PartOperands<MT> nxt;
        if (cnt <= kPartProducts && nrows <= kRowsPerPartTile) {
#pragma unroll
            for (int k = 0; k < kPartRp; ++k) {
                const int kk = k * kThreads + tid;
                if (kk <= nrows) rp[kk] = cur.rp[k] - p0;
            }
            double gv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = u * kThreads + tid;
                gv[u] = q < cnt ? op.gather(cur.j[u]) : 0.0;
            }
            // the next tile's operands, behind this tile's gathers
            if (more) part_load(dn, bp, idx, val, tid, nxt);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = u * kThreads + tid;
                if (q < cnt) prod[q] = (double)cur.v[u] * gv[u];
            }
            __syncthreads();
            {
                const int lane = tid & 63;
                for (int base = tid - lane; base < nrows; base += kThreads) {
                    const int li = base + lane;
                    const bool lng = li < nrows && rp[li + 1] - rp[li] > kTileShortRow;
                    unsigned long long m = __ballot(lng);
                    while (m) {
                        const int j = __ffsll((long long)m) - 1;
                        m &= m - 1;
                        const int bb = rp[base + j], e = rp[base + j + 1];
                        double s = 0.0;
                        for (int qq = bb + lane; qq < e; qq += 64) s += prod[qq];
                        s = group_allsum<64>(s);  // every lane has read the row's products
                        if (lane == 0) prod[bb] = s;
                    }
                }
            }
            __syncthreads();
            for (int li = tid; li < nrows; li += kThreads) {
                const int bb = rp[li], e = rp[li + 1];
                double s = 0.0;
                if (e - bb > kTileShortRow) {
                    s = prod[bb];
                } else {
                    for (int qq = bb; qq < e; ++qq) s += prod[qq];
                }
                __builtin_nontemporal_store(s, &out[r0 + li]);
            }
        } else {
            // a row longer than the staging buffer (a tile of its own)
            if (more) part_load(dn, bp, idx, val, tid, nxt);
            for (int i = d.x; i < d.y; ++i) {
                double s[1] = {0.0};
                for (int qq = bp[i] + tid; qq < bp[i + 1]; qq += kThreads) s[0] += (double)val[qq] * op.gather(idx[qq]);
                block_reduce<1>(s, sh);
                if (tid == 0) out[i] = s[0];
            }
        }
        if (!more) break;
        __syncthreads();  // the LDS is refilled for the next tile
        t = tn;
        d = dn;
        cur = nxt;
    }
}

// SELL-64.  Wavefront w of the workgroup takes kSellSlicesPerWave
// consecutive slices of its unit, lane = row.  The slices' column indices and
// values (column-major, width = the slice's longest row) are streamed with
// nontemporal loads, kSellUnroll columns of every slice in flight per lane,
// and each row is summed in storage order in a register.  The first columns
// are requested before the system's state is read, so that the state load
// overlaps the matrix stream.  No LDS and no barrier except the reduction's.
#ifndef SPAI_SELL_UNROLL
#define SPAI_SELL_UNROLL 8
#endif
constexpr int kSellUnroll = SPAI_SELL_UNROLL;  // columns in flight per slice and lane
// Ops whose gather reads several vectors (K3 forms s = r - alpha a from two,
// the fused K1 p' from four) stage what the gather of each of the
// workgroup's own rows returns in LDS, and gathers inside the unit read that
// copy.  An op opts in with kStage = true; for single-vector gathers the
// barrier costs more than the loads it saves.  SPAI_SELL_STAGE: 0 none, 1
// opted-in ops, 2 every op.
#ifndef SPAI_SELL_STAGE
#define SPAI_SELL_STAGE 1
#endif
template <class Op, class = void>
struct StageTrait : std::false_type {};
template <class Op>
struct StageTrait<Op, std::void_t<decltype(Op::kStage)>> : std::integral_constant<bool, Op::kStage> {};
template <class Op>
constexpr bool sell_stage() {
    return SPAI_SELL_STAGE == 2 || (SPAI_SELL_STAGE == 1 && StageTrait<Op>::value);
}
// optional occupancy floor (waves per SIMD) for the register allocator
#ifndef SPAI_SELL_MIN_WAVES
#define SPAI_SELL_MIN_WAVES 0
#endif
#if SPAI_SELL_MIN_WAVES > 0
#define SPAI_SELL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SPAI_SELL_MIN_WAVES)))
#else
#define SPAI_SELL_OCCUPANCY
#endif
// Column storage of a SELL-64 operator (one instance of k_sell each):
//   kIdx32    32-bit column per entry;
//   kIdx16    16-bit column - row offset per entry;
//   kIdxDiag  one column - row offset per column block of a slice, shared by
//             its 64 rows (a slice's distinct offsets in ascending order,
//             0.0 where a row has no entry at that offset).  The gathered
//             index row + offset is clamped into the system's rows, so an
//             entry of 0.0 reads a row of its own system; the index does not
//             wait for the value.
enum { kIdx32 = 0, kIdx16 = 1, kIdxDiag = 2 };
template <int IX>
struct SellIndex {
    using T = int;
};
template <>
struct SellIndex<kIdx16> {
    using T = short;
};

constexpr int kSellThreads = 64 * kSellWaves;
template <bool HAS_MATRIX, class MT, class Op, int IX>
__global__ __launch_bounds__(kSellThreads) SPAI_SELL_OCCUPANCY void k_sell(const typename SellIndex<IX>::T *__restrict__ col,
                                                                       const MT *__restrict__ val, UnitView uv, Ctx c,
                                                                       Op op) {
    using IT = typename SellIndex<IX>::T;
    constexpr int NRED = Op::NRED;
    constexpr int SPW = kSellSlicesPerWave;
    constexpr int U = kSellUnroll;
    constexpr bool STAGE = HAS_MATRIX && sell_stage<Op>();
    __shared__ double sh[(NRED > 0 ? NRED : 1) * kSellWaves];
    __shared__ double shk[(OpTake<Op>::value > 0 ? OpTake<Op>::value : 1) * kSellWaves];
    __shared__ double sv[STAGE ? kSlicesPerUnit * kSliceRows : 1];  // the unit's gathered values
    const int unit = unit_of(uv);
    if (unit < 0) return;
    const int s = uv.sys[unit];
    TakeShare<Op, kSellWaves> gs;
    take_issue<kSellWaves>(c, s, unit, op, gs);
    const int lane = threadIdx.x & (kSliceRows - 1);
    const int sl_end = uv.slice[unit + 1];
    const int sl0 = uv.slice[unit] + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kSliceRows)) * SPW;
    // the unit's rows [u_lo, u_hi) (its slices are consecutive)
    const int u_lo = STAGE ? uv.slice_row[uv.slice[unit]] : 0, u_hi = STAGE ? uv.slice_row[sl_end] : 0;
    int row[SPW], w[SPW], b0[SPW];
    bool live[SPW];
#pragma unroll
    for (int j = 0; j < SPW; ++j) {
        const int sl = sl0 + j;
        const bool have = sl < sl_end;
        const int r0 = have ? uv.slice_row[sl] : 0, r1 = have ? uv.slice_row[sl + 1] : 0;
        b0[j] = have ? uv.slice_blk[sl] : 0;
        w[j] = have ? uv.slice_blk[sl + 1] - b0[j] : 0;
        row[j] = r0 + lane;
        live[j] = row[j] < r1;
    }
    int lo = 0, hi = 0;  // the system's rows (diagonal slices)
    if constexpr (IX == kIdxDiag) {
        lo = uv.sys_row[s];
        hi = uv.sys_row[s + 1] - 1;
    }
    IT jj[SPW][U];
    MT vv[SPW][U];
    auto load_cols = [&](int k0) {
#pragma unroll
        for (int j = 0; j < SPW; ++j)
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const bool ok = k0 + k < w[j];
                const long long q = (long long)(b0[j] + k0 + k) * kSliceRows + lane;
                if constexpr (IX == kIdxDiag) jj[j][k] = ok ? col[b0[j] + k0 + k] : 0;
                else jj[j][k] = ok ? stream_load(&col[q]) : (IT)0;
                vv[j][k] = ok ? stream_load(&val[q]) : (MT)0;
            }
    };
    if (HAS_MATRIX) load_cols(0);
    if (!op.enabled(c.st[s])) return;
    op.load(c.st[s]);
    take_finish<kSellWaves>(op, gs, shk);
    double acc[NRED > 0 ? NRED : 1];
#pragma unroll
    for (int q = 0; q < (NRED > 0 ? NRED : 1); ++q) acc[q] = 0.0;
    if (!HAS_MATRIX) {
#pragma unroll
        for (int j = 0; j < SPW; ++j)
            if (live[j]) op.row(row[j], 0.0, op.pre(row[j], acc), acc);
    } else {
        typename Op::Pre pre[SPW];
#pragma unroll
        for (int j = 0; j < SPW; ++j) pre[j] = live[j] ? op.pre(row[j], acc) : typename Op::Pre{};
        if constexpr (STAGE) {
            // what a gather of each own row returns, for the gathers that
            // stay inside the unit (every wavefront reaches the barrier)
#pragma unroll
            for (int j = 0; j < SPW; ++j)
                if (live[j]) sv[row[j] - u_lo] = op.stage(row[j], pre[j]);
            __syncthreads();
        }
        double sum[SPW];
        int wmax = 0;
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            sum[j] = 0.0;
            wmax = max(wmax, w[j]);
        }
        for (int k0 = 0;;) {
            double gv[SPW][U];
#pragma unroll
            for (int j = 0; j < SPW; ++j)
#pragma unroll
                for (int k = 0; k < U; ++k) {
                    int cj;
                    if constexpr (IX == kIdx32) cj = jj[j][k];
                    else if constexpr (IX == kIdx16) cj = row[j] + (int)jj[j][k];
                    else cj = min(max(row[j] + jj[j][k], lo), hi);
                    double g = 0.0;
                    if (k0 + k < w[j]) {
                        if constexpr (!STAGE) {
                            g = op.gather(cj);
                        } else if constexpr (IX == kIdxDiag) {
                            // the wavefront's 64 rows share the offset
                            const int first = row[j] - lane + jj[j][k];
                            g = (first >= u_lo && first + kSliceRows <= u_hi) ? sv[cj - u_lo] : op.gather(cj);
                        } else {
                            const unsigned rel = (unsigned)(cj - u_lo);
                            g = rel < (unsigned)(u_hi - u_lo) ? sv[rel] : op.gather(cj);
                        }
                    }
                    gv[j][k] = g;
                }
#pragma unroll
            for (int j = 0; j < SPW; ++j)
#pragma unroll
                for (int k = 0; k < U; ++k)
                    if (k0 + k < w[j]) sum[j] += (double)vv[j][k] * gv[j][k];
            k0 += U;
            if (k0 >= wmax) break;
            load_cols(k0);  // rows longer than U: the next U columns
        }
#pragma unroll
        for (int j = 0; j < SPW; ++j)
            if (live[j]) op.row(row[j], sum[j], pre[j], acc);
    }
    finish_unit<kSellWaves>(c, unit, s, acc, op, sh);
}

// ------------------------------------------------------------- the ops ----
// The shadow residual r* is stored as RT: double with an FP64 M, float with
// an FP32 (mixed) M.  It is rounded once, when OpInit forms it, and rho and
// every later inner product read the rounded value, so the solve is BiCGSTAB
// with r* = fl32(r0), another valid shadow vector (its oracle iteration
// counts spread as those of r* = r0).  Every other vector stays FP64.
//
// r = b - A x0, r* = r, s = r, p = a = t = y = z = 0; rho = (r*, r),
// rr = (r, r); tol.  With alpha = omega = beta = 0 the first K1 reproduces r
// and forms p' = r.
template <class RT>
struct OpInit {
    static constexpr int NRED = 3;
    static constexpr bool kTree = true;
    struct Pre {
        double b;
    };
    const double *x, *b;
    double *r;
    RT *rs;
    double *p0, *p1, *a, *s, *t, *y, *z, *y1, *z1;
    double rel_tol, abs_tol;
    int restart;  // 1: restart of an active system from its current x (iterations and tol kept)
    __device__ bool enabled(const SysState &S) const { return !restart || S.active; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int i, double (&)[3]) const { return {b[i]}; }
    __device__ double gather(int j) const { return x[j]; }
    __device__ double stage(int i, const Pre &) const { return x[i]; }
    __device__ void row(int i, double ax, const Pre &pr, double (&acc)[3]) const {
        double ri = -ax + pr.b;
        const RT rsi = (RT)ri;
        r[i] = ri;
        rs[i] = rsi;
        s[i] = ri;
        p0[i] = 0.0;
        p1[i] = 0.0;
        a[i] = 0.0;
        t[i] = 0.0;
        y[i] = 0.0;
        z[i] = 0.0;
        if (y1) {  // the second y and z buffers exist only when x updates are deferred
            y1[i] = 0.0;
            z1[i] = 0.0;
        }
        acc[0] += (double)rsi * ri;  // (r*, r): (r, r) itself when RT is double
        acc[1] += ri * ri;
        acc[2] += pr.b * pr.b;
    }
    __device__ void finish(const Ctx &c, int s_, const double (&tot)[3]) const {
        SysState &S = c.st[s_];
        S.rho = tot[0];
        S.rho_next = tot[0];
        S.rr = tot[1];
        S.alpha = 0.0;
        S.omega = 0.0;
        S.beta = 0.0;
        S.alpha_prev = 0.0;
        S.omega_prev = 0.0;
        S.true_rr = 0.0;
        S.xdefer = 0;
        S.k1_count = 0;
        S.k2_count = 0;
        S.k1_defer = 0;
        S.iters_base = restart ? S.iters : 0;
        if (restart) {
            S.restarts += 1;
            if (sqrt(tot[1]) < S.tol) {
                S.status = kConverged;
                S.active = 0;
                atomicSub(&c.counter[c.nsys], 1);
            }
            return;
        }
        S.err0 = sqrt(tot[1]);
        S.tol = abs_tol > 0.0 ? abs_tol : rel_tol * sqrt(tot[2]);
        S.iters = 0;
        if (S.err0 < S.tol) {
            // r0 already below the threshold: the reference would divide 0/0
            S.status = kConvergedAtStart;
            S.active = 0;
            atomicSub(&c.counter[c.nsys], 1);
        } else if (S.max_iter <= 1) {
            // the reference's loop (iter_times = 1; while (iter_times <
            // MAX_ITER)) runs no iteration: MAX_ITER is returned with x = x0
            S.status = kMaxIter;
            S.iters = S.max_iter;
            S.active = 0;
            atomicSub(&c.counter[c.nsys], 1);
        }
    }
};

// K1.  Per row: the previous iteration's x += alpha y + omega z and
// r = s - omega t, (r, r) published; p' = beta p + r - beta omega a.  omega,
// rho' and beta come from K4's partials (every workgroup forms them alike):
// omega = (t, s) / (t, t), rho' = (r*, s) - omega (r*, t) = (r*, r),
// beta = (rho' / rho) (alpha / omega); the first K1 after OpInit has
// alpha = omega = beta = 0, reproduces r and forms p' = r.  The system's
// first unit commits omega, beta and rho' (K2 makes rho' the rho).  YMODE:
//   kYProduct  y = M p' gathered from p'(j) formed on the fly (bit-identical
//              to the stored p'): the fused K1, four gathered vectors
//   kYCopy     y = p' (no preconditioner)
//   kYNone     no y: the streaming half of the split K1, OpApplyM follows
enum { kYProduct = 0, kYCopy = 1, kYNone = 2 };
enum { kXOne = 0, kXDefer = 1, kXTwo = 2 };
#ifndef SPAI_STAGE_K1
#define SPAI_STAGE_K1 1
#endif
template <class MT, int YMODE, int XMODE>
struct OpMapplyP {
    static constexpr int NRED = 1;
    static constexpr int kPublish = 1, kSlot = kSlotRR;
    static constexpr int kTake = 4, kTakeSlot = kSlotOmega;
    static constexpr bool kCommit = true;
    static constexpr bool kStage = SPAI_STAGE_K1 != 0;
    struct Pre {
        double p;
    };
    double *x, *r, *pout, *y;  // y: this iteration's buffer (this parity)
    const double *s, *t, *pin, *a, *z;
    const double *yb, *zb;  // the previous iteration's y and z (the other parity)
    // XMODE, the x update: kXOne the previous iteration's (every K1);
    // kXDefer none, x is left for the next K1 (even iterations of a batch);
    // kXTwo the two previous iterations' (odd iterations of a batch)
    double alpha, omega, alpha_prev, omega_prev, nomega, beta, nbw, rho, rho_new;
    int first;  // no K2 since OpInit
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) {
        alpha = S.alpha;
        rho = S.rho;
        first = S.k2_count == 0;
        if constexpr (XMODE == kXTwo) {
            alpha_prev = S.alpha_prev;
            omega_prev = S.omega_prev;
        }
        omega = 0.0;
        beta = 0.0;
        rho_new = rho;
        nomega = -omega;
        nbw = -beta * omega;
    }
    __device__ bool need_take() const { return !first; }
    __device__ void take(const double (&v)[4]) {
        omega = wave_uniform(v[0] / v[1]);
        rho_new = wave_uniform(v[2] - omega * v[3]);
        beta = wave_uniform((rho_new / rho) * (alpha / omega));
        nomega = -omega;
        nbw = wave_uniform(-beta * omega);
    }
    __device__ double rval(int j) const { return s[j] + nomega * t[j]; }
    __device__ double pval(int j, double rj) const { return (beta * pin[j] + rj) + nbw * a[j]; }
    // the row's x, r and p' are stored before the product (no other row's
    // gather reads x, r or p')
    __device__ Pre pre(int i, double (&acc)[1]) const {
        const double ri = rval(i);
        const double pi = pval(i, ri);
        if constexpr (XMODE != kXDefer) {
            double xi = x[i];
            // kXTwo: first the update of two iterations back (its y and z are
            // this parity's buffers: y is overwritten only by this row's
            // product below, z by K3), in the order and rounding of one
            // update per K1
            if constexpr (XMODE == kXTwo) xi = (xi + alpha_prev * y[i]) + omega_prev * z[i];
            x[i] = (xi + alpha * yb[i]) + omega * zb[i];
        }
        r[i] = ri;
        pout[i] = pi;
        acc[0] += ri * ri;
        return {pi};
    }
    __device__ double gather(int j) const { return pval(j, rval(j)); }
    __device__ double stage(int, const Pre &pr) const { return pr.p; }
    __device__ void row(int i, double mx, const Pre &pr, double (&)[1]) const {
        if (YMODE == kYProduct) y[i] = mx;
        else if (YMODE == kYCopy) y[i] = pr.p;
    }
    __device__ void finish(const Ctx &c, int s_, const double (&)[1]) const {
        SysState &S = c.st[s_];
        S.omega = omega;
        S.beta = beta;
        S.rho_next = rho_new;
        S.k1_count += 1;
        S.k1_defer = XMODE == kXDefer;
        if constexpr (XMODE == kXDefer) {  // the update this K1 leaves for the next
            S.alpha_prev = alpha;
            S.omega_prev = omega;
        }
    }
};

// x += the update a system that stopped in a deferring K1 still owes:
// (x + alpha_prev y) + omega_prev z with the odd-parity y and z.
struct OpFlushX {
    static constexpr int NRED = 0;
    struct Pre {};
    double *x;
    const double *y1, *z1;
    double alpha, omega;
    __device__ bool enabled(const SysState &S) const { return S.xdefer != 0; }
    __device__ void load(const SysState &S) {
        alpha = S.alpha_prev;
        omega = S.omega_prev;
    }
    __device__ Pre pre(int i, double (&)[1]) const {
        x[i] = (x[i] + alpha * y1[i]) + omega * z1[i];
        return {};
    }
    __device__ double gather(int) const { return 0.0; }
    __device__ double stage(int, const Pre &) const { return 0.0; }
    __device__ void row(int, double, const Pre &, double (&)[1]) const {}
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

__global__ void k_clear_xdefer(SysState *st, int nsys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) st[s].xdefer = 0;
}

// Second half of the split K1: y = M p' with p' stored by the first half
// (one gathered vector instead of four).
struct OpApplyM {
    static constexpr int NRED = 0;
    struct Pre {};
    const double *p;
    double *y;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int, double (&)[1]) const { return {}; }
    __device__ double gather(int j) const { return p[j]; }
    __device__ double stage(int i, const Pre &) const { return p[i]; }
    __device__ void row(int i, double mx, const Pre &, double (&)[1]) const { y[i] = mx; }
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

// K2: a = A y, (r*, a) published.  The system's first unit also sums K1's
// (r, r) partials (loaded before its rows, folded into its block sums) and
// runs the stop test of the iteration K1 finished (the reference's loop
// tail, bicgstab_solver.h:224-231): converged, then non-finite, then
// MAX_ITER.  A stopping system's other workgroups still form a and their
// partials, which nothing reads.
template <class RT>
struct OpSpmvAlpha {
    static constexpr int NRED = 2;  // (r*, a); the committer's (r, r)
    static constexpr int kPublish = 1, kSlot = kSlotRsA;
    static constexpr bool kFold = true, kCommit = true;
    struct Pre {
        double rs;
    };
    const double *y;
    const RT *rs;
    double *a;
    double rrv[kTakePer];
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &) {}
    // committer only (the other workgroups keep rrv = 0 from the launch)
    __device__ void preload(const Ctx &c, int s_) {
        GroupShare<1> gs;
        gs.load(c, s_, kSlotRR);
#pragma unroll
        for (int k = 0; k < kTakePer; ++k) rrv[k] = gs.v[0][k];
    }
    __device__ void fold(double (&acc)[2]) const {
        double t = rrv[0];
#pragma unroll
        for (int k = 1; k < kTakePer; ++k) t += rrv[k];
        acc[1] = t;
    }
    __device__ Pre pre(int i, double (&)[2]) const { return {(double)rs[i]}; }
    __device__ double gather(int j) const { return y[j]; }
    __device__ double stage(int i, const Pre &) const { return y[i]; }
    __device__ void row(int i, double ax, const Pre &pr, double (&acc)[2]) const {
        a[i] = ax;
        acc[0] += pr.rs * ax;
    }
    __device__ void finish(const Ctx &c, int s, const double (&tot)[2]) const {
        SysState &S = c.st[s];
        const double rr = tot[1];
        S.rr = rr;
        S.rho = S.rho_next;
        S.k2_count += 1;
        if (S.k1_count < 2) return;  // the first K1 finished no iteration
        S.iters = S.iters_base + S.k1_count - 1;
        const double err = sqrt(rr);
        if (c.hist && S.iters <= c.hist_stride) c.hist[(size_t)s * c.hist_stride + S.iters - 1] = err;
        bool stop = false;
        if (err < S.tol) {
            S.status = kConverged;
            stop = true;
        } else if (!(err == err) || !isfinite(S.beta)) {
            S.status = kBreakdown;
            stop = true;
        } else if (S.iters >= S.max_iter - 1) {
            // the reference returns MAX_ITER after MAX_ITER - 1 iterations
            S.status = kMaxIter;
            S.iters = S.max_iter;
            stop = true;
        }
        if (stop) {
            S.active = 0;
            S.xdefer = S.k1_defer;  // OpFlushX applies the update that K1 left
            atomicSub(&c.counter[c.nsys], 1);
        }
    }
};

// alpha = rho / (r*, a) from K2's partials, formed by every workgroup of a
// K3; the system's first unit commits it.
struct TakeAlpha {
    double rho, alpha, nalpha;
    __device__ void load_rho(const SysState &S) { rho = S.rho; }
    __device__ void take(const double (&v)[1]) {
        alpha = wave_uniform(rho / v[0]);
        nalpha = -alpha;
    }
    __device__ void commit(const Ctx &c, int s) const { c.st[s].alpha = alpha; }
};

template <class MT>
struct OpMapplyS {  // K3: z = M s, s = r - alpha a (stored before the product; gathers read r and a)
    static constexpr int NRED = 0;
    static constexpr int kTake = 1, kTakeSlot = kSlotRsA;
    static constexpr bool kCommit = true;
    static constexpr bool kStage = true;
    struct Pre {
        double s;
    };
    const double *r, *a;
    double *s, *z;
    TakeAlpha ta;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) { ta.load_rho(S); }
    __device__ bool need_take() const { return true; }
    __device__ void take(const double (&v)[1]) { ta.take(v); }
    __device__ void finish(const Ctx &c, int s_, const double (&)[1]) const { ta.commit(c, s_); }
    __device__ double sval(int j) const { return r[j] + ta.nalpha * a[j]; }
    __device__ Pre pre(int i, double (&)[1]) const {
        const double si = sval(i);
        s[i] = si;
        return {si};
    }
    __device__ double gather(int j) const { return sval(j); }
    __device__ double stage(int, const Pre &pr) const { return pr.s; }
    __device__ void row(int i, double mx, const Pre &, double (&)[1]) const { z[i] = mx; }
};

// K3 split in two for row tiles: s = r - alpha a as a streaming pass, then
// z = M s (OpApplyM) gathering the stored s alone.  The random gathers of a
// row tile mostly miss L2 (a 64-byte line fetched per 8-byte value), so one
// gathered vector instead of two halves the M-apply's fetched bytes.  s is
// formed by the same expression as OpMapplyS's, so z is the same.
struct OpUpdateS {
    static constexpr int NRED = 0;
    static constexpr int kTake = 1, kTakeSlot = kSlotRsA;
    static constexpr bool kCommit = true;
    struct Pre {};
    const double *r, *a;
    double *s;
    TakeAlpha ta;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) { ta.load_rho(S); }
    __device__ bool need_take() const { return true; }
    __device__ void take(const double (&v)[1]) { ta.take(v); }
    __device__ void finish(const Ctx &c, int s_, const double (&)[1]) const { ta.commit(c, s_); }
    __device__ Pre pre(int i, double (&)[1]) const {
        s[i] = r[i] + ta.nalpha * a[i];
        return {};
    }
    __device__ double gather(int) const { return 0.0; }
    __device__ double stage(int, const Pre &) const { return 0.0; }
    __device__ void row(int, double, const Pre &, double (&)[1]) const {}
};

template <class RT>
struct OpSpmvOmega {  // K4: t = A z; (t, s), (t, t), (r*, s), (r*, t) published for the next K1
    static constexpr int NRED = 4;
    static constexpr int kPublish = 4, kSlot = kSlotOmega;
    struct Pre {
        double s, rs;
    };
    const double *z, *s;
    const RT *rs;
    double *t;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int i, double (&)[4]) const { return {s[i], (double)rs[i]}; }
    __device__ double gather(int j) const { return z[j]; }
    __device__ double stage(int i, const Pre &) const { return z[i]; }
    __device__ void row(int i, double ax, const Pre &pr, double (&acc)[4]) const {
        t[i] = ax;
        acc[0] += ax * pr.s;
        acc[1] += ax * ax;
        acc[2] += pr.rs * pr.s;
        acc[3] += pr.rs * ax;
    }
};

struct OpTrueResidual {  // ||b - A x||^2 per system
    static constexpr int NRED = 1;
    static constexpr bool kTree = true;
    struct Pre {
        double b;
    };
    const double *x, *b;
    __device__ bool enabled(const SysState &) const { return true; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int i, double (&)[1]) const { return {b[i]}; }
    __device__ double gather(int j) const { return x[j]; }
    __device__ double stage(int i, const Pre &) const { return x[i]; }
    __device__ void row(int, double ax, const Pre &pr, double (&acc)[1]) const {
        double d = pr.b - ax;
        acc[0] += d * d;
    }
    __device__ void finish(const Ctx &c, int s, const double (&tot)[1]) const { c.st[s].true_rr = tot[0]; }
};

struct OpSpmvPlain {  // y = A x (timing of the bare product)
    static constexpr int NRED = 0;
    struct Pre {};
    const double *x;
    double *y;
    __device__ bool enabled(const SysState &) const { return true; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int, double (&)[1]) const { return {}; }
    __device__ double gather(int j) const { return x[j]; }
    __device__ double stage(int i, const Pre &) const { return x[i]; }
    __device__ void row(int i, double ax, const Pre &, double (&)[1]) const { y[i] = ax; }
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

struct OpCopyS {  // K3 without a preconditioner: z = s
    static constexpr int NRED = 0;
    static constexpr int kTake = 1, kTakeSlot = kSlotRsA;
    static constexpr bool kCommit = true;
    struct Pre {
        double s;
    };
    const double *r, *a;
    double *s, *z;
    TakeAlpha ta;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) { ta.load_rho(S); }
    __device__ bool need_take() const { return true; }
    __device__ void take(const double (&v)[1]) { ta.take(v); }
    __device__ void finish(const Ctx &c, int s_, const double (&)[1]) const { ta.commit(c, s_); }
    __device__ Pre pre(int i, double (&)[1]) const {
        const double si = r[i] + ta.nalpha * a[i];
        s[i] = si;
        return {si};
    }
    __device__ double gather(int) const { return 0.0; }
    __device__ double stage(int, const Pre &) const { return 0.0; }
    __device__ void row(int i, double, const Pre &pr, double (&)[1]) const { z[i] = pr.s; }
};

// ncounters: the arrival counters after the two counts (tree nodes, groups)
__global__ void k_state_init(SysState *st, int nsys, int max_iter, int *counter, int ncounters) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) {
        SysState z{};
        z.active = 1;
        z.status = kRunning;
        z.max_iter = max_iter;
        st[s] = z;
        counter[s] = 0;
    }
    if (s == 0) {
        counter[nsys] = nsys;
        counter[nsys + 1] = 0;
    }
    for (int b = s; b < ncounters; b += gridDim.x * blockDim.x) counter[nsys + 2 + b] = 0;
}

// systems whose recursive residual converged but whose true residual did not
// are re-activated for a restart
__global__ void k_check_true(SysState *st, int nsys, int *counter) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    SysState &S = st[s];
    if ((S.status == kConverged || S.status == kConvergedAtStart) && !(sqrt(S.true_rr) < S.tol) &&
        S.iters < S.max_iter - 1 && S.restarts < 4) {
        S.status = kRunning;
        S.active = 1;
        atomicAdd(&counter[nsys], 1);
    }
}

// re-enables every system (scalars kept) so single kernels can be timed
__global__ void k_state_activate(SysState *st, int nsys, int *counter, int ncounters) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) {
        st[s].active = 1;
        counter[s] = 0;
    }
    for (int b = s; b < ncounters; b += gridDim.x * blockDim.x) counter[nsys + 2 + b] = 0;
}

// Single-kernel timing: the iteration partials all 1.0 and one K2 counted,
// so that every K1 forms its scalars from K4's partials as in a solve (finite
// values: omega = 1, rho' = 0, beta = 0)
__global__ void k_seed_partials(SysState *st, int nsys, double *gpart, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsys) st[i].k2_count = 1;
    if (i < n) gpart[i] = 1.0;
}

// The graph loop's host check: active systems and active units into
// host-coherent memory (system-scope stores), read by the host once an event
// after the graph has completed.
__device__ __forceinline__ void publish_counts(int *pub, int systems, int units) {
    __hip_atomic_store(&pub[0], systems, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&pub[1], units, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_publish_counts(const int *counter2, int *pub) {
    if (threadIdx.x == 0) publish_counts(pub, counter2[0], counter2[1]);
}

// Active-unit list, two launches.  k_compact_scan (one workgroup): base[s] =
// number of units of active systems before s, the total into counter2[1]
// (counter2[0] holds the active systems) and, with pub, both into host memory.
// k_compact_fill: every unit of an active system takes position base[s] + its
// rank inside the system of the layout order, and writes its id where
// xcd_workgroup puts that position.
constexpr int kCompactBlock = 1024;
__global__ __launch_bounds__(kCompactBlock) void k_compact_scan(const SysState *st, const int *sys_units, int nsys,
                                                                int *base, int *counter2, int *pub) {
    __shared__ int sh[kCompactBlock];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nsys; b0 += kCompactBlock) {
        const int s = b0 + (int)threadIdx.x;
        const int v = (s < nsys && st[s].active) ? sys_units[s + 1] - sys_units[s] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kCompactBlock; o <<= 1) {
            const int t = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (s < nsys) base[s] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += sh[kCompactBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counter2[1] = carry;
        if (pub) publish_counts(pub, counter2[0], carry);
    }
}

__global__ void k_compact_fill(const SysState *st, const int *unit_sys, const int *sys_units, const int *base,
                               const int *count, int nunits, int *list) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const int s = unit_sys[u];
    if (!st[s].active) return;
    const int rank = base[s] + u - sys_units[s];
    list[xcd_workgroup(rank, *count)] = u;
}

// Both steps in one workgroup when the batch has at most kCompactBlock
// systems (one launch per host check instead of two): the scan in LDS, then
// the workgroup walks every unit.
__global__ __launch_bounds__(kCompactBlock) void k_compact_small(const SysState *st, const int *unit_sys,
                                                                 const int *sys_units, int nsys, int nunits,
                                                                 int *counter2, int *list, int *pub) {
    __shared__ int sh[kCompactBlock];
    __shared__ int base[kCompactBlock];
    const int s = (int)threadIdx.x;
    const int v = (s < nsys && st[s].active) ? sys_units[s + 1] - sys_units[s] : 0;
    sh[s] = v;
    __syncthreads();
    for (int o = 1; o < kCompactBlock; o <<= 1) {
        const int t = s >= o ? sh[s - o] : 0;
        __syncthreads();
        sh[s] += t;
        __syncthreads();
    }
    base[s] = v > 0 ? sh[s] - v : -1;  // -1: inactive (or no units)
    const int total = sh[kCompactBlock - 1];
    __syncthreads();
    for (int u = s; u < nunits; u += kCompactBlock) {
        const int q = unit_sys[u];
        const int b = base[q];
        if (b < 0) continue;
        const int rank = b + u - sys_units[q];
        list[xcd_workgroup(rank, total)] = u;
    }
    if (s == 0) {
        counter2[1] = total;
        if (pub) publish_counts(pub, counter2[0], total);
    }
}

// SELL-64 construction from CSR: one thread per (slice, lane).  Lane l of
// slice sl holds row slice_row[sl] + l; its entries go to column k of the
// slice, padding entries are (row, 0) - or (first row of the slice, 0) for a
// lane past the system's last row - so every padded gather stays in bounds.
// With sell_off set, each entry's column is also stored as a 16-bit offset
// from the lane's row (slice_row[sl] + l, past the system's end too); *wide
// is set when an offset does not fit.
__global__ void k_sell_fill(int nslices, const int *slice_row, const int *slice_blk, const int *rowptr,
                            const int *colidx, int *sell_col, short *sell_off, int *wide, int *sell_pos) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nslices * kSliceRows) return;
    const int sl = (int)(t / kSliceRows), lane = (int)(t % kSliceRows);
    const int r0 = slice_row[sl], row = r0 + lane;
    const bool live = row < slice_row[sl + 1];
    const int w = slice_blk[sl + 1] - slice_blk[sl];
    const long long base = (long long)slice_blk[sl] * kSliceRows + lane;
    const int p0 = live ? rowptr[row] : 0, len = live ? rowptr[row + 1] - p0 : 0;
    bool fits = true;
    for (int k = 0; k < w; ++k) {
        const long long q = base + (long long)k * kSliceRows;
        int c;
        if (k < len) {
            c = colidx[p0 + k];
            sell_pos[p0 + k] = (int)q;
        } else {
            c = live ? row : r0;
        }
        sell_col[q] = c;
        if (sell_off) {
            const int d = c - row;
            const bool ok = d >= -32768 && d <= 32767;
            sell_off[q] = ok ? (short)d : (short)0;
            fits = fits && ok;
        }
    }
    if (!fits) atomicOr(wide, 1);
}

template <class MT_OUT, class MT_IN>
__global__ void k_scatter_values(long long nnz, const int *pos, const MT_IN *in, MT_OUT *out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) out[pos[p]] = (MT_OUT)in[p];
}

template <class MT_OUT, class MT_IN>
__global__ void k_convert_values(long long nnz, const MT_IN *in, MT_OUT *out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) out[p] = (MT_OUT)in[p];
}

template <class T>
__global__ void k_fill_zero(T *p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (T)0;
}

__global__ void k_fill_double(double *p, double v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_compose_perm(int nnz, const int *perm_csr, const int *sell_pos, int *out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) out[p] = sell_pos[perm_csr[p]];
}

__global__ void k_adjoint_scatter(int n, const int *colptr, const int *rowidx, const double *lambda, const double *x,
                                  double *dA) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double xc = x[c];
    for (int p = colptr[c]; p < colptr[c + 1]; ++p) dA[p] = -lambda[rowidx[p]] * xc;
}

__global__ void k_f32_to_f64(const float *in, double *out, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

int split_units(const TileMap &tm) { return tm.kind == kLayoutSell ? kSplitUnits : kSplitUnitsTiles; }

UnitView unit_view(const TileMap &tm, bool use_list, const int *count) {
    return UnitView{tm.unit_row.get(),
                    tm.unit_sys.get(),
                    tm.unit_slice.get(),
                    tm.slice_row.get(),
                    tm.slice_blk.get(),
                    use_list ? tm.active.get() : nullptr,
                    count,
                    tm.sys_row.get(),
                    0};
}


// One launch of op over the layout: every unit (grid = nunits, list unused)
// or the first `grid` entries of the active list; `reverse` walks each XCD's
// range of units from its end.
template <bool HM, class MT, class Op>
void launch_units(const OperatorView &O, const TileMap &tm, const Ctx &c, const Op &op, const int *list, int grid,
                  hipStream_t st, bool reverse) {
    UnitView uv = unit_view(tm, list != nullptr, c.counter + c.nsys + 1);
    uv.reverse = reverse ? 1 : 0;
    if (O.blk) uv.slice_blk = O.blk;  // an operator with its own pattern
    const int g = list ? grid : tm.nunits;
    if (g <= 0) return;
    if (tm.kind == kLayoutSell && O.diag)
        hipLaunchKernelGGL((k_sell<HM, MT, Op, kIdxDiag>), dim3(g), dim3(kSellThreads), 0, st, O.diag, (const MT *)O.val,
                           uv, c, op);
    else if (tm.kind == kLayoutSell && O.off)
        hipLaunchKernelGGL((k_sell<HM, MT, Op, kIdx16>), dim3(g), dim3(kSellThreads), 0, st, O.off, (const MT *)O.val, uv,
                           c, op);
    else if (tm.kind == kLayoutSell)
        hipLaunchKernelGGL((k_sell<HM, MT, Op, kIdx32>), dim3(g), dim3(kSellThreads), 0, st, O.idx, (const MT *)O.val, uv,
                           c, op);
    else
        hipLaunchKernelGGL((k_tiles<HM, MT, Op>), dim3(g), dim3(kThreads), 0, st, O.ptr, O.idx, (const MT *)O.val,
                           uv, c, op);
}

// Operands of a K1 launch (see OpMapplyP).
struct K1Args {
    double *x, *r, *pout, *y;
    const double *s, *t, *pin, *a, *z, *yb, *zb;
};

// One K1 launch with the x update `xmode` selected at compile time.
template <bool HM, class MT, int YMODE>
void launch_k1(const OperatorView &O, const TileMap &tm, const Ctx &c, int xmode, const K1Args &k, const int *list,
               int grid, hipStream_t st, bool reverse) {
    auto go = [&](auto xtag) {
        constexpr int XM = decltype(xtag)::value;
        OpMapplyP<MT, YMODE, XM> op{k.x, k.r, k.pout, k.y, k.s, k.t, k.pin, k.a, k.z, k.yb, k.zb,
                                    0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
        launch_units<HM, MT>(O, tm, c, op, list, grid, st, reverse);
    };
    if (xmode == kXOne) go(std::integral_constant<int, kXOne>{});
    else if (xmode == kXDefer) go(std::integral_constant<int, kXDefer>{});
    else go(std::integral_constant<int, kXTwo>{});
}

// Column-blocked product of one op (k_tiles_part) into part, and the launch
// of the op that consumes the partials (k_tiles, PS = kColBlocks).
template <class MT, class Op>
void launch_part(const TileMap &tm, const Ctx &c, const Op &op, const MT *val, double *part, int n, hipStream_t st) {
    constexpr int per = 8 / kColBlocks;  // XCDs per block
    // persistent: kPartWgPerCu workgroups on each of an XCD's 32 CUs, fewer
    // when the largest block has fewer tiles
    const int per_xcd = std::min(kPartPipe ? 32 * kPartWgPerCu : INT32_MAX, (tm.cb_units_max + per - 1) / per);
    PartTiles pt;
    pt.tile = reinterpret_cast<const int4 *>(tm.cb_tile.get());
    for (int k = 0; k <= kColBlocks; ++k) pt.off[k] = tm.cb_toff[k];
    hipLaunchKernelGGL((k_tiles_part<MT, Op, kColBlocks>), dim3(8 * per_xcd), dim3(kThreads), 0, st,
                       tm.cb_ptr.get(), tm.cb_idx.get(), val, pt, n, part, c, op);
}
template <class Op>
void launch_summed(const TileMap &tm, const Ctx &c, const Op &op, const double *part, int n, hipStream_t st,
                   bool reverse) {
    UnitView uv = unit_view(tm, false, c.counter + c.nsys + 1);
    uv.reverse = reverse ? 1 : 0;
    uv.psum = part;
    uv.pstride = n;
    hipLaunchKernelGGL((k_tiles<false, double, Op, kColBlocks>), dim3(tm.nunits), dim3(kThreads), 0, st,
                       (const int *)nullptr, (const int *)nullptr, (const double *)nullptr, uv, c, op);
}

// f(RT{}) with RT = float when r* is stored in FP32, else double.
template <class F>
void with_rt(bool rs32, F &&f) {
    if (rs32) f(float{});
    else f(double{});
}

}  // namespace

// ================================================================ host ====
void csc_to_csr_device(int m, int n, int nnz, const int *colptr, const int *rowidx, const double *cscval,
                       int *rowptr, int *colidx, double *csrval, int *perm, hipStream_t st) {
    require(m >= 0 && n >= 0 && nnz >= 0, "csc_to_csr: negative size");
    DeviceArray<int> cnt((size_t)m + 1), src((size_t)std::max(nnz, 1)), tmp;
    const int B = 256;
    hipLaunchKernelGGL(k_fill_int, dim3(std::max(1, ceil_div(m + 1, B))), dim3(B), 0, st, cnt.get(), 0,
                       (long long)m + 1);
    if (nnz > 0) hipLaunchKernelGGL(k_count_rows, dim3(ceil_div(nnz, B)), dim3(B), 0, st, rowidx, nnz, cnt.get());
    exclusive_scan(cnt.get(), rowptr, m, tmp, st);
    hipLaunchKernelGGL(k_fill_int, dim3(std::max(1, ceil_div(m + 1, B))), dim3(B), 0, st, cnt.get(), 0,
                       (long long)m + 1);
    if (n > 0)  // one thread per column appends its entries to their rows
        hipLaunchKernelGGL(k_fill_rows, dim3(ceil_div(n, B)), dim3(B), 0, st, colptr, rowidx, n, rowptr, cnt.get(),
                           colidx, src.get());
    if (m > 0)  // one thread per row orders it by column
        hipLaunchKernelGGL(k_sort_rows, dim3(ceil_div(m, B)), dim3(B), 0, st, m, rowptr, colidx, src.get(), cscval,
                           csrval, perm);
    SPAI_CHECK_LAUNCH();
    SPAI_HIP(hipStreamSynchronize(st));  // temporaries are freed on return
}

// Most compact SELL-64 column storage allowed: the environment variable
// SPAI_MI355X_SELL_INDEX = 32 or 16 caps it (32-bit columns, or 16-bit
// offsets at most); unset allows diagonal slices.  Every form gives
// bit-identical results.
static int sell_index_cap() {
    const char *env = std::getenv("SPAI_MI355X_SELL_INDEX");
    if (env && std::strcmp(env, "32") == 0) return kIdx32;
    if (env && std::strcmp(env, "16") == 0) return kIdx16;
    return kIdxDiag;
}

// Diagonal slices: the distinct column - row offsets of each slice's rows,
// ascending, become its column blocks, so a row's entries keep their CSR
// (ascending column) order and the gaps hold 0.0.  Taken when they add at
// most 10% to the blocks of the row-length padding (stencil matrices add
// none).  On success dblk (nslices + 1 block offsets), doff (one offset per
// block) and pos (CSR position -> layout position) are filled.
static bool build_diagonal_slices(const std::vector<int> &srow, const std::vector<int> &hp,
                                  const std::vector<int> &hidx, long long sell_blocks, std::vector<int> &dblk,
                                  std::vector<int> &doff, std::vector<int> &pos) {
    const int nslices = (int)srow.size() - 1;  // srow: nslices + 1 row boundaries
    dblk.assign((size_t)nslices + 1, 0);
    doff.clear();
    std::vector<int> d;
    long long blocks = 0;
    for (int k = 0; k < nslices; ++k) {
        const int r0 = srow[(size_t)k], r1 = srow[(size_t)k + 1];
        d.clear();
        for (int i = r0; i < r1; ++i)
            for (int p = hp[(size_t)i]; p < hp[(size_t)i + 1]; ++p) d.push_back(hidx[(size_t)p] - i);
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());
        blocks += (long long)d.size();
        if (blocks * 10 > sell_blocks * 11) return false;
        dblk[(size_t)k + 1] = (int)blocks;
        doff.insert(doff.end(), d.begin(), d.end());
    }
    if (blocks * kSliceRows >= (long long)INT32_MAX) return false;
    pos.assign(hidx.size(), 0);
    for (int k = 0; k < nslices; ++k) {
        const int r0 = srow[(size_t)k], r1 = srow[(size_t)k + 1];
        const int *o0 = doff.data() + dblk[(size_t)k], *o1 = doff.data() + dblk[(size_t)k + 1];
        for (int i = r0; i < r1; ++i)
            for (int p = hp[(size_t)i]; p < hp[(size_t)i + 1]; ++p) {
                const int slot = (int)(std::lower_bound(o0, o1, hidx[(size_t)p] - i) - o0);
                pos[(size_t)p] = (dblk[(size_t)k] + slot) * kSliceRows + (i - r0);
            }
    }
    return true;
}

// SELL-64 column indices (and positions) of an operator on the layout's
// slices; the 16-bit offsets are kept only when every entry's fits.
#ifndef SPAI_SELL_OFFSETS
#define SPAI_SELL_OFFSETS 1
#endif
static void fill_sell(int nslices, const int *slice_row, const int *slice_blk, const int *rowptr, const int *colidx,
                      long long len, DeviceArray<int> &col, DeviceArray<short> &off, int *pos, hipStream_t st) {
    col.resize((size_t)std::max<long long>(len, 1));
    const bool want = SPAI_SELL_OFFSETS != 0 && len > 0 && sell_index_cap() != kIdx32;
    if (want) off.resize((size_t)len);
    else off.release();
    DeviceArray<int> wide(1);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, wide.get(), 0, 1LL);
    const long long threads = (long long)nslices * kSliceRows;
    if (threads > 0)
        hipLaunchKernelGGL(k_sell_fill, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, nslices, slice_row,
                           slice_blk, rowptr, colidx, col.get(), want ? off.get() : nullptr, wide.get(), pos);
    SPAI_CHECK_LAUNCH();
    int h_wide = 0;
    SPAI_HIP(hipMemcpyAsync(&h_wide, wide.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipStreamSynchronize(st));
    if (h_wide) off.release();
}

// Row tiles over the rows [r0, r1) of one operator given by its row lengths
// (hp: prefix, n + 1): at most `target` entries and max_rows rows per tile,
// a row longer than the staging buffer alone.
static void cut_tiles(const std::vector<int> &hp, int r0, int r1, int target, std::vector<int> &urow,
                      int max_rows = kRowsPerTile, int cap = kLdsProducts) {
    int i = r0;
    while (i < r1) {
        const int start = i;
        long long nz = 0;
        if (hp[(size_t)i + 1] - hp[(size_t)i] > cap) {
            ++i;  // a row longer than the staging buffer gets a tile of its own
        } else {
            while (i < r1 && i - start < max_rows) {
                const int len = hp[(size_t)i + 1] - hp[(size_t)i];
                if (len > cap || (i > start && nz + len > target)) break;
                nz += len;
                ++i;
            }
        }
        urow.push_back(start);
    }
}

static bool column_blocks_enabled() {
    const char *env = std::getenv("SPAI_MI355X_COLUMN_BLOCKS");
    return !(env && std::atoi(env) == 0);
}

// Column blocks of a one-system row-tile layout (see TileMap): entries
// reordered block by block on the host, each block cut into tiles of its own.
static void build_column_blocks(TileMap &tm, int n, const std::vector<int> &hp, const int *idx, long long total,
                                int target, hipStream_t st) {
    constexpr int C = kColBlocks;
    std::vector<int> hidx((size_t)std::max<long long>(total, 1));
    if (total > 0)
        SPAI_HIP(hipMemcpyAsync(hidx.data(), idx, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipStreamSynchronize(st));
    const long long width = ((long long)n + C - 1) / C;  // columns per block
    auto block_of = [&](int col) { return (int)std::min<long long>(col / width, C - 1); };
    // counts per (block, row), block-major, then their prefix
    std::vector<int> bp((size_t)C * n + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int p = hp[(size_t)i]; p < hp[(size_t)i + 1]; ++p) ++bp[(size_t)block_of(hidx[(size_t)p]) * n + i + 1];
    for (size_t q = 1; q < bp.size(); ++q) bp[q] += bp[q - 1];
    std::vector<int> bidx((size_t)std::max<long long>(total, 1)), bpos((size_t)std::max<long long>(total, 1));
    {
        std::vector<int> cur(bp.begin(), bp.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int p = hp[(size_t)i]; p < hp[(size_t)i + 1]; ++p) {
                const int q = cur[(size_t)block_of(hidx[(size_t)p]) * n + i]++;
                bidx[(size_t)q] = hidx[(size_t)p];
                bpos[(size_t)p] = q;
            }
    }
    // each block's tiles, on the block's own row lengths, as descriptors
    // (first row, end row, first entry, end entry in the block numbering)
    std::vector<int> tiles;
    int umax = 0;
    tm.cb_toff[0] = 0;
    for (int c = 0; c < C; ++c) {
        std::vector<int> lens((size_t)n + 1), urow;
        for (int i = 0; i <= n; ++i) lens[(size_t)i] = bp[(size_t)c * n + i] - bp[(size_t)c * n];
        cut_tiles(lens, 0, n, kPartProducts, urow, kRowsPerPartTile, kPartProducts);
        urow.push_back(n);
        const int nt = (int)urow.size() - 1;
        umax = std::max(umax, nt);
        for (int k = 0; k < nt; ++k) {
            const int r0 = urow[(size_t)k], r1 = urow[(size_t)k + 1];
            tiles.insert(tiles.end(), {r0, r1, bp[(size_t)c * n + r0], bp[(size_t)c * n + r1]});
        }
        tm.cb_toff[c + 1] = tm.cb_toff[c] + nt;
    }
    auto up = [&](DeviceArray<int> &d, const std::vector<int> &h) {
        d.resize(std::max<size_t>(h.size(), 1));
        SPAI_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, st));
    };
    up(tm.cb_ptr, bp);
    up(tm.cb_idx, bidx);
    up(tm.cb_pos, bpos);
    up(tm.cb_tile, tiles);
    SPAI_HIP(hipStreamSynchronize(st));  // the host vectors end with this function
    tm.cblocks = C;
    tm.cb_units_max = umax;
    tm.cb_nnz = total;
}

void build_layout(TileMap &tm, const int *ptr, const int *idx, const std::vector<int> &h_row_off,
                  const std::vector<long long> &h_nnz_off, bool allow_sell, hipStream_t st, const int *ptr2,
                  bool allow_blocks) {
    const int nsys = (int)h_row_off.size() - 1;
    require(nsys >= 1, "build_layout: no systems");
    static std::atomic<unsigned long long> generations{0};
    tm.generation = ++generations;
    const int n = h_row_off[(size_t)nsys];
    std::vector<int> hp((size_t)n + 1);
    SPAI_HIP(hipMemcpyAsync(hp.data(), ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipStreamSynchronize(st));
    if (ptr2) {
        // tiles are cut on the longer of the two operators' rows: hp becomes
        // the prefix of max(len_1, len_2)
        std::vector<int> h2((size_t)n + 1);
        SPAI_HIP(hipMemcpyAsync(h2.data(), ptr2, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
        std::vector<int> m((size_t)n + 1, 0);
        for (int i = 0; i < n; ++i)
            m[(size_t)i + 1] = m[(size_t)i] + std::max(hp[(size_t)i + 1] - hp[(size_t)i], h2[(size_t)i + 1] - h2[(size_t)i]);
        hp.swap(m);
    }
    const long long total = h_nnz_off[(size_t)nsys];
    // SELL-64 candidate: slices of 64 rows inside each system
    std::vector<int> srow, sblk(1, 0), sys_slices((size_t)nsys + 1, 0);
    bool sell = false;
    if (allow_sell && total > 0) {
        long long blocks = 0;
        for (int s = 0; s < nsys; ++s) {
            const int r0 = h_row_off[(size_t)s], r1 = h_row_off[(size_t)s + 1];
            for (int r = r0; r < r1; r += kSliceRows) {
                const int e = std::min(r + kSliceRows, r1);
                int w = 0;
                for (int i = r; i < e; ++i) w = std::max(w, hp[(size_t)i + 1] - hp[(size_t)i]);
                srow.push_back(r);
                blocks += w;
                sblk.push_back((int)std::min<long long>(blocks, INT32_MAX));
            }
            sys_slices[(size_t)s + 1] = (int)srow.size();
        }
        const long long padded = blocks * kSliceRows;
        // at most 20% padding, and SELL positions must fit in int32
        sell = padded * 5 <= total * 6 && padded < (long long)INT32_MAX;
    }
    std::vector<int> urow, usys, uslice;
    tm.h_sys_units.assign((size_t)nsys + 1, 0);
    if (sell) {
        tm.kind = kLayoutSell;
        tm.target_nnz = 0;
        tm.nslices = (int)srow.size();
        tm.sell_len = (long long)sblk.back() * kSliceRows;
        for (int s = 0; s < nsys; ++s) {
            const int k0 = sys_slices[(size_t)s], k1 = sys_slices[(size_t)s + 1];
            if (k0 == k1) {  // empty system: one empty unit keeps the reductions uniform
                urow.push_back(h_row_off[(size_t)s]);
                usys.push_back(s);
                uslice.push_back(k0);
            }
            for (int k = k0; k < k1; k += kSlicesPerUnit) {
                urow.push_back(srow[(size_t)k]);
                usys.push_back(s);
                uslice.push_back(k);
            }
            tm.h_sys_units[(size_t)s + 1] = (int)usys.size();
        }
        srow.push_back(n);
        uslice.push_back(tm.nslices);
    } else {
        tm.kind = kLayoutTiles;
        tm.nslices = 0;
        tm.sell_len = 0;
        // nonzeros per tile: enough tiles to fill the GPU on large inputs, at
        // most the LDS staging buffer; rows per tile at most kRowsPerTile
        int target = 256;
        while (target < kLdsProducts && (long long)target * 4096 < total) target <<= 1;
        tm.target_nnz = target;
        for (int s = 0; s < nsys; ++s) {
            const int r0 = h_row_off[(size_t)s], r1 = h_row_off[(size_t)s + 1];
            if (r0 == r1) urow.push_back(r0);  // empty system: one empty tile keeps the reductions uniform
            cut_tiles(hp, r0, r1, target, urow);
            usys.resize(urow.size(), s);
            tm.h_sys_units[(size_t)s + 1] = (int)usys.size();
        }
        uslice.assign(usys.size() + 1, 0);
    }
    urow.push_back(n);
    tm.nunits = (int)usys.size();
    // reduction tree per system: its units, then ceil(n / kFanIn) nodes per
    // level until one remains
    std::vector<int> snodes((size_t)nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        int n = tm.h_sys_units[(size_t)s + 1] - tm.h_sys_units[(size_t)s], total = n;
        while (n > 1) {
            n = (n + kFanIn - 1) / kFanIn;
            total += n;
        }
        snodes[(size_t)s + 1] = snodes[(size_t)s] + total;
    }
    tm.nnodes = snodes[(size_t)nsys];
    // partial-sum groups of the iteration kernels
    std::vector<int> sgroups((size_t)nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        const int nu = tm.h_sys_units[(size_t)s + 1] - tm.h_sys_units[(size_t)s];
        const int G = group_units(nu);
        sgroups[(size_t)s + 1] = sgroups[(size_t)s] + (nu + G - 1) / G;
    }
    tm.ngroups = sgroups[(size_t)nsys];
    auto up = [&](DeviceArray<int> &d, const std::vector<int> &h) {
        d.resize(std::max<size_t>(h.size(), 1));
        if (!h.empty()) SPAI_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, st));
    };
    up(tm.unit_row, urow);
    up(tm.sys_row, h_row_off);
    up(tm.unit_sys, usys);
    up(tm.sys_units, tm.h_sys_units);
    up(tm.unit_slice, uslice);
    up(tm.sys_nodes, snodes);
    up(tm.sys_groups, sgroups);
    tm.active.resize((size_t)std::max(tm.nunits, 1));
    tm.sys_base.resize((size_t)nsys);
    if (sell) {
        up(tm.slice_row, srow);
        std::vector<int> dblk, doff, hpos;
        bool diag = false;
        if (sell_index_cap() == kIdxDiag && !ptr2) {
            std::vector<int> hidx((size_t)total);
            SPAI_HIP(hipMemcpyAsync(hidx.data(), idx, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, st));
            SPAI_HIP(hipStreamSynchronize(st));
            diag = build_diagonal_slices(srow, hp, hidx, (long long)sblk.back(), dblk, doff, hpos);
        }
        tm.sell_pos.resize((size_t)std::max<long long>(total, 1));
        if (diag) {
            tm.sell_len = (long long)dblk.back() * kSliceRows;
            up(tm.slice_blk, dblk);
            up(tm.sell_diag, doff);
            up(tm.sell_pos, hpos);
            SPAI_HIP(hipStreamSynchronize(st));  // the host vectors end with this block
            tm.sell_col.release();
            tm.sell_off.release();
        } else {
            tm.sell_diag.release();
            up(tm.slice_blk, sblk);
            fill_sell(tm.nslices, tm.slice_row.get(), tm.slice_blk.get(), ptr, idx, tm.sell_len, tm.sell_col,
                      tm.sell_off, tm.sell_pos.get(), st);
        }
    } else {
        tm.slice_row.resize(1);
        tm.slice_blk.resize(1);
    }
    tm.cblocks = 0;
    tm.cb_units_max = 0;
    if (!sell && allow_blocks && !ptr2 && nsys == 1 && n >= kColBlockMinRows && column_blocks_enabled())
        build_column_blocks(tm, n, hp, idx, total, tm.target_nnz, st);
    else {
        tm.cb_ptr.release();
        tm.cb_idx.release();
        tm.cb_pos.release();
        tm.cb_tile.release();
    }
    SPAI_HIP(hipStreamSynchronize(st));
}

template <class MT_OUT, class MT_IN>
void layout_values(const TileMap &tm, long long nnz, const MT_IN *in, MT_OUT *out, hipStream_t st) {
    if (tm.kind == kLayoutSell) {
        hipLaunchKernelGGL(k_fill_zero<MT_OUT>, dim3((unsigned)((tm.sell_len + 255) / 256)), dim3(256), 0, st, out,
                           tm.sell_len);
        if (nnz > 0)
            hipLaunchKernelGGL((k_scatter_values<MT_OUT, MT_IN>), dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0,
                               st, nnz, tm.sell_pos.get(), in, out);
    } else if (nnz > 0) {
        hipLaunchKernelGGL((k_convert_values<MT_OUT, MT_IN>), dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st,
                           nnz, in, out);
    }
    SPAI_CHECK_LAUNCH();
}
template void layout_values<double, double>(const TileMap &, long long, const double *, double *, hipStream_t);
template void layout_values<float, double>(const TileMap &, long long, const double *, float *, hipStream_t);
template void layout_values<float, float>(const TileMap &, long long, const float *, float *, hipStream_t);

void build_sell_operator(const TileMap &tm, const int *rowptr, const int *colidx, int nnz, SellOperator &op,
                         hipStream_t st) {
    require(tm.kind == kLayoutSell, "build_sell_operator: the layout is not SELL-64");
    const int n = [&] {
        int last = 0;
        SPAI_HIP(hipMemcpyAsync(&last, tm.slice_row.get() + tm.nslices, sizeof(int), hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
        return last;
    }();
    std::vector<int> hp((size_t)n + 1), srow((size_t)tm.nslices + 1), sblk((size_t)tm.nslices + 1, 0);
    SPAI_HIP(hipMemcpyAsync(hp.data(), rowptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipMemcpyAsync(srow.data(), tm.slice_row.get(), sizeof(int) * srow.size(), hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipStreamSynchronize(st));
    long long blocks = 0;
    for (int k = 0; k < tm.nslices; ++k) {
        int w = 0;
        for (int i = srow[(size_t)k]; i < srow[(size_t)k + 1]; ++i) w = std::max(w, hp[(size_t)i + 1] - hp[(size_t)i]);
        blocks += w;
        require(blocks * kSliceRows < (long long)INT32_MAX, "build_sell_operator: more than 2^31 padded entries");
        sblk[(size_t)k + 1] = (int)blocks;
    }
    op.len = blocks * kSliceRows;
    op.slice_blk.resize(sblk.size());
    SPAI_HIP(hipMemcpyAsync(op.slice_blk.get(), sblk.data(), sizeof(int) * sblk.size(), hipMemcpyHostToDevice, st));
    op.pos.resize((size_t)std::max(nnz, 1));
    fill_sell(tm.nslices, tm.slice_row.get(), op.slice_blk.get(), rowptr, colidx, op.len, op.col, op.off,
              op.pos.get(), st);
}

void compose_positions(int nnz, const int *perm_csr, const int *sell_pos, int *perm_out, hipStream_t st) {
    if (nnz <= 0) return;
    hipLaunchKernelGGL(k_compose_perm, dim3(ceil_div(nnz, 256)), dim3(256), 0, st, nnz, perm_csr, sell_pos, perm_out);
    SPAI_CHECK_LAUNCH();
}

void layout_perm(const TileMap &tm, int nnz, const int *perm_csr, int *perm_out, hipStream_t st) {
    if (nnz <= 0) return;
    if (tm.kind == kLayoutSell) {
        hipLaunchKernelGGL(k_compose_perm, dim3(ceil_div(nnz, 256)), dim3(256), 0, st, nnz, perm_csr,
                           tm.sell_pos.get(), perm_out);
        SPAI_CHECK_LAUNCH();
    } else {
        SPAI_HIP(hipMemcpyAsync(perm_out, perm_csr, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice, st));
    }
}

void adjoint_scatter(int n, const int *colptr, const int *rowidx, const double *lambda, const double *x, double *dA,
                     hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_adjoint_scatter, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, colptr, rowidx, lambda, x,
                       dA);
    SPAI_CHECK_LAUNCH();
}

void convert_f32_to_f64(const float *in, double *out, long long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out, n);
    SPAI_CHECK_LAUNCH();
}

BiCGStab::BiCGStab() = default;
BiCGStab::~BiCGStab() {
    drop_graph();
    for (auto &e : check_ev_)
        if (e) (void)hipEventDestroy(e);
    if (h_pub_) (void)hipHostFree(h_pub_);
    if (h_states_pin_) (void)hipHostFree(h_states_pin_);
}

void BiCGStab::drop_graph() {
    for (auto &g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
    for (auto &k : graph_key_) k = nullptr;
    pred_launches_ = 0;
}

void BiCGStab::forget_layout(unsigned long long generation) {
    if (!graphs_.empty() && graph_key_[kGraphKeyGeneration] == (const void *)(size_t)generation) drop_graph();
}

Ctx BiCGStab::make_ctx(const TileMap &tm, int nsys) {
    return Ctx{st_,       tm.sys_units.get(),  tm.sys_nodes.get(),
               cnt_,      part_,               nsys,
               hist_ptr_, hist_stride_,        tm.sys_groups.get(),
               gpart_,    upart_,              cnt_ + nsys + 2 + tm.nnodes,
               tm.ngroups, tm.nunits};
}

namespace {
// the eleven work vectors: r, r*, p (two), y, a, s, z, t, and the second y
// and z of deferred x updates
void carve_vectors(double *v, size_t N, double *&r, double *&rs, double *(&p)[2], double *(&y)[2], double *&a,
                   double *&s, double *(&z)[2], double *&t) {
    r = v;
    rs = v + N;
    p[0] = v + 2 * N;
    p[1] = v + 3 * N;
    y[0] = v + 4 * N;
    a = v + 5 * N;
    s = v + 6 * N;
    z[0] = v + 7 * N;
    t = v + 8 * N;
    y[1] = v + 9 * N;
    z[1] = v + 10 * N;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

namespace {
// doubles of the reduction partials: kMaxRed per tree node, then kSlots per
// group and kSlots per unit
size_t partial_doubles(const TileMap &tm) {
    return (size_t)kMaxRed * (size_t)std::max(tm.nnodes, 1) + (size_t)kSlots * (size_t)std::max(tm.ngroups, 1) +
           (size_t)kSlots * (size_t)std::max(tm.nunits, 1);
}
size_t counter_ints(const TileMap &tm, int nsys) { return (size_t)nsys + 2 + (size_t)tm.nnodes + (size_t)tm.ngroups; }
}  // namespace

size_t BiCGStab::workspace_bytes(int n, int nsys, const TileMap &tm) {
    const size_t N = (size_t)std::max(n, 1);
    // nine vectors (no second y and z: a bound workspace never defers x)
    return round256(sizeof(double) * 9 * N) + round256(sizeof(SysState) * (size_t)nsys) +
           round256(sizeof(int) * counter_ints(tm, nsys)) + round256(sizeof(double) * partial_doubles(tm));
}

void BiCGStab::bind_workspace(void *base, size_t bytes) {
    require(base != nullptr && ((uintptr_t)base & 255) == 0, "bind_workspace: base must be 256-byte aligned");
    drop_graph();
    ws_base_ = base;
    ws_bytes_ = bytes;
    n_ = 0;
    nsys_ = 0;
    nnodes_ = -1;
    nunits_ = -1;
    ngroups_ = -1;
}

void BiCGStab::ensure(int n, int nsys, const TileMap &tm) {
    const int nunits = tm.nunits, nnodes = tm.nnodes, ngroups = tm.ngroups;
    auto carve_partials = [&](double *p) {
        part_ = p;
        gpart_ = p + (size_t)kMaxRed * (size_t)std::max(nnodes, 1);
        upart_ = gpart_ + (size_t)kSlots * (size_t)std::max(ngroups, 1);
    };
    if (ws_base_) {
        // a bound workspace: carve, never allocate (the same pointers for
        // the same shape, so nothing is cached across shapes)
        const size_t need = workspace_bytes(n, nsys, tm);
        if (need > ws_bytes_)
            throw Error(kErrUnsupported, "the solve needs " + std::to_string(need) + " bytes of workspace, " +
                                             std::to_string(ws_bytes_) + " are available");
        const size_t N = (size_t)std::max(n, 1);
        char *p = (char *)ws_base_;
        double *v = (double *)p;
        p += round256(sizeof(double) * 9 * N);
        carve_vectors(v, N, r_, rs_, p_, y_, a_, s_, z_, t_);
        y_[1] = y_[0];  // unused: x updates are never deferred here
        z_[1] = z_[0];
        st_ = (SysState *)p;
        p += round256(sizeof(SysState) * (size_t)nsys);
        cnt_ = (int *)p;
        p += round256(sizeof(int) * counter_ints(tm, nsys));
        carve_partials((double *)p);
        n_ = n;
        nsys_ = nsys;
        nnodes_ = nnodes;
        nunits_ = nunits;
        ngroups_ = ngroups;
        return;
    }
    if (n != n_) {
        drop_graph();
        vec_.resize((size_t)11 * std::max(n, 1));
        carve_vectors(vec_.get(), (size_t)std::max(n, 1), r_, rs_, p_, y_, a_, s_, z_, t_);
        n_ = n;
    }
    if (nsys != nsys_ || nnodes != nnodes_ || nunits != nunits_ || ngroups != ngroups_) {
        drop_graph();
        state_.resize((size_t)nsys);
        counter_.resize(counter_ints(tm, nsys));
        partials_.resize(partial_doubles(tm));
        st_ = state_.get();
        cnt_ = counter_.get();
        carve_partials(partials_.get());
        nsys_ = nsys;
        nnodes_ = nnodes;
        nunits_ = nunits;
        ngroups_ = ngroups;
    }
}

void BiCGStab::prepare_blocks(const OperatorView &A, const OperatorView &M, const TileMap &tm, int n, int nsys,
                              hipStream_t st) {
    cb_on_ = tm.kind == kLayoutTiles && tm.cblocks == kColBlocks && nsys == 1 && A.ptr && A.idx && A.val &&
             (!M.val || (M.ptr == A.ptr && M.idx == A.idx));
    if (!cb_on_) return;
    const long long nnz = tm.cb_nnz;
    const unsigned g = (unsigned)((std::max<long long>(nnz, 1) + 255) / 256);
    // grow-only: a captured graph keeps these addresses
    cb_part_.reserve((size_t)kColBlocks * (size_t)n);
    cb_a_.reserve((size_t)std::max<long long>(nnz, 1));
    hipLaunchKernelGGL((k_scatter_values<double, double>), dim3(g), dim3(256), 0, st, nnz, tm.cb_pos.get(),
                       (const double *)A.val, cb_a_.get());
    if (M.val && M.fp32) {
        cb_m32_.reserve((size_t)std::max<long long>(nnz, 1));
        hipLaunchKernelGGL((k_scatter_values<float, float>), dim3(g), dim3(256), 0, st, nnz, tm.cb_pos.get(),
                           (const float *)M.val, cb_m32_.get());
    } else if (M.val) {
        cb_m64_.reserve((size_t)std::max<long long>(nnz, 1));
        hipLaunchKernelGGL((k_scatter_values<double, double>), dim3(g), dim3(256), 0, st, nnz, tm.cb_pos.get(),
                           (const double *)M.val, cb_m64_.get());
    }
    SPAI_CHECK_LAUNCH();
}

// Rebuilds the active-unit list from the system states (stream-ordered).
void BiCGStab::compact(const TileMap &tm, int nsys, hipStream_t st, int *pub) {
    if (nsys <= kCompactBlock) {
        hipLaunchKernelGGL(k_compact_small, dim3(1), dim3(kCompactBlock), 0, st, st_, tm.unit_sys.get(),
                           tm.sys_units.get(), nsys, tm.nunits, cnt_ + nsys, tm.active.get(), pub);
    } else {
        hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(kCompactBlock), 0, st, st_, tm.sys_units.get(),
                           nsys, tm.sys_base.get(), cnt_ + nsys, pub);
        hipLaunchKernelGGL(k_compact_fill, dim3(ceil_div(tm.nunits, 256)), dim3(256), 0, st, st_,
                           tm.unit_sys.get(), tm.sys_units.get(), tm.sys_base.get(), cnt_ + nsys + 1,
                           tm.nunits, tm.active.get());
    }
    SPAI_CHECK_LAUNCH();
}

// K1 of an iteration: fused (one kernel gathering four vectors) or split
// (a streaming update, then y = M p' gathering p' alone).
void BiCGStab::enqueue_k1(const OperatorView &A, const OperatorView &M, const TileMap &tm, int parity,
                          const int *list, int grid, bool split, hipStream_t st) {
    Ctx c = make_ctx(tm, nsys_);
    double *pin = p_[parity], *pout = p_[parity ^ 1];
    const K1Args k1{x_, r_, pout, ybuf(parity), s_, t_, pin, a_, zbuf(parity), ybuf(parity ^ 1), zbuf(parity ^ 1)};
    if (!M.val) {
        // no preconditioner: y = p' through the streaming path
        launch_k1<false, double, kYCopy>(A, tm, c, x_mode(parity), k1, list, grid, st, flip());
    } else if (split) {
        launch_k1<false, double, kYNone>(A, tm, c, x_mode(parity), k1, list, grid, st, flip());
        enqueue_mapply_p(M, tm, parity, list, grid, st);
    } else if (M.fp32) {
        launch_k1<true, float, kYProduct>(M, tm, c, x_mode(parity), k1, list, grid, st, flip());
    } else {
        launch_k1<true, double, kYProduct>(M, tm, c, x_mode(parity), k1, list, grid, st, flip());
    }
}

// Second half of the split K1: y = M p', column-blocked when the solve is.
void BiCGStab::enqueue_mapply_p(const OperatorView &M, const TileMap &tm, int parity, const int *list, int grid,
                                hipStream_t st) {
    Ctx c = make_ctx(tm, nsys_);
    OpApplyM om{p_[parity ^ 1], ybuf(parity)};
    if (cb_on_) {
        if (M.fp32) launch_part<float>(tm, c, om, cb_m32_.get(), cb_part_.get(), n_, st);
        else launch_part<double>(tm, c, om, cb_m64_.get(), cb_part_.get(), n_, st);
        launch_summed(tm, c, om, cb_part_.get(), n_, st, flip());
    } else if (M.fp32) {
        launch_units<true, float>(M, tm, c, om, list, grid, st, flip());
    } else {
        launch_units<true, double>(M, tm, c, om, list, grid, st, flip());
    }
}

// K3: z = M s with s = r - alpha a, fused (s formed while gathering) or, for
// row tiles while K1 is split too, a streaming s pass and an M-apply of the
// stored s (column-blocked when the solve is).
void BiCGStab::enqueue_k3(const OperatorView &M, const TileMap &tm, int parity, const int *list, int grid,
                          bool split, hipStream_t st) {
    Ctx c = make_ctx(tm, nsys_);
    if (split && tm.kind == kLayoutTiles) {
        OpUpdateS us{r_, a_, s_, {}};
        launch_units<false, double>(M, tm, c, us, list, grid, st, flip());
        OpApplyM om{s_, zbuf(parity)};
        if (cb_on_) {
            if (M.fp32) launch_part<float>(tm, c, om, cb_m32_.get(), cb_part_.get(), n_, st);
            else launch_part<double>(tm, c, om, cb_m64_.get(), cb_part_.get(), n_, st);
            launch_summed(tm, c, om, cb_part_.get(), n_, st, flip());
        } else if (M.fp32) {
            launch_units<true, float>(M, tm, c, om, list, grid, st, flip());
        } else {
            launch_units<true, double>(M, tm, c, om, list, grid, st, flip());
        }
    } else if (M.fp32) {
        OpMapplyS<float> op{r_, a_, s_, zbuf(parity), {}};
        launch_units<true, float>(M, tm, c, op, list, grid, st, flip());
    } else {
        OpMapplyS<double> op{r_, a_, s_, zbuf(parity), {}};
        launch_units<true, double>(M, tm, c, op, list, grid, st, flip());
    }
}

// K2: a = A y, (r*, a).
void BiCGStab::enqueue_k2(const OperatorView &A, const TileMap &tm, int parity, const int *list, int grid,
                          hipStream_t st) {
    with_rt(rs32_, [&](auto rt) {
        using RT = decltype(rt);
        Ctx c = make_ctx(tm, nsys_);
        OpSpmvAlpha<RT> op{ybuf(parity), reinterpret_cast<const RT *>(rs_), a_, {}};
        if (cb_on_) {
            launch_part<double>(tm, c, op, cb_a_.get(), cb_part_.get(), n_, st);
            launch_summed(tm, c, op, cb_part_.get(), n_, st, flip());
        } else {
            launch_units<true, double>(A, tm, c, op, list, grid, st, flip());
        }
    });
}

// K4: t = A z, (t, s), (t, t), (r*, s), (r*, t).
void BiCGStab::enqueue_k4(const OperatorView &A, const TileMap &tm, int parity, const int *list, int grid,
                          hipStream_t st) {
    with_rt(rs32_, [&](auto rt) {
        using RT = decltype(rt);
        Ctx c = make_ctx(tm, nsys_);
        OpSpmvOmega<RT> op{zbuf(parity), s_, reinterpret_cast<const RT *>(rs_), t_};
        if (cb_on_) {
            launch_part<double>(tm, c, op, cb_a_.get(), cb_part_.get(), n_, st);
            launch_summed(tm, c, op, cb_part_.get(), n_, st, flip());
        } else {
            launch_units<true, double>(A, tm, c, op, list, grid, st, flip());
        }
    });
}

// OpInit over every unit: a solve's start (restart = 0) or a restart of the
// systems still active from their current x (restart = 1).
void BiCGStab::enqueue_init(const OperatorView &A, const TileMap &tm, const double *x, const double *b,
                            double rel_tol, double abs_tol, int restart, hipStream_t st) {
    with_rt(rs32_, [&](auto rt) {
        using RT = decltype(rt);
        Ctx c = make_ctx(tm, nsys_);
        OpInit<RT> op{x,     b,     r_,    reinterpret_cast<RT *>(rs_), p_[0], p_[1],   a_,      s_,
                      t_,    y_[0], z_[0], y1(),                        z1(),  rel_tol, abs_tol, restart};
        launch_units<true, double>(A, tm, c, op, nullptr, 0, st, flip());
    });
    SPAI_CHECK_LAUNCH();
}

// One iteration: K1 (one or two launches), K2, K3 (one or two), K4.  grid > 0: walk the
// first `grid` entries of the active-unit list; grid == 0: every unit.
void BiCGStab::enqueue_iteration(const OperatorView &A, const OperatorView &M, const TileMap &tm, int parity,
                                 int grid, bool split, hipStream_t st) {
    Ctx c = make_ctx(tm, nsys_);
    const int *list = grid > 0 ? tm.active.get() : nullptr;
    enqueue_k1(A, M, tm, parity, list, grid, split, st);
    enqueue_k2(A, tm, parity, list, grid, st);
    if (M.val) {
        enqueue_k3(M, tm, parity, list, grid, split, st);
    } else {
        OpCopyS op{r_, a_, s_, zbuf(parity), {}};
        launch_units<false, double>(A, tm, c, op, list, grid, st, flip());
    }
    enqueue_k4(A, tm, parity, list, grid, st);
    SPAI_CHECK_LAUNCH();
}

void BiCGStab::solve(const OperatorView &A, const OperatorView &M, const TileMap &tm, int n, int nsys,
                     const double *b, double *x, const SolveOptions &opt, hipStream_t st, SolveResult *res,
                     SysState *states_out) {
    require(n > 0 && nsys >= 1 && (A.idx || A.diag) && A.val && b && x, "bicgstab: bad arguments");
    require(tm.kind == kLayoutSell || A.ptr, "bicgstab: row tiles need a pointer array");
    require(tm.nunits > 0 && (int)tm.h_sys_units.size() == nsys + 1, "bicgstab: layout does not match the systems");
    require(opt.max_iter >= 1, "bicgstab: max_iter must be >= 1");
    // a bound workspace allocates nothing: no graph, no history, no events,
    // no pinned host memory, and the states go to the caller's host array
    const bool bound = ws_base_ != nullptr;
    require(!bound || (!opt.use_graph && !opt.record_history && states_out),
            "bicgstab: a bound workspace supports neither graphs nor histories and needs a states array");
    ensure(n, nsys, tm);
    b_ = b;
    x_ = x;
    defer_x_ = nsys > 1 && !bound;
    rs32_ = M.val && M.fp32;
    if (bound) cb_on_ = false;
    else prepare_blocks(A, M, tm, n, nsys, st);
    if (opt.record_history) {
        hist_.reserve((size_t)nsys * (size_t)opt.max_iter);
        hist_ptr_ = hist_.get();
        hist_stride_ = opt.max_iter;
    } else {
        hist_ptr_ = nullptr;
        hist_stride_ = 0;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    if (!bound) {
        SPAI_HIP(hipEventCreate(&e0));
        SPAI_HIP(hipEventCreate(&e1));
        SPAI_HIP(hipEventRecord(e0, st));
    }
    Ctx c = make_ctx(tm, nsys);
    hipLaunchKernelGGL(k_state_init, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, st_, nsys, opt.max_iter,
                       cnt_, tm.nnodes + tm.ngroups);
    enqueue_init(A, tm, x, b, opt.rel_tol, opt.abs_tol, 0, st);
    // A batch walks the active-unit list, rebuilt after every host check; a
    // single system walks every unit.
    const bool use_list = nsys > 1;
    if (!bound) {
        for (auto &e : check_ev_)
            if (!e) SPAI_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        if (!h_pub_) {
            SPAI_HIP(hipHostMalloc((void **)&h_pub_, 64, hipHostMallocMapped | hipHostMallocCoherent));
            SPAI_HIP(hipHostGetDevicePointer((void **)&d_pub_, h_pub_, 0));
        }
    }
    int active = nsys;
    int units = tm.nunits;
    int checks = 0;
    int local_counts[2] = {0, 0};
    // copied counts (the graphs publish into h_pub_[0..1])
    int *h_counts = bound ? local_counts : h_pub_ + 4;
    auto read_counts = [&]() {
        SPAI_HIP(hipMemcpyAsync(h_counts, cnt_ + nsys, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
        ++checks;
        active = h_counts[0];
        units = use_list ? h_counts[1] : tm.nunits;
    };
    // every system stops by then: the K1 of iteration k finishes iteration k - 1
    const int max_launch_iters = opt.max_iter + 1;
    int done_iters = 0;
    int passes = 0;  // iteration loops run (a restart runs another)
    try {
        // `active` holds nsys, or the count read after the last restart
        // check; the unit count starts at its bound (the list is sized by
        // the compaction itself), so no host round trip precedes the loop
        auto iterate = [&]() {
            if (use_list) compact(tm, nsys, st);
            units = tm.nunits;
            struct PassCount {
                int &p;
                ~PassCount() { ++p; }
            } pass_count{passes};
            if (opt.use_graph) {
                const int gi = std::max(2, opt.graph_iters + (opt.graph_iters & 1));
                const void *key[kGraphKey] = {A.ptr, A.val, A.idx, A.off, A.diag, M.ptr, M.val, M.idx, M.off,
                                              M.diag, M.blk, (const void *)(size_t)M.fp32,
                                              (const void *)(size_t)tm.generation, tm.unit_row.get(), b, x,
                                              hist_ptr_, (const void *)(size_t)defer_x_};
                // One graph runs ahead of the host check: graph k + 1 is
                // queued before the counts graph k published are read, so the
                // device does not wait for the host between graphs.  The
                // counts only fall while iterating, so a grid sized from
                // counts one graph old covers every active unit; a graph
                // launched after the last system stopped finds every unit
                // inactive.  Graphs of other operands are dropped.
                bool same_key = true;
                for (int q = 0; q < kGraphKey; ++q) same_key = same_key && key[q] == graph_key_[q];
                if (!same_key) {
                    drop_graph();
                    for (int q = 0; q < kGraphKey; ++q) graph_key_[q] = key[q];
                }
                // Grid rungs: nunits, then ceil(5/8) of the previous rung; a
                // graph walks the smallest rung holding every active unit, so
                // at most 3/8 of its workgroups find no unit.
                auto rung = [&](int u) {
                    int g = tm.nunits;
                    while (g > 1) {
                        const int next = (g * 5 + 7) / 8;
                        if (next < u || next == g) break;
                        g = next;
                    }
                    return g;
                };
                auto find_graph = [&](int grid, bool split, int giv) -> hipGraphExec_t {
                    for (auto &g : graphs_)
                        if (g.iters == giv && g.grid == grid && g.compact == (int)use_list && g.split == (int)split)
                            return g.exec;
                    return nullptr;
                };
                auto capture_graph = [&](int grid, bool split, int giv) {
                    hipStream_t cap;
                    hipGraph_t graph = nullptr;
                    hipGraphExec_t exec = nullptr;
                    SPAI_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
                    SPAI_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                    for (int k = 0; k < giv; ++k) {
                        enqueue_iteration(A, M, tm, k & 1, grid, split, cap);
                        // long graphs also rebuild the list every kCompactEvery iterations
                        if (use_list && (k + 1) % kCompactEvery == 0 && k + 1 < giv) compact(tm, nsys, cap);
                    }
                    if (use_list) compact(tm, nsys, cap, d_pub_);
                    else
                        hipLaunchKernelGGL(k_publish_counts, dim3(1), dim3(64), 0, cap, cnt_ + nsys,
                                           d_pub_);
                    SPAI_HIP(hipStreamEndCapture(cap, &graph));
                    SPAI_HIP(hipStreamDestroy(cap));
                    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    SPAI_HIP(e);
                    graphs_.push_back(GraphEntry{giv, grid, (int)use_list, (int)split, exec});
                    return exec;
                };
                // A full-length graph is captured together with the shorter
                // even lengths that a planned solve ends with (4, 2), so that
                // a planned solve meets no capture of its own.
                auto graph_for = [&](int grid, bool split, int giv) {
                    if (hipGraphExec_t e = find_graph(grid, split, giv)) return e;
                    hipGraphExec_t e = capture_graph(grid, split, giv);
                    if (giv == gi)
                        for (int g2 = 2; g2 < gi; g2 <<= 1)
                            if (!find_graph(grid, split, g2)) capture_graph(grid, split, g2);
                    return e;
                };
                int slot = 0, pending = 0;
                auto read_published = [&]() {
                    ++checks;
                    active = __atomic_load_n(&h_pub_[0], __ATOMIC_ACQUIRE);
                    units = use_list ? __atomic_load_n(&h_pub_[1], __ATOMIC_ACQUIRE) : tm.nunits;
                };
                // The first loop of a solve whose operands and stop rule match
                // the previous graph-replayed solve's ends with shorter graphs
                // (even lengths keep the iteration parity) sized to the
                // iterations that solve launched, instead of running most of
                // a last full graph after its systems stopped.  Only launch
                // counts change; a wrong guess costs launches, not results.
                // Once the planned iterations are queued, the host waits for
                // them before queueing more (instead of one graph ahead); if
                // systems are still active the plan is dropped.
                int planned = (passes == 0 && pred_launches_ > 0 && pred_max_iter_ == opt.max_iter &&
                               pred_rel_ == opt.rel_tol && pred_abs_ == opt.abs_tol)
                                  ? pred_launches_
                                  : 0;
                try {
                    while (active > 0 && done_iters < max_launch_iters) {
                        if (planned > 0 && done_iters >= planned && pending > 0) {
                            SPAI_HIP(hipEventSynchronize(check_ev_[slot ^ 1]));  // the latest graph's
                            pending = 0;
                            read_published();
                            planned = 0;
                            continue;
                        }
                        const int grid = use_list ? rung(units) : 0;
                        const bool split = M.val && units >= split_units(tm);
                        int giv = gi;
                        if (planned > done_iters && planned - done_iters < gi) {
                            giv = 2;
                            while (giv < planned - done_iters) giv <<= 1;
                            giv = std::min(giv, gi);
                        }
                        SPAI_HIP(hipGraphLaunch(graph_for(grid, split, giv), st));
                        done_iters += giv;
                        SPAI_HIP(hipEventRecord(check_ev_[slot], st));
                        slot ^= 1;
                        if (++pending == 2) {
                            // graph k has published its counts (graph k + 1,
                            // running, may already have published newer ones:
                            // either is valid for sizing graph k + 2)
                            SPAI_HIP(hipEventSynchronize(check_ev_[slot]));
                            --pending;
                            read_published();
                        }
                    }
                } catch (...) {
                    (void)hipStreamSynchronize(st);
                    throw;
                }
                // the graph still in flight (if any) runs before whatever
                // is queued next; its counts are not needed
            } else {
                const int every = std::max(1, opt.check_every);
                int parity = 0;
                while (active > 0 && done_iters < max_launch_iters) {
                    const int grid = use_list ? units : 0;
                    const bool split = M.val && units >= split_units(tm);
                    for (int k = 0; k < every; ++k) {
                        enqueue_iteration(A, M, tm, parity, grid, split, st);
                        parity ^= 1;
                    }
                    if (use_list) compact(tm, nsys, st);
                    done_iters += every;
                    read_counts();
                }
            }
        };
        // x of a system that stopped in a deferring K1 still owes one update
        // (only batches defer)
        auto flush_x = [&]() {
            if (!defer_x_) return;
            OpFlushX op{x, y_[1], z_[1], 0.0, 0.0};
            launch_units<false, double>(A, tm, c, op, nullptr, 0, st, flip());
            hipLaunchKernelGGL(k_clear_xdefer, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, st_, nsys);
            SPAI_CHECK_LAUNCH();
        };
        iterate();
        flush_x();
        bool true_current = false;  // ||b - A x|| is in the states for the final x
        for (int round = 0; opt.true_residual_restart && round < 4; ++round) {
            OpTrueResidual tr{x, b};
            launch_units<true, double>(A, tm, c, tr, nullptr, 0, st, flip());
            hipLaunchKernelGGL(k_check_true, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, st_, nsys,
                               cnt_);
            SPAI_CHECK_LAUNCH();
            read_counts();
            if (active == 0) {
                true_current = true;
                break;
            }
            enqueue_init(A, tm, x, b, opt.rel_tol, opt.abs_tol, 1, st);
            done_iters = 0;
            iterate();
            flush_x();
        }
        if (!true_current) {
            OpTrueResidual op{x, b};
            launch_units<true, double>(A, tm, c, op, nullptr, 0, st, flip());
            SPAI_CHECK_LAUNCH();
        }
        if (!bound) {
            SPAI_HIP(hipEventRecord(e1, st));
            // the states' copy rides behind the end event (one wait for both)
            if (h_states_cap_ < nsys) {
                if (h_states_pin_) SPAI_HIP(hipHostFree(h_states_pin_));
                h_states_pin_ = nullptr;
                h_states_cap_ = 0;
                SPAI_HIP(hipHostMalloc((void **)&h_states_pin_, sizeof(SysState) * nsys, hipHostMallocDefault));
                h_states_cap_ = nsys;
            }
            SPAI_HIP(hipMemcpyAsync(h_states_pin_, st_, sizeof(SysState) * nsys, hipMemcpyDeviceToHost, st));
        }
        SPAI_HIP(hipStreamSynchronize(st));
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    float ms = 0.0f;
    if (!bound) {
        SPAI_HIP(hipEventElapsedTime(&ms, e0, e1));
        SPAI_HIP(hipEventDestroy(e0));
        SPAI_HIP(hipEventDestroy(e1));
    } else {
        // no events in a bound workspace: host time from the first enqueue
        ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    const SysState *states = states_out;
    if (states_out) {
        SPAI_HIP(hipMemcpyAsync(states_out, st_, sizeof(SysState) * nsys, hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
    }
    if (!bound) {
        // copied through pinned staging behind the end event (a copy into
        // pageable memory costs the caller tens of microseconds per solve)
        h_states_.assign(h_states_pin_, h_states_pin_ + nsys);
        states = h_states_.data();
        if (opt.use_graph) {
            // the launches this solve's single iteration loop needed: the K2
            // of iteration k + 1 stops a system after k iterations
            pred_launches_ = 0;
            if (passes == 1) {
                int nmax = 0;
                for (int q = 0; q < nsys; ++q) nmax = std::max(nmax, states[q].iters);
                pred_launches_ = std::min(nmax + 1, max_launch_iters);
                pred_max_iter_ = opt.max_iter;
                pred_rel_ = opt.rel_tol;
                pred_abs_ = opt.abs_tol;
            }
        }
    }
    h_hist_.clear();
    if (hist_ptr_) {
        h_hist_.resize((size_t)nsys * hist_stride_);
        SPAI_HIP(hipMemcpy(h_hist_.data(), hist_ptr_, sizeof(double) * h_hist_.size(), hipMemcpyDeviceToHost));
    }
    if (res) {
        res->ms = ms;
        res->host_checks = checks;
        res->iterations_max = 0;
        res->iterations_sum = 0;
        res->all_converged = 1;
        for (int q = 0; q < nsys; ++q) {
            const SysState &S = states[q];
            res->iterations_max = std::max(res->iterations_max, S.iters);
            res->iterations_sum += S.iters;
            if (!(S.status == kConverged || S.status == kConvergedAtStart)) res->all_converged = 0;
        }
    }
}

int BiCGStab::history(int s, double *out, int maxlen) const {
    if (h_hist_.empty() || s < 0 || s >= (int)h_states_.size()) return 0;
    int k = std::min(h_states_[(size_t)s].iters, hist_stride_);
    if (h_states_[(size_t)s].status == kMaxIter) k = std::min(k, hist_stride_ - 1);
    k = std::min(k, maxlen);
    for (int i = 0; i < k; ++i) out[i] = h_hist_[(size_t)s * hist_stride_ + i];
    return k;
}

double *BiCGStab::seed_timing_state(const OperatorView &A, const TileMap &tm, int n, int nsys, int max_iter,
                                     hipStream_t st) {
    tvec_.reserve((size_t)2 * n);
    double *tb = tvec_.get(), *tx = tvec_.get() + n;
    hipLaunchKernelGGL(k_fill_double, dim3(ceil_div(n, 256)), dim3(256), 0, st, tb, 1.0, (long long)n);
    hipLaunchKernelGGL(k_fill_double, dim3(ceil_div(n, 256)), dim3(256), 0, st, tx, 0.0, (long long)n);
    hipLaunchKernelGGL(k_state_init, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, st_, nsys, max_iter,
                       cnt_, tm.nnodes + tm.ngroups);
    enqueue_init(A, tm, tx, tb, 0.0, 1e-300, 0, st);
    return tx;
}

float BiCGStab::time_kernel(int which, const OperatorView &A, const OperatorView &M, const TileMap &tm, int n,
                            int nsys, int reps, hipStream_t st) {
    require(which >= 0 && which <= 5, "time_kernel: which must be 0..5");
    require((which != 1 && which != 3 && which != 5) || M.val, "time_kernel: no preconditioner");
    defer_x_ = nsys > 1;  // as the solve runs it
    rs32_ = M.val && M.fp32;
    ensure(n, nsys, tm);
    prepare_blocks(A, M, tm, n, nsys, st);
    hist_ptr_ = nullptr;
    hist_stride_ = 0;
    Ctx c = make_ctx(tm, nsys);
    const int parity = 0;  // the y and z buffers of an even iteration
    // The kernels run on the solver's own vectors and states, seeded as a
    // solve from x0 = 0, b = 1 starts (K1 updates the solver-owned x, never a
    // caller's); every system stays active.
    double *tx = seed_timing_state(A, tm, n, nsys, 1 << 30, st);
    hipLaunchKernelGGL(k_state_activate, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, st_, nsys,
                       cnt_, tm.nnodes + tm.ngroups);
    {
        const long long ng = (long long)kSlots * tm.ngroups;
        hipLaunchKernelGGL(k_seed_partials, dim3(ceil_div(std::max<long long>(ng, nsys), 256)), dim3(256), 0, st,
                           st_, nsys, gpart_, ng);
    }
    hipEvent_t e0, e1;
    SPAI_HIP(hipEventCreate(&e0));
    SPAI_HIP(hipEventCreate(&e1));
    auto once = [&]() {
        switch (which) {
        case 0: {
            OpSpmvPlain op{p_[0], y_[0]};
            launch_units<true, double>(A, tm, c, op, nullptr, 0, st, flip());
        } break;
        case 5: {
            OpSpmvPlain op{p_[0], y_[0]};
            if (M.fp32) launch_units<true, float>(M, tm, c, op, nullptr, 0, st, flip());
            else launch_units<true, double>(M, tm, c, op, nullptr, 0, st, flip());
        } break;
        case 1: {
            // an odd iteration's fused K1 (a batch's two-update form)
            const K1Args k1{tx, r_, p_[0], ybuf(1), s_, t_, p_[1], a_, zbuf(1), ybuf(0), zbuf(0)};
            if (M.fp32) {
                launch_k1<true, float, kYProduct>(M, tm, c, x_mode(1), k1, nullptr, 0, st, flip());
            } else {
                launch_k1<true, double, kYProduct>(M, tm, c, x_mode(1), k1, nullptr, 0, st, flip());
            }
        } break;
        case 2:
            enqueue_k2(A, tm, parity, nullptr, 0, st);
            break;
        case 3:  // the form the solver uses with every unit active
            enqueue_k3(M, tm, parity, nullptr, 0, tm.nunits >= split_units(tm), st);
            break;
        case 4:
            enqueue_k4(A, tm, parity, nullptr, 0, st);
            break;
        }
    };
    once();
    SPAI_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps; ++i) once();
    SPAI_HIP(hipEventRecord(e1, st));
    SPAI_HIP(hipEventSynchronize(e1));
    SPAI_CHECK_LAUNCH();
    float ms = 0.0f;
    SPAI_HIP(hipEventElapsedTime(&ms, e0, e1));
    SPAI_HIP(hipEventDestroy(e0));
    SPAI_HIP(hipEventDestroy(e1));
    return ms / std::max(1, reps);
}

void BiCGStab::time_iteration(const OperatorView &A, const OperatorView &M, const TileMap &tm, int n, int nsys,
                              int reps, int variant, hipStream_t st, float ms[5]) {
    require(A.val && M.val, "time_iteration: needs A and M");
    require(reps >= 1, "time_iteration: reps must be >= 1");
    require(variant >= -1 && variant <= 1, "time_iteration: variant must be -1, 0 or 1");
    const bool split = variant < 0 ? tm.nunits >= split_units(tm) : variant == 1;
    defer_x_ = nsys > 1;  // as the solve runs it
    rs32_ = M.val && M.fp32;
    ensure(n, nsys, tm);
    prepare_blocks(A, M, tm, n, nsys, st);
    hist_ptr_ = nullptr;
    hist_stride_ = 0;
    Ctx c = make_ctx(tm, nsys);
    // no threshold is reached: every system stays active for the timed iterations
    double *tx = seed_timing_state(A, tm, n, nsys, reps + 3, st);
    const double *b_keep = b_;
    double *x_keep = x_;
    b_ = tvec_.get();
    x_ = tx;
    constexpr int K = 5;  // event slots per iteration: K1 (or its halves), K2, K3, K4
    std::vector<hipEvent_t> ev((size_t)reps * K + 1);
    for (auto &e : ev) SPAI_HIP(hipEventCreate(&e));
    try {
        for (int it = 0; it < reps; ++it) {
            const int parity = it & 1;
            double *pin = p_[parity], *pout = p_[parity ^ 1];
            const K1Args k1{x_, r_, pout, ybuf(parity), s_, t_, pin, a_, zbuf(parity), ybuf(parity ^ 1), zbuf(parity ^ 1)};
            hipEvent_t *e = &ev[(size_t)it * K];
            SPAI_HIP(hipEventRecord(e[0], st));
            if (split) {
                launch_k1<false, double, kYNone>(A, tm, c, x_mode(parity), k1, nullptr, 0, st, flip());
                SPAI_HIP(hipEventRecord(e[1], st));
                enqueue_mapply_p(M, tm, parity, nullptr, 0, st);
            } else {
                if (M.fp32) {
                    launch_k1<true, float, kYProduct>(M, tm, c, x_mode(parity), k1, nullptr, 0, st, flip());
                } else {
                    launch_k1<true, double, kYProduct>(M, tm, c, x_mode(parity), k1, nullptr, 0, st, flip());
                }
                SPAI_HIP(hipEventRecord(e[1], st));
            }
            SPAI_HIP(hipEventRecord(e[2], st));
            enqueue_k2(A, tm, parity, nullptr, 0, st);
            SPAI_HIP(hipEventRecord(e[3], st));
            enqueue_k3(M, tm, parity, nullptr, 0, split, st);
            SPAI_HIP(hipEventRecord(e[4], st));
            enqueue_k4(A, tm, parity, nullptr, 0, st);
        }
        SPAI_HIP(hipEventRecord(ev.back(), st));
        SPAI_HIP(hipEventSynchronize(ev.back()));
        SPAI_CHECK_LAUNCH();
        // ms: [K1 or its streaming half, M-apply half (0 when fused), K2, K3, K4]
        for (int k = 0; k < K; ++k) ms[k] = 0.0f;
        for (int it = 0; it < reps; ++it)
            for (int k = 0; k < K; ++k) {
                if (k == 1 && !split) continue;
                float t = 0.0f;
                SPAI_HIP(hipEventElapsedTime(&t, ev[(size_t)it * K + k], ev[(size_t)it * K + k + 1]));
                ms[k] += t / reps;
            }
    } catch (...) {
        for (auto &e : ev) (void)hipEventDestroy(e);
        b_ = b_keep;
        x_ = x_keep;
        throw;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    b_ = b_keep;
    x_ = x_keep;
}

}  // namespace spai
