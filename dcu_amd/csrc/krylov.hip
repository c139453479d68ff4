// Sparse formats, row tiles and the fused right-preconditioned BiCGSTAB.
//
// The recurrences are those of include/bicgstab/bicgstab_solver.h:106-232
// (one iteration: y = M p, a = A y, alpha = rho / (r*, a), s = r - alpha a,
// z = M s, t = A z, omega = (t, s) / (t, t), x += alpha y + omega z,
// r = s - omega t, rho' = (r*, r), beta = (rho'/rho)(alpha/omega),
// p = r + beta (p - omega a), stop on ||r||_2 < TOL).  The reference issues
// 18 hipSPARSE/hipBLAS calls and 5 blocking scalar reads per iteration; here
// one iteration is five kernels and every scalar stays in device memory:
//
//   K1  y = M p'   with p' = beta p + r - beta omega a formed while gathering
//                  (the p update of the previous iteration, fused), p' stored
//   K2  a = A y,   (r*, a) reduced, alpha = rho / (r*, a) by the last tile
//   K3  z = M s    with s = r - alpha a formed while gathering, s stored
//   K4  t = A z,   (t, s), (t, t) reduced, omega by the last tile
//   K5  x += alpha y + omega z, r = s - omega t, (r*, r), (r, r) reduced,
//       beta, convergence and the iteration count by the last tile
//
// Every kernel walks row tiles (one workgroup per tile); a tile's nonzeros
// are streamed with fully coalesced loads into LDS as products, then each
// thread sums its rows in storage order (deterministic).  Reductions are
// per-tile partials summed in a fixed order by the last tile of each system,
// so results are reproducible run to run.  All systems of a batch share the
// launches; tiles of a converged system return at once.
#include "krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace spai {

// Reductions are counted in two levels so that no counter sees more than
// kBucket atomic increments per launch (same-address atomics serialise at
// the memory side): groups -> buckets of up to kBucket consecutive groups of
// one system -> the system.
struct Ctx {
    SysState *st;
    const int *group_bucket;  // bucket of each group
    const int *bucket_group;  // nbuckets + 1: first group of each bucket
    const int *sys_buckets;   // nsys + 1: first bucket of each system
    int *counter;             // [nsys] systems, [nsys] = n_active, [nsys + 1 + b] buckets
    double *partials;         // ngroups * 2 group partials
    double *bpartials;        // nbuckets * 2 bucket partials
    int nsys;
    double *hist;             // optional ||r_k|| history, hist_stride entries per system
    int hist_stride;
};


namespace {

constexpr int kThreads = 256;
constexpr int kLdsProducts = 2048;  // products staged per tile (16 KB)
constexpr int kBucket = 32;         // groups per first-level reduction counter

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

// Publishes this group's partials; the last group of a bucket sums the
// bucket, the last bucket of the system sums the buckets (both in a fixed
// order) and calls op.finish on its thread 0.
//
// No __threadfence(): on gfx950 an agent-scope fence writes back and
// invalidates the whole L2 (buffer_wbl2 / buffer_inv), which, issued by
// thousands of workgroups, evicts the vectors the others are gathering.
// Instead partials are agent-scope atomic stores (coherent across the XCDs'
// L2s), the thread waits until they are acknowledged (vmcnt(0)) before the
// relaxed counter increment, and the last workgroup reads them back with
// agent-scope atomic loads.
template <int NRED>
__device__ __forceinline__ bool publish_and_count(double *dst, const double (&v)[NRED], int *counter, int expect,
                                                  int *flag) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NRED; ++q) __hip_atomic_store(&dst[q], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        int prev = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *flag = (prev == expect - 1);
        if (*flag) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return *flag != 0;
}

template <int NRED>
__device__ __forceinline__ void sum_range(const double *src, int first, int last, double (&acc)[NRED], double *sh) {
#pragma unroll
    for (int q = 0; q < NRED; ++q) acc[q] = 0.0;
    for (int t = first + (int)threadIdx.x; t < last; t += blockDim.x) {
#pragma unroll
        for (int q = 0; q < NRED; ++q) acc[q] += load_coherent(&src[(size_t)t * NRED + q]);
    }
    block_reduce<NRED>(acc, sh);
}

template <int NRED, class Op>
__device__ __forceinline__ void reduce_system(const Ctx &c, int group, int s, double (&v)[NRED], Op &op, double *sh) {
    __shared__ int flag;
    block_reduce<NRED>(v, sh);
    const int b = c.group_bucket[group];
    const int g0 = c.bucket_group[b], g1 = c.bucket_group[b + 1];
    if (!publish_and_count<NRED>(&c.partials[(size_t)group * NRED], v, &c.counter[c.nsys + 1 + b], g1 - g0, &flag))
        return;
    double bsum[NRED];
    sum_range<NRED>(c.partials, g0, g1, bsum, sh);
    const int b0 = c.sys_buckets[s], b1 = c.sys_buckets[s + 1];
    if (!publish_and_count<NRED>(&c.bpartials[(size_t)b * NRED], bsum, &c.counter[s], b1 - b0, &flag)) return;
    double tot[NRED];
    sum_range<NRED>(c.bpartials, b0, b1, tot, sh);
    if (threadIdx.x == 0) op.finish(c, s, tot);
}

// ------------------------------------------------------- tile kernels ----
struct TileView {
    const int *row;         // tile row boundaries
    const int *sys;         // system per tile
    const int *group_tile;  // first tile per group
    const int *group_sys;   // system per group
};

constexpr int kRowsPerTile = 512;                  // row cap of a tile (host-built)
constexpr int kRowsPerThread = kRowsPerTile / kThreads;

// For every row i of the workgroup's tile(s): y_i = sum_p val[p] *
// op.gather(idx[p]) (stored as products in LDS, then summed per row in
// storage order), then op.row(i, y_i, pre_i, acc) where pre_i = op.pre(i) are
// the row's own vector operands, loaded before the product so that they
// overlap with it.  One tile needs two dependent global round trips: the
// index/value/row-pointer/operand loads, then the gathers.  HAS_MATRIX =
// false skips the product (streaming ops).  Ops with reductions run one
// workgroup per group of tiles (GROUPED) and publish one partial per group.
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Consecutive tiles gather overlapping vector ranges, so within every period
// of 8 * kXcdRun workgroups, XCD x is given the contiguous run of kXcdRun
// units starting at x * kXcdRun; a partial last period keeps the identity.
// Runs (not one block of G/8 units per XCD) keep every system spread over all
// XCDs when only a few systems are still iterating.
constexpr int kXcds = 8;
constexpr int kXcdRun = 8;
__device__ __forceinline__ int xcd_unit(int b, int G) {
    constexpr int P = kXcds * kXcdRun;
    const int p = b / P;
    if ((p + 1) * P > G) return b;
    const int i = b - p * P;
    return p * P + (i % kXcds) * kXcdRun + i / kXcds;
}

template <bool HAS_MATRIX, class MT, class Op>
__global__ __launch_bounds__(kThreads) void k_tiles(const int *__restrict__ ptr, const int *__restrict__ idx,
                                                    const MT *__restrict__ val, TileView tv, Ctx c, Op op) {
    constexpr int NRED = Op::NRED;
    constexpr bool GROUPED = NRED > 0;
    __shared__ double prod[HAS_MATRIX ? kLdsProducts : 1];
    __shared__ int rp[HAS_MATRIX ? kRowsPerTile + 1 : 1];
    __shared__ double sh[(NRED > 0 ? NRED : 1) * 4];
    const int unit = xcd_unit(blockIdx.x, gridDim.x);
    const int s = GROUPED ? tv.group_sys[unit] : tv.sys[unit];
    if (!op.enabled(c.st[s])) return;
    op.load(c.st[s]);
    const int ft0 = GROUPED ? tv.group_tile[unit] : unit;
    const int ft1 = GROUPED ? tv.group_tile[unit + 1] : unit + 1;
    const int tid = threadIdx.x;
    double acc[NRED > 0 ? NRED : 1];
#pragma unroll
    for (int q = 0; q < (NRED > 0 ? NRED : 1); ++q) acc[q] = 0.0;

    for (int ft = ft0; ft < ft1; ++ft) {
        const int r0 = tv.row[ft], r1 = tv.row[ft + 1];
        const int nrows = r1 - r0;
        if (!HAS_MATRIX) {
            for (int i = r0 + tid; i < r1; i += kThreads) op.row(i, 0.0, op.pre(i), acc);
            continue;
        }
        const int p0 = ptr[r0];
        const int cnt = ptr[r1] - p0;
        if (cnt <= kLdsProducts && nrows <= kRowsPerTile) {
            typename Op::Pre pre[kRowsPerThread];
#pragma unroll
            for (int k = 0; k < kRowsPerThread; ++k) {
                const int i = r0 + tid + k * kThreads;
                if (i < r1) pre[k] = op.pre(i);
            }
            for (int k = tid; k <= nrows; k += kThreads) rp[k] = ptr[r0 + k] - p0;
            constexpr int U = kLdsProducts / kThreads;
            int jj[U];
            MT vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = u * kThreads + tid;
                jj[u] = q < cnt ? idx[p0 + q] : 0;
                vv[u] = q < cnt ? val[p0 + q] : (MT)0;
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
#pragma unroll
            for (int k = 0; k < kRowsPerThread; ++k) {
                const int li = tid + k * kThreads;
                if (li < nrows) {
                    const int b = rp[li], e = rp[li + 1];
                    double sum = 0.0;
                    for (int q = b; q < e; ++q) sum += prod[q];
                    op.row(r0 + li, sum, pre[k], acc);
                }
            }
            __syncthreads();  // prod and rp are reused by the next tile
        } else {
            // a row longer than the LDS buffer (it has a tile of its own):
            // the whole workgroup reduces it
            for (int i = r0; i < r1; ++i) {
                const int b = ptr[i], e = ptr[i + 1];
                double part[1] = {0.0};
                for (int q = b + tid; q < e; q += kThreads) part[0] += (double)val[q] * op.gather(idx[q]);
                block_reduce<1>(part, sh);
                if (tid == 0) op.row(i, part[0], op.pre(i), acc);
            }
        }
    }
    if constexpr (NRED > 0) {
        double v[NRED];
#pragma unroll
        for (int q = 0; q < NRED; ++q) v[q] = acc[q];
        reduce_system<NRED>(c, unit, s, v, op, sh);
    }
}

// ------------------------------------------------------------- the ops ----
struct OpInit {  // r = b - A x0, r* = r, p = a = 0; rho = rr = (r, r); tol
    static constexpr int NRED = 2;
    struct Pre {
        double b;
    };
    const double *x, *b;
    double *r, *rs, *p0, *p1, *a;
    double rel_tol, abs_tol;
    int restart;  // 1: restart of an active system from its current x (iterations and tol kept)
    __device__ bool enabled(const SysState &S) const { return !restart || S.active; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int i) const { return {b[i]}; }
    __device__ double gather(int j) const { return x[j]; }
    __device__ void row(int i, double ax, const Pre &pr, double (&acc)[2]) const {
        double ri = -ax + pr.b;
        r[i] = ri;
        rs[i] = ri;
        p0[i] = 0.0;
        p1[i] = 0.0;
        a[i] = 0.0;
        acc[0] += ri * ri;
        acc[1] += pr.b * pr.b;
    }
    __device__ void finish(const Ctx &c, int s, const double (&tot)[2]) const {
        SysState &S = c.st[s];
        S.rho = tot[0];
        S.rr = tot[0];
        S.alpha = 0.0;
        S.omega = 0.0;
        S.beta = 0.0;
        S.true_rr = 0.0;
        if (restart) {
            S.restarts += 1;
            if (sqrt(tot[0]) < S.tol) {
                S.status = kConverged;
                S.active = 0;
                atomicSub(&c.counter[c.nsys], 1);
            }
            return;
        }
        S.err0 = sqrt(tot[0]);
        S.tol = abs_tol > 0.0 ? abs_tol : rel_tol * sqrt(tot[1]);
        S.iters = 0;
        if (S.err0 < S.tol) {
            // r0 already below the threshold: the reference would divide 0/0
            S.status = kConvergedAtStart;
            S.active = 0;
            atomicSub(&c.counter[c.nsys], 1);
        }
    }
};

template <class MT>
struct OpMapplyP {  // K1: y = M p', p' = beta p + r - beta omega a
    static constexpr int NRED = 0;
    struct Pre {
        double p;
    };
    const double *r, *pin, *a;
    double *pout, *y;
    double beta, nbw;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) {
        beta = S.beta;
        nbw = -S.beta * S.omega;
    }
    __device__ double pval(int j) const { return (beta * pin[j] + r[j]) + nbw * a[j]; }
    __device__ Pre pre(int i) const { return {pval(i)}; }
    __device__ double gather(int j) const { return pval(j); }
    __device__ void row(int i, double mx, const Pre &pr, double (&)[1]) const {
        y[i] = mx;
        pout[i] = pr.p;
    }
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

struct OpSpmvAlpha {  // K2: a = A y, alpha = rho / (r*, a)
    static constexpr int NRED = 1;
    struct Pre {
        double rs;
    };
    const double *y, *rs;
    double *a;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int i) const { return {rs[i]}; }
    __device__ double gather(int j) const { return y[j]; }
    __device__ void row(int i, double ax, const Pre &pr, double (&acc)[1]) const {
        a[i] = ax;
        acc[0] += pr.rs * ax;
    }
    __device__ void finish(const Ctx &c, int s, const double (&tot)[1]) const {
        SysState &S = c.st[s];
        S.alpha = S.rho / tot[0];
    }
};

template <class MT>
struct OpMapplyS {  // K3: z = M s, s = r - alpha a
    static constexpr int NRED = 0;
    struct Pre {
        double s;
    };
    const double *r, *a;
    double *s, *z;
    double nalpha;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) { nalpha = -S.alpha; }
    __device__ double sval(int j) const { return r[j] + nalpha * a[j]; }
    __device__ Pre pre(int i) const { return {sval(i)}; }
    __device__ double gather(int j) const { return sval(j); }
    __device__ void row(int i, double mx, const Pre &pr, double (&)[1]) const {
        z[i] = mx;
        s[i] = pr.s;
    }
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

struct OpSpmvOmega {  // K4: t = A z, omega = (t, s) / (t, t)
    static constexpr int NRED = 2;
    struct Pre {
        double s;
    };
    const double *z, *s;
    double *t;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int i) const { return {s[i]}; }
    __device__ double gather(int j) const { return z[j]; }
    __device__ void row(int i, double ax, const Pre &pr, double (&acc)[2]) const {
        t[i] = ax;
        acc[0] += ax * pr.s;
        acc[1] += ax * ax;
    }
    __device__ void finish(const Ctx &c, int s_, const double (&tot)[2]) const {
        SysState &S = c.st[s_];
        S.omega = tot[0] / tot[1];
    }
};

struct OpUpdate {  // K5: x, r update; rho', beta, ||r||, stop test
    static constexpr int NRED = 2;
    struct Pre {
        double x, y, z, s, t, rs;
    };
    double *x, *r;
    const double *y, *z, *s, *t, *rs;
    double alpha, omega;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) {
        alpha = S.alpha;
        omega = S.omega;
    }
    __device__ Pre pre(int i) const { return {x[i], y[i], z[i], s[i], t[i], rs[i]}; }
    __device__ double gather(int) const { return 0.0; }
    __device__ void row(int i, double, const Pre &pr, double (&acc)[2]) const {
        x[i] = (pr.x + alpha * pr.y) + omega * pr.z;
        double ri = pr.s + (-omega) * pr.t;
        r[i] = ri;
        acc[0] += pr.rs * ri;
        acc[1] += ri * ri;
    }
    __device__ void finish(const Ctx &c, int s, const double (&tot)[2]) const {
        SysState &S = c.st[s];
        double rho_new = tot[0];
        S.beta = (rho_new / S.rho) * (S.alpha / S.omega);
        S.rho = rho_new;
        S.rr = tot[1];
        S.iters += 1;
        double err = sqrt(tot[1]);
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
            atomicSub(&c.counter[c.nsys], 1);
        }
    }
};

struct OpTrueResidual {  // ||b - A x||^2 per system
    static constexpr int NRED = 1;
    struct Pre {
        double b;
    };
    const double *x, *b;
    __device__ bool enabled(const SysState &) const { return true; }
    __device__ void load(const SysState &) {}
    __device__ Pre pre(int i) const { return {b[i]}; }
    __device__ double gather(int j) const { return x[j]; }
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
    __device__ Pre pre(int) const { return {}; }
    __device__ double gather(int j) const { return x[j]; }
    __device__ void row(int i, double ax, const Pre &, double (&)[1]) const { y[i] = ax; }
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

struct OpCopyP {  // K1 without a preconditioner: y = p'
    static constexpr int NRED = 0;
    struct Pre {
        double p;
    };
    const double *r, *pin, *a;
    double *pout, *y;
    double beta, nbw;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) {
        beta = S.beta;
        nbw = -S.beta * S.omega;
    }
    __device__ Pre pre(int i) const { return {(beta * pin[i] + r[i]) + nbw * a[i]}; }
    __device__ double gather(int) const { return 0.0; }
    __device__ void row(int i, double, const Pre &pr, double (&)[1]) const {
        pout[i] = pr.p;
        y[i] = pr.p;
    }
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

struct OpCopyS {  // K3 without a preconditioner: z = s
    static constexpr int NRED = 0;
    struct Pre {
        double s;
    };
    const double *r, *a;
    double *s, *z;
    double nalpha;
    __device__ bool enabled(const SysState &S) const { return S.active; }
    __device__ void load(const SysState &S) { nalpha = -S.alpha; }
    __device__ Pre pre(int i) const { return {r[i] + nalpha * a[i]}; }
    __device__ double gather(int) const { return 0.0; }
    __device__ void row(int i, double, const Pre &pr, double (&)[1]) const {
        s[i] = pr.s;
        z[i] = pr.s;
    }
    __device__ void finish(const Ctx &, int, const double (&)[1]) const {}
};

__global__ void k_state_init(SysState *st, int nsys, int max_iter, int *counter, int nbuckets) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) {
        SysState z{};
        z.active = 1;
        z.status = kRunning;
        z.max_iter = max_iter;
        st[s] = z;
        counter[s] = 0;
    }
    if (s == 0) counter[nsys] = nsys;
    for (int b = s; b < nbuckets; b += gridDim.x * blockDim.x) counter[nsys + 1 + b] = 0;
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
__global__ void k_state_activate(SysState *st, int nsys, int *counter, int nbuckets) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) {
        st[s].active = 1;
        counter[s] = 0;
    }
    for (int b = s; b < nbuckets; b += gridDim.x * blockDim.x) counter[nsys + 1 + b] = 0;
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

template <bool HM, class MT, class Op>
void launch_tiles(const OperatorView &O, const TileMap &tm, const Ctx &c, const Op &op, hipStream_t st) {
    TileView tv{tm.row.get(), tm.sys.get(), tm.group_tile.get(), tm.group_sys.get()};
    const int grid = Op::NRED > 0 ? tm.ngroups : tm.ntiles;
    hipLaunchKernelGGL((k_tiles<HM, MT, Op>), dim3(grid), dim3(kThreads), 0, st, O.ptr, O.idx, (const MT *)O.val,
                       tv, c, op);
}

}  // namespace

// ================================================================ host ====
void csc_to_csr_device(int n, int nnz, const int *colptr, const int *rowidx, const double *cscval, int *rowptr,
                       int *colidx, double *csrval, int *perm, hipStream_t st) {
    require(n >= 0 && nnz >= 0, "csc_to_csr: negative size");
    DeviceArray<int> cnt((size_t)n + 1), src((size_t)std::max(nnz, 1)), tmp;
    const int B = 256;
    hipLaunchKernelGGL(k_fill_int, dim3(std::max(1, ceil_div(n + 1, B))), dim3(B), 0, st, cnt.get(), 0,
                       (long long)n + 1);
    if (nnz > 0) hipLaunchKernelGGL(k_count_rows, dim3(ceil_div(nnz, B)), dim3(B), 0, st, rowidx, nnz, cnt.get());
    exclusive_scan(cnt.get(), rowptr, n, tmp, st);
    hipLaunchKernelGGL(k_fill_int, dim3(std::max(1, ceil_div(n + 1, B))), dim3(B), 0, st, cnt.get(), 0,
                       (long long)n + 1);
    if (n > 0) {
        hipLaunchKernelGGL(k_fill_rows, dim3(ceil_div(n, B)), dim3(B), 0, st, colptr, rowidx, n, rowptr, cnt.get(),
                           colidx, src.get());
        hipLaunchKernelGGL(k_sort_rows, dim3(ceil_div(n, B)), dim3(B), 0, st, n, rowptr, colidx, src.get(), cscval,
                           csrval, perm);
    }
    SPAI_CHECK_LAUNCH();
    SPAI_HIP(hipStreamSynchronize(st));  // temporaries are freed on return
}

void build_tiles(TileMap &tm, const int *ptr, const std::vector<int> &h_row_off,
                 const std::vector<long long> &h_nnz_off, hipStream_t st) {
    const int nsys = (int)h_row_off.size() - 1;
    require(nsys >= 1, "build_tiles: no systems");
    const int n = h_row_off[(size_t)nsys];
    std::vector<int> hp((size_t)n + 1);
    SPAI_HIP(hipMemcpyAsync(hp.data(), ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipStreamSynchronize(st));
    // nonzeros per tile: enough tiles to fill the GPU on large inputs, at most
    // the LDS staging buffer; rows per tile at most kRowsPerTile
    long long total = h_nnz_off[(size_t)nsys];
    int target = 256;
    while (target < kLdsProducts && (long long)target * 4096 < total) target <<= 1;
    tm.target_nnz = target;
    std::vector<int> trow, tsys;
    tm.h_sys_tiles.assign((size_t)nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        const int r0 = h_row_off[(size_t)s], r1 = h_row_off[(size_t)s + 1];
        int i = r0;
        if (r0 == r1) {  // empty system: one empty tile keeps the reductions uniform
            trow.push_back(r0);
            tsys.push_back(s);
        }
        while (i < r1) {
            const int start = i;
            long long nz = 0;
            if (hp[(size_t)i + 1] - hp[(size_t)i] > kLdsProducts) {
                ++i;  // a row longer than the staging buffer gets a tile of its own
            } else {
                while (i < r1 && i - start < kRowsPerTile) {
                    const int len = hp[(size_t)i + 1] - hp[(size_t)i];
                    if (len > kLdsProducts || (i > start && nz + len > target)) break;
                    nz += len;
                    ++i;
                }
            }
            trow.push_back(start);
            tsys.push_back(s);
        }
        tm.h_sys_tiles[(size_t)s + 1] = (int)tsys.size();
    }
    trow.push_back(n);
    tm.ntiles = (int)tsys.size();
    // reducing kernels: groups of consecutive tiles of one system, enough of
    // them to give every CU several workgroups
    tm.group_size = (int)std::max<long long>(1, std::min<long long>(4, tm.ntiles / 2048));
    std::vector<int> gtile, gsys, sgroups((size_t)nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        for (int t = tm.h_sys_tiles[(size_t)s]; t < tm.h_sys_tiles[(size_t)s + 1]; t += tm.group_size) {
            gtile.push_back(t);
            gsys.push_back(s);
        }
        sgroups[(size_t)s + 1] = (int)gsys.size();
    }
    gtile.push_back(tm.ntiles);
    tm.ngroups = (int)gsys.size();
    // buckets of up to kBucket consecutive groups of one system
    std::vector<int> gbucket((size_t)tm.ngroups), bgroup, sbuckets((size_t)nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        for (int g = sgroups[(size_t)s]; g < sgroups[(size_t)s + 1]; g += kBucket) {
            int bidx = (int)bgroup.size();
            bgroup.push_back(g);
            for (int h = g; h < std::min(g + kBucket, sgroups[(size_t)s + 1]); ++h) gbucket[(size_t)h] = bidx;
        }
        sbuckets[(size_t)s + 1] = (int)bgroup.size();
    }
    bgroup.push_back(tm.ngroups);
    tm.nbuckets = (int)bgroup.size() - 1;
    tm.group_bucket.resize(gbucket.size());
    tm.bucket_group.resize(bgroup.size());
    tm.sys_buckets.resize(sbuckets.size());
    tm.row.resize(trow.size());
    tm.sys.resize(tsys.size());
    tm.sys_tiles.resize((size_t)nsys + 1);
    tm.group_tile.resize(gtile.size());
    tm.group_sys.resize(gsys.size());
    tm.sys_groups.resize(sgroups.size());
    auto up = [&](DeviceArray<int> &d, const std::vector<int> &h) {
        SPAI_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, st));
    };
    up(tm.row, trow);
    up(tm.sys, tsys);
    up(tm.sys_tiles, tm.h_sys_tiles);
    up(tm.group_tile, gtile);
    up(tm.group_sys, gsys);
    up(tm.sys_groups, sgroups);
    up(tm.group_bucket, gbucket);
    up(tm.bucket_group, bgroup);
    up(tm.sys_buckets, sbuckets);
    SPAI_HIP(hipStreamSynchronize(st));
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
BiCGStab::~BiCGStab() { drop_graph(); }

void BiCGStab::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    for (auto &k : graph_key_) k = nullptr;
    graph_iters_ = 0;
}

Ctx BiCGStab::make_ctx(const TileMap &tm, int nsys) {
    return Ctx{state_.get(),     tm.group_bucket.get(), tm.bucket_group.get(), tm.sys_buckets.get(),
               counter_.get(),   partials_.get(),       bpartials_.get(),      nsys,
               hist_ptr_,        hist_stride_};
}

void BiCGStab::ensure(int n, int nsys, int ntiles, int nbuckets) {
    if (n != n_) {
        drop_graph();
        vec_.resize((size_t)9 * std::max(n, 1));
        double *v = vec_.get();
        size_t N = (size_t)std::max(n, 1);
        r_ = v;
        rs_ = v + N;
        p_[0] = v + 2 * N;
        p_[1] = v + 3 * N;
        y_ = v + 4 * N;
        a_ = v + 5 * N;
        s_ = v + 6 * N;
        z_ = v + 7 * N;
        t_ = v + 8 * N;
        n_ = n;
    }
    if (nsys != nsys_ || nbuckets != nbuckets_) {
        drop_graph();
        state_.resize((size_t)nsys);
        counter_.resize((size_t)nsys + 1 + nbuckets);
        bpartials_.resize((size_t)2 * std::max(nbuckets, 1));
        nsys_ = nsys;
        nbuckets_ = nbuckets;
    }
    if (ntiles != ntiles_) {
        drop_graph();
        partials_.resize((size_t)2 * std::max(ntiles, 1));
        ntiles_ = ntiles;
    }
}

void BiCGStab::enqueue_iteration(const OperatorView &A, const OperatorView &M, const TileMap &tm, int parity,
                                 hipStream_t st) {
    Ctx c = make_ctx(tm, nsys_);
    double *pin = p_[parity], *pout = p_[parity ^ 1];
    if (M.val) {
        if (M.fp32) {
            OpMapplyP<float> op{r_, pin, a_, pout, y_, 0.0, 0.0};
            launch_tiles<true, float>(M, tm, c, op, st);
        } else {
            OpMapplyP<double> op{r_, pin, a_, pout, y_, 0.0, 0.0};
            launch_tiles<true, double>(M, tm, c, op, st);
        }
    } else {
        // no preconditioner: y = p' through the streaming path
        OpCopyP op{r_, pin, a_, pout, y_, 0.0, 0.0};
        launch_tiles<false, double>(A, tm, c, op, st);
    }
    {
        OpSpmvAlpha op{y_, rs_, a_};
        launch_tiles<true, double>(A, tm, c, op, st);
    }
    if (M.val) {
        if (M.fp32) {
            OpMapplyS<float> op{r_, a_, s_, z_, 0.0};
            launch_tiles<true, float>(M, tm, c, op, st);
        } else {
            OpMapplyS<double> op{r_, a_, s_, z_, 0.0};
            launch_tiles<true, double>(M, tm, c, op, st);
        }
    } else {
        OpCopyS op{r_, a_, s_, z_, 0.0};
        launch_tiles<false, double>(A, tm, c, op, st);
    }
    {
        OpSpmvOmega op{z_, s_, t_};
        launch_tiles<true, double>(A, tm, c, op, st);
    }
    {
        OpUpdate op{x_, r_, y_, z_, s_, t_, rs_, 0.0, 0.0};
        launch_tiles<false, double>(A, tm, c, op, st);
    }
    SPAI_CHECK_LAUNCH();
}

void BiCGStab::solve(const OperatorView &A, const OperatorView &M, const TileMap &tm, int n, int nsys,
                     const double *b, double *x, const SolveOptions &opt, hipStream_t st, SolveResult *res) {
    require(n > 0 && nsys >= 1 && A.ptr && A.idx && A.val && b && x, "bicgstab: bad arguments");
    require(tm.ntiles > 0 && tm.ngroups > 0 && (int)tm.h_sys_tiles.size() == nsys + 1,
            "bicgstab: tiles do not match the systems");
    require(opt.max_iter >= 1, "bicgstab: max_iter must be >= 1");
    ensure(n, nsys, tm.ngroups, tm.nbuckets);
    b_ = b;
    x_ = x;
    if (opt.record_history) {
        hist_.reserve((size_t)nsys * (size_t)opt.max_iter);
        hist_ptr_ = hist_.get();
        hist_stride_ = opt.max_iter;
    } else {
        hist_ptr_ = nullptr;
        hist_stride_ = 0;
    }
    hipEvent_t e0, e1;
    SPAI_HIP(hipEventCreate(&e0));
    SPAI_HIP(hipEventCreate(&e1));
    SPAI_HIP(hipEventRecord(e0, st));
    Ctx c = make_ctx(tm, nsys);
    hipLaunchKernelGGL(k_state_init, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, state_.get(), nsys, opt.max_iter,
                       counter_.get(), tm.nbuckets);
    {
        OpInit op{x, b, r_, rs_, p_[0], p_[1], a_, opt.rel_tol, opt.abs_tol, 0};
        launch_tiles<true, double>(A, tm, c, op, st);
    }
    SPAI_CHECK_LAUNCH();
    int active = nsys;
    int checks = 0;
    int *h_active = nullptr;
    SPAI_HIP(hipHostMalloc((void **)&h_active, sizeof(int), hipHostMallocDefault));
    auto read_active = [&]() {
        SPAI_HIP(hipMemcpyAsync(h_active, counter_.get() + nsys, sizeof(int), hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
        ++checks;
        return *h_active;
    };
    const int max_launch_iters = opt.max_iter;  // every system stops by then
    int done_iters = 0;
    try {
        active = read_active();
        auto iterate = [&]() {
            if (opt.use_graph) {
                int gi = std::max(2, opt.graph_iters + (opt.graph_iters & 1));
                const void *key[7] = {A.val, M.val, tm.row.get(), b, x, (const void *)(size_t)M.fp32, hist_ptr_};
                bool same = graph_exec_ && gi == graph_iters_;
                for (int q = 0; q < 7 && same; ++q) same = key[q] == graph_key_[q];
                if (!same) {
                    drop_graph();
                    hipStream_t cap;
                    SPAI_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
                    SPAI_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                    for (int k = 0; k < gi; ++k) enqueue_iteration(A, M, tm, k & 1, cap);
                    SPAI_HIP(hipStreamEndCapture(cap, &graph_));
                    SPAI_HIP(hipStreamDestroy(cap));
                    SPAI_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
                    graph_iters_ = gi;
                    for (int q = 0; q < 7; ++q) graph_key_[q] = key[q];
                }
                while (active > 0 && done_iters < max_launch_iters) {
                    SPAI_HIP(hipGraphLaunch(graph_exec_, st));
                    done_iters += gi;
                    active = read_active();
                }
            } else {
                int every = std::max(1, opt.check_every);
                int parity = 0;
                while (active > 0 && done_iters < max_launch_iters) {
                    for (int k = 0; k < every; ++k) {
                        enqueue_iteration(A, M, tm, parity, st);
                        parity ^= 1;
                    }
                    done_iters += every;
                    active = read_active();
                }
            }
        };
        iterate();
        for (int round = 0; opt.true_residual_restart && round < 4; ++round) {
            OpTrueResidual tr{x, b};
            launch_tiles<true, double>(A, tm, c, tr, st);
            hipLaunchKernelGGL(k_check_true, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, state_.get(), nsys,
                               counter_.get());
            SPAI_CHECK_LAUNCH();
            active = read_active();
            if (active == 0) break;
            OpInit op{x, b, r_, rs_, p_[0], p_[1], a_, opt.rel_tol, opt.abs_tol, 1};
            launch_tiles<true, double>(A, tm, c, op, st);
            SPAI_CHECK_LAUNCH();
            active = read_active();
            done_iters = 0;
            iterate();
        }
        {
            OpTrueResidual op{x, b};
            launch_tiles<true, double>(A, tm, c, op, st);
            SPAI_CHECK_LAUNCH();
        }
        SPAI_HIP(hipEventRecord(e1, st));
        SPAI_HIP(hipEventSynchronize(e1));
    } catch (...) {
        (void)hipHostFree(h_active);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipHostFree(h_active);
    float ms = 0.0f;
    SPAI_HIP(hipEventElapsedTime(&ms, e0, e1));
    SPAI_HIP(hipEventDestroy(e0));
    SPAI_HIP(hipEventDestroy(e1));
    h_states_.resize((size_t)nsys);
    SPAI_HIP(hipMemcpy(h_states_.data(), state_.get(), sizeof(SysState) * nsys, hipMemcpyDeviceToHost));
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
        for (auto &S : h_states_) {
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

float BiCGStab::time_kernel(int which, const OperatorView &A, const OperatorView &M, const TileMap &tm, int n,
                            int nsys, int reps, hipStream_t st) {
    // which: 0 bare A*x, 5 bare M*x, 1..4 the fused kernels K1..K4 with the
    // scalars and vectors of the last solve (systems re-activated first)
    require(which >= 0 && which <= 5, "time_kernel: which must be 0..5");
    require((which != 1 && which != 3 && which != 5) || M.val, "time_kernel: no preconditioner");
    ensure(n, nsys, tm.ngroups, tm.nbuckets);
    Ctx c = make_ctx(tm, nsys);
    hipLaunchKernelGGL(k_state_activate, dim3(ceil_div(nsys, 256)), dim3(256), 0, st, state_.get(), nsys,
                       counter_.get(), tm.nbuckets);
    hipEvent_t e0, e1;
    SPAI_HIP(hipEventCreate(&e0));
    SPAI_HIP(hipEventCreate(&e1));
    auto once = [&]() {
        switch (which) {
        case 0: {
            OpSpmvPlain op{p_[0], y_};
            launch_tiles<true, double>(A, tm, c, op, st);
        } break;
        case 5: {
            OpSpmvPlain op{p_[0], y_};
            if (M.fp32) launch_tiles<true, float>(M, tm, c, op, st);
            else launch_tiles<true, double>(M, tm, c, op, st);
        } break;
        case 1: {
            if (M.fp32) {
                OpMapplyP<float> op{r_, p_[0], a_, p_[1], y_, 0.0, 0.0};
                launch_tiles<true, float>(M, tm, c, op, st);
            } else {
                OpMapplyP<double> op{r_, p_[0], a_, p_[1], y_, 0.0, 0.0};
                launch_tiles<true, double>(M, tm, c, op, st);
            }
        } break;
        case 2: {
            OpSpmvAlpha op{y_, rs_, a_};
            launch_tiles<true, double>(A, tm, c, op, st);
        } break;
        case 3: {
            if (M.fp32) {
                OpMapplyS<float> op{r_, a_, s_, z_, 0.0};
                launch_tiles<true, float>(M, tm, c, op, st);
            } else {
                OpMapplyS<double> op{r_, a_, s_, z_, 0.0};
                launch_tiles<true, double>(M, tm, c, op, st);
            }
        } break;
        case 4: {
            OpSpmvOmega op{z_, s_, t_};
            launch_tiles<true, double>(A, tm, c, op, st);
        } break;
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

}  // namespace spai
