// Static SPAI construction on gfx950: one least-squares problem per column,
//
//     min || A(I_k, J_k) m_k - e_k(I_k) ||_2,   pattern(M) = pattern(A),
//
// the semantics of StaticSPAIv20 (src/spai_single_dcu.cpp:1224-2687) computed
// in one pass per column instead of the reference's seven passes over padded
// n*n1max*n2max workspaces:
//
//   gather   J_k = A(:,k)'s rows, then the (row, source) pairs of A(:,J_k)
//            straight from CSC into registers (no HBM tile);
//   I_k      a bitonic sort of the gathered rows across the lanes of the
//            column's group, duplicates collapsed by a ballot prefix count,
//            giving I_k sorted ascending (computeIShared_Symbol_SpMMv1,
//            src/spai_single_dcu.cpp:510-615) and E_k = position of k
//            (ComputeTildeE_SSPAIv10, :973-992);
//   A(I,J)   scattered once into LDS, each lane then owns rows of the tile;
//   QR       Householder in registers (the reference uses MGS,
//            QR_RShared_SSPAIv10 :1057-1121; both give the same least-squares
//            solution, checked to 1e-10 against the MGS oracle);
//   solve    Q^T e_E and back substitution (Sol_SSPAIv10 :1016-1055), x = 0
//            when E_k = -1;
//   store    M's values at A's CSC positions (modifyData4_SSPAIv10
//            :1123-1140) and at the matching CSR positions for the M-apply.
//
// Mixed precision: the Householder QR runs in FP32 and is followed by one
// FP64 corrected-semi-normal-equations sweep (r = e - A x, g = A^T r,
// R^T R d = g, x += d) before M is stored in FP32.
//
// Scheduling: the planner classifies every column by (T_k = entries gathered,
// n2_k = |J_k|) into size tiers.  balanced_columns runs every column through
// the one tier that fits the largest column, in contiguous column blocks per
// wavefront; balanced_nnz runs each tier's column list through its own
// kernel with an atomic work queue (smallest tier first, each register
// tier's queue ending with the columns the tier before it handed on).
#include "device_common.hpp"
#include "spai_build.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace spai {

namespace {

#ifndef SPAI_TIER_BLOCK
#define SPAI_TIER_BLOCK 256
#endif
constexpr int kBlock = SPAI_TIER_BLOCK;  // threads per workgroup for the column kernels

// Per-tier group shapes: W lanes per column, EPL gathered entries per lane,
// NJ = largest |J_k| the tier holds in registers, RPL tile rows per lane.
// A column is routed by T_k and |J_k|; its |I_k| is known only once the
// gathered rows are merged.  When |I_k| exceeds the tier's ROWS (< TMAX)
// the column goes to the next tier's overflow list, which that tier's
// queue takes after its own list.  The row tiles are sized for the common
// case instead of the largest T_k: the 3D 7-point stencil gathers T = 49
// entries into |I_k| = 25 rows, the 2D 5-point 25 into 13.  Narrow groups
// (8 or 16 lanes, 4 entries and 2 rows per lane) put 8 or 4 columns in a
// wavefront, so every cross-lane step serves that many columns.
template <int W_, int EPL_, int NJ_, int RPL_>
struct RegTier {
    static constexpr int W = W_;
    static constexpr int EPL = EPL_;
    static constexpr int NJ = NJ_;
    static constexpr int RPL = RPL_;
    static constexpr int ROWS = W * RPL;     // largest |I_k|
    static constexpr int TMAX = W * EPL;     // largest T_k
    static constexpr int TS = NJ + 1;        // LDS tile row stride (doubles)
    static constexpr int GROUPS = kBlock / W;
};

// Size buckets of the register tiers: a column goes to the first bucket
// whose largest T_k and |J_k| hold it.  Each bucket has a list of candidate
// group shapes (the first is the default); the autotuner times them on a
// system's own columns and keeps the fastest per bucket.  Every candidate of
// a bucket gathers exactly the bucket's T_k (TMAX) and has a lane per x entry
// in the refinement (W >= NJ).  A candidate's NJ may be below the bucket's
// largest |J_k| (5 columns for the 2D 5-point stencil, 7 for the 3D 7-point:
// the QR, the solve and the refinement then do not run over padded columns);
// a column with more than NJ columns, or with more rows than the candidate
// holds, goes to the next bucket, whose every candidate holds its rows (the
// last bucket's candidates hold its full |J_k|).  The shape changes the order
// of the group sums, so M differs between shapes in the last bits (well
// inside the parity bars).
constexpr int kRegTiers = kTunedTiers;
constexpr int kBucketT[kRegTiers] = {32, 64, 128};
constexpr int kBucketNJ[kRegTiers] = {8, 8, 16};
using Bucket0 = std::tuple<RegTier<8, 4, 8, 2>, RegTier<16, 2, 8, 1>, RegTier<8, 4, 8, 4>, RegTier<16, 2, 8, 2>,
                           RegTier<8, 4, 5, 2>>;
using Bucket1 = std::tuple<RegTier<16, 4, 8, 2>, RegTier<32, 2, 8, 1>, RegTier<16, 4, 8, 4>, RegTier<8, 8, 8, 4>,
                           RegTier<8, 8, 7, 4>, RegTier<16, 4, 7, 2>>;
using Bucket2 = std::tuple<RegTier<64, 2, 16, 2>, RegTier<32, 4, 16, 4>>;
template <int B>
using BucketOf = std::tuple_element_t<B, std::tuple<Bucket0, Bucket1, Bucket2>>;

constexpr bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
template <int B, class TR>
constexpr bool shape_ok() {
    return TR::TMAX == kBucketT[B] && TR::NJ <= kBucketNJ[B] && (B < kRegTiers - 1 || TR::NJ == kBucketNJ[B]) &&
           TR::W >= TR::NJ && pow2(TR::W) && TR::W <= 64 &&
           pow2(TR::EPL) && kBlock % TR::W == 0 && (B == 0 || TR::ROWS >= kBucketT[B - 1]) &&
           (B < kRegTiers - 1 || TR::ROWS >= TR::TMAX);
}
template <int B, size_t... I>
constexpr bool bucket_ok(std::index_sequence<I...>) {
    return (shape_ok<B, std::tuple_element_t<I, BucketOf<B>>>() && ...);
}
template <int B>
constexpr bool bucket_ok() {
    return bucket_ok<B>(std::make_index_sequence<std::tuple_size_v<BucketOf<B>>>{});
}
static_assert(bucket_ok<0>() && bucket_ok<1>() && bucket_ok<2>(),
              "a candidate must gather its bucket's T_k, the last bucket's must hold its |J_k|, and an overflowed "
              "column must fit the next bucket's rows");

// f(shape) for candidate c of bucket B (c is range-checked by the caller)
template <int B, class F, size_t... I>
void with_shape(int c, F &&f, std::index_sequence<I...>) {
    ((c == (int)I ? (f(std::tuple_element_t<I, BucketOf<B>>{}), 0) : 0), ...);
}
template <int B, class F>
void with_shape(int c, F &&f) {
    with_shape<B>(c, f, std::make_index_sequence<std::tuple_size_v<BucketOf<B>>>{});
}
template <int B, size_t... I>
TileShape shape_info(int c, std::index_sequence<I...>) {
    TileShape r{};
    ((c == (int)I ? (r = TileShape{std::tuple_element_t<I, BucketOf<B>>::W, std::tuple_element_t<I, BucketOf<B>>::EPL,
                                   std::tuple_element_t<I, BucketOf<B>>::NJ, std::tuple_element_t<I, BucketOf<B>>::RPL},
                     0)
                  : 0),
     ...);
    return r;
}

// LDS of one column group, in bytes: one area that holds, one after the
// other, what each step needs:
//   gather   the entry table (source position and J index of every
//            gathered entry), read into registers before the tile is
//            cleared;
//   scatter  the tile A(I_k, J_k), ROWS x TS in the QR's type (FP32 in mixed
//            mode), dead once its rows are in registers;
//   solve    mixed mode's refinement data (the entries' FP64 values, their
//            J indices and each row's first entry, in sorted order), then R,
//            y and 1 / R_cc.
// Sharing the area keeps the group to the tile's size, so the 3D shapes fit
// four workgroups per CU instead of three.
template <class TR, class QT, bool REFINE>
struct TierLds {
    static constexpr int al8(int b) { return (b + 7) / 8 * 8; }
    static constexpr int al4(int b) { return (b + 3) / 4 * 4; }
    static constexpr int TILE = TR::ROWS * TR::TS * (int)sizeof(QT);
    static constexpr int TJ_OFF = TR::TMAX * 4;  // table: int source position, then byte J index
    static constexpr int TABLE = TJ_OFF + TR::TMAX;
    static constexpr int JS_OFF = TR::TMAX * 8;  // refinement: double value, byte J index, int row start
    static constexpr int RST_OFF = al4(JS_OFF + TR::TMAX);
    static constexpr int REF = REFINE ? RST_OFF + (TR::ROWS + 1) * 4 : 0;
    static constexpr int RS_OFF = al8(REF);
    static constexpr int YS_OFF = al8(RS_OFF + TR::NJ * TR::NJ * (int)sizeof(QT));
    static constexpr int RI_OFF = YS_OFF + TR::NJ * 8;  // 1 / R_cc
    static constexpr int END = RI_OFF + TR::NJ * 8;
    static constexpr int AREA = TILE > TABLE ? (TILE > END ? TILE : END) : (TABLE > END ? TABLE : END);
    static constexpr int DOUBLES = al8(AREA) / 8;
};

// The LDS tier: one wavefront per column, tile in LDS, FP64 arithmetic.
constexpr int kLdsTierT = 1024;        // largest T_k
constexpr int kLdsTierEPL = kLdsTierT / 64;
constexpr int kLdsTierNJ = 64;         // largest |J_k|
constexpr int kLdsTierDoubles = 7168;  // tile + R + y (56 KB)

struct ColumnArgs {
    const int *colptr;   // A (CSC): the columns gathered for A(I_k, J_k)
    const int *rowidx;
    const double *val;
    const int *pcolptr;  // pattern of M (CSC): J_k = prowidx[pcolptr[k] .. pcolptr[k+1])
    const int *prowidx;
    const int *perm;     // pattern position -> position of M's solver layout
    void *m_csc;      // MT[nnz]
    void *m_csr;      // MT[nnz]
    int *n1_out;      // optional per-column |I_k|
    // scheduling
    const int *list;  // dynamic: columns of this tier; static: nullptr
    const int *count; // dynamic: number of columns in list
    int *head;        // dynamic: queue head
    int col_begin, col_end;  // static range
    int *ovf_list;    // columns with |I_k| above the tier's rows go to the next tier's list
    int *ovf_count;
    const int *ovf_in_list;   // dynamic: columns handed on by the previous tier, after list
    const int *ovf_in_count;
};

// 64-bit shuffle helpers on top of the 32-bit cross-lane permute.
template <int W>
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
    lo = __shfl_xor(lo, m, W);
    hi = __shfl_xor(hi, m, W);
    return ((uint64_t)hi << 32) | lo;
}

// Locate, for gathered entry e (0 <= e < T), the J index j with
// off_j <= e < off_j + c_j.  off is the exclusive prefix held by lane j;
// the search runs the same number of steps on every lane.
template <int W, int NJ>
__device__ __forceinline__ int locate_j(int e, int off, int n2) {
    int lo = 0;
#pragma unroll
    for (int step = NJ / 2; step >= 1; step >>= 1) {
        int cand = lo + step;
        int oc = __shfl(off, cand < W ? cand : W - 1, W);
        if (cand < n2 && oc <= e) lo = cand;
    }
    return lo;
}

template <int W>
__device__ __forceinline__ uint32_t shfl_xor_key(uint32_t v, int m) {
    return (uint32_t)__shfl_xor((unsigned)v, m, W);
}
template <int W>
__device__ __forceinline__ uint64_t shfl_xor_key(uint64_t v, int m) {
    return shfl_xor_u64<W>(v, m);
}

// (row, entry) sort keys of the register tiers.  With 32-bit keys the entry
// index takes EB = log2(T_max) low bits and the row the rest, which needs
// n < 2^(32 - EB) - 1 (the all-ones key marks an empty slot); half the compare
// and exchange work of 64-bit keys, same order.
template <bool K32, int EB>
struct SortKey {
    using T = typename std::conditional<K32, uint32_t, uint64_t>::type;
    static constexpr T kNone = ~T(0);
    static constexpr int kShift = K32 ? EB : 32;
    __device__ static T make(int row, int e) { return ((T)(unsigned)row << kShift) | (T)(unsigned)e; }
    __device__ static int row(T k) { return (int)(k >> kShift); }
    __device__ static int ent(T k) { return (int)(k & ((T(1) << kShift) - 1)); }
};
constexpr int log2_exact(int v) { return v <= 1 ? 0 : 1 + log2_exact(v / 2); }

// Bitonic sort (ascending) of W*EPL keys, lane l holding elements
// l*EPL .. l*EPL+EPL-1.
template <int W, int EPL, class KT>
__device__ __forceinline__ void bitonic_sort(KT (&key)[EPL], int lane) {
    constexpr int N = W * EPL;
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= EPL) {
                const int ls = stride / EPL;
#pragma unroll
                for (int i = 0; i < EPL; ++i) {
                    KT pk = shfl_xor_key<W>(key[i], ls);
                    int e = lane * EPL + i;
                    bool up = (e & size) == 0;
                    bool lower = (e & stride) == 0;
                    bool take_min = (lower == up);
                    KT mn = pk < key[i] ? pk : key[i];
                    KT mx = pk < key[i] ? key[i] : pk;
                    key[i] = take_min ? mn : mx;
                }
            } else {
#pragma unroll
                for (int i = 0; i < EPL; ++i) {
                    if ((i & stride) == 0) {
                        const int i2 = i | stride;
                        int e = lane * EPL + i;
                        bool up = (e & size) == 0;
                        KT a = key[i], b = key[i2];
                        bool sw = up ? (a > b) : (a < b);
                        key[i] = sw ? b : a;
                        key[i2] = sw ? a : b;
                    }
                }
            }
        }
    }
}

// Reciprocal and square root of the register-tier Householder: the hardware
// instructions in FP32 (mixed mode, whose FP64 refinement sweep corrects the
// FP32 solution), IEEE operations in FP64.
__device__ __forceinline__ float qr_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double qr_rcp(double v) { return 1.0 / v; }
__device__ __forceinline__ float qr_sqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ double qr_sqrt(double v) { return sqrt(v); }

template <class MT>
__device__ __forceinline__ void store_m(const ColumnArgs &a, int pos, double x) {
    MT *csc = (MT *)a.m_csc;
    MT *csr = (MT *)a.m_csr;
    csc[pos] = (MT)x;
    csr[a.perm[pos]] = (MT)x;
}

// ---------------------------------------------------------------------------
// Register tier: one group of W lanes per column.
// ---------------------------------------------------------------------------
// 1: a per-column LDS table maps each gathered entry to its source position
// and J index (one read each); 0: binary search over the J offsets with
// cross-lane permutes, twice per entry.
#ifndef SPAI_TIER_ENTRY_TABLE
#define SPAI_TIER_ENTRY_TABLE 1
#endif
// 1: a group keeps its previous column's sorted (row, entry) keys; when the
// new column gathers the same number of entries and each entry's row lies
// at the same offset from k as before (interior stencil columns), the new
// sorted keys are the old ones shifted by the column difference, and a
// wavefront whose groups all find that skips the sort.  M is bit-identical
// either way; the batch's register tiers 1.358 -> 1.280 ms.
#ifndef SPAI_TIER_SORT_REUSE
#define SPAI_TIER_SORT_REUSE 1
#endif
template <class KT, int EPL>
struct TierCarry {
    KT key[EPL];  // the previous column's sorted keys
    int k = -1;   // its column
    int T = -1;   // its gathered entries (-1: none yet)
};
// Column k with J_k = prowidx[c0 .. c0 + n2) and, in lane j < n2, jr = the
// row index J_k[j] (both loaded by the caller, see spai_reg_tier_kernel).
template <class TR, class QT, bool REFINE, class MT, bool K32>
__device__ __forceinline__ void solve_column_reg(const ColumnArgs &a, int k, int c0, int n2, int jr, int lane,
                                                 double *lds,
                                                 TierCarry<typename SortKey<K32, log2_exact(TR::TMAX)>::T, TR::EPL> &cy) {
    constexpr int W = TR::W, EPL = TR::EPL, NJ = TR::NJ, TS = TR::TS;
    using SK = SortKey<K32, log2_exact(TR::TMAX)>;
    using KT = typename SK::T;
    constexpr int RPL = TR::RPL;
    using L = TierLds<TR, QT, REFINE>;
    QT *tile = (QT *)lds;                                     // ROWS x TS
    QT *Rs = (QT *)((char *)lds + L::RS_OFF);                 // NJ x NJ
    double *ys = (double *)((char *)lds + L::YS_OFF);         // NJ (then x in the refinement)
    double *rinv = (double *)((char *)lds + L::RI_OFF);       // NJ: 1 / R_cc
    double *vslot = lds;                                      // refinement: entry values (sorted order)
    unsigned char *jslot = (unsigned char *)lds + L::JS_OFF;  //   their J indices
    int *rstart = (int *)((char *)lds + L::RST_OFF);          //   first entry of each row, n1 + 1

    if (n2 <= 0) return;
    if (n2 > NJ) {  // a shape narrower than its bucket: the next bucket takes it
        if (lane == 0) a.ovf_list[atomicAdd(a.ovf_count, 1)] = k;
        return;
    }

    // the extent of each column A(:, J_k[j])
    int cs = 0, cl = 0;
    if (lane < n2) {
        cs = a.colptr[jr];
        cl = a.colptr[jr + 1] - cs;
    }
    const int incl = group_inclusive_scan<W>(cl, lane);
    const int T = __shfl(incl, W - 1, W);
    const int off = incl - cl;
#if SPAI_TIER_ENTRY_TABLE
    // entry e -> its position in A's arrays and its J index, written by the
    // lane that owns the column (the area is free until the tile is cleared)
    int *tsrc = (int *)lds;
    unsigned char *tj = (unsigned char *)lds + L::TJ_OFF;
    static_assert(TR::NJ <= 255, "J indices are bytes");
    for (int q = 0; q < cl; ++q) {
        tsrc[off + q] = cs + q;
        tj[off + q] = (unsigned char)lane;
    }
    wave_lds_fence();
#endif

    // gather (row, entry) keys, blocked layout e = lane*EPL + i
    KT key[EPL];
    bool reuse = false;
#if SPAI_TIER_SORT_REUSE && SPAI_TIER_ENTRY_TABLE
    if (cy.T == T) {
        // the previous column's sorted order: entry e of this column in the
        // slot where entry e was, its row checked against the old row + dk
        const int dk = k - cy.k;
        reuse = true;
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            const KT pk = cy.key[i];
            const bool valid = pk != SK::kNone;
            const int e = valid ? SK::ent(pk) : 0;
            const int row = valid ? a.rowidx[tsrc[e]] : 0;
            key[i] = valid ? SK::make(row, e) : SK::kNone;
            reuse = reuse && (!valid || row == SK::row(pk) + dk);
        }
    } else
#endif
    {
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            int e = lane * EPL + i;
            key[i] = SK::kNone;
#if SPAI_TIER_ENTRY_TABLE
            if (e < T) key[i] = SK::make(a.rowidx[tsrc[e]], e);
#else
            int j = locate_j<W, NJ>(e, off, n2);
            int src = __shfl(cs, j, W) + (e - __shfl(off, j, W));
            if (e < T) key[i] = SK::make(a.rowidx[src], e);
#endif
        }
    }
    // either order holds the column's T keys; the wavefront sorts unless
    // every group of it found its previous order
    if (!__all(reuse)) bitonic_sort<W, EPL>(key, lane);
#if SPAI_TIER_SORT_REUSE && SPAI_TIER_ENTRY_TABLE
#pragma unroll
    for (int i = 0; i < EPL; ++i) cy.key[i] = key[i];
    cy.k = k;
    cy.T = T;
#endif

    // collapse duplicate rows: pos = (#first occurrences up to here) - 1
    int rowv[EPL];
    bool first[EPL];
    int cnt = 0;
    int prev = __shfl_up(SK::row(key[EPL - 1]), 1, W);
    if (lane == 0) prev = -1;
#pragma unroll
    for (int i = 0; i < EPL; ++i) {
        bool valid = key[i] != SK::kNone;
        rowv[i] = SK::row(key[i]);
        first[i] = valid && (rowv[i] != prev);
        prev = rowv[i];
        cnt += first[i] ? 1 : 0;
    }
    const int cincl = group_inclusive_scan<W>(cnt, lane);
    const int n1 = __shfl(cincl, W - 1, W);
    int run = cincl - cnt;
    int pos[EPL];
    int e_pos = -1;
#pragma unroll
    for (int i = 0; i < EPL; ++i) {
        run += first[i] ? 1 : 0;
        pos[i] = run - 1;
        if (key[i] != SK::kNone && rowv[i] == k) e_pos = pos[i];
    }
    const int E = group_max<W>(e_pos);
    if (TR::ROWS < TR::TMAX && n1 > TR::ROWS) {  // the next tier takes it
        if (lane == 0) a.ovf_list[atomicAdd(a.ovf_count, 1)] = k;
        return;
    }

    // the entries' FP64 values and J indices into registers (the entry
    // table dies here; the values also feed the refinement)
    double vv[EPL];
    int jv[EPL];
#pragma unroll
    for (int i = 0; i < EPL; ++i) {
        // shuffles run on every lane of the group; only the loads are masked
        const bool valid = key[i] != SK::kNone;
        const int e = valid ? SK::ent(key[i]) : 0;
#if SPAI_TIER_ENTRY_TABLE
        const int j = valid ? (int)tj[e] : 0;
        const int src = valid ? tsrc[e] : 0;
#else
        const int j = locate_j<W, NJ>(e, off, n2);
        const int src = __shfl(cs, j, W) + (e - __shfl(off, j, W));
#endif
        vv[i] = valid ? a.val[src] : 0.0;
        jv[i] = j;
    }
    wave_lds_fence();
    // zero the rows this lane owns, then scatter A(I_k, J_k) into LDS
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        int rho = lane + W * r;
        if (rho < n1) {
#pragma unroll
            for (int c = 0; c < NJ; ++c) tile[rho * TS + c] = (QT)0;
        }
    }
    wave_lds_fence();
#pragma unroll
    for (int i = 0; i < EPL; ++i)
        if (key[i] != SK::kNone) tile[pos[i] * TS + jv[i]] = (QT)vv[i];
    wave_lds_fence();

    // rows rho = lane + W*r into registers
    QT A[RPL][NJ];
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        int rho = lane + W * r;
#pragma unroll
        for (int c = 0; c < NJ; ++c) A[r][c] = (rho < n1 && c < n2) ? tile[rho * TS + c] : (QT)0;
    }
    if constexpr (REFINE) {
        // the entries in sorted (row, then column) order over the dead tile
        wave_lds_fence();
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            const int slot = lane * EPL + i;
            if (key[i] != SK::kNone) {
                vslot[slot] = vv[i];
                jslot[slot] = (unsigned char)jv[i];
                if (first[i]) rstart[pos[i]] = slot;
            }
        }
        if (lane == 0) rstart[n1] = T;
    }

    // Householder QR: H_j = I - tau_j v_j v_j^T, v_j[j] = 1 (v below the
    // diagonal kept in A), R on and above the diagonal.
    QT tau[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        tau[j] = 0;
        if (j < n2) {
            QT alpha = __shfl(A[j / W][j], j % W, W);
            QT sl = 0;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                int rho = lane + W * r;
                sl += (rho > j) ? A[r][j] * A[r][j] : (QT)0;
            }
            QT sigma = group_allsum<W>(sl);
            QT beta = alpha, t = 0, scale = 0;
            if (sigma != (QT)0) {
                QT nrm = qr_sqrt(alpha * alpha + sigma);
                beta = alpha >= (QT)0 ? -nrm : nrm;
                t = (beta - alpha) * qr_rcp(beta);
                scale = qr_rcp(alpha - beta);
            }
            tau[j] = t;
            QT v[RPL];
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                int rho = lane + W * r;
                v[r] = (rho > j) ? A[r][j] * scale : (rho == j ? (QT)1 : (QT)0);
                if (rho > j) A[r][j] = v[r];
                if (rho == j) A[r][j] = beta;
            }
            QT w[NJ];
#pragma unroll
            for (int c = j + 1; c < NJ; ++c) {
                QT s = 0;
#pragma unroll
                for (int r = 0; r < RPL; ++r) s += v[r] * A[r][c];
                w[c] = s;
            }
#pragma unroll
            for (int c = j + 1; c < NJ; ++c) w[c] = group_allsum<W>(w[c]);
#pragma unroll
            for (int c = j + 1; c < NJ; ++c) {
                if (c < n2) {
                    QT wc = w[c] * t;
#pragma unroll
                    for (int r = 0; r < RPL; ++r) A[r][c] -= wc * v[r];
                }
            }
        }
    }

    // y = Q^T e_E
    QT y[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) y[r] = (lane + W * r == E) ? (QT)1 : (QT)0;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        if (j < n2) {
            QT s = 0;
            QT v[RPL];
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                int rho = lane + W * r;
                v[r] = (rho > j) ? A[r][j] : (rho == j ? (QT)1 : (QT)0);
                s += v[r] * y[r];
            }
            s = group_allsum<W>(s) * tau[j];
#pragma unroll
            for (int r = 0; r < RPL; ++r) y[r] -= s * v[r];
        }
    }

    // R and y to LDS (rows < n2 only), 1 / R_cc once per column (lane c),
    // then a uniform back substitution
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        int rho = lane + W * r;
        if (rho < n2) {
#pragma unroll
            for (int c = 0; c < NJ; ++c) Rs[rho * NJ + c] = A[r][c];
            ys[rho] = (double)y[r];
        }
    }
    if (lane < n2) {  // W >= NJ: row c < n2 of R is lane c's first row
        QT d = 0;
#pragma unroll
        for (int c = 0; c < NJ; ++c)
            if (c == lane) d = A[0][c];
        rinv[lane] = 1.0 / (double)d;
    }
    wave_lds_fence();
    double x[NJ];
#pragma unroll
    for (int c = NJ - 1; c >= 0; --c) {
        x[c] = 0.0;
        if (c < n2) {
            double t = ys[c];
#pragma unroll
            for (int c2 = c + 1; c2 < NJ; ++c2)
                if (c2 < n2) t -= (double)Rs[c * NJ + c2] * x[c2];
            x[c] = t * rinv[c];
        }
    }
    if (E < 0) {
#pragma unroll
        for (int c = 0; c < NJ; ++c) x[c] = 0.0;
    }

    if constexpr (REFINE) {
        if (E >= 0) {
            // one FP64 corrected-semi-normal-equations sweep from the column's
            // gathered FP64 entries: a row's entries are visited in column
            // order, as a dense row of the tile would be
            double *xs = ys;  // y is dead after the back substitution
            static_assert(TR::W >= TR::NJ, "lane c holds x[c]");
            wave_lds_fence();
            {
                double xv = 0.0;
#pragma unroll
                for (int c = 0; c < NJ; ++c)
                    if (lane == c) xv = x[c];
                if (lane < n2) xs[lane] = xv;
            }
            wave_lds_fence();
            double res[RPL];
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int rho = lane + W * r;
                double t = (rho == E) ? 1.0 : 0.0;
                if (rho < n1) {
                    const int q1 = rstart[rho + 1];
                    for (int q = rstart[rho]; q < q1; ++q) t = fma(-vslot[q], xs[jslot[q]], t);
                }
                res[r] = (rho < n1) ? t : 0.0;
            }
            double g[NJ];
#pragma unroll
            for (int c = 0; c < NJ; ++c) g[c] = 0.0;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int rho = lane + W * r;
                if (rho < n1) {
                    const int q1 = rstart[rho + 1];
                    for (int q = rstart[rho]; q < q1; ++q) {
                        const int cq = jslot[q];
                        const double v = vslot[q];
#pragma unroll
                        for (int c = 0; c < NJ; ++c) g[c] = fma(c == cq ? v : 0.0, res[r], g[c]);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NJ; ++c) g[c] = group_allsum<W>(g[c]);
            // R^T u = g (forward), R d = u (backward), x += d
#pragma unroll
            for (int c = 0; c < NJ; ++c) {
                if (c < n2) {
                    double t = g[c];
#pragma unroll
                    for (int c2 = 0; c2 < c; ++c2) t -= (double)Rs[c2 * NJ + c] * g[c2];
                    g[c] = t * rinv[c];
                }
            }
#pragma unroll
            for (int c = NJ - 1; c >= 0; --c) {
                if (c < n2) {
                    double t = g[c];
#pragma unroll
                    for (int c2 = c + 1; c2 < NJ; ++c2)
                        if (c2 < n2) t -= (double)Rs[c * NJ + c2] * g[c2];
                    g[c] = t * rinv[c];
                    x[c] += g[c];
                }
            }
        }
    }

    // lane c writes x[c]
    double mine = 0.0;
#pragma unroll
    for (int c = 0; c < NJ; ++c)
        if (lane == c) mine = x[c];
    if (lane < n2) store_m<MT>(a, c0 + lane, mine);
    if (a.n1_out && lane == 0) a.n1_out[k] = n1;
    wave_lds_fence();  // LDS reused by this group's next column
}

#ifndef SPAI_TIER_GRAB
#define SPAI_TIER_GRAB 4
#endif
// optional occupancy floor (waves per SIMD) for the register allocator
#ifndef SPAI_TIER_MIN_WAVES
#define SPAI_TIER_MIN_WAVES 0
#endif
// Without it, the FP32 shapes of at most 8 columns ask for 4 wavefronts per
// SIMD (128 registers): the 3D 7-column shape otherwise lands 2 registers
// above that and runs 3.
template <class TR, class QT>
constexpr int tier_min_waves() {
    return SPAI_TIER_MIN_WAVES > 0 ? SPAI_TIER_MIN_WAVES : (sizeof(QT) == 4 && TR::NJ <= 8 ? 4 : 1);
}
#define SPAI_TIER_OCCUPANCY __attribute__((amdgpu_waves_per_eu(tier_min_waves<TR, QT>())))
// A wavefront takes its columns in batches of CH per group (G * CH per
// wavefront).  The dependent index loads that precede a column's gather (the
// list entry k, then pcolptr[k], then J_k = prowidx[...]) are issued for the
// whole batch at once, one round trip per level for CH columns, and parked in
// a small per-group LDS area; the columns then run one after another.
//
// Dynamic schedule: most of a tier's list is split statically into equal
// contiguous shares, one per wavefront of the persistent grid (columns of one
// bucket cost about the same), and only the rest goes through the atomic
// queue, so the queue head, a single address every wavefront increments, sees
// few atomics.
#ifndef SPAI_TIER_STATIC_SHARE
#define SPAI_TIER_STATIC_SHARE 3  // quarters of the list split statically
#endif
template <class TR, class QT, bool REFINE, class MT, bool DYNAMIC, bool K32>
__global__ __launch_bounds__(kBlock) SPAI_TIER_OCCUPANCY void spai_reg_tier_kernel(ColumnArgs a) {
    constexpr int W = TR::W;
    constexpr int G = 64 / W;    // groups per wavefront
    constexpr int CH = SPAI_TIER_GRAB;  // columns per group per batch
    constexpr int HEAD = CH * (W + 3);  // per group: k, c0, n2 and each lane's J_k entry, per batch column
    __shared__ double lds_all[TR::GROUPS * TierLds<TR, QT, REFINE>::DOUBLES];
    __shared__ int head_all[TR::GROUPS * HEAD];
    const int tid = threadIdx.x;
    const int lane = tid % W;
    const int grp = tid / W;
    const int wgrp = (tid % 64) / W;  // group index inside the wavefront
    double *lds = lds_all + grp * TierLds<TR, QT, REFINE>::DOUBLES;
    int *hk = head_all + grp * HEAD, *hc0 = hk + CH, *hn2 = hc0 + CH, *hjr = hn2 + CH;
    TierCarry<typename SortKey<K32, log2_exact(TR::TMAX)>::T, TR::EPL> cy;

    // one batch: column_of(it) gives the group's it-th column (or -1)
    auto batch = [&](auto column_of) {
        int kk[CH], c0v[CH], n2v[CH];
#pragma unroll
        for (int it = 0; it < CH; ++it) kk[it] = column_of(it);
#pragma unroll
        for (int it = 0; it < CH; ++it) {
            c0v[it] = kk[it] >= 0 ? a.pcolptr[kk[it]] : 0;
            n2v[it] = kk[it] >= 0 ? a.pcolptr[kk[it] + 1] - c0v[it] : 0;
        }
#pragma unroll
        for (int it = 0; it < CH; ++it) {
            // lanes past the tier's NJ load nothing (such a column moves on)
            hjr[it * W + lane] = (lane < n2v[it] && n2v[it] <= TR::NJ) ? a.prowidx[c0v[it] + lane] : 0;
            if (lane == 0) {
                hk[it] = kk[it];
                hc0[it] = c0v[it];
                hn2[it] = n2v[it];
            }
        }
        wave_lds_fence();
        for (int it = 0; it < CH; ++it) {
            const int k = hk[it];
            if (k >= 0)
                solve_column_reg<TR, QT, REFINE, MT, K32>(a, k, hc0[it], hn2[it], hjr[it * W + lane], lane, lds, cy);
        }
        wave_lds_fence();  // the head area is rewritten by the next batch
    };

    const int nwaves = gridDim.x * (kBlock / 64);
    const int wave = blockIdx.x * (kBlock / 64) + tid / 64;
    if (DYNAMIC) {
        // the tier's list, then the columns the previous tier handed on
        const int count = *a.count;
        const int total = count + (a.ovf_in_count ? *a.ovf_in_count : 0);
        if (total == 0) return;
        // static shares, whole batches each; then the rest from the queue
        // (its head counts from the end of the shares).  One loop, so that
        // the column code is inlined once.
        const int share = (int)((long long)total * SPAI_TIER_STATIC_SHARE / 4 / nwaves / (G * CH)) * (G * CH);
        const int rest0 = share * nwaves;
        int base = wave * share;
        const int share_end = base + share;
        for (;;) {
            if (base >= share_end) {
                int b = 0;
                if ((tid & 63) == 0) b = atomicAdd(a.head, G * CH);
                base = __shfl(b, 0, 64) + rest0;
                if (base >= total) break;
            }
            batch([&](int it) {
                const int idx = base + it * G + wgrp;
                return idx < total ? (idx < count ? a.list[idx] : a.ovf_in_list[idx - count]) : -1;
            });
            base = base + G * CH < share_end ? base + G * CH : share_end;
        }
    } else {
        const int ncols = a.col_end - a.col_begin;
        const int chunk = (ncols + nwaves - 1) / nwaves;
        const int start = a.col_begin + wave * chunk;
        const int stop = min(start + chunk, a.col_end);
        for (int base = start; base < stop; base += G * CH)
            batch([&](int it) {
                const int k = base + it * G + wgrp;
                return k < stop ? k : -1;
            });
    }
}

// ---------------------------------------------------------------------------
// LDS tier: one wavefront per column, tile n1 x n2 (row-major) in LDS, FP64.
// ---------------------------------------------------------------------------
template <class MT>
__device__ void solve_column_lds(const ColumnArgs &a, int k, int lane, double *lds, int *ilds) {
    constexpr int W = 64, EPL = kLdsTierEPL, NJ = kLdsTierNJ;
    const int c0 = a.pcolptr[k];
    const int n2 = a.pcolptr[k + 1] - c0;
    if (n2 <= 0) return;
    int *offs = ilds;        // NJ
    int *starts = ilds + NJ; // NJ
    int cs = 0, cl = 0;
    if (lane < n2) {
        int jr = a.prowidx[c0 + lane];
        cs = a.colptr[jr];
        cl = a.colptr[jr + 1] - cs;
    }
    const int incl = group_inclusive_scan<W>(cl, lane);
    const int T = __shfl(incl, W - 1, W);
    const int off = incl - cl;
    if (lane < n2) {
        offs[lane] = off;
        starts[lane] = cs;
    }
    wave_lds_fence();
    auto find_j = [&](int e) {
        int lo = 0, hi = n2 - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= e) lo = mid; else hi = mid - 1;
        }
        return lo;
    };
    uint64_t key[EPL];
#pragma unroll
    for (int i = 0; i < EPL; ++i) {
        int e = lane * EPL + i;
        key[i] = ~0ull;
        if (e < T) {
            int j = find_j(e);
            key[i] = ((uint64_t)(unsigned)a.rowidx[starts[j] + e - offs[j]] << 32) | (unsigned)e;
        }
    }
    bitonic_sort<W, EPL>(key, lane);
    int cnt = 0;
    int prev = __shfl_up((int)(key[EPL - 1] >> 32), 1, W);
    if (lane == 0) prev = -1;
    bool first[EPL];
#pragma unroll
    for (int i = 0; i < EPL; ++i) {
        int rv = (int)(key[i] >> 32);
        first[i] = key[i] != ~0ull && rv != prev;
        prev = rv;
        cnt += first[i] ? 1 : 0;
    }
    const int cincl = group_inclusive_scan<W>(cnt, lane);
    const int n1 = __shfl(cincl, W - 1, W);
    double *tile = lds;                       // n1 x n2
    double *Rs = lds + (size_t)n1 * n2;       // n2 x n2
    double *ys = Rs + (size_t)n2 * n2;        // n2
    for (int idx = lane; idx < n1 * n2; idx += W) tile[idx] = 0.0;
    wave_lds_fence();
    int run = cincl - cnt, e_pos = -1;
#pragma unroll
    for (int i = 0; i < EPL; ++i) {
        run += first[i] ? 1 : 0;
        if (key[i] != ~0ull) {
            int e = (int)(unsigned)key[i];
            int j = find_j(e);
            int src = starts[j] + e - offs[j];
            tile[(run - 1) * n2 + j] = a.val[src];
            if ((int)(key[i] >> 32) == k) e_pos = run - 1;
        }
    }
    const int E = group_max<W>(e_pos);
    wave_lds_fence();
    // Householder on the LDS tile; lane owns rows lane, lane+64, ...
    double taus_last = 0.0;
    for (int j = 0; j < n2; ++j) {
        double alpha = tile[j * n2 + j];
        double sl = 0.0;
        for (int rho = lane; rho < n1; rho += W)
            if (rho > j) sl += tile[rho * n2 + j] * tile[rho * n2 + j];
        double sigma = group_sum<W>(sl);
        double beta = alpha, t = 0.0, scale = 0.0;
        if (sigma != 0.0) {
            double nrm = sqrt(alpha * alpha + sigma);
            beta = alpha >= 0.0 ? -nrm : nrm;
            t = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        wave_lds_fence();
        for (int rho = lane; rho < n1; rho += W) {
            if (rho > j) tile[rho * n2 + j] *= scale;
            if (rho == j) tile[rho * n2 + j] = beta;
        }
        if (lane == 0) ys[j] = t;  // tau_j parked in ys until Q^T e is formed
        wave_lds_fence();
        for (int c = j + 1; c < n2; ++c) {
            double s = 0.0;
            for (int rho = lane; rho < n1; rho += W) {
                double v = rho > j ? tile[rho * n2 + j] : (rho == j ? 1.0 : 0.0);
                s += v * tile[rho * n2 + c];
            }
            s = group_sum<W>(s) * t;
            for (int rho = lane; rho < n1; rho += W) {
                double v = rho > j ? tile[rho * n2 + j] : (rho == j ? 1.0 : 0.0);
                if (rho >= j) tile[rho * n2 + c] -= s * v;
            }
        }
        wave_lds_fence();
        taus_last = t;
    }
    (void)taus_last;
    // y = Q^T e_E kept distributed: lane holds y for rows lane + 64 i
    constexpr int MAXR = 16;  // rows per lane for n1 <= 1024
    double y[MAXR];
#pragma unroll
    for (int i = 0; i < MAXR; ++i) y[i] = (lane + W * i == E) ? 1.0 : 0.0;
    for (int j = 0; j < n2; ++j) {
        double t = ys[j];
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < MAXR; ++i) {
            int rho = lane + W * i;
            if (rho < n1 && rho >= j) {
                double v = rho > j ? tile[rho * n2 + j] : 1.0;
                s += v * y[i];
            }
        }
        s = group_sum<W>(s) * t;
#pragma unroll
        for (int i = 0; i < MAXR; ++i) {
            int rho = lane + W * i;
            if (rho < n1 && rho >= j) {
                double v = rho > j ? tile[rho * n2 + j] : 1.0;
                y[i] -= s * v;
            }
        }
    }
    wave_lds_fence();
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
        int rho = lane + W * i;
        if (rho < n2) ys[rho] = y[i];
    }
    wave_lds_fence();
    // back substitution: lane 0 solves, then shares through LDS
    if (lane == 0) {
        for (int c = n2 - 1; c >= 0; --c) {
            double t = ys[c];
            for (int c2 = c + 1; c2 < n2; ++c2) t -= tile[c * n2 + c2] * ys[c2];
            ys[c] = (E < 0) ? 0.0 : t / tile[c * n2 + c];
        }
    }
    wave_lds_fence();
    for (int c = lane; c < n2; c += W) store_m<MT>(a, c0 + c, ys[c]);
    if (a.n1_out && lane == 0) a.n1_out[k] = n1;
    wave_lds_fence();
    (void)Rs;
}

template <class MT, bool DYNAMIC>
__global__ __launch_bounds__(64) void spai_lds_tier_kernel(ColumnArgs a) {
    __shared__ double lds[kLdsTierDoubles];
    __shared__ int ilds[2 * kLdsTierNJ];
    const int lane = threadIdx.x;
    if (DYNAMIC) {
        const int count = *a.count;
        for (;;) {
            int idx = 0;
            if (lane == 0) idx = atomicAdd(a.head, 1);
            idx = __shfl(idx, 0, 64);
            if (idx >= count) break;
            solve_column_lds<MT>(a, a.list[idx], lane, lds, ilds);
        }
    } else {
        for (int k = a.col_begin + blockIdx.x; k < a.col_end; k += gridDim.x)
            solve_column_lds<MT>(a, k, lane, lds, ilds);
    }
}

// ---------------------------------------------------------------------------
// Global tier: one 512-thread workgroup per very wide column, every buffer in
// a per-workgroup scratch slot in global memory, FP64 in both precision modes.
// ---------------------------------------------------------------------------
constexpr int kGlobThreads = 512;
constexpr int kGlobCols = 32;                           // columns per pass
constexpr int kGlobRowGroups = kGlobThreads / kGlobCols;  // 16 row groups
constexpr int kGlobNB = 16;                             // panel width of the blocked QR
constexpr int kGlobChunk = 128;                         // rows of V staged in LDS at a time
static_assert(kGlobNB <= kGlobRowGroups, "one row group per reflector when forming T^T W");

struct GlobalScratch {
    double *tile;       // slots * tile_cap: max(n1, n2) x (n2 + 1), row-major (stride n2 + 1)
    double *vec;        // slots * vec_cap: x (n2)
    uint64_t *keys;     // slots * key_cap: sorted (row, entry) keys
    int *ints;          // slots * int_cap: positions (key_cap + 1), J offsets (n2 + 1)
    long long tile_cap, vec_cap, key_cap, int_cap;
    int n1_cap, n2_cap;
};

__device__ double glob_block_sum(double v, double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kGlobThreads / 64; ++i) t += red[i];
    return t;
}

// Exclusive scan of cnt(i), i < m, into out[0..m] (out[m] = total); returns total.
template <class F>
__device__ int glob_block_scan(int m, F cnt, int *out, int *ired) {
    int carry = 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int base = 0; base < m; base += kGlobThreads) {
        const int i = base + (int)threadIdx.x;
        const int v = i < m ? cnt(i) : 0;
        int incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        __syncthreads();
        if (lane == 63) ired[w] = incl;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int q = 0; q < kGlobThreads / 64; ++q) {
            before += q < w ? ired[q] : 0;
            total += ired[q];
        }
        if (i < m) out[i] = carry + before + incl - v;
        carry += total;
    }
    __syncthreads();
    if (threadIdx.x == 0) out[m] = carry;
    __syncthreads();
    return carry;
}

__device__ __forceinline__ int glob_find_j(const int *offs, int n2, int e) {
    int lo = 0, hi = n2 - 1;  // largest j with offs[j] <= e
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Sums N values across the workgroup (each thread's v[q], q < cnt); every
// thread receives the totals in v.  Uses red[64 .. 64 + 8 N).
template <int N>
__device__ __forceinline__ void glob_block_sum_vec(double (&v)[N], int cnt, double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        if (q < cnt) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        }
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q)
            if (q < cnt) red[64 + q * 8 + w] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) {
        double t = 0.0;
        if (q < cnt) {
#pragma unroll
            for (int i = 0; i < kGlobThreads / 64; ++i) t += red[64 + q * 8 + i];
        }
        v[q] = t;
    }
    __syncthreads();
}

// Blocked Householder QR of the m x (n + 1) row-major tile (leading dimension
// ld), reflectors from the first n columns, applied to all n + 1: on return
// R is on and above the diagonal of the first n columns and column n holds
// Q^T times its input.  Panels of kGlobNB columns: each panel is factored
// column by column (the reflector, then its dot products with the panel's
// later columns and with the panel's earlier reflectors, one workgroup sum
// for all of them), which gives the compact-WY factor T (Q_panel = I - V T
// V^T, the reference LAPACK dlarft's forward columnwise recurrence); the
// trailing columns are then updated 32 at a time as A2 -= V (T^T (V^T A2)),
// reading V and A2 twice per panel instead of twice per column.
__device__ __forceinline__ void glob_householder(double *tile, int m, int n, int ld, double *red, double *wsh,
                                                 double *tsh, double *wp, double *vsh) {
    const int tid = threadIdx.x;
    constexpr int NB = kGlobNB;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = min(NB, n - j0);
        // ---- panel: columns j0 .. j0 + nb - 1
        for (int jj = 0; jj < nb; ++jj) {
            const int c = j0 + jj;
            double sl = 0.0;
            for (int r = c + 1 + tid; r < m; r += kGlobThreads) {
                const double v = tile[(size_t)r * ld + c];
                sl += v * v;
            }
            const double sigma = glob_block_sum(sl, red);
            const double alpha = tile[(size_t)c * ld + c];
            double beta = alpha, tau = 0.0, scale = 0.0;
            if (sigma != 0.0) {
                const double nrm = sqrt(alpha * alpha + sigma);
                beta = alpha >= 0.0 ? -nrm : nrm;
                tau = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            for (int r = c + 1 + tid; r < m; r += kGlobThreads) tile[(size_t)r * ld + c] *= scale;
            __syncthreads();
            // acc[q] = (v_jj, A(:, j0 + q)) over rows >= c: for q < jj the
            // stored reflector v_q (for T), for q > jj a later panel column
            double acc[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) acc[q] = 0.0;
            for (int r = c + tid; r < m; r += kGlobThreads) {
                const double *row = tile + (size_t)r * ld + j0;
                const double vr = r == c ? 1.0 : row[jj];
#pragma unroll
                for (int q = 0; q < NB; ++q)
                    if (q < nb && q != jj) acc[q] += vr * row[q];
            }
            glob_block_sum_vec<NB>(acc, nb, red);
            if (tau != 0.0) {
                for (int r = c + tid; r < m; r += kGlobThreads) {
                    double *row = tile + (size_t)r * ld + j0;
                    const double tv = tau * (r == c ? 1.0 : row[jj]);
#pragma unroll
                    for (int q = 0; q < NB; ++q)
                        if (q > jj && q < nb) row[q] -= acc[q] * tv;
                }
            }
            if (tid == 0) {
                tile[(size_t)c * ld + c] = beta;
                // T(0:jj, jj) = -tau T(0:jj, 0:jj) (V(:, 0:jj)^T v_jj), T(jj, jj) = tau;
                // the sums go through LDS (red[200..]) so that the loops index
                // memory, not registers
                double *gq = red + 200;
#pragma unroll
                for (int q = 0; q < NB; ++q)
                    if (q < jj) gq[q] = acc[q];
                for (int i = 0; i < jj; ++i) {
                    double t = 0.0;
                    for (int q = i; q < jj; ++q) t += tsh[i * NB + q] * gq[q];
                    tsh[i * NB + jj] = -tau * t;
                }
                tsh[jj * NB + jj] = tau;
            }
            __syncthreads();
        }
        // ---- trailing columns j0 + nb .. n (n: the right-hand side), 32 at a
        // time.  The rows go in chunks of kGlobChunk: the chunk's V (nb
        // columns, 0 above the diagonal, 1 on it) is staged in LDS once for the
        // whole workgroup, and each thread has its kGlobChunk / 16 rows of A2
        // in flight at once.
        const int cc = tid % kGlobCols, rg = tid / kGlobCols, lane = tid & 63, wv = tid >> 6;
        constexpr int RPT = kGlobChunk / kGlobRowGroups;  // rows per thread and chunk
        auto stage_v = [&](int r0, int rn) {
            __syncthreads();  // the previous chunk's readers are done
            for (int idx = tid; idx < kGlobChunk * NB; idx += kGlobThreads) {
                const int rr = idx / NB, i = idx - rr * NB;
                const int r = r0 + rr, d = r - (j0 + i);
                double v = 0.0;
                if (rr < rn && i < nb && d >= 0) v = d == 0 ? 1.0 : tile[(size_t)r * ld + j0 + i];
                vsh[idx] = v;
            }
            __syncthreads();
        };
        for (int cb = j0 + nb; cb <= n; cb += kGlobCols) {
            const int col = cb + cc;
            const bool live = col <= n;
            // W = V^T A2 (nb x 32)
            double w[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) w[i] = 0.0;
            for (int r0 = j0; r0 < m; r0 += kGlobChunk) {
                const int rn = min(kGlobChunk, m - r0);
                stage_v(r0, rn);
                double ar[RPT];
#pragma unroll
                for (int u = 0; u < RPT; ++u) {
                    const int rr = rg + u * kGlobRowGroups;
                    ar[u] = (live && rr < rn) ? tile[(size_t)(r0 + rr) * ld + col] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < RPT; ++u) {
                    const double *vr = vsh + (rg + u * kGlobRowGroups) * NB;
#pragma unroll
                    for (int i = 0; i < NB; ++i) w[i] += vr[i] * ar[u];
                }
            }
            // the two row groups of a wavefront, then the 8 wavefronts
#pragma unroll
            for (int i = 0; i < NB; ++i) w[i] += __shfl_xor(w[i], 32, 64);
            if (lane < 32) {
#pragma unroll
                for (int i = 0; i < NB; ++i) wsh[(wv * NB + i) * kGlobCols + cc] = w[i];
            }
            __syncthreads();
            // W' = T^T W: thread (i, cc) of the first NB row groups
            if (rg < nb) {
                const int i = rg;
                double t = 0.0;
                for (int q = 0; q <= i; ++q) {
                    double wq = 0.0;
#pragma unroll
                    for (int v8 = 0; v8 < kGlobThreads / 64; ++v8) wq += wsh[(v8 * NB + q) * kGlobCols + cc];
                    t += tsh[q * NB + i] * wq;
                }
                wp[i * kGlobCols + cc] = t;
            }
            __syncthreads();
            // A2 -= V W'
            double wc[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) wc[i] = i < nb ? wp[i * kGlobCols + cc] : 0.0;
            for (int r0 = j0; r0 < m; r0 += kGlobChunk) {
                const int rn = min(kGlobChunk, m - r0);
                stage_v(r0, rn);
                if (live) {
                    double ar[RPT];
#pragma unroll
                    for (int u = 0; u < RPT; ++u) {
                        const int rr = rg + u * kGlobRowGroups;
                        ar[u] = rr < rn ? tile[(size_t)(r0 + rr) * ld + col] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < RPT; ++u) {
                        const int rr = rg + u * kGlobRowGroups;
                        const double *vr = vsh + rr * NB;
                        double x = 0.0;
#pragma unroll
                        for (int i = 0; i < NB; ++i) x += vr[i] * wc[i];
                        if (rr < rn) tile[(size_t)(r0 + rr) * ld + col] = ar[u] - x;
                    }
                }
            }
        }
        __syncthreads();  // the next panel (or the back substitution) reads the updated columns
    }
}

template <class MT>
__device__ __forceinline__ void solve_column_global(const ColumnArgs &a, int k, const GlobalScratch &g, int slot, double *red,
                                    int *ired, int *sh, double *wsh, double *tsh, double *wp, double *vsh) {
    const int tid = threadIdx.x;
    const int c0 = a.pcolptr[k];
    const int n2 = a.pcolptr[k + 1] - c0;
    if (n2 <= 0) return;
    double *tile = g.tile + (size_t)slot * g.tile_cap;
    double *xv = g.vec + (size_t)slot * g.vec_cap;  // n2
    uint64_t *keys = g.keys + (size_t)slot * g.key_cap;
    int *pos = g.ints + (size_t)slot * g.int_cap;  // key_cap + 1 (scan total at [T])
    int *offs = pos + g.key_cap + 1;                // n2 + 1

    // J offsets: exclusive prefix of |A(:, J_k[j])|
    const int T = glob_block_scan(
        n2, [&](int j) { int jr = a.prowidx[c0 + j]; return a.colptr[jr + 1] - a.colptr[jr]; }, offs, ired);
    int Tp = 1;
    while (Tp < T) Tp <<= 1;
    for (int e = tid; e < Tp; e += kGlobThreads) {
        uint64_t key = ~0ull;
        if (e < T) {
            const int j = glob_find_j(offs, n2, e);
            const int jr = a.prowidx[c0 + j];
            key = ((uint64_t)(unsigned)a.rowidx[a.colptr[jr] + e - offs[j]] << 32) | (unsigned)e;
        }
        keys[e] = key;
    }
    __syncthreads();
    // bitonic sort of Tp keys in global memory
    for (int size = 2; size <= Tp; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int q = tid; q < Tp / 2; q += kGlobThreads) {
                const int i = 2 * stride * (q / stride) + (q % stride);
                const int l = i + stride;
                const bool up = (i & size) == 0;
                const uint64_t x = keys[i], y = keys[l];
                if ((x > y) == up) {
                    keys[i] = y;
                    keys[l] = x;
                }
            }
            __syncthreads();
        }
    }
    // positions of the unique rows: pos[e] = #first occurrences in [0, e] - 1
    const int n1 = glob_block_scan(
        T,
        [&](int e) { return (e == 0 || (keys[e] >> 32) != (keys[e - 1] >> 32)) ? 1 : 0; }, pos, ired);
    if (tid == 0) sh[0] = -1;
    __syncthreads();
    for (int e = tid; e < T; e += kGlobThreads) {
        const bool first = (e == 0 || (keys[e] >> 32) != (keys[e - 1] >> 32));
        if (first && (int)(keys[e] >> 32) == k) sh[0] = pos[e];  // positions are exclusive counts
    }
    __syncthreads();
    const int E = sh[0];
    // exclusive count of first occurrences -> position of each entry's row:
    // inclusive count - 1 (a duplicate's run starts earlier, so its exclusive
    // count already includes its own row)
    for (int e = tid; e < T; e += kGlobThreads) {
        const bool first = (e == 0 || (keys[e] >> 32) != (keys[e - 1] >> 32));
        pos[e] = pos[e] + (first ? 1 : 0) - 1;
    }
    __syncthreads();
    if (E < 0) {  // k not in I_k: M(:, k) = 0, as the reference's Sol_SSPAIv10
        for (int c = tid; c < n2; c += kGlobThreads) store_m<MT>(a, c0 + c, 0.0);
        if (a.n1_out && tid == 0) a.n1_out[k] = n1;
        __syncthreads();
        return;
    }
    // The tile: m = max(n1, n2) rows (a column with n1 < n2 still has an R_jj
    // for every j), n2 columns of A(I_k, J_k) and, as column n2, the right-
    // hand side e_E, so that every panel's reflectors reach Q^T e_E with the
    // trailing columns.  Row-major, leading dimension ld = n2 + 1.
    const int m = n1 > n2 ? n1 : n2;
    const int ld = n2 + 1;
    for (long long idx = tid; idx < (long long)m * ld; idx += kGlobThreads) {
        const int r = (int)(idx / ld), c = (int)(idx - (long long)r * ld);
        tile[idx] = (c == n2 && r == E) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < T; e += kGlobThreads) {
        const int ee = (int)(unsigned)keys[e];
        const int j = glob_find_j(offs, n2, ee);
        const int jr = a.prowidx[c0 + j];
        tile[(size_t)pos[e] * ld + j] = a.val[a.colptr[jr] + ee - offs[j]];
    }
    __syncthreads();
    glob_householder(tile, m, n2, ld, red, wsh, tsh, wp, vsh);
    // back substitution, column oriented: x_c = y_c / R_cc, then y_q -= R_qc x_c
    for (int c = n2 - 1; c >= 0; --c) {
        const double xc = tile[(size_t)c * ld + n2] / tile[(size_t)c * ld + c];
        for (int q = tid; q < c; q += kGlobThreads) tile[(size_t)q * ld + n2] -= tile[(size_t)q * ld + c] * xc;
        if (tid == 0) xv[c] = xc;
        __syncthreads();
    }
    for (int c = tid; c < n2; c += kGlobThreads) store_m<MT>(a, c0 + c, xv[c]);
    if (a.n1_out && tid == 0) a.n1_out[k] = n1;
    __syncthreads();
}

// Orders the global tier's column list by decreasing work |J_k| x T_k (ties
// by column), so that the widest columns start first, each on a workgroup of
// its own, rather than wherever they fall in the queue.  One workgroup; a
// list longer than kWideSortMax keeps the planner's order.
constexpr int kWideSortMax = 4096;
__global__ __launch_bounds__(1024) void spai_sort_wide_columns(int *list, const int *count, const int *colptr,
                                                               const int *pcolptr, const int *prowidx) {
    __shared__ unsigned long long key[kWideSortMax];
    const int cnt = *count;
    if (cnt < 2 || cnt > kWideSortMax) return;
    int P = 1;
    while (P < cnt) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        unsigned long long kk = ~0ull;
        if (i < cnt) {
            const int k = list[i];
            const int c0 = pcolptr[k], n2 = pcolptr[k + 1] - c0;
            unsigned long long t = 0ull;
            for (int j = 0; j < n2; ++j) {
                const int r = prowidx[c0 + j];
                t += (unsigned long long)(colptr[r + 1] - colptr[r]);
            }
            const unsigned long long work = min((unsigned long long)n2 * t, 0xffffffffull);
            kk = ((0xffffffffull - work) << 32) | (unsigned)k;  // ascending key: descending work
        }
        key[i] = kk;
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int q = threadIdx.x; q < P / 2; q += blockDim.x) {
                const int i = 2 * stride * (q / stride) + (q % stride);
                const int l = i + stride;
                const bool up = (i & size) == 0;
                const unsigned long long x = key[i], y = key[l];
                if ((x > y) == up) {
                    key[i] = y;
                    key[l] = x;
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) list[i] = (int)(unsigned)(key[i] & 0xffffffffull);
}

template <class MT, bool DYNAMIC>
__global__ __launch_bounds__(kGlobThreads) void spai_global_tier_kernel(ColumnArgs a, GlobalScratch g) {
    __shared__ double red[64 + kGlobThreads];
    __shared__ int ired[kGlobThreads / 64];
    __shared__ int sh[2];
    __shared__ double wsh[kGlobThreads / 64 * kGlobNB * kGlobCols];  // W partials per wavefront
    __shared__ double tsh[kGlobNB * kGlobNB];                        // T of the panel
    __shared__ double wp[kGlobNB * kGlobCols];                       // T^T W
    __shared__ double vsh[kGlobChunk * kGlobNB];                     // a chunk of V
    const int slot = blockIdx.x;
    if (DYNAMIC) {
        const int count = *a.count;
        for (;;) {
            if (threadIdx.x == 0) sh[1] = atomicAdd(a.head, 1);
            __syncthreads();
            const int idx = sh[1];
            __syncthreads();
            if (idx >= count) break;
            solve_column_global<MT>(a, a.list[idx], g, slot, red, ired, sh, wsh, tsh, wp, vsh);
        }
    } else {
        for (int k = a.col_begin + blockIdx.x; k < a.col_end; k += gridDim.x)
            solve_column_global<MT>(a, k, g, slot, red, ired, sh, wsh, tsh, wp, vsh);
    }
}

// ---------------------------------------------------------------------------
// Planner
// ---------------------------------------------------------------------------
// counters layout (ints): [0..kTiers) column count per tier,
// [kTiers..2kTiers) queue heads, [2kTiers] max T, [2kTiers+1] max n2,
// [2kTiers+2] empty-column count; stats64[0] = sum of T_k.
__device__ __forceinline__ int classify(int T, int n2, int lds_fit) {
    if (n2 <= kBucketNJ[0] && T <= kBucketT[0]) return 0;
    if (n2 <= kBucketNJ[1] && T <= kBucketT[1]) return 1;
    if (n2 <= kBucketNJ[2] && T <= kBucketT[2]) return 2;
    if (n2 <= kLdsTierNJ && T <= kLdsTierT && lds_fit) return 3;
    return 4;
}

__device__ __forceinline__ bool lds_tier_fits(int T, int n2) {
    // n1 <= T, tile + R + y must fit the LDS tier's buffer
    long long need = (long long)T * n2 + (long long)n2 * n2 + n2;
    return need <= kLdsTierDoubles;
}

__global__ void spai_zero_counters(int *c, int n, unsigned long long *s64, int n64) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
    if (i < n64) s64[i] = 0ull;
}

// n2_k, T_k = sum_j |A(:,J_k[j])| and the tier of column k.
__device__ __forceinline__ void plan_column(const int *colptr, const int *pcolptr, const int *prowidx, int k, int &T,
                                            int &n2, int &tier) {
    const int c0 = pcolptr[k];
    n2 = pcolptr[k + 1] - c0;
    T = 0;
    // the first 8 columns' loads are issued together (two round trips
    // instead of two per column of J_k); wider columns continue serially
    constexpr int kFirst = 8;
    int jr[kFirst];
#pragma unroll
    for (int j = 0; j < kFirst; ++j) jr[j] = j < n2 ? prowidx[c0 + j] : 0;
    int lo[kFirst], hi[kFirst];
#pragma unroll
    for (int j = 0; j < kFirst; ++j) {
        lo[j] = j < n2 ? colptr[jr[j]] : 0;
        hi[j] = j < n2 ? colptr[jr[j] + 1] : 0;
    }
#pragma unroll
    for (int j = 0; j < kFirst; ++j) T += hi[j] - lo[j];
    for (int j = kFirst; j < n2; ++j) {
        const int r = prowidx[c0 + j];
        T += colptr[r + 1] - colptr[r];
    }
    tier = n2 > 0 ? classify(T, n2, lds_tier_fits(T, n2)) : -1;
}

// The planner: a workgroup plans kPlanCols consecutive columns, kPlanItems
// per thread.  Slots inside a tier come from wavefront-aggregated LDS
// counts, and the maxima and sums are reduced per thread, then per
// wavefront, so each workgroup issues one global atomic per counter: with a
// workgroup per 256 columns the same-address global atomics of ~10^4
// workgroups serialised in L2 and took 0.38 ms for 2.6 M columns.
// stats64: [0] sum T_k, [1] max over global-tier columns of max(min(T_k, n), n2_k) * n2_k
// (a bound on that column's tile), [2] the same over all columns.
constexpr int kPlanBlock = 256;
constexpr int kPlanItems = 16;
constexpr int kPlanCols = kPlanBlock * kPlanItems;

template <bool DYNAMIC>
__global__ __launch_bounds__(kPlanBlock) void spai_plan_kernel(const int *colptr, const int *pcolptr,
                                                               const int *prowidx, int n, int col_begin, int col_end,
                                                               int *counters, unsigned long long *stats64, int *lists,
                                                               int list_stride) {
    __shared__ int s_cnt[kTiers], s_base[kTiers], s_maxT, s_maxN, s_empty;
    __shared__ unsigned long long s_sumT, s_tile4, s_tileAll;
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < kTiers) s_cnt[tid] = 0;
    if (tid == 0) {
        s_maxT = 0;
        s_maxN = 0;
        s_empty = 0;
        s_sumT = 0ull;
        s_tile4 = 0ull;
        s_tileAll = 0ull;
    }
    __syncthreads();
    const int first = col_begin + blockIdx.x * kPlanCols;
    int tiers[kPlanItems], local[kPlanItems];
    int maxT = 0, maxN = 0, empty = 0;
    unsigned long long sumT = 0ull, tile4 = 0ull, tileAll = 0ull;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int it = 0; it < kPlanItems; ++it) {
        const int k = first + it * kPlanBlock + tid;
        int T = 0, n2 = 0, tier = -1;
        if (k < col_end) {
            plan_column(colptr, pcolptr, prowidx, k, T, n2, tier);
            maxT = max(maxT, T);
            maxN = max(maxN, n2);
            sumT += (unsigned long long)T;
            empty += n2 == 0;
            // rows of the tile: n1 <= min(T, n); a rank-deficient column
            // (n1 < n2) still touches the diagonal R_jj for every j < n2
            const unsigned long long tb = (unsigned long long)max(min(T, n), n2) * (unsigned long long)n2;
            tileAll = max(tileAll, tb);
            if (tier == kTiers - 1) tile4 = max(tile4, tb);
        }
        // slot inside the tier, local to the workgroup: one LDS atomic per
        // wavefront and tier
        int slot = -1;
#pragma unroll
        for (int t = 0; t < kTiers; ++t) {
            const unsigned long long m = __ballot(tier == t);
            if (m == 0ull) continue;
            int b = 0;
            if (lane == __ffsll((long long)m) - 1) b = atomicAdd(&s_cnt[t], __popcll(m));
            b = __shfl(b, __ffsll((long long)m) - 1, 64);
            if (tier == t) slot = b + __popcll(m & below);
        }
        tiers[it] = tier;
        local[it] = slot;
    }
    // per-wavefront maxima and sums, then one LDS atomic per wavefront
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        maxT = max(maxT, __shfl_xor(maxT, o, 64));
        maxN = max(maxN, __shfl_xor(maxN, o, 64));
        empty += __shfl_xor(empty, o, 64);
        sumT += __shfl_xor(sumT, o, 64);
        tile4 = max(tile4, (unsigned long long)__shfl_xor(tile4, o, 64));
        tileAll = max(tileAll, (unsigned long long)__shfl_xor(tileAll, o, 64));
    }
    if (lane == 0) {
        atomicMax(&s_maxT, maxT);
        atomicMax(&s_maxN, maxN);
        if (empty) atomicAdd(&s_empty, empty);
        atomicAdd(&s_sumT, sumT);
        atomicMax(&s_tile4, tile4);
        atomicMax(&s_tileAll, tileAll);
    }
    __syncthreads();
    if (tid < kTiers) s_base[tid] = s_cnt[tid] ? atomicAdd(&counters[tid], s_cnt[tid]) : 0;
    if (tid == 0) {
        atomicMax(&counters[2 * kTiers], s_maxT);
        atomicMax(&counters[2 * kTiers + 1], s_maxN);
        if (s_empty) atomicAdd(&counters[2 * kTiers + 2], s_empty);
        atomicAdd(&stats64[0], s_sumT);
        atomicMax(&stats64[1], s_tile4);
        atomicMax(&stats64[2], s_tileAll);
    }
    __syncthreads();
    if (DYNAMIC) {
        for (int it = 0; it < kPlanItems; ++it) {
            const int tier = tiers[it];
            if (tier >= 0) lists[(size_t)tier * list_stride + s_base[tier] + local[it]] = first + it * kPlanBlock + tid;
        }
    }
}

template <class KernelT>
int persistent_grid(KernelT kernel, int block, int device_cus) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return per_cu * device_cus;
}

// counters beyond the planner's: overflow count and queue head for the
// overflow lists of tiers 1 and 2; lists: kTiers planner lists, then the two
// overflow lists
constexpr int kOvfCount = 2 * kTiers + 3;
constexpr int kOvfHead = kOvfCount + 2;
constexpr int kCounters = kOvfHead + 2;
constexpr int kLists = kTiers + 2;

bool wide_keys_forced() {
    const char *env = std::getenv("SPAI_MI355X_SORT_KEYS");
    return env && std::atoi(env) == 64;
}

template <class MT, bool REFINE>
void launch_tiers(const ColumnArgs &base, int base_n, bool dynamic, int static_tier, const int *shape, int *counters,
                  int *lists, int list_stride, int cus, hipStream_t st, hipEvent_t after_reg) {
    using QT = typename std::conditional<REFINE, float, double>::type;
    ColumnArgs a = base;
    // t: tier; ovf: a static run's follow-up over the tier's overflow list
    // alone (in a dynamic run a tier's queue takes that list after its own)
    auto setup = [&](int t, bool ovf) {
        a = base;
        if (ovf) {
            a.list = lists + (size_t)(kTiers + t - 1) * list_stride;
            a.count = counters + kOvfCount + t - 1;
            a.head = counters + kOvfHead + t - 1;
        } else if (dynamic) {
            a.list = lists + (size_t)t * list_stride;
            a.count = counters + t;
            a.head = counters + kTiers + t;
            if (t == 1 || t == 2) {
                a.ovf_in_list = lists + (size_t)(kTiers + t - 1) * list_stride;
                a.ovf_in_count = counters + kOvfCount + t - 1;
            }
        }
        if (t < 2) {  // tiers 0 and 1 hand wider row sets on
            a.ovf_list = lists + (size_t)(kTiers + t) * list_stride;
            a.ovf_count = counters + kOvfCount + t;
        }
    };
    auto run_reg = [&](auto tier_tag, int t, bool ovf = false) {
        using TR = decltype(tier_tag);
        setup(t, ovf);
        // 32-bit sort keys when every row index fits beside the entry index
        // (SPAI_MI355X_SORT_KEYS=64 keeps 64-bit keys; both give the same M).
        // Not for the FP64 16-column bucket: there the compiler trades the
        // 64-bit keys' spill for 256 registers and half the occupancy.
        constexpr bool k32_pays = REFINE || TR::NJ < 16;
        const bool k32 = k32_pays && (long long)base_n + 1 < (1LL << (32 - log2_exact(TR::TMAX))) &&
                         !wide_keys_forced();
        auto launch = [&](auto k32_tag) {
            constexpr bool K32 = decltype(k32_tag)::value;
            if (dynamic || ovf) {
                auto k = spai_reg_tier_kernel<TR, QT, REFINE, MT, true, K32>;
                hipLaunchKernelGGL(k, dim3(persistent_grid(k, kBlock, cus)), dim3(kBlock), 0, st, a);
            } else {
                auto k = spai_reg_tier_kernel<TR, QT, REFINE, MT, false, K32>;
                int ncols = base.col_end - base.col_begin;
                int waves = ceil_div(ncols, 64 / TR::W * 4);  // ~4 columns per group
                int grid = std::max(1, std::min(ceil_div(waves, kBlock / 64), persistent_grid(k, kBlock, cus) * 4));
                hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, st, a);
            }
        };
        if (k32) launch(std::true_type{});
        else launch(std::false_type{});
        SPAI_CHECK_LAUNCH();
    };
    auto run_lds = [&](int t) {
        setup(t, false);
        if (dynamic) {
            auto k = spai_lds_tier_kernel<MT, true>;
            hipLaunchKernelGGL(k, dim3(persistent_grid(k, 64, cus)), dim3(64), 0, st, a);
        } else {
            auto k = spai_lds_tier_kernel<MT, false>;
            int ncols = base.col_end - base.col_begin;
            hipLaunchKernelGGL(k, dim3(std::max(1, std::min(ncols, persistent_grid(k, 64, cus) * 8))), dim3(64), 0,
                               st, a);
        }
        SPAI_CHECK_LAUNCH();
    };
    // register tier t with the bucket's chosen shape
    auto run_bucket = [&](int t, bool ovf = false) {
        auto go = [&](auto tier_tag) { run_reg(tier_tag, t, ovf); };
        if (t == 0) with_shape<0>(shape[0], go);
        else if (t == 1) with_shape<1>(shape[1], go);
        else with_shape<2>(shape[2], go);
    };
    if (dynamic) {
        // smallest tier first: each register tier's queue ends with the
        // columns the tier before it handed on (the launches run one after
        // another on the stream either way)
        for (int t = 0; t < kRegTiers; ++t) run_bucket(t);
        SPAI_HIP(hipEventRecord(after_reg, st));
        run_lds(3);
    } else if (static_tier < kRegTiers) {
        run_bucket(static_tier);
        // the overflow lists of the tiers after the static one
        for (int t = static_tier + 1; t < kRegTiers; ++t) run_bucket(t, true);
        SPAI_HIP(hipEventRecord(after_reg, st));
    } else {
        SPAI_HIP(hipEventRecord(after_reg, st));
        run_lds(3);
    }
}

// ---------------------------------------------------------------------------
// Product C = A M and the SPAI quality ||A M - I||_F.  Column k of C is
// A(:, J_k) m_k, supported on the row set I_k of the SPAI build.  One
// 512-thread workgroup per column (persistent over columns), scratch slot in
// global memory: the gathered (row, entry) keys are bitonic-sorted, runs of
// equal rows are summed in storage order, C(:, k) is written at ccolptr[k]
// and ||C(:, k) - e_k||^2 is reduced.
// ---------------------------------------------------------------------------
struct ProductArgs {
    const int *colptr, *rowidx;  // A (CSC)
    const double *val;
    const int *pcolptr, *prowidx;  // M's pattern (CSC)
    const void *mval;              // M's values in pattern order (double or float)
    int n;
    const int *ccolptr;  // C's column offsets (prefix of |I_k|), nullptr: quality only
    int *crow;
    double *cval;
    double *colres2;  // ||A m_k - e_k||^2 per column
    int *n1_out;      // |I_k| per column (nullable)
    uint64_t *keys;   // slots * key_cap
    int *ints;        // slots * int_cap: positions (key_cap + 1), J offsets (n2_cap + 1)
    long long key_cap, int_cap;
};

template <class MT>
__global__ __launch_bounds__(kGlobThreads) void spai_product_kernel(ProductArgs a) {
    __shared__ double red[64 + kGlobThreads];
    __shared__ int ired[kGlobThreads / 64];
    __shared__ int sh[1];
    const int tid = threadIdx.x;
    const MT *m = (const MT *)a.mval;
    uint64_t *keys = a.keys + (size_t)blockIdx.x * a.key_cap;
    int *pos = a.ints + (size_t)blockIdx.x * a.int_cap;
    int *offs = pos + a.key_cap + 1;
    for (int k = blockIdx.x; k < a.n; k += gridDim.x) {
        const int c0 = a.pcolptr[k];
        const int n2 = a.pcolptr[k + 1] - c0;
        if (n2 <= 0) {  // M(:, k) = 0: the column's residual is ||e_k||^2 = 1
            if (tid == 0) {
                a.colres2[k] = 1.0;
                if (a.n1_out) a.n1_out[k] = 0;
            }
            continue;
        }
        const int T = glob_block_scan(
            n2, [&](int j) { int jr = a.prowidx[c0 + j]; return a.colptr[jr + 1] - a.colptr[jr]; }, offs, ired);
        int Tp = 1;
        while (Tp < T) Tp <<= 1;
        for (int e = tid; e < Tp; e += kGlobThreads) {
            uint64_t key = ~0ull;
            if (e < T) {
                const int j = glob_find_j(offs, n2, e);
                const int jr = a.prowidx[c0 + j];
                key = ((uint64_t)(unsigned)a.rowidx[a.colptr[jr] + e - offs[j]] << 32) | (unsigned)e;
            }
            keys[e] = key;
        }
        __syncthreads();
        for (int size = 2; size <= Tp; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int q = tid; q < Tp / 2; q += kGlobThreads) {
                    const int i = 2 * stride * (q / stride) + (q % stride);
                    const int l = i + stride;
                    const bool up = (i & size) == 0;
                    const uint64_t x = keys[i], y = keys[l];
                    if ((x > y) == up) {
                        keys[i] = y;
                        keys[l] = x;
                    }
                }
                __syncthreads();
            }
        }
        // output slot of each run of equal rows: exclusive count of run starts
        const int n1 = glob_block_scan(
            T, [&](int e) { return (e == 0 || (keys[e] >> 32) != (keys[e - 1] >> 32)) ? 1 : 0; }, pos, ired);
        if (tid == 0) sh[0] = 0;
        __syncthreads();
        double part = 0.0;
        for (int e = tid; e < T; e += kGlobThreads) {
            const int row = (int)(keys[e] >> 32);
            if (e > 0 && row == (int)(keys[e - 1] >> 32)) continue;  // not a run start
            double sum = 0.0;
            for (int f = e; f < T && (int)(keys[f] >> 32) == row; ++f) {
                const int ee = (int)(unsigned)keys[f];
                const int j = glob_find_j(offs, n2, ee);
                const int jr = a.prowidx[c0 + j];
                sum += a.val[a.colptr[jr] + ee - offs[j]] * (double)m[c0 + j];
            }
            if (a.ccolptr) {
                a.crow[a.ccolptr[k] + pos[e]] = row;
                a.cval[a.ccolptr[k] + pos[e]] = sum;
            }
            const double d = row == k ? sum - 1.0 : sum;
            part += d * d;
            if (row == k) sh[0] = 1;
        }
        const double tot = glob_block_sum(part, red);
        if (tid == 0) {
            a.colres2[k] = tot + (sh[0] ? 0.0 : 1.0);  // e_k outside I_k contributes 1
            if (a.n1_out) a.n1_out[k] = n1;
        }
        __syncthreads();
    }
}

}  // namespace

SpaiBuilder::SpaiBuilder() = default;
SpaiBuilder::~SpaiBuilder() {
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (h_counters_) (void)hipHostFree(h_counters_);
    if (h_stats_) (void)hipHostFree(h_stats_);
}

void SpaiBuilder::build(const SpaiBuildParams &p, hipStream_t st, SpaiBuildStats *stats) {
    require(p.n >= 0 && p.colptr && p.rowidx && p.val && p.perm && p.m_csc && p.m_csr, "spai build: bad arguments");
    int cb = std::max(0, p.col_begin), ce = std::min(p.n, p.col_end < 0 ? p.n : p.col_end);
    require(cb <= ce, "spai build: empty or inverted column range");
    int dev = 0;
    SPAI_HIP(hipGetDevice(&dev));
    if (cus_ == 0) SPAI_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
    const bool dynamic = p.schedule == 1;
    for (auto &e : ev_)
        if (!e) SPAI_HIP(hipEventCreate(&e));
    // J_k from M's pattern (A's own unless another pattern is given)
    const int *pc = p.pcolptr ? p.pcolptr : p.colptr;
    const int *pr = p.pcolptr ? p.prowidx : p.rowidx;
    const int nc = kCounters;
    counters_.reserve(nc);
    stats64_.reserve(3);
    const int list_stride = std::max(1, ce - cb);
    lists_.reserve((size_t)kLists * list_stride);
    SPAI_HIP(hipEventRecord(ev_[0], st));
    hipLaunchKernelGGL(spai_zero_counters, dim3(1), dim3(64), 0, st, counters_.get(), nc, stats64_.get(), 3);
    SPAI_CHECK_LAUNCH();
    if (ce > cb) {
        int grid = ceil_div(ce - cb, kPlanCols);
        if (dynamic)
            hipLaunchKernelGGL(spai_plan_kernel<true>, dim3(grid), dim3(kPlanBlock), 0, st, p.colptr, pc, pr, p.n, cb, ce,
                               counters_.get(), stats64_.get(), lists_.get(), list_stride);
        else
            hipLaunchKernelGGL(spai_plan_kernel<false>, dim3(grid), dim3(kPlanBlock), 0, st, p.colptr, pc, pr, p.n, cb,
                               ce, counters_.get(), stats64_.get(), (int *)nullptr, list_stride);
        SPAI_CHECK_LAUNCH();
    }
    SPAI_HIP(hipEventRecord(ev_[1], st));
    int hc[kCounters] = {0};
    unsigned long long h64[3] = {0, 0, 0};
    if (!h_counters_) SPAI_HIP(hipHostMalloc((void **)&h_counters_, sizeof(int) * kCounters, hipHostMallocDefault));
    if (!h_stats_) SPAI_HIP(hipHostMalloc((void **)&h_stats_, sizeof(h64), hipHostMallocDefault));
    auto copy_counters = [&]() {
        SPAI_HIP(hipMemcpyAsync(h_counters_, counters_.get(), sizeof(int) * nc, hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipMemcpyAsync(h_stats_, stats64_.get(), sizeof(h64), hipMemcpyDeviceToHost, st));
    };
    auto take_counters = [&]() {
        std::memcpy(hc, h_counters_, sizeof(int) * nc);
        std::memcpy(h64, h_stats_, sizeof(h64));
    };
    ColumnArgs a{};
    a.colptr = p.colptr;
    a.rowidx = p.rowidx;
    a.pcolptr = pc;
    a.prowidx = pr;
    a.val = p.val;
    a.perm = p.perm;
    a.m_csc = p.m_csc;
    a.m_csr = p.m_csr;
    a.n1_out = p.n1_out;
    a.col_begin = cb;
    a.col_end = ce;
    int static_tier = 0;
    if (!dynamic && ce > cb) {
        // the reference reads n2max back before choosing its kernels
        // (src/spai_single_dcu.cpp:1245-1252); static mode does the same for
        // the largest column's tier
        copy_counters();
        SPAI_HIP(hipStreamSynchronize(st));
        take_counters();
        static_tier = classify_host(hc[2 * kTiers], hc[2 * kTiers + 1]);
    } else if (ce > cb) {
        // the global tier's columns and tile bounds are final once the
        // planner has run (the register tiers hand columns on only among
        // themselves): the copy is read after the tiers are queued, so the
        // device does not wait for the host in between
        copy_counters();
        SPAI_HIP(hipEventRecord(ev_[7], st));
    }
    for (int t = 0; t < kTunedTiers; ++t)
        require(p.shape[t] >= 0 && p.shape[t] < tile_shape_count(t), "spai build: no such tile shape");
    SPAI_HIP(hipEventRecord(ev_[2], st));
    if (ce > cb && static_tier < kTiers - 1) {
        if (p.precision == 1)
            launch_tiers<float, true>(a, p.n, dynamic, static_tier, p.shape, counters_.get(), lists_.get(),
                                      list_stride, cus_, st, ev_[3]);
        else
            launch_tiers<double, false>(a, p.n, dynamic, static_tier, p.shape, counters_.get(), lists_.get(),
                                        list_stride, cus_, st, ev_[3]);
    } else {
        SPAI_HIP(hipEventRecord(ev_[3], st));
    }
    SPAI_HIP(hipEventRecord(ev_[4], st));
    if (dynamic && ce > cb) {
        SPAI_HIP(hipEventSynchronize(ev_[7]));
        take_counters();
    }
    SPAI_HIP(hipEventRecord(ev_[5], st));
    const int global_cols = dynamic ? hc[kTiers - 1] : (static_tier == kTiers - 1 ? ce - cb : 0);
    if (global_cols > 0) {
        // very wide columns: one workgroup per column with a global scratch
        // slot sized by the planner's tile bound
        GlobalScratch g{};
        const long long maxT = std::max(1, hc[2 * kTiers]);
        g.n2_cap = std::max(1, hc[2 * kTiers + 1]);
        g.n1_cap = (int)std::max<long long>(std::min<long long>(maxT, std::max(p.n, 1)), g.n2_cap);
        // max(n1, n2) x (n2 + 1) <= the planner's bound max(min(T, n), n2) x n2
        // plus one column of max(min(T, n), n2) rows
        g.tile_cap = (long long)std::max<unsigned long long>(1ull, dynamic ? h64[1] : h64[2]) +
                     std::max(g.n1_cap, g.n2_cap);
        long long kp = 1;
        while (kp < maxT) kp <<= 1;
        g.key_cap = kp;
        g.vec_cap = (long long)g.n2_cap;
        g.int_cap = g.key_cap + 1 + g.n2_cap + 1;
        const long long slot_bytes = 8 * (g.tile_cap + g.vec_cap + g.key_cap) + 4 * g.int_cap;
        const long long budget = 8LL << 30;  // scratch for concurrent wide columns
        int slots = (int)std::min<long long>(std::max<long long>(1, budget / slot_bytes),
                                             std::min<long long>(global_cols, 2LL * cus_));
        glob_tile_.reserve((size_t)(slots * g.tile_cap));
        glob_vec_.reserve((size_t)(slots * g.vec_cap));
        glob_keys_.reserve((size_t)(slots * g.key_cap));
        glob_ints_.reserve((size_t)(slots * g.int_cap));
        g.tile = glob_tile_.get();
        g.vec = glob_vec_.get();
        g.keys = glob_keys_.get();
        g.ints = glob_ints_.get();
        ColumnArgs ga = a;
        if (dynamic) {
            ga.list = lists_.get() + (size_t)(kTiers - 1) * list_stride;
            ga.count = counters_.get() + (kTiers - 1);
            ga.head = counters_.get() + kTiers + (kTiers - 1);
            if (global_cols > 1 && global_cols <= kWideSortMax) {
                hipLaunchKernelGGL(spai_sort_wide_columns, dim3(1), dim3(1024), 0, st, (int *)ga.list, ga.count,
                                   p.colptr, pc, pr);
                SPAI_CHECK_LAUNCH();
            }
        }
        auto launch = [&](auto mt_tag) {
            using MT = decltype(mt_tag);
            if (dynamic)
                hipLaunchKernelGGL((spai_global_tier_kernel<MT, true>), dim3(slots), dim3(kGlobThreads), 0, st, ga,
                                   g);
            else
                hipLaunchKernelGGL((spai_global_tier_kernel<MT, false>), dim3(slots), dim3(kGlobThreads), 0, st, ga,
                                   g);
        };
        if (p.precision == 1) launch(float{});
        else launch(double{});
        SPAI_CHECK_LAUNCH();
    }
    // the final counters (the tiers' hand-on counts) for the report
    copy_counters();
    SPAI_HIP(hipEventRecord(ev_[6], st));
    SPAI_HIP(hipEventSynchronize(ev_[6]));
    take_counters();
    if (p.tiers_ms) SPAI_HIP(hipEventElapsedTime(p.tiers_ms, ev_[2], ev_[4]));
    if (stats) {
        const int from[4] = {0, 2, 3, 5}, to[4] = {1, 3, 4, 6};
        for (int q = 0; q < 4; ++q) SPAI_HIP(hipEventElapsedTime(&stats->phase_ms[q], ev_[from[q]], ev_[to[q]]));
        for (int t = 0; t < kTiers; ++t) stats->tier_columns[t] = dynamic ? hc[t] : 0;
        if (!dynamic && ce > cb) stats->tier_columns[static_tier] = ce - cb - hc[2 * kTiers + 2];
        // columns handed on to the next tier are counted where they were solved
        for (int t = 0; t < 2; ++t) {
            stats->tier_columns[t] -= hc[kOvfCount + t];
            stats->tier_columns[t + 1] += hc[kOvfCount + t];
        }
        stats->max_gathered = hc[2 * kTiers];
        stats->max_n2 = hc[2 * kTiers + 1];
        stats->empty_columns = hc[2 * kTiers + 2];
        stats->sum_gathered = (long long)h64[0];
        stats->static_tier = dynamic ? -1 : static_tier;
    }
}

int tile_shape_count(int tier) {
    switch (tier) {
    case 0: return (int)std::tuple_size_v<Bucket0>;
    case 1: return (int)std::tuple_size_v<Bucket1>;
    case 2: return (int)std::tuple_size_v<Bucket2>;
    default: return 0;
    }
}

TileShape tile_shape(int tier, int c) {
    if (c < 0 || c >= tile_shape_count(tier)) return TileShape{};
    switch (tier) {
    case 0: return shape_info<0>(c, std::make_index_sequence<std::tuple_size_v<Bucket0>>{});
    case 1: return shape_info<1>(c, std::make_index_sequence<std::tuple_size_v<Bucket1>>{});
    default: return shape_info<2>(c, std::make_index_sequence<std::tuple_size_v<Bucket2>>{});
    }
}

int classify_host(int T, int n2) {
    for (int b = 0; b < kRegTiers; ++b)
        if (n2 <= kBucketNJ[b] && T <= kBucketT[b]) return b;
    long long need = (long long)T * n2 + (long long)n2 * n2 + n2;
    if (n2 <= kLdsTierNJ && T <= kLdsTierT && need <= kLdsTierDoubles) return 3;
    return 4;
}

void spai_product(const SpaiProductParams &p, hipStream_t st, int cus, DeviceArray<uint64_t> &keys,
                  DeviceArray<int> &ints) {
    require(p.n >= 0 && p.colptr && p.rowidx && p.val && p.pcolptr && p.prowidx && p.mval && p.colres2,
            "spai product: bad arguments");
    if (p.n == 0) return;
    ProductArgs a{};
    a.colptr = p.colptr;
    a.rowidx = p.rowidx;
    a.val = p.val;
    a.pcolptr = p.pcolptr;
    a.prowidx = p.prowidx;
    a.mval = p.mval;
    a.n = p.n;
    a.ccolptr = p.ccolptr;
    a.crow = p.crow;
    a.cval = p.cval;
    a.colres2 = p.colres2;
    a.n1_out = p.n1_out;
    long long kp = 1;
    while (kp < std::max(1, p.max_gathered)) kp <<= 1;
    a.key_cap = kp;
    a.int_cap = kp + 1 + (long long)std::max(1, p.max_n2) + 1;
    const int slots = std::max(1, std::min(p.n, 2 * std::max(cus, 1)));
    keys.reserve((size_t)slots * (size_t)a.key_cap);
    ints.reserve((size_t)slots * (size_t)a.int_cap);
    a.keys = keys.get();
    a.ints = ints.get();
    if (p.m_fp32)
        hipLaunchKernelGGL(spai_product_kernel<float>, dim3(slots), dim3(kGlobThreads), 0, st, a);
    else
        hipLaunchKernelGGL(spai_product_kernel<double>, dim3(slots), dim3(kGlobThreads), 0, st, a);
    SPAI_CHECK_LAUNCH();
    SPAI_HIP(hipStreamSynchronize(st));
}

}  // namespace spai
