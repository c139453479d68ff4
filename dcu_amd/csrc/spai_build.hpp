// Host interface of the SPAI column kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace spai {

// Size tiers: 0..2 register tiers (8, 16, 64 lanes per column), 3 the LDS
// wavefront tier, 4 the global-memory tier (one workgroup per very wide
// column, FP64, tile in a scratch slot).
constexpr int kTiers = 5;

// Group shape of a register tier: W lanes per column, EPL gathered entries
// per lane, NJ largest |J_k|, RPL tile rows per lane.  Each of the three
// register tiers (size buckets) has a list of candidates; entry 0 is the
// default and the autotuner picks one per bucket.
struct TileShape {
    int lanes = 0, entries_per_lane = 0, max_cols = 0, rows_per_lane = 0;
};
constexpr int kTunedTiers = 3;
int tile_shape_count(int tier);          // candidates of register tier `tier`, 0 if none
TileShape tile_shape(int tier, int c);   // candidate c (zeros when out of range)

struct SpaiBuildParams {
    int n = 0;
    const int *colptr = nullptr;  // device CSC of A
    const int *rowidx = nullptr;
    const double *val = nullptr;
    const int *pcolptr = nullptr;  // M's pattern (CSC) when it is not A's: J_k =
    const int *prowidx = nullptr;  // prowidx[pcolptr[k] .. pcolptr[k+1])
    const int *perm = nullptr;    // pattern position -> position in M's solver layout
    int precision = 0;            // 0: FP64 QR, FP64 M; 1: FP32 QR + FP64 sweep, FP32 M
    int schedule = 0;             // 0: balanced_columns (static), 1: balanced_nnz (queues)
    int col_begin = 0, col_end = -1;
    void *m_csc = nullptr;        // double[nnz(M)] or float[nnz(M)] (precision 1), pattern order
    void *m_csr = nullptr;        // same type, M's solver layout order
    int *n1_out = nullptr;        // optional per-column |I_k|
    int shape[kTunedTiers] = {0, 0, 0};  // candidate per register tier
    float *tiers_ms = nullptr;    // optional: device time of the tier kernels (planner and global tier excluded)
};

struct SpaiBuildStats {
    int tier_columns[kTiers] = {0, 0, 0, 0, 0};
    int max_gathered = 0;          // max_k T_k
    int max_n2 = 0;                // max_k |J_k|
    int empty_columns = 0;
    long long sum_gathered = 0;    // sum_k T_k (entries of A read by the gather)
    int static_tier = -1;
    // device ms of the phases: planner, register tiers, LDS tier, global tier
    float phase_ms[4] = {0.0f, 0.0f, 0.0f, 0.0f};
};

int classify_host(int T, int n2);

// C = A M and ||A m_k - e_k||^2 per column (the SPAI quality ||A M - I||_F).
// Column k of C is A(:, J_k) m_k on the row set I_k; ccolptr (prefix of
// |I_k|, n + 1) may be null to compute the residuals only.
struct SpaiProductParams {
    int n = 0;
    const int *colptr = nullptr, *rowidx = nullptr;  // A (CSC, device)
    const double *val = nullptr;
    const int *pcolptr = nullptr, *prowidx = nullptr;  // M's pattern (CSC, device)
    const void *mval = nullptr;                        // M's values, pattern order
    int m_fp32 = 0;
    int max_gathered = 0;  // max_k T_k and max_k |J_k| (planner statistics)
    int max_n2 = 0;
    const int *ccolptr = nullptr;
    int *crow = nullptr;
    double *cval = nullptr;
    double *colres2 = nullptr;  // n doubles
    int *n1_out = nullptr;      // |I_k| per column (n ints, nullable)
};
void spai_product(const SpaiProductParams &p, hipStream_t st, int cus, DeviceArray<uint64_t> &keys,
                  DeviceArray<int> &ints);

class SpaiBuilder {
public:
    SpaiBuilder();
    ~SpaiBuilder();
    SpaiBuilder(const SpaiBuilder &) = delete;  // owns pinned host buffers
    SpaiBuilder &operator=(const SpaiBuilder &) = delete;
    // Enqueues the build on `st` and waits for it (the planner's counters are
    // read back at the end); throws spai::Error on failure.
    void build(const SpaiBuildParams &p, hipStream_t st, SpaiBuildStats *stats);

private:
    DeviceArray<int> counters_;
    DeviceArray<unsigned long long> stats64_;
    DeviceArray<int> lists_;
    DeviceArray<double> glob_tile_, glob_vec_;
    DeviceArray<uint64_t> glob_keys_;
    DeviceArray<int> glob_ints_;
    int cus_ = 0;
    // pinned host copies of the planner's counters and statistics (read back
    // mid-build; a copy into pageable memory stalls the stream far longer)
    int *h_counters_ = nullptr;
    unsigned long long *h_stats_ = nullptr;
    // phase boundaries: planner start/end, tiers start, after the register
    // tiers, after the LDS tier, global tier start/end; then the planner
    // counters' copy to the host
    hipEvent_t ev_[8] = {};
};

}  // namespace spai
