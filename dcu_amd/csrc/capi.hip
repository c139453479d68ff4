// C ABI (spai_mi355x.h) and the reference-shaped C++ entry points
// (spai_multi_dcu_api.h) on top of the SPAI builder and the BiCGSTAB solver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "krylov.hpp"
#include "problems.hpp"
#include "spai_build.hpp"
#include "spai_mi355x.h"
#include "spai_multi_dcu_api.h"

using namespace spai;

namespace {
thread_local std::string g_last_error;

int fail(int status, const std::string &msg) {
    g_last_error = msg;
    return status;
}

template <class F>
int guarded(F &&f) {
    try {
        f();
        return SPAI_OK;
    } catch (const Error &e) {
        return fail(e.status, e.what());
    } catch (const std::bad_alloc &) {
        return fail(SPAI_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(SPAI_ERR_INVALID, e.what());
    }
}

__global__ void k_fill_f64(double *p, double v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float *p, float v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Plain copy of `n` doubles, 16 bytes per lane per step (calibration kernel).
// Streaming copy: 4 x 16 B per thread per step, nontemporal loads and stores
// (the practical HBM ceiling the solver kernels are compared with).
typedef double f64x2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy_calibration(const f64x2 *__restrict__ src, f64x2 *__restrict__ dst,
                                                            long long n2) {
    const long long stride = (long long)gridDim.x * blockDim.x * 4;
    for (long long base = (long long)blockIdx.x * blockDim.x * 4 + threadIdx.x; base < n2; base += stride) {
        f64x2 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long i = base + (long long)k * blockDim.x;
            if (i < n2) v[k] = __builtin_nontemporal_load(&src[i]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long i = base + (long long)k * blockDim.x;
            if (i < n2) __builtin_nontemporal_store(v[k], &dst[i]);
        }
    }
}

void fill(double *p, long long n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, 0.0, n);
    SPAI_CHECK_LAUNCH();
}
void fill(float *p, long long n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, 0.0f, n);
    SPAI_CHECK_LAUNCH();
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    explicit Timer(hipStream_t s) : st(s) {
        SPAI_HIP(hipEventCreate(&a));
        SPAI_HIP(hipEventCreate(&b));
        SPAI_HIP(hipEventRecord(a, st));
    }
    float stop() {
        SPAI_HIP(hipEventRecord(b, st));
        SPAI_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        SPAI_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Host-side structural checks of a CSC matrix before it is uploaded.
void validate_host_csc(const CSC_Matrix *A) {
    require(A != nullptr, "matrix is NULL");
    int n = A->nCol;
    require(n > 0, "matrix has no columns");
    require(A->nRow == n, "matrix is not square (nRow != nCol)");
    require(A->mPtr != nullptr, "mPtr is NULL");
    require(A->mPtr[0] == 0, "mPtr[0] must be 0");
    for (int c = 0; c < n; ++c) require(A->mPtr[c + 1] >= A->mPtr[c], "mPtr is not non-decreasing");
    int nnz = A->mPtr[n];
    require(A->nonzeroes == nnz, "nonzeroes does not match mPtr[nCol]");
    require(nnz == 0 || (A->mIndex && A->mData), "mIndex/mData are NULL");
    for (int p = 0; p < nnz; ++p) require(A->mIndex[p] >= 0 && A->mIndex[p] < n, "row index out of range");
}
}  // namespace

// Device scratch of the allocation-free solve (spai_system_solve_noalloc):
// the solver's work vectors, per-system states, counters and reduction
// partials for callers that cannot allocate device memory.  Each device the
// library runs on holds its own copy from its first context until the
// process ends; every allocation-free solve on a device shares that copy.
// In host code `scratch` is only a stand-in: its device address comes from
// hipGetSymbolAddress (scratch_address below), once per device.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_rc[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current; looked
// up on the first call for each device and kept.
hipError_t scratch_address(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void *ptr = nullptr;
        g_scratch_rc[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = g_scratch_rc[device] == hipSuccess ? ptr : nullptr;
        g_scratch_done[device] = true;
    }
    *out = g_scratch_ptr[device];
    return g_scratch_rc[device];
}
}  // namespace

struct spai_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // systems hold a pointer to their context: destroying the context while
    // systems are alive defers the release to the last system's destroy
    int live_systems = 0;
    bool release_pending = false;
    int cus = 0;
    std::string name;
    SpaiBuilder builder;
    BiCGStab solver;
    BiCGStab adjoint_solver;
    // the allocation-free solve: a solver bound to this device's scratch
    BiCGStab noalloc_solver;
    void *scratch = nullptr;
    std::string scratch_error;  // why the scratch address is missing
};

struct spai_system {
    spai_context *ctx = nullptr;
    int n = 0, nnz = 0, nsys = 1;
    std::vector<int> h_row_off;
    std::vector<long long> h_nnz_off;
    DeviceArray<int> colptr, rowidx, rowptr, colidx, perm, n1;
    DeviceArray<double> cscval, csrval;
    // the solver's layout (tiles: CSR order; SELL-64: padded slice order)
    DeviceArray<int> perm_lay;   // CSC position of M's pattern -> position in M's layout value order
    DeviceArray<double> sellval;  // A's values in SELL order (SELL only)
    bool sell_layout = false;     // the layout chosen for A at creation
    // M's sparsity pattern: A's own (SPAI_PATTERN_A) or one of the reference's
    // alternatives (GetSparsityPattern flags 0..2)
    int pattern = SPAI_PATTERN_A;
    int m_nnz = 0;
    DeviceArray<int> pcolptr, prowidx;        // pattern CSC (pattern != A)
    DeviceArray<int> mrowptr, mcolidx, mperm;  // its CSR and CSC -> CSR positions
    SellOperator msell;                        // its SELL-64 arrays (SELL layout)
    DeviceArray<double> m64_csc, m64_lay;
    DeviceArray<float> m32_csc, m32_lay;
    int m_precision = -1;
    TileMap tiles, tiles_t;
    bool tiles_t_ready = false;
    // the adjoint's operators A^T and M^T on a SELL-64 tiles_t: values in
    // its order, refreshed when A's values or M change
    DeviceArray<double> at_lay, mt64_lay;
    DeviceArray<float> mt32_lay;
    bool at_ready = false;
    unsigned long long m_version = 0, mt_version = 0;  // builds of M; the build mt*_lay holds
    DeviceArray<double> stage_b, stage_x, stage_g, stage_dA, stage_db;
    float setup_ms = 0.0f;
    int tile_shape[kTunedTiers] = {0, 0, 0};  // register-tier group shapes of the next builds
    std::vector<SysState> last_states;  // nsys entries from creation on
    bool have_states = false;           // a solve or adjoint has filled them
    const BiCGStab *last_solver = nullptr;
};

struct spai_problem {
    Problem p;
};

// ----------------------------------------------------------- helpers ----
namespace {
void finish_setup(spai_system *S, hipStream_t st) {
    S->rowptr.resize((size_t)S->n + 1);
    S->colidx.resize((size_t)std::max(S->nnz, 1));
    S->csrval.resize((size_t)std::max(S->nnz, 1));
    S->perm.resize((size_t)std::max(S->nnz, 1));
    csc_to_csr_device(S->n, S->n, S->nnz, S->colptr.get(), S->rowidx.get(), S->cscval.get(), S->rowptr.get(),
                      S->colidx.get(), S->csrval.get(), S->perm.get(), st);
    build_layout(S->tiles, S->rowptr.get(), S->colidx.get(), S->h_row_off, S->h_nnz_off, true, st, nullptr, true);
    S->sell_layout = S->tiles.kind == kLayoutSell;
    S->m_nnz = S->nnz;
    S->perm_lay.resize((size_t)std::max(S->nnz, 1));
    layout_perm(S->tiles, S->nnz, S->perm.get(), S->perm_lay.get(), st);
    if (S->tiles.kind == kLayoutSell) {
        S->sellval.resize((size_t)S->tiles.sell_len);
        layout_values<double>(S->tiles, S->nnz, S->csrval.get(), S->sellval.get(), st);
    }
    SPAI_HIP(hipStreamSynchronize(st));
}

// Pattern of M for the reference's GetSparsityPattern flags
// (cuFormatConversion.h:598-974), column by column, rows sorted ascending:
//   0  A^T          column k = row k of A
//   1  E + A        column k of A plus the diagonal entry
//   2  E + A + A^T  column k of A, row k of A and the diagonal entry
// Built on the host from A's CSC and CSR patterns (a set-up step).
void pattern_host(int n, const std::vector<int> &cp, const std::vector<int> &ri, const std::vector<int> &rp,
                  const std::vector<int> &ci, int flag, std::vector<int> &pcp, std::vector<int> &pri) {
    pcp.assign((size_t)n + 1, 0);
    pri.clear();
    std::vector<int> col;
    for (int k = 0; k < n; ++k) {
        col.clear();
        if (flag == SPAI_PATTERN_AT || flag == SPAI_PATTERN_I_PLUS_A_PLUS_AT)
            col.insert(col.end(), ci.begin() + rp[(size_t)k], ci.begin() + rp[(size_t)k + 1]);
        if (flag != SPAI_PATTERN_AT) {
            col.insert(col.end(), ri.begin() + cp[(size_t)k], ri.begin() + cp[(size_t)k + 1]);
            col.push_back(k);
        }
        std::sort(col.begin(), col.end());
        col.erase(std::unique(col.begin(), col.end()), col.end());
        pri.insert(pri.end(), col.begin(), col.end());
        require(pri.size() < (size_t)INT32_MAX, "pattern: more than 2^31-1 entries");
        pcp[(size_t)k + 1] = (int)pri.size();
    }
}

// A in the solver's layout
OperatorView a_view(spai_system *S) {
    if (S->tiles.kind == kLayoutSell) {
        OperatorView A{nullptr, S->tiles.sell_col.get(), S->sellval.get(), 0};
        A.off = S->tiles.sell_off.get();
        A.diag = S->tiles.sell_diag.get();
        return A;
    }
    return OperatorView{S->rowptr.get(), S->colidx.get(), S->csrval.get(), 0};
}

spai_system *create_from_host(spai_context *ctx, const std::vector<const CSC_Matrix *> &parts) {
    auto S = std::make_unique<spai_system>();
    S->ctx = ctx;
    S->nsys = (int)parts.size();
    S->h_row_off.assign((size_t)S->nsys + 1, 0);
    S->h_nnz_off.assign((size_t)S->nsys + 1, 0);
    for (int s = 0; s < S->nsys; ++s) {
        validate_host_csc(parts[(size_t)s]);
        S->h_row_off[(size_t)s + 1] = S->h_row_off[(size_t)s] + parts[(size_t)s]->nCol;
        S->h_nnz_off[(size_t)s + 1] = S->h_nnz_off[(size_t)s] + parts[(size_t)s]->nonzeroes;
    }
    require(S->h_nnz_off.back() < (1LL << 31), "batch has more than 2^31-1 nonzeros");
    S->n = S->h_row_off.back();
    S->nnz = (int)S->h_nnz_off.back();
    hipStream_t st = ctx->stream;
    SPAI_HIP(hipSetDevice(ctx->device));
    // block-diagonal concatenation with global indices
    std::vector<int> cp((size_t)S->n + 1), ri((size_t)std::max(S->nnz, 1));
    std::vector<double> va((size_t)std::max(S->nnz, 1));
    for (int s = 0; s < S->nsys; ++s) {
        const CSC_Matrix *A = parts[(size_t)s];
        int r0 = S->h_row_off[(size_t)s];
        long long z0 = S->h_nnz_off[(size_t)s];
        for (int c = 0; c < A->nCol; ++c) cp[(size_t)r0 + c] = (int)(z0 + A->mPtr[c]);
        for (int p = 0; p < A->nonzeroes; ++p) {
            ri[(size_t)(z0 + p)] = A->mIndex[p] + r0;
            va[(size_t)(z0 + p)] = A->mData[p];
        }
    }
    cp[(size_t)S->n] = S->nnz;
    Timer tm(st);
    S->colptr.resize((size_t)S->n + 1);
    S->rowidx.resize((size_t)std::max(S->nnz, 1));
    S->cscval.resize((size_t)std::max(S->nnz, 1));
    SPAI_HIP(hipMemcpyAsync(S->colptr.get(), cp.data(), sizeof(int) * cp.size(), hipMemcpyHostToDevice, st));
    SPAI_HIP(hipMemcpyAsync(S->rowidx.get(), ri.data(), sizeof(int) * (size_t)S->nnz, hipMemcpyHostToDevice, st));
    SPAI_HIP(hipMemcpyAsync(S->cscval.get(), va.data(), sizeof(double) * (size_t)S->nnz, hipMemcpyHostToDevice, st));
    finish_setup(S.get(), st);
    S->setup_ms = tm.stop();
    S->last_states.resize((size_t)S->nsys);
    return S.release();
}

// The adjoint's layout over the CSR of A^T (A's CSC arrays): SELL-64 when M
// has A's pattern (M^T then shares A^T's slices) and the rows suit it, row
// tiles otherwise.
void ensure_transposed_tiles(spai_system *S) {
    if (S->tiles_t_ready) return;
    const bool own = S->pattern != SPAI_PATTERN_A;
    build_layout(S->tiles_t, S->colptr.get(), S->rowidx.get(), S->h_row_off, S->h_nnz_off, !own, S->ctx->stream,
                 own ? S->pcolptr.get() : nullptr);
    S->tiles_t_ready = true;
    S->at_ready = false;
    S->mt_version = 0;
}

// A^T for the adjoint solve in tiles_t's layout
OperatorView at_view(spai_system *S) {
    if (S->tiles_t.kind != kLayoutSell) return OperatorView{S->colptr.get(), S->rowidx.get(), S->cscval.get(), 0};
    if (!S->at_ready) {
        S->at_lay.resize((size_t)S->tiles_t.sell_len);
        layout_values<double>(S->tiles_t, S->nnz, S->cscval.get(), S->at_lay.get(), S->ctx->stream);
        S->at_ready = true;
    }
    OperatorView A{nullptr, S->tiles_t.sell_col.get(), S->at_lay.get(), 0};
    A.off = S->tiles_t.sell_off.get();
    A.diag = S->tiles_t.sell_diag.get();
    return A;
}

SolveOptions to_options(const spai_solve_options *o) {
    SolveOptions s;
    if (o) {
        s.rel_tol = o->rel_tol;
        s.abs_tol = o->abs_tol;
        s.max_iter = o->max_iter;
        s.use_graph = o->use_graph;
        s.graph_iters = o->graph_iters;
        s.check_every = o->check_every;
        s.record_history = o->record_history;
        s.true_residual_restart = o->true_residual_restart;
    }
    return s;
}

void fill_report(spai_solve_report *r, const SolveResult &res) {
    if (!r) return;
    r->ms = res.ms;
    r->iterations_max = res.iterations_max;
    r->iterations_sum = res.iterations_sum;
    r->host_checks = res.host_checks;
    r->all_converged = res.all_converged;
}

// M in the solver's layout; transposed: M^T in tiles_t's (its CSC arrays =
// the CSR of M^T for row tiles; on SELL-64 tiles_t, which only M with A's
// pattern gets, A^T's slices with M's CSC values in their order)
OperatorView m_view(spai_system *S, bool transposed) {
    if (transposed && S->tiles_t.kind == kLayoutSell) {
        const bool fp32 = S->m_precision == SPAI_PRECISION_MIXED;
        if (S->mt_version != S->m_version) {
            const size_t len = (size_t)S->tiles_t.sell_len;
            if (fp32) {
                S->mt32_lay.resize(len);
                layout_values<float, float>(S->tiles_t, S->nnz, S->m32_csc.get(), S->mt32_lay.get(), S->ctx->stream);
            } else {
                S->mt64_lay.resize(len);
                layout_values<double>(S->tiles_t, S->nnz, S->m64_csc.get(), S->mt64_lay.get(), S->ctx->stream);
            }
            S->mt_version = S->m_version;
        }
        OperatorView M{nullptr, S->tiles_t.sell_col.get(), nullptr, fp32 ? 1 : 0};
        M.val = fp32 ? (const void *)S->mt32_lay.get() : (const void *)S->mt64_lay.get();
        M.off = S->tiles_t.sell_off.get();
        M.diag = S->tiles_t.sell_diag.get();
        return M;
    }
    OperatorView M;
    const bool sell = !transposed && S->tiles.kind == kLayoutSell;
    const bool own = S->pattern != SPAI_PATTERN_A;
    if (transposed) {
        M.ptr = own ? S->pcolptr.get() : S->colptr.get();
        M.idx = own ? S->prowidx.get() : S->rowidx.get();
    } else if (sell) {
        M.idx = own ? S->msell.col.get() : S->tiles.sell_col.get();
        M.off = own ? S->msell.off.get() : S->tiles.sell_off.get();
        M.diag = own ? nullptr : S->tiles.sell_diag.get();
        M.blk = own ? S->msell.slice_blk.get() : nullptr;
    } else {
        M.ptr = own ? S->mrowptr.get() : S->rowptr.get();
        M.idx = own ? S->mcolidx.get() : S->colidx.get();
    }
    if (S->m_precision == SPAI_PRECISION_MIXED) {
        M.val = transposed ? (const void *)S->m32_csc.get() : (const void *)S->m32_lay.get();
        M.fp32 = 1;
    } else {
        M.val = transposed ? (const void *)S->m64_csc.get() : (const void *)S->m64_lay.get();
        M.fp32 = 0;
    }
    return M;
}
}  // namespace

// ============================================================ C ABI ====
extern "C" {

const char *spai_status_string(int status) {
    switch (status) {
    case SPAI_OK: return "ok";
    case SPAI_ERR_INVALID: return "invalid argument";
    case SPAI_ERR_HIP: return "HIP runtime error";
    case SPAI_ERR_NOMEM: return "out of memory";
    case SPAI_ERR_UNSUPPORTED: return "unsupported input";
    case SPAI_ERR_BREAKDOWN: return "BiCGSTAB breakdown";
    case SPAI_ERR_NOT_BUILT: return "preconditioner not built";
    default: return "unknown status";
    }
}

const char *spai_last_error(void) { return g_last_error.c_str(); }

const char *spai_version(void) { return "spai_mi355x 0.3 (gfx950)"; }

int spai_context_create(int device, spai_context **out) {
    if (!out) return fail(SPAI_ERR_INVALID, "out is NULL");
    *out = nullptr;
    return guarded([&] {
        int count = 0;
        SPAI_HIP(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, "no HIP device " + std::to_string(device));
        auto c = std::make_unique<spai_context>();
        c->device = device;
        SPAI_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        SPAI_HIP(hipGetDeviceProperties(&prop, device));
        c->name = prop.name;
        if (c->name.empty()) c->name = prop.gcnArchName;
        c->cus = prop.multiProcessorCount;
        SPAI_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        // set-up of the allocation-free solve: this device's scratch address
        // (a failed lookup only disables that entry point)
        void *ptr = nullptr;
        const hipError_t rc = scratch_address(device, &ptr);
        if (rc == hipSuccess && ptr) {
            c->scratch = ptr;
            c->noalloc_solver.bind_workspace(ptr, kNoAllocScratchBytes);
        } else {
            c->scratch_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc);
        }
        *out = c.release();
    });
}

namespace {
void release_context(spai_context *ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    hipStream_t st = ctx->stream;
    delete ctx;
    (void)hipStreamDestroy(st);
}
}  // namespace

int spai_context_destroy(spai_context *ctx) {
    if (!ctx) return SPAI_OK;
    return guarded([&] {
        if (ctx->live_systems > 0) {
            ctx->release_pending = true;
            return;
        }
        release_context(ctx);
    });
}

int spai_context_device_info(spai_context *ctx, char *name, int name_len, int *cus) {
    if (!ctx) return fail(SPAI_ERR_INVALID, "ctx is NULL");
    if (name && name_len > 0) {
        std::strncpy(name, ctx->name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cus) *cus = ctx->cus;
    return SPAI_OK;
}

int spai_system_create(spai_context *ctx, const CSC_Matrix *host_A, spai_system **out) {
    if (!ctx || !out) return fail(SPAI_ERR_INVALID, "ctx/out is NULL");
    *out = nullptr;
    return guarded([&] {
        *out = create_from_host(ctx, {host_A});
        ctx->live_systems++;
    });
}

int spai_system_create_batched(spai_context *ctx, int nsys, const CSC_Matrix *host_systems, spai_system **out) {
    if (!ctx || !out || nsys < 1 || !host_systems) return fail(SPAI_ERR_INVALID, "bad batched arguments");
    *out = nullptr;
    return guarded([&] {
        std::vector<const CSC_Matrix *> parts;
        for (int s = 0; s < nsys; ++s) parts.push_back(host_systems + s);
        *out = create_from_host(ctx, parts);
        ctx->live_systems++;
    });
}

int spai_system_create_device(spai_context *ctx, const CSC_Matrix *dev_A, spai_system **out) {
    if (!ctx || !out || !dev_A) return fail(SPAI_ERR_INVALID, "ctx/out/A is NULL");
    *out = nullptr;
    return guarded([&] {
        require(dev_A->nCol > 0 && dev_A->nRow == dev_A->nCol, "device matrix must be square and non-empty");
        require(dev_A->mPtr && dev_A->mIndex && dev_A->mData, "device matrix pointers are NULL");
        SPAI_HIP(hipSetDevice(ctx->device));
        auto S = std::make_unique<spai_system>();
        S->ctx = ctx;
        S->n = dev_A->nCol;
        hipStream_t st = ctx->stream;
        int nnz = 0;
        SPAI_HIP(hipMemcpyAsync(&nnz, dev_A->mPtr + S->n, sizeof(int), hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
        require(nnz >= 0 && nnz == dev_A->nonzeroes, "device mPtr[n] does not match nonzeroes");
        S->nnz = nnz;
        S->h_row_off = {0, S->n};
        S->h_nnz_off = {0, (long long)nnz};
        Timer tm(st);
        S->colptr.resize((size_t)S->n + 1);
        S->rowidx.resize((size_t)std::max(nnz, 1));
        S->cscval.resize((size_t)std::max(nnz, 1));
        SPAI_HIP(hipMemcpyAsync(S->colptr.get(), dev_A->mPtr, sizeof(int) * ((size_t)S->n + 1),
                                hipMemcpyDeviceToDevice, st));
        SPAI_HIP(hipMemcpyAsync(S->rowidx.get(), dev_A->mIndex, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice,
                                st));
        SPAI_HIP(hipMemcpyAsync(S->cscval.get(), dev_A->mData, sizeof(double) * (size_t)nnz,
                                hipMemcpyDeviceToDevice, st));
        finish_setup(S.get(), st);
        S->setup_ms = tm.stop();
        S->last_states.resize(1);
        *out = S.release();
        ctx->live_systems++;
    });
}

int spai_system_destroy(spai_system *sys) {
    if (!sys) return SPAI_OK;
    return guarded([&] {
        spai_context *ctx = sys->ctx;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        // graphs captured on this system's layouts are never replayed again
        ctx->solver.forget_layout(sys->tiles.generation);
        ctx->adjoint_solver.forget_layout(sys->tiles_t.generation);
        delete sys;
        if (--ctx->live_systems == 0 && ctx->release_pending) release_context(ctx);
    });
}

int spai_system_get_info(const spai_system *sys, spai_system_info *info) {
    if (!sys || !info) return fail(SPAI_ERR_INVALID, "sys/info is NULL");
    info->n = sys->n;
    info->nnz = sys->nnz;
    info->nsys = sys->nsys;
    info->m_precision = sys->m_precision;
    info->setup_ms = sys->setup_ms;
    return SPAI_OK;
}

int spai_system_row_offsets(const spai_system *sys, int *row_offsets) {
    if (!sys || !row_offsets) return fail(SPAI_ERR_INVALID, "sys/row_offsets is NULL");
    std::memcpy(row_offsets, sys->h_row_off.data(), sizeof(int) * sys->h_row_off.size());
    return SPAI_OK;
}

int spai_system_set_values(spai_system *sys, const double *values, int values_on_device) {
    if (!sys || !values) return fail(SPAI_ERR_INVALID, "sys/values is NULL");
    return guarded([&] {
        hipStream_t st = sys->ctx->stream;
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        SPAI_HIP(hipMemcpyAsync(sys->cscval.get(), values, sizeof(double) * (size_t)sys->nnz,
                                values_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        // CSR values follow through the permutation: csr[perm[p]] = csc[p]
        std::vector<int> hp((size_t)sys->nnz);
        std::vector<double> hv((size_t)sys->nnz), hc((size_t)sys->nnz);
        SPAI_HIP(hipMemcpyAsync(hp.data(), sys->perm.get(), sizeof(int) * (size_t)sys->nnz, hipMemcpyDeviceToHost,
                                st));
        SPAI_HIP(hipMemcpyAsync(hv.data(), sys->cscval.get(), sizeof(double) * (size_t)sys->nnz,
                                hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
        for (int p = 0; p < sys->nnz; ++p) hc[(size_t)hp[(size_t)p]] = hv[(size_t)p];
        SPAI_HIP(hipMemcpyAsync(sys->csrval.get(), hc.data(), sizeof(double) * (size_t)sys->nnz,
                                hipMemcpyHostToDevice, st));
        if (sys->tiles.kind == kLayoutSell)
            layout_values<double>(sys->tiles, sys->nnz, sys->csrval.get(), sys->sellval.get(), st);
        sys->at_ready = false;  // the adjoint's A^T values follow on its next call
        SPAI_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"

namespace {
// One SPAI build of M with the register-tier shapes `shape`; tiers_ms
// (nullable) receives the tier kernels' device time.  Throws on failure.
void build_m(spai_system *sys, int precision, int schedule, int col_begin, int col_end, const int *shape,
             float *tiers_ms, spai_build_report *report) {
    hipStream_t st = sys->ctx->stream;
    SPAI_HIP(hipSetDevice(sys->ctx->device));
    const bool own = sys->pattern != SPAI_PATTERN_A;
    size_t nz = (size_t)std::max(sys->m_nnz, 1);
    const bool sell = sys->tiles.kind == kLayoutSell;
    const long long lay =
        std::max<long long>(sell ? (own ? sys->msell.len : sys->tiles.sell_len) : (long long)sys->m_nnz, 1);
    if (precision == SPAI_PRECISION_MIXED) {
        sys->m32_csc.resize(nz);
        sys->m32_lay.resize((size_t)lay);
        if (sell) fill(sys->m32_lay.get(), lay, st);  // SELL padding entries stay 0
    } else {
        sys->m64_csc.resize(nz);
        sys->m64_lay.resize((size_t)lay);
        if (sell) fill(sys->m64_lay.get(), lay, st);
    }
    sys->n1.resize((size_t)sys->n);
    int cb = std::max(0, col_begin), ce = col_end < 0 ? sys->n : std::min(col_end, sys->n);
    require(cb <= ce, "column range is inverted");
    sys->m_precision = -1;
    Timer tm(st);
    bool partial = cb > 0 || ce < sys->n;
    if (partial) {
        if (precision == SPAI_PRECISION_MIXED) {
            fill(sys->m32_csc.get(), sys->m_nnz, st);
            if (!sell) fill(sys->m32_lay.get(), sys->m_nnz, st);
        } else {
            fill(sys->m64_csc.get(), sys->m_nnz, st);
            if (!sell) fill(sys->m64_lay.get(), sys->m_nnz, st);
        }
    }
    SpaiBuildParams p;
    p.n = sys->n;
    p.colptr = sys->colptr.get();
    p.rowidx = sys->rowidx.get();
    p.val = sys->cscval.get();
    p.perm = sys->perm_lay.get();
    if (own) {
        p.pcolptr = sys->pcolptr.get();
        p.prowidx = sys->prowidx.get();
    }
    p.precision = precision;
    p.schedule = schedule;
    p.col_begin = cb;
    p.col_end = ce;
    p.m_csc = precision == SPAI_PRECISION_MIXED ? (void *)sys->m32_csc.get() : (void *)sys->m64_csc.get();
    p.m_csr = precision == SPAI_PRECISION_MIXED ? (void *)sys->m32_lay.get() : (void *)sys->m64_lay.get();
    p.n1_out = sys->n1.get();
    for (int t = 0; t < kTunedTiers; ++t) p.shape[t] = shape[t];
    p.tiers_ms = tiers_ms;
    SpaiBuildStats stats;
    sys->ctx->builder.build(p, st, &stats);
    float ms = tm.stop();
    sys->m_precision = precision;
    sys->m_version += 1;
    if (report) {
        report->ms = ms;
        for (int t = 0; t < 5; ++t) report->tier_columns[t] = stats.tier_columns[t];
        report->max_gathered = stats.max_gathered;
        report->max_n2 = stats.max_n2;
        report->empty_columns = stats.empty_columns;
        report->sum_gathered = stats.sum_gathered;
        report->static_tier = stats.static_tier;
        for (int t = 0; t < kTunedTiers; ++t) report->tile_shape[t] = shape[t];
        for (int q = 0; q < 4; ++q) report->phase_ms[q] = stats.phase_ms[q];
    }
}

int check_build_args(const spai_system *sys, int precision, int schedule) {
    if (!sys) return fail(SPAI_ERR_INVALID, "sys is NULL");
    if (precision != SPAI_PRECISION_FP64 && precision != SPAI_PRECISION_MIXED)
        return fail(SPAI_ERR_INVALID, "unknown precision");
    if (schedule != SPAI_SCHEDULE_BALANCED_COLUMNS && schedule != SPAI_SCHEDULE_BALANCED_NNZ)
        return fail(SPAI_ERR_INVALID, "unknown schedule");
    return SPAI_OK;
}
}  // namespace

extern "C" {

int spai_system_build(spai_system *sys, int precision, int schedule, int col_begin, int col_end,
                      spai_build_report *report) {
    if (int rc = check_build_args(sys, precision, schedule)) return rc;
    return guarded([&] { build_m(sys, precision, schedule, col_begin, col_end, sys->tile_shape, nullptr, report); });
}

int spai_tile_shape_count(int tier) { return tile_shape_count(tier); }

int spai_tile_shape_get(int tier, int candidate, spai_tile_shape *out) {
    if (!out || candidate < 0 || candidate >= tile_shape_count(tier))
        return fail(SPAI_ERR_INVALID, "no such tier or candidate");
    const TileShape t = tile_shape(tier, candidate);
    out->lanes = t.lanes;
    out->entries_per_lane = t.entries_per_lane;
    out->max_cols = t.max_cols;
    out->rows_per_lane = t.rows_per_lane;
    return SPAI_OK;
}

int spai_system_set_tile_shapes(spai_system *sys, const int *shape) {
    if (!sys || !shape) return fail(SPAI_ERR_INVALID, "bad arguments");
    for (int t = 0; t < kTunedTiers; ++t)
        if (shape[t] < 0 || shape[t] >= tile_shape_count(t)) return fail(SPAI_ERR_INVALID, "no such tile shape");
    for (int t = 0; t < kTunedTiers; ++t) sys->tile_shape[t] = shape[t];
    return SPAI_OK;
}

int spai_system_get_tile_shapes(const spai_system *sys, int *shape) {
    if (!sys || !shape) return fail(SPAI_ERR_INVALID, "bad arguments");
    for (int t = 0; t < kTunedTiers; ++t) shape[t] = sys->tile_shape[t];
    return SPAI_OK;
}

// Coordinate search, one register tier (size bucket) at a time from the
// smallest: every candidate of the tier is timed (one untimed build, then
// the best of `reps` timed ones) with the shapes chosen so far for the
// others; ties keep the lower index.  The tiers' kernel time is the
// measure, since the shapes change nothing else of the build.  A tier with
// no columns under the default shapes (whose rows per group are the
// smallest, so they hand on the most columns) keeps its current shape.
int spai_system_autotune(spai_system *sys, int precision, int schedule, int reps, spai_tuning *out) {
    if (int rc = check_build_args(sys, precision, schedule)) return rc;
    if (reps < 1 || reps > 100) return fail(SPAI_ERR_INVALID, "reps must be in 1..100");
    return guarded([&] {
        const auto t0 = std::chrono::steady_clock::now();
        spai_tuning res{};
        for (int t = 0; t < SPAI_TUNE_TIERS; ++t)
            for (int c = 0; c < SPAI_TUNE_MAX_CANDIDATES; ++c) res.ms[t][c] = -1.0f;
        static_assert(SPAI_TUNE_TIERS == kTunedTiers, "one entry per register tier");
        int shape[kTunedTiers] = {0, 0, 0};
        spai_build_report rep{};
        build_m(sys, precision, schedule, 0, -1, shape, nullptr, &rep);
        for (int t = 0; t < kTunedTiers; ++t) {
            res.columns[t] = rep.tier_columns[t];
            shape[t] = sys->tile_shape[t];
        }
        for (int t = 0; t < kTunedTiers; ++t) {
            const int nc = tile_shape_count(t);
            require(nc <= SPAI_TUNE_MAX_CANDIDATES, "too many tile-shape candidates");
            if (res.columns[t] == 0) continue;
            int best = -1;
            for (int c = 0; c < nc; ++c) {
                shape[t] = c;
                build_m(sys, precision, schedule, 0, -1, shape, nullptr, nullptr);  // code load, caches
                float m = 0.0f;
                for (int r = 0; r < reps; ++r) {
                    float ms = 0.0f;
                    build_m(sys, precision, schedule, 0, -1, shape, &ms, nullptr);
                    m = r == 0 ? ms : std::min(m, ms);
                }
                res.ms[t][c] = m;
                if (best < 0 || m < res.ms[t][best]) best = c;
            }
            shape[t] = best;
        }
        for (int t = 0; t < kTunedTiers; ++t) {
            sys->tile_shape[t] = shape[t];
            res.shape[t] = shape[t];
        }
        build_m(sys, precision, schedule, 0, -1, shape, nullptr, nullptr);  // M from the chosen shapes
        res.tune_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (out) *out = res;
    });
}

int spai_system_set_pattern(spai_system *sys, int pattern) {
    if (!sys) return fail(SPAI_ERR_INVALID, "sys is NULL");
    if (pattern < SPAI_PATTERN_A || pattern > SPAI_PATTERN_I_PLUS_A_PLUS_AT) return fail(SPAI_ERR_INVALID, "unknown pattern");
    return guarded([&] {
        hipStream_t st = sys->ctx->stream;
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        const int n = sys->n, nnz = sys->nnz;
        sys->m_precision = -1;  // M must be built again on the new pattern
        sys->tiles_t_ready = false;
        if (pattern == SPAI_PATTERN_A) {
            sys->pattern = pattern;
            sys->m_nnz = nnz;
            if (!sys->sell_layout)
                build_layout(sys->tiles, sys->rowptr.get(), sys->colidx.get(), sys->h_row_off, sys->h_nnz_off, false,
                             st, nullptr, true);
            layout_perm(sys->tiles, nnz, sys->perm.get(), sys->perm_lay.get(), st);
            SPAI_HIP(hipStreamSynchronize(st));
            return;
        }
        std::vector<int> cp((size_t)n + 1), ri((size_t)std::max(nnz, 1)), rp((size_t)n + 1),
            ci((size_t)std::max(nnz, 1));
        SPAI_HIP(hipMemcpyAsync(cp.data(), sys->colptr.get(), sizeof(int) * cp.size(), hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipMemcpyAsync(ri.data(), sys->rowidx.get(), sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipMemcpyAsync(rp.data(), sys->rowptr.get(), sizeof(int) * rp.size(), hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipMemcpyAsync(ci.data(), sys->colidx.get(), sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
        std::vector<int> pcp, pri;
        pattern_host(n, cp, ri, rp, ci, pattern, pcp, pri);
        const int mn = pcp[(size_t)n];
        sys->pcolptr.resize(pcp.size());
        sys->prowidx.resize((size_t)std::max(mn, 1));
        SPAI_HIP(hipMemcpyAsync(sys->pcolptr.get(), pcp.data(), sizeof(int) * pcp.size(), hipMemcpyHostToDevice, st));
        SPAI_HIP(hipMemcpyAsync(sys->prowidx.get(), pri.data(), sizeof(int) * (size_t)mn, hipMemcpyHostToDevice, st));
        sys->mrowptr.resize((size_t)n + 1);
        sys->mcolidx.resize((size_t)std::max(mn, 1));
        sys->mperm.resize((size_t)std::max(mn, 1));
        csc_to_csr_device(n, n, mn, sys->pcolptr.get(), sys->prowidx.get(), nullptr, sys->mrowptr.get(),
                          sys->mcolidx.get(), nullptr, sys->mperm.get(), st);
        sys->perm_lay.resize((size_t)std::max(mn, 1));
        if (sys->sell_layout) {
            build_sell_operator(sys->tiles, sys->mrowptr.get(), sys->mcolidx.get(), mn, sys->msell, st);
            compose_positions(mn, sys->mperm.get(), sys->msell.pos.get(), sys->perm_lay.get(), st);
        } else {
            // tiles cut so that neither A's nor M's rows overflow one
            build_layout(sys->tiles, sys->rowptr.get(), sys->colidx.get(), sys->h_row_off, sys->h_nnz_off, false, st,
                         sys->mrowptr.get());
            SPAI_HIP(hipMemcpyAsync(sys->perm_lay.get(), sys->mperm.get(), sizeof(int) * (size_t)mn,
                                    hipMemcpyDeviceToDevice, st));
        }
        SPAI_HIP(hipStreamSynchronize(st));
        sys->m_nnz = mn;
        sys->pattern = pattern;
    });
}

int spai_system_m_pattern(const spai_system *sys, int *nnz, int *colptr, int *rowidx) {
    if (!sys) return fail(SPAI_ERR_INVALID, "sys is NULL");
    return guarded([&] {
        if (nnz) *nnz = sys->m_nnz;
        const bool own = sys->pattern != SPAI_PATTERN_A;
        hipStream_t st = sys->ctx->stream;
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        if (colptr)
            SPAI_HIP(hipMemcpyAsync(colptr, own ? sys->pcolptr.get() : sys->colptr.get(), sizeof(int) * ((size_t)sys->n + 1),
                                    hipMemcpyDeviceToHost, st));
        if (rowidx)
            SPAI_HIP(hipMemcpyAsync(rowidx, own ? sys->prowidx.get() : sys->rowidx.get(), sizeof(int) * (size_t)sys->m_nnz,
                                    hipMemcpyDeviceToHost, st));
        SPAI_HIP(hipStreamSynchronize(st));
    });
}

namespace {
// Runs the product kernel over every column; with ccolptr set it also
// writes C.  Caps come from the current pattern over all columns.
void run_product(spai_system *sys, DeviceArray<double> &colres2, DeviceArray<int> &n1, const int *ccolptr,
                 int *crow, double *cval) {
    hipStream_t st = sys->ctx->stream;
    const int n = sys->n;
    const bool own = sys->pattern != SPAI_PATTERN_A;
    const int *pc = own ? sys->pcolptr.get() : sys->colptr.get();
    const int *pr = own ? sys->prowidx.get() : sys->rowidx.get();
    std::vector<int> acp((size_t)n + 1), pcp((size_t)n + 1), pri((size_t)std::max(sys->m_nnz, 1));
    SPAI_HIP(hipMemcpyAsync(acp.data(), sys->colptr.get(), sizeof(int) * acp.size(), hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipMemcpyAsync(pcp.data(), pc, sizeof(int) * pcp.size(), hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipMemcpyAsync(pri.data(), pr, sizeof(int) * (size_t)sys->m_nnz, hipMemcpyDeviceToHost, st));
    SPAI_HIP(hipStreamSynchronize(st));
    long long max_t = 0;
    int max_n2 = 0;
    for (int k = 0; k < n; ++k) {
        long long t = 0;
        for (int q = pcp[(size_t)k]; q < pcp[(size_t)k + 1]; ++q) t += acp[(size_t)pri[(size_t)q] + 1] - acp[(size_t)pri[(size_t)q]];
        max_t = std::max(max_t, t);
        max_n2 = std::max(max_n2, pcp[(size_t)k + 1] - pcp[(size_t)k]);
    }
    require(max_t < (1LL << 30), "product: a column gathers more than 2^30 entries");
    SpaiProductParams p;
    p.n = n;
    p.colptr = sys->colptr.get();
    p.rowidx = sys->rowidx.get();
    p.val = sys->cscval.get();
    p.pcolptr = pc;
    p.prowidx = pr;
    p.m_fp32 = sys->m_precision == SPAI_PRECISION_MIXED;
    p.mval = p.m_fp32 ? (const void *)sys->m32_csc.get() : (const void *)sys->m64_csc.get();
    p.max_gathered = (int)max_t;
    p.max_n2 = max_n2;
    colres2.resize((size_t)std::max(n, 1));
    n1.resize((size_t)std::max(n, 1));
    p.colres2 = colres2.get();
    p.n1_out = n1.get();
    p.ccolptr = ccolptr;
    p.crow = crow;
    p.cval = cval;
    DeviceArray<uint64_t> keys;
    DeviceArray<int> ints;
    spai_product(p, st, sys->ctx->cus, keys, ints);
}
}  // namespace

int spai_system_quality(spai_system *sys, double *frobenius, double *column_residuals) {
    if (!sys || !frobenius) return fail(SPAI_ERR_INVALID, "sys/frobenius is NULL");
    if (sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        DeviceArray<double> colres2;
        DeviceArray<int> n1;
        run_product(sys, colres2, n1, nullptr, nullptr, nullptr);
        std::vector<double> h((size_t)sys->n);
        SPAI_HIP(hipMemcpy(h.data(), colres2.get(), sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        double tot = 0.0;
        for (int k = 0; k < sys->n; ++k) {
            tot += h[(size_t)k];
            if (column_residuals) column_residuals[k] = std::sqrt(h[(size_t)k]);
        }
        *frobenius = std::sqrt(tot);
    });
}

int spai_system_product(spai_system *sys, long long *nnz, int *colptr, int *rowidx, double *values) {
    if (!sys || !nnz) return fail(SPAI_ERR_INVALID, "sys/nnz is NULL");
    if (sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        hipStream_t st = sys->ctx->stream;
        DeviceArray<double> colres2;
        DeviceArray<int> n1;
        run_product(sys, colres2, n1, nullptr, nullptr, nullptr);
        std::vector<int> hn1((size_t)sys->n);
        SPAI_HIP(hipMemcpy(hn1.data(), n1.get(), sizeof(int) * hn1.size(), hipMemcpyDeviceToHost));
        std::vector<int> cp((size_t)sys->n + 1, 0);
        long long total = 0;
        for (int k = 0; k < sys->n; ++k) {
            total += hn1[(size_t)k];
            require(total < (long long)INT32_MAX, "product: more than 2^31-1 entries");
            cp[(size_t)k + 1] = (int)total;
        }
        *nnz = total;
        if (!colptr && !rowidx && !values) return;
        DeviceArray<int> dcp(cp.size()), drow((size_t)std::max<long long>(total, 1));
        DeviceArray<double> dval((size_t)std::max<long long>(total, 1));
        SPAI_HIP(hipMemcpyAsync(dcp.get(), cp.data(), sizeof(int) * cp.size(), hipMemcpyHostToDevice, st));
        run_product(sys, colres2, n1, dcp.get(), drow.get(), dval.get());
        if (colptr) std::memcpy(colptr, cp.data(), sizeof(int) * cp.size());
        if (rowidx)
            SPAI_HIP(hipMemcpy(rowidx, drow.get(), sizeof(int) * (size_t)total, hipMemcpyDeviceToHost));
        if (values)
            SPAI_HIP(hipMemcpy(values, dval.get(), sizeof(double) * (size_t)total, hipMemcpyDeviceToHost));
    });
}

int spai_system_get_m(spai_system *sys, double *values, int values_on_device) {
    if (!sys || !values) return fail(SPAI_ERR_INVALID, "sys/values is NULL");
    if (sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        hipStream_t st = sys->ctx->stream;
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        size_t bytes = sizeof(double) * (size_t)sys->m_nnz;
        const double *src = sys->m64_csc.get();
        DeviceArray<double> tmp;
        if (sys->m_precision == SPAI_PRECISION_MIXED) {
            tmp.resize((size_t)std::max(sys->m_nnz, 1));
            convert_f32_to_f64(sys->m32_csc.get(), tmp.get(), sys->m_nnz, st);
            src = tmp.get();
        }
        SPAI_HIP(hipMemcpyAsync(values, src, bytes, values_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                st));
        SPAI_HIP(hipStreamSynchronize(st));
    });
}

int spai_system_get_row_set_sizes(spai_system *sys, int *n1) {
    if (!sys || !n1) return fail(SPAI_ERR_INVALID, "sys/n1 is NULL");
    if (sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        SPAI_HIP(hipMemcpyAsync(n1, sys->n1.get(), sizeof(int) * (size_t)sys->n, hipMemcpyDeviceToHost,
                                sys->ctx->stream));
        SPAI_HIP(hipStreamSynchronize(sys->ctx->stream));
    });
}

void spai_solve_options_default(spai_solve_options *opt) {
    if (!opt) return;
    opt->rel_tol = 1e-8;
    opt->abs_tol = 0.0;
    opt->max_iter = 1000;
    opt->use_graph = 0;
    opt->graph_iters = 8;
    opt->check_every = 1;
    opt->precondition = 1;
    opt->record_history = 0;
    opt->true_residual_restart = 0;
}

int spai_system_solve(spai_system *sys, const double *b, double *x, int on_device, const spai_solve_options *opt,
                      spai_solve_report *report) {
    if (!sys || !b || !x) return fail(SPAI_ERR_INVALID, "sys/b/x is NULL");
    spai_solve_options dflt;
    spai_solve_options_default(&dflt);
    if (!opt) opt = &dflt;
    if (opt->precondition && sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        hipStream_t st = sys->ctx->stream;
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        const double *db = b;
        double *dx = x;
        if (!on_device) {
            sys->stage_b.reserve((size_t)sys->n);
            sys->stage_x.reserve((size_t)sys->n);
            SPAI_HIP(hipMemcpyAsync(sys->stage_b.get(), b, sizeof(double) * sys->n, hipMemcpyHostToDevice, st));
            SPAI_HIP(hipMemcpyAsync(sys->stage_x.get(), x, sizeof(double) * sys->n, hipMemcpyHostToDevice, st));
            db = sys->stage_b.get();
            dx = sys->stage_x.get();
        }
        OperatorView A = a_view(sys);
        OperatorView M = opt->precondition ? m_view(sys, false) : OperatorView{};
        SolveResult res;
        sys->ctx->solver.solve(A, M, sys->tiles, sys->n, sys->nsys, db, dx, to_options(opt), st, &res);
        sys->last_states = sys->ctx->solver.states();
        sys->have_states = true;
        sys->last_solver = &sys->ctx->solver;
        if (!on_device) {
            SPAI_HIP(hipMemcpyAsync(x, dx, sizeof(double) * sys->n, hipMemcpyDeviceToHost, st));
            SPAI_HIP(hipStreamSynchronize(st));
        }
        fill_report(report, res);
    });
}

int spai_system_solve_noalloc(spai_system *sys, const double *b, double *x, const spai_solve_options *opt,
                              spai_solve_report *report) {
    if (!sys || !b || !x) return fail(SPAI_ERR_INVALID, "sys/b/x is NULL");
    spai_solve_options dflt;
    spai_solve_options_default(&dflt);
    if (!opt) opt = &dflt;
    if (opt->use_graph || opt->record_history)
        return fail(SPAI_ERR_INVALID, "the allocation-free solve runs without graphs and residual histories");
    if (opt->precondition && sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    spai_context *ctx = sys->ctx;
    if (!ctx->scratch) return fail(SPAI_ERR_HIP, "no device scratch: " + ctx->scratch_error);
    const size_t need = BiCGStab::workspace_bytes(sys->n, sys->nsys, sys->tiles);
    if (need > kNoAllocScratchBytes)
        return fail(SPAI_ERR_UNSUPPORTED, "the allocation-free solve needs " + std::to_string(need) +
                                              " bytes of scratch; the device scratch holds " +
                                              std::to_string(kNoAllocScratchBytes) + " (64 MiB)");
    return guarded([&] {
        SPAI_HIP(hipSetDevice(ctx->device));
        OperatorView A = a_view(sys);
        OperatorView M = opt->precondition ? m_view(sys, false) : OperatorView{};
        SolveResult res;
        ctx->noalloc_solver.solve(A, M, sys->tiles, sys->n, sys->nsys, b, x, to_options(opt), ctx->stream, &res,
                                  sys->last_states.data());
        sys->have_states = true;
        sys->last_solver = &ctx->noalloc_solver;
        fill_report(report, res);
    });
}

int spai_system_outcomes(spai_system *sys, spai_system_outcome *out) {
    if (!sys || !out) return fail(SPAI_ERR_INVALID, "sys/out is NULL");
    if (!sys->have_states || (int)sys->last_states.size() != sys->nsys)
        return fail(SPAI_ERR_INVALID, "no solve has run");
    for (int s = 0; s < sys->nsys; ++s) {
        const SysState &S = sys->last_states[(size_t)s];
        out[s].status = S.status;
        out[s].iterations = S.iters;
        out[s].residual = std::sqrt(S.rr);
        out[s].residual0 = S.err0;
        out[s].true_residual = std::sqrt(S.true_rr);
        out[s].tol = S.tol;
        out[s].restarts = S.restarts;
    }
    return SPAI_OK;
}

int spai_system_residual_history(spai_system *sys, int s, double *out, int maxlen) {
    if (!sys || !out || maxlen < 0) return fail(SPAI_ERR_INVALID, "sys/out is NULL");
    if (!sys->last_solver || s < 0 || s >= sys->nsys) return fail(SPAI_ERR_INVALID, "no solve or bad system index");
    return sys->last_solver->history(s, out, maxlen);
}

int spai_system_adjoint(spai_system *sys, const double *x, const double *g, double *dA, double *db, int on_device,
                        const spai_solve_options *opt, spai_solve_report *report) {
    if (!sys || !x || !g || !dA || !db) return fail(SPAI_ERR_INVALID, "sys/x/g/dA/db is NULL");
    spai_solve_options dflt;
    spai_solve_options_default(&dflt);
    if (!opt) opt = &dflt;
    if (opt->precondition && sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        hipStream_t st = sys->ctx->stream;
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        ensure_transposed_tiles(sys);
        const size_t n = (size_t)sys->n, nz = (size_t)std::max(sys->nnz, 1);
        const double *dx = x, *dg = g;
        double *ddA = dA, *ddb = db;
        if (!on_device) {
            sys->stage_x.reserve(n);
            sys->stage_g.reserve(n);
            sys->stage_dA.reserve(nz);
            sys->stage_db.reserve(n);
            SPAI_HIP(hipMemcpyAsync(sys->stage_x.get(), x, sizeof(double) * n, hipMemcpyHostToDevice, st));
            SPAI_HIP(hipMemcpyAsync(sys->stage_g.get(), g, sizeof(double) * n, hipMemcpyHostToDevice, st));
            dx = sys->stage_x.get();
            dg = sys->stage_g.get();
            ddA = sys->stage_dA.get();
            ddb = sys->stage_db.get();
        }
        fill(ddb, (long long)n, st);  // lambda_0 = 0
        OperatorView At = at_view(sys);
        OperatorView Mt = opt->precondition ? m_view(sys, true) : OperatorView{};
        SolveResult res;
        sys->ctx->adjoint_solver.solve(At, Mt, sys->tiles_t, sys->n, sys->nsys, dg, ddb, to_options(opt), st, &res);
        sys->last_states = sys->ctx->adjoint_solver.states();
        sys->have_states = true;
        sys->last_solver = &sys->ctx->adjoint_solver;
        adjoint_scatter(sys->n, sys->colptr.get(), sys->rowidx.get(), ddb, dx, ddA, st);
        if (!on_device) {
            SPAI_HIP(hipMemcpyAsync(dA, ddA, sizeof(double) * (size_t)sys->nnz, hipMemcpyDeviceToHost, st));
            SPAI_HIP(hipMemcpyAsync(db, ddb, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        }
        SPAI_HIP(hipStreamSynchronize(st));
        fill_report(report, res);
    });
}

int spai_system_time_spmv(spai_system *sys, int which, int reps, float *ms) {
    if (!sys || !ms || reps < 1) return fail(SPAI_ERR_INVALID, "bad arguments");
    if (which != 0 && sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        OperatorView A = a_view(sys);
        OperatorView M = which != 0 ? m_view(sys, false) : OperatorView{};
        *ms = sys->ctx->solver.time_kernel(which, A, M, sys->tiles, sys->n, sys->nsys, reps, sys->ctx->stream);
    });
}

int spai_system_time_iteration(spai_system *sys, int reps, int variant, float *ms) {
    if (!sys || !ms || reps < 1 || variant < -1 || variant > 1) return fail(SPAI_ERR_INVALID, "bad arguments");
    if (sys->m_precision < 0) return fail(SPAI_ERR_NOT_BUILT, "M has not been built");
    return guarded([&] {
        SPAI_HIP(hipSetDevice(sys->ctx->device));
        sys->ctx->solver.time_iteration(a_view(sys), m_view(sys, false), sys->tiles, sys->n, sys->nsys, reps,
                                        variant, sys->ctx->stream, ms);
    });
}

int spai_system_layout(const spai_system *sys, int *kind, long long *entries) {
    if (!sys) return fail(SPAI_ERR_INVALID, "sys is NULL");
    if (kind)
        *kind = sys->tiles.kind != kLayoutSell
                    ? (sys->tiles.cblocks > 0 && sys->pattern == SPAI_PATTERN_A ? SPAI_LAYOUT_ROW_TILES_BLOCKED
                                                                                  : SPAI_LAYOUT_ROW_TILES)
                : sys->tiles.sell_diag.get()   ? SPAI_LAYOUT_SELL64_DIAG
                : sys->tiles.sell_off.get()    ? SPAI_LAYOUT_SELL64_OFF16
                                               : SPAI_LAYOUT_SELL64;
    if (entries) *entries = sys->tiles.value_len(sys->nnz);
    return SPAI_OK;
}

int spai_calibration_copy(spai_context *ctx, long long bytes, int reps, float *ms) {
    if (!ctx || bytes < 16 || reps < 1 || !ms) return fail(SPAI_ERR_INVALID, "bad calibration arguments");
    return guarded([&] {
        SPAI_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        long long n2 = bytes / 16;
        DeviceArray<double> a((size_t)n2 * 2), b((size_t)n2 * 2);
        fill(a.get(), n2 * 2, st);
        int grid = ctx->cus * 16;
        hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, st, (const f64x2 *)a.get(),
                           (f64x2 *)b.get(), n2);
        Timer tm(st);
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, st, (const f64x2 *)a.get(),
                               (f64x2 *)b.get(), n2);
        SPAI_CHECK_LAUNCH();
        *ms = tm.stop() / reps;
    });
}

int spai_csc_to_csr_device(int m, int n, int nnz, const double *cscData, const int *cscIndex, const int *cscPtr,
                           double *csrData, int *csrIndex, int *csrPtr) {
    if (m <= 0 || n <= 0 || nnz < 0 || !cscIndex || !cscPtr || !csrIndex || !csrPtr)
        return fail(SPAI_ERR_INVALID, "csc_to_csr: bad arguments");
    return guarded([&] {
        DeviceArray<int> perm((size_t)std::max(nnz, 1));
        csc_to_csr_device(m, n, nnz, cscPtr, cscIndex, cscData, csrPtr, csrIndex, csrData, perm.get(), nullptr);
    });
}

// ---------------------------------------------------------- problems ----
int spai_problem_generate(int config, uint64_t seed, spai_problem **out) {
    if (!out) return fail(SPAI_ERR_INVALID, "out is NULL");
    *out = nullptr;
    return guarded([&] {
        auto P = std::make_unique<spai_problem>();
        P->p = generate_config(config, seed);
        *out = P.release();
    });
}

int spai_problem_generate_power_law(int n, uint64_t seed, spai_problem **out) {
    if (!out || n < 2) return fail(SPAI_ERR_INVALID, "bad power-law arguments");
    *out = nullptr;
    return guarded([&] {
        auto P = std::make_unique<spai_problem>();
        P->p = generate_power_law(seed, n);
        *out = P.release();
    });
}

int spai_problem_generate_batched(int nsys, uint64_t seed, spai_problem **out) {
    if (!out || nsys < 1) return fail(SPAI_ERR_INVALID, "bad batched arguments");
    *out = nullptr;
    return guarded([&] {
        auto P = std::make_unique<spai_problem>();
        P->p = generate_batched(seed, nsys);
        *out = P.release();
    });
}

int spai_problem_generate_grf(int m, uint64_t seed, spai_problem **out) {
    if (!out || m < 2) return fail(SPAI_ERR_INVALID, "bad grid size");
    *out = nullptr;
    return guarded([&] {
        auto P = std::make_unique<spai_problem>();
        P->p = generate_diffusion_grf_64(seed, m);
        *out = P.release();
    });
}

int spai_problem_convection_diffusion(int m, double peclet, double vx, double vy, spai_problem **out) {
    if (!out || m < 1 || !(peclet > 0)) return fail(SPAI_ERR_INVALID, "bad grid arguments");
    *out = nullptr;
    return guarded([&] {
        auto P = std::make_unique<spai_problem>();
        P->p.name = "convdiff" + std::to_string(m);
        P->p.systems.push_back(convection_diffusion_2d(m, peclet, vx, vy));
        P->p.rhs.emplace_back((size_t)m * m, 1.0);
        P->p.target.emplace_back();
        *out = P.release();
    });
}

int spai_problem_read_mtx(const char *path, spai_problem **out) {
    if (!out || !path) return fail(SPAI_ERR_INVALID, "path/out is NULL");
    *out = nullptr;
    return guarded([&] {
        auto P = std::make_unique<spai_problem>();
        P->p.name = path;
        P->p.systems.push_back(read_matrix_market(path));
        P->p.rhs.emplace_back((size_t)P->p.systems[0].n, 1.0);
        P->p.target.emplace_back();
        *out = P.release();
    });
}

int spai_problem_free(spai_problem *p) {
    delete p;
    return SPAI_OK;
}

int spai_problem_num_systems(const spai_problem *p) { return p ? (int)p->p.systems.size() : 0; }

int spai_problem_system(const spai_problem *p, int s, CSC_Matrix *out) {
    if (!p || !out || s < 0 || s >= (int)p->p.systems.size()) return fail(SPAI_ERR_INVALID, "bad system index");
    const HostCSC &A = p->p.systems[(size_t)s];
    out->n = out->nRow = out->nCol = A.n;
    out->nonzeroes = (int)A.nnz();
    out->mPtr = const_cast<int *>(A.colptr.data());
    out->mIndex = const_cast<int *>(A.rowidx.data());
    out->mData = const_cast<double *>(A.val.data());
    return SPAI_OK;
}

const double *spai_problem_rhs(const spai_problem *p, int s) {
    if (!p || s < 0 || s >= (int)p->p.rhs.size() || p->p.rhs[(size_t)s].empty()) return nullptr;
    return p->p.rhs[(size_t)s].data();
}

const double *spai_problem_target(const spai_problem *p, int s) {
    if (!p || s < 0 || s >= (int)p->p.target.size() || p->p.target[(size_t)s].empty()) return nullptr;
    return p->p.target[(size_t)s].data();
}

const char *spai_problem_name(const spai_problem *p) { return p ? p->p.name.c_str() : ""; }
const char *spai_problem_notes(const spai_problem *p) { return p ? p->p.notes.c_str() : ""; }

}  // extern "C"

// ===================================================== C++ drop-ins ====
struct MultiDCU_Context {
    spai_context *ctx = nullptr;
    int numDCUs = 1;
};

namespace {
[[noreturn]] void die(const char *where) {
    std::printf("HIP error in %s: %s\n", where, spai_last_error());
    std::exit(EXIT_FAILURE);
}

float run_spai_api(MultiDCU_Context *ctx, const CSC_Matrix *CSC_A, CSC_Matrix *devM, int cs, int ce, int schedule) {
    if (!ctx || !CSC_A || !devM) return 0.0f;
    if (cs < 0 || ce > CSC_A->nCol || cs >= ce) return 0.0f;  // the reference returns 0 for bad ranges
    spai_system *S = nullptr;
    if (spai_system_create(ctx->ctx, CSC_A, &S) != SPAI_OK) die("RunSPAI (upload)");
    // SPAI_MI355X_AUTOTUNE=1: pick the tile shapes on this matrix first (the
    // returned time is the build's alone)
    const char *tune = std::getenv("SPAI_MI355X_AUTOTUNE");
    if (tune && std::atoi(tune) == 1 &&
        spai_system_autotune(S, SPAI_PRECISION_FP64, schedule, 3, nullptr) != SPAI_OK)
        die("RunSPAI (autotune)");
    spai_build_report rep{};
    if (spai_system_build(S, SPAI_PRECISION_FP64, schedule, cs, ce, &rep) != SPAI_OK) die("RunSPAI (build)");
    int n = CSC_A->nCol, nnz = CSC_A->nonzeroes;
    devM->n = devM->nRow = devM->nCol = n;
    devM->nonzeroes = nnz;
    if (hipMalloc((void **)&devM->mPtr, sizeof(int) * ((size_t)n + 1)) != hipSuccess ||
        hipMalloc((void **)&devM->mIndex, sizeof(int) * (size_t)std::max(nnz, 1)) != hipSuccess ||
        hipMalloc((void **)&devM->mData, sizeof(double) * (size_t)std::max(nnz, 1)) != hipSuccess) {
        g_last_error = "hipMalloc for M failed";
        die("RunSPAI (allocate M)");
    }
    if (hipMemcpy(devM->mPtr, S->colptr.get(), sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice) !=
            hipSuccess ||
        hipMemcpy(devM->mIndex, S->rowidx.get(), sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice) != hipSuccess) {
        g_last_error = "copy of M's pattern failed";
        die("RunSPAI (M pattern)");
    }
    if (spai_system_get_m(S, devM->mData, 1) != SPAI_OK) die("RunSPAI (M values)");
    spai_system_destroy(S);
    return rep.ms;
}

spai_context *default_context() {
    static spai_context *ctx = nullptr;
    if (!ctx) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (spai_context_create(dev, &ctx) != SPAI_OK) die("default context");
    }
    return ctx;
}
}  // namespace

// The reference takes devices 0 .. numDCUs - 1 (src/spai_multi_dcu.cpp:1309-1342)
// and exits when there is none; with one device that is device 0, which is
// left current.
MultiDCU_Context *initMultiDCU(int requestedDCUs) {
    (void)requestedDCUs;  // one device
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        std::printf("initMultiDCU: no HIP device found\n");
        std::exit(EXIT_FAILURE);
    }
    auto *c = new MultiDCU_Context();
    if (spai_context_create(0, &c->ctx) != SPAI_OK) die("initMultiDCU");
    c->numDCUs = 1;
    return c;
}

void cleanupMultiDCU(MultiDCU_Context *ctx) {
    if (!ctx) return;
    spai_context_destroy(ctx->ctx);
    delete ctx;
}

float RunStaticSPAI_MultiDCU_MPI(MultiDCU_Context *ctx, const CSC_Matrix *CSC_A, CSC_Matrix *devCSC_M_global,
                                 int globalColStart, int globalColEnd) {
    return run_spai_api(ctx, CSC_A, devCSC_M_global, globalColStart, globalColEnd,
                        SPAI_SCHEDULE_BALANCED_COLUMNS);
}

float RunDynamicSPAI_MultiDCU_MPI(MultiDCU_Context *ctx, const CSC_Matrix *CSC_A, CSC_Matrix *devCSC_M_global,
                                  int globalColStart, int globalColEnd) {
    return run_spai_api(ctx, CSC_A, devCSC_M_global, globalColStart, globalColEnd, SPAI_SCHEDULE_BALANCED_NNZ);
}

int cublas2_pbicgstabv2(CSR_Matrix *devA, CSR_Matrix *devM, double *dev_b, double *dev_x, double TOL, int MAX_ITER) {
    if (!devA || !dev_b || !dev_x) return 0;
    spai_context *ctx = default_context();
    int n = devA->n;
    int status = guarded([&] {
        hipStream_t st = ctx->stream;
        TileMap ta;
        std::vector<int> roff = {0, n};
        build_layout(ta, devA->mPtr, devA->mIndex, roff, {0, (long long)devA->nonzeroes}, false, st);
        OperatorView A{devA->mPtr, devA->mIndex, devA->mData, 0};
        OperatorView M{};
        // M may have its own pattern: the kernels read each operator's own
        // row pointers inside the shared row ranges
        if (devM) M = OperatorView{devM->mPtr, devM->mIndex, devM->mData, 0};
        SolveOptions o;
        o.abs_tol = TOL;
        o.max_iter = MAX_ITER <= 0 ? n : MAX_ITER;
        SolveResult res;
        ctx->solver.solve(A, M, ta, n, 1, dev_b, dev_x, o, st, &res);
    });
    if (status != SPAI_OK) die("cublas2_pbicgstabv2");
    const SysState st0 = ctx->solver.states()[0];
    // The reference has no breakdown test: once a scalar is non-finite its
    // err < TOL never holds and it returns MAX_ITER
    // (include/bicgstab/bicgstab_solver.h:138, :228).  x holds non-finite
    // values in both cases.
    if (st0.status == kBreakdown) return MAX_ITER <= 0 ? n : MAX_ITER;
    return st0.iters;
}

void cuCSC2CSR(int m, int n, int nnz, double *cscData, int *cscIndex, int *cscPtr, double *csrData, int *csrIndex,
               int *csrPtr) {
    if (spai_csc_to_csr_device(m, n, nnz, cscData, cscIndex, cscPtr, csrData, csrIndex, csrPtr) != SPAI_OK)
        die("cuCSC2CSR");
}
