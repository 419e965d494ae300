// Internal types of libsgamd.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sg_capi.h"
#include "host_setup.h"

namespace sg {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kRedBlocks = 1024;   // fixed grid of the deterministic reductions
constexpr int kRedThreads = 256;
constexpr int kMaxRedCols = 128;   // widest multi-dot (restart + 1)

struct Error {
    sg_status st;
    std::string msg;
};

// Throws Error when a HIP call fails; the C ABI boundary converts it.
void hip_check(hipError_t e, const char* what, const char* file, int line);
#define SG_HIP(call) ::sg::hip_check((call), #call, __FILE__, __LINE__)
[[noreturn]] void fail(sg_status st, const std::string& msg);
// Runs f; exceptions become a status plus the context's last-error message.
sg_status run_guarded(sg_context* ctx, const std::function<void()>& f);

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) SG_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        if (n) SG_HIP(hipMemcpyAsync(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Step guard handed from the solver to the kernels of one Arnoldi step.  A
// kernel (or the part of it serving one system of a batch) returns at once
// when its system's flag is 0.  flag == nullptr: always run.  The flags of
// consecutive systems are `stride` ints apart.  When `batched`, each consumer
// looks the system up through its own unit -> system map (SpMV row block,
// subdomain); a consumer without such a map always runs.
struct Guard {
    const int* flag = nullptr;
    int stride = 0;
    bool batched = false;
};

struct DevGuard {
    const int* flag;
    const int* map;  // unit -> system; entries < 0 always run
    int stride;
};

__device__ inline bool guard_stop(const DevGuard& g, int unit) {
    if (!g.flag) return false;
    if (!g.map) return g.flag[0] == 0;
    const int s = g.map[unit];
    return s >= 0 && g.flag[s * g.stride] == 0;
}

inline DevGuard dev_guard(const Guard& g, const int* map) {
    if (g.batched) return DevGuard{map ? g.flag : nullptr, map, g.stride};
    return DevGuard{g.flag, nullptr, 0};
}

struct Prof {
    long long launches = 0;
    double ms = 0.0;
    double units = 0.0;  // work units of the timed launches (see ProfScope)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

// GMRES/FGMRES workspace kept by a context between solves (krylov.hip).
struct KrylovCache;
void krylov_cache_free(KrylovCache* c);
// Looks up this device's copy of the no-allocation solve's static scratch
// array (krylov.hip); called once per context with the device current.
void noalloc_init_device(int device);

}  // namespace sg

struct sg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cu = 256;
    int max_lds = 65536;  // bytes of LDS one workgroup may allocate
    // deterministic reduction scratch: partials[kMaxRedCols][kRedBlocks], result
    double* d_partials = nullptr;
    double* d_result = nullptr;  // kMaxRedCols doubles
    double* h_pinned = nullptr;  // pinned staging for small reads
    bool profiling = false;
    int autotune = 1;  // 0 default shapes, 1 time them on first use, 2+v force shape v
    // tuning cache: key -> chosen shape, mirrored in a text file when a path is set
    std::string tune_path;
    std::vector<std::pair<std::string, int>> tune_kv;
    sg::Prof prof[SG_K_COUNT];
    std::vector<hipEvent_t> event_pool;
    sg::KrylovCache* krylov = nullptr;  // reused by every sg_solve on this context
    bool no_tune = false;  // set during sg_solve_noalloc: kernel shapes are not timed

    hipEvent_t get_event();
    void collect_profiles();  // stream must be synchronised
};

struct sg_csr {
    sg_context* ctx = nullptr;
    int nrows = 0, ncols = 0, nnz = 0;
    sg::HostCsr host;
    sg::DevBuf<int> rp, ci;
    sg::DevBuf<double> v;
    // row blocks for the LDS-staged SpMV, one table per block shape:
    // block b covers rows [blk[b], blk[b+1])
    int nblk_v[3] = {0, 0, 0};
    sg::DevBuf<int> blk_v[3];
    sg::DevBuf<int> blk_sys_v[3];  // row block -> system (blocks never straddle systems)
    // independent systems of a block-diagonal batch: rows [sys_off[i], sys_off[i+1])
    std::vector<int> sys_off;
    // SELL-64 copy (built when its padding is small): slice s holds rows
    // [sell_row[s], sell_row[s+1]) (at most 64, never straddling a system), its
    // entry k of lane l at sell_ptr[s] + 64 k + l; padding entries have column -1
    int nsl = 0;
    sg::DevBuf<int> sell_ptr, sell_row, sell_sys, sell_ci;
    sg::DevBuf<double> sell_v;
    // the same columns as 16-bit offsets from the row (-32768: padding), when
    // every offset fits
    sg::DevBuf<short> sell_cd;
    int variant = 0;     // selected shape: 0..2 row-block CSR, 3 SELL-64, 4 SELL-64 16-bit
    bool tuned = false;
};

namespace sg {

// RAII timing of one launch family on the context stream.
// `units` is added to the family's work-unit count when profiling is on (the
// Gram passes record their basis columns j + 1; other families 1 per launch).
struct ProfScope {
    sg_context* ctx;
    int fam;
    double units;
    hipEvent_t a = nullptr;
    ProfScope(sg_context* c, int f, double u = 1.0);
    ~ProfScope();
};

// ---- tuning cache (context.hip) ---------------------------------------------------
// 64-bit FNV-1a over raw bytes, chained through h.
uint64_t fnv1a(const void* data, size_t bytes, uint64_t h = 1469598103934665603ull);
int tune_lookup(sg_context* ctx, const std::string& key);  // -1 when absent
void tune_store(sg_context* ctx, const std::string& key, int value);

// ---- SpMV (spmv.hip) ---------------------------------------------------------
// y = alpha*A*x + beta*y, skipping the rows of stopped systems (see Guard).
// y0 may be null (then y0 = y); beta == 0 never reads y0.
void spmv_launch(sg_csr* A, double alpha, const double* x, double beta, const double* y0,
                 double* y, const Guard& guard, hipStream_t s);
void spmv_prepare(sg_csr* A);
// Runs the block-shape autotuner now if it is due (never inside graph capture).
void spmv_ensure_tuned(sg_csr* A);

// ---- reductions / BLAS-1 (blas.hip) -------------------------------------------
// result[0] = dot(a, b) on the device (deterministic tree), stream-ordered.
void dot_launch(sg_context* ctx, int n, const double* a, const double* b, double* result);

// ---- BiCGSTAB (bicgstab.hip) ---------------------------------------------------
void bicgstab_impl(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                   const sg_solve_params* prm, sg_solve_result* res, double* hist, int hist_cap,
                   bool use_graph);

// ---- learned stage (unet.hip, learned.hip) ---------------------------------------
// stop: the forward's kernels return at once while *stop == 0 (a solver's step
// flag); null: always run
void unet_forward_dev(sg_unet* U, int B, int H, int W, const float* d_in, bool grads,
                      const int* stop);
const float* unet_output(const sg_unet* U);
// the bf16 NHWC input [B][H][W][16] that unet_forward_dev(.., d_in = nullptr, ..) reads
uint16_t* unet_input_buffer(sg_unet* U, int B, int H, int W);
void unet_prepare(sg_unet* U, int B, int H, int W);  // pack weights, size buffers
struct LearnedStage;
LearnedStage* learned_new(sg_context* ctx, sg_precond* base, sg_unet* net, int nx, int ny,
                          const double* log_kappa, const sg_learned_opts* opts);
void learned_free(LearnedStage* L);
void learned_setup(LearnedStage* L, const sg_csr* A);
void learned_apply(LearnedStage* L, const double* x, double* y, const Guard& guard);
bool learned_apply_az(LearnedStage* L, const sg_csr* A, const double* x, double* y, double* az,
                      const Guard& guard);
void learned_prepare(LearnedStage* L);
void learned_counts(const LearnedStage* L, long long* calls, long long* accepted);
sg_precond* learned_base(const LearnedStage* L);

// ---- preconditioner apply (precond.hip) ----------------------------------------
void precond_apply_launch(sg_precond* M, const double* x, double* y, const Guard& guard);
// y = M x; a preconditioner that forms A y on the way (the learned stage, with
// A its own matrix) also writes it to az and returns true; otherwise false
// and az is untouched
bool precond_apply_az(sg_precond* M, const sg_csr* A, const double* x, double* y, double* az,
                      const Guard& guard);
// Maps each subdomain to the system of a block-diagonal batch that holds it
// (-1 if it spans several) for per-system guards.
void precond_bind_systems(sg_precond* M, const std::vector<int>& sys_off);
void precond_ensure_tuned(sg_precond* M);
bool precond_is_callback(const sg_precond* M);
bool precond_is_flexible(const sg_precond* M);  // changes per application: FGMRES only
bool precond_ready(const sg_precond* M);
int precond_n(const sg_precond* M);

}  // namespace sg
