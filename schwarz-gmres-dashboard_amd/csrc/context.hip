// Context, device memory, error handling, kernel timing and BLAS-1 for
// libsgamd.so.
#include "sg_internal.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace sg {

static std::mutex g_err_mu;
static std::string g_last_err;

void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e == hipSuccess) return;
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %s (%d) in %s at %s:%d", hipGetErrorString(e),
                  (int)e, what, file, line);
    throw Error{SG_E_HIP, buf};
}

void fail(sg_status st, const std::string& msg) { throw Error{st, msg}; }

void set_global_error(const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_err = msg;
}

// Runs f, converting exceptions into a status and the context's message.
template <typename F>
sg_status guarded(sg_context* ctx, F&& f) {
    try {
        f();
        return SG_OK;
    } catch (const Error& e) {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_last_err = e.msg;
        if (ctx) ctx->err = e.msg;
        return e.st;
    } catch (const std::bad_alloc&) {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_last_err = "host out of memory";
        if (ctx) ctx->err = g_last_err;
        return SG_E_ARG;
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_last_err = e.what();
        if (ctx) ctx->err = g_last_err;
        std::string m = e.what();
        if (m.find("singular") != std::string::npos) return SG_E_SINGULAR;
        if (m.find("partition") != std::string::npos) return SG_E_PARTITION;
        return SG_E_ARG;
    }
}

ProfScope::ProfScope(sg_context* c, int f, double u) : ctx(c), fam(f), units(u) {
    if (!ctx->profiling) return;
    a = ctx->get_event();
    SG_HIP(hipEventRecord(a, ctx->stream));
}

ProfScope::~ProfScope() {
    if (!a) return;
    hipEvent_t b = ctx->get_event();
    (void)hipEventRecord(b, ctx->stream);
    ctx->prof[fam].pending.emplace_back(a, b);
    ctx->prof[fam].launches++;
    ctx->prof[fam].units += units;
}

// ---- deterministic reduction kernels ----------------------------------------

// Block partial sums of a*b (or a*a) over a grid-stride loop (k_dot_part),
// then one workgroup adds the partials in a fixed order (k_dot_fin): two
// launches instead of a last-block ticket, whose device-scope fences cost more
// than the reduction on this part.
template <bool SQUARE>
__global__ void __launch_bounds__(kRedThreads)
k_dot_part(int n, const double* __restrict__ a, const double* __restrict__ b,
           double* __restrict__ partials) {
    double s = 0.0;
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
        s += SQUARE ? a[i] * a[i] : a[i] * b[i];
    __shared__ double red[kRedThreads];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kRedThreads)
k_dot_fin(int nblk, const double* __restrict__ partials, double* __restrict__ out, int take_sqrt) {
    __shared__ double red[kRedThreads];
    double t = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kRedThreads) t += partials[i];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = take_sqrt ? sqrt(red[0]) : red[0];
}

__global__ void k_axpy(int n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] += alpha * x[i];
}

__global__ void k_scal(int n, double alpha, double* __restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] *= alpha;
}

static int red_grid(int n) {
    int g = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
    return g < 1 ? 1 : (g > kRedBlocks ? kRedBlocks : g);
}

void dot_launch(sg_context* ctx, int n, const double* a, const double* b, double* result) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(k_dot_part<false>, dim3(g), dim3(kRedThreads), 0, ctx->stream, n, a, b,
                       ctx->d_partials);
    hipLaunchKernelGGL(k_dot_fin, dim3(1), dim3(kRedThreads), 0, ctx->stream, g, ctx->d_partials,
                       result, 0);
    SG_HIP(hipGetLastError());
}

static void nrm_launch(sg_context* ctx, int n, const double* a, double* result) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(k_dot_part<true>, dim3(g), dim3(kRedThreads), 0, ctx->stream, n, a, a,
                       ctx->d_partials);
    hipLaunchKernelGGL(k_dot_fin, dim3(1), dim3(kRedThreads), 0, ctx->stream, g, ctx->d_partials,
                       result, 1);
    SG_HIP(hipGetLastError());
}

static int ew_grid(int n) {
    int g = (n + 255) / 256;
    return g < 1 ? 1 : (g > 8192 ? 8192 : g);
}

uint64_t fnv1a(const void* data, size_t bytes, uint64_t h) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

int tune_lookup(sg_context* ctx, const std::string& key) {
    for (const auto& kv : ctx->tune_kv)
        if (kv.first == key) return kv.second;
    return -1;
}

void tune_store(sg_context* ctx, const std::string& key, int value) {
    for (auto& kv : ctx->tune_kv)
        if (kv.first == key) {
            kv.second = value;
            return;
        }
    ctx->tune_kv.emplace_back(key, value);
    if (!ctx->tune_path.empty()) {
        std::ofstream out(ctx->tune_path, std::ios::app);
        out << key << " " << value << "\n";
    }
}

}  // namespace sg

using namespace sg;

hipEvent_t sg_context::get_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    SG_HIP(hipEventCreate(&e));
    return e;
}

void sg_context::collect_profiles() {
    for (auto& p : prof) {
        for (auto& ab : p.pending) {
            float ms = 0.f;
            SG_HIP(hipEventElapsedTime(&ms, ab.first, ab.second));
            p.ms += ms;
            event_pool.push_back(ab.first);
            event_pool.push_back(ab.second);
        }
        p.pending.clear();
    }
}

extern "C" {

sg_status sg_context_create(int device, sg_context** out) {
    if (!out) return SG_E_ARG;
    *out = nullptr;
    sg_context* ctx = new sg_context();
    sg_status st = guarded(nullptr, [&] {
        int ndev = 0;
        SG_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) fail(SG_E_ARG, "no such HIP device");
        ctx->device = device;
        SG_HIP(hipSetDevice(device));
        SG_HIP(hipStreamCreate(&ctx->stream));
        SG_HIP(hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, device));
        SG_HIP(hipDeviceGetAttribute(&ctx->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                     device));
        SG_HIP(hipMalloc(&ctx->d_partials, sizeof(double) * kRedBlocks * kMaxRedCols));
        SG_HIP(hipMalloc(&ctx->d_result, sizeof(double) * kMaxRedCols));
        SG_HIP(hipHostMalloc(&ctx->h_pinned, sizeof(double) * 256, hipHostMallocDefault));
        sg::noalloc_init_device(device);
    });
    if (st != SG_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return SG_OK;
}

sg_status sg_context_destroy(sg_context* ctx) {
    if (!ctx) return SG_E_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->prof)
        for (auto& ab : p.pending) {
            (void)hipEventDestroy(ab.first);
            (void)hipEventDestroy(ab.second);
        }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    sg::krylov_cache_free(ctx->krylov);
    if (ctx->d_partials) (void)hipFree(ctx->d_partials);
    if (ctx->d_result) (void)hipFree(ctx->d_result);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SG_OK;
}

sg_status sg_context_synchronize(sg_context* ctx) {
    if (!ctx) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipStreamSynchronize(ctx->stream));
        ctx->collect_profiles();
    });
}

const char* sg_last_error(const sg_context* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_last_err.c_str();
}

sg_status sg_malloc(sg_context* ctx, size_t bytes, void** d_ptr) {
    if (!ctx || !d_ptr) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipSetDevice(ctx->device));
        SG_HIP(hipMalloc(d_ptr, bytes ? bytes : 1));
    });
}

sg_status sg_free(sg_context* ctx, void* d_ptr) {
    if (!ctx) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipStreamSynchronize(ctx->stream));
        if (d_ptr) SG_HIP(hipFree(d_ptr));
    });
}

sg_status sg_memcpy_h2d(sg_context* ctx, void* d, const void* h, size_t bytes) {
    if (!ctx || (bytes && (!d || !h))) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
        SG_HIP(hipStreamSynchronize(ctx->stream));
    });
}

sg_status sg_memcpy_d2h(sg_context* ctx, void* h, const void* d, size_t bytes) {
    if (!ctx || (bytes && (!d || !h))) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        SG_HIP(hipStreamSynchronize(ctx->stream));
    });
}

sg_status sg_memcpy_d2d(sg_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    });
}

sg_status sg_memset_zero(sg_context* ctx, void* d, size_t bytes) {
    if (!ctx || (bytes && !d)) return SG_E_ARG;
    return guarded(ctx, [&] { SG_HIP(hipMemsetAsync(d, 0, bytes, ctx->stream)); });
}

sg_status sg_dot(sg_context* ctx, int n, const double* a, const double* b, double* result) {
    if (!ctx || n < 0 || !result || (n && (!a || !b))) return SG_E_ARG;
    return guarded(ctx, [&] {
        dot_launch(ctx, n, a, b, ctx->d_result);
        SG_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_result, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
        SG_HIP(hipStreamSynchronize(ctx->stream));
        *result = ctx->h_pinned[0];
    });
}

sg_status sg_nrm2(sg_context* ctx, int n, const double* a, double* result) {
    if (!ctx || n < 0 || !result || (n && !a)) return SG_E_ARG;
    return guarded(ctx, [&] {
        nrm_launch(ctx, n, a, ctx->d_result);
        SG_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_result, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
        SG_HIP(hipStreamSynchronize(ctx->stream));
        *result = ctx->h_pinned[0];
    });
}

sg_status sg_axpy(sg_context* ctx, int n, double alpha, const double* x, double* y) {
    if (!ctx || n < 0 || (n && (!x || !y))) return SG_E_ARG;
    return guarded(ctx, [&] {
        if (!n) return;
        hipLaunchKernelGGL(k_axpy, dim3(ew_grid(n)), dim3(256), 0, ctx->stream, n, alpha, x, y);
        SG_HIP(hipGetLastError());
    });
}

sg_status sg_scal(sg_context* ctx, int n, double alpha, double* x) {
    if (!ctx || n < 0 || (n && !x)) return SG_E_ARG;
    return guarded(ctx, [&] {
        if (!n) return;
        hipLaunchKernelGGL(k_scal, dim3(ew_grid(n)), dim3(256), 0, ctx->stream, n, alpha, x);
        SG_HIP(hipGetLastError());
    });
}

sg_status sg_context_set_profiling(sg_context* ctx, int enable) {
    if (!ctx) return SG_E_ARG;
    ctx->profiling = enable != 0;
    return SG_OK;
}

sg_status sg_context_set_autotune(sg_context* ctx, int enable) {
    if (!ctx || enable < 0 || enable > 11) return SG_E_ARG;
    ctx->autotune = enable;
    return SG_OK;
}

sg_status sg_context_set_tune_cache(sg_context* ctx, const char* path) {
    if (!ctx) return SG_E_ARG;
    return run_guarded(ctx, [&] {
        ctx->tune_path = path ? path : "";
        if (ctx->tune_path.empty()) return;
        std::ifstream in(ctx->tune_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string key;
            int v;
            if (ls >> key >> v) {
                bool found = false;
                for (auto& kv : ctx->tune_kv)
                    if (kv.first == key) kv.second = v, found = true;
                if (!found) ctx->tune_kv.emplace_back(key, v);
            }
        }
    });
}

sg_status sg_context_kernel_stats(sg_context* ctx, int family, long long* launches,
                                  double* total_ms) {
    if (!ctx || family < 0 || family >= SG_K_COUNT) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipStreamSynchronize(ctx->stream));
        ctx->collect_profiles();
        if (launches) *launches = ctx->prof[family].launches;
        if (total_ms) *total_ms = ctx->prof[family].ms;
    });
}

sg_status sg_context_reset_stats(sg_context* ctx) {
    if (!ctx) return SG_E_ARG;
    return guarded(ctx, [&] {
        SG_HIP(hipStreamSynchronize(ctx->stream));
        ctx->collect_profiles();
        for (auto& p : ctx->prof) {
            p.launches = 0;
            p.ms = 0.0;
            p.units = 0.0;
        }
    });
}

sg_status sg_context_kernel_units(sg_context* ctx, int family, double* units) {
    if (!ctx || !units || family < 0 || family >= SG_K_COUNT) return SG_E_ARG;
    *units = ctx->prof[family].units;
    return SG_OK;
}

}  // extern "C"

// Shared by the other translation units.
namespace sg {
sg_status run_guarded(sg_context* ctx, const std::function<void()>& f) { return guarded(ctx, f); }
}  // namespace sg
