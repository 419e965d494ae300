// FP64 CSR SpMV for CDNA4 and the sg_csr object.
//
// LDS-staged "CSR stream": a workgroup of NT threads (NT = 256, 512 or 1024)
// owns a block of up to NT consecutive rows whose nonzeros (at most 8*NT) are
// streamed with fully coalesced loads (eight values and column indices per
// thread issued together, then the eight x gathers, and the row bounds loaded
// before the pass); each product vals[k]*x[col[k]] is rounded and parked in
// LDS, then one lane per row adds its products sequentially in CSR order.  The sum therefore equals the reference's serial row loop
// (src/core/sparse_matrix.h:30-37) bit for bit, whatever NT is.  Blocks are
// remapped so that each of the 8 XCDs walks one contiguous eighth of the rows,
// which keeps the x neighbourhood of a stencil (+-nx*ny rows) inside that XCD's L2.
//
// SELL-64 (sliced ELLPACK, one wavefront per slice of 64 rows): a copy of the
// matrix with the entries of a slice stored column by column, so lane l owns
// row l and loads its k-th entry next to lane l+1's.  Value and column loads
// are fully coalesced, the x gathers of a stencil are nearly contiguous across
// the wave (consecutive rows read consecutive neighbours), and each lane adds
// its row's products in CSR order in a register: no LDS staging, no barrier,
// and the same sum as the serial row loop bit for bit.  Padding entries (column
// -1) are skipped, not added as zeros, so signed zeros and non-finite x values
// come out as in the serial loop.
#include "sg_internal.h"

#include <algorithm>
#include <cstdio>

namespace sg {

constexpr int kXcd = 8;
constexpr int kNnzPerThread = 8;  // LDS staging capacity: 8 products per thread

// Block shapes (threads = rows per block); results are identical for all of
// them, so the choice is a pure performance knob (see spmv_autotune).
constexpr int kSpmvVariants[3] = {256, 512, 1024};
constexpr int kSellVariant = 3;   // shape index of the SELL-64 kernel
constexpr int kSell16Variant = 4; // SELL-64 with 16-bit column offsets
constexpr int kNumShapes = 5;
constexpr short kSellPad16 = -32768;
constexpr int kSellC = 64;        // rows per slice: one wavefront
constexpr int kSellThreads = 256; // four slices per workgroup
// The SELL copy is built only when its padding stays below half the nonzeros
// and no row is longer than 2048 entries (a wavefront would spend most of a
// slice on one row's padding).
constexpr double kSellMaxFill = 1.5;
constexpr int kSellMaxRow = 256 * kNnzPerThread;

template <int NT>
__global__ void __launch_bounds__(NT)
k_spmv_stream(const int* __restrict__ rp, const int* __restrict__ ci,
              const double* __restrict__ v, const int* __restrict__ blk, int nblk, int per_xcd,
              double alpha, const double* __restrict__ x, double beta,
              const double* __restrict__ y0, double* __restrict__ y, DevGuard guard) {
    constexpr int kCap = NT * kNnzPerThread;
    const int b = (blockIdx.x % kXcd) * per_xcd + blockIdx.x / kXcd;
    if (b >= nblk || guard_stop(guard, b)) return;
    const int r0 = blk[b], r1 = blk[b + 1];
    const int k0 = rp[r0], k1 = rp[r1];
    const int t = threadIdx.x;
    __shared__ double prod[kCap];
    const int rr = min(r0 + t, r1 - 1);
    const int ra = rp[rr], re = rp[rr + 1];  // row bounds, loaded before the staging pass
    if (k1 - k0 <= kCap) {
        for (int kb = k0 + t; kb < k1; kb += 8 * NT) {
            double a[8];
            int c[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = min(kb + u * NT, k1 - 1);
                a[u] = __builtin_nontemporal_load(&v[k]);
                c[u] = __builtin_nontemporal_load(&ci[k]);
            }
            double xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (kb + u * NT < k1) prod[kb + u * NT - k0] = a[u] * xv[u];
        }
        __syncthreads();
        const int r = r0 + t;
        if (r < r1) {
            const int a = ra - k0, e = re - k0;
            double s = 0.0;
            for (int k = a; k < e; ++k) s += prod[k];
            y[r] = beta == 0.0 ? alpha * s : alpha * s + beta * y0[r];
        }
    } else {
        // one row longer than the staging capacity (row_blocks gives it a block
        // of its own): its products are staged kCap at a time and lane 0 adds
        // each chunk in CSR order, so the sum is the serial row loop's whatever
        // NT is
        double s = 0.0;
        for (int c0 = k0; c0 < k1; c0 += kCap) {
            const int c1 = min(k1, c0 + kCap);
            for (int k = c0 + t; k < c1; k += NT) prod[k - c0] = v[k] * x[ci[k]];
            __syncthreads();
            if (t == 0)
                for (int k = 0; k < c1 - c0; ++k) s += prod[k];
            __syncthreads();
        }
        if (t == 0) y[r0] = beta == 0.0 ? alpha * s : alpha * s + beta * y0[r0];
    }
}

template <int NT>
__global__ void __launch_bounds__(NT)
k_spmv_sell(const int* __restrict__ sp, const int* __restrict__ srow, const int* __restrict__ ci,
            const double* __restrict__ v, int nsl, int per_xcd, double alpha,
            const double* __restrict__ x, double beta, const double* __restrict__ y0,
            double* __restrict__ y, DevGuard guard) {
    constexpr int kSpb = NT / kSellC;
    const int b = (blockIdx.x % kXcd) * per_xcd + blockIdx.x / kXcd;
    const int sl = b * kSpb + (int)(threadIdx.x / kSellC), lane = threadIdx.x % kSellC;
    if (sl >= nsl || guard_stop(guard, sl)) return;
    const int p0 = sp[sl], width = (sp[sl + 1] - p0) / kSellC;
    const int r = srow[sl] + lane;
    double sum = 0.0;
    for (int k0 = 0; k0 < width; k0 += 8) {
        double a[8];
        int c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = p0 + min(k0 + u, width - 1) * kSellC + lane;
            a[u] = __builtin_nontemporal_load(&v[idx]);
            c[u] = __builtin_nontemporal_load(&ci[idx]);
        }
        double xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[max(c[u], 0)];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < width && c[u] >= 0) sum += a[u] * xv[u];
    }
    if (r < srow[sl + 1]) y[r] = beta == 0.0 ? alpha * sum : alpha * sum + beta * y0[r];
}

// SELL-64 with the columns stored as 16-bit offsets from the row: 10 bytes per
// entry instead of 12.  Same loads, products and sums as k_spmv_sell.
template <int NT>
__global__ void __launch_bounds__(NT)
k_spmv_sell16(const int* __restrict__ sp, const int* __restrict__ srow,
              const short* __restrict__ cd, const double* __restrict__ v, int nsl, int per_xcd,
              double alpha, const double* __restrict__ x, double beta,
              const double* __restrict__ y0, double* __restrict__ y, DevGuard guard) {
    constexpr int kSpb = NT / kSellC;
    const int b = (blockIdx.x % kXcd) * per_xcd + blockIdx.x / kXcd;
    const int sl = b * kSpb + (int)(threadIdx.x / kSellC), lane = threadIdx.x % kSellC;
    if (sl >= nsl || guard_stop(guard, sl)) return;
    const int p0 = sp[sl], width = (sp[sl + 1] - p0) / kSellC;
    const int r = srow[sl] + lane;
    double sum = 0.0;
    for (int k0 = 0; k0 < width; k0 += 8) {
        double a[8];
        short d[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = p0 + min(k0 + u, width - 1) * kSellC + lane;
            a[u] = __builtin_nontemporal_load(&v[idx]);
            d[u] = __builtin_nontemporal_load(&cd[idx]);
        }
        double xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[d[u] == kSellPad16 ? 0 : r + d[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < width && d[u] != kSellPad16) sum += a[u] * xv[u];
    }
    if (r < srow[sl + 1]) y[r] = beta == 0.0 ? alpha * sum : alpha * sum + beta * y0[r];
}

// SELL-64 copy of A, slices cut per system; false (nothing built) when the
// padding or a row length is beyond the limits above.
static bool sell_build(sg_csr* A) {
    const std::vector<int>& rp = A->host.rp;
    if (A->nnz == 0 || A->ncols == 0) return false;
    std::vector<int> sp(1, 0), srow, ssys;
    long long total = 0;
    for (size_t q = 0; q + 1 < A->sys_off.size(); ++q) {
        for (int r0 = A->sys_off[q]; r0 < A->sys_off[q + 1]; r0 += kSellC) {
            const int r1 = std::min(A->sys_off[q + 1], r0 + kSellC);
            int w = 0;
            for (int r = r0; r < r1; ++r) w = std::max(w, rp[r + 1] - rp[r]);
            if (w > kSellMaxRow) return false;
            total += (long long)kSellC * w;
            if (total > kSellMaxFill * A->nnz || total > INT32_MAX) return false;
            sp.push_back((int)total);
            srow.push_back(r0);
            ssys.push_back((int)q);
        }
    }
    srow.push_back(A->nrows);
    const int nsl = (int)ssys.size();
    std::vector<int> sci((size_t)total, -1);
    std::vector<double> sv((size_t)total, 0.0);
    for (int sl = 0; sl < nsl; ++sl)
        for (int r = srow[sl]; r < srow[sl + 1]; ++r)
            for (int k = rp[r]; k < rp[r + 1]; ++k) {
                const size_t idx = (size_t)sp[sl] + (size_t)(k - rp[r]) * kSellC + (r - srow[sl]);
                sci[idx] = A->host.ci[k];
                sv[idx] = A->host.v[k];
            }
    // 16-bit offsets when every entry's column is within 32767 of its row
    std::vector<short> scd((size_t)total, kSellPad16);
    bool fits16 = true;
    for (int sl = 0; sl < nsl && fits16; ++sl)
        for (int l = 0; l < kSellC; ++l) {
            const int r = srow[sl] + l;
            for (size_t idx = (size_t)sp[sl] + l; idx < (size_t)sp[sl + 1]; idx += kSellC) {
                if (sci[idx] < 0) continue;
                const int d = sci[idx] - r;
                if (d < -32767 || d > 32767) {
                    fits16 = false;
                    break;
                }
                scd[idx] = (short)d;
            }
        }
    hipStream_t s = A->ctx->stream;
    A->nsl = nsl;
    A->sell_ptr.upload(sp, s);
    A->sell_row.upload(srow, s);
    A->sell_sys.upload(ssys, s);
    A->sell_ci.upload(sci, s);
    A->sell_v.upload(sv, s);
    if (fits16) A->sell_cd.upload(scd, s);
    SG_HIP(hipStreamSynchronize(s));  // the host vectors end here
    return true;
}

// Row blocks of at most nt rows and nt*kNnzPerThread nonzeros (a longer row
// gets a block of its own); no block crosses a system boundary of a batch.
static void row_blocks(const std::vector<int>& rp, const std::vector<int>& sys_off, int nt,
                       std::vector<int>& blk, std::vector<int>& blk_sys) {
    const int cap = nt * kNnzPerThread;
    blk.assign(1, 0);
    blk_sys.clear();
    for (size_t q = 0; q + 1 < sys_off.size(); ++q) {
        int r = sys_off[q];
        const int rend = sys_off[q + 1];
        while (r < rend) {
            int start = r;
            while (r < rend && r - start < nt && rp[r + 1] - rp[start] <= cap) ++r;
            if (r == start) ++r;
            blk.push_back(r);
            blk_sys.push_back((int)q);
        }
    }
}

void spmv_prepare(sg_csr* A) {
    if (A->sys_off.empty()) A->sys_off = {0, A->nrows};
    std::vector<int> blk, bsys;
    for (int v = 0; v < 3; ++v) {
        row_blocks(A->host.rp, A->sys_off, kSpmvVariants[v], blk, bsys);
        A->nblk_v[v] = (int)blk.size() - 1;
        A->blk_v[v].upload(blk, A->ctx->stream);
        A->blk_sys_v[v].upload(bsys, A->ctx->stream);
        SG_HIP(hipStreamSynchronize(A->ctx->stream));
    }
    sell_build(A);
    A->variant = 0;
}

static bool shape_available(const sg_csr* A, int var) {
    if (var < 0 || var >= kNumShapes) return false;
    if (var == kSellVariant) return A->nsl > 0;
    if (var == kSell16Variant) return A->nsl > 0 && A->sell_cd.p;
    return true;
}

static void spmv_launch_variant(sg_csr* A, int var, double alpha, const double* x, double beta,
                                const double* y0, double* y, const Guard& g, hipStream_t s) {
    if (var == kSellVariant || var == kSell16Variant) {
        const int nb = (A->nsl + kSellThreads / kSellC - 1) / (kSellThreads / kSellC);
        const int per = (nb + kXcd - 1) / kXcd;
        if (var == kSellVariant)
            hipLaunchKernelGGL(k_spmv_sell<kSellThreads>, dim3(per * kXcd), dim3(kSellThreads), 0,
                               s, A->sell_ptr.p, A->sell_row.p, A->sell_ci.p, A->sell_v.p, A->nsl,
                               per, alpha, x, beta, y0, y, dev_guard(g, A->sell_sys.p));
        else
            hipLaunchKernelGGL(k_spmv_sell16<kSellThreads>, dim3(per * kXcd), dim3(kSellThreads),
                               0, s, A->sell_ptr.p, A->sell_row.p, A->sell_cd.p, A->sell_v.p,
                               A->nsl, per, alpha, x, beta, y0, y, dev_guard(g, A->sell_sys.p));
        SG_HIP(hipGetLastError());
        return;
    }
    const int nblk = A->nblk_v[var];
    const DevGuard guard = dev_guard(g, A->blk_sys_v[var].p);
    const int per = (nblk + kXcd - 1) / kXcd;
    const int* blk = A->blk_v[var].p;
    switch (kSpmvVariants[var]) {
        case 256:
            hipLaunchKernelGGL(k_spmv_stream<256>, dim3(per * kXcd), dim3(256), 0, s, A->rp.p,
                               A->ci.p, A->v.p, blk, nblk, per, alpha, x, beta, y0, y, guard);
            break;
        case 512:
            hipLaunchKernelGGL(k_spmv_stream<512>, dim3(per * kXcd), dim3(512), 0, s, A->rp.p,
                               A->ci.p, A->v.p, blk, nblk, per, alpha, x, beta, y0, y, guard);
            break;
        default:
            hipLaunchKernelGGL(k_spmv_stream<1024>, dim3(per * kXcd), dim3(1024), 0, s, A->rp.p,
                               A->ci.p, A->v.p, blk, nblk, per, alpha, x, beta, y0, y, guard);
            break;
    }
    SG_HIP(hipGetLastError());
}

// Times every shape on this matrix (HIP events, three launches each) and
// keeps the fastest; ties keep the earlier shape.
void spmv_autotune(sg_csr* A) {
    sg_context* ctx = A->ctx;
    hipStream_t s = ctx->stream;
    uint64_t h = fnv1a(A->host.rp.data(), sizeof(int) * A->host.rp.size());
    h = fnv1a(A->host.ci.data(), sizeof(int) * A->host.ci.size(), h);
    h = fnv1a(A->sys_off.data(), sizeof(int) * A->sys_off.size(), h);
    char key[96];
    std::snprintf(key, sizeof key, "spmv:%d:%d:%016llx", A->nrows, A->nnz, (unsigned long long)h);
    const int cached = tune_lookup(ctx, key);
    if (shape_available(A, cached)) {
        A->variant = cached;
        A->tuned = true;
        return;
    }
    DevBuf<double> x, y;
    x.alloc(std::max(A->ncols, 1));
    y.alloc(std::max(A->nrows, 1));
    SG_HIP(hipMemsetAsync(x.p, 0, sizeof(double) * x.n, s));
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    float best = 0.f;
    int bestv = 0;
    for (int v = 0; v < kNumShapes; ++v) {
        if (!shape_available(A, v)) continue;
        spmv_launch_variant(A, v, 1.0, x.p, 0.0, nullptr, y.p, Guard{}, s);  // warm
        SG_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) spmv_launch_variant(A, v, 1.0, x.p, 0.0, nullptr, y.p, Guard{}, s);
        SG_HIP(hipEventRecord(e1, s));
        SG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SG_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (v == 0 || ms < best * 0.98f) {
            best = ms;
            bestv = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    A->variant = bestv;
    A->tuned = true;
    tune_store(ctx, key, bestv);
}

void spmv_ensure_tuned(sg_csr* A) {
    if (A->nrows == 0) return;
    if (A->ctx->autotune == 1 && !A->tuned && !A->ctx->no_tune) spmv_autotune(A);
    if (A->ctx->autotune >= 2) {
        const int v = (A->ctx->autotune - 2) % kNumShapes;
        A->variant = shape_available(A, v) ? v : 0;
    } else if (A->ctx->autotune == 0) {
        A->variant = 0;
    }
}

void spmv_launch(sg_csr* A, double alpha, const double* x, double beta, const double* y0,
                 double* y, const Guard& guard, hipStream_t s) {
    if (A->nrows == 0) return;
    spmv_ensure_tuned(A);
    ProfScope ps(A->ctx, SG_K_SPMV);
    spmv_launch_variant(A, A->variant, alpha, x, beta, y0 ? y0 : y, y, guard, s);
}

}  // namespace sg

using namespace sg;

extern "C" {

static sg_status csr_create(sg_context* ctx, int nrows, int ncols, int nnz, const int* row_ptr,
                            const int* col_idx, const double* vals, const std::vector<int>& sys_off,
                            sg_csr** out) {
    if (!ctx || !out || nrows < 0 || ncols < 0 || nnz < 0 || !row_ptr ||
        (nnz && (!col_idx || !vals)))
        return SG_E_ARG;
    *out = nullptr;
    sg_csr* A = new sg_csr();
    sg_status st = run_guarded(ctx, [&] {
        if (row_ptr[0] != 0 || row_ptr[nrows] != nnz) fail(SG_E_ARG, "row_ptr inconsistent with nnz");
        for (int i = 0; i < nrows; ++i)
            if (row_ptr[i + 1] < row_ptr[i]) fail(SG_E_ARG, "row_ptr not monotone");
        for (int k = 0; k < nnz; ++k)
            if (col_idx[k] < 0 || col_idx[k] >= ncols) fail(SG_E_ARG, "column index out of range");
        if (!sys_off.empty()) {
            // offsets are validated before any of them indexes row_ptr
            if (sys_off.front() != 0 || sys_off.back() != nrows || nrows != ncols)
                fail(SG_E_ARG, "batch: system offsets must cover all rows of a square matrix");
            for (size_t q = 0; q + 1 < sys_off.size(); ++q)
                if (sys_off[q] < 0 || sys_off[q] >= sys_off[q + 1] || sys_off[q + 1] > nrows)
                    fail(SG_E_ARG, "batch: every system needs at least one row");
            for (size_t q = 0; q + 1 < sys_off.size(); ++q) {
                const int a = sys_off[q], b = sys_off[q + 1];
                for (int k = row_ptr[a]; k < row_ptr[b]; ++k)
                    if (col_idx[k] < a || col_idx[k] >= b)
                        fail(SG_E_ARG, "batch: matrix is not block diagonal over the systems");
            }
        }
        SG_HIP(hipSetDevice(ctx->device));
        A->ctx = ctx;
        A->nrows = nrows;
        A->ncols = ncols;
        A->nnz = nnz;
        A->host.n = nrows;
        A->host.ncols = ncols;
        A->host.rp.assign(row_ptr, row_ptr + nrows + 1);
        A->host.ci.assign(col_idx, col_idx + nnz);
        A->host.v.assign(vals, vals + nnz);
        A->rp.upload(A->host.rp, ctx->stream);
        A->ci.upload(A->host.ci, ctx->stream);
        A->v.upload(A->host.v, ctx->stream);
        A->sys_off = sys_off;
        spmv_prepare(A);
        SG_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (st != SG_OK) {
        delete A;
        return st;
    }
    *out = A;
    return SG_OK;
}

sg_status sg_csr_create(sg_context* ctx, int nrows, int ncols, int nnz, const int* row_ptr,
                        const int* col_idx, const double* vals, sg_csr** out) {
    return csr_create(ctx, nrows, ncols, nnz, row_ptr, col_idx, vals, {}, out);
}

sg_status sg_csr_create_batch(sg_context* ctx, int nsys, const int* sys_rows, int nnz,
                              const int* row_ptr, const int* col_idx, const double* vals,
                              sg_csr** out) {
    if (nsys < 1 || !sys_rows) return SG_E_ARG;
    std::vector<int> off(sys_rows, sys_rows + nsys + 1);
    const int n = off.back();
    return csr_create(ctx, n, n, nnz, row_ptr, col_idx, vals, off, out);
}

sg_status sg_csr_systems(const sg_csr* A, int* nsys) {
    if (!A || !nsys) return SG_E_ARG;
    *nsys = (int)A->sys_off.size() - 1;
    return SG_OK;
}

sg_status sg_csr_destroy(sg_csr* A) {
    if (!A) return SG_E_ARG;
    (void)hipStreamSynchronize(A->ctx->stream);
    delete A;
    return SG_OK;
}

sg_status sg_csr_info(const sg_csr* A, int* nrows, int* ncols, int* nnz) {
    if (!A) return SG_E_ARG;
    if (nrows) *nrows = A->nrows;
    if (ncols) *ncols = A->ncols;
    if (nnz) *nnz = A->nnz;
    return SG_OK;
}

sg_status sg_spmv(sg_csr* A, double alpha, const double* d_x, double beta, double* d_y) {
    if (!A || (A->nrows && (!d_x || !d_y))) return SG_E_ARG;
    return run_guarded(A->ctx, [&] {
        spmv_launch(A, alpha, d_x, beta, nullptr, d_y, Guard{}, A->ctx->stream);
    });
}

}  // extern "C"
