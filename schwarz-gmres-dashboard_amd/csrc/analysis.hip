// Convergence analytics on the device: the Lanczos condition estimate of
// M^{-1}A (src/analysis/condition_estimator.h:94-191), the power-iteration
// spectral radius of I - M^{-1}A (src/analysis/spectral_analysis.h:33-95), and
// the per-run residual-history metrics (src/analysis/convergence_analysis.h).
//
// Both iterations use the product's SpMV and preconditioner apply; their
// scalars stay on the device (one-workgroup "finish" kernels over per-block
// partial sums, as in the Krylov solvers), and a device flag stops every later
// launch once an iteration decides to stop.  The start vectors are the
// reference's: a 32-bit LCG (multiplier 1664525, increment 1013904223), each
// state read as a signed int / 2^31 and normalised on the host in sequential
// order.  The Ritz extremes come from the Lanczos tridiagonal by Sturm-count
// bisection on the host.
#include "segs.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace sg {

namespace {

constexpr int kAT = 256;
constexpr int kATile = kAT * 4;
constexpr int kAF = 1024;

struct AState {
    int active;  // first field: SpMV / preconditioner guards read it
    int steps;   // Lanczos: alpha entries; power: iterations
    int scale;   // power: normalise v in this iteration
    int pad;
    double alpha, beta, mu_prev, en;
};

struct AWork {
    AState* st;
    double* part;
    double* alpha;  // Lanczos coefficients
    double* beta;
    int pstride;
};

// ---- Lanczos -------------------------------------------------------------------
// partials of v.z
__global__ void __launch_bounds__(kAT)
k_an_dot(Segs sp, const double* __restrict__ v, const double* __restrict__ z, AWork W) {
    if (!W.st->active) return;
    const BlockRange br = sp.blk[blockIdx.x];
    double a = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kAT) a += v[e] * z[e];
    block_partials2<kAT>(a, 0.0, W.part, W.pstride, 1);
}

__global__ void __launch_bounds__(kAF) k_an_alpha(Segs sp, AWork W, int j) {
    if (!W.st->active) return;
    double a;
    finish_sums2(W.part, W.pstride, sp.sys[0], 1, &a);
    if (threadIdx.x != 0) return;
    W.st->alpha = a;
    W.alpha[j] = a;
    W.st->steps = j + 1;
}

// z -= alpha v + beta v_prev (condition_estimator.h:139-141); partials of z.z
__global__ void __launch_bounds__(kAT)
k_an_orth(Segs sp, double* __restrict__ z, const double* __restrict__ v,
          const double* __restrict__ vp, AWork W) {
    if (!W.st->active) return;
    const BlockRange br = sp.blk[blockIdx.x];
    const double al = W.st->alpha, be = W.st->beta;
    double a = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kAT) {
        const double ze = z[e] - (al * v[e] + be * vp[e]);
        z[e] = ze;
        a += ze * ze;
    }
    block_partials2<kAT>(a, 0.0, W.part, W.pstride, 1);
}

// beta_{j+1} and the invariant-subspace stop (condition_estimator.h:143-156)
__global__ void __launch_bounds__(kAF) k_an_beta(Segs sp, AWork W, int j, int k) {
    if (!W.st->active) return;
    double zz;
    finish_sums2(W.part, W.pstride, sp.sys[0], 1, &zz);
    if (threadIdx.x != 0) return;
    AState* st = W.st;
    const double bn = sqrt(zz);
    if (bn < 1e-10 * (1.0 + fabs(st->alpha)) || j >= k - 1) {
        st->active = 0;
        return;
    }
    W.beta[j] = bn;
    st->beta = bn;
}

// v_next = z * (1 / beta) (condition_estimator.h:150-152)
__global__ void __launch_bounds__(kAT)
k_an_next(Segs sp, const double* __restrict__ z, double* __restrict__ vn, AWork W) {
    if (!W.st->active) return;
    const BlockRange br = sp.blk[blockIdx.x];
    const double inv = 1.0 / W.st->beta;
    for (int e = br.start + threadIdx.x; e < br.end; e += kAT) vn[e] = z[e] * inv;
}

// ---- power iteration -------------------------------------------------------------
// w = v - z; partials of v.w and w.w (spectral_analysis.h:64-76)
__global__ void __launch_bounds__(kAT)
k_pw_diff(Segs sp, const double* __restrict__ v, const double* z, double* w, AWork W) {
    if (!W.st->active) return;
    const BlockRange br = sp.blk[blockIdx.x];
    double a = 0.0, b = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kAT) {
        const double we = v[e] - z[e];
        w[e] = we;
        a += v[e] * we;
        b += we * we;
    }
    block_partials2<kAT>(a, b, W.part, W.pstride, 2);
}

// Rayleigh quotient, norm and the stopping tests (spectral_analysis.h:73-89)
__global__ void __launch_bounds__(kAF) k_pw_finish(Segs sp, AWork W, int it, double tol) {
    AState* st = W.st;
    if (!st->active) {
        if (threadIdx.x == 0) st->scale = 0;
        return;
    }
    double d[2];
    finish_sums2(W.part, W.pstride, sp.sys[0], 2, d);
    if (threadIdx.x != 0) return;
    const double mu = d[0], en = sqrt(d[1]);
    st->steps = it + 1;
    if (en < 1e-14) {
        st->scale = 0;
        st->active = 0;
        return;
    }
    st->en = en;
    st->scale = 1;
    if (fabs(mu - st->mu_prev) < tol * (1.0 + fabs(mu))) {
        st->active = 0;
        return;
    }
    st->mu_prev = mu;
}

// v = w / ||w||
__global__ void __launch_bounds__(kAT)
k_pw_scale(Segs sp, const double* __restrict__ w, double* __restrict__ v, AWork W) {
    if (!W.st->scale) return;
    const BlockRange br = sp.blk[blockIdx.x];
    const double en = W.st->en;
    for (int e = br.start + threadIdx.x; e < br.end; e += kAT) v[e] = w[e] / en;
}

std::vector<double> lcg_unit(int n, unsigned seed, bool guard_zero) {
    std::vector<double> v(n);
    unsigned st = seed;
    double nn = 0.0;
    for (int i = 0; i < n; ++i) {
        st = st * 1664525u + 1013904223u;
        v[i] = (double)(int)st / 2147483648.0;
        nn += v[i] * v[i];
    }
    nn = std::sqrt(nn);
    if (guard_zero && nn < 1e-14) nn = 1.0;
    for (int i = 0; i < n; ++i) v[i] /= nn;
    return v;
}

// Number of eigenvalues of the symmetric tridiagonal (d, e) below x (Sturm
// sequence of the LDL^T pivots).
int sturm_below(const std::vector<double>& d, const std::vector<double>& e, double x) {
    const int m = (int)d.size();
    int cnt = 0;
    double q = 1.0;
    for (int i = 0; i < m; ++i) {
        const double off = i ? e[i - 1] * e[i - 1] / q : 0.0;
        q = d[i] - x - off;
        if (q == 0.0) q = -1e-300;
        if (q < 0.0) ++cnt;
    }
    return cnt;
}

// The idx-th smallest eigenvalue by bisection to adjacent doubles.
double tridiag_eig(const std::vector<double>& d, const std::vector<double>& e, int idx) {
    const int m = (int)d.size();
    double lo = d[0], hi = d[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    const double pad = 1e-12 * std::max(1.0, std::max(std::fabs(lo), std::fabs(hi)));
    lo -= pad;
    hi += pad;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_below(d, e, mid) > idx) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

struct AnWs {
    DevBuf<double> vecs, small;
    DevBuf<AState> st;
    SegsHost seg;
    AWork W{};
    Segs sp{};
    size_t ld = 0;
};

void an_setup(sg_csr* A, sg_precond* M, int nvec, int ncoef, AnWs& ws) {
    sg_context* ctx = A->ctx;
    if (A->nrows != A->ncols) fail(SG_E_ARG, "analysis needs a square matrix");
    if (A->sys_off.size() != 2) fail(SG_E_ARG, "analysis runs on one system, not a batch");
    if (M && !precond_ready(M)) fail(SG_E_STATE, "preconditioner used before setup");
    if (M && precond_n(M) != A->nrows) fail(SG_E_ARG, "preconditioner size mismatch");
    if (M && precond_is_flexible(M)) fail(SG_E_ARG, "analysis needs a fixed preconditioner");
    const int n = A->nrows;
    hipStream_t s = ctx->stream;
    SG_HIP(hipSetDevice(ctx->device));
    ws.ld = ((size_t)n + 63) / 64 * 64;
    ws.vecs.alloc(ws.ld * nvec);
    SG_HIP(hipMemsetAsync(ws.vecs.p, 0, sizeof(double) * ws.ld * nvec, s));
    build_segs(A->sys_off, 4 * ctx->num_cu, kATile, ws.seg, s);
    ws.sp = ws.seg.dev();
    ws.small.alloc((size_t)2 * ws.seg.nblk + 2 * (size_t)std::max(ncoef, 1));
    ws.st.alloc(1);
    AState init{};
    init.active = 1;
    SG_HIP(hipMemcpyAsync(ws.st.p, &init, sizeof(init), hipMemcpyHostToDevice, s));
    ws.W = AWork{ws.st.p, ws.small.p, ws.small.p + 2 * ws.seg.nblk,
                 ws.small.p + 2 * ws.seg.nblk + std::max(ncoef, 1), ws.seg.nblk};
}

}  // namespace

}  // namespace sg

using namespace sg;

extern "C" sg_status sg_estimate_condition(sg_csr* A, sg_precond* M, int k, unsigned seed,
                                           sg_condition_estimate* out, double* alpha,
                                           double* beta) {
    if (!A || !out) return SG_E_ARG;
    return run_guarded(A->ctx, [&] {
        *out = sg_condition_estimate{};
        const int n = A->nrows;
        k = std::min(k, n);
        if (k < 2) return;  // condition_estimator.h:100-101
        sg_context* ctx = A->ctx;
        hipStream_t s = ctx->stream;
        AnWs ws;
        an_setup(A, M, 5, k, ws);
        double* V[3] = {ws.vecs.p, ws.vecs.p + ws.ld, ws.vecs.p + 2 * ws.ld};
        double* w = ws.vecs.p + 3 * ws.ld;
        double* z = M ? ws.vecs.p + 4 * ws.ld : w;
        std::vector<double> v0 = lcg_unit(n, seed, true);
        SG_HIP(hipMemcpyAsync(V[0], v0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        const Guard g{&ws.st.p->active, 0, false};
        const dim3 gv(ws.seg.nblk), bv(kAT), gf(1), bf(kAF);
        // v_prev (V[2] at j = 0) is zero and beta = 0 in the first step
        for (int j = 0; j < k; ++j) {
            double* vc = V[j % 3];
            double* vp = V[(j + 2) % 3];
            double* vn = V[(j + 1) % 3];
            spmv_launch(A, 1.0, vc, 0.0, nullptr, w, g, s);
            if (M) precond_apply_launch(M, w, z, g);
            hipLaunchKernelGGL(k_an_dot, gv, bv, 0, s, ws.sp, vc, z, ws.W);
            hipLaunchKernelGGL(k_an_alpha, gf, bf, 0, s, ws.sp, ws.W, j);
            hipLaunchKernelGGL(k_an_orth, gv, bv, 0, s, ws.sp, z, vc, vp, ws.W);
            hipLaunchKernelGGL(k_an_beta, gf, bf, 0, s, ws.sp, ws.W, j, k);
            hipLaunchKernelGGL(k_an_next, gv, bv, 0, s, ws.sp, z, vn, ws.W);
            SG_HIP(hipGetLastError());
            if ((j & 7) == 7) {  // stop launching once the recurrence has stopped
                AState h;
                SG_HIP(hipMemcpyAsync(&h, ws.st.p, sizeof(h), hipMemcpyDeviceToHost, s));
                SG_HIP(hipStreamSynchronize(s));
                if (!h.active) break;
            }
        }
        AState h;
        SG_HIP(hipMemcpyAsync(&h, ws.st.p, sizeof(h), hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        const int m = h.steps;
        std::vector<double> al(std::max(m, 1)), be(std::max(m, 1), 0.0);
        if (m) SG_HIP(hipMemcpy(al.data(), ws.W.alpha, sizeof(double) * m, hipMemcpyDeviceToHost));
        if (m > 1)
            SG_HIP(hipMemcpy(be.data(), ws.W.beta, sizeof(double) * (m - 1), hipMemcpyDeviceToHost));
        al.resize(m);
        be.resize(std::max(m - 1, 0));
        if (alpha) std::copy(al.begin(), al.end(), alpha);
        if (beta) std::copy(be.begin(), be.end(), beta);
        if (m < 2) return;  // condition_estimator.h:160-161
        for (double a : al)
            if (!std::isfinite(a)) return;
        for (double b : be)
            if (!std::isfinite(b)) return;
        const double lmin = tridiag_eig(al, be, 0), lmax = tridiag_eig(al, be, m - 1);
        out->lambda_min = lmin;
        out->lambda_max = lmax;
        out->kappa = std::fabs(lmin) > 1e-14 ? lmax / lmin : 1e30;
        out->lanczos_steps = m;
        out->valid = lmin > 0.0;
    });
}

extern "C" sg_status sg_estimate_spectral_radius(sg_csr* A, sg_precond* M, int max_iter,
                                                 double tol, unsigned seed,
                                                 sg_spectral_result* out, double* d_eigvec) {
    if (!A || !out || max_iter < 0) return SG_E_ARG;
    return run_guarded(A->ctx, [&] {
        *out = sg_spectral_result{};
        const int n = A->nrows;
        sg_context* ctx = A->ctx;
        hipStream_t s = ctx->stream;
        AnWs ws;
        an_setup(A, M, 3, 1, ws);
        double* v = ws.vecs.p;
        double* w = v + ws.ld;
        double* z = M ? w + ws.ld : w;
        std::vector<double> v0 = lcg_unit(n, seed, false);
        SG_HIP(hipMemcpyAsync(v, v0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        const Guard g{&ws.st.p->active, 0, false};
        const dim3 gv(ws.seg.nblk), bv(kAT), gf(1), bf(kAF);
        for (int it = 0; it < max_iter; ++it) {
            spmv_launch(A, 1.0, v, 0.0, nullptr, w, g, s);
            if (M) precond_apply_launch(M, w, z, g);
            // with M == null, z aliases w: w = v - w reads each element before writing it
            hipLaunchKernelGGL(k_pw_diff, gv, bv, 0, s, ws.sp, v, z, w, ws.W);
            hipLaunchKernelGGL(k_pw_finish, gf, bf, 0, s, ws.sp, ws.W, it, tol);
            hipLaunchKernelGGL(k_pw_scale, gv, bv, 0, s, ws.sp, w, v, ws.W);
            SG_HIP(hipGetLastError());
            if ((it & 7) == 7) {
                AState h;
                SG_HIP(hipMemcpyAsync(&h, ws.st.p, sizeof(h), hipMemcpyDeviceToHost, s));
                SG_HIP(hipStreamSynchronize(s));
                if (!h.active) break;
            }
        }
        AState h;
        SG_HIP(hipMemcpyAsync(&h, ws.st.p, sizeof(h), hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        out->spectral_radius = std::fabs(h.mu_prev);
        out->convergence_ok = out->spectral_radius < 1.0;
        out->power_iters = h.steps;
        if (d_eigvec)
            SG_HIP(hipMemcpyAsync(d_eigvec, v, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        SG_HIP(hipStreamSynchronize(s));
    });
}
