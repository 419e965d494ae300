// The learned FGMRES stage: a base (two-level Schwarz) preconditioner plus a
// network correction of its remaining residual, kept only when it reduces the
// residual.  Every scalar stays on the device, so an application launches
// without host synchronisation (and can sit inside a captured solve):
//
//   z_b = base(x);  s = x - A z_b;  ss = s.s
//   e   = sqrt(ss/n) * net(s / sqrt(ss/n), log kappa)
//   a = s.(A e), c = (A e).(A e)            (||s - alpha A e||^2 = ss - a^2 / c)
//   z = z_b + (a / c) e  if a^2 >= min_gain * c * ss,  else  z = z_b
//
// The stage changes from call to call, which FGMRES allows
// (src/core/fgmres.cpp:28-172, FlexApply src/core/fgmres.h:37).
#include "sg_internal.h"

#include <cmath>

namespace sg {

struct LearnedStage {
    sg_context* ctx = nullptr;
    sg_precond* base = nullptr;
    sg_unet* net = nullptr;
    const sg_csr* A = nullptr;
    int nx = 0, ny = 0, n = 0;
    double min_gain = 0.01;
    int every = 1;
    long long calls_host = 0;
    DevBuf<double> s, e, Ae, scal;  // scal: ss, a, c
    DevBuf<float> lk, nin;
    DevBuf<int> cnt;                // applications, accepted corrections
};

__global__ void k_stage_in(int n, const double* __restrict__ s, const double* __restrict__ ss,
                           const float* __restrict__ lk, float* __restrict__ nin) {
    const double r = *ss > 0.0 ? 1.0 / sqrt(*ss / n) : 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        nin[i] = (float)(s[i] * r);
        nin[n + i] = lk[i];
    }
}

__global__ void k_stage_out(int n, const float* __restrict__ o, const double* __restrict__ ss,
                            double* __restrict__ e) {
    const double r = sqrt(*ss / n);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        e[i] = (double)o[i] * r;
}

__global__ void k_stage_final(int n, const double* __restrict__ sc, const double* __restrict__ e,
                              double* __restrict__ y, double min_gain, int* __restrict__ cnt,
                              DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    const double ss = sc[0], a = sc[1], c = sc[2];
    const bool ok = c > 0.0 && ss > 0.0 && a * a >= min_gain * c * ss;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[0] += 1;
        cnt[1] += ok ? 1 : 0;
    }
    if (!ok) return;
    const double al = a / c;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] += al * e[i];
}

__global__ void k_count_base(int* __restrict__ cnt, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    cnt[0] += 1;
}

static int grid_n(int n) { return std::max(1, std::min((n + 255) / 256, 8192)); }

LearnedStage* learned_new(sg_context* ctx, sg_precond* base, sg_unet* net, int nx, int ny,
                          const double* log_kappa, const sg_learned_opts* o) {
    if (nx < 8 || ny < 8 || nx % 8 || ny % 8)
        fail(SG_E_ARG, "learned stage: grid sides must be multiples of 8");
    LearnedStage* L = new LearnedStage();
    L->ctx = ctx;
    L->base = base;
    L->net = net;
    L->nx = nx;
    L->ny = ny;
    L->n = nx * ny;
    if (o) {
        if (o->min_gain > 0.0) L->min_gain = o->min_gain;
        if (o->every > 1) L->every = o->every;
    }
    std::vector<float> lk(log_kappa, log_kappa + L->n);
    L->lk.upload(lk, ctx->stream);
    L->s.alloc(L->n);
    L->e.alloc(L->n);
    L->Ae.alloc(L->n);
    L->scal.alloc(3);
    L->nin.alloc((size_t)2 * L->n);
    L->cnt.alloc(2);
    SG_HIP(hipMemsetAsync(L->cnt.p, 0, sizeof(int) * 2, ctx->stream));
    SG_HIP(hipStreamSynchronize(ctx->stream));
    return L;
}

void learned_free(LearnedStage* L) { delete L; }

sg_precond* learned_base(const LearnedStage* L) { return L->base; }

void learned_setup(LearnedStage* L, const sg_csr* A) {
    if (A->nrows != L->n || A->ncols != L->n) fail(SG_E_ARG, "learned stage: matrix size != nx * ny");
    if (!precond_ready(L->base)) fail(SG_E_STATE, "learned stage: base preconditioner not set up");
    if (precond_n(L->base) != L->n) fail(SG_E_ARG, "learned stage: base preconditioner size mismatch");
    L->A = A;
    learned_prepare(L);
}

void learned_prepare(LearnedStage* L) {
    precond_ensure_tuned(L->base);
    spmv_ensure_tuned(const_cast<sg_csr*>(L->A));
    unet_prepare(L->net, 1, L->ny, L->nx);
}

void learned_apply(LearnedStage* L, const double* x, double* y, const Guard& g) {
    hipStream_t st = L->ctx->stream;
    sg_csr* A = const_cast<sg_csr*>(L->A);
    const int n = L->n;
    precond_apply_launch(L->base, x, y, g);
    const bool run_net = (L->calls_host++ % L->every) == 0;
    if (!run_net) {
        hipLaunchKernelGGL(k_count_base, dim3(1), dim3(64), 0, st, L->cnt.p, dev_guard(g, nullptr));
        SG_HIP(hipGetLastError());
        return;
    }
    spmv_launch(A, -1.0, y, 1.0, x, L->s.p, g, st);  // s = x - A z_b
    dot_launch(L->ctx, n, L->s.p, L->s.p, L->scal.p);
    hipLaunchKernelGGL(k_stage_in, dim3(grid_n(n)), dim3(256), 0, st, n, L->s.p, L->scal.p, L->lk.p,
                       L->nin.p);
    unet_forward_dev(L->net, 1, L->ny, L->nx, L->nin.p, false);
    hipLaunchKernelGGL(k_stage_out, dim3(grid_n(n)), dim3(256), 0, st, n, unet_output(L->net),
                       L->scal.p, L->e.p);
    spmv_launch(A, 1.0, L->e.p, 0.0, nullptr, L->Ae.p, g, st);
    dot_launch(L->ctx, n, L->s.p, L->Ae.p, L->scal.p + 1);
    dot_launch(L->ctx, n, L->Ae.p, L->Ae.p, L->scal.p + 2);
    hipLaunchKernelGGL(k_stage_final, dim3(grid_n(n)), dim3(256), 0, st, n, L->scal.p, L->e.p, y,
                       L->min_gain, L->cnt.p, dev_guard(g, nullptr));
    SG_HIP(hipGetLastError());
}

void learned_counts(const LearnedStage* L, long long* calls, long long* accepted) {
    int c[2] = {0, 0};
    SG_HIP(hipMemcpyAsync(c, L->cnt.p, sizeof(c), hipMemcpyDeviceToHost, L->ctx->stream));
    SG_HIP(hipStreamSynchronize(L->ctx->stream));
    if (calls) *calls = c[0];
    if (accepted) *accepted = c[1];
}

}  // namespace sg
