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
// With coarsen = c > 1 the network sees the c x c block averages of s and of
// log kappa (normalised by the averaged residual's own rms) and its output is
// interpolated bilinearly (cell-centred) back to the grid; the A-scaling and
// the acceptance test are the same.
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
    int coarsen = 1, cnx = 0, cny = 0, cn = 0;  // the network's grid
    long long calls_host = 0;
    DevBuf<double> s, e, Ae, scal;  // scal: ss, a, c, ss of the averaged residual
    DevBuf<double> sc;              // block averages of s (coarsen > 1)
    DevBuf<float> lk, lkc, nin;
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

// c x c block averages: sc (double) of s, lkc of log kappa (coarse grid cnx x cny)
__global__ void k_stage_pool(int cnx, int cny, int c, int nx, const double* __restrict__ s,
                             const float* __restrict__ lk, double* __restrict__ sc,
                             float* __restrict__ lkc) {
    const int cn = cnx * cny;
    const double w = 1.0 / ((double)c * c);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cn; i += gridDim.x * blockDim.x) {
        const int I = i % cnx, J = i / cnx;
        double a = 0.0, l = 0.0;
        for (int dy = 0; dy < c; ++dy)
            for (int dx = 0; dx < c; ++dx) {
                const int f = (J * c + dy) * nx + I * c + dx;
                a += s[f];
                l += lk[f];
            }
        sc[i] = a * w;
        lkc[i] = (float)(l * w);
    }
}

// e = rms * bilinear interpolation (cell-centred, clamped at the edges) of the
// coarse network output o (cnx x cny) onto the nx x ny grid
__global__ void k_stage_up(int nx, int ny, int c, int cnx, int cny, const float* __restrict__ o,
                           const double* __restrict__ ss, double* __restrict__ e) {
    const int n = nx * ny;
    const double r = sqrt(*ss / ((double)cnx * cny));
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = i % nx, y = i / nx;
        const double fx = (x + 0.5) / c - 0.5, fy = (y + 0.5) / c - 0.5;
        int x0 = (int)floor(fx), y0 = (int)floor(fy);
        const double tx = fx - x0, ty = fy - y0;
        const int x1 = min(max(x0 + 1, 0), cnx - 1), y1 = min(max(y0 + 1, 0), cny - 1);
        x0 = min(max(x0, 0), cnx - 1);
        y0 = min(max(y0, 0), cny - 1);
        const double v = (1.0 - ty) * ((1.0 - tx) * o[y0 * cnx + x0] + tx * o[y0 * cnx + x1]) +
                         ty * ((1.0 - tx) * o[y1 * cnx + x0] + tx * o[y1 * cnx + x1]);
        e[i] = v * r;
    }
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
        if (o->coarsen > 1) L->coarsen = o->coarsen;
    }
    const int c = L->coarsen;
    if (c != 1 && c != 2 && c != 4) {
        delete L;
        fail(SG_E_ARG, "learned stage: coarsen must be 1, 2 or 4");
    }
    if (nx % (8 * c) || ny % (8 * c)) {
        delete L;
        fail(SG_E_ARG, "learned stage: nx / coarsen and ny / coarsen must be multiples of 8");
    }
    L->cnx = nx / c;
    L->cny = ny / c;
    L->cn = L->cnx * L->cny;
    std::vector<float> lk(log_kappa, log_kappa + L->n);
    L->lk.upload(lk, ctx->stream);
    L->s.alloc(L->n);
    L->e.alloc(L->n);
    L->Ae.alloc(L->n);
    L->scal.alloc(4);
    L->nin.alloc((size_t)2 * L->cn);
    if (c > 1) {
        L->sc.alloc(L->cn);
        L->lkc.alloc(L->cn);
    }
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
    unet_prepare(L->net, 1, L->cny, L->cnx);
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
    if (L->coarsen == 1) {
        hipLaunchKernelGGL(k_stage_in, dim3(grid_n(n)), dim3(256), 0, st, n, L->s.p, L->scal.p,
                           L->lk.p, L->nin.p);
        unet_forward_dev(L->net, 1, L->ny, L->nx, L->nin.p, false);
        hipLaunchKernelGGL(k_stage_out, dim3(grid_n(n)), dim3(256), 0, st, n, unet_output(L->net),
                           L->scal.p, L->e.p);
    } else {
        const int cn = L->cn;
        hipLaunchKernelGGL(k_stage_pool, dim3(grid_n(cn)), dim3(256), 0, st, L->cnx, L->cny,
                           L->coarsen, L->nx, L->s.p, L->lk.p, L->sc.p, L->lkc.p);
        dot_launch(L->ctx, cn, L->sc.p, L->sc.p, L->scal.p + 3);
        hipLaunchKernelGGL(k_stage_in, dim3(grid_n(cn)), dim3(256), 0, st, cn, L->sc.p,
                           L->scal.p + 3, L->lkc.p, L->nin.p);
        unet_forward_dev(L->net, 1, L->cny, L->cnx, L->nin.p, false);
        hipLaunchKernelGGL(k_stage_up, dim3(grid_n(n)), dim3(256), 0, st, L->nx, L->ny, L->coarsen,
                           L->cnx, L->cny, unet_output(L->net), L->scal.p + 3, L->e.p);
    }
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
