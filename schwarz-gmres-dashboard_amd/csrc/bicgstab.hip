// Right-preconditioned BiCGSTAB (Van der Vorst 1992) with every scalar on the
// device, for one system or for a batch of independent systems stored as one
// block-diagonal matrix.
//
// Semantics follow src/core/bicgstab.h:57-204: r = b - A x0, fixed shadow
// residual r^ = r0, rho = alpha = omega = 1; relative tests ||s|| < tol ||r0||
// (half step, x += alpha y) and ||r|| < tol ||r0|| (full step); breakdown when
// |rho| or |(r^, v)| < 1e-300, (t, t) < 1e-300 ends with x += alpha y, and
// |omega| < 1e-300 stops after the full step; ||r0|| < 1e-14 returns at once
// with no iteration.  Vector updates keep the reference's operation order:
// p = r + beta (p - omega v), s = r - alpha v, x += (alpha y + omega z),
// r = s - omega t.
//
// Each iteration is a fixed chain of launches:
//   p-update | y = M^-1 p | v = A y | (r^,v) | alpha | s, (s,s) | ||s|| test |
//   z = M^-1 s | t = A z | (t,t),(t,s) | omega | x, r, (r,r),(r^,r) | next
// Vector kernels write per-block partial sums; one-workgroup-per-system
// "finish" kernels add them in a fixed order and run the scalar logic, so no
// scalar ever travels to the host inside the solve.  A system that stops drops
// out of every later launch (vector kernels check its state, SpMV and
// preconditioner take the step guard).  The host keeps `lookahead` iterations
// in flight and stops launching once every system is done; with a graph the
// iteration is captured once and replayed.
#include "segs.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace sg {

namespace {

constexpr int kBT = 256;             // threads of the vector kernels
constexpr int kBTile = kBT * 4;      // elements per block and pass (chunk granularity)
constexpr int kBF = 1024;            // threads of the finish kernels

struct BState {
    int active;     // iterating; first field: SpMV / preconditioner guards read it
    int converged;
    int iters;      // GMRESResult::iterations
    int it;         // current iteration (1-based)
    int nhist;
    int applies;
    int pend;       // 1: x += alpha y is still to be applied (half-step exits)
    int has_m;
    int maxiter;
    int pad;
    double tol;
    double r0;      // ||r_0||
    double rho, rho_prev, alpha, omega, beta;
    double snorm;   // ||s|| of the current iteration
    double rnorm;   // ||r|| of the current residual
    double res;     // GMRESResult::residual_norm
    double nrm;     // scratch: a norm for the host (||b||, ||b - A x||)
};

struct BWork {
    BState* st;
    double* part;   // part[c * pstride + block], c < 2
    double* hist;   // hcap per system
    int pstride, hcap;
};

__device__ inline void block_partials(double a, double b, const BWork& W, int ncols) {
    block_partials2<kBT>(a, b, W.part, W.pstride, ncols);
}

__device__ inline void finish_sums(const BWork& W, const SysBlocks& sb, int ncols, double* out) {
    finish_sums2(W.part, W.pstride, sb, ncols, out);
}

// ---- start -------------------------------------------------------------------
// r^ = r and the partials of r.r (r = b - A x0 computed by the SpMV before).
__global__ void __launch_bounds__(kBT)
k_bi_start(Segs sp, const double* __restrict__ r, double* __restrict__ rh, BWork W) {
    const BlockRange br = sp.blk[blockIdx.x];
    double a = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kBT) {
        const double v = r[e];
        rh[e] = v;
        a += v * v;
    }
    block_partials(a, 0.0, W, 1);
}

// ||r0||, the early return and rho_1 = (r^, r0) (bicgstab.h:84-102).
__global__ void __launch_bounds__(kBF) k_bi_start_finish(Segs sp, BWork W) {
    const int q = blockIdx.x;
    double rr;
    finish_sums(W, sp.sys[q], 1, &rr);
    if (threadIdx.x != 0) return;
    BState* st = W.st + q;
    const double r0 = sqrt(rr);
    st->r0 = r0;
    st->rnorm = r0;
    st->it = 1;
    st->active = 0;
    if (r0 < 1e-14) {  // already converged at x0: {0, ||r0||, true}
        st->converged = 1;
        st->res = r0;
        return;
    }
    if (st->maxiter < 1) return;  // loop never runs: residual_norm stays 0
    st->rho_prev = 1.0;
    st->rho = rr;  // (r^, r0) with r^ = r0: the same products in the same order
    st->alpha = 1.0;
    st->omega = 1.0;
    if (fabs(st->rho) < 1e-300) {
        st->iters = 1;
        st->res = r0;
        return;
    }
    st->active = 1;
}

// ---- one iteration -------------------------------------------------------------
// p = r (first iteration) or p = r + beta (p - omega v) (bicgstab.h:110-117).
__global__ void __launch_bounds__(kBT)
k_bi_p(Segs sp, const double* __restrict__ r, const double* __restrict__ v,
       double* __restrict__ p, BWork W) {
    const BlockRange br = sp.blk[blockIdx.x];
    const BState* st = W.st + br.sys;
    if (!st->active) return;
    if (st->it == 1) {
        for (int e = br.start + threadIdx.x; e < br.end; e += kBT) p[e] = r[e];
        return;
    }
    const double beta = st->beta, om = st->omega;
    for (int e = br.start + threadIdx.x; e < br.end; e += kBT)
        p[e] = r[e] + beta * (p[e] - om * v[e]);
}

// Partials of a.b (and c.d when c is given) for the active systems.
__global__ void __launch_bounds__(kBT)
k_bi_dots(Segs sp, const double* __restrict__ a, const double* __restrict__ b,
          const double* __restrict__ c, const double* __restrict__ d, BWork W) {
    const BlockRange br = sp.blk[blockIdx.x];
    if (!W.st[br.sys].active) return;
    double s0 = 0.0, s1 = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kBT) {
        s0 += a[e] * b[e];
        if (c) s1 += c[e] * d[e];
    }
    block_partials(s0, s1, W, c ? 2 : 1);
}

// alpha = rho / (r^, v) with the breakdown test (bicgstab.h:126-134).
__global__ void __launch_bounds__(kBF) k_bi_alpha(Segs sp, BWork W) {
    const int q = blockIdx.x;
    BState* st = W.st + q;
    if (!st->active) return;
    double rv;
    finish_sums(W, sp.sys[q], 1, &rv);
    if (threadIdx.x != 0) return;
    st->applies += st->has_m;  // y = M^-1 p
    if (fabs(rv) < 1e-300) {
        st->iters = st->it;
        st->res = st->rnorm;  // norm2(r) of the current residual
        st->active = 0;
        return;
    }
    st->alpha = st->rho / rv;
}

// s = r - alpha v and the partials of s.s (bicgstab.h:136-139).
__global__ void __launch_bounds__(kBT)
k_bi_s(Segs sp, const double* __restrict__ r, const double* __restrict__ v,
       double* __restrict__ s, BWork W) {
    const BlockRange br = sp.blk[blockIdx.x];
    const BState* st = W.st + br.sys;
    if (!st->active) return;
    const double al = st->alpha;
    double a = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kBT) {
        const double se = r[e] - al * v[e];
        s[e] = se;
        a += se * se;
    }
    block_partials(a, 0.0, W, 1);
}

// Half-step convergence test on ||s|| (bicgstab.h:139-148).
__global__ void __launch_bounds__(kBF) k_bi_snorm(Segs sp, BWork W) {
    const int q = blockIdx.x;
    BState* st = W.st + q;
    if (!st->active) return;
    double ss;
    finish_sums(W, sp.sys[q], 1, &ss);
    if (threadIdx.x != 0) return;
    const double sn = sqrt(ss);
    st->snorm = sn;
    if (sn < st->tol * st->r0) {
        st->pend = 1;
        st->iters = st->it;
        st->res = sn;
        st->converged = 1;
        if (st->nhist < W.hcap) W.hist[(size_t)q * W.hcap + st->nhist++] = sn;
        st->active = 0;
    }
}

// omega = (t, s) / (t, t), with the t = 0 exit (bicgstab.h:157-167).
__global__ void __launch_bounds__(kBF) k_bi_omega(Segs sp, BWork W) {
    const int q = blockIdx.x;
    BState* st = W.st + q;
    if (!st->active) return;
    double d[2];
    finish_sums(W, sp.sys[q], 2, d);
    if (threadIdx.x != 0) return;
    st->applies += st->has_m;  // z = M^-1 s
    const double tt = d[0];
    if (tt < 1e-300) {
        st->pend = 1;
        st->iters = st->it;
        st->res = st->snorm;
        st->converged = st->snorm < st->tol * st->r0;
        st->active = 0;
        return;
    }
    st->omega = d[1] / tt;
}

// x += alpha y + omega z, r = s - omega t and the partials of r.r and r^.r
// (bicgstab.h:169-177); a system with a pending half step only does
// x += alpha y (bicgstab.h:142,160).
__global__ void __launch_bounds__(kBT)
k_bi_xr(Segs sp, double* __restrict__ x, const double* __restrict__ y,
        const double* __restrict__ z, const double* __restrict__ s, const double* __restrict__ t,
        double* __restrict__ r, const double* __restrict__ rh, BWork W) {
    const BlockRange br = sp.blk[blockIdx.x];
    const BState* st = W.st + br.sys;
    const double al = st->alpha;
    if (st->pend) {
        for (int e = br.start + threadIdx.x; e < br.end; e += kBT) x[e] += al * y[e];
        return;
    }
    if (!st->active) return;
    const double om = st->omega;
    double a = 0.0, b = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kBT) {
        x[e] += al * y[e] + om * z[e];
        const double re = s[e] - om * t[e];
        r[e] = re;
        a += re * re;
        b += rh[e] * re;
    }
    block_partials(a, b, W, 2);
}

// Full-step test, omega breakdown, the iteration bound and the start of the
// next iteration: rho = (r^, r), its breakdown test and beta
// (bicgstab.h:177-199, 99-114).
__global__ void __launch_bounds__(kBF) k_bi_next(Segs sp, BWork W) {
    const int q = blockIdx.x;
    BState* st = W.st + q;
    if (st->pend) {
        if (threadIdx.x == 0) st->pend = 0;
        return;
    }
    if (!st->active) return;
    double d[2];
    finish_sums(W, sp.sys[q], 2, d);
    if (threadIdx.x != 0) return;
    const double rn = sqrt(d[0]);
    st->rnorm = rn;
    if (st->nhist < W.hcap) W.hist[(size_t)q * W.hcap + st->nhist++] = rn;
    st->iters = st->it;
    st->res = rn;
    if (rn < st->tol * st->r0) {
        st->converged = 1;
        st->active = 0;
        return;
    }
    if (fabs(st->omega) < 1e-300 || st->it >= st->maxiter) {
        st->active = 0;
        return;
    }
    st->it += 1;
    st->rho_prev = st->rho;
    st->rho = d[1];
    if (fabs(st->rho) < 1e-300) {
        st->iters = st->it;
        st->res = rn;
        st->active = 0;
        return;
    }
    st->beta = (st->rho / st->rho_prev) * (st->alpha / st->omega);
}

// ---- host control and the exit residual ---------------------------------------
// any[0] = 1 if some system is still iterating or has a pending half step.
__global__ void k_bi_any(const BState* __restrict__ st, int nsys, int* __restrict__ any) {
    __shared__ int f;
    if (threadIdx.x == 0) f = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < nsys; q += blockDim.x)
        if (st[q].active || st[q].pend) f = 1;  // benign race: every writer stores 1
    __syncthreads();
    if (threadIdx.x == 0) *any = f;
}

__global__ void __launch_bounds__(kBT)
k_bi_norm(Segs sp, const double* __restrict__ u, BWork W) {
    const BlockRange br = sp.blk[blockIdx.x];
    double a = 0.0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kBT) a += u[e] * u[e];
    block_partials(a, 0.0, W, 1);
}

__global__ void __launch_bounds__(kBF) k_bi_norm_finish(Segs sp, BWork W) {
    double uu;
    finish_sums(W, sp.sys[blockIdx.x], 1, &uu);
    if (threadIdx.x == 0) W.st[blockIdx.x].nrm = sqrt(uu);
}

__global__ void k_bi_init(BState* st, int nsys, int has_m, int maxiter, double tol) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsys) return;
    BState b{};
    b.has_m = has_m;
    b.maxiter = maxiter;
    b.tol = tol;
    b.rho = b.rho_prev = b.alpha = b.omega = 1.0;
    st[q] = b;
}

struct BiWs {
    DevBuf<double> vecs, small;
    DevBuf<BState> st;
    DevBuf<int> any;
    int* pinned = nullptr;
    BState* h_st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BiWs() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pinned) (void)hipHostFree(pinned);
        if (h_st) (void)hipHostFree(h_st);
    }
};

}  // namespace

void bicgstab_impl(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                   const sg_solve_params* prm, sg_solve_result* res, double* hist, int hist_cap,
                   bool use_graph) {
    sg_context* ctx = A->ctx;
    if (A->nrows != A->ncols) fail(SG_E_ARG, "BiCGSTAB needs a square matrix");
    if (M && !precond_ready(M)) fail(SG_E_STATE, "preconditioner used before setup");
    if (M && precond_n(M) != A->nrows) fail(SG_E_ARG, "preconditioner size mismatch");
    if (prm->max_iter < 0) fail(SG_E_ARG, "max_iter must be >= 0");
    if (prm->rtol > 0.0) fail(SG_E_ARG, "rtol is a GMRES/FGMRES option; BiCGSTAB's tol is relative");
    if (M && precond_is_flexible(M))
        fail(SG_E_ARG, "BiCGSTAB needs a fixed preconditioner (a learned stage needs FGMRES)");
    const int n = A->nrows;
    if (n < 1) fail(SG_E_ARG, "empty system");
    const std::vector<int>& off = A->sys_off;
    const int nsys = (int)off.size() - 1;
    if (M && precond_is_callback(M) && (use_graph || nsys > 1))
        fail(SG_E_ARG, "a host-callback preconditioner runs only in unbatched, uncaptured solves");
    const int L = prm->lookahead > 0 ? prm->lookahead : 2;
    hipStream_t s = ctx->stream;
    SG_HIP(hipSetDevice(ctx->device));
    if (M && nsys > 1) precond_bind_systems(M, off);
    auto t0 = std::chrono::steady_clock::now();

    const size_t ld = ((size_t)n + 63) / 64 * 64;
    // r, r^, p, v, s, t, and y, z when preconditioned
    const int nvec = M ? 8 : 6;
    BiWs ws;
    ws.vecs.alloc(ld * nvec);
    double* r = ws.vecs.p;
    double* rh = r + ld;
    double* p = rh + ld;
    double* v = p + ld;
    double* sv = v + ld;
    double* t = sv + ld;
    double* y = M ? t + ld : p;
    double* z = M ? y + ld : sv;
    SG_HIP(hipMemsetAsync(ws.vecs.p, 0, sizeof(double) * ld * nvec, s));  // v = p = 0

    SegsHost seg;
    build_segs(off, 4 * ctx->num_cu, kBTile, seg, s);
    const Segs sp = seg.dev();
    const int hcap = std::max(1, prm->max_iter);
    ws.small.alloc((size_t)2 * seg.nblk + (size_t)hcap * nsys);
    ws.st.alloc(nsys);
    ws.any.alloc(1);
    SG_HIP(hipHostMalloc((void**)&ws.pinned, sizeof(int) * L, hipHostMallocDefault));
    SG_HIP(hipHostMalloc((void**)&ws.h_st, sizeof(BState) * nsys, hipHostMallocDefault));
    BWork W{ws.st.p, ws.small.p, ws.small.p + (size_t)2 * seg.nblk, seg.nblk, hcap};
    const Guard step_guard{&ws.st.p->active, (int)(sizeof(BState) / sizeof(int)), nsys > 1};
    const dim3 gv(seg.nblk), bv(kBT), gf(nsys), bf(kBF);

    hipLaunchKernelGGL(k_bi_init, dim3((nsys + 63) / 64), dim3(64), 0, s, ws.st.p, nsys,
                       M ? 1 : 0, prm->max_iter, prm->tol_abs);
    // ||b|| for the relative true residual reported at exit
    hipLaunchKernelGGL(k_bi_norm, gv, bv, 0, s, sp, d_b, W);
    hipLaunchKernelGGL(k_bi_norm_finish, gf, bf, 0, s, sp, W);
    SG_HIP(hipMemcpyAsync(ws.h_st, ws.st.p, sizeof(BState) * nsys, hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    std::vector<double> bnorm(nsys);
    for (int q = 0; q < nsys; ++q) bnorm[q] = ws.h_st[q].nrm;

    {
        ProfScope ps(ctx, SG_K_UPDATE);
        spmv_launch(A, -1.0, d_x, 1.0, d_b, r, Guard{}, s);  // r = b - A x0
        hipLaunchKernelGGL(k_bi_start, gv, bv, 0, s, sp, r, rh, W);
        hipLaunchKernelGGL(k_bi_start_finish, gf, bf, 0, s, sp, W);
        SG_HIP(hipGetLastError());
    }

    auto launch_iteration = [&]() {
        {
            ProfScope ps(ctx, SG_K_UPDATE);
            hipLaunchKernelGGL(k_bi_p, gv, bv, 0, s, sp, r, v, p, W);
        }
        if (M) precond_apply_launch(M, p, y, step_guard);
        spmv_launch(A, 1.0, y, 0.0, nullptr, v, step_guard, s);
        {
            ProfScope ps(ctx, SG_K_ORTHO);
            hipLaunchKernelGGL(k_bi_dots, gv, bv, 0, s, sp, rh, v, (const double*)nullptr,
                               (const double*)nullptr, W);
            hipLaunchKernelGGL(k_bi_alpha, gf, bf, 0, s, sp, W);
            hipLaunchKernelGGL(k_bi_s, gv, bv, 0, s, sp, r, v, sv, W);
            hipLaunchKernelGGL(k_bi_snorm, gf, bf, 0, s, sp, W);
        }
        if (M) precond_apply_launch(M, sv, z, step_guard);
        spmv_launch(A, 1.0, z, 0.0, nullptr, t, step_guard, s);
        {
            ProfScope ps(ctx, SG_K_ORTHO);
            hipLaunchKernelGGL(k_bi_dots, gv, bv, 0, s, sp, t, t, t, sv, W);
            hipLaunchKernelGGL(k_bi_omega, gf, bf, 0, s, sp, W);
        }
        {
            ProfScope ps(ctx, SG_K_UPDATE);
            hipLaunchKernelGGL(k_bi_xr, gv, bv, 0, s, sp, d_x, y, z, sv, t, r, rh, W);
            hipLaunchKernelGGL(k_bi_next, gf, bf, 0, s, sp, W);
        }
        hipLaunchKernelGGL(k_bi_any, dim3(1), dim3(256), 0, s, ws.st.p, nsys, ws.any.p);
        SG_HIP(hipGetLastError());
    };

    if (use_graph) {
        spmv_ensure_tuned(A);
        if (M) precond_ensure_tuned(M);
        SG_HIP(hipStreamSynchronize(s));
        ProfOff po(ctx);
        CaptureGuard cg{s};
        SG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        cg.on = true;
        launch_iteration();
        cg.on = false;
        SG_HIP(hipStreamEndCapture(s, &ws.graph));
        SG_HIP(hipGraphInstantiate(&ws.exec, ws.graph, nullptr, nullptr, 0));
    }

    std::vector<hipEvent_t> evs(L);
    for (int j = 0; j < L; ++j) evs[j] = ctx->get_event();
    struct EvBack {
        sg_context* c;
        std::vector<hipEvent_t>& e;
        ~EvBack() {
            for (auto x : e) c->event_pool.push_back(x);
        }
    } evback{ctx, evs};

    // the start may already have stopped every system
    SG_HIP(hipMemcpyAsync(ws.h_st, ws.st.p, sizeof(BState) * nsys, hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    bool any0 = false;
    for (int q = 0; q < nsys; ++q) any0 = any0 || ws.h_st[q].active;
    // Iteration i+1 may run past the end of iteration i: every launch of a
    // stopped system returns at once, and the host waits for iteration i - L
    // before launching i, stopping once nothing is left to do.
    for (int i = 0; any0 && i < prm->max_iter; ++i) {
        if (i >= L) {
            SG_HIP(hipEventSynchronize(evs[i % L]));
            if (!ws.pinned[i % L]) break;
        }
        if (use_graph) SG_HIP(hipGraphLaunch(ws.exec, s));
        else launch_iteration();
        SG_HIP(hipMemcpyAsync(&ws.pinned[i % L], ws.any.p, sizeof(int), hipMemcpyDeviceToHost, s));
        SG_HIP(hipEventRecord(evs[i % L], s));
    }
    SG_HIP(hipMemcpyAsync(ws.h_st, ws.st.p, sizeof(BState) * nsys, hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    ctx->collect_profiles();
    const double elapsed = std::chrono::duration<double>(t1 - t0).count();

    std::vector<BState> fin(ws.h_st, ws.h_st + nsys);
    for (int q = 0; q < nsys; ++q) {
        const BState& h = fin[q];
        if (h.active || h.pend) fail(SG_E_STATE, "BiCGSTAB: iteration bound reached with a system still running");
        sg_solve_result& R = res[q];
        R.iterations = h.iters;
        R.converged = h.converged;
        R.residual_norm = h.res;
        R.cycles = 0;
        R.precond_applies = h.applies;
        R.hist_len = h.nhist;
        R.time_solve_s = elapsed;
        if (hist && h.nhist) {
            std::vector<double> hv(h.nhist);
            SG_HIP(hipMemcpy(hv.data(), W.hist + (size_t)q * hcap, sizeof(double) * h.nhist,
                             hipMemcpyDeviceToHost));
            for (int i = 0; i < h.nhist && i < hist_cap; ++i) hist[(size_t)q * hist_cap + i] = hv[i];
        }
        if (hist)
            for (int i = h.nhist; i < hist_cap; ++i) hist[(size_t)q * hist_cap + i] = 0.0;
    }
    // FP64 true residual at exit
    spmv_launch(A, -1.0, d_x, 1.0, d_b, r, Guard{}, s);
    hipLaunchKernelGGL(k_bi_norm, gv, bv, 0, s, sp, r, W);
    hipLaunchKernelGGL(k_bi_norm_finish, gf, bf, 0, s, sp, W);
    SG_HIP(hipMemcpyAsync(ws.h_st, ws.st.p, sizeof(BState) * nsys, hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < nsys; ++q) {
        const double rn = ws.h_st[q].nrm;
        res[q].rel_true_residual = bnorm[q] > 0.0 ? rn / bnorm[q] : rn;
    }
}

}  // namespace sg
