// The learned FGMRES stage: a base (two-level Schwarz) preconditioner plus a
// network correction of its remaining residual, kept only when it reduces the
// residual.  Every scalar stays on the device, so an application launches
// without host synchronisation (and can sit inside a captured solve):
//
//   z_b = base(x);  s = x - A z_b;  ss = s.s
//   sc  = c x c block averages of s (sc = s when c = 1); q = sc.sc
//   o   = net(sc / sqrt(q/cn), <log kappa>_c)        (cn = cells of the net grid)
//   e   = sqrt(q/cn) * P o                            (P: cell-centred bilinear
//                                                      interpolation, identity at c = 1)
//   a = s.(A e), cc = (A e).(A e)          (||s - alpha A e||^2 = ss - a^2 / cc)
//   z = z_b + (a / cc) e  if a^2 >= min_gain * cc * ss,  else  z = z_b
//
// The grid work outside the network is four kernels: the residual with its
// block averages and the partial sums of ss and q; the network input, written
// in bf16 straight into the network's input tensor (the block averages of
// log kappa are formed once at setup); A e with the
// partial sums of a and cc, e interpolated on the fly from the network output
// for every column the row touches (e is never stored); and the update, which
// interpolates e again.  Every reduction is per-block partials summed in one
// fixed order by each reader, so applications are reproducible bit for bit.
// Every kernel, the network's included, returns at once when the solver's
// step guard says the system has stopped.
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
    int nblk_r = 0, nblk_e = 0;      // blocks of the residual / A e kernels
    bool sell = false, tile = false; // SELL-64 rows / 64 x 16 tiles usable (set at setup)
    DevBuf<double> s, sc, scal;      // scal: ss, a, cc, q
    DevBuf<double> ae, azb;          // A e and A z_b, kept when the solver takes A z from the stage
    DevBuf<double> part_r, part_e;   // block partials {ss, q}, {a, cc}
    DevBuf<float> lk;                // c x c block averages of log kappa [cny][cnx]
    DevBuf<int> cnt;                 // applications, accepted corrections
};

// float -> bfloat16 bits, round to nearest even (NaN/Inf kept), as the
// network's own input conversion
__device__ inline unsigned bf16_bits(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return u >> 16;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

constexpr int kRT = 256;  // threads of the residual kernel: 64 columns x 4 net-grid rows
constexpr int kET = 256;  // threads of the A e and update kernels

// Deterministic block sum of two doubles (a, b) over kRT threads; thread 0
// gets the totals.
__device__ inline void block_sum2(double& a, double& b) {
    __shared__ double ra[kRT / kWave], rb[kRT / kWave];
    for (int o = kWave / 2; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        ra[w] = a;
        rb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = ra[0];
        b = rb[0];
        for (int k = 1; k < (int)(blockDim.x / kWave); ++k) {
            a += ra[k];
            b += rb[k];
        }
    }
}

// Sum of part[2 k + j] over k < nblk in one fixed order (every block that
// calls it gets the same bits); result valid in thread 0.
__device__ inline void sum_partials(const double* __restrict__ part, int nblk, double& s0,
                                    double& s1) {
    s0 = 0.0;
    s1 = 0.0;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
        s0 += part[2 * k];
        s1 += part[2 * k + 1];
    }
    block_sum2(s0, s1);
}

// sum_k a_ik u(col_k) of row i in CSR order.  SELL: from the SELL-64 copy of
// A (slice i / 64, entry k of lane i % 64 at sp[slice] + 64 k + lane, padding
// column -1; the same entries in the same order), eight entries in flight.
template <bool SELL, typename U>
__device__ inline double row_sum(int i, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const double* __restrict__ v, const int* __restrict__ sp,
                                 const int* __restrict__ sci, const double* __restrict__ sv,
                                 U&& u) {
    double acc = 0.0;
    if (SELL) {
        const int sl = i >> 6, lane = i & 63;
        const int p0 = sp[sl], width = (sp[sl + 1] - p0) >> 6;
        for (int k0 = 0; k0 < width; k0 += 8) {
            double a[8], xv[8];
            int c[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int idx = p0 + min(k0 + q, width - 1) * 64 + lane;
                a[q] = sv[idx];
                c[q] = sci[idx];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) xv[q] = (k0 + q < width && c[q] >= 0) ? u(c[q]) : 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (k0 + q < width && c[q] >= 0) acc += a[q] * xv[q];
        }
    } else {
        const int k0 = rp[i], k1 = rp[i + 1];
        for (int k = k0; k < k1; ++k) acc += v[k] * u(ci[k]);
    }
    return acc;
}

// s = x - A z for the rows of a 64-column x (4 C)-row tile, the C x C block
// averages sc of s, block partials {s.s, sc.sc}; azb (optional) keeps A z.
// Thread (tx, ty): fine column bx * 64 + tx, net-grid row by * 4 + ty (its C
// fine rows); the C lanes of one net-grid cell add their column sums with
// shuffles.
template <int C, bool SELL>
__global__ void __launch_bounds__(kRT)
k_stage_res(int nx, int ny, const int* __restrict__ rp, const int* __restrict__ ci,
            const double* __restrict__ v, const int* __restrict__ sp, const int* __restrict__ sci,
            const double* __restrict__ sv, const double* __restrict__ x,
            const double* __restrict__ z, double* __restrict__ s, double* __restrict__ sc,
            double* __restrict__ part, double* __restrict__ azb, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int X = blockIdx.x * 64 + tx, Y = blockIdx.y * 4 + ty;
    const int cnx = nx / C, cny = ny / C;
    const bool live = X < nx && Y < cny;
    double col = 0.0, ss = 0.0;
    if (live) {
#pragma unroll
        for (int dy = 0; dy < C; ++dy) {
            const int i = (Y * C + dy) * nx + X;
            const double acc = row_sum<SELL>(i, rp, ci, v, sp, sci, sv, [&](int j) { return z[j]; });
            if (azb) azb[i] = acc;  // the bits of the solver's own SpMV of z_b
            const double si = x[i] - acc;
            s[i] = si;
            ss += si * si;
            col += si;
        }
    }
#pragma unroll
    for (int o = 1; o < C; o <<= 1) col += __shfl_xor(col, o);
    double q = 0.0;
    if (live && (tx % C) == 0) {
        const double a = col * (1.0 / (C * C));
        sc[Y * cnx + X / C] = a;
        q = a * a;
    }
    block_sum2(ss, q);
    if (threadIdx.x == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        part[2 * b] = ss;
        part[2 * b + 1] = q;
    }
}

// The network input, bf16 NHWC [cny][cnx][16]: channel 0 = sc / sqrt(q / cn),
// channel 1 = the averaged log kappa (lk), the rest zero; block 0 also stores
// ss and q in scal[0], scal[3]
__global__ void __launch_bounds__(kRT)
k_stage_in(int cn, int nblk, const double* __restrict__ part, const double* __restrict__ sc,
           const float* __restrict__ lk, double* __restrict__ scal, uint16_t* __restrict__ x,
           DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    __shared__ double sh_r;
    double ss, q;
    sum_partials(part, nblk, ss, q);
    if (threadIdx.x == 0) {
        sh_r = q > 0.0 ? 1.0 / sqrt(q / cn) : 0.0;
        if (blockIdx.x == 0) {
            scal[0] = ss;
            scal[3] = q;
        }
    }
    __syncthreads();
    const double r = sh_r;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cn; i += gridDim.x * blockDim.x) {
        const unsigned c0 = bf16_bits((float)(sc[i] * r)), c1 = bf16_bits(lk[i]);
        uint4* o = reinterpret_cast<uint4*>(x + (size_t)i * 16);
        o[0] = uint4{c0 | (c1 << 16), 0u, 0u, 0u};
        o[1] = uint4{0u, 0u, 0u, 0u};
    }
}

// P o at fine cell j: cell-centred bilinear interpolation from the cnx x cny
// net grid, clamped at the edges (C = 1: o itself).  Cell x of the grid sits
// at (x + 1/2) / C - 1/2 = (2x + 1 - C) / (2C) on the net grid: its floor and
// fraction come from an integer shift and mask (exact, the same values as the
// floating-point form); the row of j from a reciprocal with a one-step fix.
template <int C>
__device__ inline double interp_xy(const float* __restrict__ o, int x, int y, int cnx, int cny) {
    constexpr int kSh = C == 2 ? 2 : 3;  // log2(2C)
    const int tx2 = 2 * x + 1 - C, ty2 = 2 * y + 1 - C;
    int x0 = tx2 >> kSh, y0 = ty2 >> kSh;  // arithmetic shifts: floor
    const double tx = (double)(tx2 & (2 * C - 1)) * (1.0 / (2 * C));
    const double ty = (double)(ty2 & (2 * C - 1)) * (1.0 / (2 * C));
    const int x1 = min(max(x0 + 1, 0), cnx - 1), y1 = min(max(y0 + 1, 0), cny - 1);
    x0 = min(max(x0, 0), cnx - 1);
    y0 = min(max(y0, 0), cny - 1);
    return (1.0 - ty) * ((1.0 - tx) * o[y0 * cnx + x0] + tx * o[y0 * cnx + x1]) +
           ty * ((1.0 - tx) * o[y1 * cnx + x0] + tx * o[y1 * cnx + x1]);
}

// grid coordinates of cell j: the row from a reciprocal with a one-step fix
__device__ inline void cell_xy(int j, int nx, double rnx, int& x, int& y) {
    y = (int)((double)j * rnx);
    x = j - y * nx;
    if (x < 0) {
        --y;
        x += nx;
    } else if (x >= nx) {
        ++y;
        x -= nx;
    }
}

template <int C>
__device__ inline double interp(const float* __restrict__ o, int j, int nx, double rnx, int cnx,
                                int cny) {
    if (C == 1) return (double)o[j];
    int x, y;
    cell_xy(j, nx, rnx, x, y);
    return interp_xy<C>(o, x, y, cnx, cny);
}

// block partials {s.(A e), (A e).(A e)} with e = sqrt(q/cn) P o, row i per thread
template <int C, bool SELL>
__global__ void __launch_bounds__(kET)
k_stage_ae(int nx, int ny, const int* __restrict__ rp, const int* __restrict__ ci,
           const double* __restrict__ v, const int* __restrict__ sp, const int* __restrict__ sci,
           const double* __restrict__ sv, const float* __restrict__ o,
           const double* __restrict__ scal, const double* __restrict__ s,
           double* __restrict__ part, double* __restrict__ ae_out, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    const int n = nx * ny, cnx = nx / C, cny = ny / C;
    const double r = sqrt(scal[3] / ((double)cnx * cny)), rnx = 1.0 / nx;
    double a = 0.0, c = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ae = row_sum<SELL>(i, rp, ci, v, sp, sci, sv, [&](int j) {
            return r * interp<C>(o, j, nx, rnx, cnx, cny);
        });
        if (ae_out) ae_out[i] = ae;
        a += s[i] * ae;
        c += ae * ae;
    }
    block_sum2(a, c);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = c;
    }
}

// The same partials for a grid whose SELL-64 slices are aligned 64-column
// rows (nx a multiple of 64, ny of 16): a workgroup takes a 64 x 16 tile,
// first interpolates e on the tile and its one-cell halo into LDS (once per
// cell instead of once per matrix entry), then forms A e of its rows from
// there (columns outside the halo, which a 5-point stencil never has, are
// interpolated directly).  Each row's sum has the same terms in the same
// order as k_stage_ae's.
constexpr int kAX = 64, kAY = 16;
template <int C>
__global__ void __launch_bounds__(kET)
k_stage_ae_tile(int nx, int ny, const int* __restrict__ sp, const int* __restrict__ sci,
                const double* __restrict__ sv, const float* __restrict__ o,
                const double* __restrict__ scal, const double* __restrict__ s,
                double* __restrict__ part, double* __restrict__ ae_out, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    constexpr int kLX = kAX + 2, kLY = kAY + 2;
    __shared__ double et[kLY * kLX];
    const int cnx = nx / C, cny = ny / C;
    const double r = sqrt(scal[3] / ((double)cnx * cny)), rnx = 1.0 / nx;
    const int X0 = blockIdx.x * kAX, Y0 = blockIdx.y * kAY;
    for (int e = threadIdx.x; e < kLY * kLX; e += blockDim.x) {
        const int gx = X0 - 1 + e % kLX, gy = Y0 - 1 + e / kLX;
        double v = 0.0;
        if (gx >= 0 && gx < nx && gy >= 0 && gy < ny)
            v = r * (C == 1 ? (double)o[gy * nx + gx] : interp_xy<C>(o, gx, gy, cnx, cny));
        et[e] = v;
    }
    __syncthreads();
    const int tx = threadIdx.x & 63, w = threadIdx.x >> 6;
    double a = 0.0, c = 0.0;
    for (int k = 0; k < kAY / 4; ++k) {
        const int ly = w + 4 * k, gx = X0 + tx, gy = Y0 + ly;
        const int i = gy * nx + gx;
        const double ae = row_sum<true>(i, nullptr, nullptr, nullptr, sp, sci, sv, [&](int j) {
            const int d = j - i;
            int jx, jy;
            if (d == 0) {
                jx = gx, jy = gy;
            } else if (d == 1 && gx + 1 < nx) {
                jx = gx + 1, jy = gy;
            } else if (d == -1 && gx > 0) {
                jx = gx - 1, jy = gy;
            } else if (d == nx) {
                jx = gx, jy = gy + 1;
            } else if (d == -nx) {
                jx = gx, jy = gy - 1;
            } else {
                cell_xy(j, nx, rnx, jx, jy);
            }
            const int lx = jx - X0 + 1, lyy = jy - Y0 + 1;
            if (lx >= 0 && lx < kLX && lyy >= 0 && lyy < kLY) return et[lyy * kLX + lx];
            return r * interp<C>(o, j, nx, rnx, cnx, cny);
        });
        if (ae_out) ae_out[i] = ae;
        a += s[i] * ae;
        c += ae * ae;
    }
    block_sum2(a, c);
    if (threadIdx.x == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        part[2 * b] = a;
        part[2 * b + 1] = c;
    }
}

// the acceptance test and z = z_b + (a / cc) e; with az, also the solver's next
// product A z: A z_b + (a / cc) A e by linearity from the stored products, or
// A z_b itself (the bits of an SpMV of z_b) when the correction is refused
template <int C>
__global__ void __launch_bounds__(kET)
k_stage_final(int nx, int ny, int nblk, const double* __restrict__ part,
              double* __restrict__ scal, const float* __restrict__ o, double* __restrict__ y,
              double min_gain, int* __restrict__ cnt, const double* __restrict__ azb,
              const double* __restrict__ ae, double* __restrict__ az, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    __shared__ double sh_al;
    __shared__ int sh_ok;
    double a, c;
    sum_partials(part, nblk, a, c);
    if (threadIdx.x == 0) {
        const double ss = scal[0];
        const bool ok = c > 0.0 && ss > 0.0 && a * a >= min_gain * c * ss;
        sh_ok = ok ? 1 : 0;
        sh_al = ok ? a / c : 0.0;
        if (blockIdx.x == 0) {
            scal[1] = a;
            scal[2] = c;
            cnt[0] += 1;
            cnt[1] += ok ? 1 : 0;
        }
    }
    __syncthreads();
    const int n = nx * ny, cnx = nx / C, cny = ny / C;
    if (az) {
        const double aa = sh_al;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            az[i] = sh_ok ? azb[i] + aa * ae[i] : azb[i];
    }
    if (!sh_ok) return;
    const double al = sh_al * sqrt(scal[3] / ((double)cnx * cny)), rnx = 1.0 / nx;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] += al * interp<C>(o, i, nx, rnx, cnx, cny);
}

__global__ void k_count_base(int* __restrict__ cnt, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    cnt[0] += 1;
}

// blocks of the A e kernel (each leaves a partial pair) and of the update
// kernel (each block adds all of those partials itself)
static int grid_ae(int n) { return std::max(1, std::min((n + kET - 1) / kET, 1024)); }
static int grid_final(int n) { return std::max(1, std::min((n + kET - 1) / kET, 1024)); }

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
    // the network's second input channel: c x c block averages of log kappa
    std::vector<float> lk_h((size_t)L->cn, 0.f);
    for (int J = 0; J < L->cny; ++J)
        for (int I = 0; I < L->cnx; ++I) {
            double a = 0.0;
            for (int dy = 0; dy < c; ++dy)
                for (int dx = 0; dx < c; ++dx) a += log_kappa[(size_t)(J * c + dy) * nx + I * c + dx];
            lk_h[(size_t)J * L->cnx + I] = (float)(a / ((double)c * c));
        }
    L->lk.upload(lk_h, ctx->stream);
    L->nblk_r = ((nx + 63) / 64) * ((L->cny + 3) / 4);
    L->s.alloc(L->n);
    L->ae.alloc(L->n);
    L->azb.alloc(L->n);
    L->sc.alloc(L->cn);
    L->scal.alloc(4);
    L->part_r.alloc((size_t)2 * L->nblk_r);
    // enough for either A e kernel (the one used is chosen at setup)
    L->part_e.alloc((size_t)2 * std::max(grid_ae(L->n), (nx / kAX) * (ny / kAY)));
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
    // the SELL-64 copy serves whole wavefronts of 64 grid columns when its
    // slices are the aligned 64-row blocks of one system
    L->sell = A->nsl > 0 && A->sys_off.size() == 2 && L->nx % 64 == 0;
    L->tile = L->sell && L->ny % kAY == 0;
    L->nblk_e = L->tile ? (L->nx / kAX) * (L->ny / kAY) : grid_ae(L->n);
    learned_prepare(L);
}

void learned_prepare(LearnedStage* L) {
    precond_ensure_tuned(L->base);
    unet_prepare(L->net, 1, L->cny, L->cnx);
}

template <int C, bool SELL>
static void stage_launch(LearnedStage* L, const double* x, double* y, double* az,
                         const DevGuard& dg) {
    double* ae = az ? L->ae.p : nullptr;
    double* azb = az ? L->azb.p : nullptr;
    hipStream_t st = L->ctx->stream;
    const sg_csr* A = L->A;
    hipLaunchKernelGGL((k_stage_res<C, SELL>), dim3((L->nx + 63) / 64, (L->cny + 3) / 4), dim3(kRT),
                       0, st, L->nx, L->ny, A->rp.p, A->ci.p, A->v.p, A->sell_ptr.p, A->sell_ci.p,
                       A->sell_v.p, x, y, L->s.p, L->sc.p, L->part_r.p, azb, dg);
    hipLaunchKernelGGL(k_stage_in, dim3(std::max(1, std::min((L->cn + kRT - 1) / kRT, 256))),
                       dim3(kRT), 0, st, L->cn, L->nblk_r, L->part_r.p, L->sc.p, L->lk.p,
                       L->scal.p, unet_input_buffer(L->net, 1, L->cny, L->cnx), dg);
    unet_forward_dev(L->net, 1, L->cny, L->cnx, nullptr, false, dg.flag);
    if (SELL && L->tile)
        hipLaunchKernelGGL((k_stage_ae_tile<C>), dim3(L->nx / kAX, L->ny / kAY), dim3(kET), 0, st,
                           L->nx, L->ny, A->sell_ptr.p, A->sell_ci.p, A->sell_v.p, unet_output(L->net),
                           L->scal.p, L->s.p, L->part_e.p, ae, dg);
    else
        hipLaunchKernelGGL((k_stage_ae<C, SELL>), dim3(L->nblk_e), dim3(kET), 0, st, L->nx, L->ny,
                           A->rp.p, A->ci.p, A->v.p, A->sell_ptr.p, A->sell_ci.p, A->sell_v.p,
                           unet_output(L->net), L->scal.p, L->s.p, L->part_e.p, ae, dg);
    hipLaunchKernelGGL(k_stage_final<C>, dim3(grid_final(L->n)), dim3(kET), 0, st, L->nx, L->ny,
                       L->nblk_e, L->part_e.p, L->scal.p, unet_output(L->net), y, L->min_gain,
                       L->cnt.p, azb, ae, az, dg);
}

// One application; with az (and A the stage's own matrix) it also writes
// az = A z from the stage's products and returns true, false when the caller
// must form A z itself (base-only applications, another matrix, no az).
bool learned_apply_az(LearnedStage* L, const sg_csr* A, const double* x, double* y, double* az,
                      const Guard& g) {
    hipStream_t st = L->ctx->stream;
    precond_apply_launch(L->base, x, y, g);
    const DevGuard dg = dev_guard(g, nullptr);
    const bool run_net = (L->calls_host++ % L->every) == 0;
    if (!run_net) {
        hipLaunchKernelGGL(k_count_base, dim3(1), dim3(64), 0, st, L->cnt.p, dg);
        SG_HIP(hipGetLastError());
        return false;
    }
    double* w = (az && A == L->A) ? az : nullptr;
    const bool sell = L->sell;
    if (L->coarsen == 4)
        sell ? stage_launch<4, true>(L, x, y, w, dg) : stage_launch<4, false>(L, x, y, w, dg);
    else if (L->coarsen == 2)
        sell ? stage_launch<2, true>(L, x, y, w, dg) : stage_launch<2, false>(L, x, y, w, dg);
    else
        sell ? stage_launch<1, true>(L, x, y, w, dg) : stage_launch<1, false>(L, x, y, w, dg);
    SG_HIP(hipGetLastError());
    return w != nullptr;
}

void learned_apply(LearnedStage* L, const double* x, double* y, const Guard& g) {
    learned_apply_az(L, nullptr, x, y, nullptr, g);
}

void learned_counts(const LearnedStage* L, long long* calls, long long* accepted) {
    int c[2] = {0, 0};
    SG_HIP(hipMemcpyAsync(c, L->cnt.p, sizeof(c), hipMemcpyDeviceToHost, L->ctx->stream));
    SG_HIP(hipStreamSynchronize(L->ctx->stream));
    if (calls) *calls = c[0];
    if (accepted) *accepted = c[1];
}

}  // namespace sg
