// Restarted GMRES(m) (left preconditioning) and FGMRES(m) (flexible right
// preconditioning) with every scalar on the device, for one system or for a
// batch of independent systems stored as one block-diagonal matrix.
//
// Semantics follow src/core/gmres.cpp:34-160 and src/core/fgmres.cpp:28-172:
// absolute test on beta at each restart and on |g_{j+1}| after each Arnoldi
// step, breakdown when h_{j+1,j} < 1e-30 (no rotation for that column),
// tangent-form Givens rotations (gmres.cpp:11-23), back substitution and the
// x update in the reference's order, max_iter bounding cycles with the total
// step count checked after each cycle.  Gram-Schmidt runs either as MGS in the
// reference's order or as a two-pass CGS2 whose re-orthogonalisation
// coefficients come from the Gram matrix.
//
// Reductions: every vector kernel runs over a table of blocks, each block one
// contiguous range inside one system; blocks write per-block partial sums and
// a "finish" kernel (one workgroup per system) adds them in a fixed order and
// runs that system's scalar epilogue.  No grid-wide atomics or fences.  A
// system's ranges are the same whether it is solved alone or in a batch, so
// its iterates are bitwise identical either way.
//
// Control: each system has a device state (KState) whose `active` flag every
// kernel of a step checks, so a system that converged or broke down drops out
// of the rest of the cycle.  The host reads the states back once per cycle and
// decides per system (next cycle, true-residual verification, done).  A cycle
// launches either directly (with `lookahead`: the host reads the flag of one
// system asynchronously and stops launching) or as one hipGraph captured
// before the first cycle and replayed (default for batches).
#include "segs.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

namespace sg {

sg_context* precond_ctx(const sg_precond* M);
int precond_n(const sg_precond* M);
bool precond_ready(const sg_precond* M);

constexpr int kT = 256;                 // threads per reduction block
constexpr int kMaxCols = 128;           // restart + 1
constexpr int kRedCols = 2 * kMaxCols;  // widest reduction (two dots per basis vector)
constexpr int kEo = 2;                  // elements per thread per tile
constexpr int kTile = kT * kEo;
constexpr int kCBG = 8;                 // basis vectors per register block of the Gram passes
constexpr int kEoG = 4;                 // elements per thread per tile of the Gram passes
constexpr int kTGD = 256;               // threads per block, Gram dots
constexpr int kTGU = 512;               // threads per block, Gram update
constexpr int kTileGD = kTGD * kEoG;
constexpr int kTileGU = kTGU * kEoG;
constexpr int kTF = 1024;               // threads of the finish kernels

// Column stride of the Krylov bases: n rounded up to 64 elements, plus 2112
// elements (16.5 KiB) for long vectors.  The Gram passes stream every column
// at the same element offsets; with strides that are multiples of a large
// power of two (n = 4096000 gives 31.25 x 1 MiB) those streams fall on the
// same HBM channels, and the pad spreads them (measured: 3.5% faster Gram
// passes on 4.1M-row systems).
static size_t basis_ld(int n) {
    const size_t ld = ((size_t)n + 63) / 64 * 64;
    return ld >= 65536 ? ld + 2112 : ld;
}

struct KState {
    int active;     // current cycle continues
    int k;          // Arnoldi columns completed in this cycle
    int converged;
    int total;      // Arnoldi steps over all cycles (GMRESResult::iterations)
    int est_conv;   // rtol mode: the estimate met the internal tolerance
    int first;      // no restart done yet
    int method;     // 0 left GMRES, 1 FGMRES
    int m;
    int has_m;      // a preconditioner is applied
    int applies;    // preconditioner applications so far
    double tol;     // test threshold on beta and |g|
    double rtol;    // > 0: true-residual mode
    double bnorm;
    double rel_true;
    double scale;   // divisor (GMRES) or multiplier (FGMRES) for the next basis vector
    double res;
};

// Per-system control written by the host before each cycle.
enum : int { kRun = 0, kVerify = 1, kDone = 2 };

struct Work {
    double* H;      // (m+1) x m, column-major H[i + j*(m+1)]
    double* Hr;     // the same columns before the Givens rotations (H-bar of the cycle)
    double* g;      // m+1
    double* cs;     // m
    double* sn;     // m
    double* y;      // m
    double* h1;     // kRedCols: projection coefficients h
    double* h2;     // kRedCols: reduced dots
    double* G;      // kMaxCols^2: Gram matrix V^T V, row-major, stride kMaxCols
    double* hist;   // residual history (hcap)
    double* part;   // block partial sums, part[c * pstride + block]
    KState* st;
    const int* mode;
    int* hflag;     // single system, direct launches: host-mapped active flag per step
    int m, hcap, pstride;
};

// The view of system s (pointers advanced to its slices; part is shared).
__device__ inline Work sys_work(const Work& W, int s) {
    Work o = W;
    const size_t m = W.m;
    o.H += (size_t)s * (m + 1) * m;
    o.Hr += (size_t)s * (m + 1) * m;
    o.g += (size_t)s * (m + 1);
    o.cs += (size_t)s * m;
    o.sn += (size_t)s * m;
    o.y += (size_t)s * m;
    o.h1 += (size_t)s * kRedCols;
    o.h2 += (size_t)s * kRedCols;
    o.G += (size_t)s * kMaxCols * kMaxCols;
    o.hist += (size_t)s * W.hcap;
    o.st += s;
    o.mode += s;
    return o;
}

// Givens rotation, tangent form (src/core/gmres.cpp:11-23, fgmres.cpp:12-20).
__device__ inline void givens_compute(double a, double b, double& c, double& s) {
    if (fabs(b) < 1e-30) {
        c = 1.0;
        s = 0.0;
    } else if (fabs(b) > fabs(a)) {
        double t = -a / b;
        s = 1.0 / sqrt(1.0 + t * t);
        c = s * t;
    } else {
        double t = -b / a;
        c = 1.0 / sqrt(1.0 + t * t);
        s = c * t;
    }
}

__device__ inline void givens_apply(double c, double s, double& h1, double& h2) {
    double t1 = c * h1 - s * h2;
    double t2 = s * h1 + c * h2;
    h1 = t1;
    h2 = t2;
}

// End of Arnoldi step j once h_{j+1,j} is known (gmres.cpp:99-127); W is one
// system's view.
// Step epilogue of one system (H[j+1, j] = ||w||, the rotations of column j,
// the new rotation, the residual estimate), run by a whole block: column j of
// H, the rotations so far, g[j..j+1] and the state are loaded together, then
// thread 0 rotates in LDS and registers -- a chain of LDS accesses, not of
// global round trips (0.33 us per column at j ~ 40 otherwise).  The
// operations and their order are those of one thread on global memory, so
// the values are the same bit for bit.  Every thread of the block calls it;
// threads other than 0 return after the loads.
__device__ void step_epilogue(const Work& W, int j, double hn) {
    __shared__ double eh[kMaxCols + 1], ec[kMaxCols], es[kMaxCols];
    const int m = W.m;
    double* H = W.H + (size_t)j * (m + 1);
    for (int i = threadIdx.x; i <= j; i += blockDim.x) eh[i] = H[i];
    for (int i = threadIdx.x; i < j; i += blockDim.x) {
        ec[i] = W.cs[i];
        es[i] = W.sn[i];
    }
    KState k;
    double g0 = 0.0, g1 = 0.0;
    if (threadIdx.x == 0) {
        k = *W.st;
        g0 = W.g[j];
        g1 = W.g[j + 1];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    KState* st = W.st;
    st->applies = k.applies + k.has_m;
    H[j + 1] = hn;
    eh[j + 1] = hn;
    for (int i = 0; i <= j + 1; ++i) W.Hr[i + (size_t)j * (m + 1)] = eh[i];
    if (hn < 1e-30) {  // breakdown: column j is used unrotated
        st->k = j + 1;
        st->active = 0;
        return;
    }
    st->scale = k.method == 0 ? hn : 1.0 / hn;
    for (int i = 0; i < j; ++i) {
        double a = eh[i], b = eh[i + 1];
        givens_apply(ec[i], es[i], a, b);
        eh[i] = a;
        eh[i + 1] = b;
    }
    double c, sn;
    double hj = eh[j], hj1 = eh[j + 1];
    givens_compute(hj, hj1, c, sn);
    W.cs[j] = c;
    W.sn[j] = sn;
    givens_apply(c, sn, hj, hj1);
    eh[j] = hj;
    eh[j + 1] = hj1;
    givens_apply(c, sn, g0, g1);
    for (int i = 0; i <= j + 1; ++i) H[i] = eh[i];
    W.g[j] = g0;
    W.g[j + 1] = g1;
    const double res = fabs(g1);
    W.hist[k.total] = res;
    st->total = k.total + 1;
    st->res = res;
    st->k = j + 1;
    if (res < k.tol) {
        if (k.rtol > 0.0) st->est_conv = 1;
        else st->converged = 1;
        st->active = 0;
    }
}

// Adds the per-wave values red[wave][c] and stores column c's block partial.
template <int NT = kT>
__device__ inline void store_partials(double (*red)[kRedCols], int ncols, double* part,
                                      int pstride) {
    __syncthreads();
    for (int i = threadIdx.x; i < ncols; i += NT) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) s += red[w][i];
        part[(size_t)i * pstride + blockIdx.x] = s;
    }
}

// out[c] = sum over a system's blocks of part[c][blk0 + b].  One wavefront per
// column: each lane loads its partials (16 loads in flight), adds them in
// order, then across lanes.  Ends with __syncthreads.
__device__ void reduce_partials(const Work& W, const SysBlocks& sb, int ncols, double* out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int i = wv; i < ncols; i += nw) {
        const double* pi = W.part + (size_t)i * W.pstride + sb.blk0;
        double s = 0.0;
        for (int b0 = 0; b0 < sb.nblk; b0 += 16 * 64) {
            double v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int b = b0 + q * 64 + lane;
                v[q] = b < sb.nblk ? pi[b] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) s += v[q];
        }
        s = wave_sum(s);
        if (lane == 0) out[i] = s;
    }
    __syncthreads();
}

// ---- CGS2 with Gram-matrix re-orthogonalisation ----------------------------
// Pass 1 (k_gram_dot + k_gram_finish): p = V^T w and, lagged by one step,
// g = V^T v_j -- the actual Gram row of the newest basis vector -- in one
// read of V; then h = (2I - G) p, which equals G^{-1} p to first order in the
// loss of orthogonality, i.e. the coefficients of projecting twice.
// Pass 2 (k_gram_update + k_norm_finish): w -= V h, ||w||, step epilogue.
// Basis vectors go in register blocks of kCBG, the remainder in blocks of 4, 2
// and 1, so no load is spent on a column beyond ncols.  Elements beyond the
// block are clamped to its last element and masked; every load of a tile is
// issued before the scheduling barrier, so a wave keeps a whole register
// block of loads in flight.

// Pass 1 stages a block's range of w and v_j (up to kPassGD elements) in LDS
// once and streams the other basis columns past it, so w and v_j are read
// from memory once per step and column j is not read a second time:
// 8 n (j + 2) bytes.  Each thread adds its products tile by tile (kEoG
// elements of a kTileGD tile per thread), the order a tile loop over the range
// gives, whatever the column blocking.
constexpr int kTilesGD = 2;
constexpr int kPassGD = kTileGD * kTilesGD;

// Dots of columns [cb, cb + CB) with w and v_j over the staged range; lane 0
// of each wave stores (first pass) or adds its wave's sums into redw.
template <int CB>
__device__ __forceinline__ void gram_dot_cols(int cb, int ncols, int p0, int p1, bool first,
                                              const double* __restrict__ V, size_t ld,
                                              const double* sw, const double* sv, double* redw) {
    const int tid = threadIdx.x;
    double aw[CB], av[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) aw[c] = av[c] = 0.0;
    for (int base = p0 + tid; base < p1; base += kTileGD) {
        int ee[kEoG];
        double a[CB][kEoG];
#pragma unroll
        for (int r = 0; r < kEoG; ++r) ee[r] = min(base + r * kTGD, p1 - 1);
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int r = 0; r < kEoG; ++r)
                a[c][r] = __builtin_nontemporal_load(&V[(size_t)(cb + c) * ld + ee[r]]);
        double wr[kEoG], vr[kEoG];
#pragma unroll
        for (int r = 0; r < kEoG; ++r) {
            wr[r] = sw[base + r * kTGD - p0];
            vr[r] = sv[base + r * kTGD - p0];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < CB; ++c) {
#pragma unroll
            for (int r = 0; r < kEoG; ++r) {
                aw[c] += a[c][r] * wr[r];
                av[c] += a[c][r] * vr[r];
            }
        }
    }
    const int lane = tid & 63;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        const double sw_ = wave_sum(aw[c]);
        const double sv_ = wave_sum(av[c]);
        if (lane == 0) {
            redw[cb + c] = first ? sw_ : redw[cb + c] + sw_;
            redw[ncols + cb + c] = first ? sv_ : redw[ncols + cb + c] + sv_;
        }
    }
}

__global__ void __launch_bounds__(kTGD)
k_gram_dot(Segs sp, int ncols, const double* __restrict__ V, size_t ld,
           const double* __restrict__ w, Work W, int j) {
    const BlockRange br = sp.blk[blockIdx.x];
    if (!W.st[br.sys].active) return;
    __shared__ double red[kTGD / 64][kRedCols];
    // slots past the range (up to the end of its last tile) hold a clamped
    // copy scaled by zero, as a multiply by a 0/1 mask would leave them
    __shared__ double sw[kPassGD], sv[kPassGD];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double* redw = red[wv];  // only lane 0 of wave wv touches it
    const double* vj = V + (size_t)j * ld;
    const int nv = ncols - 1;  // columns 0..j-1 are streamed; column j is v_j itself
    for (int p0 = br.start; p0 < br.end; p0 += kPassGD) {
        const int p1 = min(br.end, p0 + kPassGD);
        const bool first = p0 == br.start;
        if (!first) __syncthreads();  // the previous pass is done with sw / sv
        const int ntile = (p1 - p0 + kTileGD - 1) / kTileGD;
        for (int t = 0; t < ntile; ++t) {
            double a[kEoG], b[kEoG];
            int ee[kEoG];
#pragma unroll
            for (int r = 0; r < kEoG; ++r) {
                ee[r] = min(p0 + t * kTileGD + r * kTGD + tid, p1 - 1);
                a[r] = w[ee[r]];
                b[r] = vj[ee[r]];
            }
#pragma unroll
            for (int r = 0; r < kEoG; ++r) {
                const int e = p0 + t * kTileGD + r * kTGD + tid;
                sw[e - p0] = e < p1 ? a[r] : a[r] * 0.0;
                sv[e - p0] = e < p1 ? b[r] : b[r] * 0.0;
            }
        }
        __syncthreads();
        int cb = 0;
        for (; cb + kCBG <= nv; cb += kCBG)
            gram_dot_cols<kCBG>(cb, ncols, p0, p1, first, V, ld, sw, sv, redw);
        if (nv - cb >= 4) {
            gram_dot_cols<4>(cb, ncols, p0, p1, first, V, ld, sw, sv, redw);
            cb += 4;
        }
        if (nv - cb >= 2) {
            gram_dot_cols<2>(cb, ncols, p0, p1, first, V, ld, sw, sv, redw);
            cb += 2;
        }
        if (nv - cb >= 1) gram_dot_cols<1>(cb, ncols, p0, p1, first, V, ld, sw, sv, redw);
        // column j: v_j . w and v_j . v_j from the staged values, in the same
        // per-thread order
        double s1 = 0.0, s2 = 0.0;
        for (int base = p0 + tid; base < p1; base += kTileGD) {
#pragma unroll
            for (int r = 0; r < kEoG; ++r) {
                const double wv_ = sw[base + r * kTGD - p0], vv = sv[base + r * kTGD - p0];
                s1 += vv * wv_;
                s2 += vv * vv;
            }
        }
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        if (lane == 0) {
            redw[j] = first ? s1 : redw[j] + s1;
            redw[ncols + j] = first ? s2 : redw[ncols + j] + s2;
        }
    }
    // columns [0, ncols) hold V^T w, [ncols, 2 ncols) hold V^T v_j
    store_partials<kTGD>(red, 2 * ncols, W.part, W.pstride);
}

__global__ void __launch_bounds__(kTF) k_gram_finish(Work W0, Segs sp, int ncols, int j) {
    const Work W = sys_work(W0, blockIdx.x);
    if (!W.st->active) return;
    reduce_partials(W, sp.sys[blockIdx.x], 2 * ncols, W.h2);
    const int m = W.st->m;
    double* Gm = W.G;
    for (int i = threadIdx.x; i < ncols; i += kTF) {  // Gram row of v_j
        const double g = W.h2[ncols + i];
        Gm[j * kMaxCols + i] = g;
        Gm[i * kMaxCols + j] = g;
    }
    __syncthreads();
    // h = 2p - G p, one wavefront per row
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = wv; i < ncols; i += kTF / 64) {
        double gp = 0.0;
        for (int k = lane; k < ncols; k += 64) gp += Gm[i * kMaxCols + k] * W.h2[k];
        gp = wave_sum(gp);
        if (lane == 0) {
            const double h = 2.0 * W.h2[i] - gp;
            W.h1[i] = h;
            W.H[i + j * (m + 1)] = h;
        }
    }
}

// Column c of system blockIdx.y: the sum of its block partials in block order
// (the order of reduce_partials), into h2[c]; columns ncols.. are the Gram row
// of v_j and also go into G.  One wavefront per column, so the reduction is
// spread over 2 ncols workgroups instead of one.
__global__ void __launch_bounds__(64) k_gram_sums(Work W0, Segs sp, int ncols, int j) {
    const Work W = sys_work(W0, blockIdx.y);
    if (!W.st->active) return;
    const SysBlocks sb = sp.sys[blockIdx.y];
    const int c = blockIdx.x, lane = threadIdx.x;
    const double* pi = W.part + (size_t)c * W.pstride + sb.blk0;
    double s = 0.0;
    for (int b0 = 0; b0 < sb.nblk; b0 += 16 * 64) {
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int b = b0 + q * 64 + lane;
            v[q] = b < sb.nblk ? pi[b] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) s += v[q];
    }
    s = wave_sum(s);
    if (lane == 0) {
        W.h2[c] = s;
        if (c >= ncols) {
            const int i = c - ncols;
            W.G[j * kMaxCols + i] = s;
            W.G[i * kMaxCols + j] = s;
        }
    }
}

// wr -= V[:, cb:cb+CB] h[cb:cb+CB] on the thread's elements ee.
template <int CB>
__device__ __forceinline__ void gram_update_cols(const double* __restrict__ V, size_t ld, int cb,
                                                 const int* ee, double* wr, const double* hs) {
    double a[CB][kEoG];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int r = 0; r < kEoG; ++r) a[c][r] = __builtin_nontemporal_load(&V[(size_t)(cb + c) * ld + ee[r]]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        const double hi = hs[cb + c];
#pragma unroll
        for (int r = 0; r < kEoG; ++r) wr[r] -= hi * a[c][r];
    }
}

// h = (2I - G) p is formed by every block from the sums of k_gram_sums (one
// wavefront per row, lanes over the columns, as k_gram_finish does); the
// system's first block also stores it as column j of H.
__global__ void __launch_bounds__(kTGU)
k_gram_update(Segs sp, int ncols, const double* __restrict__ V, size_t ld,
              double* __restrict__ w, Work W0, int j) {
    const BlockRange br = sp.blk[blockIdx.x];
    if (!W0.st[br.sys].active) return;
    const Work W = sys_work(W0, br.sys);
    __shared__ double red[kTGU / 64][kRedCols];
    __shared__ double hs[kMaxCols];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const bool first = (int)blockIdx.x == sp.sys[br.sys].blk0;
    const int m = W.m;
    for (int i = wv; i < ncols; i += kTGU / 64) {
        double gp = 0.0;
        for (int k = lane; k < ncols; k += 64) gp += W.G[i * kMaxCols + k] * W.h2[k];
        gp = wave_sum(gp);
        if (lane == 0) {
            const double h = 2.0 * W.h2[i] - gp;
            hs[i] = h;
            if (first) W.H[i + j * (m + 1)] = h;
        }
    }
    __syncthreads();
    const int start = br.start, end = br.end;
    const int nfull = ncols / kCBG * kCBG;
    double acc = 0.0;
    for (int base = start + tid; base < end; base += kTileGU) {
        int ee[kEoG];
        double wr[kEoG];
#pragma unroll
        for (int r = 0; r < kEoG; ++r) {
            ee[r] = min(base + r * kTGU, end - 1);
            wr[r] = w[ee[r]];
        }
        int cb = 0;
        for (; cb < nfull; cb += kCBG) gram_update_cols<kCBG>(V, ld, cb, ee, wr, hs);
        if (ncols - cb >= 4) {
            gram_update_cols<4>(V, ld, cb, ee, wr, hs);
            cb += 4;
        }
        if (ncols - cb >= 2) {
            gram_update_cols<2>(V, ld, cb, ee, wr, hs);
            cb += 2;
        }
        if (ncols - cb >= 1) gram_update_cols<1>(V, ld, cb, ee, wr, hs);
#pragma unroll
        for (int r = 0; r < kEoG; ++r) {
            if (base + r * kTGU < end) {
                w[ee[r]] = wr[r];
                acc += wr[r] * wr[r];
            }
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) red[wv][0] = acc;
    store_partials<kTGU>(red, 1, W.part, W.pstride);
}

__global__ void __launch_bounds__(kTF) k_norm_finish(Work W0, Segs sp, int j) {
    const Work W = sys_work(W0, blockIdx.x);
    // the host reads hflag[j] after the step's event, whether or not the
    // system was still active
    if (!W.st->active) {
        if (W.hflag && threadIdx.x == 0) W.hflag[j] = 0;
        return;
    }
    double* out = W.h2 + kRedCols - 1;
    reduce_partials(W, sp.sys[blockIdx.x], 1, out);
    step_epilogue(W, j, sqrt(*out));
    if (threadIdx.x == 0 && W.hflag) W.hflag[j] = W.st->active;
}

// ||w|| from the update pass's partials, the step epilogue (system's first
// block) and v_{j+1} = w scaled, in one kernel: every block reduces the same
// partials in the same order (reduce_partials), so all of them scale by the
// same value, the one the epilogue records.  Replaces k_norm_finish followed
// by k_scale_basis for the steps that have a next basis vector.
__global__ void __launch_bounds__(kT)
k_norm_scale(Segs sp, Segs sp_norm, const double* __restrict__ u, double* __restrict__ v,
             Work W0, int j) {
    const BlockRange br = sp.blk[blockIdx.x];
    const Work W = sys_work(W0, br.sys);
    const bool first = (int)blockIdx.x == sp.sys[br.sys].blk0;
    if (!W.st->active) {
        if (first && W.hflag && threadIdx.x == 0) W.hflag[j] = 0;
        return;
    }
    // this block's first elements are loaded before the reduction (they do
    // not depend on it), then the rest eight per thread at a time
    constexpr int U = 8;
    const int tid = threadIdx.x;
    double pre[U];
#pragma unroll
    for (int r = 0; r < U; ++r) pre[r] = __builtin_nontemporal_load(&u[min(br.start + tid + r * kT, br.end - 1)]);
    __shared__ double nrm;
    const SysBlocks sb = sp_norm.sys[br.sys];
    if (threadIdx.x < 64) {
        const double* pi = W.part + sb.blk0;
        double s = 0.0;
        for (int b0 = 0; b0 < sb.nblk; b0 += 16 * 64) {
            double t[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int b = b0 + q * 64 + (int)threadIdx.x;
                t[q] = b < sb.nblk ? pi[b] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) s += t[q];
        }
        s = wave_sum(s);
        if (threadIdx.x == 0) nrm = s;
    }
    __syncthreads();
    const double hn = sqrt(nrm);
    const bool div = W.st->method == 0;
    if (first) {  // block-uniform: the whole block takes part in the epilogue
        step_epilogue(W, j, hn);
        if (threadIdx.x == 0 && W.hflag) W.hflag[j] = W.st->active;
    }
    if (hn < 1e-30) return;  // breakdown: the epilogue ends the cycle, no v_{j+1}
    const double sc = div ? hn : 1.0 / hn;
#pragma unroll
    for (int r = 0; r < U; ++r) {
        const int e = br.start + tid + r * kT;
        if (e < br.end) v[e] = div ? pre[r] / sc : pre[r] * sc;
    }
    for (int e0 = br.start + tid + U * kT; e0 < br.end; e0 += U * kT) {
        double a[U];
#pragma unroll
        for (int r = 0; r < U; ++r) a[r] = __builtin_nontemporal_load(&u[min(e0 + r * kT, br.end - 1)]);
#pragma unroll
        for (int r = 0; r < U; ++r)
            if (e0 + r * kT < br.end) v[e0 + r * kT] = div ? a[r] / sc : a[r] * sc;
    }
}

// ---- MGS in the reference's order (gmres.cpp:93-97) -------------------------
// k_mgs: w -= h_{i-1} V_{i-1} (when i > 0), then partial sums of w . V_i, or
// of w . w when Vi is null; k_mgs_finish stores H[i,j] or runs the epilogue.
__global__ void __launch_bounds__(kT)
k_mgs(Segs sp, const double* __restrict__ Vprev, const double* __restrict__ Vi,
      double* __restrict__ w, Work W0, int i, int j) {
    const BlockRange br = sp.blk[blockIdx.x];
    if (!W0.st[br.sys].active) return;
    const Work W = sys_work(W0, br.sys);
    __shared__ double red[kT / 64][kRedCols];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m = W.st->m;
    const double hprev = Vprev ? W.H[(i - 1) + j * (m + 1)] : 0.0;
    double s = 0.0;
    for (int e = br.start + tid; e < br.end; e += kT) {
        double we = w[e];
        if (Vprev) {
            we -= hprev * Vprev[e];
            w[e] = we;
        }
        s += Vi ? we * Vi[e] : we * we;
    }
    s = wave_sum(s);
    if (lane == 0) red[wv][0] = s;
    store_partials(red, 1, W.part, W.pstride);
}

__global__ void __launch_bounds__(kTF) k_mgs_finish(Work W0, Segs sp, int i, int j, int is_norm) {
    const Work W = sys_work(W0, blockIdx.x);
    if (!W.st->active) return;
    double* out = W.h2 + kRedCols - 1;
    reduce_partials(W, sp.sys[blockIdx.x], 1, out);
    if (is_norm) step_epilogue(W, j, sqrt(*out));
    else if (threadIdx.x == 0) W.H[i + j * (W.st->m + 1)] = *out;
}

// ---- restart ---------------------------------------------------------------
// Partial sums of u.u and (when r is given) r.r; blocks of finished systems
// skip when skip_done is set.
__global__ void __launch_bounds__(kT)
k_norms(Segs sp, const double* __restrict__ u, const double* __restrict__ r, Work W,
        int skip_done) {
    const BlockRange br = sp.blk[blockIdx.x];
    if (skip_done && W.mode[br.sys] == kDone) return;
    __shared__ double red[kT / 64][kRedCols];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double su = 0.0, sr = 0.0;
    for (int e = br.start + tid; e < br.end; e += kT) {
        const double a = u[e];
        su += a * a;
        if (r) {
            const double b = r[e];
            sr += b * b;
        }
    }
    su = wave_sum(su);
    sr = wave_sum(sr);
    if (lane == 0) {
        red[wv][0] = su;
        red[wv][1] = sr;
    }
    store_partials(red, 2, W.part, W.pstride);
}

// ||u|| from k_norms into st->bnorm (also used for the exit residual).
__global__ void __launch_bounds__(kTF) k_bnorm_finish(Work W0, Segs sp) {
    const Work W = sys_work(W0, blockIdx.x);
    double* out = W.h2 + kRedCols - 2;
    reduce_partials(W, sp.sys[blockIdx.x], 2, out);
    if (threadIdx.x == 0) W.st->bnorm = sqrt(out[0]);
}

// Beta test and start of the cycle (gmres.cpp:59-81, fgmres.cpp:54-71); in
// true-residual mode the relative residual ||r||/||b|| decides convergence.
// A system in verification mode only checks; a finished one does nothing.
__global__ void __launch_bounds__(kTF) k_restart_finish(Work W0, Segs sp, int has_r) {
    const Work W = sys_work(W0, blockIdx.x);
    KState* st = W.st;
    const int md = *W.mode;
    if (md == kDone) {
        if (threadIdx.x == 0) {
            st->active = 0;
            st->k = 0;
        }
        return;
    }
    double* out = W.h2 + kRedCols - 2;
    reduce_partials(W, sp.sys[blockIdx.x], 2, out);
    if (threadIdx.x != 0) return;
    if (st->method == 0) st->applies += st->has_m;  // M^{-1} r of the restart
    const double beta = sqrt(out[0]);
    const double rn = has_r ? sqrt(out[1]) : beta;
    st->k = 0;
    if (st->rtol > 0.0) {
        // b = 0: the absolute residual decides (x = 0 is exact), as at exit
        const double rel = st->bnorm > 0.0 ? rn / st->bnorm : rn;
        st->rel_true = rel;
        if (rel <= st->rtol || !(beta > 0.0)) {
            st->converged = rel <= st->rtol;
            st->active = 0;
            return;
        }
        if (st->first) {
            st->tol = st->method == 0 ? st->rtol * beta
                                      : st->rtol * (st->bnorm > 0.0 ? st->bnorm : 1.0);
        } else if (st->est_conv) {
            double f = 0.5 * st->rtol / rel;
            st->tol *= f < 1e-3 ? 1e-3 : (f > 0.9 ? 0.9 : f);
        }
        st->est_conv = 0;
    } else if (beta < st->tol) {
        st->converged = 1;
        st->active = 0;
        return;
    }
    st->first = 0;
    if (md == kVerify) {
        st->active = 0;
        return;
    }
    st->active = 1;
    const int m = st->m;
    W.g[0] = beta;
    for (int i = 1; i <= m; ++i) W.g[i] = 0.0;
    st->scale = st->method == 0 ? beta : 1.0 / beta;
}

// v = u / scale (GMRES, gmres.cpp:76-77,107-108) or u * scale (FGMRES,
// fgmres.cpp:66-67,97-98).
__global__ void __launch_bounds__(kT)
k_scale_basis(Segs sp, const double* __restrict__ u, double* __restrict__ v, Work W) {
    const BlockRange br = sp.blk[blockIdx.x];
    const KState* st = W.st + br.sys;
    if (!st->active) return;
    const double sc = st->scale;
    const bool div = st->method == 0;
    for (int e = br.start + threadIdx.x; e < br.end; e += kT) v[e] = div ? u[e] / sc : u[e] * sc;
}

// Back substitution H y = g (gmres.cpp:131-137), one workgroup per system:
// the k x k upper triangle is staged in LDS, one thread solves in the
// reference's order.
__global__ void __launch_bounds__(64) k_backsolve(Work W0) {
    extern __shared__ double hl[];  // packed rows: row i holds H[i][i..k-1]
    const Work W = sys_work(W0, blockIdx.x);
    const KState* st = W.st;
    const int k = st->k, m = st->m;
    if (k == 0) return;
    // row i starts at i*k - i*(i-1)/2
    for (int i = 0; i < k; ++i)
        for (int l = i + threadIdx.x; l < k; l += blockDim.x)
            hl[i * k - i * (i - 1) / 2 + (l - i)] = W.H[i + l * (m + 1)];
    double* ys = hl + k * (k + 1) / 2;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int i = k - 1; i >= 0; --i) {
        const double* row = hl + i * k - i * (i - 1) / 2;
        double yi = W.g[i];
        for (int l = i + 1; l < k; ++l) yi -= row[l - i] * ys[l];
        ys[i] = yi / row[0];
        W.y[i] = ys[i];
    }
}

// x += sum_i y_i B_i in the reference's axpy order (gmres.cpp:140-141).
__global__ void __launch_bounds__(kT)
k_update(Segs sp, double* __restrict__ x, const double* __restrict__ B, size_t ld, Work W0) {
    const BlockRange br = sp.blk[blockIdx.x];
    const Work W = sys_work(W0, br.sys);
    const int k = W.st->k;
    if (k == 0) return;
    __shared__ double ys[kMaxCols];
    for (int i = threadIdx.x; i < k; i += blockDim.x) ys[i] = W.y[i];
    __syncthreads();
    for (int e = br.start + threadIdx.x; e < br.end; e += kT) {
        double xe = x[e];
        for (int i = 0; i < k; ++i) xe += ys[i] * __builtin_nontemporal_load(&B[(size_t)i * ld + e]);
        x[e] = xe;
    }
}

__global__ void k_init_state(KState* st, int nsys, int method, int m, int has_m, double tol,
                             double rtol) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    KState k{};
    k.active = 0;
    k.k = 0;
    k.converged = 0;
    k.total = 0;
    k.est_conv = 0;
    k.first = 1;
    k.method = method;
    k.m = m;
    k.has_m = has_m;
    k.applies = 0;
    k.tol = tol;
    k.rtol = rtol;
    k.bnorm = 0.0;
    k.rel_true = -1.0;
    k.scale = 1.0;
    k.res = 0.0;
    st[s] = k;
}

// Device and pinned buffers of sg_solve, kept by the context and grown on
// demand, so a repeated solve allocates nothing; also the record of the last
// single-system solve's basis for sg_solve_last_basis.
struct KrylovCache {
    DevBuf<double> vecs;  // V, Z, r, z, w, t
    DevBuf<double> small;
    DevBuf<KState> st;
    DevBuf<int> mode;
    int* pinned = nullptr;   // lookahead flags (m)
    int* h_mode = nullptr;   // per-system control (nsys)
    KState* h_st = nullptr;  // per-system state read back (nsys)
    size_t n_pinned = 0, n_hmode = 0, n_hst = 0;
    // reduction tables, rebuilt when the system sizes change
    std::vector<int> seg_off;
    int seg_cu = 0;
    SegsHost seg_red, seg_gd, seg_gu;
    bool busy = false;
    // last single-system solve (valid until the next solve on the context)
    bool basis_valid = false;
    int b_n = 0, b_m = 0, b_k = 0, b_kv = 0, b_hasz = 0;
    size_t b_ld = 0;
    const double* b_V = nullptr;
    const double* b_Z = nullptr;
    const double* b_Hr = nullptr;
    ~KrylovCache() {
        if (pinned) (void)hipHostFree(pinned);
        if (h_mode) (void)hipHostFree(h_mode);
        if (h_st) (void)hipHostFree(h_st);
    }
};

void krylov_cache_free(KrylovCache* c) { delete c; }

template <typename T>
static void grow(DevBuf<T>& b, size_t n) {
    if (b.n < n) b.alloc(n);
}

template <typename T>
static T* grow_pinned(T*& p, size_t& cap, size_t n) {
    if (cap < n) {
        if (p) SG_HIP(hipHostFree(p));
        p = nullptr;
        cap = 0;
        SG_HIP(hipHostMalloc((void**)&p, sizeof(T) * n, hipHostMallocDefault));
        cap = n;
    }
    return p;
}

// A solve nested inside another one on the same context (e.g. from a host
// callback preconditioner) gets a workspace of its own.
struct CacheLease {
    KrylovCache* c;
    bool own;
    explicit CacheLease(sg_context* ctx) {
        if (!ctx->krylov) ctx->krylov = new KrylovCache();
        if (ctx->krylov->busy) {
            c = new KrylovCache();
            own = true;
        } else {
            c = ctx->krylov;
            own = false;
        }
        c->busy = true;
        c->basis_valid = false;
    }
    ~CacheLease() {
        c->busy = false;
        if (own) delete c;
    }
};

// Graph objects of one solve.
struct GraphHold {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphHold() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// The launches of one restart cycle over a laid-out workspace: restart
// residual and v_0, Arnoldi step j, back substitution and x update.
struct Cycle {
    sg_context* ctx;
    sg_csr* A;
    sg_precond* M;
    const double* d_b;
    double* d_x;
    int method, m, nsys;
    bool cgs2, rmode;
    double *V, *Z, *r, *z, *w, *t;
    size_t ld;
    Work W;
    Segs sp_red, sp_gd, sp_gu;
    int nb_red, nb_gd, nb_gu;
    Guard step_guard;
    size_t bs_lds;
    hipStream_t s;

    void restart() const {
        {
            ProfScope ps(ctx, SG_K_UPDATE);
            spmv_launch(A, -1.0, d_x, 1.0, d_b, r, Guard{}, s);  // r = b - A x
        }
        const double* u = r;
        if (method == 0 && M) {
            precond_apply_launch(M, r, z, Guard{});
            u = z;
        }
        ProfScope ps(ctx, SG_K_UPDATE);
        const bool has_r = rmode && u != r;
        hipLaunchKernelGGL(k_norms, dim3(nb_red), dim3(kT), 0, s, sp_red, u,
                           has_r ? (const double*)r : nullptr, W, 1);
        hipLaunchKernelGGL(k_restart_finish, dim3(nsys), dim3(kTF), 0, s, W, sp_red,
                           has_r ? 1 : 0);
        hipLaunchKernelGGL(k_scale_basis, dim3(nb_red), dim3(kT), 0, s, sp_red, u, V, W);
        SG_HIP(hipGetLastError());
    }

    void step(int j) const {
        double* Vj = V + (size_t)j * ld;
        if (method == 0) {
            if (M) {
                spmv_launch(A, 1.0, Vj, 0.0, nullptr, t, step_guard, s);
                precond_apply_launch(M, t, w, step_guard);
            } else {
                spmv_launch(A, 1.0, Vj, 0.0, nullptr, w, step_guard, s);
            }
        } else {
            double* Zj = M ? Z + (size_t)j * ld : Vj;
            // w = A z_j, unless the preconditioner formed it already (learned stage)
            if (!(M && precond_apply_az(M, A, Vj, Zj, w, step_guard)))
                spmv_launch(A, 1.0, Zj, 0.0, nullptr, w, step_guard, s);
        }
        {
            ProfScope ps(ctx, SG_K_ORTHO);
            if (cgs2) {
                {
                    ProfScope pd(ctx, SG_K_GRAM_DOT, j + 1);
                    hipLaunchKernelGGL(k_gram_dot, dim3(nb_gd), dim3(kTGD), 0, s, sp_gd, j + 1, V,
                                       ld, w, W, j);
                }
                hipLaunchKernelGGL(k_gram_sums, dim3(2 * (j + 1), nsys), dim3(64), 0, s, W, sp_gd,
                                   j + 1, j);
                {
                    ProfScope pu(ctx, SG_K_GRAM_UPDATE, j + 1);
                    hipLaunchKernelGGL(k_gram_update, dim3(nb_gu), dim3(kTGU), 0, s, sp_gu, j + 1,
                                       V, ld, w, W, j);
                }
                if (j + 1 == m)
                    hipLaunchKernelGGL(k_norm_finish, dim3(nsys), dim3(kTF), 0, s, W, sp_gu, j);
            } else {
                for (int i = 0; i <= j + 1; ++i) {
                    const double* Vp = i ? V + (size_t)(i - 1) * ld : nullptr;
                    const double* Vi = i <= j ? V + (size_t)i * ld : nullptr;
                    hipLaunchKernelGGL(k_mgs, dim3(nb_red), dim3(kT), 0, s, sp_red, Vp, Vi, w, W,
                                       i, j);
                    hipLaunchKernelGGL(k_mgs_finish, dim3(nsys), dim3(kTF), 0, s, W, sp_red, i, j,
                                       i > j ? 1 : 0);
                }
            }
            SG_HIP(hipGetLastError());
        }
        if (j + 1 < m) {
            ProfScope ps(ctx, SG_K_UPDATE);
            if (cgs2)
                hipLaunchKernelGGL(k_norm_scale, dim3(nb_red), dim3(kT), 0, s, sp_red, sp_gu, w,
                                   V + (size_t)(j + 1) * ld, W, j);
            else
                hipLaunchKernelGGL(k_scale_basis, dim3(nb_red), dim3(kT), 0, s, sp_red, w,
                                   V + (size_t)(j + 1) * ld, W);
        }
    }

    void tail() const {
        ProfScope ps(ctx, SG_K_UPDATE);
        hipLaunchKernelGGL(k_backsolve, dim3(nsys), dim3(64), bs_lds, s, W);
        hipLaunchKernelGGL(k_update, dim3(nb_red), dim3(kT), 0, s, sp_red, d_x,
                           (method && M) ? Z : V, ld, W);
        SG_HIP(hipGetLastError());
    }
};

static void solve_impl(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                       const sg_solve_params* prm, sg_solve_result* res, double* hist,
                       int hist_cap, bool use_graph) {
    sg_context* ctx = A->ctx;
    if (A->nrows != A->ncols) fail(SG_E_ARG, "GMRES needs a square matrix");
    if (M && !precond_ready(M)) fail(SG_E_STATE, "preconditioner used before setup");
    if (M && precond_n(M) != A->nrows) fail(SG_E_ARG, "preconditioner size mismatch");
    const int n = A->nrows;
    const int m = prm->restart;
    const int method = prm->method == SG_FGMRES ? 1 : 0;
    if (m < 1 || m + 1 > kMaxCols) fail(SG_E_ARG, "restart must be in [1, 127]");
    if (prm->max_iter < 0) fail(SG_E_ARG, "max_iter must be >= 0");
    if (n < 1) fail(SG_E_ARG, "empty system");
    const std::vector<int>& off = A->sys_off;
    const int nsys = (int)off.size() - 1;
    if (M && method == 0 && precond_is_flexible(M))
        fail(SG_E_ARG, "a flexible (learned) stage needs FGMRES");
    if (M && precond_is_callback(M) && (use_graph || nsys > 1))
        fail(SG_E_ARG, "a host-callback preconditioner runs only in unbatched, uncaptured solves");
    const bool cgs2 = prm->ortho == SG_ORTHO_CGS2;
    const int L = prm->lookahead > 0 ? prm->lookahead : 2;
    const bool rmode = prm->rtol > 0.0;
    hipStream_t s = ctx->stream;
    SG_HIP(hipSetDevice(ctx->device));
    if (M && nsys > 1) precond_bind_systems(M, off);
    auto t0 = std::chrono::steady_clock::now();

    const size_t ld = basis_ld(n);
    const int nz = (method && M) ? m : 0;
    const int nvec = (m + 1) + nz + 4;
    CacheLease lease(ctx);
    KrylovCache& ws = *lease.c;
    GraphHold gh;
    grow(ws.vecs, ld * nvec);
    double* V = ws.vecs.p;
    double* Z = nz ? V + ld * (m + 1) : nullptr;
    double* r = V + ld * ((m + 1) + nz);
    double* z = r + ld;
    double* w = z + ld;
    double* t = w + ld;
    if (nsys > 1)  // stopped systems of a batch keep computing, on finite values
        SG_HIP(hipMemsetAsync(V, 0, sizeof(double) * ld * nvec, s));

    if (ws.seg_off != off || ws.seg_cu != ctx->num_cu) {
        build_segs(off, 4 * ctx->num_cu, kTile, ws.seg_red, s);
        build_segs(off, 4 * ctx->num_cu, kTileGD, ws.seg_gd, s, kPassGD);
        build_segs(off, 4 * ctx->num_cu, kTileGU, ws.seg_gu, s);
        ws.seg_off = off;
        ws.seg_cu = ctx->num_cu;
    }
    const SegsHost &seg_red = ws.seg_red, &seg_gd = ws.seg_gd, &seg_gu = ws.seg_gu;
    const Segs sp_red = seg_red.dev(), sp_gd = seg_gd.dev(), sp_gu = seg_gu.dev();
    const int pstride = std::max(seg_red.nblk, std::max(seg_gd.nblk, seg_gu.nblk));

    const int hcap = prm->max_iter + m + 2;
    const size_t per_sys = 2 * (size_t)(m + 1) * m + (m + 1) + 3 * (size_t)m + 2 * kRedCols +
                           hcap + (size_t)kMaxCols * kMaxCols;
    grow(ws.small, per_sys * nsys + (size_t)kRedCols * pstride);
    grow(ws.st, nsys);
    grow(ws.mode, nsys);
    grow_pinned(ws.pinned, ws.n_pinned, (size_t)m + 1);
    grow_pinned(ws.h_mode, ws.n_hmode, (size_t)nsys);
    grow_pinned(ws.h_st, ws.n_hst, (size_t)nsys);
    Work W;
    W.H = ws.small.p;
    W.Hr = W.H + (size_t)(m + 1) * m * nsys;
    W.g = W.Hr + (size_t)(m + 1) * m * nsys;
    W.cs = W.g + (size_t)(m + 1) * nsys;
    W.sn = W.cs + (size_t)m * nsys;
    W.y = W.sn + (size_t)m * nsys;
    W.h1 = W.y + (size_t)m * nsys;
    W.h2 = W.h1 + (size_t)kRedCols * nsys;
    W.hist = W.h2 + (size_t)kRedCols * nsys;
    W.G = W.hist + (size_t)hcap * nsys;
    W.part = W.G + (size_t)kMaxCols * kMaxCols * nsys;
    W.st = ws.st.p;
    W.mode = ws.mode.p;
    W.m = m;
    W.hcap = hcap;
    W.pstride = pstride;
    W.hflag = nullptr;
    if (nsys == 1 && !use_graph && cgs2)
        SG_HIP(hipHostGetDevicePointer((void**)&W.hflag, ws.pinned, 0));
    KState* st = ws.st.p;
    int* hmode = ws.h_mode;
    const Guard step_guard{&st->active, (int)(sizeof(KState) / sizeof(int)), nsys > 1};

    const size_t bs_lds = sizeof(double) * ((size_t)m * (m + 1) / 2 + m + 1);
    SG_HIP(hipFuncSetAttribute((const void*)k_backsolve,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bs_lds));
    const double tol = rmode ? 0.0 : prm->tol_abs;
    hipLaunchKernelGGL(k_init_state, dim3((nsys + 63) / 64), dim3(64), 0, s, st, nsys, method, m,
                       M ? 1 : 0, tol, prm->rtol);
    for (int q = 0; q < nsys; ++q) hmode[q] = kRun;
    SG_HIP(hipMemcpyAsync(ws.mode.p, hmode, sizeof(int) * nsys, hipMemcpyHostToDevice, s));
    // ||b|| for the relative test and the report
    hipLaunchKernelGGL(k_norms, dim3(seg_red.nblk), dim3(kT), 0, s, sp_red, d_b,
                       (const double*)nullptr, W, 0);
    hipLaunchKernelGGL(k_bnorm_finish, dim3(nsys), dim3(kTF), 0, s, W, sp_red);
    SG_HIP(hipGetLastError());

    const Cycle cyc{ctx, A, M, d_b, d_x, method, m, nsys, cgs2, rmode, V, Z, r, z, w, t, ld, W,
                    sp_red, sp_gd, sp_gu, seg_red.nblk, seg_gd.nblk, seg_gu.nblk, step_guard,
                    bs_lds, s};
    auto launch_restart = [&]() { cyc.restart(); };
    auto launch_step = [&](int j) { cyc.step(j); };
    auto launch_tail = [&]() { cyc.tail(); };

    if (use_graph) {
        // kernel shapes are picked before capture (tuning times real launches)
        spmv_ensure_tuned(A);
        if (M) precond_ensure_tuned(M);
        SG_HIP(hipStreamSynchronize(s));
        ProfOff po(ctx);
        CaptureGuard cg{s};
        SG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        cg.on = true;
        launch_restart();
        for (int j = 0; j < m; ++j) launch_step(j);
        launch_tail();
        cg.on = false;
        SG_HIP(hipStreamEndCapture(s, &gh.graph));
        SG_HIP(hipGraphInstantiate(&gh.exec, gh.graph, nullptr, nullptr, 0));
    }

    std::vector<hipEvent_t> evs(m);
    for (int j = 0; j < m; ++j) evs[j] = ctx->get_event();
    struct EvBack {
        sg_context* c;
        std::vector<hipEvent_t>& e;
        ~EvBack() {
            for (auto x : e) c->event_pool.push_back(x);
        }
    } evback{ctx, evs};

    std::vector<char> verify(nsys, 0), done(nsys, 0);
    std::vector<int> cycles(nsys, 0);
    KState* hst = ws.h_st;
    std::memset((void*)hst, 0, sizeof(KState) * nsys);
    int last_k = 0;  // columns of the last cycle that ran Arnoldi steps (single system)
    for (int cycle = 0;; ++cycle) {
        // per-system decisions before the cycle (gmres.cpp:57 bounds cycles)
        bool any = false, any_run = false;
        for (int q = 0; q < nsys; ++q) {
            if (!done[q] && !verify[q] && cycle >= prm->max_iter) {
                if (rmode && hst[q].est_conv) verify[q] = 1;
                else done[q] = 1;
            }
            hmode[q] = done[q] ? kDone : (verify[q] ? kVerify : kRun);
            any = any || !done[q];
            any_run = any_run || hmode[q] == kRun;
        }
        if (!any) break;
        SG_HIP(hipMemcpyAsync(ws.mode.p, hmode, sizeof(int) * nsys, hipMemcpyHostToDevice, s));
        if (use_graph) {
            SG_HIP(hipGraphLaunch(gh.exec, s));
        } else {
            launch_restart();
            for (int j = 0; j < m && any_run; ++j) {
                if (nsys == 1 && j >= L) {
                    SG_HIP(hipEventSynchronize(evs[j - L]));
                    if (!ws.pinned[j - L]) break;
                }
                launch_step(j);
                if (nsys == 1) {
                    // the CGS2 epilogue stores the flag into host memory itself;
                    // MGS copies it
                    if (!W.hflag)
                        SG_HIP(hipMemcpyAsync(&ws.pinned[j], &st->active, sizeof(int),
                                              hipMemcpyDeviceToHost, s));
                    SG_HIP(hipEventRecord(evs[j], s));
                }
            }
            if (any_run) launch_tail();
        }
        SG_HIP(hipMemcpyAsync(hst, st, sizeof(KState) * nsys, hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        if (nsys == 1 && hmode[0] == kRun && hst[0].k > 0) last_k = hst[0].k;
        for (int q = 0; q < nsys; ++q) {
            if (hmode[q] == kDone) continue;
            ++cycles[q];
            if (hst[q].converged || verify[q]) {
                done[q] = 1;
            } else if (rmode && hst[q].est_conv) {
                if (hst[q].total >= prm->max_iter) verify[q] = 1;  // next restart checks
            } else if (hst[q].total >= prm->max_iter) {
                done[q] = 1;  // gmres.cpp:146-147
            }
        }
    }
    SG_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    ctx->collect_profiles();
    if (nsys == 1) {
        ws.basis_valid = true;
        ws.b_n = n;
        ws.b_m = m;
        ws.b_k = last_k;
        ws.b_ld = ld;
        ws.b_V = V;
        ws.b_Z = Z;
        ws.b_hasz = Z ? 1 : 0;
        ws.b_Hr = W.Hr;
        // v_k exists when the last step was not the cycle's m-th and did not
        // break down (k_norm_scale forms v_{j+1} only then)
        ws.b_kv = last_k;
        if (last_k > 0 && last_k < m) {
            double hk = 0.0;
            SG_HIP(hipMemcpy(&hk, W.Hr + (size_t)last_k + (size_t)(last_k - 1) * (m + 1),
                             sizeof(double), hipMemcpyDeviceToHost));
            if (hk >= 1e-30) ws.b_kv = last_k + 1;
        }
    }

    const double elapsed = std::chrono::duration<double>(t1 - t0).count();
    std::vector<double> bn(nsys);
    bool need_final = false;
    for (int q = 0; q < nsys; ++q) {
        sg_solve_result& R = res[q];
        const KState& h = hst[q];
        R.iterations = h.total;
        R.converged = h.converged;
        R.cycles = cycles[q];
        R.precond_applies = h.applies;
        R.hist_len = h.total;
        R.time_solve_s = elapsed;
        bn[q] = h.bnorm;
        std::vector<double> hv(h.total);
        if (h.total)
            SG_HIP(hipMemcpy(hv.data(), W.hist + (size_t)q * hcap, sizeof(double) * h.total,
                             hipMemcpyDeviceToHost));
        R.residual_norm = h.total ? hv.back() : 0.0;  // gmres.cpp:150-155
        if (hist)
            for (int i = 0; i < h.total && i < hist_cap; ++i)
                hist[(size_t)q * hist_cap + i] = hv[i];
        if (rmode && h.converged) R.rel_true_residual = h.rel_true;
        else need_final = true;
    }
    // FP64 true residual at exit (reported; inside the solve in rtol mode)
    if (need_final) {
        for (int q = 0; q < nsys; ++q) hmode[q] = kRun;
        SG_HIP(hipMemcpyAsync(ws.mode.p, hmode, sizeof(int) * nsys, hipMemcpyHostToDevice, s));
        spmv_launch(A, -1.0, d_x, 1.0, d_b, r, Guard{}, s);
        hipLaunchKernelGGL(k_norms, dim3(seg_red.nblk), dim3(kT), 0, s, sp_red, r,
                           (const double*)nullptr, W, 0);
        hipLaunchKernelGGL(k_bnorm_finish, dim3(nsys), dim3(kTF), 0, s, W, sp_red);
        SG_HIP(hipMemcpyAsync(hst, st, sizeof(KState) * nsys, hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < nsys; ++q) {
            if (rmode && res[q].converged) continue;
            const double rn = hst[q].bnorm;  // k_bnorm_finish stored ||r|| in the bnorm slot
            res[q].rel_true_residual = bn[q] > 0.0 ? rn / bn[q] : rn;
        }
    }
}

// ---- no-allocation solve ------------------------------------------------------
// sg_solve_noalloc keeps its whole device workspace (state, control word,
// reduction tables, Hessenberg and Gram storage, partial sums, Krylov basis)
// in one static device array per GPU, so a call makes no device or pinned
// allocation.  Host code never passes `scratch` itself to a kernel: it uses
// the device address that hipGetSymbolAddress returns for the current device
// (g_scratch[device], looked up once per device at context creation).
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[kMaxScratchDevices];
static std::string g_scratch_err[kMaxScratchDevices];

void noalloc_init_device(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_scratch[device] = static_cast<double*>(p);
}

static double* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxScratchDevices)
        fail(SG_E_ARG, "sg_solve_noalloc: device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device])
        fail(SG_E_HIP, "sg_solve_noalloc: scratch array unavailable on this device (" +
                           g_scratch_err[device] + ")");
    return g_scratch[device];
}

// Block table of one system of n elements into a caller array (build_segs for
// one system, without allocating); returns the block count.
static int segs_into(int n, int max_blocks, int tile, int pass, BlockRange* out) {
    int nblk, chunk;
    const int mb = pass > 0 ? std::max(max_blocks, (n + pass - 1) / pass) : max_blocks;
    split_shape(n, mb, tile, nblk, chunk);
    for (int b = 0; b < nblk; ++b) out[b] = BlockRange{0, b * chunk, std::min(n, (b + 1) * chunk), 0};
    return nblk;
}

static void solve_noalloc_impl(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                               const sg_solve_params* prm, sg_solve_result* res, double* hist,
                               int hist_cap) {
    sg_context* ctx = A->ctx;
    if (A->sys_off.size() != 2) fail(SG_E_ARG, "sg_solve_noalloc: single systems only");
    if (A->nrows != A->ncols) fail(SG_E_ARG, "GMRES needs a square matrix");
    if (prm->method != SG_GMRES_LEFT && prm->method != SG_FGMRES)
        fail(SG_E_ARG, "sg_solve_noalloc: GMRES or FGMRES only");
    if (prm->flags & SG_SOLVE_GRAPH) fail(SG_E_ARG, "sg_solve_noalloc: no graph capture");
    if (M && !precond_ready(M)) fail(SG_E_STATE, "preconditioner used before setup");
    if (M && precond_n(M) != A->nrows) fail(SG_E_ARG, "preconditioner size mismatch");
    if (M && (precond_is_callback(M) || precond_is_flexible(M)))
        fail(SG_E_ARG, "sg_solve_noalloc: device preconditioners only (no callback or learned stage)");
    const int n = A->nrows;
    const int m = prm->restart;
    const int method = prm->method == SG_FGMRES ? 1 : 0;
    if (m < 1 || m + 1 > kMaxCols) fail(SG_E_ARG, "restart must be in [1, 127]");
    if (prm->max_iter < 0) fail(SG_E_ARG, "max_iter must be >= 0");
    if (n < 1) fail(SG_E_ARG, "empty system");
    const bool cgs2 = prm->ortho == SG_ORTHO_CGS2;
    const bool rmode = prm->rtol > 0.0;
    hipStream_t s = ctx->stream;
    SG_HIP(hipSetDevice(ctx->device));
    double* base = scratch_ptr(ctx->device);

    // reduction tables on the host (no heap), then into the scratch array
    thread_local BlockRange hb[3][kMaxBlocks];
    const int nb_red = segs_into(n, 4 * ctx->num_cu, kTile, 0, hb[0]);
    const int nb_gd = segs_into(n, 4 * ctx->num_cu, kTileGD, kPassGD, hb[1]);
    const int nb_gu = segs_into(n, 4 * ctx->num_cu, kTileGU, 0, hb[2]);
    const int pstride = std::max(nb_red, std::max(nb_gd, nb_gu));
    const int nbs[3] = {nb_red, nb_gd, nb_gu};

    // layout (byte offsets, 256-aligned): state | mode | 3 x (blocks, system) |
    // small | basis
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t ld = basis_ld(n);
    const int nz = (method && M) ? m : 0;
    const int nvec = (m + 1) + nz + 4;
    const size_t hcap = (size_t)prm->max_iter + m + 2;
    const size_t small_n = 2 * (size_t)(m + 1) * m + (m + 1) + 3 * (size_t)m + 2 * kRedCols +
                           hcap + (size_t)kMaxCols * kMaxCols + (size_t)kRedCols * pstride;
    size_t off = 0;
    const size_t o_st = off;
    off += al(sizeof(KState));
    const size_t o_mode = off;
    off += al(sizeof(int));
    size_t o_blk[3], o_sys[3];
    for (int q = 0; q < 3; ++q) {
        o_blk[q] = off;
        off += al(sizeof(BlockRange) * nbs[q]);
        o_sys[q] = off;
        off += al(sizeof(SysBlocks));
    }
    const size_t o_small = off;
    off += al(sizeof(double) * small_n);
    const size_t o_vecs = off;
    const double need = (double)off + (double)sizeof(double) * (double)ld * nvec;
    if (need > (double)kScratchBytes)
        fail(SG_E_ARG, "sg_solve_noalloc: the system needs " +
                           std::to_string((long long)need) + " bytes of workspace; the static "
                           "scratch array holds " + std::to_string(kScratchBytes));
    char* cb = reinterpret_cast<char*>(base);
    KState* st = reinterpret_cast<KState*>(cb + o_st);
    int* dmode = reinterpret_cast<int*>(cb + o_mode);
    Segs sp[3];
    for (int q = 0; q < 3; ++q) {
        BlockRange* db = reinterpret_cast<BlockRange*>(cb + o_blk[q]);
        SysBlocks* ds = reinterpret_cast<SysBlocks*>(cb + o_sys[q]);
        const SysBlocks hs{0, nbs[q]};
        SG_HIP(hipMemcpyAsync(db, hb[q], sizeof(BlockRange) * nbs[q], hipMemcpyHostToDevice, s));
        SG_HIP(hipMemcpyAsync(ds, &hs, sizeof(SysBlocks), hipMemcpyHostToDevice, s));
        sp[q] = Segs{db, ds};
    }
    SG_HIP(hipStreamSynchronize(s));  // the host tables are reused by the next call
    double* small = reinterpret_cast<double*>(cb + o_small);
    double* V = reinterpret_cast<double*>(cb + o_vecs);
    double* Z = nz ? V + ld * (m + 1) : nullptr;
    double* r = V + ld * ((m + 1) + nz);
    double* z = r + ld;
    double* w = z + ld;
    double* t = w + ld;
    Work W;
    W.H = small;
    W.Hr = W.H + (size_t)(m + 1) * m;
    W.g = W.Hr + (size_t)(m + 1) * m;
    W.cs = W.g + (m + 1);
    W.sn = W.cs + m;
    W.y = W.sn + m;
    W.h1 = W.y + m;
    W.h2 = W.h1 + kRedCols;
    W.hist = W.h2 + kRedCols;
    W.G = W.hist + hcap;
    W.part = W.G + (size_t)kMaxCols * kMaxCols;
    W.st = st;
    W.mode = dmode;
    W.m = m;
    W.hcap = (int)hcap;
    W.pstride = pstride;
    W.hflag = nullptr;

    // no kernel-shape timing and no profiling events inside this call
    struct NoTune {
        sg_context* c;
        bool was;
        ~NoTune() { c->no_tune = was; }
    } nt{ctx, ctx->no_tune};
    ctx->no_tune = true;
    ProfOff po(ctx);

    auto t0 = std::chrono::steady_clock::now();
    const size_t bs_lds = sizeof(double) * ((size_t)m * (m + 1) / 2 + m + 1);
    SG_HIP(hipFuncSetAttribute((const void*)k_backsolve,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bs_lds));
    const double tol = rmode ? 0.0 : prm->tol_abs;
    // every value read before it is written in this call starts here: the
    // state (k_init_state) and the control word; the rest of the array keeps
    // the previous call's contents and is overwritten before use
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, st, 1, method, m, M ? 1 : 0, tol,
                       prm->rtol);
    int hmode = kRun;
    SG_HIP(hipMemcpyAsync(dmode, &hmode, sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_norms, dim3(nb_red), dim3(kT), 0, s, sp[0], d_b, (const double*)nullptr,
                       W, 0);
    hipLaunchKernelGGL(k_bnorm_finish, dim3(1), dim3(kTF), 0, s, W, sp[0]);
    SG_HIP(hipGetLastError());

    const Guard step_guard{&st->active, (int)(sizeof(KState) / sizeof(int)), false};
    const Cycle cyc{ctx, A, M, d_b, d_x, method, m, 1, cgs2, rmode, V, Z, r, z, w, t, ld, W,
                    sp[0], sp[1], sp[2], nb_red, nb_gd, nb_gu, step_guard, bs_lds, s};
    KState hst;
    std::memset((void*)&hst, 0, sizeof(hst));
    bool verify = false, done = false;
    int cycles = 0;
    for (int cycle = 0;; ++cycle) {
        if (!done && !verify && cycle >= prm->max_iter) {
            if (rmode && hst.est_conv) verify = true;
            else done = true;
        }
        if (done) break;
        hmode = verify ? kVerify : kRun;
        SG_HIP(hipMemcpyAsync(dmode, &hmode, sizeof(int), hipMemcpyHostToDevice, s));
        cyc.restart();
        if (hmode == kRun) {
            // every step is launched; after convergence its kernels return at once
            for (int j = 0; j < m; ++j) cyc.step(j);
            cyc.tail();
        }
        SG_HIP(hipMemcpyAsync(&hst, st, sizeof(KState), hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        ++cycles;
        if (hst.converged || verify) {
            done = true;
        } else if (rmode && hst.est_conv) {
            if (hst.total >= prm->max_iter) verify = true;  // next restart checks
        } else if (hst.total >= prm->max_iter) {
            done = true;  // gmres.cpp:146-147
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    sg_solve_result& R = *res;
    R.iterations = hst.total;
    R.converged = hst.converged;
    R.cycles = cycles;
    R.precond_applies = hst.applies;
    R.hist_len = hst.total;
    R.time_solve_s = std::chrono::duration<double>(t1 - t0).count();
    R.residual_norm = 0.0;  // gmres.cpp:150-155: the last estimate
    if (hst.total) {
        SG_HIP(hipMemcpy(&R.residual_norm, W.hist + (hst.total - 1), sizeof(double),
                         hipMemcpyDeviceToHost));
        if (hist && hist_cap > 0)
            SG_HIP(hipMemcpy(hist, W.hist, sizeof(double) * std::min(hst.total, hist_cap),
                             hipMemcpyDeviceToHost));
    }
    const double bn = hst.bnorm;
    if (rmode && hst.converged) {
        R.rel_true_residual = hst.rel_true;
    } else {
        // FP64 true residual at exit (reported)
        spmv_launch(A, -1.0, d_x, 1.0, d_b, r, Guard{}, s);
        hipLaunchKernelGGL(k_norms, dim3(nb_red), dim3(kT), 0, s, sp[0], r,
                           (const double*)nullptr, W, 0);
        hipLaunchKernelGGL(k_bnorm_finish, dim3(1), dim3(kTF), 0, s, W, sp[0]);
        SG_HIP(hipMemcpyAsync(&hst, st, sizeof(KState), hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        R.rel_true_residual = bn > 0.0 ? hst.bnorm / bn : hst.bnorm;
    }
}

}  // namespace sg

using namespace sg;

extern "C" sg_status sg_solve(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                              const sg_solve_params* prm, sg_solve_result* res, double* hist,
                              int hist_cap) {
    if (!A || !prm || !res || (A->nrows && (!d_b || !d_x))) return SG_E_ARG;
    return run_guarded(A->ctx, [&] {
        if (A->sys_off.size() != 2)
            fail(SG_E_ARG, "matrix holds a batch of systems: use sg_solve_batch");
        if (prm->method == SG_BICGSTAB)
            return bicgstab_impl(A, M, d_b, d_x, prm, res, hist, hist_cap,
                                 (prm->flags & SG_SOLVE_GRAPH) != 0);
        solve_impl(A, M, d_b, d_x, prm, res, hist, hist_cap, (prm->flags & SG_SOLVE_GRAPH) != 0);
    });
}

extern "C" sg_status sg_solve_batch(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                                    const sg_solve_params* prm, sg_solve_result* res,
                                    double* hist, int hist_cap) {
    if (!A || !prm || !res || (A->nrows && (!d_b || !d_x))) return SG_E_ARG;
    return run_guarded(A->ctx, [&] {
        if (prm->method == SG_BICGSTAB)
            return bicgstab_impl(A, M, d_b, d_x, prm, res, hist, hist_cap,
                                 (prm->flags & SG_SOLVE_NO_GRAPH) == 0);
        solve_impl(A, M, d_b, d_x, prm, res, hist, hist_cap,
                   (prm->flags & SG_SOLVE_NO_GRAPH) == 0);
    });
}

extern "C" sg_status sg_solve_last_basis(sg_context* ctx, int* n, int* m, int* k, int* kv,
                                         double* V, double* Z, double* Hbar) {
    if (!ctx || !n || !m || !k || !kv) return SG_E_ARG;
    return run_guarded(ctx, [&] {
        KrylovCache* c = ctx->krylov;
        if (!c || !c->basis_valid || c->busy)
            fail(SG_E_STATE, "no single-system GMRES/FGMRES solve on this context yet");
        *n = c->b_n;
        *m = c->b_m;
        *k = c->b_k;
        *kv = c->b_kv;
        SG_HIP(hipSetDevice(ctx->device));
        SG_HIP(hipStreamSynchronize(ctx->stream));
        const size_t nb = sizeof(double) * (size_t)c->b_n;
        if (V)
            for (int i = 0; i < c->b_kv; ++i)
                SG_HIP(hipMemcpy(V + (size_t)i * c->b_n, c->b_V + (size_t)i * c->b_ld, nb,
                                 hipMemcpyDeviceToHost));
        if (Z && c->b_hasz)
            for (int i = 0; i < c->b_k; ++i)
                SG_HIP(hipMemcpy(Z + (size_t)i * c->b_n, c->b_Z + (size_t)i * c->b_ld, nb,
                                 hipMemcpyDeviceToHost));
        if (Hbar)
            SG_HIP(hipMemcpy(Hbar, c->b_Hr, sizeof(double) * (size_t)(c->b_m + 1) * c->b_m,
                             hipMemcpyDeviceToHost));
    });
}

extern "C" sg_status sg_solve_noalloc(sg_csr* A, sg_precond* M, const double* d_b, double* d_x,
                                      const sg_solve_params* prm, sg_solve_result* res,
                                      double* hist, int hist_cap) {
    if (!A || !prm || !res || (A->nrows && (!d_b || !d_x))) return SG_E_ARG;
    return run_guarded(A->ctx, [&] {
        solve_noalloc_impl(A, M, d_b, d_x, prm, res, hist, hist_cap);
    });
}
