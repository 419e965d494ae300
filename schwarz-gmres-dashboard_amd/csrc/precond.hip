// Preconditioners on the device: Jacobi, ILU(0), ASM, RAS and two-level
// Schwarz, plus host-callback preconditioners.
//
// Every ILU(0)-based preconditioner is a batch of "subdomains" (one for the
// global ILU(0)).  Setup (host): partition, overlap sets, extract_submatrix,
// ILU level schedules and the packed sweep layout.  Setup (device): IKJ
// ILU(0) factorisation level by level in FP64, then packing of the factor
// values (FP64 or FP32) into the sweep layout.  Apply (device): one workgroup
// per subdomain, taken from a queue sorted by subdomain nonzeros, runs the
// restriction gather, the unit-L and U sweeps level by level with the solution
// vector in LDS (or global memory when it does not fit), and the restricted
// (RAS) or additive (ASM) prolongation, in one kernel.
//
// Sweep layout: the rows of one level are cut into chunks of 64 (one per
// wavefront lane).  Entry k of every row in a chunk is stored contiguously
// (64-wide ELL slices), so the factor stream is read with coalesced loads.
// Padding entries point at a zero slot z[ns] with value 0.  The entries of a
// row keep CSR order, so each row's sum has the reference's sequential order
// (src/precond/ilu0.h:62-76) and the FP64 sweep reproduces it bit for bit.
#include "sg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <exception>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace sg {

constexpr int kChunk = 64;
constexpr int kTrsvThreads = 512;

struct SubInfo {
    int ns;      // rows of the (overlapped) subdomain
    int row0;    // first row in the concatenated local numbering
    int zoff;    // offset of z in the global scratch (ns + 1 slots)
    int flev0;   // first forward level (index into f_lev)
    int nflev;
    int blev0;
    int nblev;
    int own0;    // first owned entry (coarse restriction list)
    int nown;
};

struct Plan {
    // host-side staging, moved to the device at setup
    std::vector<SubInfo> subs;
    std::vector<int> order;                        // subdomains by descending nnz
    std::vector<int> f_lev, f_row, f_gid, f_off, f_K, f_col, f_src;
    std::vector<int> b_lev, b_row, b_out, b_off, b_K, b_col, b_src, b_dsrc;
    std::vector<int> lrp, lci, ldiag;              // concatenated local CSR
    std::vector<double> lval;
    std::vector<int> own_idx;                      // owned global ids, per subdomain ascending
    int total_rows = 0;
    int max_ns = 0;
};

}  // namespace sg

struct sg_precond {
    sg_context* ctx = nullptr;
    sg_precond_kind kind = SG_PC_NONE;
    sg_precond_opts opts{};
    int n = 0;
    bool ready = false;
    // callback
    sg_apply_fn fn = nullptr;
    void* user = nullptr;
    int device_pointers = 0;
    std::vector<double> hx, hy;
    // jacobi
    sg::DevBuf<double> inv_diag;
    // decomposition (host)
    int nparts = 0;
    std::vector<int> part;
    std::vector<std::vector<int>> owned, dofs;
    std::vector<int> sub_nnz;  // local nnz per subdomain
    // sweep plan (device)
    int nsub = 0;
    bool z_in_lds = false;
    size_t lds_bytes = 0;
    sg::DevBuf<sg::SubInfo> d_subs;
    sg::DevBuf<int> d_order;
    sg::DevBuf<int> f_lev, f_row, f_gid, f_off, f_K, f_col, f_src;
    sg::DevBuf<int> b_lev, b_row, b_out, b_off, b_K, b_col, b_src, b_dsrc;
    sg::DevBuf<int> lrp, lci, ldiag;
    sg::DevBuf<double> lval;  // factored FP64 values in local CSR order
    sg::DevBuf<double> f_v64, b_v64, b_d64;
    sg::DevBuf<float> f_v32, b_v32, b_d32;
    sg::DevBuf<double> z64;   // global z scratch when z does not fit in LDS
    sg::DevBuf<float> z32;
    sg::DevBuf<double> zout64;  // ASM: per-subdomain solutions, summed by k_asm_gather
    sg::DevBuf<float> zout32;
    std::vector<int> lrp_host;  // for introspection
    // ASM: deterministic accumulation y[g] = sum over (s ascending) zbuf[pos]
    sg::DevBuf<int> cov_ptr, cov_pos;
    sg::DevBuf<int> d_part;
    // two-level
    int nc = 0;
    std::vector<double> A0, LU;
    std::vector<int> piv;
    sg::DevBuf<double> d_LU64, d_LUt64, d_inv64, d_c64, d_r0_64, d_cz64;
    sg::DevBuf<float> d_LU32, d_LUt32, d_inv32, d_c32, d_r0_32, d_cz32;
    sg::DevBuf<int> d_piv, d_own_ptr, d_own_idx;
    sg::DevBuf<int> d_lu_status;
};

namespace sg {

// ---------------------------------------------------------------------------
// host: per-subdomain sweep packing
// ---------------------------------------------------------------------------

struct SubPack {
    std::vector<int> f_lev, f_row, f_gid, f_K, f_col, f_src;  // f_src local CSR positions
    std::vector<int> b_lev, b_row, b_out, b_K, b_col, b_src, b_dsrc;
    std::vector<int> f_off, b_off;  // relative to this subdomain's ELL start
};

// Packs one sweep.  lower=true: strictly-lower entries in CSR order;
// lower=false: entries right of the diagonal in CSR order plus the diagonal.
static void pack_sweep(const HostCsr& L, const std::vector<int>& diag, const std::vector<int>& lev,
                       bool lower, const std::vector<int>& gid_of_row, std::vector<int>& out_lev,
                       std::vector<int>& out_row, std::vector<int>& out_gid,
                       std::vector<int>& out_off, std::vector<int>& out_K,
                       std::vector<int>& out_col, std::vector<int>& out_src,
                       std::vector<int>* out_dsrc) {
    int ns = L.n;
    int nlev = 0;
    for (int l : lev) nlev = std::max(nlev, l + 1);
    // rows bucketed by level, ascending row order inside a level (forward) or
    // descending levels processed first for backward (level 0 = no deps)
    std::vector<int> cnt(nlev + 1, 0);
    for (int i = 0; i < ns; ++i) cnt[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
    std::vector<int> rows(ns);
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int i = 0; i < ns; ++i) rows[pos[lev[i]]++] = i;
    }
    int chunk = 0, ell = 0;
    out_lev.push_back(0);
    for (int l = 0; l < nlev; ++l) {
        for (int a = cnt[l]; a < cnt[l + 1]; a += kChunk) {
            int b = std::min(cnt[l + 1], a + kChunk);
            int K = 0;
            for (int p = a; p < b; ++p) {
                int i = rows[p];
                int c = 0;
                if (lower) {
                    for (int k = L.rp[i]; k < L.rp[i + 1] && L.ci[k] < i; ++k) ++c;
                } else {
                    c = L.rp[i + 1] - (diag[i] + 1);
                }
                K = std::max(K, c);
            }
            out_off.push_back(ell);
            out_K.push_back(K);
            for (int lane = 0; lane < kChunk; ++lane) {
                int p = a + lane;
                if (p < b) {
                    int i = rows[p];
                    out_row.push_back(i);
                    out_gid.push_back(gid_of_row[i]);
                    if (out_dsrc) out_dsrc->push_back(diag[i]);
                } else {
                    out_row.push_back(-1);
                    out_gid.push_back(-1);
                    if (out_dsrc) out_dsrc->push_back(-1);
                }
            }
            // ELL slices: entry k of all 64 lanes contiguous
            size_t base = out_col.size();
            out_col.resize(base + (size_t)K * kChunk, ns);
            out_src.resize(base + (size_t)K * kChunk, -1);
            for (int lane = 0; lane < kChunk && a + lane < b; ++lane) {
                int i = rows[a + lane];
                int k0 = lower ? L.rp[i] : diag[i] + 1;
                for (int e = 0; e < K; ++e) {
                    int k = k0 + e;
                    bool ok = lower ? (k < L.rp[i + 1] && L.ci[k] < i) : (k < L.rp[i + 1]);
                    if (!ok) break;
                    out_col[base + (size_t)e * kChunk + lane] = L.ci[k];
                    out_src[base + (size_t)e * kChunk + lane] = k;
                }
            }
            ell += K * kChunk;
            ++chunk;
        }
        out_lev.push_back(chunk);
    }
}

static void run_parallel(int count, const std::function<void(int)>& f) {
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    int nt = std::min<int>(count, (int)hw);
    if (nt <= 1) {
        for (int i = 0; i < count; ++i) f(i);
        return;
    }
    std::vector<std::thread> ts;
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> errs(nt);
    for (int t = 0; t < nt; ++t)
        ts.emplace_back([&, t] {
            try {
                for (int i = next++; i < count; i = next++) f(i);
            } catch (...) {
                errs[t] = std::current_exception();
            }
        });
    for (auto& t : ts) t.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------

__global__ void k_jacobi(int n, const double* __restrict__ inv, const double* __restrict__ x,
                         double* __restrict__ y, const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] = inv[i] * x[i];  // src/precond/jacobi.h:22-25
}

// IKJ ILU(0) (src/precond/ilu0.h:31-51), one workgroup per subdomain, rows of
// one level in parallel.  Per-row operation order equals the reference's.
__global__ void __launch_bounds__(kTrsvThreads)
k_ilu0_factor(const SubInfo* __restrict__ subs, const int* __restrict__ f_lev,
              const int* __restrict__ f_row, const int* __restrict__ lrp,
              const int* __restrict__ lci, const int* __restrict__ ldiag,
              double* __restrict__ lu) {
    const SubInfo si = subs[blockIdx.x];
    const int wave = threadIdx.x / kChunk, lane = threadIdx.x % kChunk;
    const int nw = blockDim.x / kChunk;
    for (int l = si.flev0; l < si.flev0 + si.nflev; ++l) {
        for (int c = f_lev[l] + wave; c < f_lev[l + 1]; c += nw) {
            const int row = f_row[c * kChunk + lane];
            if (row < 0) continue;
            const int gi = si.row0 + row;
            const int a = lrp[gi], e = lrp[gi + 1];
            for (int jp = a; jp < e; ++jp) {
                const int k = lci[jp];
                if (k >= row) break;
                const int dk = ldiag[si.row0 + k];
                if (dk < 0 || fabs(lu[dk]) < 1e-30) continue;
                const double aik = lu[jp] / lu[dk];
                lu[jp] = aik;
                const int ke = lrp[si.row0 + k + 1];
                for (int kp = dk + 1; kp < ke; ++kp) {
                    const int ck = lci[kp];
                    for (int ip = a; ip < e; ++ip) {
                        if (lci[ip] == ck) {
                            lu[ip] -= aik * lu[kp];
                            break;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void k_pack(int n, const int* __restrict__ src, const double* __restrict__ lu,
                       T* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int s = src[i];
        out[i] = s >= 0 ? (T)lu[s] : (T)0;
    }
}

// Fused restriction gather + unit-L sweep + U sweep + prolongation.
// MODE 0: restricted (RAS / global ILU0): y[out] = z (+ coarse term cz[s]).
// MODE 1: additive (ASM): zbuf[row0 + row] = z, summed by k_asm_gather.
template <typename T, bool ZLDS, int MODE>
__global__ void __launch_bounds__(kTrsvThreads)
k_schwarz_sweep(const SubInfo* __restrict__ subs, const int* __restrict__ order,
                const int* __restrict__ f_lev, const int* __restrict__ f_row,
                const int* __restrict__ f_gid, const int* __restrict__ f_off,
                const int* __restrict__ f_K, const int* __restrict__ f_col,
                const T* __restrict__ f_val, const int* __restrict__ b_lev,
                const int* __restrict__ b_row, const int* __restrict__ b_out,
                const int* __restrict__ b_off, const int* __restrict__ b_K,
                const int* __restrict__ b_col, const T* __restrict__ b_val,
                const T* __restrict__ b_diag, const double* __restrict__ x,
                double* __restrict__ y, T* __restrict__ zglob, T* __restrict__ zout,
                const T* __restrict__ cz, const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    extern __shared__ unsigned char lds_raw[];
    const int s = order[blockIdx.x];
    const SubInfo si = subs[s];
    T* z = ZLDS ? reinterpret_cast<T*>(lds_raw) : zglob + si.zoff;
    const int wave = threadIdx.x / kChunk, lane = threadIdx.x % kChunk;
    const int nw = blockDim.x / kChunk;
    if (threadIdx.x == 0) z[si.ns] = (T)0;  // zero slot for padding entries
    __syncthreads();
    // forward: z = L^{-1} R_s x
    for (int l = si.flev0; l < si.flev0 + si.nflev; ++l) {
        const int c1 = f_lev[l + 1];
        for (int c = f_lev[l] + wave; c < c1; c += nw) {
            const int q = c * kChunk + lane;
            const int row = f_row[q];
            const int off = f_off[c] + lane, K = f_K[c];
            if (row >= 0) {
                T t = (T)x[f_gid[q]];
                for (int k = 0; k < K; ++k) t -= f_val[off + k * kChunk] * z[f_col[off + k * kChunk]];
                z[row] = t;
            }
        }
        __syncthreads();
    }
    const T add = (MODE == 0 && cz) ? cz[s] : (T)0;
    // backward: z = U^{-1} z, then prolongation
    for (int l = si.blev0; l < si.blev0 + si.nblev; ++l) {
        const int c1 = b_lev[l + 1];
        for (int c = b_lev[l] + wave; c < c1; c += nw) {
            const int q = c * kChunk + lane;
            const int row = b_row[q];
            const int off = b_off[c] + lane, K = b_K[c];
            if (row >= 0) {
                T t = z[row];
                for (int k = 0; k < K; ++k) t -= b_val[off + k * kChunk] * z[b_col[off + k * kChunk]];
                t = t / b_diag[q];
                z[row] = t;
                if (MODE == 0) {
                    const int g = b_out[q];
                    if (g >= 0) {
                        if (cz) y[g] = (double)(t + add);
                        else y[g] = (double)t;
                    }
                } else {
                    zout[si.row0 + row] = t;
                }
            }
        }
        __syncthreads();
    }
}

// ASM prolongation: y[g] = sum over subdomains containing g, ascending, of z.
template <typename T>
__global__ void k_asm_gather(int n, const int* __restrict__ ptr, const int* __restrict__ pos,
                             const T* __restrict__ zbuf, const int* __restrict__ part,
                             const T* __restrict__ cz, double* __restrict__ y,
                             const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int k = ptr[g]; k < ptr[g + 1]; ++k) acc += (double)zbuf[pos[k]];
        if (cz) {
            if (sizeof(T) == 8) y[g] = acc + (double)cz[part[g]];
            else y[g] = (double)((T)acc + cz[part[g]]);
        } else {
            y[g] = acc;
        }
    }
}

// Coarse restriction r0[s] = sum_{k in owned_s} c_s x[k]
// (src/precond/coarse_space.h:63-70), one workgroup per subdomain.
template <typename T>
__global__ void __launch_bounds__(256)
k_coarse_restrict(const int* __restrict__ own_ptr, const int* __restrict__ own_idx,
                  const T* __restrict__ c, const double* __restrict__ x, T* __restrict__ r0,
                  const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    const int s = blockIdx.x;
    const T cs = c[s];
    T acc = 0;
    for (int k = own_ptr[s] + threadIdx.x; k < own_ptr[s + 1]; k += 256) acc += cs * (T)x[own_idx[k]];
    __shared__ T red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) r0[s] = red[0];
}

// Coarse solve z0 = A0^{-1} r0 and cz[s] = c_s z0[s].
// mode 0: pivots + forward unit-L + backward U with the LU factors
// (src/precond/coarse_solve.h:50-67), column-oriented, one workgroup.
// mode 1: one dense mat-vec with the precomputed inverse (stored transposed).
template <typename T>
__global__ void __launch_bounds__(1024)
k_coarse_solve(int nc, int mode, const T* __restrict__ LUt, const int* __restrict__ piv,
               const T* __restrict__ invT, const T* __restrict__ r0, const T* __restrict__ c,
               T* __restrict__ cz, const int* __restrict__ guard) {
    if (guard && *guard == 0) return;
    extern __shared__ unsigned char lds_raw[];
    T* xs = reinterpret_cast<T*>(lds_raw);
    for (int i = threadIdx.x; i < nc; i += blockDim.x) xs[i] = r0[i];
    __syncthreads();
    if (mode == 1) {
        for (int i = threadIdx.x; i < nc; i += blockDim.x) {
            T acc = 0;
            for (int j = 0; j < nc; ++j) acc += invT[(size_t)j * nc + i] * xs[j];
            cz[i] = c[i] * acc;
        }
        return;
    }
    for (int k = 0; k < nc; ++k) {
        if (threadIdx.x == 0 && piv[k] != k) {
            T t = xs[k];
            xs[k] = xs[piv[k]];
            xs[piv[k]] = t;
        }
        __syncthreads();
        const T xk = xs[k];
        for (int i = k + 1 + threadIdx.x; i < nc; i += blockDim.x) xs[i] -= LUt[(size_t)k * nc + i] * xk;
        __syncthreads();
    }
    for (int k = nc - 1; k >= 0; --k) {
        if (threadIdx.x == 0) xs[k] = xs[k] / LUt[(size_t)k * nc + k];
        __syncthreads();
        const T xk = xs[k];
        for (int i = threadIdx.x; i < k; i += blockDim.x) xs[i] -= LUt[(size_t)k * nc + i] * xk;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nc; i += blockDim.x) cz[i] = c[i] * xs[i];
}

// Right-looking LU with partial pivoting on the first maximum |.|
// (src/precond/coarse_solve.h:15-47), one workgroup, row-major A in place.
// Each element sees the same update sequence as the reference, so the
// factors match it bit for bit.
__global__ void __launch_bounds__(1024)
k_dense_lu(int n, double* __restrict__ A, int* __restrict__ piv, int* __restrict__ status) {
    __shared__ double bv[1024];
    __shared__ int bi[1024];
    __shared__ int prow;
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bidx = n;
        for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
            double v = fabs(A[(size_t)i * n + k]);
            if (v > best) { best = v; bidx = i; }
        }
        bv[threadIdx.x] = best;
        bi[threadIdx.x] = bidx;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                double ov = bv[threadIdx.x + w];
                int oi = bi[threadIdx.x + w];
                if (ov > bv[threadIdx.x] || (ov == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
                    bv[threadIdx.x] = ov;
                    bi[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            // first maximum: candidates with value equal to the max and the
            // smallest index; k itself wins ties (reference starts at k).
            double dk = fabs(A[(size_t)k * n + k]);
            int r = (bv[0] > dk) ? bi[0] : k;
            piv[k] = r;
            prow = r;
            if (!(bv[0] >= 1e-30) && !(dk >= 1e-30)) *status = 1;
        }
        __syncthreads();
        if (*status) return;
        const int r = prow;
        if (r != k)
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                double t = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)r * n + j];
                A[(size_t)r * n + j] = t;
            }
        __syncthreads();
        const double akk = A[(size_t)k * n + k];
        for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) A[(size_t)i * n + k] /= akk;
        __syncthreads();
        const int m = n - k - 1;
        for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
            int i = k + 1 + e / m, j = k + 1 + e % m;
            A[(size_t)i * n + j] -= A[(size_t)i * n + k] * A[(size_t)k * n + j];
        }
        __syncthreads();
    }
}

// Inverse from the LU factors: column j of A0^{-1} solves LU x = P e_j.
// Stored transposed (invT[j][i] = inv[i][j]) for coalesced mat-vec reads.
__global__ void __launch_bounds__(256)
k_lu_inverse(int n, const double* __restrict__ LU, const int* __restrict__ piv,
             double* __restrict__ invT) {
    extern __shared__ unsigned char lds_raw[];
    double* xs = reinterpret_cast<double*>(lds_raw);
    const int j = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = (i == j) ? 1.0 : 0.0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x == 0 && piv[k] != k) {
            double t = xs[k];
            xs[k] = xs[piv[k]];
            xs[piv[k]] = t;
        }
        __syncthreads();
        double xk = xs[k];
        for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) xs[i] -= LU[(size_t)i * n + k] * xk;
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        if (threadIdx.x == 0) xs[k] = xs[k] / LU[(size_t)k * n + k];
        __syncthreads();
        double xk = xs[k];
        for (int i = threadIdx.x; i < k; i += blockDim.x) xs[i] -= LU[(size_t)i * n + k] * xk;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) invT[(size_t)j * n + i] = xs[i];
}

template <typename Tin, typename Tout>
__global__ void k_convert(int n, const Tin* __restrict__ a, Tout* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        b[i] = (Tout)a[i];
}

__global__ void k_transpose(int n, const double* __restrict__ a, double* __restrict__ at) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
        int i = e / n, j = e % n;
        at[(size_t)j * n + i] = a[e];
    }
}

static int grid_for(long long n) {
    long long g = (n + 255) / 256;
    return (int)std::max<long long>(1, std::min<long long>(g, 16384));
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------

static void setup_jacobi(sg_precond* M, const sg_csr* A) {
    std::vector<double> inv(A->nrows);
    for (int i = 0; i < A->nrows; ++i) {
        double d = 0.0;
        for (int k = A->host.rp[i]; k < A->host.rp[i + 1]; ++k)
            if (A->host.ci[k] == i) { d = A->host.v[k]; break; }
        inv[i] = std::fabs(d) > 1e-15 ? 1.0 / d : 1.0;  // src/precond/jacobi.h:14-18
    }
    M->inv_diag.upload(inv, M->ctx->stream);
}

static void setup_ilu_batch(sg_precond* M, const sg_csr* A, bool restricted, bool additive) {
    sg_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const HostCsr& H = A->host;
    const int ns_total = M->nparts;
    std::vector<HostCsr> local(ns_total);
    std::vector<std::vector<int>> diag(ns_total);
    std::vector<SubPack> pk(ns_total);
    std::vector<std::vector<int>> owned_local(ns_total);
    const bool global_ilu = (M->kind == SG_PC_ILU0);

    run_parallel(ns_total, [&](int s) {
        const std::vector<int>& d = M->dofs[s];
        local[s] = global_ilu ? H : extract_submatrix(H, d);
        HostCsr& L = local[s];
        for (int i = 0; i < L.n; ++i)
            for (int k = L.rp[i] + 1; k < L.rp[i + 1]; ++k)
                if (L.ci[k] <= L.ci[k - 1])
                    throw std::runtime_error("ILU(0) needs strictly ascending column indices");
        diag[s] = diag_positions(L);
        for (int i = 0; i < L.n; ++i)
            if (diag[s][i] < 0) throw std::runtime_error("ILU(0) needs a diagonal entry in every row");
        std::vector<int> lf = lower_levels(L), lb = upper_levels(L);
        std::vector<int> gid(L.n), out(L.n, -1);
        // owned flags: ras.h:34-41 (owned DOFs mapped into subdomain_dofs)
        std::vector<char> own(L.n, 0);
        if (global_ilu) {
            for (int i = 0; i < L.n; ++i) { gid[i] = i; own[i] = 1; }
        } else {
            for (int i = 0; i < L.n; ++i) {
                gid[i] = d[i];
                own[i] = M->part[d[i]] == s;
            }
        }
        for (int i = 0; i < L.n; ++i) out[i] = own[i] ? gid[i] : -1;
        SubPack& P = pk[s];
        pack_sweep(L, diag[s], lf, true, gid, P.f_lev, P.f_row, P.f_gid, P.f_off, P.f_K, P.f_col,
                   P.f_src, nullptr);
        std::vector<int> dummy_gid;
        pack_sweep(L, diag[s], lb, false, out, P.b_lev, P.b_row, P.b_out, P.b_off, P.b_K, P.b_col,
                   P.b_src, &P.b_dsrc);
    });
    (void)restricted;

    // concatenate
    Plan pl;
    pl.subs.resize(ns_total);
    int row0 = 0, nnz0 = 0, zoff = 0, fchunk0 = 0, bchunk0 = 0, fell0 = 0, bell0 = 0;
    pl.lrp.push_back(0);
    M->sub_nnz.assign(ns_total, 0);
    for (int s = 0; s < ns_total; ++s) {
        HostCsr& L = local[s];
        SubPack& P = pk[s];
        SubInfo& si = pl.subs[s];
        si.ns = L.n;
        si.row0 = row0;
        si.zoff = zoff;
        si.flev0 = (int)pl.f_lev.size();
        si.nflev = (int)P.f_lev.size() - 1;
        for (int l = 0; l < si.nflev; ++l) pl.f_lev.push_back(P.f_lev[l] + fchunk0);
        si.blev0 = (int)pl.b_lev.size();
        si.nblev = (int)P.b_lev.size() - 1;
        for (int l = 0; l < si.nblev; ++l) pl.b_lev.push_back(P.b_lev[l] + bchunk0);
        // each level list needs its closing boundary: store [start..] and the
        // end of the last level as the start of the next subdomain's first
        // level; append explicit terminators instead to keep levels self-contained
        pl.f_lev.push_back(P.f_lev[si.nflev] + fchunk0);
        pl.b_lev.push_back(P.b_lev[si.nblev] + bchunk0);
        for (size_t c = 0; c < P.f_K.size(); ++c) {
            pl.f_off.push_back(P.f_off[c] + fell0);
            pl.f_K.push_back(P.f_K[c]);
        }
        for (size_t c = 0; c < P.b_K.size(); ++c) {
            pl.b_off.push_back(P.b_off[c] + bell0);
            pl.b_K.push_back(P.b_K[c]);
        }
        pl.f_row.insert(pl.f_row.end(), P.f_row.begin(), P.f_row.end());
        pl.f_gid.insert(pl.f_gid.end(), P.f_gid.begin(), P.f_gid.end());
        pl.b_row.insert(pl.b_row.end(), P.b_row.begin(), P.b_row.end());
        for (int g : P.b_out) pl.b_out.push_back(additive ? -1 : g);
        pl.f_col.insert(pl.f_col.end(), P.f_col.begin(), P.f_col.end());
        pl.b_col.insert(pl.b_col.end(), P.b_col.begin(), P.b_col.end());
        for (int k : P.f_src) pl.f_src.push_back(k >= 0 ? k + nnz0 : -1);
        for (int k : P.b_src) pl.b_src.push_back(k >= 0 ? k + nnz0 : -1);
        for (int k : P.b_dsrc) pl.b_dsrc.push_back(k >= 0 ? k + nnz0 : -1);
        for (int i = 0; i < L.n; ++i) {
            pl.lrp.push_back(L.rp[i + 1] + nnz0);
            pl.ldiag.push_back(diag[s][i] + nnz0);
        }
        pl.lci.insert(pl.lci.end(), L.ci.begin(), L.ci.end());
        pl.lval.insert(pl.lval.end(), L.v.begin(), L.v.end());
        M->sub_nnz[s] = L.nnz();
        pl.max_ns = std::max(pl.max_ns, L.n);
        row0 += L.n;
        nnz0 += L.nnz();
        zoff += L.n + 1;
        fchunk0 += (int)P.f_K.size();
        bchunk0 += (int)P.b_K.size();
        fell0 += P.f_col.size();
        bell0 += P.b_col.size();
        // the f_lev/b_lev arrays per subdomain have nflev+1 entries
    }
    pl.total_rows = row0;
    // queue: largest subdomains first
    pl.order.resize(ns_total);
    std::iota(pl.order.begin(), pl.order.end(), 0);
    std::stable_sort(pl.order.begin(), pl.order.end(),
                     [&](int a, int b) { return M->sub_nnz[a] > M->sub_nnz[b]; });

    M->nsub = ns_total;
    M->lrp_host = pl.lrp;
    M->d_subs.upload(pl.subs, st);
    M->d_order.upload(pl.order, st);
    M->f_lev.upload(pl.f_lev, st);
    M->f_row.upload(pl.f_row, st);
    M->f_gid.upload(pl.f_gid, st);
    M->f_off.upload(pl.f_off, st);
    M->f_K.upload(pl.f_K, st);
    M->f_col.upload(pl.f_col, st);
    M->f_src.upload(pl.f_src, st);
    M->b_lev.upload(pl.b_lev, st);
    M->b_row.upload(pl.b_row, st);
    M->b_out.upload(pl.b_out, st);
    M->b_off.upload(pl.b_off, st);
    M->b_K.upload(pl.b_K, st);
    M->b_col.upload(pl.b_col, st);
    M->b_src.upload(pl.b_src, st);
    M->b_dsrc.upload(pl.b_dsrc, st);
    M->lrp.upload(pl.lrp, st);
    M->lci.upload(pl.lci, st);
    M->ldiag.upload(pl.ldiag, st);
    M->lval.upload(pl.lval, st);

    // device factorisation in FP64
    hipLaunchKernelGGL(k_ilu0_factor, dim3(ns_total), dim3(kTrsvThreads), 0, st, M->d_subs.p,
                       M->f_lev.p, M->f_row.p, M->lrp.p, M->lci.p, M->ldiag.p, M->lval.p);
    SG_HIP(hipGetLastError());
    // pack values for the sweeps
    const bool fp32 = M->opts.precision == SG_FP32;
    int nf = (int)pl.f_src.size(), nb = (int)pl.b_src.size(), nd = (int)pl.b_dsrc.size();
    if (fp32) {
        M->f_v32.alloc(nf);
        M->b_v32.alloc(nb);
        M->b_d32.alloc(nd);
        if (nf) hipLaunchKernelGGL(k_pack<float>, dim3(grid_for(nf)), dim3(256), 0, st, nf, M->f_src.p, M->lval.p, M->f_v32.p);
        if (nb) hipLaunchKernelGGL(k_pack<float>, dim3(grid_for(nb)), dim3(256), 0, st, nb, M->b_src.p, M->lval.p, M->b_v32.p);
        if (nd) hipLaunchKernelGGL(k_pack<float>, dim3(grid_for(nd)), dim3(256), 0, st, nd, M->b_dsrc.p, M->lval.p, M->b_d32.p);
    } else {
        M->f_v64.alloc(nf);
        M->b_v64.alloc(nb);
        M->b_d64.alloc(nd);
        if (nf) hipLaunchKernelGGL(k_pack<double>, dim3(grid_for(nf)), dim3(256), 0, st, nf, M->f_src.p, M->lval.p, M->f_v64.p);
        if (nb) hipLaunchKernelGGL(k_pack<double>, dim3(grid_for(nb)), dim3(256), 0, st, nb, M->b_src.p, M->lval.p, M->b_v64.p);
        if (nd) hipLaunchKernelGGL(k_pack<double>, dim3(grid_for(nd)), dim3(256), 0, st, nd, M->b_dsrc.p, M->lval.p, M->b_d64.p);
    }
    SG_HIP(hipGetLastError());

    // z in LDS when the largest subdomain fits
    size_t zbytes = (size_t)(pl.max_ns + 1) * (fp32 ? 4 : 8);
    M->z_in_lds = zbytes <= (size_t)ctx->max_lds;
    M->lds_bytes = M->z_in_lds ? zbytes : 0;
    if (!M->z_in_lds) {
        if (fp32) M->z32.alloc((size_t)zoff);
        else M->z64.alloc((size_t)zoff);
    }
    if (additive) {
        if (fp32) M->zout32.alloc((size_t)pl.total_rows + 1);
        else M->zout64.alloc((size_t)pl.total_rows + 1);
    }
    if (M->z_in_lds) {
        const void* fns[4] = {(const void*)k_schwarz_sweep<float, true, 0>,
                              (const void*)k_schwarz_sweep<float, true, 1>,
                              (const void*)k_schwarz_sweep<double, true, 0>,
                              (const void*)k_schwarz_sweep<double, true, 1>};
        for (const void* f : fns)
            SG_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)M->lds_bytes));
    }

    if (additive) {
        // cover lists: for global g, positions row0_s + local index for every
        // subdomain containing g, ascending s (asm.h:54-58 single-thread order)
        std::vector<int> cnt(M->n + 1, 0);
        for (int s = 0; s < ns_total; ++s)
            for (int g : M->dofs[s]) cnt[g + 1]++;
        for (int g = 0; g < M->n; ++g) cnt[g + 1] += cnt[g];
        std::vector<int> pos(cnt[M->n]);
        std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
        for (int s = 0; s < ns_total; ++s)
            for (int i = 0; i < (int)M->dofs[s].size(); ++i)
                pos[fillp[M->dofs[s][i]]++] = pl.subs[s].row0 + i;
        M->cov_ptr.upload(cnt, st);
        M->cov_pos.upload(pos, st);
    }
}

// Coarse space of src/precond/coarse_space.h:26-60 in sparse form: row s of R0
// is c_s = 1/sqrt(|owned_s|) on owned_s.  A0[s][t] accumulates, for k in
// owned_s ascending, c_s * (sum over row k in CSR order of a_kj c_t for j owned
// by t) -- the nonzero terms of the reference's dense loops in their order.
static void setup_coarse(sg_precond* M, const sg_csr* A) {
    sg_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const HostCsr& H = A->host;
    int nc = M->nparts;
    M->nc = nc;
    std::vector<double> c(nc);
    for (int s = 0; s < nc; ++s) c[s] = 1.0 / std::sqrt((double)M->owned[s].size());
    M->A0.assign((size_t)nc * nc, 0.0);
    std::vector<double> aphi(nc, 0.0);
    std::vector<char> seen(nc, 0);
    std::vector<int> touched;
    for (int s = 0; s < nc; ++s)
        for (int k : M->owned[s]) {
            touched.clear();
            for (int j = H.rp[k]; j < H.rp[k + 1]; ++j) {
                int t = M->part[H.ci[j]];
                if (!seen[t]) { seen[t] = 1; touched.push_back(t); aphi[t] = 0.0; }
                aphi[t] += H.v[j] * c[t];
            }
            for (int t : touched) {
                M->A0[(size_t)s * nc + t] += c[s] * aphi[t];
                seen[t] = 0;
            }
        }
    // device LU (FP64) with partial pivoting
    DevBuf<double> dA;
    dA.upload(M->A0, st);
    M->d_piv.alloc(nc);
    M->d_lu_status.alloc(1);
    SG_HIP(hipMemsetAsync(M->d_lu_status.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_dense_lu, dim3(1), dim3(1024), 0, st, nc, dA.p, M->d_piv.p,
                       M->d_lu_status.p);
    SG_HIP(hipGetLastError());
    int status = 0;
    SG_HIP(hipMemcpyAsync(&status, M->d_lu_status.p, sizeof(int), hipMemcpyDeviceToHost, st));
    SG_HIP(hipStreamSynchronize(st));
    if (status) fail(SG_E_SINGULAR, "Coarse matrix is singular");
    M->LU.resize((size_t)nc * nc);
    M->piv.resize(nc);
    SG_HIP(hipMemcpyAsync(M->LU.data(), dA.p, sizeof(double) * nc * nc, hipMemcpyDeviceToHost, st));
    SG_HIP(hipMemcpyAsync(M->piv.data(), M->d_piv.p, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
    // transposed factors and inverse
    M->d_LUt64.alloc((size_t)nc * nc);
    hipLaunchKernelGGL(k_transpose, dim3(grid_for((long long)nc * nc)), dim3(256), 0, st, nc, dA.p,
                       M->d_LUt64.p);
    M->d_inv64.alloc((size_t)nc * nc);
    hipLaunchKernelGGL(k_lu_inverse, dim3(nc), dim3(256), (size_t)nc * 8, st, nc, dA.p,
                       M->d_piv.p, M->d_inv64.p);
    SG_HIP(hipGetLastError());
    M->d_c64.upload(c, st);
    // owned lists for the restriction
    std::vector<int> own_ptr(nc + 1, 0), own_idx;
    own_idx.reserve(M->n);
    for (int s = 0; s < nc; ++s) {
        own_idx.insert(own_idx.end(), M->owned[s].begin(), M->owned[s].end());
        own_ptr[s + 1] = (int)own_idx.size();
    }
    M->d_own_ptr.upload(own_ptr, st);
    M->d_own_idx.upload(own_idx, st);
    if (M->opts.precision == SG_FP32) {
        int nn = nc * nc;
        M->d_LUt32.alloc(nn);
        M->d_inv32.alloc(nn);
        M->d_c32.alloc(nc);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(nn)), dim3(256), 0, st, nn,
                           M->d_LUt64.p, M->d_LUt32.p);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(nn)), dim3(256), 0, st, nn,
                           M->d_inv64.p, M->d_inv32.p);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(1), dim3(256), 0, st, nc, M->d_c64.p,
                           M->d_c32.p);
        M->d_r0_32.alloc(nc);
        M->d_cz32.alloc(nc);
    } else {
        M->d_r0_64.alloc(nc);
        M->d_cz64.alloc(nc);
    }
    SG_HIP(hipGetLastError());
    SG_HIP(hipStreamSynchronize(st));
}

static void setup(sg_precond* M, const sg_csr* A) {
    sg_context* ctx = M->ctx;
    SG_HIP(hipSetDevice(ctx->device));
    if (A->nrows != A->ncols) fail(SG_E_ARG, "preconditioners need a square matrix");
    M->n = A->nrows;
    switch (M->kind) {
        case SG_PC_NONE:
            break;
        case SG_PC_JACOBI:
            setup_jacobi(M, A);
            break;
        case SG_PC_ILU0:
            M->nparts = 1;
            M->part.assign(M->n, 0);
            M->owned.assign(1, {});
            M->dofs.assign(1, {});
            M->owned[0].resize(M->n);
            std::iota(M->owned[0].begin(), M->owned[0].end(), 0);
            M->dofs[0] = M->owned[0];
            setup_ilu_batch(M, A, true, false);
            break;
        case SG_PC_ASM:
        case SG_PC_RAS:
        case SG_PC_TWOLEVEL_ASM:
        case SG_PC_TWOLEVEL_RAS: {
            int np = std::max(1, M->opts.nparts);
            if (np > M->n) fail(SG_E_PARTITION, "partitioning failed: more subdomains than rows");
            M->nparts = np;
            if (M->opts.partition) {
                M->part.assign(M->opts.partition, M->opts.partition + M->n);
                for (int p : M->part)
                    if (p < 0 || p >= np) fail(SG_E_PARTITION, "partition entry out of range");
            } else {
                M->part = partition_graph(A->host, np, M->opts.ufactor);
            }
            build_sets(A->host, M->part, np, std::max(0, M->opts.overlap), M->owned, M->dofs);
            for (int s = 0; s < np; ++s)
                if (M->owned[s].empty()) fail(SG_E_PARTITION, "partitioning failed: empty subdomain");
            bool additive = (M->kind == SG_PC_ASM || M->kind == SG_PC_TWOLEVEL_ASM);
            setup_ilu_batch(M, A, !additive, additive);
            if (M->kind == SG_PC_TWOLEVEL_ASM || M->kind == SG_PC_TWOLEVEL_RAS) setup_coarse(M, A);
            if (additive) {
                M->d_part.upload(M->part, ctx->stream);
            }
            break;
        }
        default:
            fail(SG_E_ARG, "unknown preconditioner kind");
    }
    SG_HIP(hipStreamSynchronize(ctx->stream));
    M->ready = true;
}

template <typename T, bool ZLDS, int MODE>
static void launch_sweep(sg_precond* M, const double* x, double* y, const int* guard,
                         const T* fv, const T* bv, const T* bd, T* zg, T* zo, const T* cz) {
    hipLaunchKernelGGL((k_schwarz_sweep<T, ZLDS, MODE>), dim3(M->nsub), dim3(kTrsvThreads),
                       ZLDS ? M->lds_bytes : 0, M->ctx->stream, M->d_subs.p, M->d_order.p,
                       M->f_lev.p, M->f_row.p, M->f_gid.p, M->f_off.p, M->f_K.p, M->f_col.p, fv,
                       M->b_lev.p, M->b_row.p, M->b_out.p, M->b_off.p, M->b_K.p, M->b_col.p, bv,
                       bd, x, y, zg, zo, cz, guard);
}

template <typename T>
static void apply_batch(sg_precond* M, const double* x, double* y, const int* guard,
                        const T* fv, const T* bv, const T* bd, T* zg, T* zo, const T* cz) {
    hipStream_t st = M->ctx->stream;
    bool additive = (M->kind == SG_PC_ASM || M->kind == SG_PC_TWOLEVEL_ASM);
    {
        ProfScope ps(M->ctx, SG_K_TRSV);
        if (additive) {
            if (M->z_in_lds) launch_sweep<T, true, 1>(M, x, y, guard, fv, bv, bd, zg, zo, cz);
            else launch_sweep<T, false, 1>(M, x, y, guard, fv, bv, bd, zg, zo, cz);
        } else {
            if (M->z_in_lds) launch_sweep<T, true, 0>(M, x, y, guard, fv, bv, bd, zg, zo, cz);
            else launch_sweep<T, false, 0>(M, x, y, guard, fv, bv, bd, zg, zo, cz);
        }
        SG_HIP(hipGetLastError());
    }
    if (additive) {
        hipLaunchKernelGGL(k_asm_gather<T>, dim3(grid_for(M->n)), dim3(256), 0, st, M->n,
                           M->cov_ptr.p, M->cov_pos.p, zo, M->d_part.p, cz, y, guard);
        SG_HIP(hipGetLastError());
    }
}

template <typename T>
static const T* coarse_launch(sg_precond* M, const double* x, const int* guard, const T* LUt,
                              const T* inv, const T* c, T* r0, T* cz) {
    hipStream_t st = M->ctx->stream;
    int nc = M->nc;
    hipLaunchKernelGGL(k_coarse_restrict<T>, dim3(nc), dim3(256), 0, st, M->d_own_ptr.p,
                       M->d_own_idx.p, c, x, r0, guard);
    hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(1024), (size_t)nc * sizeof(T), st, nc,
                       M->opts.coarse_solve, LUt, M->d_piv.p, inv, r0, c, cz, guard);
    SG_HIP(hipGetLastError());
    return cz;
}

}  // namespace sg

using namespace sg;

namespace sg {

void precond_apply_launch(sg_precond* M, const double* x, double* y, const int* guard) {
    sg_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    ProfScope ps(ctx, SG_K_PRECOND);
    switch (M->kind) {
        case SG_PC_NONE:
            SG_HIP(hipMemcpyAsync(y, x, sizeof(double) * M->n, hipMemcpyDeviceToDevice, st));
            return;
        case SG_PC_JACOBI:
            hipLaunchKernelGGL(k_jacobi, dim3(grid_for(M->n)), dim3(256), 0, st, M->n,
                               M->inv_diag.p, x, y, guard);
            SG_HIP(hipGetLastError());
            return;
        default:
            break;
    }
    if (M->fn) {
        // host callback (guard honoured on the host)
        if (guard) {
            int g = 0;
            SG_HIP(hipMemcpyAsync(&g, guard, sizeof(int), hipMemcpyDeviceToHost, st));
            SG_HIP(hipStreamSynchronize(st));
            if (!g) return;
        }
        if (M->device_pointers) {
            SG_HIP(hipStreamSynchronize(st));
            M->fn(M->user, x, y, M->n);
            return;
        }
        M->hx.resize(M->n);
        M->hy.resize(M->n);
        SG_HIP(hipMemcpyAsync(M->hx.data(), x, sizeof(double) * M->n, hipMemcpyDeviceToHost, st));
        SG_HIP(hipStreamSynchronize(st));
        M->fn(M->user, M->hx.data(), M->hy.data(), M->n);
        SG_HIP(hipMemcpyAsync(y, M->hy.data(), sizeof(double) * M->n, hipMemcpyHostToDevice, st));
        SG_HIP(hipStreamSynchronize(st));
        return;
    }
    const bool two = (M->kind == SG_PC_TWOLEVEL_ASM || M->kind == SG_PC_TWOLEVEL_RAS);
    if (M->opts.precision == SG_FP32) {
        const float* cz = two ? coarse_launch<float>(M, x, guard, M->d_LUt32.p, M->d_inv32.p,
                                                      M->d_c32.p, M->d_r0_32.p, M->d_cz32.p)
                              : nullptr;
        apply_batch<float>(M, x, y, guard, M->f_v32.p, M->b_v32.p, M->b_d32.p, M->z32.p, M->zout32.p, cz);
    } else {
        const double* cz = two ? coarse_launch<double>(M, x, guard, M->d_LUt64.p, M->d_inv64.p,
                                                        M->d_c64.p, M->d_r0_64.p, M->d_cz64.p)
                               : nullptr;
        apply_batch<double>(M, x, y, guard, M->f_v64.p, M->b_v64.p, M->b_d64.p, M->z64.p, M->zout64.p, cz);
    }
}

int precond_n(const sg_precond* M) { return M->n; }
sg_context* precond_ctx(const sg_precond* M) { return M->ctx; }
bool precond_ready(const sg_precond* M) { return M->ready; }

}  // namespace sg

extern "C" {

sg_status sg_precond_create(sg_context* ctx, sg_precond_kind kind, const sg_precond_opts* opts,
                            sg_precond** out) {
    if (!ctx || !out || (int)kind < 0 || (int)kind > SG_PC_TWOLEVEL_RAS) return SG_E_ARG;
    sg_precond* M = new sg_precond();
    M->ctx = ctx;
    M->kind = kind;
    if (opts) M->opts = *opts;
    else {
        M->opts.nparts = 4;
        M->opts.overlap = 1;
    }
    if (M->opts.precision != SG_FP64 && M->opts.precision != SG_FP32) {
        delete M;
        return SG_E_ARG;
    }
    M->opts.partition = nullptr;  // copied at setup
    if (opts && opts->partition) M->opts.partition = opts->partition;
    *out = M;
    return SG_OK;
}

sg_status sg_precond_create_callback(sg_context* ctx, int n, sg_apply_fn fn, void* user,
                                     int device_pointers, sg_precond** out) {
    if (!ctx || !out || !fn || n < 0) return SG_E_ARG;
    sg_precond* M = new sg_precond();
    M->ctx = ctx;
    M->kind = SG_PC_ILU0;  // placeholder kind; fn takes precedence
    M->fn = fn;
    M->user = user;
    M->device_pointers = device_pointers;
    M->n = n;
    M->ready = true;
    *out = M;
    return SG_OK;
}

sg_status sg_precond_setup(sg_precond* M, const sg_csr* A) {
    if (!M || !A) return SG_E_ARG;
    if (M->fn) return A->nrows == M->n ? SG_OK : SG_E_ARG;
    sg_status st = run_guarded(M->ctx, [&] { setup(M, A); });
    M->opts.partition = nullptr;  // the caller's array need not outlive setup
    return st;
}

sg_status sg_precond_apply(sg_precond* M, const double* d_x, double* d_y) {
    if (!M) return SG_E_ARG;
    if (!M->ready) {
        M->ctx->err = "preconditioner used before setup";
        return SG_E_STATE;
    }
    if (M->n && (!d_x || !d_y)) return SG_E_ARG;
    return run_guarded(M->ctx, [&] { precond_apply_launch(M, d_x, d_y, nullptr); });
}

sg_status sg_precond_destroy(sg_precond* M) {
    if (!M) return SG_E_ARG;
    (void)hipStreamSynchronize(M->ctx->stream);
    delete M;
    return SG_OK;
}

sg_status sg_precond_partition(const sg_precond* M, int* part) {
    if (!M || !part) return SG_E_ARG;
    if (!M->ready || M->part.empty()) return SG_E_STATE;
    std::copy(M->part.begin(), M->part.end(), part);
    return SG_OK;
}

sg_status sg_precond_subdomains(const sg_precond* M, int* total, int* ptr, int* idx) {
    if (!M || !total) return SG_E_ARG;
    if (!M->ready || M->dofs.empty()) return SG_E_STATE;
    int t = 0;
    if (ptr) ptr[0] = 0;
    for (size_t s = 0; s < M->dofs.size(); ++s) {
        if (idx) std::copy(M->dofs[s].begin(), M->dofs[s].end(), idx + t);
        t += (int)M->dofs[s].size();
        if (ptr) ptr[s + 1] = t;
    }
    *total = t;
    return SG_OK;
}

sg_status sg_precond_ilu_values(const sg_precond* M, int s, int* nnz, double* vals) {
    if (!M || !nnz) return SG_E_ARG;
    if (!M->ready || M->nsub == 0) return SG_E_STATE;
    if (s < 0 || s >= M->nsub) return SG_E_ARG;
    return run_guarded(M->ctx, [&] {
        int row0 = 0;
        for (int t = 0; t < s; ++t) row0 += (int)M->dofs[t].size();
        int ns = (int)M->dofs[s].size();
        int a = M->lrp_host[row0], b = M->lrp_host[row0 + ns];
        *nnz = b - a;
        if (vals) {
            SG_HIP(hipMemcpyAsync(vals, M->lval.p + a, sizeof(double) * (b - a),
                                  hipMemcpyDeviceToHost, M->ctx->stream));
            SG_HIP(hipStreamSynchronize(M->ctx->stream));
        }
    });
}

sg_status sg_precond_coarse(const sg_precond* M, int* nc, double* A0, double* LU, int* piv) {
    if (!M || !nc) return SG_E_ARG;
    if (!M->ready || M->nc == 0) return SG_E_STATE;
    *nc = M->nc;
    if (A0) std::copy(M->A0.begin(), M->A0.end(), A0);
    if (LU) std::copy(M->LU.begin(), M->LU.end(), LU);
    if (piv) std::copy(M->piv.begin(), M->piv.end(), piv);
    return SG_OK;
}

}  // extern "C"
