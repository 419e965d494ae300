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

#include <climits>

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <atomic>
#include <exception>
#include <numeric>
#include <type_traits>
#include <stdexcept>
#include <thread>

#ifndef SG_SF_DEPTH
#define SG_SF_DEPTH 3
#endif

namespace sg {

constexpr int kTrsvThreads = 512;   // workgroup of the factorisation
constexpr int kPadRows = 1024;      // padding of the sweep arrays (largest sweep workgroup)
// Sweep kernel shapes, all bit-identical: level-synchronous with 256/512/1024
// threads, and sync-free (per-row LDS flags, no barrier per level) with
// 512/1024 threads when z and the flags fit in LDS.
// depth: levels of row data in flight per thread in the barrier sweeps (the
// row data of level l + depth - 1 is loaded while level l computes)
// staged: factor data copied chunk by chunk into LDS ahead of the levels
// that use them (k_sweep_staged), 1024, 512 or 256 threads.
struct SweepShape {
    int threads;
    bool syncfree;
    int depth;
    bool staged;
};
constexpr int kNumSweepShapes = 10;
constexpr SweepShape kSweepShapes[kNumSweepShapes] = {
    {256, false, 3, false}, {512, false, 3, false}, {1024, false, 3, false},
    {512, true, 3, false},  {1024, true, 3, false}, {256, false, 6, false},
    {512, false, 6, false}, {1024, false, 3, true}, {512, false, 3, true},
    {256, false, 3, true}};
constexpr int kStageBytes = 32768;  // largest LDS buffer of the staged sweep (two are used)
constexpr int kKreg = 4;  // factor entries per row prefetched in registers
constexpr int kRestrictSplit = 8;  // workgroups per subdomain in the coarse restriction
constexpr int kSfDepth = SG_SF_DEPTH;  // items of row data in flight per wave (sync-free sweep)

struct SubInfo {
    int ns;      // rows of the (overlapped) subdomain
    int row0;    // first row in the concatenated local numbering
    int zoff;    // offset of z in the global scratch (ns + 1 slots)
    int flev0;   // first forward level (index into the level table)
    int nflev;
    int blev0;
    int nblev;
    int own0;    // first entry in the owned-row list
    int nown;
};

// One level of a sweep: rows [start, start + size) of the level-ordered row
// list; entry k of the row at position start + q sits at ell + k * size + q.
struct Level {
    int start, size, K, ell;
};

// A chunk of one sweep of one subdomain, staged into LDS as a unit: the
// level-ordered positions [row0, row0 + nrow), which cover levels [lev0, lev0 +
// nlev) of the subdomain's own level table (the first and last of them possibly
// in part: a level wider than a buffer spans chunks); bwd: backward sweep.
struct SChunk {
    int lev0, nlev, row0, nrow, bwd, img;  // img: offset in the chunk image, 16-byte units
    int pad[2];                            // 32 bytes: two 16-byte LDS words
};

// A chunk of the staged ring sweep: records at img, the rows' own values at own
// (both in 16-byte units; own values of a sweep's positions are written there
// by the gather (forward) or the forward sweep (backward)).
struct SChunkR {
    int lev0, nlev, row0, nrow, img, own, bwd, pad;
};

struct Plan {
    // host-side staging, moved to the device at setup
    std::vector<SubInfo> subs;
    std::vector<int> order;  // subdomains by descending nnz
    std::vector<Level> flev, blev;
    std::vector<int> f_row, f_col, f_src, f_bpos;
    std::vector<int> b_row, b_col, b_src, b_dsrc;
    std::vector<int> lrp, lci, ldiag;  // concatenated local CSR
    std::vector<double> lval;
    std::vector<int> dofs;   // global id of every local row, concatenated
    std::vector<int> own;    // owned local rows, per subdomain ascending
    std::vector<int> own_gid;  // global ids of those rows
    std::vector<int> f_item, b_item;  // per level: 64-row items of the earlier levels
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
    long long levels_fwd = 0, levels_bwd = 0;
    int max_levels = 0;
    // sweep plan (device)
    int nsub = 0;
    bool z_in_lds = false;
    size_t lds_bytes = 0;      // dynamic LDS of the sweep with z in LDS
    size_t lds_tab_bytes = 0;  // level tables only (16-byte aligned); 0 when in global memory
    bool tab_in_lds = true;
    // ring sweep (z too large for LDS): R slots, a full barrier every F levels
    bool ring = false;
    int ring_R = 0, ring_F = 0;
    int ring_kp = sg::kKreg;      // entries prefetched per row: 8 when some level has K > 4
    size_t ring_bytes = 0, zstride = 0;
    sg::DevBuf<int> f_bpos;
    int sweep_variant = 1;     // index into kSweepShapes
    bool sweep_tuned = false;
    // sync-free sweep: LDS = level tables | item prefixes | z | row flags
    bool sf_ok = false;
    int max_ns = 0;
    size_t sf_item_off = 0, sf_z_off = 0, sf_flag_off = 0, sf_bytes = 0;
    sg::DevBuf<int> d_fitem, d_bitem;
    // staged sweep: chunks of every subdomain (forward ones first); LDS =
    // level tables | z | chunk descriptors | two buffers of st_buf bytes
    bool st_ok = false;
    int st_kr = 4;  // record format: 4, or (FP32) 3 / 2 when every level has K <= 3 / 2
    int st_buf = 0;
    size_t st_ct_off = 0, st_buf_off = 0, st_bytes = 0;
    sg::DevBuf<sg::SChunk> d_chunks;
    sg::DevBuf<uint4> d_img;   // every chunk in its LDS layout, 16-byte aligned
    // diagnostic clock of the staged sweep: 12 values per workgroup while armed
    sg::DevBuf<long long> d_clk;
    bool clk_on = false;
    // staged ring sweep (vector too large for LDS): ring of R values + two
    // chunk buffers; chunks per subdomain: forward first/count, backward first/count
    bool sr_ok = false;
    bool sr_cr = false;  // compact two-word records (sring_k entries per row at most)
    int sr_R = 0, sr_buf = 0;
    size_t sr_ring_off = 0, sr_buf_off = 0, sr_bytes = 0;
    sg::DevBuf<sg::SChunkR> d_rchunks;
    sg::DevBuf<int> d_rchunk0;   // 4 ints per subdomain
    sg::DevBuf<uint4> d_rimg;    // records
    sg::DevBuf<uint4> d_rown;    // own values, per chunk segment
    sg::DevBuf<int> d_fslot, d_fgid;  // forward position -> own element index, global row
    sg::DevBuf<int> d_chunk0;  // per subdomain: first chunk, then the count (2 ints)
    sg::DevBuf<sg::SubInfo> d_subs;
    sg::DevBuf<int> d_order;
    sg::DevBuf<sg::Level> flev, blev;
    sg::DevBuf<int> f_row, f_col, f_src;
    sg::DevBuf<int> b_row, b_col, b_src, b_dsrc;
    sg::DevBuf<int> d_dofs, d_own, d_own_gid;
    sg::DevBuf<int> d_own2;  // (local row, global id) pairs of the owned rows (staged sweeps)
    sg::DevBuf<int> lrp, lci, ldiag;
    sg::DevBuf<double> lval;  // factored FP64 values in local CSR order
    long long ilut_l_nnz = 0, ilut_u_nnz = 0;  // sizes of the reference's L / U arrays
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
    sg::DevBuf<double> d_LUt64, d_inv64, d_c64, d_r0_64, d_cz64;
    sg::DevBuf<float> d_LUt32, d_inv32, d_c32, d_r0_32, d_cz32;
    sg::DevBuf<int> d_piv;
    int restrict_split = 1;
    sg::DevBuf<int> d_lu_status;
    // batch binding: subdomain -> system (-1: spans systems) for sys_off
    std::vector<int> bound_sys;
    sg::DevBuf<int> d_sub_sys;
    // learned FGMRES stage around a base preconditioner (not owned)
    sg::LearnedStage* learned = nullptr;
    ~sg_precond() {
        if (learned) sg::learned_free(learned);
    }
};

namespace sg {

// ---------------------------------------------------------------------------
// host: per-subdomain sweep packing
// ---------------------------------------------------------------------------

// Packs one sweep of a subdomain's local matrix L into level order.
// lower = true: the strictly-lower entries of each row in CSR order (unit-L
// forward sweep); lower = false: the entries right of the diagonal in CSR
// order plus the diagonal (backward sweep).  Rows of a level are ascending.
// Padding entries point at the zero slot (column ns) with source -1.
struct SweepPack {
    std::vector<Level> lev;  // start / ell relative to this subdomain
    std::vector<int> row, col, src, dsrc;
};

static void pack_levels(const HostCsr& L, const std::vector<int>& diag,
                        const std::vector<int>& lev, bool lower, SweepPack& P) {
    const int ns = L.n;
    int nlev = 0;
    for (int l : lev) nlev = std::max(nlev, l + 1);
    std::vector<int> cnt(nlev + 1, 0);
    for (int i = 0; i < ns; ++i) cnt[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
    std::vector<int> rows(ns);
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int i = 0; i < ns; ++i) rows[pos[lev[i]]++] = i;
    }
    auto first = [&](int i) { return lower ? L.rp[i] : diag[i] + 1; };
    auto count = [&](int i) {
        if (!lower) return L.rp[i + 1] - (diag[i] + 1);
        int c = 0;
        for (int k = L.rp[i]; k < L.rp[i + 1] && L.ci[k] < i; ++k) ++c;
        return c;
    };
    int ell = 0;
    for (int l = 0; l < nlev; ++l) {
        const int a = cnt[l], size = cnt[l + 1] - cnt[l];
        int K = 0;
        for (int q = 0; q < size; ++q) K = std::max(K, count(rows[a + q]));
        P.lev.push_back(Level{a, size, K, ell});
        P.col.resize((size_t)ell + (size_t)K * size, ns);
        P.src.resize((size_t)ell + (size_t)K * size, -1);
        for (int q = 0; q < size; ++q) {
            const int i = rows[a + q];
            P.row.push_back(i);
            if (!lower) P.dsrc.push_back(diag[i]);
            const int k0 = first(i), c = count(i);
            for (int e = 0; e < c; ++e) {
                P.col[(size_t)ell + (size_t)e * size + q] = L.ci[k0 + e];
                P.src[(size_t)ell + (size_t)e * size + q] = k0 + e;
            }
        }
        ell += K * size;
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
                         double* __restrict__ y, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] = inv[i] * x[i];  // src/precond/jacobi.h:22-25
}

// IKJ ILU(0) (src/precond/ilu0.h:31-51), one workgroup per subdomain, rows of
// one forward level in parallel.  Per-row operation order equals the
// reference's, so the factors match it bit for bit.
__global__ void __launch_bounds__(kTrsvThreads)
k_ilu0_factor(const SubInfo* __restrict__ subs, const Level* __restrict__ flev,
              const int* __restrict__ f_row, const int* __restrict__ lrp,
              const int* __restrict__ lci, const int* __restrict__ ldiag,
              double* __restrict__ lu) {
    const SubInfo si = subs[blockIdx.x];
    for (int l = si.flev0; l < si.flev0 + si.nflev; ++l) {
        const Level L = flev[l];
        for (int q = threadIdx.x; q < L.size; q += blockDim.x) {
            const int row = f_row[L.start + q];
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

// LDS flag bytes with an explicit address space: the accesses stay ds_read /
// ds_write (a generic pointer would make them flat accesses, whose waits also
// drain every global load in flight).
typedef __attribute__((address_space(3))) volatile unsigned char lds_flag_t;

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for all
// outstanding global loads (vmcnt(0)), which would drain the row-data prefetch
// at every level.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Row data of one thread for one level, loaded two levels ahead.  The raw
// loaded values are kept as they are; `use` (bit k: entry k is real, bit 31:
// the thread has a row) is applied only when the level is computed, so no
// instruction touches a prefetched register before it is needed.
template <typename T>
struct RowPre {
    int row;
    int col[kKreg];
    T val[kKreg];
    T diag;
    unsigned use;
};

// Branch-free load of thread q's row of level L: indices are clamped into the
// level (the row / ELL arrays carry kTrsvThreads padding entries at the end),
// so every load is unconditional and stays in flight across levels.
template <typename T, bool BWD>
__device__ __forceinline__ void row_prefetch(RowPre<T>& p, const Level& L, int q,
                                             const int* __restrict__ rows,
                                             const int* __restrict__ cols,
                                             const T* __restrict__ vals,
                                             const T* __restrict__ diags) {
    const bool live = q < L.size;
    const int qq = live ? q : 0;
    p.row = rows[L.start + qq];
    unsigned use = live ? 0x80000000u : 0u;
#pragma unroll
    for (int k = 0; k < kKreg; ++k) {
        const int idx = L.ell + (k < L.K ? k : 0) * L.size + qq;
        p.col[k] = cols[idx];
        p.val[k] = vals[idx];
        if (live && k < L.K) use |= 1u << k;
    }
    if (BWD) p.diag = diags[L.start + qq];
    p.use = use;
}

template <typename T, bool BWD>
__device__ __forceinline__ void row_compute(T* z, const RowPre<T>& p, const Level& L, int zslot,
                                            const int* __restrict__ rows,
                                            const int* __restrict__ cols,
                                            const T* __restrict__ vals,
                                            const T* __restrict__ diags) {
    const int t = threadIdx.x, B = blockDim.x;
    if (p.use & 0x80000000u) {
        T acc = z[p.row];
#pragma unroll
        for (int k = 0; k < kKreg; ++k) {
            const bool u = (p.use >> k) & 1u;
            acc -= (u ? p.val[k] : (T)0) * z[u ? p.col[k] : zslot];  // unused: - 0*0
        }
        for (int k = kKreg; k < L.K; ++k)
            acc -= vals[L.ell + k * L.size + t] * z[cols[L.ell + k * L.size + t]];
        if (BWD) acc = acc / p.diag;
        z[p.row] = acc;
    }
    for (int q = t + B; q < L.size; q += B) {  // levels wider than the block
        const int row = rows[L.start + q];
        T acc = z[row];
        for (int k = 0; k < L.K; ++k)
            acc -= vals[L.ell + k * L.size + q] * z[cols[L.ell + k * L.size + q]];
        if (BWD) acc = acc / diags[L.start + q];
        z[row] = acc;
    }
}

// One level-scheduled sweep over a subdomain: forward (unit L) or backward
// (U, division by the diagonal), z in place.  Each row subtracts its entries
// in CSR order, as src/precond/ilu0.h:62-76 does.  Row data is loaded D - 1
// levels ahead into a ring of D register buffers; the loop is unrolled by D so
// the ring never needs register moves (a move would wait for the loads in
// flight).  lev points at the subdomain's level table.
template <typename T, bool BWD, bool ZLDS, int D>
__device__ __forceinline__ void sweep(T* z, const Level* lev, int nlev, int zslot,
                                      const int* __restrict__ rows,
                                      const int* __restrict__ cols,
                                      const T* __restrict__ vals,
                                      const T* __restrict__ diags) {
    if (nlev == 0) return;
    const int t = threadIdx.x;
    const Level none{0, 0, 0, 0};
    auto level = [&](int l) { return l < nlev ? lev[l] : none; };
    auto level_barrier = [] {
        if (ZLDS) lds_barrier();
        else __syncthreads();
    };
    RowPre<T> b[D];
#pragma unroll
    for (int j = 0; j < D - 1; ++j) row_prefetch<T, BWD>(b[j], level(j), t, rows, cols, vals, diags);
    for (int l = 0; l < nlev; l += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            row_prefetch<T, BWD>(b[(j + D - 1) % D], level(l + j + D - 1), t, rows, cols, vals,
                                 diags);
            row_compute<T, BWD>(z, b[j], level(l + j), zslot, rows, cols, vals, diags);
            level_barrier();
            if (l + j + 1 >= nlev) return;
        }
    }
}

// R_s x of one subdomain: z[i] = x[d[i]], i < ns.  Every index and value
// load is issued unconditionally (indices past the end are clamped to the last
// row), so each thread keeps eight of each in flight.
template <int NT, typename T, typename ZP, int U = 8>
__device__ __forceinline__ void gather_x(const int* __restrict__ d, int ns,
                                         const double* __restrict__ x, ZP z) {
    for (int i0 = threadIdx.x; i0 < ns; i0 += U * NT) {
        int g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) g[u] = d[min(i0 + u * NT, ns - 1)];
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[g[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (i0 + u * NT < ns) z[i0 + u * NT] = (T)xv[u];
    }
}

// RAS prolongation of one subdomain: y[og[k]] = z[ow[k]] (+ add) over its
// owned rows, eight per thread in flight.
template <int NT, typename T, typename ZP>
__device__ __forceinline__ void scatter_owned(const int* __restrict__ ow,
                                              const int* __restrict__ og, int nown, ZP z,
                                              bool has_cz, T add, double* __restrict__ y) {
    constexpr int U = 8;
    for (int k0 = threadIdx.x; k0 < nown; k0 += U * NT) {
        int li[U], g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * NT;
            li[u] = k < nown ? ow[k] : -1;
            g[u] = k < nown ? og[k] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (g[u] >= 0) {
                const T v = z[li[u]];
                y[g[u]] = has_cz ? (double)(v + add) : (double)v;
            }
    }
}

// The same for the staged sweeps, from (local row, global id) pairs, U per
// thread in flight.  The caller loads the first batch (own_batch) before the
// last chunk's levels run, so that its latency is hidden; entries past nown
// (nown >= 1) repeat the last pair and are not stored.
typedef int own_i2 __attribute__((ext_vector_type(2)));
template <int NT, int U>
__device__ __forceinline__ void own_batch(own_i2 (&b)[U], const own_i2* __restrict__ o2, int nown,
                                          int k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) b[u] = o2[min(k0 + u * NT, nown - 1)];
}
template <int NT, int U, typename T, typename ZP>
__device__ __forceinline__ void scatter_owned2(own_i2 (&b)[U], const own_i2* __restrict__ o2,
                                               int nown, ZP z, bool has_cz, T add,
                                               double* __restrict__ y) {
    for (int k0 = threadIdx.x; k0 < nown; k0 += U * NT) {
        if (k0 != (int)threadIdx.x) own_batch<NT, U>(b, o2, nown, k0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u * NT < nown) {
                const T v = z[b[u].x];
                y[b[u].y] = has_cz ? (double)(v + add) : (double)v;
            }
    }
}

// Fused restriction gather + unit-L sweep + U sweep + prolongation, one
// workgroup per subdomain taken from a queue sorted by subdomain nonzeros.
// MODE 0: restricted (RAS / global ILU0): y[g] = z (+ coarse term cz[s]) for
//         the owned rows (src/precond/ras.h:55-68).
// MODE 1: additive (ASM): the whole local solution goes to zout, summed by
//         k_asm_gather (src/precond/asm.h:54-58).
template <typename T, bool ZLDS, int MODE, int NT, int D, bool TAB>
__global__ void __launch_bounds__(NT)
k_schwarz_sweep(const SubInfo* __restrict__ subs, const int* __restrict__ order,
                const Level* __restrict__ flev, const int* __restrict__ f_row,
                const int* __restrict__ f_col, const T* __restrict__ f_val,
                const Level* __restrict__ blev, const int* __restrict__ b_row,
                const int* __restrict__ b_col, const T* __restrict__ b_val,
                const T* __restrict__ b_diag, const int* __restrict__ dofs,
                const int* __restrict__ own, const int* __restrict__ own_gid,
                const double* __restrict__ x,
                double* __restrict__ y, T* __restrict__ zglob, T* __restrict__ zout,
                const T* __restrict__ cz, DevGuard guard, int lds_z_off) {
    extern __shared__ unsigned char lds_raw[];
    const int s = order[blockIdx.x];
    if (guard_stop(guard, s)) return;
    const SubInfo si = subs[s];
    // LDS: this subdomain's level tables (TAB; else they are too long for LDS
    // and are read from global memory), then (ZLDS) the solution vector.  The
    // choice is a template parameter so every LDS pointer is known as one and
    // its accesses stay ds_read / ds_write (a pointer that may be either would
    // make them flat accesses, whose waits also drain the global prefetch).
    const Level* lf = flev + si.flev0;
    const Level* lb = blev + si.blev0;
    if (TAB) {
        Level* tf = reinterpret_cast<Level*>(lds_raw);
        Level* tb = tf + si.nflev;
        for (int l = threadIdx.x; l < si.nflev; l += blockDim.x) tf[l] = flev[si.flev0 + l];
        for (int l = threadIdx.x; l < si.nblev; l += blockDim.x) tb[l] = blev[si.blev0 + l];
        lf = tf;
        lb = tb;
    }
    T* z = ZLDS ? reinterpret_cast<T*>(lds_raw + lds_z_off) : zglob + si.zoff;
    gather_x<NT, T>(dofs + si.row0, si.ns, x, z);
    if (threadIdx.x == 0) z[si.ns] = (T)0;  // zero slot for padding entries
    __syncthreads();
    sweep<T, false, ZLDS, D>(z, lf, si.nflev, si.ns, f_row, f_col, f_val, nullptr);
    __syncthreads();
    sweep<T, true, ZLDS, D>(z, lb, si.nblev, si.ns, b_row, b_col, b_val, b_diag);
    __syncthreads();
    if (MODE == 0) {
        scatter_owned<NT>(own + si.own0, own_gid + si.own0, si.nown, z, cz != nullptr,
                          cz ? cz[s] : (T)0, y);
    } else {
        for (int i = threadIdx.x; i < si.ns; i += blockDim.x) zout[si.row0 + i] = z[i];
    }
}

// ---- ring sweep: z too large for LDS --------------------------------------------
// For subdomains whose solution vector does not fit in LDS, the recently
// computed values live in an LDS ring of R entries indexed by the row's
// position in level order (slot = position mod R, slot R holds zero for
// padding).  Column references were encoded on the host: c >= 0 is a ring
// slot still live when the row's level runs, c < 0 is ~index into the global
// copy (forward results in backward order, zb; backward results by local row,
// zr).  Every value a row reads from global memory was written at least F
// levels earlier, and a full barrier (global stores complete) closes every
// F-th level, so those reads see it; the other levels end with an LDS-only
// barrier.  The own value of a row comes from a level-ordered array (gathered
// input zf for the forward sweep, zb for the backward one) and is prefetched
// with the row data, so nothing on the level-to-level chain waits for global
// memory.  Operation order per row equals the ILU(0) sweep's.
template <typename T, int KP>
struct RowPreR {
    int row;  // forward: backward position of the row; backward: the local row
    int col[KP];
    T val[KP];
    T diag;
    T own;
    unsigned use;
};

template <typename T, bool BWD, int KP>
__device__ __forceinline__ void row_prefetch_r(RowPreR<T, KP>& p, const Level& L, int q, int pos0,
                                               const int* __restrict__ rows,
                                               const int* __restrict__ cols,
                                               const T* __restrict__ vals,
                                               const T* __restrict__ diags,
                                               const int* __restrict__ dst,
                                               const T* __restrict__ ownv) {
    const bool live = q < L.size;
    const int qq = live ? q : 0;
    p.row = BWD ? rows[L.start + qq] : dst[L.start + qq];
    p.own = ownv[L.start - pos0 + qq];
    unsigned use = live ? 0x80000000u : 0u;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int idx = L.ell + (k < L.K ? k : 0) * L.size + qq;
        p.col[k] = cols[idx];
        p.val[k] = vals[idx];
        if (live && k < L.K) use |= 1u << k;
    }
    if (BWD) p.diag = diags[L.start + qq];
    p.use = use;
}

template <typename T>
using lds_t = __attribute__((address_space(3))) T;

// The two loads stay apart (an LDS-typed and a global pointer), so the common
// LDS read is a ds_read.  GLOB = false (no reference leaves the ring in this
// preconditioner) drops the global branch: its wait would drain the prefetch.
template <typename T, bool GLOB>
__device__ __forceinline__ T ring_read(const lds_t<T>* ring, const T* __restrict__ zglob, int c) {
    if (!GLOB || c >= 0) return ring[c];
    return zglob[~c];
}

template <typename T, bool BWD, bool GLOB, int KP>
__device__ __forceinline__ void row_compute_r(lds_t<T>* ring, const T* zdep, T* zout,
                                              const RowPreR<T, KP>& p, const Level& L, int pos0,
                                              int R, const int* __restrict__ rows,
                                              const int* __restrict__ cols,
                                              const T* __restrict__ vals,
                                              const T* __restrict__ diags,
                                              const int* __restrict__ dst,
                                              const T* __restrict__ ownv) {
    const int t = threadIdx.x, B = blockDim.x;
    int slot0 = (L.start - pos0) % R;
    if (p.use & 0x80000000u) {
        T acc = p.own;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const bool u = (p.use >> k) & 1u;
            acc -= (u ? p.val[k] : (T)0) * ring_read<T, GLOB>(ring, zdep, u ? p.col[k] : R);
        }
        for (int k = KP; k < L.K; ++k)
            acc -= vals[L.ell + k * L.size + t] *
                   ring_read<T, GLOB>(ring, zdep, cols[L.ell + k * L.size + t]);
        if (BWD) acc = acc / p.diag;
        int sl = slot0 + t;
        if (sl >= R) sl -= R;
        ring[sl] = acc;
        zout[p.row] = acc;
    }
    for (int q = t + B; q < L.size; q += B) {  // levels wider than the block
        T acc = ownv[L.start - pos0 + q];
        for (int k = 0; k < L.K; ++k)
            acc -= vals[L.ell + k * L.size + q] *
                   ring_read<T, GLOB>(ring, zdep, cols[L.ell + k * L.size + q]);
        if (BWD) acc = acc / diags[L.start + q];
        int sl = slot0 + q;
        while (sl >= R) sl -= R;
        ring[sl] = acc;
        zout[BWD ? rows[L.start + q] : dst[L.start + q]] = acc;
    }
}

template <typename T, bool BWD, int D, bool GLOB, int KP>
__device__ __forceinline__ void sweep_r(lds_t<T>* ring, const T* zdep, T* zout, const Level* lev,
                                        int nlev, int pos0, int R, int F,
                                        const int* __restrict__ rows,
                                        const int* __restrict__ cols,
                                        const T* __restrict__ vals,
                                        const T* __restrict__ diags,
                                        const int* __restrict__ dst,
                                        const T* __restrict__ ownv) {
    if (nlev == 0) return;
    const int t = threadIdx.x;
    auto level_barrier = [&](int l) {
        if (GLOB && (l + 1) % F == 0) __syncthreads();
        else lds_barrier();
    };
    // prefetches past the last level read position pos0 of an empty level
    auto level = [&](int l) { return l < nlev ? lev[l] : Level{pos0, 0, 0, 0}; };
    RowPreR<T, KP> b[D];
#pragma unroll
    for (int j = 0; j < D - 1; ++j)
        row_prefetch_r<T, BWD, KP>(b[j], level(j), t, pos0, rows, cols, vals, diags, dst, ownv);
    for (int l = 0; l < nlev; l += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            row_prefetch_r<T, BWD, KP>(b[(j + D - 1) % D], level(l + j + D - 1), t, pos0, rows,
                                   cols, vals, diags, dst, ownv);
            row_compute_r<T, BWD, GLOB, KP>(ring, zdep, zout, b[j], level(l + j), pos0, R, rows,
                                        cols, vals, diags, dst, ownv);
            level_barrier(l + j);
            if (l + j + 1 >= nlev) return;
        }
    }
}

// Gather into forward level order, forward sweep (results to zb in backward
// order), backward sweep (results to zr by local row), prolongation as
// k_schwarz_sweep does.  z3 holds zf | zb | zr, each ns + 1 per subdomain.
// KP: factor entries per row prefetched with the row (4, or 8 for ILUT
// factors); entries past KP are read inside their level.
template <typename T, int MODE, int NT, int D, bool TAB, int KP = kKreg>
__global__ void __launch_bounds__(NT)
k_schwarz_sweep_ring(const SubInfo* __restrict__ subs, const int* __restrict__ order,
                     const Level* __restrict__ flev, const int* __restrict__ f_row,
                     const int* __restrict__ f_col, const T* __restrict__ f_val,
                     const int* __restrict__ f_bpos, const Level* __restrict__ blev,
                     const int* __restrict__ b_row, const int* __restrict__ b_col,
                     const T* __restrict__ b_val, const T* __restrict__ b_diag,
                     const int* __restrict__ dofs, const int* __restrict__ own,
                     const int* __restrict__ own_gid, const double* __restrict__ x,
                     double* __restrict__ y, T* __restrict__ z3, size_t zstride,
                     T* __restrict__ zout, const T* __restrict__ cz, DevGuard guard,
                     int lds_ring_off, int R, int F) {
    extern __shared__ unsigned char lds_raw[];
    const int s = order[blockIdx.x];
    if (guard_stop(guard, s)) return;
    const SubInfo si = subs[s];
    const Level* lf = flev + si.flev0;
    const Level* lb = blev + si.blev0;
    if (TAB) {
        Level* tf = reinterpret_cast<Level*>(lds_raw);
        Level* tb = tf + si.nflev;
        for (int l = threadIdx.x; l < si.nflev; l += blockDim.x) tf[l] = flev[si.flev0 + l];
        for (int l = threadIdx.x; l < si.nblev; l += blockDim.x) tb[l] = blev[si.blev0 + l];
        lf = tf;
        lb = tb;
    }
    lds_t<T>* ring = (lds_t<T>*)(lds_raw + lds_ring_off);
    T* zf = z3 + si.zoff;
    T* zb = zf + zstride;
    T* zr = zb + zstride;
    const int pf0 = si.nflev ? flev[si.flev0].start : 0;
    const int pb0 = si.nblev ? blev[si.blev0].start : 0;
    const int* d = dofs + si.row0;
    // R_s x in forward level order: zf[p] = x[dofs[row(p)]], eight positions
    // per thread in flight (every load issued; positions past the end are
    // clamped to the last one)
    constexpr int U = 8;
    for (int p0 = threadIdx.x; p0 < si.ns; p0 += U * NT) {
        int r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) r[u] = f_row[pf0 + min(p0 + u * NT, si.ns - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u) r[u] = d[r[u]];
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[r[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (p0 + u * NT < si.ns) zf[p0 + u * NT] = (T)xv[u];
    }
    if (threadIdx.x == 0) ring[R] = (T)0;
    __syncthreads();
    if (F == INT_MAX) {  // every reference stays in the ring
        sweep_r<T, false, D, false, KP>(ring, zb, zb, lf, si.nflev, pf0, R, F, f_row, f_col, f_val,
                                    nullptr, f_bpos, zf);
        __syncthreads();
        sweep_r<T, true, D, false, KP>(ring, zr, zr, lb, si.nblev, pb0, R, F, b_row, b_col, b_val,
                                   b_diag, nullptr, zb);
    } else {
        sweep_r<T, false, D, true, KP>(ring, zb, zb, lf, si.nflev, pf0, R, F, f_row, f_col, f_val,
                                   nullptr, f_bpos, zf);
        __syncthreads();
        sweep_r<T, true, D, true, KP>(ring, zr, zr, lb, si.nblev, pb0, R, F, b_row, b_col, b_val,
                                  b_diag, nullptr, zb);
    }
    __syncthreads();
    if (MODE == 0) {
        scatter_owned<NT>(own + si.own0, own_gid + si.own0, si.nown, zr, cz != nullptr,
                          cz ? cz[s] : (T)0, y);
    } else {
        for (int i = threadIdx.x; i < si.ns; i += blockDim.x) zout[si.row0 + i] = zr[i];
    }
}

// ---- sync-free sweep ---------------------------------------------------------
// The rows of a sweep, in level order, are cut into items of at most 64 rows
// of one level (item0[l] = items of the levels before l).  Wave w of the
// workgroup runs items w, w + NW, w + 2 NW, ...; a row first waits until every
// row it reads has raised its LDS flag to `tgt`, then computes exactly as the
// level-synchronous sweep does (same entries, same order), stores z and raises
// its own flag.  An item only reads rows of earlier levels, so the smallest
// unfinished item can always proceed: no deadlock, and no barrier per level.
// Row data is prefetched two items ahead per wave.
struct ItemPos {
    Level L;
    int q;
};

__device__ __forceinline__ ItemPos locate_item(const Level* lev, const int* item0, int nlev,
                                               int nitems, int it, int& cur) {
    ItemPos p;
    if (it >= nitems) {
        p.L = Level{0, 0, 0, 0};
        p.q = 0;
        return p;
    }
    while (cur + 1 < nlev && item0[cur + 1] <= it) ++cur;
    p.L = lev[cur];
    p.q = (it - item0[cur]) * 64 + (threadIdx.x & 63);
    return p;
}

template <typename T, bool BWD>
__device__ __forceinline__ void row_compute_sf(T* z, lds_flag_t* fl,
                                               const RowPre<T>& p, const ItemPos& ip,
                                               int zslot, unsigned char tgt,
                                               const int* __restrict__ cols,
                                               const T* __restrict__ vals) {
    if (!(p.use & 0x80000000u)) return;
    const Level& L = ip.L;
    for (;;) {
        bool ok = true;
#pragma unroll
        for (int k = 0; k < kKreg; ++k)
            if ((p.use >> k) & 1u) ok = ok && fl[p.col[k]] >= tgt;
        for (int k = kKreg; k < L.K && ok; ++k) ok = fl[cols[L.ell + k * L.size + ip.q]] >= tgt;
        if (ok) break;
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    T acc = z[p.row];
#pragma unroll
    for (int k = 0; k < kKreg; ++k) {
        const bool u = (p.use >> k) & 1u;
        acc -= (u ? p.val[k] : (T)0) * z[u ? p.col[k] : zslot];  // unused: - 0*0
    }
    for (int k = kKreg; k < L.K; ++k)
        acc -= vals[L.ell + k * L.size + ip.q] * z[cols[L.ell + k * L.size + ip.q]];
    if (BWD) acc = acc / p.diag;
    z[p.row] = acc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    fl[p.row] = tgt;
}

template <typename T, bool BWD, int NT, int D>
__device__ __forceinline__ void sweep_syncfree(T* z, lds_flag_t* fl, const Level* lev,
                                               const int* item0, int nlev, int zslot,
                                               unsigned char tgt, const int* __restrict__ rows,
                                               const int* __restrict__ cols,
                                               const T* __restrict__ vals,
                                               const T* __restrict__ diags) {
    if (nlev == 0) return;
    constexpr int NW = NT / 64;
    const int wv = threadIdx.x >> 6;
    const int nitems = item0[nlev - 1] + (lev[nlev - 1].size + 63) / 64;
    int cur = 0;
    // ring of D buffers: item it + d*NW lives in buffer d (mod D); every index
    // is a compile-time constant after unrolling, so no register moves
    ItemPos pos[D];
    RowPre<T> buf[D];
#pragma unroll
    for (int d = 0; d < D - 1; ++d) {
        pos[d] = locate_item(lev, item0, nlev, nitems, wv + d * NW, cur);
        row_prefetch<T, BWD>(buf[d], pos[d].L, pos[d].q, rows, cols, vals, diags);
    }
    for (int it = wv; it < nitems; it += D * NW) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int ci = it + d * NW;
            if (ci >= nitems) break;
            const int nb = (d + D - 1) % D;
            pos[nb] = locate_item(lev, item0, nlev, nitems, ci + (D - 1) * NW, cur);
            row_prefetch<T, BWD>(buf[nb], pos[nb].L, pos[nb].q, rows, cols, vals, diags);
            row_compute_sf<T, BWD>(z, fl, buf[d], pos[d], zslot, tgt, cols, vals);
        }
    }
}

// The fused gather + sweeps + prolongation of k_schwarz_sweep (z in LDS) with
// sync-free sweeps.  LDS: level tables | item prefixes | z | row flags.
template <typename T, int MODE, int NT, int D = kSfDepth>
__global__ void __launch_bounds__(NT)
k_sweep_syncfree(const SubInfo* __restrict__ subs, const int* __restrict__ order,
                 const Level* __restrict__ flev, const int* __restrict__ fitem,
                 const int* __restrict__ f_row, const int* __restrict__ f_col,
                 const T* __restrict__ f_val, const Level* __restrict__ blev,
                 const int* __restrict__ bitem, const int* __restrict__ b_row,
                 const int* __restrict__ b_col, const T* __restrict__ b_val,
                 const T* __restrict__ b_diag, const int* __restrict__ dofs,
                 const int* __restrict__ own, const int* __restrict__ own_gid,
                 const double* __restrict__ x, double* __restrict__ y, T* __restrict__ zout,
                 const T* __restrict__ cz, DevGuard guard, int item_off, int z_off,
                 int flag_off) {
    extern __shared__ unsigned char lds_raw[];
    const int s = order[blockIdx.x];
    if (guard_stop(guard, s)) return;
    const SubInfo si = subs[s];
    Level* lf = reinterpret_cast<Level*>(lds_raw);
    Level* lb = lf + si.nflev;
    int* itf = reinterpret_cast<int*>(lds_raw + item_off);
    int* itb = itf + si.nflev;
    T* z = reinterpret_cast<T*>(lds_raw + z_off);
    lds_flag_t* fl = (lds_flag_t*)(lds_raw + flag_off);
    for (int l = threadIdx.x; l < si.nflev; l += NT) {
        lf[l] = flev[si.flev0 + l];
        itf[l] = fitem[si.flev0 + l];
    }
    for (int l = threadIdx.x; l < si.nblev; l += NT) {
        lb[l] = blev[si.blev0 + l];
        itb[l] = bitem[si.blev0 + l];
    }
    gather_x<NT, T>(dofs + si.row0, si.ns, x, z);
    for (int i = threadIdx.x; i < si.ns; i += NT) fl[i] = 0;
    if (threadIdx.x == 0) {
        z[si.ns] = (T)0;  // zero slot for padding entries, always ready
        fl[si.ns] = 3;
    }
    __syncthreads();
    sweep_syncfree<T, false, NT, D>(z, fl, lf, itf, si.nflev, si.ns, 1, f_row, f_col, f_val,
                                    nullptr);
    __syncthreads();
    sweep_syncfree<T, true, NT, D>(z, fl, lb, itb, si.nblev, si.ns, 2, b_row, b_col, b_val, b_diag);
    __syncthreads();
    if (MODE == 0) {
        scatter_owned<NT>(own + si.own0, own_gid + si.own0, si.nown, z, cz != nullptr,
                          cz ? cz[s] : (T)0, y);
    } else {
        for (int i = threadIdx.x; i < si.ns; i += NT) zout[si.row0 + i] = z[i];
    }
}

// ---- staged sweep: factor data through LDS, chunk by chunk ---------------------
// For subdomains whose solution vector fits in LDS (at most 65534 rows) and
// whose levels have at most kKreg entries per row.  Setup packs every row of
// a sweep into one record in level order (16-bit local row and columns, the
// values, the diagonal: 32 bytes in FP32, 64 in FP64) and cuts each sweep into
// chunks of consecutive levels whose records fit one LDS buffer.  While the
// levels of chunk c run entirely out of LDS (an LDS-only barrier closes each
// level), chunks c + 1 and c + 2 are in flight as 16-byte loads; chunk c + 1 is
// written to the other buffer afterwards.  A row's data is two (FP64: four)
// 16-byte LDS reads, issued a level ahead.  Rows and the operations of a row
// run in the order of the level-synchronous sweep, so the values are the same
// bit for bit.

// 16 bytes as clang vector types (valid in every address space, loads and
// stores of one dwordx4 / ds_read_b128 / ds_write_b128).
typedef unsigned stage_q __attribute__((ext_vector_type(4)));
typedef int stage_i4 __attribute__((ext_vector_type(4)));
typedef unsigned stage_d __attribute__((ext_vector_type(2)));

// 16-byte words per row record: [row | c0<<16, c1 | c2<<16, c3, diag (FP32)]
// then the four values (FP32: one quad; FP64: two quads, then the diagonal).
template <typename T>
__host__ __device__ constexpr int stage_rec_quads() {
    return sizeof(T) == 4 ? 2 : 4;
}
// Record formats (KR).  KR = 4, any T: the records above.  KR = 3, FP32 only,
// when every level of the plan has at most three entries per row: per chunk an
// index block of 8 bytes per row [row | c0<<16, c1 | c2<<16] (padded to 16
// bytes), then a value block of 16 bytes per row [v0, v1, v2, diag]: 24 instead
// of 32 bytes per row, and one z read and one product fewer (the dropped
// padding term subtracts +0, so the values do not change).
// KR = 2 (FP32, at most two entries per row): forward rows are one 16-byte
// word [row | c0<<16, c1, v0, v1]; backward chunks use the KR = 3 blocks with
// values [v0, v1, 0, diag].
template <typename T, int KR>
__host__ __device__ inline int stage_quads_n(int nrow, bool bwd) {
    return KR == 3 || (KR == 2 && bwd) ? (nrow + 1) / 2 + nrow
         : KR == 2                     ? nrow
                                       : nrow * stage_rec_quads<T>();
}

template <typename T, int NT>
struct StageRegs {
    static constexpr int kV = kStageBytes / 16 / NT;  // 16-byte words per thread
    stage_q v[kV];
};

// Issues the 16-byte loads of chunk c from its image (nothing touches the
// registers until stage_store); indices past the chunk are clamped to its
// last word, so every load is issued and the waits stay exact.
template <typename T, int NT, int KR>
__device__ __forceinline__ void stage_load(StageRegs<T, NT>& R, const SChunk& c,
                                           const stage_q* __restrict__ img) {
    const __attribute__((address_space(1))) stage_q* src =
        (const __attribute__((address_space(1))) stage_q*)(img + c.img);
    const int n = stage_quads_n<T, KR>(c.nrow, c.bwd);
#pragma unroll
    for (int i = 0; i < StageRegs<T, NT>::kV; ++i)
        R.v[i] = __builtin_nontemporal_load(&src[min((int)threadIdx.x + i * NT, n - 1)]);
}

// Words past the chunk go to a trash word (straight-line code: a branch here
// would make the compiler wait for every load in flight, the other set's too).
template <typename T, int NT, int KR>
__device__ __forceinline__ void stage_store(const StageRegs<T, NT>& R, const SChunk& c,
                                            lds_t<stage_q>* buf, lds_t<stage_q>* trash) {
    const int n = stage_quads_n<T, KR>(c.nrow, c.bwd);
#pragma unroll
    for (int i = 0; i < StageRegs<T, NT>::kV; ++i) {
        const int u = (int)threadIdx.x + i * NT;
        *(u < n ? buf + u : trash) = R.v[i];
    }
}

// Setup: packs the rows of every chunk into records (one workgroup per chunk,
// one thread per row).  Entries past a row's count are column zs (the zero
// slot of z) with value 0, as in the other sweeps' padding.
template <typename T, int KR>
__global__ void k_stage_image(const SChunk* __restrict__ chunks, const SubInfo* __restrict__ subs,
                              const int* __restrict__ chunk_sub, const Level* __restrict__ flev,
                              const Level* __restrict__ blev, const int* __restrict__ f_row,
                              const int* __restrict__ f_col, const T* __restrict__ f_val,
                              const int* __restrict__ b_row, const int* __restrict__ b_col,
                              const T* __restrict__ b_val, const T* __restrict__ b_diag,
                              uint4* __restrict__ img) {
    const SChunk c = chunks[blockIdx.x];
    const SubInfo si = subs[chunk_sub[blockIdx.x]];
    const Level* lev = c.bwd ? blev + si.blev0 : flev + si.flev0;
    const int* rows = c.bwd ? b_row : f_row;
    const int* cols = c.bwd ? b_col : f_col;
    const T* vals = c.bwd ? b_val : f_val;
    const unsigned zs = (unsigned)si.ns;
    for (int l = c.lev0; l < c.lev0 + c.nlev; ++l) {
        const Level L = lev[l];
        // only the level's rows inside the chunk (a level may span chunks)
        const int qa = max(c.row0 - L.start, 0);
        const int qb = min(c.row0 + c.nrow - L.start, L.size);
        for (int q = qa + (int)threadIdx.x; q < qb; q += blockDim.x) {
            unsigned cc[kKreg];
            T vv[kKreg];
            for (int k = 0; k < kKreg; ++k) {
                cc[k] = k < L.K ? (unsigned)cols[L.ell + k * L.size + q] : zs;
                vv[k] = k < L.K ? vals[L.ell + k * L.size + q] : (T)0;
            }
            const T dg = c.bwd ? b_diag[L.start + q] : (T)0;
            unsigned dbits[2] = {0u, 0u};
            __builtin_memcpy(dbits, &dg, sizeof(T));
            if constexpr (KR <= 3) {
                const int pr = L.start - c.row0 + q;  // row within the chunk
                unsigned vb[kKreg];
                __builtin_memcpy(vb, vv, sizeof(vv));
                const unsigned r0 = (unsigned)rows[L.start + q] | (cc[0] << 16);
                if (KR == 2 && !c.bwd) {
                    img[c.img + pr] = make_uint4(r0, cc[1], vb[0], vb[1]);
                    continue;
                }
                unsigned* ib = (unsigned*)(img + c.img);
                ib[2 * pr] = r0;
                ib[2 * pr + 1] = KR == 2 ? cc[1] : cc[1] | (cc[2] << 16);
                img[c.img + (c.nrow + 1) / 2 + pr] =
                    make_uint4(vb[0], vb[1], KR == 2 ? 0u : vb[2], dbits[0]);
                continue;
            }
            uint4* r = img + c.img + (size_t)(L.start - c.row0 + q) * stage_rec_quads<T>();
            r[0] = make_uint4((unsigned)rows[L.start + q] | (cc[0] << 16), cc[1] | (cc[2] << 16),
                              cc[3], sizeof(T) == 4 ? dbits[0] : 0u);
            unsigned vb[2 * kKreg];
            __builtin_memcpy(vb, vv, sizeof(vv));
            if (sizeof(T) == 4) {
                r[1] = make_uint4(vb[0], vb[1], vb[2], vb[3]);
            } else {
                r[1] = make_uint4(vb[0], vb[1], vb[2], vb[3]);
                r[2] = make_uint4(vb[4], vb[5], vb[6], vb[7]);
                r[3] = make_uint4(dbits[0], dbits[1], 0u, 0u);
            }
        }
    }
}

// One row of a level as read from the chunk buffer (raw 16-byte words).
template <typename T, int KR>
struct StageRow {
    stage_q q[stage_rec_quads<T>()];
};
template <>
struct StageRow<float, 3> {
    stage_d ix;  // row | c0<<16, c1 | c2<<16
    stage_q v;   // v0, v1, v2, diag
};
template <>
struct StageRow<float, 2> {
    stage_d ix;  // row | c0<<16, c1
    stage_q v;   // v0, v1, -, diag
};

// vb: first 16-byte word of the value block (KR = 3, backward KR = 2)
template <typename T, int KR, bool BWD>
__device__ __forceinline__ void stage_fetch(StageRow<T, KR>& r, const lds_t<stage_q>* buf, int p,
                                            int vb) {
    if constexpr (KR == 2 && !BWD) {
        const stage_q w = buf[p];
        r.ix = stage_d{w.x, w.y};
        r.v = stage_q{w.z, w.w, 0u, 0u};
    } else if constexpr (KR <= 3) {
        r.ix = ((const lds_t<stage_d>*)buf)[p];
        r.v = buf[vb + p];
    } else {
#pragma unroll
        for (int i = 0; i < stage_rec_quads<T>(); ++i) r.q[i] = buf[p * stage_rec_quads<T>() + i];
    }
}

__device__ __forceinline__ float stage_bits_f(unsigned a) { return __uint_as_float(a); }
__device__ __forceinline__ double stage_bits_d(unsigned lo, unsigned hi) {
    return __hiloint2double((int)hi, (int)lo);
}

template <typename T, int KR>
__device__ __forceinline__ int stage_row(const StageRow<T, KR>& r) {
    if constexpr (KR <= 3) return r.ix.x & 0xffffu;
    else return r.q[0].x & 0xffffu;
}
template <typename T, int KR>
__device__ __forceinline__ int stage_col(const StageRow<T, KR>& r, int k) {
    if constexpr (KR <= 3)
        return k == 0 ? (int)(r.ix.x >> 16) : k == 1 ? (int)(r.ix.y & 0xffffu) : (int)(r.ix.y >> 16);
    else
        return k == 0 ? (int)(r.q[0].x >> 16)
             : k == 1 ? (int)(r.q[0].y & 0xffffu)
             : k == 2 ? (int)(r.q[0].y >> 16)
                      : (int)(r.q[0].z & 0xffffu);
}
template <typename T, int KR>
__device__ __forceinline__ T stage_val(const StageRow<T, KR>& r, int k) {
    if constexpr (KR <= 3) {
        return stage_bits_f(k == 0 ? r.v.x : k == 1 ? r.v.y : r.v.z);
    } else if constexpr (sizeof(T) == 4) {
        return stage_bits_f(k == 0 ? r.q[1].x : k == 1 ? r.q[1].y : k == 2 ? r.q[1].z : r.q[1].w);
    } else {
        const stage_q& w = r.q[1 + k / 2];
        return k % 2 == 0 ? stage_bits_d(w.x, w.y) : stage_bits_d(w.z, w.w);
    }
}
template <typename T, int KR>
__device__ __forceinline__ T stage_diag(const StageRow<T, KR>& r) {
    if constexpr (KR <= 3) return stage_bits_f(r.v.w);
    else if constexpr (sizeof(T) == 4) return stage_bits_f(r.q[0].w);
    else return stage_bits_d(r.q[3].x, r.q[3].y);
}

// The levels of chunk c out of LDS; each level ends with an LDS-only barrier.
// Thread t takes row t of a level (rows t + NT, ... of a wider level in a
// second loop).  The next level's descriptor and the thread's record of it
// are read while the current level computes (they do not depend on z).  A
// wavefront with no row in a level skips it (a uniform branch): a level then
// costs the SIMDs only the instructions of the wavefronts that hold its rows.
// Threads past a level's end inside an active wavefront compute a clamped row
// and write it to the trash slot zs + 1.
template <typename T, bool BWD, int NT, int KR>
__device__ __forceinline__ void stage_levels(lds_t<T>* z, const lds_t<int>* tab,
                                             const SChunk& c, const lds_t<stage_q>* buf,
                                             int zs) {
    const int vb = (c.nrow + 1) / 2;  // value block (KR = 3)
    const int t = threadIdx.x;
    const int wbase = __builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63;
    const int lend = c.lev0 + c.nlev;
    // level descriptors: d (current, uniform), dn (next, uniform), and the raw
    // LDS values of the one after, read while the current level computes
    // a chunk may start or end inside a level (rows of one level are
    // independent): each level runs its rows inside [row0, row0 + nrow)
    struct D {
        int size, p;  // rows, first record of the level in the chunk
    };
    auto uni = [&](int start, int size) {
        const int a = __builtin_amdgcn_readfirstlane(start);
        const int st = max(a, c.row0);
        const int en = min(a + __builtin_amdgcn_readfirstlane(size), c.row0 + c.nrow);
        return D{en - st, st - c.row0};
    };
    const int l1 = min(c.lev0 + 1, lend - 1);
    D d = uni(tab[4 * c.lev0], tab[4 * c.lev0 + 1]);
    D dn = uni(tab[4 * l1], tab[4 * l1 + 1]);
    StageRow<T, KR> cur;
    if (wbase < d.size) stage_fetch<T, KR, BWD>(cur, buf, d.p + min(t, d.size - 1), vb);
    for (int l = c.lev0; l < lend; ++l) {
        // this level's z values first: they are the critical path
        T zk[KR], acc = (T)0;
        if (wbase < d.size) {
#pragma unroll
            for (int k = 0; k < KR; ++k) zk[k] = z[stage_col<T, KR>(cur, k)];
            acc = z[stage_row<T, KR>(cur)];
        }
        // then the next level's records and the descriptor after it
        StageRow<T, KR> nxt;
        if (wbase < dn.size) stage_fetch<T, KR, BWD>(nxt, buf, dn.p + min(t, dn.size - 1), vb);
        const stage_i4 dq = ((const lds_t<stage_i4>*)tab)[min(l + 2, lend - 1)];
        if (wbase < d.size) {
#pragma unroll
            for (int k = 0; k < KR; ++k) acc -= stage_val<T, KR>(cur, k) * zk[k];
            if (BWD) acc = acc / stage_diag<T, KR>(cur);
            z[t < d.size ? stage_row<T, KR>(cur) : zs + 1] = acc;
            if (d.size > NT) {
                for (int q = t + NT; q < d.size; q += NT) {  // levels wider than the workgroup
                    StageRow<T, KR> r;
                    stage_fetch<T, KR, BWD>(r, buf, d.p + q, vb);
                    T a = z[stage_row<T, KR>(r)];
#pragma unroll
                    for (int k = 0; k < KR; ++k) a -= stage_val<T, KR>(r, k) * z[stage_col<T, KR>(r, k)];
                    if (BWD) a = a / stage_diag<T, KR>(r);
                    z[stage_row<T, KR>(r)] = a;
                }
            }
        }
        lds_barrier();
        d = dn;
        dn = uni(dq.x, dq.y);
        cur = nxt;
    }
}

// The coarse term of one subdomain folded into its staged sweep (two-level
// RAS with the precomputed coarse inverse): cz_s = c_s (row s of the inverse
// . r0), r0[t] the sum of the restriction's split partials of coarse row t.
// The operations and their order are those of k_coarse_inv for row s, so the
// value is the same bit for bit; every wavefront computes it from loads issued
// before the gather (nc <= kFoldRows, split == kRestrictSplit).
constexpr int kFoldRows = 256;
template <typename T>
struct CoarseFold {
    const T* inv;  // nullptr: no fold
    const T* r0;
    const T* c;
    int nc;
};
template <typename T>
struct FoldRegs {
    T r[kFoldRows / 64][kRestrictSplit], i[kFoldRows / 64], c;
};
template <typename T>
__device__ __forceinline__ void fold_load(FoldRegs<T>& f, const CoarseFold<T>& cf, int s) {
    const int lane = threadIdx.x & 63;
    const T* ri = cf.inv + (size_t)s * cf.nc;
#pragma unroll
    for (int k = 0; k < kFoldRows / 64; ++k) {
        const int j = min(lane + 64 * k, cf.nc - 1);
        f.i[k] = ri[j];
#pragma unroll
        for (int q = 0; q < kRestrictSplit; ++q) f.r[k][q] = cf.r0[j * kRestrictSplit + q];
    }
    f.c = cf.c[s];
}
template <typename T>
__device__ __forceinline__ T fold_value(const FoldRegs<T>& f, int nc) {
    const int lane = threadIdx.x & 63;
    T acc = 0;
#pragma unroll
    for (int k = 0; k < kFoldRows / 64; ++k)
        if (lane + 64 * k < nc) {
            T r = 0;
#pragma unroll
            for (int q = 0; q < kRestrictSplit; ++q) r += f.r[k][q];
            acc += f.i[k] * r;
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    return f.c * acc;
}

// Diagnostic clock of the staged sweep, per workgroup: start, gather done,
// sweeps done, end (device wall clock), subdomain, hardware id, rows, forward
// levels, backward levels, chunks, owned rows, 0.
constexpr int kClockWords = 12;

// LDS: level tables (forward | backward, four ints per level) | z (ns values,
// zero slot, trash slot) | chunk descriptors | buffer 0 | buffer 1 | trash word.
template <typename T, int MODE, int NT, int KR>
__global__ void __launch_bounds__(NT)
k_sweep_staged(const SubInfo* __restrict__ subs, const int* __restrict__ order,
               const int* __restrict__ chunk0, const SChunk* __restrict__ chunks,
               const stage_q* __restrict__ img, const Level* __restrict__ flev,
               const Level* __restrict__ blev, const int* __restrict__ dofs,
               const own_i2* __restrict__ own2, const double* __restrict__ x,
               double* __restrict__ y, T* __restrict__ zout, const T* __restrict__ cz,
               CoarseFold<T> cf, DevGuard guard, int z_off, int ct_off, int buf_off,
               int buf_words, long long* __restrict__ clk) {
    extern __shared__ unsigned char lds_raw[];
    const long long t_start = clk ? wall_clock64() : 0;
    const int s = order[blockIdx.x];
    if (guard_stop(guard, s)) return;
    const SubInfo si = subs[s];
    const int c0 = chunk0[2 * s], nch = chunk0[2 * s + 1];
    if (nch == 0) return;  // an empty subdomain: no rows, nothing owned
    lds_t<int>* lf = (lds_t<int>*)lds_raw;
    lds_t<int>* lb = lf + 4 * si.nflev;
    lds_t<T>* z = (lds_t<T>*)(lds_raw + z_off);
    lds_t<stage_i4>* ct = (lds_t<stage_i4>*)(lds_raw + ct_off);
    lds_t<unsigned>* bufs = (lds_t<unsigned>*)(lds_raw + buf_off);
    {
        const int* tf = (const int*)(flev + si.flev0);
        const int* tb = (const int*)(blev + si.blev0);
        for (int l = threadIdx.x; l < 4 * si.nflev; l += NT) lf[l] = tf[l];
        for (int l = threadIdx.x; l < 4 * si.nblev; l += NT) lb[l] = tb[l];
        const stage_i4* tc = (const stage_i4*)(chunks + c0);
        for (int l = threadIdx.x; l < 2 * nch; l += NT) ct[l] = tc[l];
    }
    // two register sets: while chunk c computes, chunks c + 1 and c + 2 are in
    // flight (set (c + 1) & 1 and set c & 1).  Loads past the last chunk reload
    // the last one, so every step issues the same loads and the waits stay exact.
    // The first three descriptors come from global memory, the others from the
    // LDS copy (a scalar load at every chunk step would stall on its latency).
    StageRegs<T, NT> RA, RB;
    auto chunk_g = [&](int k) { return chunks[c0 + min(k, nch - 1)]; };
    auto chunk_at = [&](int k) {
        const int kk = min(k, nch - 1);
        const stage_i4 a = ct[2 * kk], b = ct[2 * kk + 1];
        SChunk c;
        c.lev0 = __builtin_amdgcn_readfirstlane(a.x);
        c.nlev = __builtin_amdgcn_readfirstlane(a.y);
        c.row0 = __builtin_amdgcn_readfirstlane(a.z);
        c.nrow = __builtin_amdgcn_readfirstlane(a.w);
        c.bwd = __builtin_amdgcn_readfirstlane(b.x);
        c.img = __builtin_amdgcn_readfirstlane(b.y);
        return c;
    };
    // the coarse term's loads go first: the gather's waits cover them (the
    // host folds only into shapes of at most 512 threads)
    constexpr bool kFold = MODE == 0 && NT <= 512;
    FoldRegs<T> fr;
    if constexpr (kFold)
        if (cf.inv) fold_load<T>(fr, cf, s);
    // descriptors of the chunks held in RA and RB, and in buffers q0 and q1
    SChunk dA = chunk_g(0), dB = chunk_g(1);
    stage_load<T, NT, KR>(RA, dA, img);
    stage_load<T, NT, KR>(RB, dB, img);
    gather_x<NT, T, lds_t<T>*, 24>(dofs + si.row0, si.ns, x, z);
    if (threadIdx.x == 0) z[si.ns] = (T)0;  // zero slot for padding entries
    T cfold = (T)0;
    if constexpr (kFold)
        if (cf.inv) cfold = fold_value<T>(fr, cf.nc);
    lds_t<stage_q>* trash = (lds_t<stage_q>*)(bufs + 2 * buf_words);
    lds_t<stage_q>* q0 = (lds_t<stage_q>*)bufs;
    lds_t<stage_q>* q1 = (lds_t<stage_q>*)(bufs + buf_words);
    SChunk c0q = dA, c1q;
    stage_store<T, NT, KR>(RA, dA, q0, trash);
    dA = chunk_g(2);
    stage_load<T, NT, KR>(RA, dA, img);
    __syncthreads();
    const long long t_gathered = clk ? wall_clock64() : 0;
    auto run = [&](const SChunk& c, const lds_t<stage_q>* buf) {
        if (c.bwd) stage_levels<T, true, NT, KR>(z, lb, c, buf, si.ns);
        else stage_levels<T, false, NT, KR>(z, lf, c, buf, si.ns);
    };
    // each step stores the next chunk into the other buffer (free since the
    // barrier that ended the previous step) and reloads its register set
    // before computing: the LDS writes drain while the levels run, and the
    // levels' last barrier publishes them.  Pairs of steps with no exit
    // inside, so the compiler's waits stay exact (set B's loads, then set A's)
    int ci = 0;
    for (; ci + 1 < nch; ci += 2) {
        const SChunk nb = chunk_at(ci + 3);
        stage_store<T, NT, KR>(RB, dB, q1, trash);
        c1q = dB;
        stage_load<T, NT, KR>(RB, nb, img);
        dB = nb;
        run(c0q, q0);
        const SChunk na = chunk_at(ci + 4);
        stage_store<T, NT, KR>(RA, dA, q0, trash);
        c0q = dA;
        stage_load<T, NT, KR>(RA, na, img);
        dA = na;
        run(c1q, q1);
    }
    // owned rows per thread in flight: a whole subdomain of up to 24K owned rows
    // in one batch (1024 threads: 16K, within the register budget)
    constexpr int US = NT >= 1024 ? 16 : NT >= 512 ? 48 : 64;
    own_i2 ob[US];
    if (MODE == 0 && si.nown > 0) own_batch<NT, US>(ob, own2 + si.own0, si.nown, threadIdx.x);
    if (ci < nch) run(c0q, q0);  // odd count: the last chunk is in buffer 0
    __syncthreads();
    const long long t_swept = clk ? wall_clock64() : 0;
    if (MODE == 0) {
        const bool has_c = cf.inv != nullptr || cz != nullptr;
        const T add = cf.inv ? cfold : cz ? cz[s] : (T)0;
        scatter_owned2<NT, US>(ob, own2 + si.own0, si.nown, z, has_c, add, y);
    } else {
        for (int i = threadIdx.x; i < si.ns; i += NT) zout[si.row0 + i] = z[i];
    }
    if (clk) {  // diagnostic clock: the workgroup's span and what it solved
        __syncthreads();
        if (threadIdx.x == 0) {
            long long* r = clk + kClockWords * (size_t)blockIdx.x;
            r[0] = t_start;
            r[1] = t_gathered;
            r[2] = t_swept;
            r[3] = wall_clock64();
            r[4] = s;
            r[5] = (long long)__smid();
            r[6] = si.ns;
            r[7] = si.nflev;
            r[8] = si.nblev;
            r[9] = nch;
            r[10] = si.nown;
            r[11] = 0;
        }
    }
}

// ---- staged ring sweep ------------------------------------------------------------
// The staged sweep for subdomains whose vector exceeds LDS, built on the ring
// encoding (every column reference is a live slot of an LDS ring of R recent
// values: level-order position mod R; padding is slot R, which holds zero).
// Setup packs each row into a record: destination, four ring references, four
// values, diagonal (FP32 48 bytes, FP64 64 bytes).  A row's own value comes
// from a level-ordered array that the gather fills (forward) or the forward
// sweep writes (backward, at the row's backward position); each chunk's share
// of it is 16-byte aligned and is staged with the records.  Forward and
// backward sweeps are two pipelines (the backward own values exist only once
// the forward sweep has finished).  Operation order per row equals the ring
// sweep's, so the results are the same bit for bit.
// Compact records (CR), when every level has at most sring_k entries per row
// and the ring slots fit 16 bits: two 16-byte words per row.  FP64 (two
// entries): [dst, c0 | c1<<16, diag lo, diag hi], [v0, v1]; FP32 (three):
// [dst, c0 | c1<<16, c2, diag], [v0, v1, v2, 0].  Dropped padding terms
// subtract +0, so the values do not change.
template <typename T, bool CR>
__host__ __device__ constexpr int sring_k() {
    return CR ? (sizeof(T) == 8 ? 2 : 3) : kKreg;
}
template <typename T, bool CR = false>
__host__ __device__ constexpr int sring_rec_quads() {
    return CR ? 2 : sizeof(T) == 4 ? 3 : 4;
}
template <typename T>
__host__ __device__ inline int sring_own_quads(int nrow) {
    return (int)((nrow * sizeof(T) + 15) / 16);
}
template <typename T, bool CR>
__host__ __device__ inline int sring_quads(const SChunkR& c) {
    return c.nrow * sring_rec_quads<T, CR>() + sring_own_quads<T>(c.nrow);
}

template <typename T, int NT, bool CR>
__device__ __forceinline__ void sring_load(StageRegs<T, NT>& R, const SChunkR& c,
                                           const stage_q* __restrict__ img,
                                           const stage_q* __restrict__ own) {
    typedef const __attribute__((address_space(1))) stage_q gq;
    const int nrec = c.nrow * sring_rec_quads<T, CR>();
    const int n = nrec + sring_own_quads<T>(c.nrow);
    const size_t bi = (size_t)(img + c.img), bo = (size_t)(own + c.own) - 16 * (size_t)nrec;
#pragma unroll
    for (int i = 0; i < StageRegs<T, NT>::kV; ++i) {
        const int u = min((int)threadIdx.x + i * NT, n - 1);
        R.v[i] = __builtin_nontemporal_load((gq*)((u < nrec ? bi : bo) + 16 * (size_t)u));
    }
}

template <typename T, int NT, bool CR>
__device__ __forceinline__ void sring_store(const StageRegs<T, NT>& R, const SChunkR& c,
                                            lds_t<stage_q>* buf, lds_t<stage_q>* trash) {
    const int n = sring_quads<T, CR>(c);
#pragma unroll
    for (int i = 0; i < StageRegs<T, NT>::kV; ++i) {
        const int u = (int)threadIdx.x + i * NT;
        *(u < n ? buf + u : trash) = R.v[i];
    }
}

template <typename T, bool CR>
struct SRingRow {
    stage_q q[sring_rec_quads<T, CR>()];
};

template <typename T, bool CR>
__device__ __forceinline__ T sring_val(const SRingRow<T, CR>& r, int k) {
    if constexpr (CR && sizeof(T) == 4) {
        return stage_bits_f(k == 0 ? r.q[1].x : k == 1 ? r.q[1].y : r.q[1].z);
    } else if constexpr (CR) {
        return k == 0 ? stage_bits_d(r.q[1].x, r.q[1].y) : stage_bits_d(r.q[1].z, r.q[1].w);
    } else if constexpr (sizeof(T) == 4) {
        return stage_bits_f(k == 0 ? r.q[2].x : k == 1 ? r.q[2].y : k == 2 ? r.q[2].z : r.q[2].w);
    } else {
        const stage_q& w = r.q[2 + k / 2];
        return k % 2 == 0 ? stage_bits_d(w.x, w.y) : stage_bits_d(w.z, w.w);
    }
}
template <typename T, bool CR>
__device__ __forceinline__ int sring_col(const SRingRow<T, CR>& r, int k) {
    if constexpr (CR)
        return (int)(k == 0 ? r.q[0].y & 0xffffu : k == 1 ? r.q[0].y >> 16 : r.q[0].z);
    else
        return (int)(k == 0 ? r.q[0].y : k == 1 ? r.q[0].z : k == 2 ? r.q[0].w : r.q[1].x);
}
template <typename T, bool CR>
__device__ __forceinline__ T sring_diag(const SRingRow<T, CR>& r) {
    if constexpr (CR && sizeof(T) == 4) return stage_bits_f(r.q[0].w);
    else if constexpr (CR) return stage_bits_d(r.q[0].z, r.q[0].w);
    else if constexpr (sizeof(T) == 4) return stage_bits_f(r.q[1].y);
    else return stage_bits_d(r.q[1].y, r.q[1].z);
}

template <typename T, bool CR>
__device__ __forceinline__ void sring_fetch(SRingRow<T, CR>& r, const lds_t<stage_q>* buf, int p) {
#pragma unroll
    for (int i = 0; i < sring_rec_quads<T, CR>(); ++i)
        r.q[i] = buf[p * sring_rec_quads<T, CR>() + i];
}

// One chunk's levels: row t of a level computes own - sum v_k ring[c_k]
// (/ diag backward), writes ring[(position - pos0) mod R] and zout[dst].
template <typename T, bool BWD, int NT, bool CR>
__device__ __forceinline__ void sring_levels(lds_t<T>* ring, int R, const lds_t<int>* tab,
                                             const SChunkR& c, const lds_t<stage_q>* buf,
                                             int pos0, T* __restrict__ zout) {
    const int t = threadIdx.x;
    const int wbase = __builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63;
    const lds_t<T>* own = (const lds_t<T>*)(buf + c.nrow * sring_rec_quads<T, CR>());
    const int lend = c.lev0 + c.nlev;
    // a chunk may start or end inside a level (rows of one level are
    // independent): each level runs its rows inside [row0, row0 + nrow)
    struct D {
        int size, p, slot0;  // rows, first record in the chunk, ring slot of the first row
    };
    auto uni = [&](int start, int size) {
        const int st = max(__builtin_amdgcn_readfirstlane(start), c.row0);
        const int en = min(__builtin_amdgcn_readfirstlane(start) + __builtin_amdgcn_readfirstlane(size),
                           c.row0 + c.nrow);
        return D{en - st, st - c.row0, (st - pos0) % R};
    };
    const int l1 = min(c.lev0 + 1, lend - 1);
    D d = uni(tab[4 * c.lev0], tab[4 * c.lev0 + 1]);
    D dn = uni(tab[4 * l1], tab[4 * l1 + 1]);
    SRingRow<T, CR> cur;
    T cown = (T)0;
    if (wbase < d.size) {
        sring_fetch<T, CR>(cur, buf, d.p + min(t, d.size - 1));
        cown = own[d.p + min(t, d.size - 1)];
    }
    for (int l = c.lev0; l < lend; ++l) {
        T zk[sring_k<T, CR>()];
        if (wbase < d.size) {
#pragma unroll
            for (int k = 0; k < sring_k<T, CR>(); ++k) zk[k] = ring[sring_col<T, CR>(cur, k)];
        }
        SRingRow<T, CR> nxt;
        T nown = (T)0;
        if (wbase < dn.size) {
            sring_fetch<T, CR>(nxt, buf, dn.p + min(t, dn.size - 1));
            nown = own[dn.p + min(t, dn.size - 1)];
        }
        const stage_i4 dq = ((const lds_t<stage_i4>*)tab)[min(l + 2, lend - 1)];
        if (wbase < d.size) {
            T acc = cown;
#pragma unroll
            for (int k = 0; k < sring_k<T, CR>(); ++k) acc -= sring_val<T, CR>(cur, k) * zk[k];
            if (BWD) acc = acc / sring_diag<T, CR>(cur);
            int sl = d.slot0 + t;
            if (sl >= R) sl -= R;
            ring[t < d.size ? sl : R + 1] = acc;
            if (t < d.size) zout[cur.q[0].x] = acc;
            if (d.size > NT) {
                for (int q = t + NT; q < d.size; q += NT) {  // levels wider than the workgroup
                    SRingRow<T, CR> r;
                    sring_fetch<T, CR>(r, buf, d.p + q);
                    T a = own[d.p + q];
#pragma unroll
                    for (int k = 0; k < sring_k<T, CR>(); ++k) a -= sring_val<T, CR>(r, k) * ring[sring_col<T, CR>(r, k)];
                    if (BWD) a = a / sring_diag<T, CR>(r);
                    int s2 = d.slot0 + q;
                    while (s2 >= R) s2 -= R;
                    ring[s2] = a;
                    zout[r.q[0].x] = a;
                }
            }
        }
        lds_barrier();
        d = dn;
        dn = uni(dq.x, dq.y);
        cur = nxt;
        cown = nown;
    }
}

// LDS: level tables (forward | backward, four ints per level) | chunk
// descriptors | ring (R values, zero slot, trash slot) | buffer 0 | buffer 1 |
// trash word.
template <typename T, int MODE, int NT, bool CR>
__global__ void __launch_bounds__(NT)
k_sweep_staged_ring(const SubInfo* __restrict__ subs, const int* __restrict__ order,
                    const int* __restrict__ chunk0, const SChunkR* __restrict__ chunks,
                    const stage_q* __restrict__ img, stage_q* __restrict__ own,
                    const int* __restrict__ fslot, const int* __restrict__ fgid,
                    const Level* __restrict__ flev, const Level* __restrict__ blev,
                    const int* __restrict__ own_l, const int* __restrict__ own_gid,
                    const double* __restrict__ x, double* __restrict__ y, T* __restrict__ z3,
                    size_t zstride, T* __restrict__ zout, const T* __restrict__ cz,
                    DevGuard guard, int ct_off, int ring_off, int buf_off, int buf_quads,
                    int R) {
    extern __shared__ unsigned char lds_raw[];
    const int s = order[blockIdx.x];
    if (guard_stop(guard, s)) return;
    const SubInfo si = subs[s];
    lds_t<int>* lf = (lds_t<int>*)lds_raw;
    lds_t<int>* lb = lf + 4 * si.nflev;
    lds_t<stage_i4>* ct = (lds_t<stage_i4>*)(lds_raw + ct_off);
    lds_t<T>* ring = (lds_t<T>*)(lds_raw + ring_off);
    lds_t<stage_q>* q0 = (lds_t<stage_q>*)(lds_raw + buf_off);
    lds_t<stage_q>* q1 = q0 + buf_quads;
    lds_t<stage_q>* trash = q0 + 2 * buf_quads;
    // this subdomain's chunks: forward ones, then backward ones
    const int nfc = chunk0[4 * s + 1], nbc = chunk0[4 * s + 3];
    {
        const int* tf = (const int*)(flev + si.flev0);
        const int* tb = (const int*)(blev + si.blev0);
        for (int l = threadIdx.x; l < 4 * si.nflev; l += NT) lf[l] = tf[l];
        for (int l = threadIdx.x; l < 4 * si.nblev; l += NT) lb[l] = tb[l];
        const stage_i4* tc = (const stage_i4*)(chunks + chunk0[4 * s]);
        for (int l = threadIdx.x; l < 2 * (nfc + nbc); l += NT) ct[l] = tc[l];
    }
    const int pf0 = si.nflev ? flev[si.flev0].start : 0;
    const int pb0 = si.nblev ? blev[si.blev0].start : 0;
    T* ownT = (T*)own;
    T* zr = z3 + 2 * zstride + si.zoff;
    // forward own values: x gathered in forward level order into the chunks'
    // own segments, eight positions per thread in flight
    {
        constexpr int U = 8;
        for (int p0 = threadIdx.x; p0 < si.ns; p0 += U * NT) {
            int g[U], sl[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = pf0 + min(p0 + u * NT, si.ns - 1);
                g[u] = fgid[p];
                sl[u] = fslot[p];
            }
            double xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[g[u]];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (p0 + u * NT < si.ns) ownT[sl[u]] = (T)xv[u];
        }
    }
    if (threadIdx.x == 0) ring[R] = (T)0;
    __syncthreads();
    // one sweep: chunks [lc0, lc0 + nch) of the LDS descriptor table through
    // the two buffers, as k_sweep_staged
    auto pipeline = [&](int lc0, int nch, const lds_t<int>* tab, bool bwd, int pos0, T* out) {
        StageRegs<T, NT> RA, RB;
        auto chunk_at = [&](int k) {
            const int kk = lc0 + min(k, nch - 1);
            const stage_i4 a = ct[2 * kk], b = ct[2 * kk + 1];
            SChunkR c;
            c.lev0 = __builtin_amdgcn_readfirstlane(a.x);
            c.nlev = __builtin_amdgcn_readfirstlane(a.y);
            c.row0 = __builtin_amdgcn_readfirstlane(a.z);
            c.nrow = __builtin_amdgcn_readfirstlane(a.w);
            c.img = __builtin_amdgcn_readfirstlane(b.x);
            c.own = __builtin_amdgcn_readfirstlane(b.y);
            c.bwd = __builtin_amdgcn_readfirstlane(b.z);
            c.pad = 0;
            return c;
        };
        auto run = [&](const SChunkR& c, const lds_t<stage_q>* buf) {
            if (bwd) sring_levels<T, true, NT, CR>(ring, R, tab, c, buf, pos0, out);
            else sring_levels<T, false, NT, CR>(ring, R, tab, c, buf, pos0, out);
        };
        // descriptors of the chunks held in RA and RB, and in buffers q0 and q1
        SChunkR dA = chunk_at(0), dB = chunk_at(1);
        sring_load<T, NT, CR>(RA, dA, img, own);
        sring_load<T, NT, CR>(RB, dB, img, own);
        SChunkR c0q = dA, c1q;
        sring_store<T, NT, CR>(RA, dA, q0, trash);
        dA = chunk_at(2);
        sring_load<T, NT, CR>(RA, dA, img, own);
        lds_barrier();
        int ci = 0;
        for (; ci + 1 < nch; ci += 2) {
            const SChunkR nb = chunk_at(ci + 3);
            sring_store<T, NT, CR>(RB, dB, q1, trash);
            c1q = dB;
            sring_load<T, NT, CR>(RB, nb, img, own);
            dB = nb;
            run(c0q, q0);
            const SChunkR na = chunk_at(ci + 4);
            sring_store<T, NT, CR>(RA, dA, q0, trash);
            c0q = dA;
            sring_load<T, NT, CR>(RA, na, img, own);
            dA = na;
            run(c1q, q1);
        }
        if (ci < nch) run(c0q, q0);
    };
    if (si.nflev) pipeline(0, nfc, lf, false, pf0, ownT);
    __syncthreads();  // backward own values (the forward results) are in global memory
    if (si.nblev) pipeline(nfc, nbc, lb, true, pb0, zr);
    __syncthreads();
    if (MODE == 0) {
        scatter_owned<NT>(own_l + si.own0, own_gid + si.own0, si.nown, zr, cz != nullptr,
                          cz ? cz[s] : (T)0, y);
    } else {
        for (int i = threadIdx.x; i < si.ns; i += NT) zout[si.row0 + i] = zr[i];
    }
}

// Setup: packs the rows of every staged-ring chunk into records (one workgroup
// per chunk).  dst: forward, the own-value index of the row's backward
// position; backward, the local row.
template <typename T, bool CR>
__global__ void k_stage_image_ring(const SChunkR* __restrict__ chunks,
                                   const SubInfo* __restrict__ subs,
                                   const int* __restrict__ chunk_sub,
                                   const Level* __restrict__ flev, const Level* __restrict__ blev,
                                   const int* __restrict__ fdst, const int* __restrict__ bdst,
                                   const int* __restrict__ fcol, const int* __restrict__ bcol,
                                   const T* __restrict__ f_val, const T* __restrict__ b_val,
                                   const T* __restrict__ b_diag, int R, uint4* __restrict__ img) {
    const SChunkR c = chunks[blockIdx.x];
    const SubInfo si = subs[chunk_sub[blockIdx.x]];
    const Level* lev = c.bwd ? blev + si.blev0 : flev + si.flev0;
    const int* cols = c.bwd ? bcol : fcol;
    const T* vals = c.bwd ? b_val : f_val;
    for (int l = c.lev0; l < c.lev0 + c.nlev; ++l) {
        const Level L = lev[l];
        // only the level's rows inside the chunk (a level may span chunks)
        const int qa = max(c.row0 - L.start, 0);
        const int qb = min(c.row0 + c.nrow - L.start, L.size);
        for (int q = qa + (int)threadIdx.x; q < qb; q += blockDim.x) {
            unsigned cc[kKreg];
            T vv[kKreg];
            for (int k = 0; k < kKreg; ++k) {
                cc[k] = k < L.K ? (unsigned)cols[L.ell + k * L.size + q] : (unsigned)R;
                vv[k] = k < L.K ? vals[L.ell + k * L.size + q] : (T)0;
            }
            const int p = L.start + q;
            const T dg = c.bwd ? b_diag[p] : (T)0;
            unsigned db[2] = {0u, 0u};
            __builtin_memcpy(db, &dg, sizeof(T));
            unsigned vb[2 * kKreg];
            __builtin_memcpy(vb, vv, sizeof(vv));
            uint4* r = img + c.img + (size_t)(p - c.row0) * sring_rec_quads<T, CR>();
            const unsigned dst = (unsigned)(c.bwd ? bdst[p] : fdst[p]);
            if constexpr (CR) {
                if (sizeof(T) == 4) {
                    r[0] = make_uint4(dst, cc[0] | (cc[1] << 16), cc[2], db[0]);
                    r[1] = make_uint4(vb[0], vb[1], vb[2], 0u);
                } else {
                    r[0] = make_uint4(dst, cc[0] | (cc[1] << 16), db[0], db[1]);
                    r[1] = make_uint4(vb[0], vb[1], vb[2], vb[3]);
                }
                continue;
            }
            r[0] = make_uint4(dst, cc[0], cc[1], cc[2]);
            r[1] = make_uint4(cc[3], db[0], sizeof(T) == 8 ? db[1] : 0u, 0u);
            if (sizeof(T) == 4) {
                r[2] = make_uint4(vb[0], vb[1], vb[2], vb[3]);
            } else {
                r[2] = make_uint4(vb[0], vb[1], vb[2], vb[3]);
                r[3] = make_uint4(vb[4], vb[5], vb[6], vb[7]);
            }
        }
    }
}

// ASM prolongation: y[g] = sum over subdomains containing g, ascending, of z.
template <typename T>
__global__ void k_asm_gather(int n, const int* __restrict__ ptr, const int* __restrict__ pos,
                             const T* __restrict__ zbuf, const int* __restrict__ part,
                             const T* __restrict__ cz, double* __restrict__ y,
                             DevGuard guard) {
    if (guard_stop(guard, 0)) return;
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
// (src/precond/coarse_space.h:63-70).  Subdomain s is split over `split`
// workgroups; partial sums go to part[s * split + j] and are added in fixed
// order by the coarse solve.
template <typename T>
__global__ void __launch_bounds__(256)
k_coarse_restrict(int split, const SubInfo* __restrict__ subs, const int* __restrict__ own_gid,
                  const T* __restrict__ c, const double* __restrict__ x, T* __restrict__ part,
                  DevGuard guard) {
    const int s = blockIdx.x / split, j = blockIdx.x % split;
    if (guard_stop(guard, s)) {
        // a stopped system of a batch contributes an exact zero to the
        // (block-diagonal) coarse solve; otherwise the whole launch is skipped
        if (guard.map && threadIdx.x == 0) part[blockIdx.x] = 0;
        return;
    }
    const SubInfo si = subs[s];
    const int per = (si.nown + split - 1) / split;
    const int a = j * per, b = min(si.nown, a + per);
    const T cs = c[s];
    const int* og = own_gid + si.own0;
    T acc = 0;
    // eight index and value loads in flight per thread; the terms are added
    // in the order of the one-at-a-time loop
    for (int k0 = a + threadIdx.x; k0 < b; k0 += 8 * 256) {
        int g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) g[u] = og[min(k0 + u * 256, b - 1)];
        double xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[g[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u * 256 < b) acc += cs * (T)xv[u];
    }
    __shared__ T red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Coarse solve with the precomputed inverse (row-major): one wavefront per
// coarse row; every block first adds the restriction partials of all rows.
template <typename T>
__global__ void __launch_bounds__(256)
k_coarse_inv(int nc, int split, const T* __restrict__ inv, const T* __restrict__ r0part,
             const T* __restrict__ c, T* __restrict__ cz, DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    extern __shared__ unsigned char lds_raw[];
    T* xs = reinterpret_cast<T*>(lds_raw);
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        T r = 0;
        for (int j = 0; j < split; ++j) r += r0part[i * split + j];
        xs[i] = r;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (row >= nc) return;
    const T* ri = inv + (size_t)row * nc;
    T acc = 0;
    for (int j = lane; j < nc; j += 64) acc += ri[j] * xs[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) cz[row] = c[row] * acc;
}

// Coarse solve z0 = A0^{-1} r0 and cz[s] = c_s z0[s].
// mode 0: pivots + forward unit-L + backward U with the LU factors
// (src/precond/coarse_solve.h:50-67), column-oriented, one workgroup.
// mode 1: one dense mat-vec with the precomputed inverse (stored transposed).
template <typename T>
__global__ void __launch_bounds__(1024)
k_coarse_solve(int nc, int mode, int split, const T* __restrict__ LUt,
               const int* __restrict__ piv, const T* __restrict__ invT,
               const T* __restrict__ r0part, const T* __restrict__ c, T* __restrict__ cz,
               DevGuard guard) {
    if (guard_stop(guard, 0)) return;
    extern __shared__ unsigned char lds_raw[];
    T* xs = reinterpret_cast<T*>(lds_raw);
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        T r = 0;
        for (int j = 0; j < split; ++j) r += r0part[i * split + j];
        xs[i] = r;
    }
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

// Ring encoding of one subdomain's sweeps (see k_schwarz_sweep_ring): column
// references become live ring slots (level-order position mod R) or ~index
// into the global copy (forward: the dependency's backward position; backward:
// its local row); padding becomes R.  bpos_out[p] = backward position of the
// row at forward position p.  Returns the smallest level distance of a global
// reference (INT_MAX when none).
static int ring_encode(const SweepPack& fw, const SweepPack& bw, int ns, int R,
                       std::vector<int>& fcol, std::vector<int>& bcol,
                       std::vector<int>& bpos_out) {
    std::vector<int> fpos(ns), bposv(ns), flv(ns), blv(ns);
    for (int l = 0; l < (int)fw.lev.size(); ++l)
        for (int q = 0; q < fw.lev[l].size; ++q) {
            const int r = fw.row[fw.lev[l].start + q];
            fpos[r] = fw.lev[l].start + q;
            flv[r] = l;
        }
    for (int l = 0; l < (int)bw.lev.size(); ++l)
        for (int q = 0; q < bw.lev[l].size; ++q) {
            const int r = bw.row[bw.lev[l].start + q];
            bposv[r] = bw.lev[l].start + q;
            blv[r] = l;
        }
    int F = INT_MAX;
    auto encode = [&](const SweepPack& P, bool fwd, std::vector<int>& out) {
        out = P.col;
        const std::vector<int>& pos = fwd ? fpos : bposv;
        const std::vector<int>& lv = fwd ? flv : blv;
        for (int l = 0; l < (int)P.lev.size(); ++l) {
            const Level& L = P.lev[l];
            for (int k = 0; k < L.K; ++k)
                for (int q = 0; q < L.size; ++q) {
                    const size_t idx = (size_t)L.ell + (size_t)k * L.size + q;
                    const int c = P.col[idx];
                    if (c == ns) {
                        out[idx] = R;
                        continue;
                    }
                    const int pc = pos[c];
                    if (pc >= L.start + L.size - R) {
                        out[idx] = pc % R;
                    } else {
                        out[idx] = ~(fwd ? bposv[c] : c);
                        F = std::min(F, l - lv[c]);
                    }
                }
        }
    };
    encode(fw, true, fcol);
    encode(bw, false, bcol);
    bpos_out.resize(ns);
    for (int p = 0; p < ns; ++p) bpos_out[p] = bposv[fw.row[p]];
    return F;
}

// pre: an already factored global matrix (ILUT) with its diagonal positions;
// its sweeps are built the same way and the device factorisation is skipped.
static void setup_ilu_batch(sg_precond* M, const sg_csr* A, bool additive,
                            const HostCsr* pre = nullptr, const std::vector<int>* pre_diag = nullptr) {
    sg_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const HostCsr& H = A->host;
    const int nsub = M->nparts;
    const bool global_ilu = (M->kind == SG_PC_ILU0 || pre);
    std::vector<HostCsr> local(nsub);
    std::vector<std::vector<int>> diag(nsub), own(nsub);
    std::vector<SweepPack> fw(nsub), bw(nsub);

    run_parallel(nsub, [&](int s) {
        const std::vector<int>& d = M->dofs[s];
        HostCsr& L = local[s];
        if (pre) {
            L = *pre;
            diag[s] = *pre_diag;
        } else {
            L = global_ilu ? H : extract_submatrix(H, d);
            for (int i = 0; i < L.n; ++i)
                for (int k = L.rp[i] + 1; k < L.rp[i + 1]; ++k)
                    if (L.ci[k] <= L.ci[k - 1])
                        throw std::runtime_error("ILU(0) needs strictly ascending column indices");
            diag[s] = diag_positions(L);
            for (int i = 0; i < L.n; ++i)
                if (diag[s][i] < 0) throw std::runtime_error("ILU(0) needs a diagonal entry in every row");
        }
        // owned rows in local numbering (src/precond/ras.h:34-41)
        for (int i = 0; i < L.n; ++i)
            if (global_ilu || M->part[d[i]] == s) own[s].push_back(i);
        pack_levels(L, diag[s], lower_levels(L), true, fw[s]);
        pack_levels(L, diag[s], upper_levels(L), false, bw[s]);
    });

    // LDS plan: level tables (when they take at most half the LDS), then z when
    // the largest subdomain fits, else a ring of recent values (every level
    // must fit in it)
    const bool fp32 = M->opts.precision == SG_FP32;
    const size_t tsz = fp32 ? 4 : 8;
    int maxlev = 0, max_ns = 0, max_w = 0;
    for (int s = 0; s < nsub; ++s) {
        maxlev = std::max(maxlev, (int)(fw[s].lev.size() + bw[s].lev.size()));
        max_ns = std::max(max_ns, local[s].n);
        for (const Level& lv : fw[s].lev) max_w = std::max(max_w, lv.size);
        for (const Level& lv : bw[s].lev) max_w = std::max(max_w, lv.size);
    }
    M->lds_tab_bytes = ((size_t)maxlev * sizeof(Level) + 15) / 16 * 16;
    M->tab_in_lds = M->lds_tab_bytes <= (size_t)ctx->max_lds / 2;
    if (!M->tab_in_lds) M->lds_tab_bytes = 0;
    const size_t zbytes = (size_t)(max_ns + 1) * tsz;
    M->z_in_lds = M->lds_tab_bytes + zbytes <= (size_t)ctx->max_lds;
    M->ring = false;
    M->ring_F = INT_MAX;
    if (!M->z_in_lds) {
        const long long R = (long long)(((size_t)ctx->max_lds - M->lds_tab_bytes) / tsz) - 1;
        if (R >= 64 && max_w <= R) {
            M->ring = true;
            M->ring_R = (int)R;
        }
    }

    // concatenate: every subdomain's offsets first (in subdomain order), then
    // each subdomain's segments filled in parallel
    Plan pl;
    pl.subs.resize(nsub);
    M->sub_nnz.assign(nsub, 0);
    M->levels_fwd = M->levels_bwd = 0;
    M->max_levels = 0;
    std::vector<std::vector<int>> rfc, rbc, rbp;
    if (M->ring) {
        rfc.resize(nsub);
        rbc.resize(nsub);
        rbp.resize(nsub);
        std::vector<int> rF(nsub);
        run_parallel(nsub, [&](int s) {
            rF[s] = ring_encode(fw[s], bw[s], local[s].n, M->ring_R, rfc[s], rbc[s], rbp[s]);
        });
        for (int s = 0; s < nsub; ++s) M->ring_F = std::min(M->ring_F, rF[s]);
    }
    struct Off {
        long long flev, blev, frow, brow, fcol, bcol, fsrc, bsrc, bdsrc, fbpos, rows, nnz, own;
    };
    std::vector<Off> off(nsub + 1, Off{});
    long long zoff = 0;
    for (int s = 0; s < nsub; ++s) {
        const HostCsr& L = local[s];
        if (!global_ilu && (int)M->dofs[s].size() != L.n) fail(SG_E_ARG, "subdomain size mismatch");
        const Off& o = off[s];
        Off& e = off[s + 1];
        e.flev = o.flev + (long long)fw[s].lev.size();
        e.blev = o.blev + (long long)bw[s].lev.size();
        e.frow = o.frow + (long long)fw[s].row.size();
        e.brow = o.brow + (long long)bw[s].row.size();
        e.fcol = o.fcol + (long long)(M->ring ? rfc[s].size() : fw[s].col.size());
        e.bcol = o.bcol + (long long)(M->ring ? rbc[s].size() : bw[s].col.size());
        e.fsrc = o.fsrc + (long long)fw[s].src.size();
        e.bsrc = o.bsrc + (long long)bw[s].src.size();
        e.bdsrc = o.bdsrc + (long long)bw[s].dsrc.size();
        e.fbpos = o.fbpos + (long long)(M->ring ? rbp[s].size() : 0);
        e.rows = o.rows + L.n;
        e.nnz = o.nnz + L.nnz();
        e.own = o.own + (long long)own[s].size();
        SubInfo& si = pl.subs[s];
        si.ns = L.n;
        si.row0 = (int)o.rows;
        si.zoff = (int)zoff;
        si.flev0 = (int)o.flev;
        si.nflev = (int)fw[s].lev.size();
        si.blev0 = (int)o.blev;
        si.nblev = (int)bw[s].lev.size();
        si.own0 = (int)o.own;
        si.nown = (int)own[s].size();
        M->sub_nnz[s] = L.nnz();
        M->levels_fwd += si.nflev;
        M->levels_bwd += si.nblev;
        M->max_levels = std::max(M->max_levels, std::max(si.nflev, si.nblev));
        pl.max_ns = std::max(pl.max_ns, L.n);
        zoff += L.n + 1;
    }
    const Off& tot = off[nsub];
    const long long row0 = tot.rows, nnz0 = tot.nnz;
    // the 64-bit offsets (the level fields are 32-bit) are checked before use
    if (nnz0 > INT32_MAX || tot.fcol > INT32_MAX || tot.bcol > INT32_MAX || zoff > INT32_MAX ||
        tot.frow > INT32_MAX || tot.brow > INT32_MAX)
        fail(SG_E_ARG, "subdomain data exceeds 32-bit indexing");
    pl.flev.resize(tot.flev);
    pl.blev.resize(tot.blev);
    pl.f_item.resize(tot.flev);
    pl.b_item.resize(tot.blev);
    pl.f_row.resize(tot.frow);
    pl.b_row.resize(tot.brow);
    pl.f_col.resize(tot.fcol);
    pl.b_col.resize(tot.bcol);
    pl.f_src.resize(tot.fsrc);
    pl.b_src.resize(tot.bsrc);
    pl.b_dsrc.resize(tot.bdsrc);
    if (M->ring) pl.f_bpos.resize(tot.fbpos);
    pl.lrp.resize(row0 + 1);
    pl.lrp[0] = 0;
    pl.ldiag.resize(row0);
    pl.lci.resize(nnz0);
    pl.lval.resize(nnz0);
    pl.dofs.resize(row0);
    pl.own.resize(tot.own);
    pl.own_gid.resize(tot.own);
    run_parallel(nsub, [&](int s) {
        const HostCsr& L = local[s];
        const Off& o = off[s];
        const int fpos = (int)o.frow, bpos = (int)o.brow, fell = (int)o.fcol, bell = (int)o.bcol;
        const int nz0 = (int)o.nnz;
        for (size_t l = 0; l < fw[s].lev.size(); ++l) {
            const Level lv = fw[s].lev[l];
            pl.flev[o.flev + l] = Level{lv.start + fpos, lv.size, lv.K, lv.ell + fell};
        }
        for (size_t l = 0; l < bw[s].lev.size(); ++l) {
            const Level lv = bw[s].lev[l];
            pl.blev[o.blev + l] = Level{lv.start + bpos, lv.size, lv.K, lv.ell + bell};
        }
        for (int dir = 0; dir < 2; ++dir) {
            const auto& lev = dir ? bw[s].lev : fw[s].lev;
            int* item = (dir ? pl.b_item.data() + o.blev : pl.f_item.data() + o.flev);
            int acc = 0;
            for (size_t l = 0; l < lev.size(); ++l) {
                item[l] = acc;
                acc += (lev[l].size + 63) / 64;
            }
        }
        std::copy(fw[s].row.begin(), fw[s].row.end(), pl.f_row.begin() + o.frow);
        std::copy(bw[s].row.begin(), bw[s].row.end(), pl.b_row.begin() + o.brow);
        if (M->ring) {
            std::copy(rfc[s].begin(), rfc[s].end(), pl.f_col.begin() + o.fcol);
            std::copy(rbc[s].begin(), rbc[s].end(), pl.b_col.begin() + o.bcol);
            std::copy(rbp[s].begin(), rbp[s].end(), pl.f_bpos.begin() + o.fbpos);
        } else {
            std::copy(fw[s].col.begin(), fw[s].col.end(), pl.f_col.begin() + o.fcol);
            std::copy(bw[s].col.begin(), bw[s].col.end(), pl.b_col.begin() + o.bcol);
        }
        for (size_t i = 0; i < fw[s].src.size(); ++i) {
            const int k = fw[s].src[i];
            pl.f_src[o.fsrc + i] = k >= 0 ? k + nz0 : -1;
        }
        for (size_t i = 0; i < bw[s].src.size(); ++i) {
            const int k = bw[s].src[i];
            pl.b_src[o.bsrc + i] = k >= 0 ? k + nz0 : -1;
        }
        for (size_t i = 0; i < bw[s].dsrc.size(); ++i) pl.b_dsrc[o.bdsrc + i] = bw[s].dsrc[i] + nz0;
        for (int i = 0; i < L.n; ++i) {
            pl.lrp[o.rows + 1 + i] = L.rp[i + 1] + nz0;
            pl.ldiag[o.rows + i] = diag[s][i] + nz0;
            pl.dofs[o.rows + i] = global_ilu ? i : M->dofs[s][i];
        }
        std::copy(L.ci.begin(), L.ci.begin() + L.nnz(), pl.lci.begin() + o.nnz);
        std::copy(L.v.begin(), L.v.begin() + L.nnz(), pl.lval.begin() + o.nnz);
        for (size_t q = 0; q < own[s].size(); ++q) {
            const int li = own[s][q];
            pl.own[o.own + q] = li;
            pl.own_gid[o.own + q] = global_ilu ? li : M->dofs[s][li];
        }
    });
    // padding read (never used) by the branch-free prefetch of idle threads
    pl.f_row.resize(pl.f_row.size() + kPadRows, 0);
    if (M->ring) pl.f_bpos.resize(pl.f_row.size(), 0);
    pl.b_row.resize(pl.b_row.size() + kPadRows, 0);
    pl.f_col.resize(pl.f_col.size() + kPadRows, 0);
    pl.b_col.resize(pl.b_col.size() + kPadRows, 0);
    pl.f_src.resize(pl.f_src.size() + kPadRows, -1);
    pl.b_src.resize(pl.b_src.size() + kPadRows, -1);
    pl.b_dsrc.resize(pl.b_dsrc.size() + kPadRows, -1);
    pl.total_rows = (int)row0;
    pl.order.resize(nsub);
    std::iota(pl.order.begin(), pl.order.end(), 0);
    std::stable_sort(pl.order.begin(), pl.order.end(),
                     [&](int a, int b) { return M->sub_nnz[a] > M->sub_nnz[b]; });

    M->nsub = nsub;
    M->max_ns = pl.max_ns;
    M->lrp_host = pl.lrp;
    {
        int kmax = 0;
        for (const Level& lv : pl.flev) kmax = std::max(kmax, lv.K);
        for (const Level& lv : pl.blev) kmax = std::max(kmax, lv.K);
        M->ring_kp = kmax > kKreg ? 2 * kKreg : kKreg;
    }
    M->d_subs.upload(pl.subs, st);
    M->d_order.upload(pl.order, st);
    M->flev.upload(pl.flev, st);
    M->blev.upload(pl.blev, st);
    M->d_fitem.upload(pl.f_item, st);
    M->d_bitem.upload(pl.b_item, st);
    M->f_row.upload(pl.f_row, st);
    M->f_col.upload(pl.f_col, st);
    M->f_src.upload(pl.f_src, st);
    M->b_row.upload(pl.b_row, st);
    M->b_col.upload(pl.b_col, st);
    M->b_src.upload(pl.b_src, st);
    M->b_dsrc.upload(pl.b_dsrc, st);
    M->d_dofs.upload(pl.dofs, st);
    if (M->ring) M->f_bpos.upload(pl.f_bpos, st);
    M->d_own.upload(pl.own, st);
    M->d_own_gid.upload(pl.own_gid, st);
    {
        std::vector<int> o2(2 * pl.own.size());
        for (size_t k = 0; k < pl.own.size(); ++k) {
            o2[2 * k] = pl.own[k];
            o2[2 * k + 1] = pl.own_gid[k];
        }
        M->d_own2.upload(o2, st);
    }
    M->lrp.upload(pl.lrp, st);
    M->lci.upload(pl.lci, st);
    M->ldiag.upload(pl.ldiag, st);
    M->lval.upload(pl.lval, st);

    // device factorisation in FP64
    if (!pre) {
        hipLaunchKernelGGL(k_ilu0_factor, dim3(nsub), dim3(kTrsvThreads), 0, st, M->d_subs.p,
                           M->flev.p, M->f_row.p, M->lrp.p, M->lci.p, M->ldiag.p, M->lval.p);
        SG_HIP(hipGetLastError());
    }
    // pack values for the sweeps
    int nf = (int)pl.f_src.size(), nb = (int)pl.b_src.size(), nd = (int)pl.b_dsrc.size();
    auto pack = [&](auto* dummy, sg::DevBuf<int>& src, auto& out, int count) {
        using T = std::remove_pointer_t<decltype(dummy)>;
        out.alloc(count);
        if (count)
            hipLaunchKernelGGL(k_pack<T>, dim3(grid_for(count)), dim3(256), 0, st, count, src.p,
                               M->lval.p, out.p);
    };
    if (fp32) {
        pack((float*)nullptr, M->f_src, M->f_v32, nf);
        pack((float*)nullptr, M->b_src, M->b_v32, nb);
        pack((float*)nullptr, M->b_dsrc, M->b_d32, nd);
    } else {
        pack((double*)nullptr, M->f_src, M->f_v64, nf);
        pack((double*)nullptr, M->b_src, M->b_v64, nb);
        pack((double*)nullptr, M->b_dsrc, M->b_d64, nd);
    }
    SG_HIP(hipGetLastError());

    M->lds_bytes = M->lds_tab_bytes + zbytes;
    M->ring_bytes = M->lds_tab_bytes + (size_t)(M->ring_R + 1) * tsz;
    M->zstride = (size_t)zoff;
    if (!M->z_in_lds) {  // global z: one copy, or zf | zb | zr for the ring sweep
        const size_t zn = (size_t)zoff * (M->ring ? 3 : 1);
        if (fp32) M->z32.alloc(zn);
        else M->z64.alloc(zn);
    }
    {
        auto al = [](size_t b) { return (b + 15) / 16 * 16; };
        M->sf_item_off = M->lds_tab_bytes;
        M->sf_z_off = M->sf_item_off + al((size_t)maxlev * sizeof(int));
        M->sf_flag_off = M->sf_z_off + al(zbytes);
        M->sf_bytes = M->sf_flag_off + al((size_t)pl.max_ns + 1);
        M->sf_ok = M->tab_in_lds && M->sf_bytes <= (size_t)ctx->max_lds;
    }
    {
        // staged sweep: chunk descriptors and two chunk buffers after the tables and z
        auto al = [](size_t b) { return (b + 15) / 16 * 16; };
        M->st_ok = false;
        M->st_ct_off = al(M->lds_tab_bytes + zbytes + tsz);  // z, zero slot, trash slot
        // records need 16-bit local rows and columns (ns + 1, the trash slot,
        // below 65536) and at most kKreg entries per row in every level
        bool fits = M->tab_in_lds && M->z_in_lds && pl.max_ns <= 65534;
        for (const Level& lv : pl.flev) fits = fits && lv.K <= kKreg;
        for (const Level& lv : pl.blev) fits = fits && lv.K <= kKreg;
        const int rq = fp32 ? stage_rec_quads<float>() : stage_rec_quads<double>();
        int kmax = 0;
        for (const Level& lv : pl.flev) kmax = std::max(kmax, lv.K);
        for (const Level& lv : pl.blev) kmax = std::max(kmax, lv.K);
        const int kr = !fp32 ? 4 : kmax <= 2 ? 2 : kmax <= 3 ? 3 : 4;
        const bool k3 = kr == 3;
        auto chunk_quads = [&](int nrow, int dir) {
            return kr == 2   ? (long long)stage_quads_n<float, 2>(nrow, dir != 0)
                 : kr == 3 ? (long long)stage_quads_n<float, 3>(nrow, dir != 0)
                           : (long long)nrow * rq;
        };
        std::vector<SChunk> ch;
        std::vector<int> c0(2 * (size_t)nsub), csub;
        long long quads = 0;
        int cap = 0;
        // the descriptor table's size depends on the buffer size and the
        // other way round: grow the table's share until the plan fits it
        size_t ct_bytes = 0;
        bool planned = false;
        for (int pass = 0; pass < 8 && fits && !planned; ++pass) {
            const long long room = (long long)ctx->max_lds - (long long)(M->st_ct_off + ct_bytes);
            cap = (int)std::min<long long>(kStageBytes, (room - 16) / 2 / 16 * 16);
            if (cap < 1024) break;
            const int capq = cap / 16;
            ch.clear();
            csub.clear();
            quads = 0;
            size_t max_nch = 0;
            for (int s = 0; s < nsub; ++s) {
                const SubInfo& si = pl.subs[s];
                c0[2 * s] = (int)ch.size();
                for (int dir = 0; dir < 2; ++dir) {
                    const int nl = dir ? si.nblev : si.nflev;
                    const Level* lv = dir ? &pl.blev[si.blev0] : &pl.flev[si.flev0];
                    // chunks of whole levels while they fit; a level larger than
                    // a buffer is cut into pieces of rows (its rows are independent)
                    int maxr = capq / rq;
                    while (chunk_quads(maxr + 1, dir) <= capq) ++maxr;
                    const int pos_end = nl ? lv[nl - 1].start + lv[nl - 1].size : 0;
                    int l = 0, pos = nl ? lv[0].start : 0;
                    while (pos < pos_end) {
                        while (lv[l].start + lv[l].size <= pos) ++l;
                        int end = pos, lk = l;
                        while (lk < nl && lv[lk].start + lv[lk].size - pos <= maxr)
                            end = lv[lk].start + lv[lk].size, ++lk;
                        if (end == pos) end = std::min(pos + maxr, lv[l].start + lv[l].size);
                        int la = l;
                        while (la < nl && lv[la].start < end) ++la;
                        ch.push_back(SChunk{l, la - l, pos, end - pos, dir, (int)quads, {0, 0}});
                        csub.push_back(s);
                        quads += chunk_quads(end - pos, dir);
                        pos = end;
                    }
                }
                c0[2 * s + 1] = (int)ch.size() - c0[2 * s];
                max_nch = std::max(max_nch, (size_t)c0[2 * s + 1]);
            }
            const size_t need = al(max_nch * sizeof(SChunk));
            if (need <= ct_bytes) planned = true;
            else ct_bytes = need;
        }
        if (planned && quads <= INT32_MAX) {
            M->st_ok = true;
            M->st_kr = kr;
            M->st_buf = cap;
            M->st_buf_off = M->st_ct_off + ct_bytes;
            M->st_bytes = M->st_buf_off + 2 * (size_t)cap + 16;  // + trash word
            M->d_chunks.upload(ch, st);
            M->d_chunk0.upload(c0, st);
            DevBuf<int> d_csub;
            d_csub.upload(csub, st);
            M->d_img.alloc((size_t)quads);
            if (kr == 2)
                hipLaunchKernelGGL((k_stage_image<float, 2>), dim3((int)ch.size()), dim3(256), 0,
                                   st, M->d_chunks.p, M->d_subs.p, d_csub.p, M->flev.p,
                                   M->blev.p, M->f_row.p, M->f_col.p, M->f_v32.p, M->b_row.p,
                                   M->b_col.p, M->b_v32.p, M->b_d32.p, M->d_img.p);
            else if (k3)
                hipLaunchKernelGGL((k_stage_image<float, 3>), dim3((int)ch.size()), dim3(256), 0,
                                   st, M->d_chunks.p, M->d_subs.p, d_csub.p, M->flev.p,
                                   M->blev.p, M->f_row.p, M->f_col.p, M->f_v32.p, M->b_row.p,
                                   M->b_col.p, M->b_v32.p, M->b_d32.p, M->d_img.p);
            else if (fp32)
                hipLaunchKernelGGL((k_stage_image<float, 4>), dim3((int)ch.size()), dim3(256), 0,
                                   st, M->d_chunks.p, M->d_subs.p, d_csub.p, M->flev.p,
                                   M->blev.p, M->f_row.p, M->f_col.p, M->f_v32.p, M->b_row.p,
                                   M->b_col.p, M->b_v32.p, M->b_d32.p, M->d_img.p);
            else
                hipLaunchKernelGGL((k_stage_image<double, 4>), dim3((int)ch.size()), dim3(256), 0,
                                   st, M->d_chunks.p, M->d_subs.p, d_csub.p, M->flev.p,
                                   M->blev.p, M->f_row.p, M->f_col.p, M->f_v64.p, M->b_row.p,
                                   M->b_col.p, M->b_v64.p, M->b_d64.p, M->d_img.p);
            SG_HIP(hipGetLastError());
            SG_HIP(hipStreamSynchronize(st));  // d_csub ends here
        }
    }
    {
        // staged ring sweep: the chunk descriptors and a ring of R values after
        // the tables, then two chunk buffers; every reference must stay in the
        // ring (no global copies)
        auto al = [](size_t b) { return (b + 15) / 16 * 16; };
        M->sr_ok = false;
        bool fits = M->ring && M->tab_in_lds;
        for (const Level& lv : pl.flev) fits = fits && lv.K <= kKreg;
        for (const Level& lv : pl.blev) fits = fits && lv.K <= kKreg;
        const int cap = kStageBytes;
        if (fits) {
            bool cr = true;
            for (const Level& lv : pl.flev) cr = cr && lv.K <= (fp32 ? 3 : 2);
            for (const Level& lv : pl.blev) cr = cr && lv.K <= (fp32 ? 3 : 2);
            const int rq = cr ? 2 : fp32 ? sring_rec_quads<float>() : sring_rec_quads<double>();
            auto ownq = [&](int nrow) {
                return fp32 ? sring_own_quads<float>(nrow) : sring_own_quads<double>(nrow);
            };
            const int capq = cap / 16, per_q = (int)(16 / tsz);
            std::vector<SChunkR> ch;
            std::vector<int> c0(4 * (size_t)nsub), csub;
            std::vector<int> fslot(pl.f_row.size(), 0), bslot(pl.b_row.size(), 0);
            long long qimg = 0, qown = 0;
            bool ok = true;
            for (int s = 0; s < nsub && ok; ++s) {
                const SubInfo& si = pl.subs[s];
                for (int dir = 0; dir < 2 && ok; ++dir) {
                    const int nl = dir ? si.nblev : si.nflev;
                    const Level* lv = dir ? &pl.blev[si.blev0] : &pl.flev[si.flev0];
                    c0[4 * s + 2 * dir] = (int)ch.size();
                    // chunks of whole levels while they fit; a level larger than
                    // a buffer is cut into pieces of rows (its rows are independent)
                    int maxr = 1;
                    while ((long long)(maxr + 1) * rq + ownq(maxr + 1) <= capq) ++maxr;
                    const int pos_beg = nl ? lv[0].start : 0;
                    const int pos_end = nl ? lv[nl - 1].start + lv[nl - 1].size : 0;
                    int l = 0, pos = pos_beg;
                    while (pos < pos_end && ok) {
                        while (lv[l].start + lv[l].size <= pos) ++l;
                        SChunkR c{l, 0, pos, 0, (int)qimg, (int)qown, dir, 0};
                        int end = pos;
                        int lk = l;
                        while (lk < nl && lv[lk].start + lv[lk].size - pos <= maxr) {
                            end = lv[lk].start + lv[lk].size;
                            ++lk;
                        }
                        if (end == pos) end = std::min(pos + maxr, lv[l].start + lv[l].size);
                        c.nrow = end - pos;
                        int la = l;
                        while (la < nl && lv[la].start < end) ++la;
                        c.nlev = la - l;
                        pos = end;
                        std::vector<int>& slot = dir ? bslot : fslot;
                        for (int p = 0; p < c.nrow; ++p)
                            slot[(size_t)c.row0 + p] = (int)(qown * per_q + p);
                        ch.push_back(c);
                        csub.push_back(s);
                        qimg += (long long)c.nrow * rq;
                        qown += ownq(c.nrow);
                    }
                    c0[4 * s + 2 * dir + 1] = (int)ch.size() - c0[4 * s + 2 * dir];
                }
            }
            ok = ok && qimg <= INT32_MAX && qown * per_q <= INT32_MAX;
            size_t max_nch = 0;
            for (int s = 0; s < nsub; ++s)
                max_nch = std::max(max_nch, (size_t)c0[4 * s + 1] + (size_t)c0[4 * s + 3]);
            const size_t ct_bytes = al(max_nch * sizeof(SChunkR));
            const size_t ring_off = M->lds_tab_bytes + ct_bytes;
            const long long room = (long long)ctx->max_lds - (long long)ring_off - 2LL * cap - 32;
            // compact records keep ring slots (and the zero slot R) in 16 bits
            const int R = (int)std::min<long long>(room / (long long)tsz - 2, cr ? 65534 : INT32_MAX);
            std::vector<std::vector<int>> fcs(nsub), bcs(nsub), bps(nsub);
            std::vector<int> Fs(nsub, 0);
            ok = ok && R >= 64 && max_w <= R;
            if (ok)
                run_parallel(nsub, [&](int s) {
                    Fs[s] = ring_encode(fw[s], bw[s], local[s].n, R, fcs[s], bcs[s], bps[s]);
                });
            for (int f : Fs) ok = ok && f == INT_MAX;
            if (ok) {
                std::vector<int> fdst(pl.f_row.size(), 0), fgid(pl.f_row.size(), 0),
                    bdst(pl.b_row.size(), 0), fc, bc;
                for (int s = 0; s < nsub; ++s) {
                    const SubInfo& si = pl.subs[s];
                    const int pf0 = si.nflev ? pl.flev[si.flev0].start : 0;
                    const int pb0 = si.nblev ? pl.blev[si.blev0].start : 0;
                    for (int p = 0; p < si.ns; ++p) {
                        const int row = fw[s].row[p];
                        fdst[(size_t)pf0 + p] = bslot[(size_t)pb0 + bps[s][p]];
                        fgid[(size_t)pf0 + p] = global_ilu ? row : M->dofs[s][row];
                        bdst[(size_t)pb0 + p] = bw[s].row[p];
                    }
                    fc.insert(fc.end(), fcs[s].begin(), fcs[s].end());
                    bc.insert(bc.end(), bcs[s].begin(), bcs[s].end());
                }
                M->sr_ok = true;
                M->sr_cr = cr;
                M->sr_R = R;
                M->sr_ring_off = ring_off;
                M->sr_buf_off = al(ring_off + (size_t)(R + 2) * tsz);
                M->sr_buf = cap;
                M->sr_bytes = M->sr_buf_off + 2 * (size_t)cap + 16;
                M->d_rchunks.upload(ch, st);
                M->d_rchunk0.upload(c0, st);
                M->d_fslot.upload(fslot, st);
                M->d_fgid.upload(fgid, st);
                M->d_rimg.alloc((size_t)qimg);
                M->d_rown.alloc((size_t)qown);
                DevBuf<int> d_csub, d_fdst, d_bdst, d_fc, d_bc;
                d_csub.upload(csub, st);
                d_fdst.upload(fdst, st);
                d_bdst.upload(bdst, st);
                d_fc.upload(fc, st);
                d_bc.upload(bc, st);
#define SG_IMG_RING(T, CR, V, B, D)                                                               \
    hipLaunchKernelGGL((k_stage_image_ring<T, CR>), dim3((int)ch.size()), dim3(256), 0, st,      \
                       M->d_rchunks.p, M->d_subs.p, d_csub.p, M->flev.p, M->blev.p, d_fdst.p,    \
                       d_bdst.p, d_fc.p, d_bc.p, V, B, D, R, M->d_rimg.p)
                if (fp32 && cr) {
                    SG_IMG_RING(float, true, M->f_v32.p, M->b_v32.p, M->b_d32.p);
                } else if (fp32) {
                    SG_IMG_RING(float, false, M->f_v32.p, M->b_v32.p, M->b_d32.p);
                } else if (cr) {
                    SG_IMG_RING(double, true, M->f_v64.p, M->b_v64.p, M->b_d64.p);
                } else {
                    SG_IMG_RING(double, false, M->f_v64.p, M->b_v64.p, M->b_d64.p);
                }
#undef SG_IMG_RING
                SG_HIP(hipGetLastError());
                SG_HIP(hipStreamSynchronize(st));  // the temporaries end here
            }
        }
    }
    if (additive) {
        if (fp32) M->zout32.alloc((size_t)pl.total_rows + 1);
        else M->zout64.alloc((size_t)pl.total_rows + 1);
    }
    {
        const int bytes = (int)std::min<size_t>((size_t)ctx->max_lds,
                                                std::max(M->lds_bytes, M->lds_tab_bytes));
#define SG_SET_LDS1(T, Z, MD, NT, D, TB)                                                          \
    SG_HIP(hipFuncSetAttribute((const void*)k_schwarz_sweep<T, Z, MD, NT, D, TB>,              \
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
#define SG_SET_LDS(T, Z, MD)                                                                      \
    SG_SET_LDS1(T, Z, MD, 256, 3, true) SG_SET_LDS1(T, Z, MD, 512, 3, true)                     \
    SG_SET_LDS1(T, Z, MD, 1024, 3, true) SG_SET_LDS1(T, Z, MD, 256, 6, true)                    \
    SG_SET_LDS1(T, Z, MD, 512, 6, true) SG_SET_LDS1(T, Z, MD, 256, 3, false)                    \
    SG_SET_LDS1(T, Z, MD, 512, 3, false) SG_SET_LDS1(T, Z, MD, 1024, 3, false)
        if (fp32) {
            SG_SET_LDS(float, true, 0) SG_SET_LDS(float, true, 1)
            SG_SET_LDS(float, false, 0) SG_SET_LDS(float, false, 1)
        } else {
            SG_SET_LDS(double, true, 0) SG_SET_LDS(double, true, 1)
            SG_SET_LDS(double, false, 0) SG_SET_LDS(double, false, 1)
        }
#undef SG_SET_LDS1
#undef SG_SET_LDS
        if (M->ring) {
            const int rb = (int)M->ring_bytes;
#define SG_SET_RING1(T, MD, NT, D, TB)                                                            \
    SG_HIP(hipFuncSetAttribute((const void*)k_schwarz_sweep_ring<T, MD, NT, D, TB>,             \
                               hipFuncAttributeMaxDynamicSharedMemorySize, rb));
#define SG_SET_RING8(T, MD, NT, D)                                                                \
    SG_HIP(hipFuncSetAttribute((const void*)k_schwarz_sweep_ring<T, MD, NT, D, true, 8>,        \
                               hipFuncAttributeMaxDynamicSharedMemorySize, rb));
#define SG_SET_RING(T, MD)                                                                        \
    SG_SET_RING1(T, MD, 256, 3, true) SG_SET_RING1(T, MD, 512, 3, true)                         \
    SG_SET_RING1(T, MD, 1024, 3, true) SG_SET_RING1(T, MD, 256, 6, true)                        \
    SG_SET_RING1(T, MD, 512, 6, true) SG_SET_RING1(T, MD, 256, 3, false)                        \
    SG_SET_RING1(T, MD, 512, 3, false) SG_SET_RING1(T, MD, 1024, 3, false)                      \
    SG_SET_RING8(T, MD, 256, 3) SG_SET_RING8(T, MD, 512, 3) SG_SET_RING8(T, MD, 1024, 3)        \
    SG_SET_RING8(T, MD, 256, 6) SG_SET_RING8(T, MD, 512, 6)
            if (fp32) {
                SG_SET_RING(float, 0) SG_SET_RING(float, 1)
            } else {
                SG_SET_RING(double, 0) SG_SET_RING(double, 1)
            }
#undef SG_SET_RING1
#undef SG_SET_RING8
#undef SG_SET_RING
        }
        if (M->sr_ok) {
            const int srb = (int)M->sr_bytes;
#define SG_SET_SR(T, MD, CR)                                                                      \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_staged_ring<T, MD, 256, CR>,                 \
                               hipFuncAttributeMaxDynamicSharedMemorySize, srb));               \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_staged_ring<T, MD, 512, CR>,                 \
                               hipFuncAttributeMaxDynamicSharedMemorySize, srb));               \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_staged_ring<T, MD, 1024, CR>,                \
                               hipFuncAttributeMaxDynamicSharedMemorySize, srb));
            if (fp32) {
                if (M->sr_cr) { SG_SET_SR(float, 0, true) SG_SET_SR(float, 1, true) }
                else { SG_SET_SR(float, 0, false) SG_SET_SR(float, 1, false) }
            } else {
                if (M->sr_cr) { SG_SET_SR(double, 0, true) SG_SET_SR(double, 1, true) }
                else { SG_SET_SR(double, 0, false) SG_SET_SR(double, 1, false) }
            }
#undef SG_SET_SR
        }
        if (M->st_ok) {
            const int stb = (int)M->st_bytes;
#define SG_SET_ST(T, MD, KR)                                                                      \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_staged<T, MD, 256, KR>,                      \
                               hipFuncAttributeMaxDynamicSharedMemorySize, stb));               \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_staged<T, MD, 512, KR>,                      \
                               hipFuncAttributeMaxDynamicSharedMemorySize, stb));               \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_staged<T, MD, 1024, KR>,                     \
                               hipFuncAttributeMaxDynamicSharedMemorySize, stb));
            if (M->st_kr == 2) {
                SG_SET_ST(float, 0, 2) SG_SET_ST(float, 1, 2)
            } else if (M->st_kr == 3) {
                SG_SET_ST(float, 0, 3) SG_SET_ST(float, 1, 3)
            } else if (fp32) {
                SG_SET_ST(float, 0, 4) SG_SET_ST(float, 1, 4)
            } else {
                SG_SET_ST(double, 0, 4) SG_SET_ST(double, 1, 4)
            }
#undef SG_SET_ST
        }
        if (M->sf_ok) {
            const int sfb = (int)M->sf_bytes;
#define SG_SET_SF(T, MD)                                                                          \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_syncfree<T, MD, 512>,                        \
                               hipFuncAttributeMaxDynamicSharedMemorySize, sfb));               \
    SG_HIP(hipFuncSetAttribute((const void*)k_sweep_syncfree<T, MD, 1024>,                       \
                               hipFuncAttributeMaxDynamicSharedMemorySize, sfb));
            if (fp32) {
                SG_SET_SF(float, 0) SG_SET_SF(float, 1)
            } else {
                SG_SET_SF(double, 0) SG_SET_SF(double, 1)
            }
#undef SG_SET_SF
        }
    }

    if (additive) {
        // cover lists: for global g, positions row0_s + local index for every
        // subdomain containing g, ascending s (asm.h:54-58 single-thread order)
        std::vector<int> cnt(M->n + 1, 0);
        for (int s = 0; s < nsub; ++s)
            for (int g : M->dofs[s]) cnt[g + 1]++;
        for (int g = 0; g < M->n; ++g) cnt[g + 1] += cnt[g];
        std::vector<int> pos(cnt[M->n]);
        std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
        for (int s = 0; s < nsub; ++s)
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
    // row s of A0 comes from subdomain s's owned rows alone: rows in parallel,
    // each summed in the owned rows' order (the same values on any thread count)
    run_parallel(nc, [&](int s) {
        std::vector<double> aphi(nc, 0.0);
        std::vector<char> seen(nc, 0);
        std::vector<int> touched;
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
    });
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
    DevBuf<double> invT;  // column j of A0^{-1} stored contiguously
    invT.alloc((size_t)nc * nc);
    hipLaunchKernelGGL(k_lu_inverse, dim3(nc), dim3(256), (size_t)nc * 8, st, nc, dA.p,
                       M->d_piv.p, invT.p);
    M->d_inv64.alloc((size_t)nc * nc);  // row-major A0^{-1}
    hipLaunchKernelGGL(k_transpose, dim3(grid_for((long long)nc * nc)), dim3(256), 0, st, nc,
                       invT.p, M->d_inv64.p);
    SG_HIP(hipGetLastError());
    M->d_c64.upload(c, st);
    // the restriction splits each subdomain over a fixed number of workgroups,
    // so its summation order does not depend on how many subdomains there are
    // (a system gives the same coarse right-hand side alone or in a batch)
    M->restrict_split = kRestrictSplit;
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
        M->d_r0_32.alloc((size_t)nc * M->restrict_split);
        M->d_cz32.alloc(nc);
    } else {
        M->d_r0_64.alloc((size_t)nc * M->restrict_split);
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
    M->bound_sys.clear();
    M->sweep_tuned = false;
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
            setup_ilu_batch(M, A, false);
            break;
        case SG_PC_ILUT: {
            // host factorisation, sequential over rows as the reference's
            // (src/precond/ilut.h:28-149); device sweeps for the apply
            if (!(M->opts.drop_tol >= 0.0) || M->opts.fill_factor < 1)
                fail(SG_E_ARG, "ILUT needs drop_tol >= 0 and fill_factor >= 1");
            M->nparts = 1;
            M->part.assign(M->n, 0);
            M->owned.assign(1, {});
            M->dofs.assign(1, {});
            M->owned[0].resize(M->n);
            std::iota(M->owned[0].begin(), M->owned[0].end(), 0);
            M->dofs[0] = M->owned[0];
            HostCsr F;
            std::vector<int> fd;
            ilut_factor(A->host, M->opts.drop_tol, M->opts.fill_factor, F, fd, M->ilut_l_nnz,
                        M->ilut_u_nnz);
            setup_ilu_batch(M, A, false, &F, &fd);
            break;
        }
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
                M->part = partition_graph(A->host, np, M->opts.ufactor, M->opts.weighted_partition != 0);
            }
            build_sets(A->host, M->part, np, std::max(0, M->opts.overlap), M->owned, M->dofs);
            for (int s = 0; s < np; ++s)
                if (M->owned[s].empty()) fail(SG_E_PARTITION, "partitioning failed: empty subdomain");
            bool additive = (M->kind == SG_PC_ASM || M->kind == SG_PC_TWOLEVEL_ASM);
            setup_ilu_batch(M, A, additive);
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
static void launch_sweep(sg_precond* M, const double* x, double* y, const DevGuard& guard,
                         const T* fv, const T* bv, const T* bd, T* zg, T* zo, const T* cz,
                         const CoarseFold<T>& fold) {
    const SweepShape sh = kSweepShapes[M->sweep_variant];
    if (ZLDS && sh.staged && M->st_ok) {
#define SG_ST(NT, KR)                                                                             \
    hipLaunchKernelGGL((k_sweep_staged<T, MODE, NT, KR>), dim3(M->nsub), dim3(NT), M->st_bytes, \
                       M->ctx->stream, M->d_subs.p, M->d_order.p, M->d_chunk0.p, M->d_chunks.p, \
                       (const stage_q*)M->d_img.p, M->flev.p, M->blev.p, M->d_dofs.p,          \
                       (const own_i2*)M->d_own2.p, x, y, zo, cz, fold, guard, (int)M->lds_tab_bytes, \
                       (int)M->st_ct_off, (int)M->st_buf_off, M->st_buf / 4,              \
                       M->clk_on ? M->d_clk.p : nullptr)
        int kr = 4;
        if constexpr (std::is_same<T, float>::value) kr = M->st_kr;
        if (kr == 2) {
            if constexpr (std::is_same<T, float>::value) {
                if (sh.threads == 1024) SG_ST(1024, 2);
                else if (sh.threads == 512) SG_ST(512, 2);
                else SG_ST(256, 2);
            }
        } else if (kr == 3) {
            if constexpr (std::is_same<T, float>::value) {
                if (sh.threads == 1024) SG_ST(1024, 3);
                else if (sh.threads == 512) SG_ST(512, 3);
                else SG_ST(256, 3);
            }
        } else if (sh.threads == 1024) SG_ST(1024, 4);
        else if (sh.threads == 512) SG_ST(512, 4);
        else SG_ST(256, 4);
#undef SG_ST
        return;
    }
    if (ZLDS && sh.syncfree && M->sf_ok) {
#define SG_SF(NT)                                                                                 \
    hipLaunchKernelGGL((k_sweep_syncfree<T, MODE, NT>), dim3(M->nsub), dim3(NT), M->sf_bytes,   \
                       M->ctx->stream, M->d_subs.p, M->d_order.p, M->flev.p, M->d_fitem.p,      \
                       M->f_row.p, M->f_col.p, fv, M->blev.p, M->d_bitem.p, M->b_row.p,         \
                       M->b_col.p, bv, bd, M->d_dofs.p, M->d_own.p, M->d_own_gid.p, x, y, zo,   \
                       cz, guard, (int)M->sf_item_off, (int)M->sf_z_off, (int)M->sf_flag_off)
        if (sh.threads == 1024) SG_SF(1024);
        else SG_SF(512);
#undef SG_SF
        return;
    }
    if (!ZLDS && sh.staged && M->sr_ok) {
#define SG_SR(NT, CR)                                                                             \
    hipLaunchKernelGGL((k_sweep_staged_ring<T, MODE, NT, CR>), dim3(M->nsub), dim3(NT), M->sr_bytes, \
                       M->ctx->stream, M->d_subs.p, M->d_order.p, M->d_rchunk0.p,               \
                       M->d_rchunks.p, (const stage_q*)M->d_rimg.p, (stage_q*)M->d_rown.p,      \
                       M->d_fslot.p, M->d_fgid.p, M->flev.p, M->blev.p, M->d_own.p,             \
                       M->d_own_gid.p, x, y, zg, M->zstride, zo, cz, guard,                    \
                       (int)M->lds_tab_bytes, (int)M->sr_ring_off, (int)M->sr_buf_off,          \
                       M->sr_buf / 16, M->sr_R)
        if (M->sr_cr) {
            if (sh.threads == 1024) SG_SR(1024, true);
            else if (sh.threads == 512) SG_SR(512, true);
            else SG_SR(256, true);
        } else if (sh.threads == 1024) SG_SR(1024, false);
        else if (sh.threads == 512) SG_SR(512, false);
        else SG_SR(256, false);
#undef SG_SR
        return;
    }
    if (!ZLDS && M->ring) {
        const int F = M->ring_F == INT_MAX ? INT_MAX : std::max(1, M->ring_F);
#define SG_RING_KP(NT, D, TB, KP)                                                                 \
    hipLaunchKernelGGL((k_schwarz_sweep_ring<T, MODE, NT, D, TB, KP>), dim3(M->nsub), dim3(NT),  \
                       M->ring_bytes, M->ctx->stream, M->d_subs.p, M->d_order.p, M->flev.p,     \
                       M->f_row.p, M->f_col.p, fv, M->f_bpos.p, M->blev.p, M->b_row.p,         \
                       M->b_col.p, bv, bd, M->d_dofs.p, M->d_own.p, M->d_own_gid.p, x, y, zg,   \
                       M->zstride, zo, cz, guard, (int)M->lds_tab_bytes, M->ring_R, F)
// levels with more than four entries per row (ILUT factors) prefetch eight
#define SG_RING(NT, D, TB)                                                                        \
    do {                                                                                          \
        if constexpr (TB) {                                                                       \
            if (M->ring_kp > kKreg) SG_RING_KP(NT, D, true, 8);                                   \
            else SG_RING_KP(NT, D, true, kKreg);                                                  \
        } else {                                                                                  \
            SG_RING_KP(NT, D, false, kKreg);                                                      \
        }                                                                                         \
    } while (0)
        const int nt = sh.syncfree ? 512 : sh.threads;
        if (!M->tab_in_lds) {
            if (nt == 256) SG_RING(256, 3, false);
            else if (nt == 1024) SG_RING(1024, 3, false);
            else SG_RING(512, 3, false);
        } else if (sh.depth == 6 && nt == 256) SG_RING(256, 6, true);
        else if (sh.depth == 6) SG_RING(512, 6, true);
        else if (nt == 256) SG_RING(256, 3, true);
        else if (nt == 1024) SG_RING(1024, 3, true);
        else SG_RING(512, 3, true);
#undef SG_RING
#undef SG_RING_KP
        return;
    }
    const size_t lds = ZLDS ? M->lds_bytes : M->lds_tab_bytes;
#define SG_SWEEP(NT, D, TB)                                                                       \
    hipLaunchKernelGGL((k_schwarz_sweep<T, ZLDS, MODE, NT, D, TB>), dim3(M->nsub), dim3(NT), lds, \
                       M->ctx->stream, M->d_subs.p, M->d_order.p, M->flev.p, M->f_row.p,        \
                       M->f_col.p, fv, M->blev.p, M->b_row.p, M->b_col.p, bv, bd, M->d_dofs.p,  \
                       M->d_own.p, M->d_own_gid.p, x, y, zg, zo, cz, guard,                     \
                       (int)M->lds_tab_bytes)
    const int nt = sh.syncfree ? 512 : sh.threads;
    if (!M->tab_in_lds) {
        if (nt == 256) SG_SWEEP(256, 3, false);
        else if (nt == 1024) SG_SWEEP(1024, 3, false);
        else SG_SWEEP(512, 3, false);
    } else if (sh.depth == 6 && nt == 256) SG_SWEEP(256, 6, true);
    else if (sh.depth == 6) SG_SWEEP(512, 6, true);
    else if (nt == 256) SG_SWEEP(256, 3, true);
    else if (nt == 1024) SG_SWEEP(1024, 3, true);
    else SG_SWEEP(512, 3, true);
#undef SG_SWEEP
}

template <typename T>
static void launch_any_sweep(sg_precond* M, const double* x, double* y, const DevGuard& guard,
                             const T* fv, const T* bv, const T* bd, T* zg, T* zo, const T* cz,
                             const CoarseFold<T>& fold = CoarseFold<T>{}) {
    const bool additive = (M->kind == SG_PC_ASM || M->kind == SG_PC_TWOLEVEL_ASM);
    if (additive) {
        if (M->z_in_lds) launch_sweep<T, true, 1>(M, x, y, guard, fv, bv, bd, zg, zo, cz, fold);
        else launch_sweep<T, false, 1>(M, x, y, guard, fv, bv, bd, zg, zo, cz, fold);
    } else {
        if (M->z_in_lds) launch_sweep<T, true, 0>(M, x, y, guard, fv, bv, bd, zg, zo, cz, fold);
        else launch_sweep<T, false, 0>(M, x, y, guard, fv, bv, bd, zg, zo, cz, fold);
    }
    SG_HIP(hipGetLastError());
}

// Whether shape v runs as itself on this preconditioner (a sync-free or staged
// shape whose LDS plan does not fit falls back to the barrier sweep).
static bool sweep_shape_available(const sg_precond* M, int v) {
    const SweepShape& sh = kSweepShapes[v];
    if (sh.syncfree) return M->sf_ok && M->z_in_lds;
    if (sh.staged) return M->z_in_lds ? M->st_ok : M->sr_ok;
    return true;
}

// Times each sweep shape (three launches each, on a zero input into scratch)
// and keeps the fastest; every shape computes the same values.
template <typename T>
static void sweep_autotune(sg_precond* M, const T* fv, const T* bv, const T* bd, T* zg, T* zo) {
    hipStream_t s = M->ctx->stream;
    uint64_t h = fnv1a(M->sub_nnz.data(), sizeof(int) * M->sub_nnz.size());
    h = fnv1a(M->lrp_host.data(), sizeof(int) * M->lrp_host.size(), h);
    char key[160];
    std::snprintf(key, sizeof key, "sweep:%d:%d:%d:%lld:%lld:%d:%016llx", (int)M->kind,
                  (int)M->opts.precision, M->nsub, M->levels_fwd, M->levels_bwd, M->max_ns,
                  (unsigned long long)h);
    const int cached = tune_lookup(M->ctx, key);
    if (cached >= 0 && cached < kNumSweepShapes) {
        M->sweep_variant = cached;
        M->sweep_tuned = true;
        return;
    }
    DevBuf<double> xin, yout;
    xin.alloc(std::max(M->n, 1));
    yout.alloc(std::max(M->n, 1));
    SG_HIP(hipMemsetAsync(xin.p, 0, sizeof(double) * xin.n, s));
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    float best = 0.f;
    int bestv = 1;
    for (int v = 0; v < kNumSweepShapes; ++v) {
        if (!sweep_shape_available(M, v)) continue;
        M->sweep_variant = v;
        const DevGuard none{nullptr, nullptr, 0};
        launch_any_sweep<T>(M, xin.p, yout.p, none, fv, bv, bd, zg, zo, nullptr);
        SG_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r)
            launch_any_sweep<T>(M, xin.p, yout.p, none, fv, bv, bd, zg, zo, nullptr);
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
    M->sweep_variant = bestv;
    M->sweep_tuned = true;
    tune_store(M->ctx, key, bestv);
}

template <typename T>
static void tune_sweep(sg_precond* M, const T* fv, const T* bv, const T* bd, T* zg, T* zo) {
    if (M->ctx->autotune == 1 && !M->sweep_tuned && !M->ctx->no_tune)
        sweep_autotune<T>(M, fv, bv, bd, zg, zo);
    if (M->ctx->autotune >= 2) M->sweep_variant = M->ctx->autotune - 2;
    else if (M->ctx->autotune == 0) M->sweep_variant = 1;
}

template <typename T>
static void apply_batch(sg_precond* M, const double* x, double* y, const Guard& guard,
                        const T* fv, const T* bv, const T* bd, T* zg, T* zo, const T* cz,
                        const CoarseFold<T>& fold = CoarseFold<T>{}) {
    hipStream_t st = M->ctx->stream;
    bool additive = (M->kind == SG_PC_ASM || M->kind == SG_PC_TWOLEVEL_ASM);
    tune_sweep<T>(M, fv, bv, bd, zg, zo);
    {
        ProfScope ps(M->ctx, SG_K_TRSV);
        launch_any_sweep<T>(M, x, y, dev_guard(guard, M->d_sub_sys.p), fv, bv, bd, zg, zo, cz,
                            fold);
    }
    if (additive) {
        hipLaunchKernelGGL(k_asm_gather<T>, dim3(grid_for(M->n)), dim3(256), 0, st, M->n,
                           M->cov_ptr.p, M->cov_pos.p, zo, M->d_part.p, cz, y,
                           dev_guard(guard, nullptr));
        SG_HIP(hipGetLastError());
    }
}

// Whether the staged sweep forms the coarse term itself (k_coarse_inv is
// skipped): two-level RAS with the coarse inverse, staged sweep of at most 512
// threads (the fold's registers), at most kFoldRows coarse rows.
static bool coarse_fold(const sg_precond* M) {
    const SweepShape& sh = kSweepShapes[M->sweep_variant];
    return M->kind == SG_PC_TWOLEVEL_RAS && M->opts.coarse_solve == 1 && M->z_in_lds &&
           M->st_ok && sh.staged && sh.threads <= 512 && M->nc <= kFoldRows &&
           M->restrict_split == kRestrictSplit;
}

template <typename T>
static const T* coarse_launch(sg_precond* M, const double* x, const Guard& guard, const T* LUt,
                              const T* inv, const T* c, T* r0, T* cz, bool fold) {
    hipStream_t st = M->ctx->stream;
    int nc = M->nc, split = M->restrict_split;
    hipLaunchKernelGGL(k_coarse_restrict<T>, dim3(nc * split), dim3(256), 0, st, split,
                       M->d_subs.p, M->d_own_gid.p, c, x, r0, dev_guard(guard, M->d_sub_sys.p));
    const DevGuard g0 = dev_guard(guard, nullptr);
    if (fold) {
        // the staged sweep forms cz itself
    } else if (M->opts.coarse_solve == 1)
        hipLaunchKernelGGL(k_coarse_inv<T>, dim3((nc + 3) / 4), dim3(256), (size_t)nc * sizeof(T),
                           st, nc, split, inv, r0, c, cz, g0);
    else
        hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(1024), (size_t)nc * sizeof(T), st, nc,
                           0, split, LUt, M->d_piv.p, inv, r0, c, cz, g0);
    SG_HIP(hipGetLastError());
    return cz;
}

}  // namespace sg

using namespace sg;

namespace sg {

bool precond_apply_az(sg_precond* M, const sg_csr* A, const double* x, double* y, double* az,
                      const Guard& guard) {
    if (M->learned) return learned_apply_az(M->learned, A, x, y, az, guard);
    precond_apply_launch(M, x, y, guard);
    return false;
}

void precond_apply_launch(sg_precond* M, const double* x, double* y, const Guard& guard) {
    if (M->learned) {
        learned_apply(M->learned, x, y, guard);
        return;
    }
    sg_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    ProfScope ps(ctx, SG_K_PRECOND);
    switch (M->kind) {
        case SG_PC_NONE:
            SG_HIP(hipMemcpyAsync(y, x, sizeof(double) * M->n, hipMemcpyDeviceToDevice, st));
            return;
        case SG_PC_JACOBI:
            hipLaunchKernelGGL(k_jacobi, dim3(grid_for(M->n)), dim3(256), 0, st, M->n,
                               M->inv_diag.p, x, y, dev_guard(guard, nullptr));
            SG_HIP(hipGetLastError());
            return;
        default:
            break;
    }
    if (M->fn) {
        // host callback (guard honoured on the host; never batched)
        if (guard.flag) {
            int g = 0;
            SG_HIP(hipMemcpyAsync(&g, guard.flag, sizeof(int), hipMemcpyDeviceToHost, st));
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
    // the sweep shape is settled first: it decides whether the coarse term is folded
    precond_ensure_tuned(M);
    const bool fold = two && coarse_fold(M);
    if (M->opts.precision == SG_FP32) {
        const float* cz = two ? coarse_launch<float>(M, x, guard, M->d_LUt32.p, M->d_inv32.p,
                                                      M->d_c32.p, M->d_r0_32.p, M->d_cz32.p, fold)
                              : nullptr;
        const CoarseFold<float> cf =
            fold ? CoarseFold<float>{M->d_inv32.p, M->d_r0_32.p, M->d_c32.p, M->nc} : CoarseFold<float>{};
        apply_batch<float>(M, x, y, guard, M->f_v32.p, M->b_v32.p, M->b_d32.p, M->z32.p,
                           M->zout32.p, fold ? nullptr : cz, cf);
    } else {
        const double* cz = two ? coarse_launch<double>(M, x, guard, M->d_LUt64.p, M->d_inv64.p,
                                                        M->d_c64.p, M->d_r0_64.p, M->d_cz64.p, fold)
                               : nullptr;
        const CoarseFold<double> cf =
            fold ? CoarseFold<double>{M->d_inv64.p, M->d_r0_64.p, M->d_c64.p, M->nc} : CoarseFold<double>{};
        apply_batch<double>(M, x, y, guard, M->f_v64.p, M->b_v64.p, M->b_d64.p, M->z64.p,
                            M->zout64.p, fold ? nullptr : cz, cf);
    }
}

void precond_bind_systems(sg_precond* M, const std::vector<int>& sys_off) {
    if (M->learned) fail(SG_E_ARG, "the learned stage does not run in batched solves");
    if (M->bound_sys == sys_off) return;
    std::vector<int> map(std::max<size_t>(M->dofs.size(), 1), -1);
    for (size_t q = 0; q < M->dofs.size(); ++q) {
        const auto& d = M->dofs[q];
        if (d.empty()) continue;
        // dofs are ascending: the subdomain lies in one system iff its first
        // and last DOF do
        const int a = (int)(std::upper_bound(sys_off.begin(), sys_off.end(), d.front()) -
                            sys_off.begin()) - 1;
        const int b = (int)(std::upper_bound(sys_off.begin(), sys_off.end(), d.back()) -
                            sys_off.begin()) - 1;
        map[q] = a == b ? a : -1;
    }
    M->d_sub_sys.upload(map, M->ctx->stream);
    SG_HIP(hipStreamSynchronize(M->ctx->stream));
    M->bound_sys = sys_off;
}

void precond_ensure_tuned(sg_precond* M) {
    if (M->learned) {
        if (M->ready) learned_prepare(M->learned);
        return;
    }
    if (M->fn || M->kind == SG_PC_NONE || M->kind == SG_PC_JACOBI || !M->ready) return;
    if (M->opts.precision == SG_FP32)
        tune_sweep<float>(M, M->f_v32.p, M->b_v32.p, M->b_d32.p, M->z32.p, M->zout32.p);
    else
        tune_sweep<double>(M, M->f_v64.p, M->b_v64.p, M->b_d64.p, M->z64.p, M->zout64.p);
}

bool precond_is_callback(const sg_precond* M) {
    return M->fn != nullptr || (M->learned && precond_is_callback(learned_base(M->learned)));
}

bool precond_is_flexible(const sg_precond* M) { return M->learned != nullptr; }

int precond_n(const sg_precond* M) { return M->n; }
sg_context* precond_ctx(const sg_precond* M) { return M->ctx; }
bool precond_ready(const sg_precond* M) { return M->ready; }

}  // namespace sg

extern "C" {

sg_status sg_precond_create(sg_context* ctx, sg_precond_kind kind, const sg_precond_opts* opts,
                            sg_precond** out) {
    if (!ctx || !out || (int)kind < 0 || (int)kind > SG_PC_ILUT || kind == SG_PC_LEARNED)
        return SG_E_ARG;
    sg_precond* M = new sg_precond();
    M->ctx = ctx;
    M->kind = kind;
    if (opts) M->opts = *opts;
    else {
        M->opts.nparts = 4;
        M->opts.overlap = 1;
    }
    if (kind == SG_PC_ILUT && M->opts.fill_factor == 0 && M->opts.drop_tol == 0.0) {
        M->opts.drop_tol = 1e-4;  // ILUTPrecond() defaults (src/precond/ilut.h:25)
        M->opts.fill_factor = 10;
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

sg_status sg_precond_create_learned(sg_context* ctx, sg_precond* base, sg_unet* net,
                                    const double* log_kappa, const sg_learned_opts* opts,
                                    sg_precond** out) {
    if (!ctx || !base || !net || !log_kappa || !opts || !out) return SG_E_ARG;
    *out = nullptr;
    if (base->learned) return SG_E_ARG;
    sg_precond* M = new sg_precond();
    sg_status st = run_guarded(ctx, [&] {
        M->ctx = ctx;
        M->kind = SG_PC_LEARNED;
        M->learned = learned_new(ctx, base, net, opts->nx, opts->ny, log_kappa, opts);
        M->n = opts->nx * opts->ny;
    });
    if (st != SG_OK) {
        delete M;
        return st;
    }
    *out = M;
    return SG_OK;
}

sg_status sg_precond_learned_stats(const sg_precond* M, long long* calls, long long* accepted) {
    if (!M || !M->learned) return SG_E_ARG;
    return run_guarded(M->ctx, [&] { learned_counts(M->learned, calls, accepted); });
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
    if (M->learned)
        return run_guarded(M->ctx, [&] {
            learned_setup(M->learned, A);
            M->n = A->nrows;
            M->ready = true;
        });
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
    return run_guarded(M->ctx, [&] { precond_apply_launch(M, d_x, d_y, Guard{}); });
}

sg_status sg_precond_sweep_clock(sg_precond* M, const double* d_x, double* d_y,
                                 long long* rec, int cap, int* nrec, double* rate_khz) {
    if (!M || !rec || !nrec || !rate_khz) return SG_E_ARG;
    if (M->learned || !M->ready || M->nsub == 0) {
        if (M->ctx) M->ctx->err = "sweep clock: needs a set-up subdomain preconditioner";
        return SG_E_STATE;
    }
    if (!d_x || !d_y || cap < kClockWords * M->nsub) return SG_E_ARG;
    return run_guarded(M->ctx, [&] {
        const size_t words = kClockWords * (size_t)M->nsub;
        if (M->d_clk.n < words) M->d_clk.alloc(words);
        SG_HIP(hipMemsetAsync(M->d_clk.p, 0, words * sizeof(long long), M->ctx->stream));
        M->clk_on = true;
        try {
            precond_apply_launch(M, d_x, d_y, Guard{});
        } catch (...) {
            M->clk_on = false;
            throw;
        }
        M->clk_on = false;
        SG_HIP(hipMemcpyAsync(rec, M->d_clk.p, words * sizeof(long long), hipMemcpyDeviceToHost,
                              M->ctx->stream));
        SG_HIP(hipStreamSynchronize(M->ctx->stream));
        int cnt = 0;
        for (int b = 0; b < M->nsub; ++b) cnt += rec[kClockWords * (size_t)b + 3] != 0;
        *nrec = cnt;
        int khz = 0;
        SG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, M->ctx->device));
        *rate_khz = khz;
    });
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

sg_status sg_precond_ilut_nnz(const sg_precond* M, long long* l_nnz, long long* u_nnz) {
    if (!M || M->kind != SG_PC_ILUT) return SG_E_ARG;
    if (!M->ready) return SG_E_STATE;
    if (l_nnz) *l_nnz = M->ilut_l_nnz;
    if (u_nnz) *u_nnz = M->ilut_u_nnz;
    return SG_OK;
}

sg_status sg_partition(int n, const int* row_ptr, const int* col_idx, const double* vals,
                       int nparts, int ufactor, int weighted, int* part) {
    if (n < 0 || !row_ptr || !part || (row_ptr[n] > 0 && !col_idx) ||
        (weighted && row_ptr[n] > 0 && !vals))
        return SG_E_ARG;
    return run_guarded(nullptr, [&] {
        HostCsr H;
        H.n = n;
        H.ncols = n;
        H.rp.assign(row_ptr, row_ptr + n + 1);
        H.ci.assign(col_idx, col_idx + row_ptr[n]);
        if (weighted) H.v.assign(vals, vals + row_ptr[n]);
        std::vector<int> p = partition_graph(H, nparts, ufactor, weighted != 0);
        std::copy(p.begin(), p.end(), part);
    });
}

sg_status sg_precond_stats(const sg_precond* M, sg_precond_info* info) {
    if (!M || !info) return SG_E_ARG;
    if (M->learned) return sg_precond_stats(learned_base(M->learned), info);
    if (!M->ready) return SG_E_STATE;
    *info = sg_precond_info{};
    info->nsub = M->nsub;
    info->z_in_lds = M->z_in_lds ? 1 : 0;
    info->nc = M->nc;
    info->max_ns = M->max_ns;
    // a shape that falls back runs the barrier sweep of its thread count (512
    // for the sync-free shapes)
    int shape = M->sweep_variant;
    if (!sweep_shape_available(M, shape))
        shape = kSweepShapes[shape].syncfree || kSweepShapes[shape].threads == 512
                    ? 1
                    : (kSweepShapes[shape].threads == 256 ? 0 : 2);
    info->sweep_shape = M->nsub && (M->sweep_tuned || M->ctx->autotune != 1) ? shape : -1;
    for (size_t s = 0; s < M->dofs.size() && M->nsub; ++s) info->n_ext += (long long)M->dofs[s].size();
    for (int v : M->sub_nnz) info->nnz_ext += v;
    info->levels_fwd = M->levels_fwd;
    info->levels_bwd = M->levels_bwd;
    info->max_levels = M->max_levels;
    return SG_OK;
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
