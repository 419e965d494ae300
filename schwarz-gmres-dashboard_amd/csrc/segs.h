// Block tables of the deterministic reductions shared by the Krylov solvers.
//
// A vector kernel runs over a table of blocks; each block covers one
// contiguous range inside one system of a (possibly batched) block-diagonal
// matrix and writes per-block partial sums.  A "finish" kernel with one
// workgroup per system adds its blocks' partials in a fixed order.  The
// ranges of a system depend only on that system's size, so a system's sums
// are the same whether it is solved alone or in a batch.
#pragma once

#include "sg_internal.h"

#include <algorithm>
#include <vector>

namespace sg {

constexpr int kMaxBlocks = 4096;  // blocks per system and split

// One block of a split: elements [start, end) of system sys.
struct BlockRange {
    int sys, start, end, pad;
};
struct SysBlocks {
    int blk0, nblk;
};
struct Segs {
    const BlockRange* blk;
    const SysBlocks* sys;
};

// Block count and chunk of one system of n elements (the chunk is a multiple
// of the tile so every block but the last runs whole tiles).
inline void split_shape(int n, int max_blocks, int tile, int& nblk, int& chunk) {
    int want = std::max(1, (n + tile - 1) / tile);
    nblk = std::min(want, std::min(max_blocks, kMaxBlocks));
    int c = (n + nblk - 1) / nblk;
    chunk = (c + tile - 1) / tile * tile;
    nblk = std::max(1, (n + chunk - 1) / chunk);
}

struct SegsHost {
    DevBuf<BlockRange> blk;
    DevBuf<SysBlocks> sys;
    int nblk = 0;
    Segs dev() const { return Segs{blk.p, sys.p}; }
};

// pass > 0: a system of n elements may use up to ceil(n / pass) blocks when
// that exceeds max_blocks, so its blocks stay within `pass` elements (up to
// kMaxBlocks blocks).  The table of a system depends only on its size.
inline void build_segs(const std::vector<int>& off, int max_blocks, int tile, SegsHost& out,
                       hipStream_t s, int pass = 0) {
    std::vector<BlockRange> br;
    std::vector<SysBlocks> sb;
    for (size_t q = 0; q + 1 < off.size(); ++q) {
        const int n = off[q + 1] - off[q];
        int nblk, chunk;
        const int mb = pass > 0 ? std::max(max_blocks, (n + pass - 1) / pass) : max_blocks;
        split_shape(n, mb, tile, nblk, chunk);
        sb.push_back(SysBlocks{(int)br.size(), nblk});
        for (int b = 0; b < nblk; ++b)
            br.push_back(BlockRange{(int)q, off[q] + b * chunk,
                                    off[q] + std::min(n, (b + 1) * chunk), 0});
    }
    out.nblk = (int)br.size();
    out.blk.upload(br, s);
    out.sys.upload(sb, s);
    SG_HIP(hipStreamSynchronize(s));
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block partial sums of up to two columns: column c of block blockIdx.x goes to
// part[c * pstride + blockIdx.x].  Every thread of the NT-thread block calls it.
template <int NT>
__device__ inline void block_partials2(double a, double b, double* part, int pstride,
                                       int ncols) {
    __shared__ double red[NT / 64][2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    a = wave_sum(a);
    if (ncols > 1) b = wave_sum(b);
    if (lane == 0) {
        red[wv][0] = a;
        red[wv][1] = b;
    }
    __syncthreads();
    if ((int)threadIdx.x < ncols) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) s += red[w][threadIdx.x];
        part[(size_t)threadIdx.x * pstride + blockIdx.x] = s;
    }
}

// out[c] (c < ncols <= 2) = the sum of one system's block partials of column
// c in block order: one wavefront per column, 16 loads in flight per lane.
// Every thread of the block calls it and receives the sums.
__device__ inline void finish_sums2(const double* part, int pstride, const SysBlocks& sb,
                                    int ncols, double* out) {
    __shared__ double res[2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (wv < ncols) {
        const double* pc = part + (size_t)wv * pstride + sb.blk0;
        double s = 0.0;
        for (int b0 = 0; b0 < sb.nblk; b0 += 16 * 64) {
            double v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int b = b0 + q * 64 + lane;
                v[q] = b < sb.nblk ? pc[b] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) s += v[q];
        }
        s = wave_sum(s);
        if (lane == 0) res[wv] = s;
    }
    __syncthreads();
    for (int c = 0; c < ncols; ++c) out[c] = res[c];
}

// Ends a stream capture that an exception interrupted, so the stream stays usable.
struct CaptureGuard {
    hipStream_t s;
    bool on = false;
    ~CaptureGuard() {
        if (on) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
        }
    }
};

// Profiling off for the scope (event records cannot join a captured graph).
struct ProfOff {
    sg_context* c;
    bool was;
    explicit ProfOff(sg_context* ctx) : c(ctx), was(ctx->profiling) { c->profiling = false; }
    ~ProfOff() { c->profiling = was; }
};

}  // namespace sg
