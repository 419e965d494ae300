// The learned residual-to-error network of the learned FGMRES stage: forward,
// backward and Adam on the GPU.  The architecture (a 4-level U-Net, channels
// 16-32-64-64, 3x3 convolutions, ReLU, 2x2 average pooling, nearest
// upsampling with skip concatenation, 1x1 head) and the parameter packing are
// those of the CPU oracle that the tests compare against.
//
// Storage: activations and activation gradients in bfloat16, NHWC (the
// channels of a pixel contiguous, so a 16-byte load carries 8 channels); a
// decoder's concatenation buffer holds the upsampled and the skip channels of
// each pixel side by side (training; inference reads the upsampled half
// straight from the half-resolution tensor).  Weights as an fp32 master copy
// plus bfloat16 copies laid out for the kernels; every sum in fp32.
//
// 3x3 convolutions (forward, and the data gradient as a convolution with the
// flipped, transposed weights) are implicit GEMMs on the matrix cores with
// v_mfma_f32_16x16x32_bf16: A = weights (16 output channels x 32 of K), B =
// input (32 of K x 16 pixels of one image row), K running over (tap, input
// channel) of a staged chunk.  A workgroup computes a 4 RW x 32 pixel tile
// for up to 64 output channels (each wavefront RW = 1 or 2 rows: RW x 2 pixel
// blocks x 4 channel blocks of accumulators); the layer's whole padded input
// (16..128 channels, with the one-pixel halo) and its weights are staged in
// LDS with 16-byte loads as one chunk.  2x2 pooling and the 1x1 head run in
// the epilogues of the convolutions that feed them.
// The weight gradient runs on the vector ALUs in fp32 (each thread one
// (output, input) channel pair, a 3x3 window slid along the tile rows) and
// writes per-tile partial sums that a second kernel adds in a fixed order.
#include "sg_internal.h"

#include <cmath>
#include <cstring>

namespace sg {

typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

struct UConv {
    int cin, cout;
};
constexpr int kUConvs = 14;
constexpr UConv kUSpec[kUConvs] = {{2, 16},  {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                   {64, 64}, {64, 64}, {64, 64}, {128, 64}, {64, 64},
                                   {96, 32}, {32, 32}, {48, 16}, {16, 16}};

inline int uconv_offset(int i) {
    int o = 0;
    for (int k = 0; k < i; ++k) o += kUSpec[k].cout * kUSpec[k].cin * 9 + kUSpec[k].cout;
    return o;
}
inline int unet_nparams_() { return uconv_offset(kUConvs) + 17; }
inline int cin_pad(int c) { return (c + 15) / 16 * 16; }

// A channel window [c0, c0 + C) of a [B][H][W][ctot] bfloat16 tensor (NHWC;
// ctot and c0 are multiples of 8).
struct View {
    uint16_t* p;
    int ctot, c0;
    __device__ size_t at(int b, int c, int y, int x, int H, int W) const {
        return (((size_t)b * H + y) * W + x) * ctot + c0 + c;
    }
};

// ---- 3x3 convolution on the matrix cores ------------------------------------
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
constexpr int kTX = 32;                   // output tile: 4 RW rows x 32 pixels (RW rows per wave)
constexpr int kHX = kTX + 2;

// Input channels per staged chunk (CC): the whole padded input of every layer
// of this network (16, 32, 48, 64, 96 or 128) is one chunk; the output
// channels per workgroup are capped so that the chunk and its weights fit the
// LDS (64 up to CC = 64, 32 at 96, 16 at 128).
__host__ __device__ constexpr int cob_max(int cc) { return cc >= 128 ? 16 : (cc >= 96 ? 32 : 64); }

__device__ __forceinline__ v8bf lds_v8(const uint16_t* p) {
    return __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(p));
}

// LDS of one workgroup: the staged input chunk of C channels (pixel stride
// sp = C + 8) and the chunk's weights for cob output channels (row stride swp).
struct ConvLds {
    int sp, swp, bytes;
};
__host__ __device__ constexpr ConvLds conv_lds(int C, int cob, int rw) {
    return ConvLds{C + 8, (9 * C + 31) / 32 * 32 + 8,
                   2 * ((4 * rw + 2) * kHX * (C + 8) + cob * ((9 * C + 31) / 32 * 32 + 8))};
}

// MODE 0: out = relu(acc + bias); MODE 1: out = acc (data gradient).
// Weights wt: [9][cout][cinp] bfloat16 (cinp = cin rounded up to 16; the
// input view holds at least cinp channels, zero beyond cin).  CC divides cinp.
// Grid: x = B * tilesY * tilesX, y = output-channel blocks of cob (16, 32 or
// 64, at most cob_max(CC)); dynamic LDS conv_lds(CC, cob, RW).bytes.  A tile
// is 4 RW rows x 32 pixels: wavefront w computes rows RW w .. RW w + RW - 1.
//
// Fusions (forward, MODE 0): channels [0, cup) of the input are read from the
// half-resolution view `up` at (y/2, x/2) (nearest upsampling of a decoder
// input, cup a multiple of 8); EPI 1 also writes the 2x2
// average pooling of the (bf16-rounded) output to `pout`; EPI 2 also writes
// the 1x1 head hout = hw[16] + sum_c hw[c] out[c] over the bf16-rounded
// outputs (cout = 16).  The pooling adds the four outputs of a 2x2 block in
// row-major order and scales by 1/4 before rounding to bf16.
template <int MODE, int EPI, int CC, int RW>
__global__ void __launch_bounds__(256)
k_conv3x3(View in, View up, int cup, int cin, int cinp, const uint16_t* __restrict__ wt,
          const float* __restrict__ bias, View out, int cout, int H, int W, int cob, View pout,
          const float* __restrict__ hw, float* __restrict__ hout, const int* __restrict__ stop) {
    if (stop && *stop == 0) return;
    extern __shared__ __attribute__((aligned(16))) uint16_t lds_u16[];
    constexpr int kTY = 4 * RW, kHY = kTY + 2;
    constexpr ConvLds L = conv_lds(CC, cob_max(CC), RW);
    constexpr int kSP = L.sp, kSWP = L.swp, OCT = CC / 8, KC = (9 * CC + 31) / 32 * 32;
    uint16_t* s_in = lds_u16;
    uint16_t* s_w = lds_u16 + kHY * kHX * kSP;
    const int tilesX = (W + kTX - 1) / kTX, tilesY = (H + kTY - 1) / kTY;
    int t = blockIdx.x;
    const int tx = t % tilesX;
    t /= tilesX;
    const int ty = t % tilesY;
    const int b = t / tilesY;
    const int x0 = tx * kTX, y0 = ty * kTY;
    const int co0 = blockIdx.y * cob;
    const int ncot = min(cob, cout - co0) / 16;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    v4f acc[RW][2][4];
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int n = 0; n < 4; ++n) acc[r][p][n] = v4f{0.f, 0.f, 0.f, 0.f};
    // input chunk c0 with its halo and the chunk's weights, every 16-byte load
    // issued before the first LDS store; branch-free: every lane loads (an
    // invalid element from the tensor's first word) and zeroes what lies
    // outside the image or the channels
    constexpr int kJI = (kHY * kHX * OCT + 255) / 256;
    constexpr int kJW = (cob_max(CC) * 9 * OCT + 255) / 256;
    for (int c0 = 0; c0 < cinp; c0 += CC) {
        uint4 ri[kJI], rw[kJW];
#pragma unroll
        for (int jj = 0; jj < kJI; ++jj) {
            const int e = tid + 256 * jj;
            const int q = e % OCT, pix = e / OCT;
            const int gy = y0 + pix / kHX - 1, gx = x0 + pix % kHX - 1;
            const int c = c0 + 8 * q;
            const bool ok = e < kHY * kHX * OCT && c < cin && gy >= 0 && gy < H && gx >= 0 && gx < W;
            const bool fu = c < cup;  // from the half-resolution view
            const uint16_t* src = fu ? up.p : in.p;
            const int sct = fu ? up.ctot : in.ctot, sc0 = fu ? up.c0 : in.c0, sh = fu ? 1 : 0;
            const size_t off = ok ? (((size_t)b * (H >> sh) + (gy >> sh)) * (W >> sh) + (gx >> sh)) * sct +
                                        sc0 + c
                                  : 0;
            const uint4 v = *reinterpret_cast<const uint4*>(src + off);
            ri[jj] = ok ? v : uint4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int jj = 0; jj < kJW; ++jj) {
            const int e = tid + 256 * jj;
            const int q = e % OCT, r = e / OCT;
            const bool ok = e < ncot * 16 * 9 * OCT;
            const size_t off = ok ? ((size_t)(r % 9) * cout + co0 + r / 9) * cinp + c0 + 8 * q : 0;
            const uint4 v = *reinterpret_cast<const uint4*>(wt + off);
            rw[jj] = ok ? v : uint4{0u, 0u, 0u, 0u};
        }
        __syncthreads();  // the previous chunk's MFMAs are done with the LDS
#pragma unroll
        for (int jj = 0; jj < kJI; ++jj) {
            const int e = tid + 256 * jj;
            if (e < kHY * kHX * OCT)
                *reinterpret_cast<uint4*>(&s_in[(e / OCT) * kSP + 8 * (e % OCT)]) = ri[jj];
        }
        // weights: s_w[co][tap * CC + ch]; K beyond 9 CC is zero
#pragma unroll
        for (int jj = 0; jj < kJW; ++jj) {
            const int e = tid + 256 * jj;
            if (e < ncot * 16 * 9 * OCT) {
                const int q = e % OCT, r = e / OCT;
                *reinterpret_cast<uint4*>(&s_w[(r / 9) * kSWP + (r % 9) * CC + 8 * q]) = rw[jj];
            }
        }
        constexpr int kPer = (KC - 9 * CC) / 8;
        if (kPer > 0)
            for (int e = tid; e < ncot * 16 * kPer; e += 256)
                *reinterpret_cast<uint4*>(&s_w[(e / kPer) * kSWP + 9 * CC + 8 * (e % kPer)]) =
                    uint4{0u, 0u, 0u, 0u};
        __syncthreads();
        for (int ks = 0; ks < KC; ks += 32) {
            // this lane's 8 K values: one tap, channels ch..ch+7 (CC is a multiple of 8)
            const int k = ks + 8 * g;
            const int tap = k / CC, ch = k - tap * CC;
            const bool live = tap < 9;  // the zero-padded tail of K
            const int tc = live ? tap : 8;
            const int ky = tc / 3, kx = tc % 3;
            v8bf bfr[2][2];
#pragma unroll
            for (int r = 0; r < RW; ++r)
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const int py = RW * wv + r + ky, px = 16 * p + li + kx;
                    const v8bf v = lds_v8(&s_in[(py * kHX + px) * kSP + (live ? ch : 0)]);
                    bfr[r][p] = live ? v : v8bf{};
                }
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                if (n >= ncot) break;
                const v8bf afr = lds_v8(&s_w[(n * 16 + li) * kSWP + k]);
#pragma unroll
                for (int r = 0; r < RW; ++r)
#pragma unroll
                    for (int p = 0; p < 2; ++p)
                        acc[r][p][n] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[r][p], acc[r][p][n], 0, 0, 0);
            }
        }
    }
    // D[co][pixel]: lane holds pixel li, channels 4 g + i of each block
    uint32_t hv[RW][2][4][2];  // bf16 outputs, two per word: [r][p][n][i / 2]
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                if (n >= ncot) break;
                const int co = co0 + n * 16 + 4 * g;
                uint16_t h[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float v = acc[r][p][n][i];
                    if (MODE == 0) {
                        v += bias[co + i];
                        v = v > 0.f ? v : 0.f;
                    }
                    h[i] = f2bf(v);
                }
                hv[r][p][n][0] = (unsigned)h[0] | ((unsigned)h[1] << 16);
                hv[r][p][n][1] = (unsigned)h[2] | ((unsigned)h[3] << 16);
                const int y = y0 + RW * wv + r, x = x0 + 16 * p + li;
                if (y < H && x < W)
                    *reinterpret_cast<uint2*>(out.p + out.at(b, co, y, x, H, W)) =
                        uint2{hv[r][p][n][0], hv[r][p][n][1]};
            }
    auto unpack = [](uint32_t w, int k) { return bf2f((uint16_t)(k ? (w >> 16) : (w & 0xffffu))); };
    if (EPI == 1 && RW == 2) {
        // rows 2 wv, 2 wv + 1 are in this lane, columns x, x + 1 in lanes li, li ^ 1
        const int y = y0 + 2 * wv;  // even
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                if (n >= ncot) break;
                uint16_t o[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float a = unpack(hv[0][p][n][i / 2], i & 1);
                    const float d = unpack(hv[1][p][n][i / 2], i & 1);
                    const float c = __shfl_xor(a, 1), f = __shfl_xor(d, 1);
                    o[i] = f2bf(0.25f * (a + c + d + f));
                }
                const int x = x0 + 16 * p + li;
                if ((li & 1) == 0 && y + 1 < H && x + 1 < W)
                    *reinterpret_cast<uint2*>(pout.p + pout.at(b, co0 + n * 16 + 4 * g, y >> 1, x >> 1,
                                                              H >> 1, W >> 1)) =
                        uint2{(unsigned)o[0] | ((unsigned)o[1] << 16),
                              (unsigned)o[2] | ((unsigned)o[3] << 16)};
            }
    }
    if (EPI == 1 && RW == 1) {
        // the four rows of the tile sit in four wavefronts: the bf16 outputs go
        // through the LDS (free now), then each thread pools 4 channels of one
        // output pixel, adding in the same order as above
        __syncthreads();  // every MFMA operand read of the LDS is done
        uint16_t* s_o = lds_u16;  // [4 rows][32 pixels][cob channels]
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                if (n >= ncot) break;
                *reinterpret_cast<uint2*>(&s_o[(wv * kTX + 16 * p + li) * cob + n * 16 + 4 * g]) =
                    uint2{hv[0][p][n][0], hv[0][p][n][1]};
            }
        __syncthreads();
        const int q4 = ncot * 4;  // 4-channel groups per pixel
        for (int e = tid; e < 2 * (kTX / 2) * q4; e += 256) {
            const int cq = e % q4, px = (e / q4) % (kTX / 2), pr = e / (q4 * (kTX / 2));
            const int y = y0 + 2 * pr, x = x0 + 2 * px;
            if (y + 1 >= H || x + 1 >= W) continue;
            auto at = [&](int rr, int xx) {
                return *reinterpret_cast<const uint2*>(&s_o[((2 * pr + rr) * kTX + 2 * px + xx) * cob + 4 * cq]);
            };
            const uint2 A0 = at(0, 0), C0 = at(0, 1), D0 = at(1, 0), F0 = at(1, 1);
            uint16_t o[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t wa = i < 2 ? A0.x : A0.y, wc = i < 2 ? C0.x : C0.y;
                const uint32_t wd = i < 2 ? D0.x : D0.y, wf = i < 2 ? F0.x : F0.y;
                const float a = unpack(wa, i & 1), c = unpack(wc, i & 1);
                const float d = unpack(wd, i & 1), f = unpack(wf, i & 1);
                o[i] = f2bf(0.25f * (a + c + d + f));
            }
            *reinterpret_cast<uint2*>(pout.p + pout.at(b, co0 + 4 * cq, y >> 1, x >> 1, H >> 1, W >> 1)) =
                uint2{(unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16)};
        }
    }
    if (EPI == 2) {
        // cout = 16: channels 4 g + i of pixel li sit in lane li + 16 g
#pragma unroll
        for (int r = 0; r < RW; ++r)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                float ch[16];
#pragma unroll
                for (int gg = 0; gg < 4; ++gg)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        ch[4 * gg + i] = __shfl(unpack(hv[r][p][0][i / 2], i & 1), li + 16 * gg);
                const int y = y0 + RW * wv + r, x = x0 + 16 * p + li;
                if (g == 0 && y < H && x < W) {
                    float sum = hw[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) sum += hw[k] * ch[k];
                    hout[((size_t)b * H + y) * W + x] = sum;
                }
            }
    }
}

// ---- weight gradient (fp32 vector ALUs) ---------------------------------------
// Tile: 8 rows x 32 columns of one image.  Block (co16, ci16, tile): thread
// (co, ci) = (t >> 4, t & 15) accumulates the 9 taps over the tile; thread
// ci = 0 also sums the bias gradient.  part[tile][tap][cout][cin] (+ bias part).
constexpr int kWY = 8, kWX = 32;
__global__ void __launch_bounds__(256)
k_conv_dw(View in, int cin, View dp, int cout, int B, int H, int W, float* __restrict__ part,
          float* __restrict__ part_b) {
    __shared__ float s_g[16][kWY][kWX];
    __shared__ float s_x[16][kWY + 2][kWX + 2 + 1];
    const int tilesX = (W + kWX - 1) / kWX, tilesY = (H + kWY - 1) / kWY;
    int t = blockIdx.x;
    const int tile = t;
    const int tx = t % tilesX;
    t /= tilesX;
    const int ty = t % tilesY;
    const int b = t / tilesY;
    const int x0 = tx * kWX, y0 = ty * kWY;
    const int cob = blockIdx.y * 16, cib = blockIdx.z * 16;
    for (int e = threadIdx.x; e < 16 * kWY * kWX; e += 256) {
        const int c = e % 16, x = (e / 16) % kWX, y = e / (16 * kWX);
        const int gy = y0 + y, gx = x0 + x;
        s_g[c][y][x] = (gy < H && gx < W) ? bf2f(dp.p[dp.at(b, cob + c, gy, gx, H, W)]) : 0.f;
    }
    for (int e = threadIdx.x; e < 16 * (kWY + 2) * (kWX + 2); e += 256) {
        const int c = e % 16, x = (e / 16) % (kWX + 2), y = e / (16 * (kWX + 2));
        const int gy = y0 + y - 1, gx = x0 + x - 1;
        float v = 0.f;
        if (cib + c < cin && gy >= 0 && gy < H && gx >= 0 && gx < W)
            v = bf2f(in.p[in.at(b, cib + c, gy, gx, H, W)]);
        s_x[c][y][x] = v;
    }
    __syncthreads();
    const int co = threadIdx.x >> 4, ci = threadIdx.x & 15;
    float acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.f;
    float bsum = 0.f;
    for (int y = 0; y < kWY; ++y) {
        float w0[3], w1[3];  // window columns x, x+1 of rows y, y+1, y+2
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            w0[k] = s_x[ci][y + k][0];
            w1[k] = s_x[ci][y + k][1];
        }
        for (int x = 0; x < kWX; ++x) {
            float w2[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) w2[k] = s_x[ci][y + k][x + 2];
            const float g = s_g[co][y][x];
            bsum += g;
#pragma unroll
            for (int ky = 0; ky < 3; ++ky) {
                acc[ky * 3 + 0] += g * w0[ky];
                acc[ky * 3 + 1] += g * w1[ky];
                acc[ky * 3 + 2] += g * w2[ky];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                w0[k] = w1[k];
                w1[k] = w2[k];
            }
        }
    }
    if (cib + ci < cin) {
        float* pt = part + (size_t)tile * 9 * cout * cin;
#pragma unroll
        for (int k = 0; k < 9; ++k) pt[((size_t)k * cout + cob + co) * cin + cib + ci] = acc[k];
    }
    if (blockIdx.z == 0 && ci == 0) part_b[(size_t)tile * cout + cob + co] = bsum;
}

// grad (packed W[co][ci][ky][kx], then b) = sum over tiles, in tile order.
__global__ void k_dw_reduce(int ntile, int cin, int cout, const float* __restrict__ part,
                            const float* __restrict__ part_b, float* __restrict__ grad) {
    const int nw = 9 * cout * cin;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nw + cout; e += gridDim.x * blockDim.x) {
        float s = 0.f;
        if (e < nw) {
            const int tap = e % 9, ci = (e / 9) % cin, co = e / (9 * cin);
            for (int t = 0; t < ntile; ++t) s += part[((size_t)t * 9 + tap) * cout * cin + (size_t)co * cin + ci];
        } else {
            const int co = e - nw;
            for (int t = 0; t < ntile; ++t) s += part_b[(size_t)t * cout + co];
        }
        grad[e] = s;
    }
}

// ---- elementwise / layout kernels ---------------------------------------------
// network input: [B][2][H][W] fp32 -> bf16 [B][H][W][16] (channels 2..15
// zero), one pixel (two 16-byte words) per thread
__global__ void k_input(int B, int H, int W, const float* __restrict__ in, View x,
                        const int* __restrict__ stop) {
    if (stop && *stop == 0) return;
    const size_t hw = (size_t)H * W, n = (size_t)B * hw;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const size_t p = e % hw;
        const int b = (int)(e / hw);
        const unsigned c0 = f2bf(in[(size_t)b * 2 * hw + p]), c1 = f2bf(in[((size_t)b * 2 + 1) * hw + p]);
        uint4* o = reinterpret_cast<uint4*>(x.p + e * 16);
        o[0] = uint4{c0 | (c1 << 16), 0u, 0u, 0u};
        o[1] = uint4{0u, 0u, 0u, 0u};
    }
}

// dx (H x W) += 0.25 * dy(y/2, x/2), rounded
__global__ void k_pool_bwd(int B, int C, int H, int W, View dy, View dx) {
    const size_t n = (size_t)B * C * H * W;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C), x = (int)((e / C) % W), y = (int)((e / ((size_t)C * W)) % H);
        const int b = (int)(e / ((size_t)W * H * C));
        const size_t i = dx.at(b, c, y, x, H, W);
        dx.p[i] = f2bf(bf2f(dx.p[i]) + 0.25f * bf2f(dy.p[dy.at(b, c, y / 2, x / 2, H / 2, W / 2)]));
    }
}

// out channels [0, C) at H x W = nearest upsampling of a (H/2 x W/2), eight
// channels (one 16-byte word) per thread
__global__ void k_up(int B, int C, int H, int W, View a, View out, const int* __restrict__ stop) {
    if (stop && *stop == 0) return;
    const int oc = C / 8;
    const size_t n = (size_t)B * oc * H * W;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int q = (int)(e % oc), x = (int)((e / oc) % W), y = (int)((e / ((size_t)oc * W)) % H);
        const int b = (int)(e / ((size_t)W * H * oc));
        *reinterpret_cast<uint4*>(out.p + out.at(b, 8 * q, y, x, H, W)) =
            *reinterpret_cast<const uint4*>(a.p + a.at(b, 8 * q, y / 2, x / 2, H / 2, W / 2));
    }
}

// da (H/2 x W/2) = sum of each 2x2 block of dy, rounded
__global__ void k_up_bwd(int B, int C, int H, int W, View dy, View da) {
    const int h = H / 2, w = W / 2;
    const size_t n = (size_t)B * C * h * w;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C), x = (int)((e / C) % w), y = (int)((e / ((size_t)C * w)) % h);
        const int b = (int)(e / ((size_t)w * h * C));
        // row-major order within the block, as the oracle accumulates
        float s = bf2f(dy.p[dy.at(b, c, 2 * y, 2 * x, H, W)]);
        s += bf2f(dy.p[dy.at(b, c, 2 * y, 2 * x + 1, H, W)]);
        s += bf2f(dy.p[dy.at(b, c, 2 * y + 1, 2 * x, H, W)]);
        s += bf2f(dy.p[dy.at(b, c, 2 * y + 1, 2 * x + 1, H, W)]);
        da.p[da.at(b, c, y, x, h, w)] = f2bf(s);
    }
}

// dy *= (y > 0)
__global__ void k_relu_mask(int B, int C, int H, int W, View y, View dy) {
    const size_t n = (size_t)B * C * H * W;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C), x = (int)((e / C) % W), yy = (int)((e / ((size_t)C * W)) % H);
        const int b = (int)(e / ((size_t)W * H * C));
        if (!(bf2f(y.p[y.at(b, c, yy, x, H, W)]) > 0.f)) dy.p[dy.at(b, c, yy, x, H, W)] = 0;
    }
}

// per-sample sums t.t and (out - t).(out - t): block partials [b][blk][2]
constexpr int kLB = 64;  // blocks per sample
__global__ void __launch_bounds__(256)
k_loss_part(int H, int W, const float* __restrict__ out, const float* __restrict__ tg,
            double* __restrict__ part) {
    const int b = blockIdx.y;
    const size_t hwn = (size_t)H * W;
    double tt = 0.0, ee = 0.0;
    for (size_t p = blockIdx.x * 256 + threadIdx.x; p < hwn; p += kLB * 256) {
        const double t = tg[b * hwn + p], d = (double)out[b * hwn + p] - t;
        tt += t * t;
        ee += d * d;
    }
    __shared__ double r0[256], r1[256];
    r0[threadIdx.x] = tt;
    r1[threadIdx.x] = ee;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            r0[threadIdx.x] += r0[threadIdx.x + s];
            r1[threadIdx.x] += r1[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((size_t)b * kLB + blockIdx.x) * 2] = r0[0];
        part[((size_t)b * kLB + blockIdx.x) * 2 + 1] = r1[0];
    }
}

// sums[b] = {tt, ee}; loss = mean_b ee/tt
__global__ void k_loss_fin(int B, const double* __restrict__ part, double* __restrict__ sums,
                           double* __restrict__ loss) {
    if (threadIdx.x != 0) return;
    double l = 0.0;
    for (int b = 0; b < B; ++b) {
        double tt = 0.0, ee = 0.0;
        for (int k = 0; k < kLB; ++k) {
            tt += part[((size_t)b * kLB + k) * 2];
            ee += part[((size_t)b * kLB + k) * 2 + 1];
        }
        sums[2 * b] = tt;
        sums[2 * b + 1] = ee;
        l += ee / tt / B;
    }
    *loss = l;
}

// loss kind 0: dout = 2 (out - t) / (tt B)
__global__ void k_mse_dout(int B, int H, int W, const float* __restrict__ out,
                           const float* __restrict__ tg, const double* __restrict__ sums,
                           float* __restrict__ dout) {
    const size_t hwn = (size_t)H * W, n = (size_t)B * hwn;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(e / hwn);
        dout[e] = (float)(2.0 * ((double)out[e] - (double)tg[e]) / sums[2 * b] / B);
    }
}

// loss kind 1: u = A v per sample, A the five-point finite-volume diffusion
// operator of kappa = exp(in[b][1]) (harmonic faces; a boundary face adds k_p)
__global__ void k_stencil(int B, int H, int W, const float* __restrict__ in,
                          const float* __restrict__ v, float* __restrict__ u) {
    const size_t hwn = (size_t)H * W, n = (size_t)B * hwn;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const size_t p = e % hwn;
        const int b = (int)(e / hwn);
        const int y = (int)(p / W), x = (int)(p % W);
        const float* lk = in + ((size_t)b * 2 + 1) * hwn;
        const float* vb = v + (size_t)b * hwn;
        const float kp = expf(lk[p]), vp = vb[p];
        float acc = 0.f;
        const int dy[4] = {-1, 0, 0, 1}, dx[4] = {0, -1, 1, 0};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int yy = y + dy[f], xx = x + dx[f];
            if (yy < 0 || yy >= H || xx < 0 || xx >= W) {
                acc += kp * vp;
            } else {
                const size_t q = (size_t)yy * W + xx;
                const float kq = expf(lk[q]);
                acc += 2.f * kp * kq / (kp + kq) * (vp - vb[q]);
            }
        }
        u[e] = acc;
    }
}

// per-sample block partials of s.s, s.u, u.u (s = in[b][0])
__global__ void __launch_bounds__(256)
k_rloss_part(int H, int W, const float* __restrict__ in, const float* __restrict__ u,
             double* __restrict__ part) {
    const int b = blockIdx.y;
    const size_t hwn = (size_t)H * W;
    const float* sb = in + (size_t)b * 2 * hwn;
    double ss = 0.0, a = 0.0, c = 0.0;
    for (size_t p = blockIdx.x * 256 + threadIdx.x; p < hwn; p += kLB * 256) {
        const double sv = sb[p], uv = u[b * hwn + p];
        ss += sv * sv;
        a += sv * uv;
        c += uv * uv;
    }
    __shared__ double r[3][256];
    r[0][threadIdx.x] = ss;
    r[1][threadIdx.x] = a;
    r[2][threadIdx.x] = c;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int k = 0; k < 3; ++k) r[k][threadIdx.x] += r[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[((size_t)b * kLB + blockIdx.x) * 3 + k] = r[k][0];
}

// loss = mean_b (1 - a^2/(ss c)) [+ mu ln(c / ss)^2 for loss kind 2];
// coef[b] = {f, q, h}: dL/du = f (s - q u) + h u
constexpr double kScaleMu = 0.01;  // weight of the scale term of loss kind 2
__global__ void k_rloss_fin(int B, int kind, const double* __restrict__ part,
                            double* __restrict__ coef, double* __restrict__ loss) {
    if (threadIdx.x != 0) return;
    double l = 0.0;
    for (int b = 0; b < B; ++b) {
        double ss = 0.0, a = 0.0, c = 0.0;
        for (int k = 0; k < kLB; ++k) {
            ss += part[((size_t)b * kLB + k) * 3];
            a += part[((size_t)b * kLB + k) * 3 + 1];
            c += part[((size_t)b * kLB + k) * 3 + 2];
        }
        const bool ok = ss > 0.0 && c > 0.0;
        l += ok ? (1.0 - a * a / (ss * c)) / B : 1.0 / B;
        coef[3 * b] = ok ? -2.0 * a / (ss * c) / B : 0.0;
        coef[3 * b + 1] = ok ? a / c : 0.0;
        coef[3 * b + 2] = 0.0;
        if (kind == 2 && ok) {
            const double lg = log(c / ss);
            l += kScaleMu * lg * lg / B;
            coef[3 * b + 2] = 4.0 * kScaleMu * lg / c / B;
        }
    }
    *loss = l;
}

__global__ void k_rloss_g(int B, int H, int W, const float* __restrict__ in,
                          const float* __restrict__ u, const double* __restrict__ coef,
                          float* __restrict__ g) {
    const size_t hwn = (size_t)H * W, n = (size_t)B * hwn;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const size_t p = e % hwn;
        const int b = (int)(e / hwn);
        const double sv = in[(size_t)b * 2 * hwn + p], uv = u[e];
        g[e] = (float)(coef[3 * b] * (sv - coef[3 * b + 1] * uv) + coef[3 * b + 2] * uv);
    }
}

// dd0[b][c][p] = bf16(dout * hw[c]); head partial gradients per block [blk][17]
__global__ void __launch_bounds__(256)
k_head_bwd(int B, int H, int W, const float* __restrict__ dout, View d0,
           const float* __restrict__ hw, View dd0, float* __restrict__ hpart) {
    const size_t hwn = (size_t)H * W, n = (size_t)B * hwn;
    float acc[17];
    for (int c = 0; c < 17; ++c) acc[c] = 0.f;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const size_t p = e % hwn;
        const int b = (int)(e / hwn);
        const float g = dout[e];
        acc[16] += g;
        const uint16_t* px = d0.p + (((size_t)b * hwn + p) * d0.ctot + d0.c0);
        uint16_t* qx = dd0.p + (((size_t)b * hwn + p) * dd0.ctot + dd0.c0);
        for (int c = 0; c < 16; ++c) {
            acc[c] += g * bf2f(px[c]);
            qx[c] = f2bf(g * hw[c]);
        }
    }
    __shared__ float red[256];
    for (int c = 0; c < 17; ++c) {
        red[threadIdx.x] = acc[c];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) hpart[(size_t)blockIdx.x * 17 + c] = red[0];
        __syncthreads();
    }
}

__global__ void k_head_reduce(int nblk, const float* __restrict__ hpart, float* __restrict__ grad) {
    const int c = threadIdx.x;
    if (c >= 17) return;
    float s = 0.f;
    for (int k = 0; k < nblk; ++k) s += hpart[(size_t)k * 17 + c];
    grad[c] = s;
}

// weights for the kernels from the fp32 master:
// fwd[tap][co][cinp] = W[co][ci][tap]; bwd[tap][ci][cout] = W[co][ci][8 - tap]
__global__ void k_pack_weights(int cin, int cinp, int cout, const float* __restrict__ w,
                               uint16_t* __restrict__ fwd, uint16_t* __restrict__ bwd) {
    const int n = 9 * cout * cinp;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int ci = e % cinp, co = (e / cinp) % cout, tap = e / (cinp * cout);
        fwd[e] = ci < cin ? f2bf(w[((size_t)co * cin + ci) * 9 + tap]) : (uint16_t)0;
    }
    if (!bwd) return;
    const int m = 9 * cin * cout;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const int co = e % cout, ci = (e / cout) % cin, tap = e / (cout * cin);
        bwd[e] = f2bf(w[((size_t)co * cin + ci) * 9 + (8 - tap)]);
    }
}

__global__ void k_head_pack(const float* __restrict__ w, float* __restrict__ hw) {
    const int c = threadIdx.x;
    if (c < 16) hw[c] = bf2f(f2bf(w[c]));
    if (c == 16) hw[16] = w[16];
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1,
                       float c2, float wd) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = wd > 0.f ? p[i] - lr * wd * p[i] : p[i];
        p[i] = pi - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

static int grid1(size_t n) {
    size_t g = (n + 255) / 256;
    return (int)std::max<size_t>(1, std::min<size_t>(g, 16384));
}

}  // namespace sg

using namespace sg;

struct sg_unet {
    sg_context* ctx = nullptr;
    int np = 0;
    DevBuf<float> p, g, m, v;           // master params, gradient, Adam moments
    DevBuf<uint16_t> wf[kUConvs], wb[kUConvs];
    DevBuf<float> hw;                   // head weights (bf16-rounded) + bias
    bool packed = false;
    long long adam_t = 0;
    // activations for the current shape
    int B = 0, H = 0, W = 0;
    DevBuf<uint16_t> x, a0, cat0, p0, a2, cat1, p1, a4, cat2, p2, a6, a7, a8, a9, a10, a11, a12, a13;
    DevBuf<uint16_t> gx, ga0, gcat0, gp0, ga2, gcat1, gp1, ga4, gcat2, gp2, ga6, ga7, ga8, ga9, ga10,
        ga11, ga12, ga13;
    DevBuf<float> out, dwpart, dwpart_b, hpart, dout, lu, lg;
    DevBuf<double> lpart, lsums, lloss;
    bool grads_alloc = false;
    int loss_kind = 0;
    float wd = 0.f;  // decoupled weight decay of the Adam step
};

namespace sg {

template <int MODE, int EPI>
static void conv_allow_lds();

static void unet_pack(sg_unet* U) {
    if (U->packed) return;
    conv_allow_lds<0, 0>();
    conv_allow_lds<0, 1>();
    conv_allow_lds<0, 2>();
    conv_allow_lds<1, 0>();
    hipStream_t s = U->ctx->stream;
    for (int i = 0; i < kUConvs; ++i) {
        const int ci = kUSpec[i].cin, co = kUSpec[i].cout, cp = cin_pad(ci);
        if (!U->wf[i].p) U->wf[i].alloc((size_t)9 * co * cp);
        if (i > 0 && !U->wb[i].p) U->wb[i].alloc((size_t)9 * ci * co);
        hipLaunchKernelGGL(k_pack_weights, dim3(grid1((size_t)9 * co * cp)), dim3(256), 0, s, ci, cp, co,
                           U->p.p + uconv_offset(i), U->wf[i].p, i > 0 ? U->wb[i].p : nullptr);
    }
    if (!U->hw.p) U->hw.alloc(17);
    hipLaunchKernelGGL(k_head_pack, dim3(1), dim3(64), 0, s, U->p.p + uconv_offset(kUConvs), U->hw.p);
    SG_HIP(hipGetLastError());
    U->packed = true;
}

static void unet_shape(sg_unet* U, int B, int H, int W, bool grads) {
    if (B < 1 || H < 8 || W < 8 || H % 8 || W % 8) fail(SG_E_ARG, "network input needs H, W multiples of 8");
    const size_t hw0 = (size_t)H * W, hw1 = hw0 / 4, hw2 = hw0 / 16, hw3 = hw0 / 64;
    if (U->B != B || U->H != H || U->W != W) {
        U->B = B, U->H = H, U->W = W;
        U->grads_alloc = false;
        auto A = [&](DevBuf<uint16_t>& d, int c, size_t hw) { d.alloc((size_t)B * c * hw); };
        A(U->x, 16, hw0), A(U->a0, 16, hw0), A(U->cat0, 48, hw0), A(U->p0, 16, hw1);
        A(U->a2, 32, hw1), A(U->cat1, 96, hw1), A(U->p1, 32, hw2), A(U->a4, 64, hw2);
        A(U->cat2, 128, hw2), A(U->p2, 64, hw3), A(U->a6, 64, hw3), A(U->a7, 64, hw3);
        A(U->a8, 64, hw2), A(U->a9, 64, hw2), A(U->a10, 32, hw1), A(U->a11, 32, hw1);
        A(U->a12, 16, hw0), A(U->a13, 16, hw0);
        U->out.alloc((size_t)B * hw0);
        U->lpart.alloc((size_t)B * kLB * 3);
        U->lsums.alloc((size_t)B * 3);
        U->lloss.alloc(1);
    }
    if (grads && !U->grads_alloc) {
        auto A = [&](DevBuf<uint16_t>& d, int c, size_t hw) { d.alloc((size_t)B * c * hw); };
        A(U->ga0, 16, hw0), A(U->gcat0, 48, hw0), A(U->gp0, 16, hw1);
        A(U->ga2, 32, hw1), A(U->gcat1, 96, hw1), A(U->gp1, 32, hw2), A(U->ga4, 64, hw2);
        A(U->gcat2, 128, hw2), A(U->gp2, 64, hw3), A(U->ga6, 64, hw3), A(U->ga7, 64, hw3);
        A(U->ga8, 64, hw2), A(U->ga9, 64, hw2), A(U->ga10, 32, hw1), A(U->ga11, 32, hw1);
        A(U->ga12, 16, hw0), A(U->ga13, 16, hw0);
        // weight-gradient partials: the largest tiles x 9 cout cin over all convs
        size_t mx = 0, mxb = 0;
        int h = H, w = W;
        const int lvl[kUConvs] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0};
        for (int i = 0; i < kUConvs; ++i) {
            h = H >> lvl[i], w = W >> lvl[i];
            const size_t tiles = (size_t)B * ((h + kWY - 1) / kWY) * ((w + kWX - 1) / kWX);
            mx = std::max(mx, tiles * 9 * kUSpec[i].cout * kUSpec[i].cin);
            mxb = std::max(mxb, tiles * kUSpec[i].cout);
        }
        U->dwpart.alloc(mx);
        U->dwpart_b.alloc(mxb);
        U->hpart.alloc((size_t)1024 * 17);
        U->dout.alloc((size_t)B * hw0);
        U->lu.alloc((size_t)B * hw0);
        U->lg.alloc((size_t)B * hw0);
        U->grads_alloc = true;
    }
}

static View V(DevBuf<uint16_t>& d, int ctot, int c0 = 0) { return View{d.p, ctot, c0}; }

// Output channels per workgroup: 64, halved (down to 16) while the grid has
// fewer than 2048 workgroups, so the small deep layers still fill the CUs;
// at most cob_max(cc).
static int conv_cob(int tiles, int cout, int cc) {
    int cob = cout >= 64 ? 64 : (cout >= 32 ? 32 : 16);
    while (cob > 16 && ((long long)tiles * ((cout + cob - 1) / cob) < 2048 || cob > cob_max(cc)))
        cob /= 2;
    return cob;
}

// Chunk width for a padded input of cinp channels: all of it when it is one
// of the instantiated widths (every layer of this network), else 16.
static int conv_cc(int cinp) {
    for (int cc : {128, 96, 64, 48, 32, 16})
        if (cinp % cc == 0 && cinp / cc == 1) return cc;
    return 16;
}

template <int MODE, int EPI, int CC, int RW>
static void conv_go(dim3 grid, int cob, hipStream_t st, View in, View up, int cup, int ci, int cinp,
                    const uint16_t* wt, const float* bias, View out, int co, int H, int W, View pout,
                    const float* hw, float* hout, const int* stop) {
    hipLaunchKernelGGL((k_conv3x3<MODE, EPI, CC, RW>), grid, dim3(256), conv_lds(CC, cob, RW).bytes, st,
                       in, up, cup, ci, cinp, wt, bias, out, co, H, W, cob, pout, hw, hout, stop);
}

template <int MODE, int EPI, int RW>
static void conv_dispatch_rw(int cc, dim3 grid, int cob, hipStream_t st, View in, View up, int cup,
                             int ci, int cinp, const uint16_t* wt, const float* bias, View out, int co,
                             int H, int W, View pout, const float* hw, float* hout, const int* stop) {
#define SG_CONV_CC(C)                                                                                 \
    case C:                                                                                           \
        conv_go<MODE, EPI, C, RW>(grid, cob, st, in, up, cup, ci, cinp, wt, bias, out, co, H, W, pout, \
                                  hw, hout, stop);                                                    \
        break;
    switch (cc) {
        SG_CONV_CC(16)
        SG_CONV_CC(32)
        SG_CONV_CC(48)
        SG_CONV_CC(64)
        SG_CONV_CC(96)
        SG_CONV_CC(128)
        default: fail(SG_E_ARG, "convolution: unsupported chunk width");
    }
#undef SG_CONV_CC
}

// LDS beyond the default 64 KB for the wide chunks (once per process and device
// is enough; repeating it is harmless)
template <int MODE, int EPI, int RW>
static void conv_allow_lds_rw() {
#define SG_CONV_LDS(C)                                                                             \
    SG_HIP(hipFuncSetAttribute((const void*)k_conv3x3<MODE, EPI, C, RW>,                          \
                               hipFuncAttributeMaxDynamicSharedMemorySize,                        \
                               conv_lds(C, cob_max(C), RW).bytes));
    SG_CONV_LDS(16)
    SG_CONV_LDS(32)
    SG_CONV_LDS(48)
    SG_CONV_LDS(64)
    SG_CONV_LDS(96)
    SG_CONV_LDS(128)
#undef SG_CONV_LDS
}

template <int MODE, int EPI>
static void conv_allow_lds() {
    conv_allow_lds_rw<MODE, EPI, 2>();
    conv_allow_lds_rw<MODE, EPI, 1>();
}

// Launch shape of a convolution: rows per wavefront (2, or 1 when the layer
// has fewer than 512 workgroups of 8-row tiles, so small layers spread over
// twice the workgroups), tiles and output channels per workgroup.
struct ConvShape {
    int rw, tiles, cob;
};
static ConvShape conv_shape(int B, int H, int W, int cout, int cc) {
    auto tiles_of = [&](int rw) { return B * ((H + 4 * rw - 1) / (4 * rw)) * ((W + kTX - 1) / kTX); };
    int t = tiles_of(2), cob = conv_cob(t, cout, cc);
    if ((long long)t * ((cout + cob - 1) / cob) >= 512) return {2, t, cob};
    t = tiles_of(1);
    return {1, t, conv_cob(t, cout, cc)};
}

template <int MODE, int EPI>
static void conv_dispatch(const ConvShape& sh, int cc, hipStream_t st, View in, View up, int cup,
                          int ci, int cinp, const uint16_t* wt, const float* bias, View out, int co,
                          int H, int W, View pout, const float* hw, float* hout, const int* stop) {
    const dim3 grid(sh.tiles, (co + sh.cob - 1) / sh.cob);
    if (sh.rw == 1) {
        conv_dispatch_rw<MODE, EPI, 1>(cc, grid, sh.cob, st, in, up, cup, ci, cinp, wt, bias, out, co,
                                       H, W, pout, hw, hout, stop);
        return;
    }
    conv_dispatch_rw<MODE, EPI, 2>(cc, grid, sh.cob, st, in, up, cup, ci, cinp, wt, bias, out, co, H,
                                   W, pout, hw, hout, stop);
}

// Forward convolution i; `up`/`cup`: the input's first cup channels read from
// the half-resolution view up; EPI 1 with pooled output pout, EPI 2 with the head.
template <int EPI = 0>
static void conv_fwd(sg_unet* U, int i, View in, View out, int H, int W, const int* stop,
                     View up = View{nullptr, 0, 0}, int cup = 0, View pout = View{nullptr, 0, 0}) {
    const int ci = kUSpec[i].cin, co = kUSpec[i].cout, cp = cin_pad(ci), cc = conv_cc(cp);
    const ConvShape sh = conv_shape(U->B, H, W, co, cc);
    conv_dispatch<0, EPI>(sh, cc, U->ctx->stream, in, up, cup, ci, cp, U->wf[i].p,
                          U->p.p + uconv_offset(i) + co * ci * 9, out, co, H, W, pout, U->hw.p,
                          U->out.p, stop);
}

static void conv_dx(sg_unet* U, int i, View dp, View dx, int H, int W) {
    const int ci = kUSpec[i].cin, co = kUSpec[i].cout, cc = conv_cc(co);  // dx has ci channels
    const ConvShape sh = conv_shape(U->B, H, W, ci, cc);
    const View none{nullptr, 0, 0};
    conv_dispatch<1, 0>(sh, cc, U->ctx->stream, dp, none, 0, co, co, U->wb[i].p, nullptr, dx, ci, H, W,
                        none, nullptr, nullptr, nullptr);
}

static void conv_dw(sg_unet* U, int i, View in, View dp, int H, int W) {
    const int ci = kUSpec[i].cin, co = kUSpec[i].cout;
    const int tiles = U->B * ((H + kWY - 1) / kWY) * ((W + kWX - 1) / kWX);
    hipLaunchKernelGGL(k_conv_dw, dim3(tiles, co / 16, (ci + 15) / 16), dim3(256), 0, U->ctx->stream,
                       in, ci, dp, co, U->B, H, W, U->dwpart.p, U->dwpart_b.p);
    hipLaunchKernelGGL(k_dw_reduce, dim3(grid1((size_t)9 * co * ci + co)), dim3(256), 0,
                       U->ctx->stream, tiles, ci, co, U->dwpart.p, U->dwpart_b.p,
                       U->g.p + uconv_offset(i));
}

// d_in == nullptr: the caller has written the bf16 input (unet_input_buffer).
// Pooling and the head run in the epilogues of the convolutions that produce
// their inputs; without gradients the decoders read their upsampled channels
// straight from the half-resolution tensors (training keeps the concatenation
// buffers, which the backward pass reads).
void unet_forward_dev(sg_unet* U, int B, int H, int W, const float* d_in, bool grads,
                      const int* stop) {
    unet_pack(U);
    unet_shape(U, B, H, W, grads);
    hipStream_t s = U->ctx->stream;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    const size_t hw0 = (size_t)H * W;
    if (d_in)
        hipLaunchKernelGGL(k_input, dim3(grid1(B * hw0)), dim3(256), 0, s, B, H, W, d_in, V(U->x, 16),
                           stop);
    conv_fwd(U, 0, V(U->x, 16), V(U->a0, 16), H, W, stop);
    conv_fwd<1>(U, 1, V(U->a0, 16), V(U->cat0, 48, 32), H, W, stop, {}, 0, V(U->p0, 16));  // s0
    conv_fwd(U, 2, V(U->p0, 16), V(U->a2, 32), H1, W1, stop);
    conv_fwd<1>(U, 3, V(U->a2, 32), V(U->cat1, 96, 64), H1, W1, stop, {}, 0, V(U->p1, 32));  // s1
    conv_fwd(U, 4, V(U->p1, 32), V(U->a4, 64), H2, W2, stop);
    conv_fwd<1>(U, 5, V(U->a4, 64), V(U->cat2, 128, 64), H2, W2, stop, {}, 0, V(U->p2, 64));  // s2
    conv_fwd(U, 6, V(U->p2, 64), V(U->a6, 64), H3, W3, stop);
    conv_fwd(U, 7, V(U->a6, 64), V(U->a7, 64), H3, W3, stop);
    if (grads) {
        hipLaunchKernelGGL(k_up, dim3(grid1(B * 8 * hw0 / 16)), dim3(256), 0, s, B, 64, H2, W2,
                           V(U->a7, 64), V(U->cat2, 128, 0), stop);
        conv_fwd(U, 8, V(U->cat2, 128), V(U->a8, 64), H2, W2, stop);
    } else {
        conv_fwd(U, 8, V(U->cat2, 128), V(U->a8, 64), H2, W2, stop, V(U->a7, 64), 64);
    }
    conv_fwd(U, 9, V(U->a8, 64), V(U->a9, 64), H2, W2, stop);
    if (grads) {
        hipLaunchKernelGGL(k_up, dim3(grid1(B * 8 * hw0 / 4)), dim3(256), 0, s, B, 64, H1, W1,
                           V(U->a9, 64), V(U->cat1, 96, 0), stop);
        conv_fwd(U, 10, V(U->cat1, 96), V(U->a10, 32), H1, W1, stop);
    } else {
        conv_fwd(U, 10, V(U->cat1, 96), V(U->a10, 32), H1, W1, stop, V(U->a9, 64), 64);
    }
    conv_fwd(U, 11, V(U->a10, 32), V(U->a11, 32), H1, W1, stop);
    if (grads) {
        hipLaunchKernelGGL(k_up, dim3(grid1(B * 4 * hw0)), dim3(256), 0, s, B, 32, H, W,
                           V(U->a11, 32), V(U->cat0, 48, 0), stop);
        conv_fwd(U, 12, V(U->cat0, 48), V(U->a12, 16), H, W, stop);
    } else {
        conv_fwd(U, 12, V(U->cat0, 48), V(U->a12, 16), H, W, stop, V(U->a11, 32), 32);
    }
    conv_fwd<2>(U, 13, V(U->a12, 16), V(U->a13, 16), H, W, stop);  // + head -> out
    SG_HIP(hipGetLastError());
}

// The bf16 NHWC input tensor [B][H][W][16] of the next forward at this shape
// (channels 0, 1 the network's two inputs, the rest zero).
uint16_t* unet_input_buffer(sg_unet* U, int B, int H, int W) {
    unet_pack(U);
    unet_shape(U, B, H, W, false);
    return U->x.p;
}

void unet_backward_dev(sg_unet* U, const float* d_in, const float* d_target) {
    const int B = U->B, H = U->H, W = U->W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    const size_t hw0 = (size_t)H * W;
    hipStream_t s = U->ctx->stream;
    const int ge = grid1(B * hw0);
    if (U->loss_kind == 0) {
        hipLaunchKernelGGL(k_loss_part, dim3(kLB, B), dim3(256), 0, s, H, W, U->out.p, d_target,
                           U->lpart.p);
        hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(64), 0, s, B, U->lpart.p, U->lsums.p, U->lloss.p);
        hipLaunchKernelGGL(k_mse_dout, dim3(ge), dim3(256), 0, s, B, H, W, U->out.p, d_target,
                           U->lsums.p, U->dout.p);
    } else {
        hipLaunchKernelGGL(k_stencil, dim3(ge), dim3(256), 0, s, B, H, W, d_in, U->out.p, U->lu.p);
        hipLaunchKernelGGL(k_rloss_part, dim3(kLB, B), dim3(256), 0, s, H, W, d_in, U->lu.p, U->lpart.p);
        hipLaunchKernelGGL(k_rloss_fin, dim3(1), dim3(64), 0, s, B, U->loss_kind, U->lpart.p, U->lsums.p,
                           U->lloss.p);
        hipLaunchKernelGGL(k_rloss_g, dim3(ge), dim3(256), 0, s, B, H, W, d_in, U->lu.p, U->lsums.p,
                           U->lg.p);
        hipLaunchKernelGGL(k_stencil, dim3(ge), dim3(256), 0, s, B, H, W, d_in, U->lg.p, U->dout.p);
    }
    const int hb = std::min(1024, grid1(B * hw0));
    hipLaunchKernelGGL(k_head_bwd, dim3(hb), dim3(256), 0, s, B, H, W, U->dout.p, V(U->a13, 16),
                       U->hw.p, V(U->ga13, 16), U->hpart.p);
    hipLaunchKernelGGL(k_head_reduce, dim3(1), dim3(64), 0, s, hb, U->hpart.p, U->g.p + uconv_offset(kUConvs));
    auto mask = [&](DevBuf<uint16_t>& y, int yc, int yc0, DevBuf<uint16_t>& dy, int dc, int dc0, int C,
                    int h, int w) {
        hipLaunchKernelGGL(k_relu_mask, dim3(grid1((size_t)B * C * h * w)), dim3(256), 0, s, B, C, h, w,
                           V(y, yc, yc0), V(dy, dc, dc0));
    };
    // conv i: dy (gradient w.r.t. its ReLU output) is masked in place, then the
    // weight gradient from its input, then the data gradient into dx
    // dec0
    mask(U->a13, 16, 0, U->ga13, 16, 0, 16, H, W);
    conv_dw(U, 13, V(U->a12, 16), V(U->ga13, 16), H, W);
    conv_dx(U, 13, V(U->ga13, 16), V(U->ga12, 16), H, W);
    mask(U->a12, 16, 0, U->ga12, 16, 0, 16, H, W);
    conv_dw(U, 12, V(U->cat0, 48), V(U->ga12, 16), H, W);
    conv_dx(U, 12, V(U->ga12, 16), V(U->gcat0, 48), H, W);
    hipLaunchKernelGGL(k_up_bwd, dim3(grid1(B * 32 * hw0 / 4)), dim3(256), 0, s, B, 32, H, W,
                       V(U->gcat0, 48, 0), V(U->ga11, 32));
    // dec1
    mask(U->a11, 32, 0, U->ga11, 32, 0, 32, H1, W1);
    conv_dw(U, 11, V(U->a10, 32), V(U->ga11, 32), H1, W1);
    conv_dx(U, 11, V(U->ga11, 32), V(U->ga10, 32), H1, W1);
    mask(U->a10, 32, 0, U->ga10, 32, 0, 32, H1, W1);
    conv_dw(U, 10, V(U->cat1, 96), V(U->ga10, 32), H1, W1);
    conv_dx(U, 10, V(U->ga10, 32), V(U->gcat1, 96), H1, W1);
    hipLaunchKernelGGL(k_up_bwd, dim3(grid1(B * 64 * hw0 / 16)), dim3(256), 0, s, B, 64, H1, W1,
                       V(U->gcat1, 96, 0), V(U->ga9, 64));
    // dec2
    mask(U->a9, 64, 0, U->ga9, 64, 0, 64, H2, W2);
    conv_dw(U, 9, V(U->a8, 64), V(U->ga9, 64), H2, W2);
    conv_dx(U, 9, V(U->ga9, 64), V(U->ga8, 64), H2, W2);
    mask(U->a8, 64, 0, U->ga8, 64, 0, 64, H2, W2);
    conv_dw(U, 8, V(U->cat2, 128), V(U->ga8, 64), H2, W2);
    conv_dx(U, 8, V(U->ga8, 64), V(U->gcat2, 128), H2, W2);
    hipLaunchKernelGGL(k_up_bwd, dim3(grid1(B * 64 * hw0 / 64)), dim3(256), 0, s, B, 64, H2, W2,
                       V(U->gcat2, 128, 0), V(U->ga7, 64));
    // bottleneck
    mask(U->a7, 64, 0, U->ga7, 64, 0, 64, H3, W3);
    conv_dw(U, 7, V(U->a6, 64), V(U->ga7, 64), H3, W3);
    conv_dx(U, 7, V(U->ga7, 64), V(U->ga6, 64), H3, W3);
    mask(U->a6, 64, 0, U->ga6, 64, 0, 64, H3, W3);
    conv_dw(U, 6, V(U->p2, 64), V(U->ga6, 64), H3, W3);
    conv_dx(U, 6, V(U->ga6, 64), V(U->gp2, 64), H3, W3);
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(B * 64 * hw0 / 16)), dim3(256), 0, s, B, 64, H2, W2,
                       V(U->gp2, 64), V(U->gcat2, 128, 64));
    // enc2 (s2 = cat2[64..128), its gradient gcat2[64..128))
    mask(U->cat2, 128, 64, U->gcat2, 128, 64, 64, H2, W2);
    conv_dw(U, 5, V(U->a4, 64), V(U->gcat2, 128, 64), H2, W2);
    conv_dx(U, 5, V(U->gcat2, 128, 64), V(U->ga4, 64), H2, W2);
    mask(U->a4, 64, 0, U->ga4, 64, 0, 64, H2, W2);
    conv_dw(U, 4, V(U->p1, 32), V(U->ga4, 64), H2, W2);
    conv_dx(U, 4, V(U->ga4, 64), V(U->gp1, 32), H2, W2);
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(B * 32 * hw0 / 4)), dim3(256), 0, s, B, 32, H1, W1,
                       V(U->gp1, 32), V(U->gcat1, 96, 64));
    // enc1
    mask(U->cat1, 96, 64, U->gcat1, 96, 64, 32, H1, W1);
    conv_dw(U, 3, V(U->a2, 32), V(U->gcat1, 96, 64), H1, W1);
    conv_dx(U, 3, V(U->gcat1, 96, 64), V(U->ga2, 32), H1, W1);
    mask(U->a2, 32, 0, U->ga2, 32, 0, 32, H1, W1);
    conv_dw(U, 2, V(U->p0, 16), V(U->ga2, 32), H1, W1);
    conv_dx(U, 2, V(U->ga2, 32), V(U->gp0, 16), H1, W1);
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid1(B * 16 * hw0)), dim3(256), 0, s, B, 16, H, W,
                       V(U->gp0, 16), V(U->gcat0, 48, 32));
    // enc0
    mask(U->cat0, 48, 32, U->gcat0, 48, 32, 16, H, W);
    conv_dw(U, 1, V(U->a0, 16), V(U->gcat0, 48, 32), H, W);
    conv_dx(U, 1, V(U->gcat0, 48, 32), V(U->ga0, 16), H, W);
    mask(U->a0, 16, 0, U->ga0, 16, 0, 16, H, W);
    conv_dw(U, 0, V(U->x, 16), V(U->ga0, 16), H, W);
    SG_HIP(hipGetLastError());
}

void unet_prepare(sg_unet* U, int B, int H, int W) {
    unet_pack(U);
    unet_shape(U, B, H, W, false);
}

// the network output of the last forward pass (device, [B][H][W] fp32)
const float* unet_output(const sg_unet* U) { return U->out.p; }
sg_context* unet_ctx(const sg_unet* U) { return U->ctx; }

}  // namespace sg

extern "C" {

int sg_unet_nparams(void) { return unet_nparams_(); }

sg_status sg_unet_create(sg_context* ctx, sg_unet** out) {
    if (!ctx || !out) return SG_E_ARG;
    *out = nullptr;
    sg_unet* U = new sg_unet();
    sg_status st = run_guarded(ctx, [&] {
        SG_HIP(hipSetDevice(ctx->device));
        U->ctx = ctx;
        U->np = unet_nparams_();
        U->p.alloc(U->np);
        U->g.alloc(U->np);
        U->m.alloc(U->np);
        U->v.alloc(U->np);
        SG_HIP(hipMemsetAsync(U->p.p, 0, sizeof(float) * U->np, ctx->stream));
        SG_HIP(hipMemsetAsync(U->g.p, 0, sizeof(float) * U->np, ctx->stream));
        SG_HIP(hipMemsetAsync(U->m.p, 0, sizeof(float) * U->np, ctx->stream));
        SG_HIP(hipMemsetAsync(U->v.p, 0, sizeof(float) * U->np, ctx->stream));
        SG_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (st != SG_OK) {
        delete U;
        return st;
    }
    *out = U;
    return SG_OK;
}

sg_status sg_unet_destroy(sg_unet* U) {
    if (!U) return SG_E_ARG;
    (void)hipStreamSynchronize(U->ctx->stream);
    delete U;
    return SG_OK;
}

sg_status sg_unet_set_params(sg_unet* U, const float* params) {
    if (!U || !params) return SG_E_ARG;
    return run_guarded(U->ctx, [&] {
        SG_HIP(hipMemcpyAsync(U->p.p, params, sizeof(float) * U->np, hipMemcpyHostToDevice, U->ctx->stream));
        SG_HIP(hipMemsetAsync(U->m.p, 0, sizeof(float) * U->np, U->ctx->stream));
        SG_HIP(hipMemsetAsync(U->v.p, 0, sizeof(float) * U->np, U->ctx->stream));
        U->adam_t = 0;
        U->packed = false;
        SG_HIP(hipStreamSynchronize(U->ctx->stream));
    });
}

sg_status sg_unet_get_params(const sg_unet* U, float* params) {
    if (!U || !params) return SG_E_ARG;
    return run_guarded(U->ctx, [&] {
        SG_HIP(hipMemcpyAsync(params, U->p.p, sizeof(float) * U->np, hipMemcpyDeviceToHost, U->ctx->stream));
        SG_HIP(hipStreamSynchronize(U->ctx->stream));
    });
}

sg_status sg_unet_forward(sg_unet* U, int B, int H, int W, const float* d_in, float* d_out) {
    if (!U || !d_in || !d_out) return SG_E_ARG;
    return run_guarded(U->ctx, [&] {
        unet_forward_dev(U, B, H, W, d_in, false, nullptr);
        SG_HIP(hipMemcpyAsync(d_out, U->out.p, sizeof(float) * B * H * W, hipMemcpyDeviceToDevice,
                              U->ctx->stream));
    });
}

sg_status sg_unet_loss_grad(sg_unet* U, int B, int H, int W, const float* d_in, const float* d_target,
                            double* loss, float* grad) {
    if (!U || !d_in) return SG_E_ARG;
    return run_guarded(U->ctx, [&] {
        if (U->loss_kind == 0 && !d_target) fail(SG_E_ARG, "this loss needs a target");
        unet_forward_dev(U, B, H, W, d_in, true, nullptr);
        unet_backward_dev(U, d_in, d_target);
        if (loss)
            SG_HIP(hipMemcpyAsync(loss, U->lloss.p, sizeof(double), hipMemcpyDeviceToHost, U->ctx->stream));
        if (grad)
            SG_HIP(hipMemcpyAsync(grad, U->g.p, sizeof(float) * U->np, hipMemcpyDeviceToHost, U->ctx->stream));
        SG_HIP(hipStreamSynchronize(U->ctx->stream));
    });
}

sg_status sg_unet_set_weight_decay(sg_unet* U, float wd) {
    if (!U || !(wd >= 0.f)) return SG_E_ARG;
    U->wd = wd;
    return SG_OK;
}

sg_status sg_unet_set_loss(sg_unet* U, int kind) {
    if (!U || kind < 0 || kind > 2) return SG_E_ARG;
    U->loss_kind = kind;
    return SG_OK;
}

sg_status sg_unet_adam_step(sg_unet* U, float lr) {
    if (!U || !(lr > 0.f)) return SG_E_ARG;
    return run_guarded(U->ctx, [&] {
        U->adam_t += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)U->adam_t), c2 = 1.f - std::pow(b2, (float)U->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(grid1(U->np)), dim3(256), 0, U->ctx->stream, U->np, U->p.p, U->g.p,
                           U->m.p, U->v.p, lr, b1, b2, 1e-8f, c1, c2, U->wd);
        SG_HIP(hipGetLastError());
        U->packed = false;
    });
}

}  // extern "C"
