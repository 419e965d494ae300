// CPU oracle of the learned residual-to-error network (test infrastructure).
//
// The network has no counterpart in the reference (it is an addition of this
// project's solver); this file fixes its definition, and the GPU kernels are
// checked against it:
//
//   input  [B][2][H][W]   channel 0 = residual / rms(residual), channel 1 = log kappa
//   enc0   conv3x3(2->16)+ReLU, conv3x3(16->16)+ReLU           -> s0 [16][H][W]
//   enc1   avgpool2, conv3x3(16->32)+ReLU, conv3x3(32->32)+ReLU -> s1 [32][H/2]
//   enc2   avgpool2, conv3x3(32->64)+ReLU, conv3x3(64->64)+ReLU -> s2 [64][H/4]
//   bott   avgpool2, conv3x3(64->64)+ReLU, conv3x3(64->64)+ReLU -> b  [64][H/8]
//   dec2   up2(b) ++ s2 (128 ch), conv3x3(128->64)+ReLU, conv3x3(64->64)+ReLU
//   dec1   up2 ++ s1 (96 ch),     conv3x3(96->32)+ReLU,  conv3x3(32->32)+ReLU
//   dec0   up2 ++ s0 (48 ch),     conv3x3(48->16)+ReLU,  conv3x3(16->16)+ReLU
//   head   conv1x1(16->1) + bias  -> output [B][H][W]
//
// 3x3 convolutions use zero padding; avgpool2 is the mean of each 2x2 block;
// up2 repeats each value into a 2x2 block; "++" concatenates channels with the
// upsampled tensor first.  Parameters are packed conv by conv in the order
// above as W[Cout][Cin][3][3] then b[Cout], then the head's w[16], b.
//
// Loss kind 0 (per batch): mean over samples of sum_p (out - t)^2 / sum_p t^2.
// Loss kind 1 (residual, no target): with s = input channel 0, kappa =
// exp(input channel 1) and A the 2D five-point finite-volume diffusion
// operator (harmonic-mean faces 2 k_p k_q / (k_p + k_q), a boundary face adds
// k_p to the diagonal), u = A out: mean over samples of 1 - (s.u)^2 / (|s|^2 |u|^2),
// the fraction of |s|^2 that the best multiple of out leaves.
// Loss kind 2: kind 1 plus mu ln(|u|^2 / |s|^2)^2 per sample (mean), mu =
// kScaleMu: kind 1 does not see the scale of out, this term holds |A out| near
// |s| so the scale cannot drift during training.
//
// With round_bf16 = 1 the oracle rounds to bfloat16 (round to nearest even)
// exactly where the GPU stores bfloat16: every weight used by a convolution,
// the network input, every activation (after bias + ReLU), the pooled
// tensors and every activation gradient; sums are accumulated in double here
// and in fp32 on the GPU.  round_bf16 = 0 is plain double arithmetic, which
// the finite-difference checks use.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct ConvSpec {
    int cin, cout;
};

constexpr int kNumConv = 14;
constexpr double kScaleMu = 0.01;  // weight of the scale term of loss kind 2
const ConvSpec kConvs[kNumConv] = {{2, 16},  {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                   {64, 64}, {64, 64}, {64, 64}, {128, 64}, {64, 64},
                                   {96, 32}, {32, 32}, {48, 16}, {16, 16}};

int conv_offset(int i) {
    int o = 0;
    for (int k = 0; k < i; ++k) o += kConvs[k].cout * kConvs[k].cin * 9 + kConvs[k].cout;
    return o;
}

int nparams() { return conv_offset(kNumConv) + 16 + 1; }

double bf16(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;  // inf / nan unchanged
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct T4 {  // [B][C][H][W] in double
    int B = 0, C = 0, H = 0, W = 0;
    std::vector<double> v;
    void init(int b, int c, int h, int w) {
        B = b, C = c, H = h, W = w;
        v.assign((size_t)b * c * h * w, 0.0);
    }
    double& at(int b, int c, int y, int x) { return v[(((size_t)b * C + c) * H + y) * W + x]; }
    double at(int b, int c, int y, int x) const { return v[(((size_t)b * C + c) * H + y) * W + x]; }
};

void round_all(T4& t, bool r) {
    if (r)
        for (double& e : t.v) e = bf16(e);
}

// y = relu(conv3x3(x; w, bias))
void conv_fwd(const T4& x, const double* w, const double* bias, int cout, T4& y, bool rnd) {
    y.init(x.B, cout, x.H, x.W);
    const int H = x.H, W = x.W;
    for (int b = 0; b < x.B; ++b)
        for (int co = 0; co < cout; ++co)
            for (int yy = 0; yy < H; ++yy)
                for (int xx = 0; xx < W; ++xx) {
                    double s = bias[co];
                    for (int ci = 0; ci < x.C; ++ci)
                        for (int ky = 0; ky < 3; ++ky) {
                            const int sy = yy + ky - 1;
                            if (sy < 0 || sy >= H) continue;
                            for (int kx = 0; kx < 3; ++kx) {
                                const int sx = xx + kx - 1;
                                if (sx < 0 || sx >= W) continue;
                                s += w[((co * x.C + ci) * 3 + ky) * 3 + kx] * x.at(b, ci, sy, sx);
                            }
                        }
                    y.at(b, co, yy, xx) = s > 0.0 ? s : 0.0;
                }
    round_all(y, rnd);
}

// Given dy (gradient w.r.t. the ReLU output y), accumulate dw, db and return
// dx (gradient w.r.t. the conv input x).
void conv_bwd(const T4& x, const double* w, const T4& y, const T4& dy, double* dw, double* db,
              T4& dx, bool rnd) {
    const int H = x.H, W = x.W, cout = y.C;
    T4 dp;  // gradient w.r.t. the pre-activation
    dp.init(y.B, cout, H, W);
    for (size_t i = 0; i < dp.v.size(); ++i) dp.v[i] = y.v[i] > 0.0 ? dy.v[i] : 0.0;
    dx.init(x.B, x.C, H, W);
    for (int b = 0; b < x.B; ++b)
        for (int co = 0; co < cout; ++co)
            for (int yy = 0; yy < H; ++yy)
                for (int xx = 0; xx < W; ++xx) {
                    const double g = dp.at(b, co, yy, xx);
                    if (g == 0.0) continue;
                    db[co] += g;
                    for (int ci = 0; ci < x.C; ++ci)
                        for (int ky = 0; ky < 3; ++ky) {
                            const int sy = yy + ky - 1;
                            if (sy < 0 || sy >= H) continue;
                            for (int kx = 0; kx < 3; ++kx) {
                                const int sx = xx + kx - 1;
                                if (sx < 0 || sx >= W) continue;
                                const int wi = ((co * x.C + ci) * 3 + ky) * 3 + kx;
                                dw[wi] += g * x.at(b, ci, sy, sx);
                                dx.at(b, ci, sy, sx) += g * w[wi];
                            }
                        }
                }
    round_all(dx, rnd);
}

void pool_fwd(const T4& x, T4& y, bool rnd) {
    y.init(x.B, x.C, x.H / 2, x.W / 2);
    for (int b = 0; b < x.B; ++b)
        for (int c = 0; c < x.C; ++c)
            for (int yy = 0; yy < y.H; ++yy)
                for (int xx = 0; xx < y.W; ++xx)
                    y.at(b, c, yy, xx) = 0.25 * (x.at(b, c, 2 * yy, 2 * xx) + x.at(b, c, 2 * yy, 2 * xx + 1) +
                                                 x.at(b, c, 2 * yy + 1, 2 * xx) +
                                                 x.at(b, c, 2 * yy + 1, 2 * xx + 1));
    round_all(y, rnd);
}

// dx += pool backward of dy (dx already holds the skip-path gradient).
void pool_bwd(const T4& dy, T4& dx, bool rnd) {
    for (int b = 0; b < dx.B; ++b)
        for (int c = 0; c < dx.C; ++c)
            for (int yy = 0; yy < dx.H; ++yy)
                for (int xx = 0; xx < dx.W; ++xx) dx.at(b, c, yy, xx) += 0.25 * dy.at(b, c, yy / 2, xx / 2);
    round_all(dx, rnd);
}

// y = up2(a) ++ s
void up_cat(const T4& a, const T4& s, T4& y) {
    y.init(s.B, a.C + s.C, s.H, s.W);
    for (int b = 0; b < s.B; ++b) {
        for (int c = 0; c < a.C; ++c)
            for (int yy = 0; yy < s.H; ++yy)
                for (int xx = 0; xx < s.W; ++xx) y.at(b, c, yy, xx) = a.at(b, c, yy / 2, xx / 2);
        for (int c = 0; c < s.C; ++c)
            for (int yy = 0; yy < s.H; ++yy)
                for (int xx = 0; xx < s.W; ++xx) y.at(b, a.C + c, yy, xx) = s.at(b, c, yy, xx);
    }
}

// split the gradient of up2(a) ++ s into da (summed over each 2x2 block) and ds
void up_cat_bwd(const T4& dy, int ca, T4& da, T4& ds, bool rnd) {
    const int B = dy.B, H = dy.H, W = dy.W, cs = dy.C - ca;
    da.init(B, ca, H / 2, W / 2);
    ds.init(B, cs, H, W);
    for (int b = 0; b < B; ++b) {
        for (int c = 0; c < ca; ++c)
            for (int yy = 0; yy < H; ++yy)
                for (int xx = 0; xx < W; ++xx) da.at(b, c, yy / 2, xx / 2) += dy.at(b, c, yy, xx);
        for (int c = 0; c < cs; ++c)
            for (int yy = 0; yy < H; ++yy)
                for (int xx = 0; xx < W; ++xx) ds.at(b, c, yy, xx) = dy.at(b, ca + c, yy, xx);
    }
    round_all(da, rnd);
}

struct Net {
    std::vector<double> p;  // parameters (bf16-rounded weights when rounding)
    bool rnd;
    const double* w(int i) const { return p.data() + conv_offset(i); }
    const double* b(int i) const { return w(i) + kConvs[i].cout * kConvs[i].cin * 9; }
    const double* hw() const { return p.data() + conv_offset(kNumConv); }
    // activations
    T4 x, a[kNumConv], pool[3], cat[3];
    std::vector<double> out;
};

void load(Net& n, const float* params, bool rnd) {
    const int np = nparams();
    n.p.assign(params, params + np);
    n.rnd = rnd;
    if (!rnd) return;
    for (int i = 0; i < kNumConv; ++i) {  // weights (not biases) in bf16
        double* wi = n.p.data() + conv_offset(i);
        for (int k = 0; k < kConvs[i].cout * kConvs[i].cin * 9; ++k) wi[k] = bf16(wi[k]);
    }
    double* h = n.p.data() + conv_offset(kNumConv);
    for (int k = 0; k < 16; ++k) h[k] = bf16(h[k]);
}

void forward(Net& n, int B, int H, int W, const float* in) {
    n.x.init(B, 2, H, W);
    for (size_t i = 0; i < n.x.v.size(); ++i) n.x.v[i] = in[i];
    round_all(n.x, n.rnd);
    const bool r = n.rnd;
    auto conv = [&](int i, const T4& src) { conv_fwd(src, n.w(i), n.b(i), kConvs[i].cout, n.a[i], r); };
    conv(0, n.x);
    conv(1, n.a[0]);  // s0
    pool_fwd(n.a[1], n.pool[0], r);
    conv(2, n.pool[0]);
    conv(3, n.a[2]);  // s1
    pool_fwd(n.a[3], n.pool[1], r);
    conv(4, n.pool[1]);
    conv(5, n.a[4]);  // s2
    pool_fwd(n.a[5], n.pool[2], r);
    conv(6, n.pool[2]);
    conv(7, n.a[6]);  // bottleneck
    up_cat(n.a[7], n.a[5], n.cat[2]);
    conv(8, n.cat[2]);
    conv(9, n.a[8]);
    up_cat(n.a[9], n.a[3], n.cat[1]);
    conv(10, n.cat[1]);
    conv(11, n.a[10]);
    up_cat(n.a[11], n.a[1], n.cat[0]);
    conv(12, n.cat[0]);
    conv(13, n.a[12]);
    const T4& d0 = n.a[13];
    n.out.assign((size_t)B * H * W, 0.0);
    const double* hw = n.hw();
    for (int b = 0; b < B; ++b)
        for (int yy = 0; yy < H; ++yy)
            for (int xx = 0; xx < W; ++xx) {
                double s = hw[16];
                for (int c = 0; c < 16; ++c) s += hw[c] * d0.at(b, c, yy, xx);
                n.out[((size_t)b * H + yy) * W + xx] = s;
            }
}

// u = A v for one sample (A symmetric), kappa from log kappa
void stencil(const float* logk, const double* v, double* u, int H, int W) {
    auto k = [&](int y, int x) { return std::exp((double)logk[y * W + x]); };
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const double kp = k(y, x), vp = v[y * W + x];
            double acc = 0.0;
            const int dy[4] = {-1, 0, 0, 1}, dx[4] = {0, -1, 1, 0};
            for (int f = 0; f < 4; ++f) {
                const int yy = y + dy[f], xx = x + dx[f];
                if (yy < 0 || yy >= H || xx < 0 || xx >= W) {
                    acc += kp * vp;
                } else {
                    const double kq = k(yy, xx);
                    acc += 2.0 * kp * kq / (kp + kq) * (vp - v[yy * W + xx]);
                }
            }
            u[y * W + x] = acc;
        }
}

}  // namespace

extern "C" {

int or_unet_nparams() { return nparams(); }

// Offsets of the 14 convolutions' W in the packed parameter vector (and the
// head at index 14); used by tests to index parameters.
int or_unet_param_offset(int i) { return conv_offset(i); }

void or_unet_forward(const float* params, int B, int H, int W, const float* in, float* out,
                     int round_bf16) {
    Net n;
    load(n, params, round_bf16 != 0);
    forward(n, B, H, W, in);
    for (size_t i = 0; i < n.out.size(); ++i) out[i] = (float)n.out[i];
}

double or_unet_loss_grad_kind(const float* params, int B, int H, int W, const float* in,
                              const float* target, int kind, int round_bf16, double* grad);

// Loss and its gradient with respect to the packed parameters.
double or_unet_loss_grad(const float* params, int B, int H, int W, const float* in,
                         const float* target, int round_bf16, double* grad) {
    return or_unet_loss_grad_kind(params, B, H, W, in, target, 0, round_bf16, grad);
}

double or_unet_loss_grad_kind(const float* params, int B, int H, int W, const float* in,
                              const float* target, int kind, int round_bf16, double* grad) {
    Net n;
    const bool r = round_bf16 != 0;
    load(n, params, r);
    forward(n, B, H, W, in);
    const size_t hw = (size_t)H * W;
    double loss = 0.0;
    std::vector<double> dout(n.out.size());
    if (kind == 0) {
        for (int b = 0; b < B; ++b) {
            double tt = 0.0, ee = 0.0;
            for (size_t i = 0; i < hw; ++i) {
                const double t = target[b * hw + i], d = n.out[b * hw + i] - t;
                tt += t * t;
                ee += d * d;
            }
            loss += ee / tt / B;
            for (size_t i = 0; i < hw; ++i)
                dout[b * hw + i] = 2.0 * (n.out[b * hw + i] - target[b * hw + i]) / tt / B;
        }
    } else {
        std::vector<double> u(hw), g(hw), v(hw);
        for (int b = 0; b < B; ++b) {
            const float* sb = in + (size_t)b * 2 * hw;
            const float* lk = sb + hw;
            for (size_t i = 0; i < hw; ++i) v[i] = n.out[b * hw + i];
            stencil(lk, v.data(), u.data(), H, W);
            double ss = 0.0, a = 0.0, c = 0.0;
            for (size_t i = 0; i < hw; ++i) {
                ss += (double)sb[i] * sb[i];
                a += (double)sb[i] * u[i];
                c += u[i] * u[i];
            }
            loss += (1.0 - a * a / (ss * c)) / B;
            const double f = -2.0 * a / (ss * c) / B, q = a / c;
            // kind 2 adds mu ln(c / ss)^2: it pins the output's scale (|A out| ~ |s|),
            // to which the kind-1 loss is blind; d/du = 4 mu ln(c / ss) u / c
            double h = 0.0;
            if (kind == 2) {
                const double l = std::log(c / ss);
                loss += kScaleMu * l * l / B;
                h = 4.0 * kScaleMu * l / c / B;
            }
            for (size_t i = 0; i < hw; ++i) g[i] = f * ((double)sb[i] - q * u[i]) + h * u[i];
            stencil(lk, g.data(), u.data(), H, W);  // dL/dout = A dL/du
            for (size_t i = 0; i < hw; ++i) dout[b * hw + i] = u[i];
        }
    }
    const int np = nparams();
    std::fill(grad, grad + np, 0.0);
    auto gw = [&](int i) { return grad + conv_offset(i); };
    auto gb = [&](int i) { return gw(i) + kConvs[i].cout * kConvs[i].cin * 9; };
    // head
    double* ghw = grad + conv_offset(kNumConv);
    const double* hwp = n.hw();
    T4 d;
    d.init(B, 16, H, W);
    for (int b = 0; b < B; ++b)
        for (size_t i = 0; i < hw; ++i) {
            const double g = dout[b * hw + i];
            ghw[16] += g;
            for (int c = 0; c < 16; ++c) {
                ghw[c] += g * n.a[13].v[(b * 16 + c) * hw + i];
                d.v[(b * 16 + c) * hw + i] = g * hwp[c];
            }
        }
    round_all(d, r);
    T4 dx, da, ds, dskip[3];
    auto cb = [&](int i, const T4& src, const T4& dy, T4& dsrc) {
        conv_bwd(src, n.w(i), n.a[i], dy, gw(i), gb(i), dsrc, r);
    };
    // dec0
    cb(13, n.a[12], d, dx);
    cb(12, n.cat[0], dx, d);
    up_cat_bwd(d, 32, da, dskip[0], r);
    // dec1
    cb(11, n.a[10], da, dx);
    cb(10, n.cat[1], dx, d);
    up_cat_bwd(d, 64, da, dskip[1], r);
    // dec2
    cb(9, n.a[8], da, dx);
    cb(8, n.cat[2], dx, d);
    up_cat_bwd(d, 64, da, dskip[2], r);
    // bottleneck
    cb(7, n.a[6], da, dx);
    cb(6, n.pool[2], dx, d);
    pool_bwd(d, dskip[2], r);  // s2 gets skip + pooled gradient
    // enc2
    cb(5, n.a[4], dskip[2], dx);
    cb(4, n.pool[1], dx, d);
    pool_bwd(d, dskip[1], r);
    // enc1
    cb(3, n.a[2], dskip[1], dx);
    cb(2, n.pool[0], dx, d);
    pool_bwd(d, dskip[0], r);
    // enc0
    cb(1, n.a[0], dskip[0], dx);
    cb(0, n.x, dx, d);
    return loss;
}

}  // extern "C"
