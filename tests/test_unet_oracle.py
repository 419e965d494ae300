"""The learned stage's CPU oracle (oracle/unet_oracle.cpp) pinned on the CPU:

* an independent float64 PyTorch-CPU restatement of the same network (conv2d,
  avg_pool2d, nearest upsampling, concatenation) gives the same output and,
  through autograd, the same parameter gradient;
* central finite differences of the oracle's own loss agree with its
  analytic gradient;
* the bfloat16 rounding mode stays close to the exact mode.
"""
import numpy as np
import pytest

import oracle_bind as O

CONVS = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 64),
         (128, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]


def init_params(seed):
    rng = np.random.default_rng(seed)
    p = []
    for cin, cout in CONVS:
        p.append(rng.standard_normal(cout * cin * 9) * np.sqrt(2.0 / (9 * cin)))
        p.append(rng.standard_normal(cout) * 0.05)
    p.append(rng.standard_normal(16) * 0.25)
    p.append(np.array([0.01]))
    return np.concatenate(p).astype(np.float32)


def sample(seed, B=2, H=8, W=8):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, 2, H, W)).astype(np.float32)
    t = rng.standard_normal((B, H, W)).astype(np.float32)
    return x, t


def test_parameter_count_and_layout():
    n = sum(ci * co * 9 + co for ci, co in CONVS) + 17
    assert O.unet_nparams() == n == 302705
    off = 0
    for i, (ci, co) in enumerate(CONVS):
        assert O.unet_param_offset(i) == off
        off += ci * co * 9 + co
    assert O.unet_param_offset(14) == off


def torch_net(params):
    import torch
    import torch.nn.functional as F
    p = torch.tensor(params.astype(np.float64), requires_grad=True)
    ws, off = [], 0
    for ci, co in CONVS:
        w = p[off:off + co * ci * 9].view(co, ci, 3, 3)
        off += co * ci * 9
        b = p[off:off + co]
        off += co
        ws.append((w, b))
    hw, hb = p[off:off + 16], p[off + 16]

    def f(x):
        def conv(i, t):
            return F.relu(F.conv2d(t, ws[i][0], ws[i][1], padding=1))

        def up_cat(a, s):
            return torch.cat([F.interpolate(a, scale_factor=2, mode="nearest"), s], 1)

        s0 = conv(1, conv(0, x))
        s1 = conv(3, conv(2, F.avg_pool2d(s0, 2)))
        s2 = conv(5, conv(4, F.avg_pool2d(s1, 2)))
        b = conv(7, conv(6, F.avg_pool2d(s2, 2)))
        d2 = conv(9, conv(8, up_cat(b, s2)))
        d1 = conv(11, conv(10, up_cat(d2, s1)))
        d0 = conv(13, conv(12, up_cat(d1, s0)))
        return (d0 * hw.view(1, 16, 1, 1)).sum(1) + hb

    return p, f


def test_forward_and_gradient_match_torch_float64():
    torch = pytest.importorskip("torch")
    params = init_params(1)
    x, t = sample(2, B=2, H=16, W=8)
    p, f = torch_net(params)
    xt = torch.tensor(x.astype(np.float64))
    out = f(xt)
    ours = O.unet_forward(params, x, round_bf16=False)
    assert np.allclose(out.detach().numpy(), ours, rtol=1e-5, atol=1e-6)  # float32 output
    tt = torch.tensor(t.astype(np.float64))
    loss = (((out - tt) ** 2).sum((1, 2)) / (tt ** 2).sum((1, 2))).mean()
    loss.backward()
    l2, g = O.unet_loss_grad(params, x, t, round_bf16=False)
    assert abs(l2 - loss.item()) <= 1e-12 * abs(loss.item())
    gt = p.grad.numpy()
    assert np.linalg.norm(g - gt) <= 1e-10 * np.linalg.norm(gt)


def test_finite_difference_gradient():
    params = init_params(3)
    x, t = sample(4)
    _, g = O.unet_loss_grad(params, x, t, round_bf16=False)
    rng = np.random.default_rng(5)
    idx = [O.unet_param_offset(i) + k for i in range(14) for k in (0, 7)]
    idx += list(rng.choice(len(params), 40, replace=False)) + [len(params) - 1, len(params) - 5]
    gmax = np.abs(g).max()
    bad = 0
    for i in idx:
        h = 1e-4 * max(abs(float(params[i])), 1e-2)
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = O.unet_loss_grad(pp, x, t, round_bf16=False)
        lm, _ = O.unet_loss_grad(pm, x, t, round_bf16=False)
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        if abs(fd - g[i]) > 1e-4 * (abs(g[i]) + 1e-3 * gmax):
            bad += 1  # a ReLU kink inside [p - h, p + h]
    assert bad <= 2, bad


def test_bf16_mode_close_to_exact():
    params = init_params(6)
    x, t = sample(7, B=2, H=16, W=16)
    e = O.unet_forward(params, x, round_bf16=False)
    r = O.unet_forward(params, x, round_bf16=True)
    assert np.linalg.norm(r - e) <= 3e-2 * np.linalg.norm(e)
    _, ge = O.unet_loss_grad(params, x, t, round_bf16=False)
    _, gr = O.unet_loss_grad(params, x, t, round_bf16=True)
    assert np.linalg.norm(gr - ge) <= 5e-2 * np.linalg.norm(ge)


def torch_stencil(logk, v):
    """A v for the 2D finite-volume diffusion operator of exp(logk) (torch)."""
    import torch
    k = torch.exp(logk)
    H, W = k.shape[-2:]
    out = torch.zeros_like(v)
    for dy, dx in ((-1, 0), (0, -1), (0, 1), (1, 0)):
        kq = torch.roll(k, shifts=(-dy, -dx), dims=(-2, -1))
        vq = torch.roll(v, shifts=(-dy, -dx), dims=(-2, -1))
        f = 2 * k * kq / (k + kq)
        yy = torch.arange(H).view(H, 1) + dy
        xx = torch.arange(W).view(1, W) + dx
        inside = ((yy >= 0) & (yy < H) & (xx >= 0) & (xx < W)).to(v.dtype)
        out = out + inside * f * (v - vq) + (1 - inside) * k * v
    return out


def test_residual_loss_matches_torch_and_scipy_matrix():
    torch = pytest.importorskip("torch")
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "sgamd_problems", os.path.join(os.path.dirname(__file__), "..",
                                       "schwarz-gmres-dashboard_amd", "problems.py"))
    P = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(P)
    params = init_params(8)
    x, _ = sample(9, B=2, H=16, W=8)
    x[:, 1] *= 0.7
    # the torch stencil is the problems.py diffusion2d matrix
    v = np.random.default_rng(1).standard_normal((16, 8))
    rp, ci, vals = P.diffusion2d(np.exp(x[0, 1].astype(np.float64)))
    from scipy.sparse import csr_matrix
    Am = csr_matrix((vals, ci, rp), shape=(128, 128))
    ts = torch_stencil(torch.tensor(x[0, 1].astype(np.float64)), torch.tensor(v)).numpy()
    assert np.allclose(ts.ravel(), Am @ v.ravel(), rtol=1e-12, atol=1e-12)
    p, f = torch_net(params)
    xt = torch.tensor(x.astype(np.float64))
    out = f(xt)
    s = xt[:, 0]
    u = torch_stencil(xt[:, 1], out)
    a = (s * u).sum((1, 2))
    loss = (1 - a * a / ((s * s).sum((1, 2)) * (u * u).sum((1, 2)))).mean()
    loss.backward()
    l2, g = O.unet_loss_grad(params, x, None, round_bf16=False, kind=1)
    assert abs(l2 - loss.item()) <= 1e-10 * abs(loss.item())
    assert np.linalg.norm(g - p.grad.numpy()) <= 1e-9 * np.linalg.norm(p.grad.numpy())
    # kind 2: the same plus 0.01 ln(|u|^2 / |s|^2)^2 per sample (the scale term)
    p.grad = None
    out = f(xt)
    u = torch_stencil(xt[:, 1], out)
    a = (s * u).sum((1, 2))
    ss, cc = (s * s).sum((1, 2)), (u * u).sum((1, 2))
    loss2 = (1 - a * a / (ss * cc) + 0.01 * torch.log(cc / ss) ** 2).mean()
    loss2.backward()
    assert abs(torch.log(cc / ss)).min().item() > 0.1  # the scale term is active
    l3, g3 = O.unet_loss_grad(params, x, None, round_bf16=False, kind=2)
    assert abs(l3 - loss2.item()) <= 1e-10 * abs(loss2.item())
    assert np.linalg.norm(g3 - p.grad.numpy()) <= 1e-9 * np.linalg.norm(p.grad.numpy())
