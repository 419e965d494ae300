"""The learned stage's network on the GPU (sg_unet) against the CPU oracle in
its bfloat16 mode (same rounding points; sums in fp32 on the GPU and in double
on the CPU, so a few bfloat16 roundings differ): output within 2e-2 and the
parameter gradient within 5e-2 in relative L2 norm, every layer's gradient
aligned with the oracle's (cosine > 0.99); Adam steps lower the loss."""
import numpy as np
import pytest

import oracle_bind as O
from test_unet_oracle import CONVS, init_params, sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="module")
def ctx(sg):
    c = sg.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape", [(1, 8, 8), (2, 32, 16), (1, 24, 40), (1, 128, 96)])
def test_forward_matches_oracle(sg, ctx, shape):
    B, H, W = shape
    params = init_params(11)
    x, _ = sample(12, B=B, H=H, W=W)
    net = sg.UNet(ctx, params)
    y = net.forward(x)
    ref = O.unet_forward(params, x, round_bf16=True)
    assert rel(y, ref) <= 2e-2, rel(y, ref)


def test_gradient_matches_oracle(sg, ctx):
    params = init_params(13)
    x, t = sample(14, B=2, H=32, W=16)
    net = sg.UNet(ctx, params)
    loss, g = net.loss_grad(x, t)
    lref, gref = O.unet_loss_grad(params, x, t, round_bf16=True)
    assert abs(loss - lref) <= 2e-2 * lref
    assert rel(g, gref) <= 5e-2, rel(g, gref)
    for i in range(14):
        a, b = O.unet_param_offset(i), O.unet_param_offset(i + 1)
        ga, gb = g[a:b], gref[a:b]
        cos = ga @ gb / (np.linalg.norm(ga) * np.linalg.norm(gb) + 1e-30)
        assert cos > 0.99, (i, cos)
    head = slice(O.unet_param_offset(14), None)
    assert rel(g[head], gref[head]) <= 2e-2


@pytest.mark.parametrize("kind", [1, 2])
def test_residual_loss_gradient_matches_oracle(sg, ctx, kind):
    params = init_params(17)
    x, _ = sample(18, B=2, H=32, W=16)
    x[:, 1] *= 0.7
    net = sg.UNet(ctx, params)
    net.set_loss(kind)
    loss, g = net.loss_grad(x, None)
    lref, gref = O.unet_loss_grad(params, x, None, round_bf16=True, kind=kind)
    assert abs(loss - lref) <= 2e-2 * lref
    assert rel(g, gref) <= 5e-2, rel(g, gref)


def test_params_roundtrip_and_adam_lowers_loss(sg, ctx):
    params = init_params(15)
    net = sg.UNet(ctx, params)
    assert np.array_equal(net.get_params(), params)
    rng = np.random.default_rng(16)
    x = rng.standard_normal((4, 2, 16, 16)).astype(np.float32)
    # a smooth target the network can fit: a local average of channel 0
    k = np.ones((3, 3)) / 9.0
    from scipy.signal import convolve2d
    t = np.stack([convolve2d(x[b, 0], k, mode="same") for b in range(4)]).astype(np.float32)
    dx = sg.DeviceArray.from_host(ctx, x.ravel())
    dt = sg.DeviceArray.from_host(ctx, t.ravel())
    l0 = net.loss_grad_dev(4, 16, 16, dx, dt)
    for _ in range(60):
        net.loss_grad_dev(4, 16, 16, dx, dt)
        net.adam(2e-3)
    l1 = net.loss_grad_dev(4, 16, 16, dx, dt)
    assert l1 < 0.5 * l0, (l0, l1)
