"""GPU checks of the learned-start net (BF16 MFMA convolution kernels,
include/trlnet.h) against a plain PyTorch FP32 reference of the same net on
the CPU: forward outputs, loss and parameter gradients, and that Adam steps
on the device lower the loss."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from nutrlan_amd import learned

pytestmark = pytest.mark.gpu
C, T = 32, 9


def bf16(a):
    """Round-to-nearest-even to BF16, returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    o = 0
    w1 = p[o:o + C * T].reshape(C, T); o += C * T
    b1 = p[o:o + C]; o += C
    w2 = p[o:o + C * T * C].reshape(C, T, C); o += C * T * C
    b2 = p[o:o + C]; o += C
    w3 = p[o:o + C * T * C].reshape(C, T, C); o += C * T * C
    b3 = p[o:o + C]; o += C
    w4 = p[o:o + T * C].reshape(T, C); o += T * C
    b4 = p[o:o + 1]
    return w1, b1, w2, b2, w3, b3, w4, b4


def torch_weights(p, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float32))
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    tw = [torch.tensor(r(w1).reshape(C, 1, 3, 3)),
          torch.tensor(r(w2).reshape(C, 3, 3, C).transpose(0, 3, 1, 2).copy()),
          torch.tensor(r(w3).reshape(C, 3, 3, C).transpose(0, 3, 1, 2).copy()),
          torch.tensor(r(w4).reshape(3, 3, C).transpose(2, 0, 1).reshape(1, C, 3, 3).copy())]
    tb = [torch.tensor(np.asarray(b, np.float32)) for b in (b1, b2, b3, b4)]
    return tw, tb


def torch_forward(x, tw, tb, round_act=True):
    a = torch.tensor(bf16(x) if round_act else x).unsqueeze(1)
    for i in range(4):
        a = F.conv2d(F.pad(a, (1, 1, 1, 1), mode="circular"), tw[i], tb[i])
        if i < 3:
            a = F.relu(a)
            if round_act:
                a = torch.tensor(bf16(a.numpy()))
    return a.squeeze(1)


def loss_t(o, t):
    o = o.flatten(1)
    t = t.flatten(1)
    c = (o * t).sum(1) ** 2 / ((o * o).sum(1) * (t * t).sum(1))
    return (1 - c).mean()


@pytest.mark.parametrize("shape", [(2, 20, 24), (3, 16, 16), (1, 64, 40)])
def test_forward_matches_torch(gpu_lib, shape):
    net = learned.Net(seed=3)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, shape).astype(np.float32)
    y = net.forward(x)
    tw, tb = torch_weights(net.params())
    ref = torch_forward(x, tw, tb).numpy()
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err <= 2e-2, err
    net.close()


def test_loss_and_gradients_match_torch(gpu_lib):
    net = learned.Net(seed=5)
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (4, 18, 18)).astype(np.float32)
    t = np.abs(rng.standard_normal((4, 18, 18))).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    p = net.params()
    # FP32 reference with the device's BF16 weights as leaf tensors
    tw, tb = torch_weights(p)
    for v in tw + tb:
        v.requires_grad_(True)
    o = torch_forward(x, tw, tb, round_act=False)
    lt = loss_t(o, torch.tensor(t))
    lt.backward()
    assert abs(loss - lt.item()) <= 2e-2 * max(lt.item(), 1e-3)
    gw = [v.grad.numpy() for v in tw]
    gb = [v.grad.numpy() for v in tb]
    ref = np.concatenate([
        gw[0].reshape(C, T).ravel(), gb[0],
        gw[1].transpose(0, 2, 3, 1).reshape(C, T, C).ravel(), gb[1],
        gw[2].transpose(0, 2, 3, 1).reshape(C, T, C).ravel(), gb[2],
        gw[3].reshape(C, T).T.ravel(), gb[3]])
    # per layer: relative error of the gradient block (BF16 activations and
    # gradients on the device)
    bounds = np.cumsum([0, C * T, C, C * T * C, C, C * T * C, C, T * C, 1])
    for k in range(8):
        a, b = bounds[k], bounds[k + 1]
        rel = np.linalg.norm(g[a:b] - ref[a:b]) / max(np.linalg.norm(ref[a:b]), 1e-12)
        assert rel <= 8e-2, (k, rel)
    net.close()


def test_adam_steps_lower_the_loss(gpu_lib):
    net = learned.Net(seed=7)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (32, 16, 16)).astype(np.float32)
    # a target the net can represent: a smoothed positive function of x
    t = np.exp(-2.0 * x)
    t = (t + np.roll(t, 1, 1) + np.roll(t, -1, 1) + np.roll(t, 1, 2) + np.roll(t, -1, 2)) / 5
    l0, _ = net.loss_grad(x, t)
    for _ in range(60):
        net.train_step(x, t, 3e-3)
    l1, _ = net.loss_grad(x, t)
    assert l1 < 0.5 * l0, (l0, l1)
    net.close()


def test_forward_many_matches_forward(gpu_lib):
    """trlnet_forward_many over groups of different sizes gives exactly what
    one trlnet_forward per group gives."""
    net = learned.Net(seed=11)
    rng = np.random.default_rng(12)
    groups = [rng.uniform(-1, 1, s).astype(np.float32) for s in ((3, 16, 16), (1, 40, 24), (2, 64, 64))]
    many = net.forward_many(groups)
    for g, y in zip(groups, many):
        assert np.array_equal(y, net.forward(g))
    net.close()
