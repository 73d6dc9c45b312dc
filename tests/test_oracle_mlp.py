"""Pins the MLP oracle: exact-mode forward and gradients against torch
autograd in float64 on the CPU (an independent implementation), a finite
difference check, and the BF16 rounding helper against torch's bfloat16."""
import numpy as np
import pytest

from oracle import mlp as OM

torch = pytest.importorskip("torch")


def init(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in OM.SHAPES:
        bound = 1.0 / np.sqrt(i)
        parts += [rng.uniform(-bound, bound, o * i), rng.uniform(-bound, bound, o)]
    return np.concatenate(parts)


def torch_net(p):
    layers = OM.unflatten(p)
    mods = []
    for k, (W, b) in enumerate(layers):
        lin = torch.nn.Linear(W.shape[1], W.shape[0]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(W))
            lin.bias.copy_(torch.from_numpy(b))
        mods.append(lin)
        if k < 4:
            mods.append(torch.nn.SiLU())
    return torch.nn.Sequential(*mods)


def test_param_count():
    assert OM.n_params() == 203277


def test_exact_forward_and_gradients_match_torch_autograd():
    p = init(1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(37, 9))
    gy = rng.normal(size=(37, 13))
    y, cache = OM.forward(p, x, bf=False)
    net = torch_net(p)
    xt = torch.from_numpy(x)
    yt = net(xt)
    assert np.allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    (yt * torch.from_numpy(gy)).sum().backward()
    ref = []
    for m in net:
        if isinstance(m, torch.nn.Linear):
            ref += [m.weight.grad.numpy().ravel(), m.bias.grad.numpy()]
    ref = np.concatenate(ref)
    g = OM.backward(p, cache, gy, bf=False)
    assert np.allclose(g, ref, rtol=1e-10, atol=1e-12)


def test_finite_difference_gradient():
    p = init(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(5, 9))
    gy = rng.normal(size=(5, 13))
    y, cache = OM.forward(p, x, bf=False)
    g = OM.backward(p, cache, gy, bf=False)
    for idx in rng.choice(len(p), size=12, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = ((OM.forward(pp, x, bf=False)[0] - OM.forward(pm, x, bf=False)[0]) * gy).sum() / (2 * h)
        assert abs(fd - g[idx]) < 1e-6 * max(1.0, abs(fd)), idx


def test_bf16_rounding_matches_torch():
    v = np.random.default_rng(5).normal(size=1000) * 10.0 ** np.random.default_rng(6).integers(-8, 8, 1000)
    ref = torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(OM.bf16(v), ref)


def test_bf16_mode_is_close_to_exact():
    p = init(7)
    x = np.random.default_rng(8).normal(size=(64, 9))
    ye, _ = OM.forward(p, x, bf=False)
    yb, _ = OM.forward(p, x, bf=True)
    assert np.linalg.norm(yb - ye) / np.linalg.norm(ye) < 3e-2
