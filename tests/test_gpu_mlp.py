"""Surrogate MLP kernels (BF16 MFMA, FP32 accumulate) against the FP64
oracle with the same BF16 rounding points.

Tolerances: forward outputs within 5e-3 of the oracle in relative L2 norm;
each layer's weight and bias gradients within 3e-2 in relative L2 norm.  The
residual differences come from FP32 vs FP64 accumulation and from BF16
re-rounding of activations / gradients that sit near a rounding boundary.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import mlp as OM  # noqa: E402
from pystarc_amd import _lib  # noqa: E402
from pystarc_amd._lib import ptr  # noqa: E402


def init(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in OM.SHAPES:
        bound = 1.0 / np.sqrt(i)
        parts += [rng.uniform(-bound, bound, o * i), rng.uniform(-bound, bound, o)]
    return np.concatenate(parts).astype(np.float32)


@pytest.fixture(scope="module")
def mlp():
    ctx = _lib.Context(0)
    h = C.c_void_p()
    ctx.check(ctx.lib.pst_mlp_create(ctx.h, 4096, C.byref(h)), "pst_mlp_create")
    yield ctx, h
    ctx.lib.pst_mlp_destroy(h)
    ctx.close()


def run_fwd(ctx, h, p, x, train):
    dp = torch.as_tensor(p, device="cuda")
    ctx.check(ctx.lib.pst_mlp_load(h, ptr(dp)), "load")
    dx = torch.as_tensor(x.astype(np.float32), device="cuda")
    y = torch.empty((x.shape[0], 13), dtype=torch.float32, device="cuda")
    ctx.check(ctx.lib.pst_mlp_forward(h, x.shape[0], ptr(dx), ptr(y), int(train)), "forward")
    torch.cuda.synchronize()
    return y.cpu().numpy().astype(np.float64), dp


def test_param_count(mlp):
    ctx, h = mlp
    assert ctx.lib.pst_mlp_param_count() == OM.n_params() == 203277


@pytest.mark.parametrize("B", [1, 64, 100, 1000])
def test_forward_matches_oracle(mlp, B):
    ctx, h = mlp
    p = init(1)
    x = np.random.default_rng(B).normal(size=(B, 9))
    y, _ = run_fwd(ctx, h, p, x, train=False)
    yo, _ = OM.forward(p.astype(np.float64), x, bf=True)
    assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 5e-3


def test_backward_matches_oracle(mlp):
    ctx, h = mlp
    p = init(2)
    B = 300
    rng = np.random.default_rng(3)
    x = rng.normal(size=(B, 9))
    gy = rng.normal(size=(B, 13)) / B
    y, dp = run_fwd(ctx, h, p, x, train=True)
    dgy = torch.as_tensor(gy.astype(np.float32), device="cuda")
    g = torch.zeros(OM.n_params(), dtype=torch.float32, device="cuda")
    ctx.check(ctx.lib.pst_mlp_backward(h, B, ptr(dgy), ptr(g)), "backward")
    torch.cuda.synchronize()
    g = g.cpu().numpy().astype(np.float64)
    _, cache = OM.forward(p.astype(np.float64), x, bf=True)
    go = OM.backward(p.astype(np.float64), cache, gy, bf=True)
    k = 0
    for o, i in OM.SHAPES:
        for n in (o * i, o):
            a, b = g[k:k + n], go[k:k + n]
            assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-2, (o, i, n)
            k += n
    # and the exact FP64 network agrees to BF16 accuracy
    _, ce = OM.forward(p.astype(np.float64), x, bf=False)
    ge = OM.backward(p.astype(np.float64), ce, gy, bf=False)
    assert np.linalg.norm(g - ge) / np.linalg.norm(ge) < 5e-2


def test_training_fits_a_smooth_map():
    """The kernels' gradients drive Adam to fit a smooth 9 -> 13 function."""
    from pystarc_amd.surrogate.model import Surrogate

    class _Eng:  # minimal engine stand-in exposing the context and device
        pass

    ctx = _lib.Context(0)
    eng = _Eng()
    eng.ctx, eng.device, eng._debye, eng.torque_mode = ctx, torch.device("cuda"), 7.86, "reference"
    s = Surrogate(eng, r_on=40, r_max=120, b=60, max_rows=8192)
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(65536, 9)).astype(np.float32)
    A = rng.normal(size=(9, 13)).astype(np.float32) / 3
    Y = np.tanh(X @ A) + 0.3 * np.sin(2 * X[:, :1])
    Xt = torch.as_tensor(X, device="cuda")
    Yt = torch.as_tensor(Y.astype(np.float32), device="cuda")
    hist = s.train(Xt, Yt, epochs=6, batch=2048, lr=2e-3)
    assert hist[-1] < 0.05 * float(np.var(Y))
    s.close()


@pytest.mark.parametrize("B", [1, 63, 64, 127, 128, 129, 1000, 4000, 4096])
def test_fused_inference_equals_layer_gemms(mlp, B):
    """The fused inference kernel (all layers per 128-row tile, activations in
    LDS, the transposed product W A^T) forms every output from the same
    products in the layer GEMMs' k order with the same roundings:
    bit-identical outputs to the train-mode forward's five layer launches,
    also for batches that end inside a tile."""
    ctx, h = mlp
    p = init(4)
    x = np.random.default_rng(10 + B).normal(size=(B, 9))
    y_fused, _ = run_fwd(ctx, h, p, x, train=False)
    y_layers, _ = run_fwd(ctx, h, p, x, train=True)
    assert np.array_equal(y_fused, y_layers)


def test_time_infer_measures_the_kernel_and_rejects_bad_sizes(mlp):
    """pst_mlp_time_infer times the fused kernel alone on the features the
    last forward packed; its launches recompute the same outputs."""
    ctx, h = mlp
    p = init(5)
    x = np.random.default_rng(5).normal(size=(1000, 9))
    y, _ = run_fwd(ctx, h, p, x, train=False)
    ms = C.c_double()
    ctx.check(ctx.lib.pst_mlp_time_infer(h, 1000, 3, C.byref(ms)), "pst_mlp_time_infer")
    assert 0.0 < ms.value < 100.0
    y2, _ = run_fwd(ctx, h, p, x, train=False)
    assert np.array_equal(y, y2)
    assert ctx.lib.pst_mlp_time_infer(h, 0, 3, C.byref(ms)) != 0
    assert ctx.lib.pst_mlp_time_infer(h, 1 << 20, 3, C.byref(ms)) != 0
    assert ctx.lib.pst_mlp_time_infer(h, 10, 0, C.byref(ms)) != 0
