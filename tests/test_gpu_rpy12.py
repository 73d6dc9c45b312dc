"""Two-body RPY mobility and move on the GPU (pst_rpy12_eval / pst_rpy12_move,
the __device__ code of the production step) against the FP64 oracle
(oracle/rpy12.py) on identical inputs.

Tolerance (FP64 state contract): |gpu - oracle| <= 1e-12 * max|oracle| per
matrix / per displacement vector; poses and quaternions to 1e-12 absolute.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import rpy12  # noqa: E402
from oracle.physics import shoemake  # noqa: E402
from pystarc_amd import _lib  # noqa: E402
from pystarc_amd.hydrodynamics.rotne_prager import rpy12_mobility  # noqa: E402

DEV = "cuda:0"
RADII = [(17.0, 15.0), (15.0, 17.0), (1.0, 5.0), (3.0, 3.0)]


def _poses(n, rng, rmax):
    u = rng.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = rng.uniform(0.05, rmax, (n, 1))
    p = u * r
    p[:4] = [[0, 0, 9.0], [0, 0, -9.0], [0, 1e-3, 7.0], [5.0, 0, 0]]  # poles of the basis
    return p


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("a0,a1", RADII)
def test_mobility_and_factor_match_oracle(ctx, a0, a1):
    rng = np.random.default_rng(int(a0 * 10 + a1))
    pos = _poses(257, rng, 2.5 * (a0 + a1))
    M, G = rpy12_mobility(pos, a0, a1, ctx)
    M, G = M.cpu().numpy(), G.cpu().numpy()
    Mo = rpy12.grand_mobility_v(pos, a0, a1)
    Go = rpy12.aligned_factor_v(pos, a0, a1)
    for i in range(len(pos)):
        s = np.abs(Mo[i]).max()
        assert np.abs(M[i] - Mo[i]).max() <= 1e-12 * s, i
        assert np.abs(G[i] - Go[i]).max() <= 1e-12 * np.abs(Go[i]).max(), i
        assert np.abs(G[i] @ G[i].T - Mo[i]).max() <= 1e-12 * s, i


@pytest.mark.parametrize("torque_mode", [0, 1])
def test_move_matches_oracle(ctx, torque_mode):
    rng = np.random.default_rng(5 + torque_mode)
    n, a0, a1, Df = 513, 17.0, 15.0, 1.0 / 0.243
    pos = _poses(n, rng, 80.0)
    q = shoemake(rng.random((n, 3)))
    F = rng.standard_normal((n, 3)) * 3.0
    T = rng.standard_normal((n, 3)) * 10.0 if torque_mode else np.zeros((n, 3))
    xi = rng.standard_normal((n, 12))
    dt = rng.uniform(0.05, 20.0, n)
    t = {k: torch.tensor(v, dtype=torch.float64, device=DEV)
         for k, v in dict(pos=pos, q=q, F=F, T=T, xi=xi, dt=dt).items()}
    dX = torch.empty((n, 12), dtype=torch.float64, device=DEV)
    P = _lib.ptr
    ctx.check(ctx.lib.pst_rpy12_move(ctx.h, n, a0, a1, Df, torque_mode, P(t["dt"]), P(t["pos"]),
                                     P(t["q"]), P(t["F"]), P(t["T"]), P(t["xi"]), P(dX)),
              "pst_rpy12_move")
    FR, TR = rpy12.receptor_load_v(pos, F, T, torque_mode)
    dXo = rpy12.displacement_v(pos, F, T, FR, TR, xi, dt, Df, a0, a1)
    po, qo = rpy12.move_v(pos, q, dXo)
    dXg = dX.cpu().numpy()
    for i in range(n):
        assert np.abs(dXg[i] - dXo[i]).max() <= 1e-12 * np.abs(dXo[i]).max(), i
    np.testing.assert_allclose(t["pos"].cpu().numpy(), po, rtol=0, atol=1e-12)
    np.testing.assert_allclose(t["q"].cpu().numpy(), qo, rtol=0, atol=1e-12)


def test_bad_arguments_raise(ctx):
    p = torch.zeros((4, 3), dtype=torch.float64, device=DEV)
    assert ctx.lib.pst_rpy12_eval(ctx.h, 4, _lib.ptr(p), -1.0, 2.0, _lib.ptr(p), None) != 0
    assert "bad arguments" in ctx.error()
