"""pst_forces_positions_noalloc (the force operator that allocates nothing)
against pst_forces_positions on the same inputs, through the C ABI.

The reverse-Born chunk skip of the reference (gpu_batch_engine.py:626-639)
is the part with scratch memory: per-chunk minimum centroid distances, which
the no-alloc entry point keeps in a static device array shared by every call
on the GPU.  Two calls in a row, the second smaller, on the 3,220-atom
protein-ligand receptor (6,784 trajectories per reference chunk): the first
call's first chunk is skipped (all poses far), the second call's first chunk
is not, so a stale scratch entry would change the second call's forces.
Results must be bit-identical (same kernels, same lanes, same sums).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.physics import quats_to_rotmats, shoemake  # noqa: E402
from pystarc_amd import systems as S  # noqa: E402
from pystarc_amd._lib import Context, ptr  # noqa: E402


def _poses(rng, n, rmin, rmax):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * rng.uniform(rmin, rmax, size=(n, 1)), shoemake(rng.random((n, 3)))


@pytest.fixture(scope="module")
def c2_engine():
    sysm = S.protein_ligand(n_traj=1000)
    ctx = Context(0)
    grids = S.build_grids(sysm, ctx)
    ctx.close()
    return sysm, S.make_engine(sysm, grids)


def _call(eng, fn, lab, q, R, c, prec):
    n, m = lab.shape[0], lab.shape[1]
    dev = eng.device
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    pos, qq, RR, cc = t(lab), t(q), t(R), t(c)
    F = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    T = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    E = torch.zeros((n,), dtype=torch.float64, device=dev)
    rc = getattr(eng.ctx.lib, fn)(eng.ctx.h, n, m, ptr(pos), ptr(qq), ptr(RR), ptr(cc), ptr(F),
                                  ptr(T), ptr(E), prec, 1)
    eng.ctx.check(rc, fn)
    torch.cuda.synchronize()
    return F.cpu().numpy(), T.cpu().numpy(), E.cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_allocating_entry_point(c2_engine, prec):
    sysm, eng = c2_engine
    rng = np.random.default_rng(77 + prec)
    body = sysm.mol2.positions_array() - sysm.mol2.centroid()
    lq = sysm.mol2.charges_array()
    # call 1: 16,000 poses = chunks of 6,784 / 6,784 / 2,432; chunk 0 and 2 far
    # (skipped), chunk 1 mixed
    c_far0, q_far0 = _poses(rng, 6784, 45.0, 60.0)
    c_mid, q_mid = _poses(rng, 6784, 20.0, 60.0)
    c_far2, q_far2 = _poses(rng, 2432, 45.0, 60.0)
    c1 = np.concatenate([c_far0, c_mid, c_far2])
    q1 = np.concatenate([q_far0, q_mid, q_far2])
    # call 2: 5,000 near poses, one chunk that must not be skipped
    c2, q2 = _poses(rng, 5000, 22.0, 40.0)
    for c, q in ((c1, q1), (c2, q2)):
        print(f"noalloc prec={prec} n={len(c)}", flush=True)
        R = quats_to_rotmats(q)
        lab = np.einsum("nij,mj->nmi", R, body) + c[:, None, :]
        ref = _call(eng, "pst_forces_positions", lab, lq, R, c, prec)
        got = _call(eng, "pst_forces_positions_noalloc", lab, lq, R, c, prec)
        for a, b, what in zip(got, ref, "FTE"):
            assert np.array_equal(a, b), f"{what} differs (n={len(c)}, prec={prec})"
        assert np.abs(ref[0]).max() > 0
