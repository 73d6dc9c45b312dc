"""Host side of the Python API on the CPU: argument checks that raise before
any device work, no silent fallback without a device, and the simulator's
k_b / return probability (gpu_batch_simulator.py:320-404), which the host
computes by Romberg quadrature, against the analytic two-spheres numbers and
the oracle."""
import math
import types

import numpy as np
import pytest

from oracle import physics as ph
from pystarc_amd import systems as S
from pystarc_amd.simulation.gpu_batch_simulator import GPUBatchSimulator


def stub_engine(debye=7.86, precision="fp64"):
    # what the simulator's constructor and quadrature read from the engine
    return types.SimpleNamespace(precision=precision, _debye=debye)


def make(sysm, **kw):
    eng = stub_engine(sysm.engine_kwargs.get("debye_length", 7.86))
    return GPUBatchSimulator(sysm.mol1, sysm.mol2, sysm.mob, sysm.pathway_set, sysm.params, eng,
                             **kw)


def test_two_spheres_return_probability_is_analytic():
    """b = 30, q = 60, +1 / -1 e, lambda 7.86: return probability 0.500406 and
    k_b / D = 49.423 / 0.130992 (analytic)."""
    s = S.two_spheres(n_traj=100)
    sim = make(s)
    rp = sim._compute_return_prob()
    D = s.mob.relative_translational_diffusion()
    assert rp == pytest.approx(0.500406, abs=2e-5)
    assert sim._k_b / D == pytest.approx(49.423 / 0.130992, rel=5e-5)


def test_protein_pair_with_hydrodynamics_matches_the_oracle():
    s = S.protein_protein(n_traj=100)
    sim = make(s)
    assert sim._use_hi
    rp = sim._compute_return_prob()
    rp_o, kb_o = ph.return_prob_and_kb(
        float(s.params.r_start), float(s.params.r_escape), float(s.mol1.total_charge()),
        float(s.mol2.total_charge()), s.engine_kwargs["debye_length"],
        float(s.mob.relative_translational_diffusion()), use_hi=True,
        a0=float(s.mob.radius1), a1=float(s.mob.radius2))
    assert rp == pytest.approx(rp_o, rel=1e-9)
    assert sim._k_b == pytest.approx(kb_o, rel=1e-9)


def test_neutral_pair_is_free_diffusion():
    s = S.two_spheres(n_traj=100)
    s.mol2 = S.Molecule.from_arrays(s.mol2.positions_array(), np.zeros(len(s.mol2.atoms)),
                                    s.mol2.radii_array(), name="neutral")
    sim = make(s)
    rp = sim._compute_return_prob()
    b, q = float(s.params.r_start), float(s.params.r_escape)
    assert rp == pytest.approx(b / q)
    assert sim._k_b == pytest.approx(4 * math.pi * s.mob.relative_translational_diffusion() * b)


@pytest.mark.parametrize("kw,err", [
    ({"rng": "mersenne"}, ValueError),
    ({"traj_range": (-1, 10)}, ValueError),
    ({"traj_range": (0, 10), "rng": "numpy"}, ValueError),
    ({"outer": "nam"}, ValueError),
    ({"outer": "lmz", "rng": "numpy"}, ValueError),
    ({"hydro": "oseen"}, ValueError),
    ({"hydro": "rpy12", "rng": "numpy"}, ValueError),
    ({"gather": "texture"}, ValueError),
    ({"record": "everything"}, ValueError),
    ({"precision": "fp16"}, ValueError),
    ({"surrogate": object(), "rng": "numpy"}, ValueError),
])
def test_simulator_rejects_bad_arguments(kw, err):
    with pytest.raises(err):
        make(S.two_spheres(n_traj=100), **kw)


def test_engine_rejects_bad_arguments_before_the_device():
    from pystarc_amd.forces.gpu_batch_engine import GPUBatchForceEngine
    with pytest.raises(NotImplementedError):
        GPUBatchForceEngine([], [], use_lj=True)
    with pytest.raises(ValueError):
        GPUBatchForceEngine([], [], precision="fp16")
    with pytest.raises(ValueError):
        GPUBatchForceEngine([], [], torque_mode="lever")


def test_engine_without_a_device_raises():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    from pystarc_amd.forces.gpu_batch_engine import GPUBatchForceEngine
    with pytest.raises(RuntimeError):
        GPUBatchForceEngine([], [])
