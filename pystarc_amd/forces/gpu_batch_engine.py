"""Drop-in GPUBatchForceEngine backed by the gfx950 kernels of libpystarc_hip.

Same constructor keywords and call signature as
pystarc/forces/gpu_batch_engine.py:139-154 and :531-605:

    engine = GPUBatchForceEngine(elec_grids, born_grids, alpha=..., receptor_charge=...,
                                 debye_length=..., sdie=..., lig_born_grids=...,
                                 rec_positions=..., rec_charges=..., multipole_expansion=...)
    F, T, E = engine(lig_positions, lig_charges, R_matrices=R, centroids=c)

Grids are uploaded once and stay resident in HBM.  Grid selection follows the
reference constructor: finest first by delta[0,0]; with a Yukawa far-field
only the finest electrostatic grid is kept; only the finest receptor Born and
ligand Born grids are kept (:179-211).  Inputs may be NumPy arrays or torch
tensors on the device; outputs are torch float64 tensors on the device.

Extra keyword-only options (defaults give the reference behaviour):
  precision   "fp64" (FP64 gather and sums) or "fp32" (FP32 gather, FP64
              cross-lane reduction); trajectory state is FP64 either way.
  torque_mode "reference" (zero torque, :597-599) or "physical" (lever arm
              about the centroid of charged atoms, pystarc/forces/engine.py:144-183).
  grid_dtype  None keeps each grid's dtype, or "float32"/"float64".
  receptor_from  another engine on the same device whose receptor grids this
              engine uses without a copy (several ligands against one
              receptor; the other engine must stay alive).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .._lib import Context, ptr
from ..constants import EPS0

_KIND = {"elec": 0, "born": 1, "lig_born": 2}
# force-evaluation modes of the library, by number (pst_set_gather_mode)
GATHER_MODES = ("direct", "binned", "zpair", "split", "split_zpair")


def _grid_parts(g):
    data = np.asarray(g.data)
    delta = np.asarray(g.delta, dtype=float)
    sp = np.array([delta[i, i] for i in range(3)], dtype=float)
    return data, np.asarray(g.origin, dtype=float), sp


def _bounds(data, origin, sp):
    dims = np.array(data.shape)
    return origin + 3.0 * sp, origin + (dims - 1 - 3.0) * sp


class GPUBatchForceEngine:
    def __init__(self, elec_grids, born_grids, alpha: float = 0.07957747,
                 receptor_charge: float = 0.0, debye_length: float = 7.858, sdie: float = 78.0,
                 lig_born_grids=None, rec_positions=None, rec_charges=None,
                 multipole_expansion=None, rec_radii=None, lig_radii=None, use_lj=False, *,
                 precision: str = "fp64", torque_mode: str = "reference", grid_dtype=None,
                 device: int = 0, force_group: int = 0, receptor_from=None):
        if use_lj:
            raise NotImplementedError("WCA repulsion (use_lj=True) is not part of this build")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if torque_mode not in ("reference", "physical"):
            raise ValueError("torque_mode must be 'reference' or 'physical'")
        import torch
        self._torch = torch
        self.device = torch.device("cuda", device)
        self.ctx = Context(device)
        self.alpha = float(alpha)
        self.precision = precision
        self.torque_mode = torque_mode
        self._rec_charge = float(receptor_charge)
        self._debye = float(debye_length)
        self._sdie = float(sdie)
        eps_s = sdie * EPS0
        self._V_factor = receptor_charge / (4.0 * math.pi * eps_s) if eps_s > 0 else 0.0
        self._has_yukawa = abs(receptor_charge) > 1e-9 and debye_length > 0
        self._multipole = multipole_expansion
        self._grid_dtype = grid_dtype
        lib, h = self.ctx.lib, self.ctx.h

        elec = sorted(list(elec_grids or []), key=lambda g: float(np.asarray(g.delta)[0, 0]))
        born = sorted(list(born_grids or []), key=lambda g: float(np.asarray(g.delta)[0, 0]))
        if self._has_yukawa and len(elec) > 1:
            elec = elec[:1]
        if len(born) > 1:
            born = born[:1]
        lig_born = []
        if lig_born_grids and rec_positions is not None and rec_charges is not None:
            lig_born = sorted(list(lig_born_grids),
                              key=lambda g: float(np.asarray(g.delta)[0, 0]))[:1]
        self._receptor_src = receptor_from
        if receptor_from is not None:
            # one resident copy of the receptor grids for several ligand engines
            self.ctx.check(lib.pst_share_receptor_grids(h, receptor_from.ctx.h),
                           "pst_share_receptor_grids")
            self._elec_grids_gpu = receptor_from._elec_grids_gpu
            self._born_grids_gpu = receptor_from._born_grids_gpu
        else:
            self._elec_grids_gpu = [self._upload("elec", g) for g in elec]
            self._born_grids_gpu = [self._upload("born", g) for g in born]
        self._lig_born_grids_gpu = [self._upload("lig_born", g) for g in lig_born]

        dip = quad = None
        if multipole_expansion is not None:
            dip = np.ascontiguousarray(multipole_expansion.dipole, dtype=np.float64)
            quad = np.ascontiguousarray(multipole_expansion.quadrupole, dtype=np.float64).ravel()
        self.ctx.check(lib.pst_set_farfield(h, self._rec_charge, self._debye, self._sdie,
                                            ptr(dip), ptr(quad)), "pst_set_farfield")
        self.ctx.check(lib.pst_set_alpha(h, self.alpha), "pst_set_alpha")
        self._rec_pos = None
        if rec_positions is not None and rec_charges is not None:
            rp = np.ascontiguousarray(rec_positions, dtype=np.float64).reshape(-1, 3)
            rq = np.ascontiguousarray(rec_charges, dtype=np.float64).ravel()
            self._rec_pos, self._rec_q = rp, rq
            self.ctx.check(lib.pst_set_receptor(h, len(rq), ptr(rp), ptr(rq)), "pst_set_receptor")
        self.ctx.check(lib.pst_set_torque_mode(h, 0 if torque_mode == "reference" else 1),
                       "pst_set_torque_mode")
        self.ctx.check(lib.pst_set_force_group(h, int(force_group)), "pst_set_force_group")
        self._lig_key = None

    # -- set-up helpers ---------------------------------------------------
    def _upload(self, kind, g):
        data, origin, sp = _grid_parts(g)
        if self._grid_dtype is not None:
            data = data.astype(self._grid_dtype)
        if data.dtype == np.float32:
            dt = 0
        else:
            data = data.astype(np.float64)
            dt = 1
        data = np.ascontiguousarray(data)
        nx, ny, nz = data.shape
        self.ctx.check(self.ctx.lib.pst_add_grid(self.ctx.h, _KIND[kind], ptr(data), dt, nx, ny, nz,
                                                 ptr(origin), ptr(sp)), "pst_add_grid")
        lo, hi = _bounds(data, origin, sp)
        return {"nx": nx, "ny": ny, "nz": nz, "origin": origin, "sp": sp, "lo": lo, "hi": hi,
                "dtype": data.dtype}

    def set_gather(self, mode="direct", max_traj=0):
        """Force evaluation of forces_pose (and the default of new simulators):
        "direct" - one lane group per trajectory gathers from the grids;
        "binned" - gather requests sorted by 8^3-cell grid brick, each brick
        staged once in LDS; identical results; buffers for max_traj
        trajectories.  Without the float2 receptor pair grid the direct
        kernel runs.
        "zpair"  - the direct kernel fetching the z nodes of every stencil
        column with wide loads (the same node values, so the same results
        to the rounding of the arithmetic; both precisions).
        "split" / "split_zpair" - production steps only: the ligand-atom
        terms over the grid-path list, then the reverse Born over the list
        re-sorted by the ligand-frame direction of the receptor (identical
        results); elsewhere "direct" / "zpair" run."""
        if mode not in GATHER_MODES:
            raise ValueError(f"mode must be one of {GATHER_MODES}")
        m = GATHER_MODES.index(mode)
        self.ctx.check(self.ctx.lib.pst_set_gather_mode(self.ctx.h, m,
                                                        int(max_traj) if m == 1 else 0),
                       "pst_set_gather_mode")
        self.gather = mode

    def last_eval(self):
        """(mode, G) of the last force evaluation: a GATHER_MODES name (None before
        the first), lanes."""
        m, g = C.c_int(), C.c_int()
        self.ctx.check(self.ctx.lib.pst_last_force_eval(self.ctx.h, C.byref(m), C.byref(g)),
                       "pst_last_force_eval")
        return (GATHER_MODES[m.value] if m.value >= 0 else None), g.value

    def set_ligand(self, body, charges):
        """Body-frame ligand coordinates (about the centroid) for the fused path."""
        body = np.ascontiguousarray(body, dtype=np.float64).reshape(-1, 3)
        charges = np.ascontiguousarray(charges, dtype=np.float64).ravel()
        key = (body.tobytes(), charges.tobytes())
        if key != self._lig_key:
            self.ctx.check(self.ctx.lib.pst_set_ligand(self.ctx.h, len(charges), ptr(body),
                                                       ptr(charges)), "pst_set_ligand")
            self._lig_key = key

    def _dev(self, a, shape=None):
        torch = self._torch
        if a is None:
            return None
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.device, dtype=torch.float64)
        else:
            t = torch.as_tensor(np.asarray(a, dtype=np.float64), device=self.device)
        t = t.contiguous()
        if shape is not None:
            t = t.reshape(shape)
        return t

    def _prec(self, precision):
        p = self.precision if precision is None else precision
        return 0 if p == "fp64" else 1

    # -- operator ------------------------------------------------------------
    def __call__(self, lig_positions, lig_charges, R_matrices=None, centroids=None, *,
                 precision=None, rev_skip="reference"):
        """F (N,3), T (N,3), E (N,) for (N,M,3) lab-frame ligand atoms."""
        torch = self._torch
        pos = self._dev(lig_positions)
        if pos.dim() != 3 or pos.shape[2] != 3:
            raise ValueError("lig_positions must have shape (N, M, 3)")
        N, M = int(pos.shape[0]), int(pos.shape[1])
        q = self._dev(lig_charges, (M,))
        R = self._dev(R_matrices, (N, 3, 3)) if R_matrices is not None else None
        c = self._dev(centroids, (N, 3)) if centroids is not None else None
        if (R is None) != (c is None):
            R = c = None
        F = torch.zeros((N, 3), dtype=torch.float64, device=self.device)
        T = torch.zeros((N, 3), dtype=torch.float64, device=self.device)
        E = torch.zeros((N,), dtype=torch.float64, device=self.device)
        if N == 0:
            return F, T, E
        rc = self.ctx.lib.pst_forces_positions(
            self.ctx.h, N, M, ptr(pos), ptr(q), ptr(R), ptr(c), ptr(F), ptr(T), ptr(E),
            self._prec(precision), 1 if rev_skip == "reference" else 0)
        self.ctx.check(rc, "pst_forces_positions")
        return F, T, E

    def forces_pose(self, centroids, quats, *, precision=None):
        """Fused path: atoms placed from (centroid, quaternion) in registers."""
        torch = self._torch
        if self._lig_key is None:
            raise RuntimeError("call set_ligand(body, charges) before forces_pose")
        c = self._dev(centroids)
        qt = self._dev(quats)
        N = int(c.shape[0])
        F = torch.zeros((N, 3), dtype=torch.float64, device=self.device)
        T = torch.zeros((N, 3), dtype=torch.float64, device=self.device)
        E = torch.zeros((N,), dtype=torch.float64, device=self.device)
        if N:
            self.ctx.check(self.ctx.lib.pst_forces_pose(self.ctx.h, N, ptr(c), ptr(qt), ptr(F),
                                                        ptr(T), ptr(E), self._prec(precision)),
                           "pst_forces_pose")
        return F, T, E


def forces_pose_group(engines, centroids, quats, *, precision=None):
    """Forces of several ligand engines (receptor grids shared through
    ``receptor_from``) in one launch of the grouped force kernel, as the
    several-ligand step evaluates them (pst_forces_pose_group).  centroids /
    quats: one (n_l,3) / (n_l,4) array per engine; returns one (F, T, E)
    tuple of device tensors per engine."""
    if not engines or len(engines) != len(centroids) or len(engines) != len(quats):
        raise ValueError("one centroid and one quaternion array per engine")
    torch = engines[0]._torch
    prec = engines[0]._prec(precision)
    L = len(engines)
    keep, outs = [], []
    for e, c, q in zip(engines, centroids, quats):
        if e._lig_key is None:
            raise RuntimeError("call set_ligand(body, charges) on every engine first")
        ct, qt = e._dev(c), e._dev(q)
        n = int(ct.shape[0])
        F = torch.zeros((n, 3), dtype=torch.float64, device=e.device)
        T = torch.zeros((n, 3), dtype=torch.float64, device=e.device)
        E = torch.zeros((n,), dtype=torch.float64, device=e.device)
        keep += [ct, qt]
        outs.append((F, T, E))
    P = C.c_void_p * L
    ctxs = P(*[e.ctx.h.value for e in engines])
    ns = (C.c_int * L)(*[int(o[0].shape[0]) for o in outs])
    pos = P(*[keep[2 * l].data_ptr() for l in range(L)])
    qua = P(*[keep[2 * l + 1].data_ptr() for l in range(L)])
    Fp, Tp, Ep = (P(*[o[k].data_ptr() for o in outs]) for k in range(3))
    engines[0].ctx.check(engines[0].ctx.lib.pst_forces_pose_group(ctxs, L, ns, pos, qua, Fp, Tp, Ep,
                                                                   prec), "pst_forces_pose_group")
    return outs
