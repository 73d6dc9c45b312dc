"""Learned surrogate of the pose -> (force, torque, energy) map outside a
contact shell (no reference counterpart).

Network: 9 inputs (centroid in the receptor frame / b, first two columns of
R(q)) -> 4 hidden layers of 256 with SiLU -> 13 outputs (ligand force and
torque, receptor force and torque, energy).  Forward and backward run in the
library's BF16-MFMA kernels (pst_mlp_forward / pst_mlp_backward); PyTorch
holds the FP32 master parameters and the Adam state only.

Targets are the plain FP64 grid path's own outputs on sampled poses (the
plain path is the reference), divided by a screened-Coulomb envelope
env(r) = e^{-(r-r0)/lam} (r0/r)^2 (1 + r/lam)/(1 + r0/lam) and a per-group
scale, so every target is O(1) across the shell.  The receptor force and
torque follow from momentum and angular-momentum balance:
F_rec = -F_lig, T_rec = -(T_lig + c x F_lig).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from .._lib import ptr

N_PARAMS = 203_277
SHAPES = [(256, 9), (256, 256), (256, 256), (256, 256), (13, 256)]


def envelope(r, r0, lam):
    x = r0 / r
    return np.exp(-(r - r0) / lam) * x * x * (1.0 + r / lam) / (1.0 + r0 / lam)


def _shoemake(u):
    return np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]),
                     np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                     np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]),
                     np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], axis=1)


def _rot(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                     2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                     2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
                    axis=1).reshape(-1, 3, 3)


class Surrogate:
    def __init__(self, engine, r_on, r_max, b, lam=None, max_rows=65536, seed=0):
        import torch
        self.torch = torch
        self.engine = engine
        self.lib, self.ctx = engine.ctx.lib, engine.ctx
        if self.lib.pst_mlp_param_count() != N_PARAMS:
            raise RuntimeError("parameter layout mismatch with the library")
        self.r_on, self.r_max, self.b = float(r_on), float(r_max), float(b)
        self.r0 = float(r_on)
        self.lam = float(lam if lam is not None else engine._debye)
        self.device = engine.device
        h = C.c_void_p()
        self.ctx.check(self.lib.pst_mlp_create(self.ctx.h, int(max_rows), C.byref(h)),
                       "pst_mlp_create")
        self.h = h
        self.max_rows = int(max_rows)
        g = torch.Generator().manual_seed(seed)
        parts = []
        for o, i in SHAPES:
            bound = 1.0 / math.sqrt(i)
            parts += [(torch.rand(o * i, generator=g) * 2 - 1) * bound,
                      (torch.rand(o, generator=g) * 2 - 1) * bound]
        self.params = torch.nn.Parameter(torch.cat(parts).to(self.device, torch.float32))
        self.scale = np.ones(5)   # F_lig, T_lig, F_rec, T_rec, E
        self.use_torque = engine.torque_mode == "physical"
        self.history = []
        self._load()

    def close(self):
        if getattr(self, "h", None):
            self.lib.pst_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _load(self):
        self.ctx.check(self.lib.pst_mlp_load(self.h, ptr(self.params.data)), "pst_mlp_load")

    # -- data -----------------------------------------------------------------------
    def sample_poses(self, n, rng, r_lo=None, r_hi=None, near_frac=0.0, near_width=20.0):
        """Uniform directions and orientations; centre distance uniform in
        [r_lo, r_hi], with a fraction `near_frac` drawn from the first
        `near_width` A beyond r_lo, where the force (and its effect on k_on)
        is largest."""
        r_lo = self.r_on if r_lo is None else r_lo
        r_hi = self.r_max if r_hi is None else r_hi
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        r = rng.uniform(r_lo, r_hi, size=(n, 1))
        near = rng.random(n) < near_frac
        r[near, 0] = rng.uniform(r_lo, min(r_hi, r_lo + near_width), size=int(near.sum()))
        return d * r, _shoemake(rng.random((n, 3)))

    def features(self, c, q):
        R = _rot(q)
        return np.concatenate([c / self.b, R[:, :, 0], R[:, :, 1]], axis=1).astype(np.float32)

    def label(self, c, q, chunk=262144):
        """Plain FP64 path on the poses: (F, T, E) float64 arrays."""
        F, T, E = [], [], []
        for i in range(0, len(c), chunk):
            f, t, e = self.engine.forces_pose(c[i:i + chunk], q[i:i + chunk], precision="fp64")
            F.append(f.cpu().numpy()); T.append(t.cpu().numpy()); E.append(e.cpu().numpy())
        return np.concatenate(F), np.concatenate(T), np.concatenate(E)

    def raw_targets(self, c, F, T, E):
        env = envelope(np.linalg.norm(c, axis=1), self.r0, self.lam)[:, None]
        Trec = -(T + np.cross(c, F))
        return env, [F / env, T / env, -F / env, Trec / env, E[:, None] / env]

    def fit_scales(self, c, F, T, E):
        _, parts = self.raw_targets(c, F, T, E)
        for k, a in enumerate(parts):
            s = float(np.sqrt(np.mean(a ** 2)))
            self.scale[k] = s if s > 0 else 1.0

    def targets(self, c, F, T, E):
        _, parts = self.raw_targets(c, F, T, E)
        return np.concatenate([a / self.scale[k] for k, a in enumerate(parts)], axis=1).astype(np.float32)

    def sample_weights(self, c):
        """Per-pose loss weight: the force envelope, normalised to mean 1, so
        the fit is weighted like the physical force (as the step's audit is)."""
        env = envelope(np.linalg.norm(c, axis=1), self.r0, self.lam)
        return (env / env.mean()).astype(np.float32)

    def make_dataset(self, n, seed=0, extra_poses=None, near_frac=0.5):
        rng = np.random.default_rng(seed)
        c, q = self.sample_poses(n, rng, near_frac=near_frac)
        if extra_poses is not None:
            ce, qe = extra_poses
            keep = np.linalg.norm(ce, axis=1) > self.r_on
            c = np.concatenate([c, ce[keep]])
            q = np.concatenate([q, qe[keep]])
        F, T, E = self.label(c, q)
        return c, q, F, T, E

    # -- network (library kernels) -----------------------------------------------------
    def forward(self, X, train=False):
        torch = self.torch
        B = X.shape[0]
        Y = torch.empty((B, 13), dtype=torch.float32, device=self.device)
        self.ctx.check(self.lib.pst_mlp_forward(self.h, B, ptr(X), ptr(Y), int(train)),
                       "pst_mlp_forward")
        return Y

    def backward(self, GY):
        torch = self.torch
        G = torch.zeros(N_PARAMS, dtype=torch.float32, device=self.device)
        self.ctx.check(self.lib.pst_mlp_backward(self.h, GY.shape[0], ptr(GY), ptr(G)),
                       "pst_mlp_backward")
        return G

    def predict(self, X, chunk=None):
        torch = self.torch
        chunk = chunk or self.max_rows
        out = [self.forward(X[i:i + chunk]) for i in range(0, X.shape[0], chunk)]
        return torch.cat(out)

    def loss_weights(self):
        """Per-output MSE weights: the ligand force (and torque when it is
        applied) carries the loss; the balance outputs and energy are kept
        as weakly weighted auxiliary targets."""
        w = np.array([1.0] * 3 + [1.0 if self.use_torque else 0.05] * 3 + [0.25] * 3
                     + [0.05] * 3 + [0.05])
        return w / w.sum() * 13.0

    def train(self, X, Y, epochs=4, batch=8192, lr=2e-3, lr_end=5e-5, log=None, W=None):
        """Adam on weighted MSE of normalised targets; X, Y (and per-sample
        weights W) are device float32 tensors."""
        torch = self.torch
        n = X.shape[0]
        # one fused kernel per step (the foreach form launches several per step)
        opt = torch.optim.Adam([self.params], lr=lr, fused=self.params.is_cuda)
        steps = max(1, epochs * (n // batch))
        wts = torch.as_tensor(self.loss_weights(), dtype=torch.float32, device=self.device)
        k = 0
        t0 = time.time()
        for ep in range(epochs):
            perm = torch.randperm(n, device=self.device)
            # summed on the device: no host synchronisation inside an epoch
            tot = torch.zeros((), dtype=torch.float32, device=self.device)
            for i in range(0, n - batch + 1, batch):
                idx = perm[i:i + batch]
                xb, yb = X[idx].contiguous(), Y[idx].contiguous()
                y = self.forward(xb, train=True)
                diff = y - yb
                gy = (2.0 / (batch * 13)) * diff * wts
                if W is not None:
                    gy = gy * W[idx][:, None]
                self.params.grad = self.backward(gy.contiguous())
                for gparam in opt.param_groups:
                    gparam["lr"] = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * k / steps))
                opt.step()
                self._load()
                tot += (diff * diff).mean()
                k += 1
            loss = float(tot) / max(1, n // batch)
            self.history.append(loss)
            if log:
                log(f"surrogate epoch {ep}: mse {loss:.3e} ({time.time() - t0:.1f}s)")
        return self.history

    def force_error(self, c, q, F):
        """Relative RMS force error of the surrogate against F (FP64 plain path)."""
        torch = self.torch
        X = torch.as_tensor(self.features(c, q), device=self.device)
        Y = self.predict(X).double().cpu().numpy()
        env = envelope(np.linalg.norm(c, axis=1), self.r0, self.lam)[:, None]
        Fp = Y[:, 0:3] * self.scale[0] * env
        err = np.linalg.norm(Fp - F, axis=1)
        return float(np.sqrt((err ** 2).sum() / (np.linalg.norm(F, axis=1) ** 2).sum())), Fp

    def torque_error(self, c, q, T):
        """Relative RMS torque error of the surrogate against T (FP64 plain
        path, physical torque mode)."""
        torch = self.torch
        X = torch.as_tensor(self.features(c, q), device=self.device)
        Y = self.predict(X).double().cpu().numpy()
        env = envelope(np.linalg.norm(c, axis=1), self.r0, self.lam)[:, None]
        Tp = Y[:, 3:6] * self.scale[1] * env
        err = np.linalg.norm(Tp - T, axis=1)
        ref = (np.linalg.norm(T, axis=1) ** 2).sum()
        return float(np.sqrt((err ** 2).sum() / ref)) if ref > 0 else 0.0

    def sim_params(self, audit_n=64):
        return np.array([self.r_on, self.r0, self.lam, 1.0 / self.b, self.scale[0], self.scale[1],
                         self.scale[4], 1.0 if self.use_torque else 0.0, float(audit_n)])


def build_surrogate(engine, system, n_train=2_000_000, epochs=4, batch=8192, contact=None,
                    shell=8.0, seed=0, log=None, max_rows=65536, extra_poses=None, lr=2e-3):
    """Train a surrogate for `system`: active beyond contact + shell (the
    contact distance defaults to the sum of the hydrodynamic radii)."""
    import torch
    mob = system.mob
    r_contact = contact if contact is not None else float(mob.radius1 + mob.radius2)
    p = system.params
    s = Surrogate(engine, r_on=r_contact + shell, r_max=p.r_escape, b=p.r_start,
                  max_rows=max_rows, seed=seed)
    body = system.mol2.positions_array() - system.mol2.centroid()
    engine.set_ligand(body, system.mol2.charges_array())
    t0 = time.time()
    c, q, F, T, E = s.make_dataset(n_train, seed=seed, extra_poses=extra_poses)
    s.fit_scales(c, F, T, E)
    X = torch.as_tensor(s.features(c, q), device=engine.device)
    Y = torch.as_tensor(s.targets(c, F, T, E), device=engine.device)
    Wt = torch.as_tensor(s.sample_weights(c), device=engine.device)
    if log:
        log(f"surrogate data: {len(c)} poses labelled by the FP64 grid path in {time.time() - t0:.1f}s")
    s.train(X, Y, epochs=epochs, batch=batch, log=log, W=Wt, lr=lr)
    cv, qv, Fv, Tv, _ = s.make_dataset(20000, seed=seed + 1, near_frac=0.0)
    s.val_error, _ = s.force_error(cv, qv, Fv)
    s.val_torque_error = s.torque_error(cv, qv, Tv) if s.use_torque else None
    if log:
        log(f"surrogate validation force RMS relative error {s.val_error:.3e}"
            + (f", torque {s.val_torque_error:.3e}" if s.use_torque else ""))
    return s
