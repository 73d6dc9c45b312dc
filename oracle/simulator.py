"""Batch BD step restated in NumPy FP64 (oracle, test infrastructure only).

Follows GPUBatchSimulator.run (pystarc/simulation/gpu_batch_simulator.py:462-1594)
in the flattened-pairs mode, consuming ``numpy.random.default_rng(seed)`` in
the reference's order and shapes:

  start   normals (N,3)            :509      quaternion uniforms (N,3)  :513
  step    escape uniforms (n_esc)  :884      return quats (n_ret,3)     :899-901
          translation normals (N_sr,3) :1086-1088
          rotation normals (N_sr,3)    :1096-1098
          bridge uniforms (N_sr,P) only if any pair is above on both ends :1161, :1179-1181

Per-step physics: reaction check :855-871, escape/return :879-911, force
batches :931-943 (engine oracle), old pair distances :950-960, adaptive dt
:961-1067, HI D_parallel :1074-1084, Ermak-McCammon :1085-1101, overlap
rejection :1106-1128, Brownian bridge :1142-1252, bookkeeping :1253-1254,
early stop :1342-1356, final counts :1486-1489.
"""
from __future__ import annotations

import math

import numpy as np

from . import forces as F_
from .physics import (apply_rotation, bridge_crossing_prob, bridge_d_eff, em_translate,
                      overlap_rejected, quats_to_rotmats, return_prob_and_kb, shoemake)


class OracleSim:
    """State + one reference step at a time.

    ``system`` keys: lig_body (M,3), lig_q (M,), rec_pos (Nr,3), pairs
    [(rec_idx, lig_idx, cutoff)], n_needed, q_rec, q_lig, D_t, D_r, a_rec,
    a_lig, engine (oracle engine dict), debye.
    ``params`` keys: n_trajectories, seed, r_start, r_escape, max_steps,
    hydrodynamic_interactions, minimum_core_dt, max_dt, overlap_check,
    gpu_force_batch, kT_scale.
    """

    def __init__(self, system, params, torque_mode="reference"):
        self.s = system
        self.p = params
        self.torque_mode = torque_mode
        self.N = int(params["n_trajectories"])
        self.rng = np.random.default_rng(params["seed"])
        s = system
        self.body = np.asarray(s["lig_body"], dtype=np.float64)
        self.q_lig_atoms = np.asarray(s["lig_q"], dtype=np.float64)
        pairs = s["pairs"]
        self.rec_idx = [p[0] for p in pairs]
        self.lig_idx = [p[1] for p in pairs]
        self.cut = np.array([p[2] for p in pairs], dtype=np.float64)
        P = len(pairs)
        nn = s.get("n_needed", -1)
        self.n_needed = P if (nn < 0 or nn > P) else nn
        if not pairs:
            self.n_needed = 1
        rec_pos = np.asarray(s["rec_pos"], dtype=np.float64)
        self.rec_gho = rec_pos[self.rec_idx] if P else None
        self.lig_gho_body = self.body[self.lig_idx] if P else None
        self.D_t = float(s["D_t"])
        self.D_r = float(s["D_r"])
        self.r_b = float(params["r_start"])
        self.r_esc = float(params["r_escape"])
        self.debye = float(s.get("debye", 7.828))
        self.return_prob, self.k_b = return_prob_and_kb(
            self.r_b, self.r_esc, s["q_rec"], s["q_lig"], self.debye, self.D_t,
            use_hi=bool(params.get("hydrodynamic_interactions", False)),
            a0=s["a_rec"], a1=s["a_lig"], kT_scale=params.get("kT_scale", 1.0))
        # adaptive dt constants (:622-636)
        self.V0 = s["q_rec"] * s["q_lig"] / (4.0 * math.pi * 78.0 * 0.000142)
        self.has_force = abs(self.V0) > 1e-9
        self.min_core_dt = float(params.get("minimum_core_dt", 0.0))
        self.max_dt = float(params.get("max_dt", 0.0))
        # Brownian-bridge D_eff (:643-660)
        self.bb_D_eff = None
        if P:
            self.bb_D_eff = bridge_d_eff(self.D_t, float(s["a_rec"]), float(s["a_lig"]),
                                         np.linalg.norm(self.rec_gho, axis=1),
                                         np.linalg.norm(self.lig_gho_body, axis=1))
        # HI constants (:741-751)
        self.use_hi = bool(params.get("hydrodynamic_interactions", False))
        if self.use_hi:
            self.hi_Df = 1.0 / 0.243
            a0, a1 = float(s["a_rec"]), float(s["a_lig"])
            self.hi_a2 = 0.5 * (a0 ** 2 + a1 ** 2)
            pi6 = 6.0 * math.pi
            self.hi_ainv = (self.hi_Df / (pi6 * a0) + self.hi_Df / (pi6 * a1)) / self.hi_Df
        self.use_overlap = bool(params.get("overlap_check", True))
        self.rec_pos = rec_pos
        M = self.body.shape[0]
        cfg_b = int(params.get("gpu_force_batch", 0))
        self.force_batch_cfg = cfg_b if cfg_b > 0 else max(1000, int(4 * 1024 ** 3 / max(1, M * 150)))
        # initial state (:509-516)
        v = self.rng.standard_normal((self.N, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        self.pos = v * self.r_b
        self.q = shoemake(self.rng.random((self.N, 3)).astype(np.float64))
        self.status = np.zeros(self.N, dtype=np.int32)
        self.n_steps = np.zeros(self.N, dtype=np.int64)
        self.total_steps = 0
        self.n_overlap_rejected = 0
        self.step_index = 0
        self.last = {}

    # --- helpers -------------------------------------------------------
    def pair_dists(self, R, cen):
        g = np.einsum("nij,kj->nki", R, self.lig_gho_body) + cen[:, None, :]
        return np.linalg.norm(g - self.rec_gho[None, :, :], axis=2)

    def forces(self, R, cen):
        n = R.shape[0]
        Fo = np.zeros((n, 3))
        To = np.zeros((n, 3))
        fb = min(self.force_batch_cfg, n) if n else 1
        for b0 in range(0, n, fb):
            b1 = min(b0 + fb, n)
            lab = F_.place_ligand(R[b0:b1], cen[b0:b1], self.body)
            f, t, _ = F_.engine_forces(self.s["engine"], lab, self.q_lig_atoms,
                                       R[b0:b1], cen[b0:b1], torque_mode=self.torque_mode)
            Fo[b0:b1] = f
            To[b0:b1] = t
        return Fo, To

    def adaptive_dt(self, r):
        """dt per still-running trajectory (:975-1067, flattened mode)."""
        D_t = self.D_t
        in_outer = r > 1.1 * self.r_b
        dt_pair = (0.1 * 0.1 / 2.0) * r * r / D_t
        if self.has_force:
            e = np.exp(-r / self.debye)
            Fr0 = self.V0 * e * (1.0 / r ** 2 + 1.0 / (r * self.debye))
            Fr1 = -self.V0 * e / self.debye ** 2 - 2.0 * Fr0 / r
            DFr1 = np.abs(D_t * Fr1)
            with np.errstate(divide="ignore"):
                dt_force = np.where((DFr1 > 1e-15) & (r < 3.0 * self.debye),
                                    0.01 / np.where(DFr1 > 0, DFr1, 1.0), dt_pair)
        else:
            dt_force = dt_pair
        dist_b = np.maximum(r - self.r_b, 1e-3)
        dist_e = np.maximum(self.r_esc - r, 1e-3)
        dt_edge = np.minimum(dist_b, dist_e) ** 2 / (18.0 * D_t)
        dt_outer = np.minimum(np.minimum(dt_force, dt_edge), dt_pair)
        dt_inner = np.minimum(dt_force, dt_pair)
        dt = np.where(in_outer, dt_outer, dt_inner)
        if self.max_dt > 0:
            dt = np.minimum(dt, self.max_dt)
        if self.min_core_dt > 0:
            dt = np.maximum(dt, self.min_core_dt)
        return dt

    def diffusion(self, r):
        if self.use_hi:
            rm1 = 1.0 / np.maximum(r, 0.01)
            rm3 = rm1 ** 3
            Di = -2.0 * self.hi_Df * (2.0 * rm1 - (4.0 / 3.0) * self.hi_a2 * rm3) / (8.0 * math.pi)
            return np.maximum(self.hi_Df * self.hi_ainv + Di, 1e-6)
        return np.full(r.shape[0], self.D_t)

    # --- one reference step --------------------------------------------
    def step(self):
        """Advance one loop iteration; returns False when the loop would break."""
        rng = self.rng
        status, pos, q = self.status, self.pos, self.q
        running = status == 0
        if not running.any():
            return False
        run_idx = np.where(running)[0]
        pos_run = pos[run_idx]
        R = quats_to_rotmats(q[run_idx])
        P = len(self.cut)
        # reaction check (:855-871)
        if P:
            d = self.pair_dists(R, pos_run)
            reacted = (d < self.cut[None, :]).sum(axis=1) >= self.n_needed
        else:
            reacted = np.zeros(len(run_idx), dtype=bool)
        newly = run_idx[reacted]
        status[newly] = 1
        # escape / return (:879-911)
        r_mag = np.linalg.norm(pos_run, axis=1)
        at_esc = (r_mag >= self.r_esc) & ~reacted
        n_ret = n_esc = 0
        if at_esc.any():
            esc_idx = run_idx[at_esc]
            u = rng.random(int(at_esc.sum()))
            ret = u < self.return_prob
            if ret.any():
                ri = esc_idx[ret]
                ri = ri[status[ri] == 0]
                if len(ri):
                    dirs = pos[ri] / np.linalg.norm(pos[ri], axis=1, keepdims=True)
                    pos[ri] = dirs * self.r_b
                    q[ri] = shoemake(rng.random((len(ri), 3)).astype(np.float64))
                    n_ret = len(ri)
            if (~ret).any():
                te = esc_idx[~ret]
                te = te[status[te] == 0]
                status[te] = 2
                n_esc = len(te)
        still = status[run_idx] == 0
        pos_run = pos[run_idx]
        R = quats_to_rotmats(q[run_idx])
        r_mag = np.linalg.norm(pos_run, axis=1)
        sr_idx = run_idx[still]
        self.last = {"n_run": len(run_idx), "n_reacted_start": int(reacted.sum()),
                     "n_return": n_ret, "n_escape": n_esc, "n_sr": len(sr_idx),
                     "bridge": False, "n_bridge": 0}
        if len(sr_idx) == 0:
            self.step_index += 1
            return self._continue()
        R_sr = R[still]
        cen = pos_run[still]
        Fo, To = self.forces(R_sr, cen)
        self.last["F"] = Fo
        old_d = self.pair_dists(R_sr, cen) if P else None
        r_sr = r_mag[still]
        dt = self.adaptive_dt(r_sr)
        D = self.diffusion(r_sr)
        self.last["dt"] = dt
        n_sr = len(sr_idx)
        nt = rng.standard_normal((n_sr, 3))
        old_pos = pos[sr_idx].copy()
        old_q = q[sr_idx].copy()
        pos[sr_idx] = em_translate(old_pos, D, dt, Fo, nt)
        nr = rng.standard_normal((n_sr, 3))
        sig_r = np.sqrt(2.0 * self.D_r * dt)[:, None]
        omega = self.D_r * dt[:, None] * To + sig_r * nr
        q[sr_idx] = apply_rotation(q[sr_idx], omega)
        if self.use_overlap:
            inside = overlap_rejected(pos[sr_idx], self.rec_pos)
            if inside.any():
                ii = np.where(inside)[0]
                pos[sr_idx[ii]] = old_pos[ii]
                q[sr_idx[ii]] = old_q[ii]
                self.n_overlap_rejected += len(ii)
            self.last["n_overlap"] = int(inside.sum())
        if P:
            Rn = quats_to_rotmats(q[sr_idx])
            new_d = self.pair_dists(Rn, pos[sr_idx])
            x0 = old_d - self.cut[None, :]
            x1 = new_d - self.cut[None, :]
            both = (x0 > 0) & (x1 > 0)
            if both.any():
                if self.bb_D_eff is not None:
                    Ddt = self.bb_D_eff[None, :] * dt[:, None]
                else:
                    Ddt = self.D_t * dt[:, None]
                pc = bridge_crossing_prob(x0, x1, Ddt)
                u = rng.random(pc.shape)
                fired = (u < pc) | (new_d < self.cut[None, :])
                bb = fired.sum(axis=1) >= self.n_needed
                g = sr_idx[bb]
                g = g[status[g] == 0]
                status[g] = 1
                self.last["bridge"] = True
                self.last["n_bridge"] = len(g)
        self.n_steps[sr_idx] += 1
        self.total_steps += n_sr
        self.step_index += 1
        return self._continue()

    def _continue(self):
        done = int((self.status != 0).sum())
        if done >= int(self.N * 0.995):
            return False
        return self.step_index < int(self.p.get("max_steps", 10 ** 9))

    def run(self, max_loop=None):
        n = 0
        while self.step():
            n += 1
            if max_loop is not None and n >= max_loop:
                break
        return self.counts()

    def counts(self):
        return {"n_reacted": int((self.status == 1).sum()),
                "n_escaped": int((self.status == 2).sum()),
                "n_max_steps": int((self.status == 0).sum()),
                "total_steps": int(self.total_steps)}
