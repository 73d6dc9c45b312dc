"""Per-trajectory restatement of the production step (oracle, test
infrastructure only).

The production kernels (refilling slot pool, Philox noise keyed by
(trajectory, its own step, stream)) make every trajectory's fate a function
of its id alone.  This oracle advances all trajectories in lock-step by their
own step counter with the same physics as the reference's flattened-pairs
step (pystarc/simulation/gpu_batch_simulator.py):

  start          b-sphere position from normals, Shoemake quaternion (:509-516)
  max steps      finish when the step counter reached max_steps (:1486)
  reaction       pair distances at the current pose (:855-871)
  escape/return  r >= r_esc: u < return_prob -> back to b with a new
                 orientation, else escaped (:879-911)
  force          oracle engine at the pose (:931-943)
  dt, D          adaptive dt (:961-1067), D_parallel with HI (:1074-1084)
  move           Ermak-McCammon + right-multiplied rotation (:1085-1101), or
                 the two-body RPY move (oracle/rpy12.py) in hydro mode "rpy12"
  overlap        centroid within 1.5 A of a receptor atom -> restore (:1106-1128)
  bridge         p = exp(-x0 x1 / (D_eff dt)) for pairs above on both ends;
                 fired = u < p or new distance < cutoff (:1142-1252)

Draws, per trajectory id t at its step n (oracle/philox.py):
  start   normal2(INIT, step 0xFFFFFFFF, sub 0, 1), uniform2(INIT, sub 2, 3)
  escape  uniform2(ESC, n, 0)[0];  return quaternion uniform2(RETQ, n, 0), (.., 1)[0]
  noise   normal2(NOISE, n, sub 0..2) -> (t0 t1) (t2 r0) (r1 r2); rpy12: sub 0..5 -> xi[0..11]
  bridge  uniform2(BRIDGE, n, sub p >> 1) -> pairs p, p + 1

The only coupling between trajectories in the device run is the batch-wide
"any pair above on both ends" gate of the bridge (:1161); a trajectory that
the gate skips and that reached its cutoffs is finished at the start of its
next step instead, with the same step count, unless that next step hits
max_steps first.  This oracle always evaluates the bridge.
"""
from __future__ import annotations

import math

import numpy as np

from . import forces as F_
from . import philox as PH
from . import physics as PHY
from . import rpy12
from .physics import apply_rotation, quats_to_rotmats, return_prob_and_kb, shoemake
from .simulator import OracleSim

INIT_STEP = 0xFFFFFFFF


def sm_tables(reactions, first_state=None):
    """State indices of the reference (gpu_batch_simulator.py:255-306) for
    ``reactions`` = [(state_before, state_after)]: first_state (else the first
    state_before) is 0, labels numbered in order of appearance, a label that
    is no reaction's state_before is terminal."""
    idx = {}

    def get(label):
        if label is None:
            return -1
        return idx.setdefault(label, len(idx))
    first = first_state if first_state is not None else reactions[0][0]
    get(first)
    before = np.array([get(b) for b, _ in reactions])
    after = np.array([get(a) for _, a in reactions])
    terminal = np.array([i not in set(before.tolist()) for i in range(len(idx))])
    return before, after, terminal


class ProductionOracle(OracleSim):
    """``system`` / ``params`` as OracleSim; hydro "scalar" or "rpy12".
    ``system["sm_reactions"]`` = [(state_before, state_after)] (pair r is
    reaction r's pair0) switches to state-machine reactions
    (gpu_batch_simulator.py:818-854 check, :1014-1060 dt refinement,
    :1185-1234 bridge)."""

    def __init__(self, system, params, torque_mode="reference", hydro="scalar",
                 kT_over_eta=1.0 / 0.243, observe=False, traj_base=0, outer=None):
        super().__init__(system, params, torque_mode)
        # LMZ outer propagator (oracle/outer.OuterModel) replacing the return
        # probability at r_esc: trajectories at r >= 1.1 b take an excursion
        self.outer = outer
        # random-stream id of trajectory t is traj_base + t (a shard of a run)
        self.traj_base = int(traj_base)
        self.observe = bool(observe)
        self.hydro = hydro
        self.Df = float(kT_over_eta)
        self.a0, self.a1 = float(system["a_rec"]), float(system["a_lig"])
        self.seed = int(params["seed"]) & ((1 << 64) - 1)
        self.max_steps = int(params.get("max_steps", 10 ** 9))
        N = self.N
        t = np.arange(N)
        z0, z1 = PH.normal2(self.seed, t + self.traj_base, INIT_STEP, PH.RNG_INIT, 0)
        z2, _ = PH.normal2(self.seed, t + self.traj_base, INIT_STEP, PH.RNG_INIT, 1)
        u0, u1 = PH.uniform2(self.seed, t + self.traj_base, INIT_STEP, PH.RNG_INIT, 2)
        u2, _ = PH.uniform2(self.seed, t + self.traj_base, INIT_STEP, PH.RNG_INIT, 3)
        v = np.stack([z0, z1, z2], axis=1)
        self.pos = v / np.linalg.norm(v, axis=1, keepdims=True) * self.r_b
        self.q = _shoemake(u0, u1, u2)
        self.status = np.zeros(N, dtype=np.int32)
        self.n_steps = np.zeros(N, dtype=np.int64)
        self.out_steps = np.full(N, -1, dtype=np.int64)
        self.total_steps = 0
        self.n_returns = 0
        self.n_overlap_rejected = 0
        self.sm = None
        if system.get("sm_reactions"):
            b, a, term = sm_tables(system["sm_reactions"], system.get("first_state"))
            self.sm = {"before": b, "after": a, "terminal": term,
                       "min_core_rxn_dt": float(params.get("minimum_core_reaction_dt", 0.0))}
            self.sm_state = np.zeros(N, dtype=np.int64)
            nr = len(b)
            self.sm_fired = np.zeros(nr, dtype=np.int64)
            self.sm_fired_bridge = np.zeros(nr, dtype=np.int64)
            self.sm_sets = [set() for _ in range(nr)]
        if self.observe:
            # the reference's sim_data (gpu_batch_simulator.py:576-620,
            # :1255-1340), per trajectory and run-wide; step_at_min is the
            # trajectory's own step
            P = len(self.cut)
            self.obs_traj = np.zeros((N, 27))
            self.obs_traj[:, 0:3] = self.pos
            self.obs_traj[:, 3:7] = self.q
            self.obs_traj[:, 7] = 1e30
            self.obs_traj[:, 9:12] = self.pos
            self.obs_traj[:, 12:16] = self.q
            # on the b-sphere by construction: exactly b, so the milestone at
            # b is not crossed by the rounding of |pos|
            self.prev_r = np.full(N, self.r_b)
            self.h_rad = np.zeros(200, dtype=np.int64)
            self.h_ang = np.zeros((36, 72), dtype=np.int64)
            self.h_ms_out = np.zeros(11, dtype=np.int64)
            self.h_ms_in = np.zeros(11, dtype=np.int64)
            self.h_trans = np.zeros((50, 50), dtype=np.int64)
            self.h_contact = np.zeros(max(P, 1), dtype=np.int64)
            self.h_contact_steps = 0

    def _finish(self, idx, code, steps, bridge=0):
        self.status[idx] = code
        self.out_steps[idx] = steps
        if self.observe and len(idx):
            self._record_end(idx, bridge)

    def _record_end(self, idx, bridge=0):
        """Trajectory records of idx as they stand (end pose, pairs within)."""
        t = self.obs_traj
        t[idx, 18] = bridge
        t[idx, 19:22] = self.pos[idx]
        t[idx, 22:26] = self.q[idx]
        if len(self.cut):
            d = self.pair_dists(quats_to_rotmats(self.q[idx]), self.pos[idx])
            t[idx, 26] = (d < self.cut[None, :]).sum(axis=1)

    @staticmethod
    def _bin(r, rmax, nb):
        return np.clip(np.floor(r / (rmax / nb)).astype(np.int64), 0, nb - 1)

    def observables(self):
        """Records of every trajectory; running ones as they stand."""
        run = np.where(self.status == 0)[0]
        if len(run):
            self._record_end(run)
        return self.obs_traj

    def step(self):
        st = self.status
        run = np.where(st == 0)[0]
        if len(run) == 0:
            return False
        ns = self.n_steps[run]
        # max steps
        mx = ns >= self.max_steps
        self._finish(run[mx], 3, ns[mx])
        run, ns = run[~mx], ns[~mx]
        P = len(self.cut)
        fired = np.zeros(len(run), dtype=bool)
        if P and len(run) and self.sm is not None:
            R = quats_to_rotmats(self.q[run])
            d = self.pair_dists(R, self.pos[run])
            rx = self._sm_lowest(run, d < self.cut[None, :])
            fired = rx >= 0
            term = self._sm_fire(run[fired], rx[fired], bridge=False)
            done = np.zeros(len(run), dtype=bool)
            done[np.where(fired)[0][term]] = True
            self._finish(run[done], 1, ns[done])
            run, ns, fired = run[~done], ns[~done], fired[~done]
        elif P and len(run):
            R = quats_to_rotmats(self.q[run])
            d = self.pair_dists(R, self.pos[run])
            re = (d < self.cut[None, :]).sum(axis=1) >= self.n_needed
            self._finish(run[re], 1, ns[re])
            run, ns = run[~re], ns[~re]
            fired = np.zeros(len(run), dtype=bool)
        # escape / return (a trajectory that fired a reaction this step is not
        # tested, :879-884)
        r = np.linalg.norm(self.pos[run], axis=1)
        if self.outer is not None:
            from .outer import QB_FACTOR, Draws, propagate
            at = (r >= QB_FACTOR * self.r_b) & ~fired
            keep = np.ones(len(run), dtype=bool)
            for k in np.where(at)[0]:
                i = run[k]
                dr = Draws(self.seed, int(i) + self.traj_base, int(ns[k]))
                back, p, qq = propagate(self.outer, dr, self.pos[i], self.q[i])
                if back:
                    self.pos[i], self.q[i] = p, qq
                    self.n_returns += 1
                    if self.observe:
                        self.obs_traj[i, 17] += 1
                else:
                    self._finish(np.array([i]), 2, np.array([ns[k]]))
                    keep[k] = False
            run, ns = run[keep], ns[keep]
            at = np.zeros(len(run), dtype=bool)
        else:
            at = (r >= self.r_esc) & ~fired
        if at.any():
            ia, na = run[at], ns[at]
            u, _ = PH.uniform2(self.seed, ia + self.traj_base, na, PH.RNG_ESC, 0)
            ret = u < self.return_prob
            ir, nr_ = ia[ret], na[ret]
            if len(ir):
                p = self.pos[ir]
                self.pos[ir] = p / np.linalg.norm(p, axis=1, keepdims=True) * self.r_b
                a0, a1 = PH.uniform2(self.seed, ir + self.traj_base, nr_, PH.RNG_RETQ, 0)
                a2, _ = PH.uniform2(self.seed, ir + self.traj_base, nr_, PH.RNG_RETQ, 1)
                self.q[ir] = _shoemake(a0, a1, a2)
                self.n_returns += len(ir)
                if self.observe:
                    self.obs_traj[ir, 17] += 1
            self._finish(ia[~ret], 2, na[~ret])
            keep = np.ones(len(run), dtype=bool)
            keep[np.where(at)[0][~ret]] = False
            run, ns = run[keep], ns[keep]
        n = len(run)
        if n == 0:
            return True
        pos0 = self.pos[run].copy()
        q0 = self.q[run].copy()
        R = quats_to_rotmats(q0)
        Fo, To = self.forces(R, pos0)
        old_d = self.pair_dists(R, pos0) if P else None
        r = np.linalg.norm(pos0, axis=1)
        if self.sm is not None:
            dt = self._sm_dt(r, run, old_d)
        else:
            dt = self.adaptive_dt(r)
        if self.hydro == "rpy12":
            xi = np.empty((n, 12))
            for k in range(6):
                xi[:, 2 * k], xi[:, 2 * k + 1] = PH.normal2(self.seed, run + self.traj_base, ns, PH.RNG_NOISE, k)
            tm = 0 if self.torque_mode == "reference" else 1
            FR, TR = rpy12.receptor_load_v(pos0, Fo, To, tm)
            dX = rpy12.displacement_v(pos0, Fo, To, FR, TR, xi, dt, self.Df, self.a0, self.a1)
            newp, newq = rpy12.move_v(pos0, q0, dX)
        else:
            D = self.diffusion(r)
            a, b = PH.normal2(self.seed, run + self.traj_base, ns, PH.RNG_NOISE, 0)
            c, e = PH.normal2(self.seed, run + self.traj_base, ns, PH.RNG_NOISE, 1)
            f, g = PH.normal2(self.seed, run + self.traj_base, ns, PH.RNG_NOISE, 2)
            nt = np.stack([a, b, c], axis=1)
            nrot = np.stack([e, f, g], axis=1)
            newp = PHY.em_translate(pos0, D, dt, Fo, nt)
            sr = np.sqrt(2.0 * self.D_r * dt)[:, None]
            newq = apply_rotation(q0, self.D_r * dt[:, None] * To + sr * nrot)
        if self.use_overlap:
            bad = PHY.overlap_rejected(newp, self.rec_pos)
            newp[bad] = pos0[bad]
            newq[bad] = q0[bad]
            self.n_overlap_rejected += int(bad.sum())
        self.pos[run] = newp
        self.q[run] = newq
        self.total_steps += n
        if self.observe:
            self._observe_move(run, ns, newp, newq, dt, old_d)
        if P:
            new_d = self.pair_dists(quats_to_rotmats(newq), newp)
            x0 = old_d - self.cut[None, :]
            x1 = new_d - self.cut[None, :]
            u = np.empty((n, P))
            for p in range(0, P, 2):
                ua, ub = PH.uniform2(self.seed, run + self.traj_base, ns, PH.RNG_BRIDGE, p >> 1)
                u[:, p] = ua
                if p + 1 < P:
                    u[:, p + 1] = ub
            Dv = self.bb_D_eff if self.bb_D_eff is not None else np.full(P, self.D_t)
            pc = PHY.bridge_crossing_prob(x0, x1, Dv[None, :] * dt[:, None])
            pf = (u < pc) | (x1 < 0.0)
            if self.sm is not None:
                rx = self._sm_lowest(run, pf)
                f = rx >= 0
                term = self._sm_fire(run[f], rx[f], bridge=True)
                bb = np.zeros(n, dtype=bool)
                bb[np.where(f)[0][term]] = True
            else:
                bb = pf.sum(axis=1) >= self.n_needed
            self._finish(run[bb], 1, ns[bb] + 1, bridge=1)
        self.n_steps[run] = ns + 1
        return True

    # --- state-machine reactions ------------------------------------------
    def _sm_lowest(self, run, mask):
        """Lowest reaction r with mask[:, r] and state == before[r], else -1."""
        ok = mask & (self.sm_state[run][:, None] == self.sm["before"][None, :])
        return np.where(ok.any(axis=1), ok.argmax(axis=1), -1)

    def _sm_fire(self, idx, rx, bridge):
        """New states; counts; returns which moved to a terminal state."""
        after = self.sm["after"][rx]
        self.sm_state[idx] = after
        for t, r in zip(idx.tolist(), rx.tolist()):
            (self.sm_fired_bridge if bridge else self.sm_fired)[r] += 1
            self.sm_sets[r].add(t)
        return np.array([a >= 0 and bool(self.sm["terminal"][a]) for a in after], dtype=bool)

    def _sm_dt(self, r, run, old_d):
        """dt refined near a reachable reaction surface, before cap and floor."""
        saved = (self.max_dt, self.min_core_dt)
        self.max_dt = self.min_core_dt = 0.0
        dt = self.adaptive_dt(r)
        self.max_dt, self.min_core_dt = saved
        reach = self.sm_state[run][:, None] == self.sm["before"][None, :]
        coord = np.where(reach, np.maximum(old_d - self.cut[None, :], 0.0), 1e30).min(axis=1)
        need = 12.0 * np.sqrt(2.0 * self.D_t * dt) > coord
        crit = coord * coord / (288.0 * self.D_t)
        if self.sm["min_core_rxn_dt"] > 0:
            crit = np.maximum(crit, self.sm["min_core_rxn_dt"])
        dt = np.where(need, np.minimum(dt, crit), dt)
        if self.max_dt > 0:
            dt = np.minimum(dt, self.max_dt)
        if self.min_core_dt > 0:
            dt = np.maximum(dt, self.min_core_dt)
        return dt

    def _observe_move(self, run, ns, newp, newq, dt, old_d):
        t = self.obs_traj
        rn = np.linalg.norm(newp, axis=1)
        t[run, 16] += dt
        closer = rn < t[run, 7]
        ic = run[closer]
        t[ic, 7] = rn[closer]
        t[ic, 8] = ns[closer]
        t[ic, 9:12] = newp[closer]
        t[ic, 12:16] = newq[closer]
        rmax = 1.2 * self.r_esc
        np.add.at(self.h_rad, self._bin(rn, rmax, 200), 1)
        rs = np.maximum(rn, 1e-30)
        th = np.arccos(np.clip(newp[:, 2] / rs, -1.0, 1.0))
        ph = np.arctan2(newp[:, 1], newp[:, 0]) + np.pi
        ti = np.clip((th / np.pi * 36).astype(np.int64), 0, 35)
        pj = np.clip((ph / (2.0 * np.pi) * 72).astype(np.int64), 0, 71)
        np.add.at(self.h_ang, (ti, pj), 1)
        rp = self.prev_r[run]
        for m in range(11):
            mr = self.r_b + (self.r_esc - self.r_b) * m / 10
            self.h_ms_out[m] += int(((rp < mr) & (rn >= mr)).sum())
            self.h_ms_in[m] += int(((rp >= mr) & (rn < mr)).sum())
        np.add.at(self.h_trans, (self._bin(rp, rmax, 50), self._bin(rn, rmax, 50)), 1)
        if len(self.cut):
            self.h_contact += (old_d < self.cut[None, :]).sum(axis=0)
        self.h_contact_steps += len(run)
        self.prev_r[run] = rn

    def run(self, max_loop=None):
        k = 0
        while self.step():
            k += 1
            if max_loop is not None and k >= max_loop:
                break
        return self.counts()

    def counts(self):
        return {"n_reacted": int((self.status == 1).sum()),
                "n_escaped": int((self.status == 2).sum()),
                "n_max_steps": int((self.status == 3).sum()),
                "total_steps": int(self.total_steps)}


def _shoemake(u0, u1, u2):
    return shoemake(np.stack([u0, u1, u2], axis=1))
