"""Batch force engine restated in NumPy FP64 (oracle, test infrastructure only).

Follows pystarc/forces/gpu_batch_engine.py:

* ``prepare_grid``   - _upload_grids :248-287 (3-spacing valid box)
* ``interp``         - the trilinear lambda inside batch_force_kernel :95-110
* ``kernel_forces``  - batch_force_kernel :65-123 (central differences at
                       +-sp/2, F = -q grad phi / -alpha q^2 grad phi_B)
* ``eval_batch``     - _eval_batch :289-391 (finest-first assignment, Yukawa
                       fallback for unassigned atoms, elec only)
* ``yukawa``         - _yukawa_forces_gpu :393-472 (radial-only gradient)
* ``born_reverse``   - _eval_born_reverse :607-657 (500 MB chunk skip)
* ``engine_forces``  - __call__ :531-605 (torque placeholder = 0)
* ``physical_torque``- lever arm from the centroid of charged atoms,
                       pystarc/forces/engine.py:144-183
"""
from __future__ import annotations

import math

import numpy as np

from .physics import EPS0


def prepare_grid(data, origin, spacing):
    """Grid record with the BC-safe box lo = o + 3 sp, hi = o + (n-4) sp (:267-273)."""
    data = np.asarray(data)
    sp = np.asarray(spacing, dtype=np.float64)
    origin = np.asarray(origin, dtype=np.float64)
    dims = np.array(data.shape)
    return {
        "data": data.astype(np.float64).ravel(),
        "origin": origin, "sp": sp, "inv_sp": 1.0 / sp,
        "n": tuple(int(d) for d in dims),
        "lo": origin + 3.0 * sp, "hi": origin + (dims - 1 - 3.0) * sp,
    }


def sort_finest_first(grids):
    """Sorted by delta[0,0] ascending (:259); input grids are (data, origin, spacing)."""
    return sorted(grids, key=lambda g: float(g["sp"][0]))


def interp(g, x, y, z):
    """Trilinear value; 0 outside cells [0, n-2] on any axis (:95-110)."""
    nx, ny, nz = g["n"]
    o, inv = g["origin"], g["inv_sp"]
    ix = (x - o[0]) * inv[0]
    iy = (y - o[1]) * inv[1]
    iz = (z - o[2]) * inv[2]
    i0 = np.floor(ix).astype(np.int64)
    j0 = np.floor(iy).astype(np.int64)
    k0 = np.floor(iz).astype(np.int64)
    ok = (i0 >= 0) & (i0 < nx - 1) & (j0 >= 0) & (j0 < ny - 1) & (k0 >= 0) & (k0 < nz - 1)
    i0s = np.where(ok, i0, 0)
    j0s = np.where(ok, j0, 0)
    k0s = np.where(ok, k0, 0)
    fx, fy, fz = ix - i0, iy - j0, iz - k0
    d = g["data"]

    def at(i, j, k):
        return d[(i * ny + j) * nz + k]

    v = (at(i0s, j0s, k0s) * (1 - fx) * (1 - fy) * (1 - fz)
         + at(i0s + 1, j0s, k0s) * fx * (1 - fy) * (1 - fz)
         + at(i0s, j0s + 1, k0s) * (1 - fx) * fy * (1 - fz)
         + at(i0s, j0s, k0s + 1) * (1 - fx) * (1 - fy) * fz
         + at(i0s + 1, j0s + 1, k0s) * fx * fy * (1 - fz)
         + at(i0s + 1, j0s, k0s + 1) * fx * (1 - fy) * fz
         + at(i0s, j0s + 1, k0s + 1) * (1 - fx) * fy * fz
         + at(i0s + 1, j0s + 1, k0s + 1) * fx * fy * fz)
    return np.where(ok, v, 0.0)


def kernel_forces(pos, q, g, alpha, is_born):
    """Per-(traj, atom) force and energy (:81-121).  pos (...,3), q (...)."""
    px, py, pz = pos[..., 0], pos[..., 1], pos[..., 2]
    h0, h1, h2 = g["sp"] * 0.5
    gx = (interp(g, px + h0, py, pz) - interp(g, px - h0, py, pz)) / (2 * h0)
    gy = (interp(g, px, py + h1, pz) - interp(g, px, py - h1, pz)) / (2 * h1)
    gz = (interp(g, px, py, pz + h2) - interp(g, px, py, pz - h2)) / (2 * h2)
    phi = interp(g, px, py, pz)
    coeff = -alpha * q * q if is_born else -q
    ecoef = alpha * q * q if is_born else q
    live = np.abs(q) >= 1e-9
    F = np.stack([coeff * gx, coeff * gy, coeff * gz], axis=-1) * live[..., None]
    E = ecoef * phi * live
    return F, E


def farfield_params(receptor_charge, debye, sdie=78.0, multipole=None):
    """Yukawa constants (:161-175): V_fac = Q/(4 pi sdie eps0)."""
    eps_s = sdie * EPS0
    ff = {
        "V_fac": receptor_charge / (4.0 * math.pi * eps_s) if eps_s > 0 else 0.0,
        "has_yukawa": abs(receptor_charge) > 1e-9 and debye > 0,
        "debye": debye, "dipole": None, "quad": None, "fpe": None,
    }
    if multipole is not None:
        ff["dipole"] = np.asarray(multipole["dipole"], dtype=np.float64)
        ff["quad"] = np.asarray(multipole["quadrupole"], dtype=np.float64)
        ff["fpe"] = 4.0 * math.pi * sdie * EPS0
    return ff


def multipole_moments(positions, charges):
    """Q, p = sum q r, Q_ab = 1/2 sum q (3 r_a r_b - r^2 delta) (multipole_farfield.py:86-108)."""
    positions = np.asarray(positions, dtype=np.float64)
    charges = np.asarray(charges, dtype=np.float64)
    r2 = np.sum(positions ** 2, axis=1)
    quad = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            quad[a, b] = np.sum(charges * (3.0 * positions[:, a] * positions[:, b]
                                           - (r2 if a == b else 0.0)))
    return {"Q": float(np.sum(charges)),
            "dipole": np.sum(charges[:, None] * positions, axis=0),
            "quadrupole": 0.5 * quad}


def yukawa(pos, charges, ff, mask=None):
    """Monopole (+dipole+quadrupole) far-field with radial-only gradient (:393-472).

    pos (B,M,3); charges (M,); returns per-trajectory sums F (B,3), E (B,).
    """
    lam = ff["debye"]
    r = np.linalg.norm(pos, axis=2)
    sr = np.maximum(r, 1.0)
    e = np.exp(-sr / lam)
    rhat = pos / sr[:, :, None]
    V = ff["V_fac"]
    phi = V * e / sr
    dphi = V * e * (-1.0 / sr ** 2 - 1.0 / (sr * lam))
    if ff["dipole"] is not None:
        p = ff["dipole"]
        Q = ff["quad"]
        fpe = ff["fpe"]
        if float(np.linalg.norm(p)) > 1e-9:
            pdr = np.sum(rhat * p[None, None, :], axis=2)
            phi = phi + pdr / (fpe * sr ** 2) * (1.0 + sr / lam) * e
            dphi = dphi + pdr / fpe * e * (-2.0 / sr ** 3 - 2.0 / (sr ** 2 * lam)
                                           - 1.0 / (sr * lam ** 2))
        if float(np.linalg.norm(Q)) > 1e-9:
            rQr = np.sum(rhat * np.einsum("ij,...j->...i", Q, rhat), axis=2)
            phi = phi + rQr / (fpe * sr ** 3) * (1.0 + sr / lam + sr ** 2 / (3.0 * lam ** 2)) * e
            dphi = dphi + rQr / fpe * e * (-3.0 / sr ** 4 - 3.0 / (sr ** 3 * lam)
                                           - 1.0 / (sr ** 2 * lam ** 2)
                                           - 1.0 / (3.0 * sr * lam ** 3))
    Fa = -charges[None, :, None] * (dphi[:, :, None] * rhat)
    Ea = charges[None, :] * phi
    if mask is not None:
        Fa = Fa * mask[:, :, None]
        Ea = Ea * mask
    return Fa.sum(axis=1), Ea.sum(axis=1), Fa, Ea


def eval_batch(pos, charges, grids, alpha, is_born, ff, per_atom=False):
    """Finest-first grid assignment + Yukawa fallback for the rest (:289-391)."""
    B, M, _ = pos.shape
    Fa = np.zeros((B, M, 3))
    Ea = np.zeros((B, M))
    if not grids:
        if not is_born and ff["has_yukawa"]:
            _, _, fa, ea = yukawa(pos, charges, ff)
            Fa += fa
            Ea += ea
        return (Fa.sum(axis=1), Ea.sum(axis=1), Fa, Ea) if per_atom else (Fa.sum(axis=1), Ea.sum(axis=1))
    assigned = np.zeros((B, M), dtype=bool)
    for g in grids:
        in_grid = np.all(pos > g["lo"], axis=2) & np.all(pos < g["hi"], axis=2)
        todo = in_grid & ~assigned
        assigned |= todo
        if not todo.any():
            continue
        q = np.where(todo, charges[None, :], 0.0)
        p = np.where(todo[:, :, None], pos, 0.0)
        f, e = kernel_forces(p, q, g, alpha, is_born)
        Fa += f
        Ea += e
    if not is_born and ff["has_yukawa"]:
        na = ~assigned
        if na.any():
            _, _, fa, ea = yukawa(pos, charges, ff, mask=na.astype(np.float64))
            Fa += fa
            Ea += ea
    if per_atom:
        return Fa.sum(axis=1), Ea.sum(axis=1), Fa, Ea
    return Fa.sum(axis=1), Ea.sum(axis=1)


def reverse_chunk_size(n_rec):
    """Trajectories per reverse-Born chunk: 500 MiB / (N_rec*3*8) (:627-629)."""
    return max(1, int(500 * 1024 * 1024 / (n_rec * 3 * 8)))


def born_reverse(R, cen, rec_pos, rec_q, lig_grid, alpha, chunk=None, per_atom=False):
    """Ligand Born grid at receptor atoms, Newton's third law (:607-657).

    The chunk is skipped when min|c| - max|r_rec| exceeds
    max(|lo_x|, |hi_x|) of the ligand grid (:633-639), exactly as the
    reference does; ``chunk`` defaults to the reference's 500 MB chunk.
    Returns the force on the ligand (B,3) (and optionally the per-receptor-atom
    lab-frame force on the receptor and the skip mask).
    """
    B = R.shape[0]
    out = np.zeros((B, 3))
    fa_all = np.zeros((B, rec_pos.shape[0], 3))
    radius = float(max(abs(lig_grid["lo"][0]), abs(lig_grid["hi"][0])))
    no_skip = chunk == 0            # chunk=0: every trajectory evaluated, no skip
    if no_skip:
        cs = max(B, 1)
    else:
        cs = reverse_chunk_size(rec_pos.shape[0]) if chunk is None else max(chunk, 1)
    cs = min(cs, B) if B > 0 else cs
    rec_max = float(np.linalg.norm(rec_pos, axis=1).max())
    for c0 in range(0, B, cs):
        c1 = min(c0 + cs, B)
        rc = np.linalg.norm(cen[c0:c1], axis=1)
        if not no_skip and float(rc.min()) - rec_max > radius:
            continue
        rel = rec_pos[None, :, :] - cen[c0:c1, None, :]
        RT = np.swapaxes(R[c0:c1], 1, 2)
        x = np.einsum("nij,nkj->nki", RT, rel)
        _, _, fa, _ = eval_batch(x, rec_q, [lig_grid], alpha, True,
                                 {"has_yukawa": False}, per_atom=True)
        fl = fa.sum(axis=1)
        out[c0:c1] = -np.einsum("nij,nj->ni", R[c0:c1], fl)
        fa_all[c0:c1] = np.einsum("nij,nkj->nki", R[c0:c1], fa)
    if per_atom:
        return out, fa_all
    return out


def place_ligand(R, cen, body):
    """lab = R b + c (gpu_batch_simulator.py:935-937)."""
    return np.einsum("nij,kj->nki", R, body) + cen[:, None, :]


def physical_torque(pos_lab, charges, Fa, R=None, cen=None, rec_pos=None, fa_rec=None):
    """T = sum (p_i - c_q) x F_i with c_q the centroid of charged atoms
    (engine.py:144-183); reverse-Born reaction forces -f_j act at receptor
    atom r_j with the same lever point."""
    chg = np.abs(charges) > 1e-9
    if chg.any():
        cq = pos_lab[:, chg, :].mean(axis=1)
    else:
        cq = pos_lab.mean(axis=1)
    T = np.cross(pos_lab - cq[:, None, :], Fa).sum(axis=1)
    if rec_pos is not None and fa_rec is not None:
        T -= np.cross(rec_pos[None, :, :] - cq[:, None, :], fa_rec).sum(axis=1)
    return T


def engine_forces(eng, lig_positions, lig_charges, R=None, cen=None, torque_mode="reference"):
    """F, T, E for every trajectory, __call__ :531-605.

    ``eng`` holds prepared grids ('elec', 'born', 'lig_born' lists already
    reduced as the constructor does :186-211), 'alpha', 'ff', 'rec_pos',
    'rec_q'.  torque_mode 'reference' returns zeros (:597-599).
    """
    B, M, _ = lig_positions.shape
    F = np.zeros((B, 3))
    E = np.zeros(B)
    Fa = np.zeros((B, M, 3))
    ff = eng["ff"]
    if eng["elec"]:
        f, e, fa, _ = eval_batch(lig_positions, lig_charges, eng["elec"], 0.0, False, ff, per_atom=True)
        F += f
        E += e
        Fa += fa
    elif ff["has_yukawa"]:
        f, e, fa, _ = yukawa(lig_positions, lig_charges, ff)
        F += f
        E += e
        Fa += fa
    if eng["born"]:
        f, e, fa, _ = eval_batch(lig_positions, lig_charges, eng["born"], eng["alpha"], True, ff, per_atom=True)
        F += f
        E += e
        Fa += fa
    fa_rec = None
    if eng.get("lig_born") and eng.get("rec_pos") is not None and R is not None \
            and cen is not None and eng["alpha"] > 1e-12:
        f2, fa_rec = born_reverse(R, cen, eng["rec_pos"], eng["rec_q"], eng["lig_born"][0],
                                  eng["alpha"], chunk=eng.get("chunk"), per_atom=True)
        F += f2
    if torque_mode == "reference":
        T = np.zeros((B, 3))
    else:
        T = physical_torque(lig_positions, lig_charges, Fa, R, cen,
                            eng.get("rec_pos"), fa_rec)
    return F, T, E


def build_engine(elec, born, alpha, receptor_charge, debye, sdie=78.0, lig_born=None,
                 rec_pos=None, rec_q=None, multipole=None, chunk=None):
    """Constructor logic of GPUBatchForceEngine (:139-246) on prepared grids."""
    ff = farfield_params(receptor_charge, debye, sdie, multipole)
    elec = sort_finest_first(list(elec or []))
    born = sort_finest_first(list(born or []))
    if ff["has_yukawa"] and len(elec) > 1:
        elec = elec[:1]
    if len(born) > 1:
        born = born[:1]
    lb = []
    if lig_born and rec_pos is not None and rec_q is not None:
        lb = sort_finest_first(list(lig_born))[:1]
    return {"elec": elec, "born": born, "lig_born": lb, "alpha": alpha, "ff": ff,
            "rec_pos": None if rec_pos is None else np.asarray(rec_pos, dtype=np.float64),
            "rec_q": None if rec_q is None else np.asarray(rec_q, dtype=np.float64),
            "chunk": chunk}
