"""Synthetic benchmark systems for the batch BD step.

Five workloads are defined on the path (two charged spheres; a 3,220-atom
receptor with an 18-atom ligand; a 1,700 / 1,400-atom protein pair; the same
pair or receptor with a learned surrogate).  This module builds the first
three as the reference's own objects would carry them: receptor / ligand
``Molecule``s, a ``MobilityTensor``, a ``PathwaySet`` of contact pairs,
``NAMParameters`` with the pipeline's extra attributes, the keyword arguments
of ``GPUBatchForceEngine`` and the grid specifications.

Atom clouds: a 2.8 A Poisson-disc minimum separation cannot hold 3,220 atoms
in a 22x20x18 A ellipsoid (nor 1,700 / 1,400 atoms in 17 / 15 A spheres), so
atoms are drawn from a seeded jittered cubic lattice filling the stated
volume; the minimum separation that results is recorded in ``meta``.
Grids are evaluated on the device by ``pst_make_grid``: a Debye-Hueckel sum
(eps_out 78, debye length 7.86 A at 0.15 M) for potentials and an analytic
Born field (eps_in 4, eps_out 78) for desolvation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from .constants import ALPHA_DEFAULT
from .forces.electrostatic.grid_force import DXGrid
from .forces.multipole_farfield import MultipoleExpansion
from .hydrodynamics.rotne_prager import MobilityTensor
from .pathways.reaction_interface import PathwaySet, ReactionInterface
from .simulation.nam_simulator import NAMParameters
from .structures.molecules import ContactPair, Molecule, ReactionCriteria

DEBYE_015M = 7.86


@dataclass
class GridSpec:
    role: str              # "elec" | "born" | "lig_born"
    kind: int              # 0 = Debye-Hueckel potential, 1 = Born field
    shape: tuple
    origin: np.ndarray
    spacing: float
    dtype: str             # "float32" | "float64"
    atoms: np.ndarray      # (n,5): x, y, z, charge, radius
    debye: float = DEBYE_015M
    eps_out: float = 78.0
    eps_in: float = 4.0


@dataclass
class BDSystem:
    name: str
    mol1: Molecule
    mol2: Molecule
    mob: MobilityTensor
    pathway_set: PathwaySet
    params: NAMParameters
    engine_kwargs: Dict
    grid_specs: List[GridSpec]
    meta: Dict = field(default_factory=dict)


def lattice_cloud(n, semi_axes, rng, jitter=0.2):
    """n points in an ellipsoid from a jittered cubic lattice, centroid at 0."""
    a = np.asarray(semi_axes, dtype=float)
    vol = 4.0 / 3.0 * math.pi * a.prod()
    s = (vol / (n * 1.12)) ** (1.0 / 3.0)
    for _ in range(60):
        ext = [np.arange(-ai, ai + s, s) for ai in a]
        g = np.stack(np.meshgrid(*ext, indexing="ij"), axis=-1).reshape(-1, 3)
        g = g - g.mean(axis=0)
        inside = ((g / a) ** 2).sum(axis=1) <= 1.0
        if inside.sum() >= n:
            break
        s *= 0.97
    sites = g[inside]
    pick = rng.choice(len(sites), size=n, replace=False)
    pts = sites[pick] + rng.uniform(-jitter, jitter, size=(n, 3))
    pts -= pts.mean(axis=0)
    return pts, s


def fcc_ball(n, nn, rng, jitter=0.03):
    """n points of a face-centred cubic lattice with nearest-neighbour
    distance nn, the n sites closest to the centre (every interior site
    filled), jittered by U(-jitter, jitter) per coordinate, centroid at 0.
    The largest empty sphere of the lattice is its octahedral hole, nn / sqrt 2
    (1.45 A at nn = 2.05), so with a small jitter no ligand centroid fits
    inside the ball under the 1.5 A overlap rule."""
    a = nn * math.sqrt(2.0)
    R = (n / (4.0 / a ** 3) * 3.0 / (4.0 * math.pi)) ** (1.0 / 3.0) + 2.0 * a
    m = int(math.ceil(R / a)) + 1
    ax = np.arange(-m, m + 1) * a
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), axis=-1).reshape(-1, 3)
    basis = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]) * a
    pts = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    order = np.argsort(np.linalg.norm(pts, axis=1), kind="stable")[:n]
    p = pts[order] + rng.uniform(-jitter, jitter, size=(n, 3))
    return p - p.mean(axis=0)


def charges_with_net(n, sigma, net, rng):
    q = rng.normal(0.0, sigma, size=n)
    return q + (net - q.sum()) / n


def min_separation(p):
    """Smallest pairwise distance (cell-free O(n^2) in blocks)."""
    best = np.inf
    for i in range(0, len(p), 512):
        d = np.linalg.norm(p[i:i + 512, None, :] - p[None, :, :], axis=2)
        for k in range(d.shape[0]):
            d[k, i + k] = np.inf
        best = min(best, float(d.min()))
    return best


def _atoms5(pos, q, rad):
    return np.column_stack([pos, q, rad]).astype(np.float64)


def _cube_origin(n, h):
    return np.full(3, -0.5 * (n - 1) * h)


def _params(n_traj, b, seed, max_steps, overlap, hi=False, min_core_dt=0.2):
    p = NAMParameters(n_trajectories=n_traj, r_start=b, r_escape=2.0 * b, seed=seed,
                      max_steps=max_steps, hydrodynamic_interactions=hi,
                      minimum_core_dt=min_core_dt)
    p.max_dt = 0.0
    p._overlap_check = overlap
    p._kT_scale = 1.0
    p.gpu_force_batch = 0
    return p


def two_spheres(n_traj=10_000, seed=11111113, grid_n=97, grid_h=0.5, max_steps=1_000_000,
                debye=DEBYE_015M):
    """Config 1: +1 e sphere (R 10 A) and -1 e sphere (R 2 A), reaction at
    centre distance < 12.5 A, b = 30, r_esc = 60, FP64 grid of the analytic
    screened Coulomb field, 0.2 ps dt floor.  Each molecule carries a charge-free
    ghost atom at its centre that defines the reaction pair."""
    z = np.zeros((2, 3))
    mol1 = Molecule.from_arrays(z, [1.0, 0.0], [10.0, 0.0], name="receptor")
    mol2 = Molecule.from_arrays(z, [-1.0, 0.0], [2.0, 0.0], name="ligand")
    crit = ReactionCriteria(name="contact", pairs=[ContactPair(1, 1, 12.5)], n_needed=1)
    ps = PathwaySet([ReactionInterface("rxn", crit)])
    mob = MobilityTensor.from_radii(10.0, 2.0)
    params = _params(n_traj, 30.0, seed, max_steps, overlap=False)
    rec_atoms = _atoms5(np.zeros((1, 3)), [1.0], [0.0])
    specs = [GridSpec("elec", 0, (grid_n,) * 3, _cube_origin(grid_n, grid_h), grid_h, "float64",
                      rec_atoms, debye=debye)]
    mp = MultipoleExpansion(mol1.positions_array(), mol1.charges_array(), debye, 78.0)
    kw = dict(alpha=0.0, receptor_charge=1.0, debye_length=debye, sdie=78.0,
              multipole_expansion=mp)
    return BDSystem("two_spheres", mol1, mol2, mob, ps, params, kw, specs,
                    {"config": "two charged spheres", "a_react": 12.5})


def _contact_pairs(rec_pos, lig_pos, n_pairs, cutoff, gap=3.0):
    """Contact pairs of a docked pose, as reaction criteria are taken from a
    complex structure: the ligand (identity orientation) sits on the receptor's
    +z face with its lowest atom `gap` A above the receptor's highest atom; the
    n_pairs closest receptor/ligand atom pairs across that interface (each atom
    used once) become the pairs."""
    zc = rec_pos[:, 2].max() - lig_pos[:, 2].min() + gap
    docked = lig_pos + np.array([0.0, 0.0, zc])
    rec_face = np.argsort(-rec_pos[:, 2])[: max(4 * n_pairs, 20)]
    lig_face = np.argsort(lig_pos[:, 2])[: max(4 * n_pairs, 20)]
    d = np.linalg.norm(rec_pos[rec_face][:, None, :] - docked[lig_face][None, :, :], axis=2)
    pairs, used_r, used_l = [], set(), set()
    for flat in np.argsort(d, axis=None):
        a, b = np.unravel_index(flat, d.shape)
        i, j = int(rec_face[a]), int(lig_face[b])
        if i in used_r or j in used_l:
            continue
        pairs.append(ContactPair(i, j, cutoff))
        used_r.add(i)
        used_l.add(j)
        if len(pairs) == n_pairs:
            break
    return pairs


def protein_ligand(n_traj=1_000_000, seed=2012, geom_seed=2002, n_rec=3220,
                   rec_axes=(22.0, 20.0, 18.0), rec_grid=(161, 0.5), lig_grid=(65, 0.4),
                   max_steps=1_000_000):
    """Config 2: 3,220-atom receptor (net +6 e) and 18-atom planar ligand
    (net +1 e); receptor potential + Born grids 161^3 at 0.5 A (FP32), ligand
    Born grid 65^3 at 0.4 A; 2 of 3 contact pairs < 4.0 A; b = 40."""
    rng = np.random.default_rng(geom_seed)
    rp, s = lattice_cloud(n_rec, rec_axes, rng)
    rq = charges_with_net(n_rec, 0.35, 6.0, rng)
    rr = rng.uniform(1.4, 2.0, n_rec)
    xs, ys = np.meshgrid(np.linspace(-3.0, 3.0, 6), np.linspace(-2.0, 2.0, 3), indexing="ij")
    lp = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(18)]) + rng.uniform(-0.1, 0.1, (18, 3))
    lp -= lp.mean(axis=0)
    lq = charges_with_net(18, 0.3, 1.0, rng)
    lr = rng.uniform(1.4, 2.0, 18)
    mol1 = Molecule.from_arrays(rp, rq, rr, name="receptor")
    mol2 = Molecule.from_arrays(lp, lq, lr, name="ligand")
    pairs = _contact_pairs(rp, lp, 3, 4.0)
    ps = PathwaySet([ReactionInterface("rxn", ReactionCriteria("contact", pairs, n_needed=2))])
    a_rec = float(np.prod(rec_axes) ** (1.0 / 3.0))
    mob = MobilityTensor.from_radii(a_rec, 3.5)
    params = _params(n_traj, 40.0, seed, max_steps, overlap=True)
    ra = _atoms5(rp, rq, rr)
    la = _atoms5(lp, lq, lr)
    gn, gh = rec_grid
    ln, lh = lig_grid
    specs = [GridSpec("elec", 0, (gn,) * 3, _cube_origin(gn, gh), gh, "float32", ra),
             GridSpec("born", 1, (gn,) * 3, _cube_origin(gn, gh), gh, "float32", ra),
             GridSpec("lig_born", 1, (ln,) * 3, _cube_origin(ln, lh), lh, "float32", la)]
    mp = MultipoleExpansion(rp, rq, DEBYE_015M, 78.0)
    kw = dict(alpha=ALPHA_DEFAULT, receptor_charge=float(rq.sum()), debye_length=DEBYE_015M,
              sdie=78.0, rec_positions=rp, rec_charges=rq, multipole_expansion=mp)
    return BDSystem("protein_ligand", mol1, mol2, mob, ps, params, kw, specs,
                    {"lattice_spacing": s, "geom_seed": geom_seed})


def protein_protein(n_traj=10_000_000, seed=3013, n_rec=1700, n_lig=1400, r_rec=17.0, r_lig=15.0,
                    grid=(193, 0.5), max_steps=1_000_000, geom_seeds=(3003, 3004)):
    """Config 3: 1,700 / 1,400-atom globules (radius 17 / 15 A, net +2 / -6 e);
    receptor potential + Born and ligand Born grids 193^3 at 0.5 A (FP32);
    bidirectional Born; scalar-HI with hydrodynamic radii 17 / 15 A; 2 of 5
    contact pairs < 5.5 A; b = 60, r_esc = 120."""
    r1 = np.random.default_rng(geom_seeds[0])
    r2 = np.random.default_rng(geom_seeds[1])
    rp, s1 = lattice_cloud(n_rec, (r_rec,) * 3, r1)
    rq = charges_with_net(n_rec, 0.35, 2.0, r1)
    rr = r1.uniform(1.4, 2.0, n_rec)
    lp, s2 = lattice_cloud(n_lig, (r_lig,) * 3, r2)
    lq = charges_with_net(n_lig, 0.35, -6.0, r2)
    lr = r2.uniform(1.4, 2.0, n_lig)
    mol1 = Molecule.from_arrays(rp, rq, rr, name="receptor")
    mol2 = Molecule.from_arrays(lp, lq, lr, name="ligand")
    pairs = _contact_pairs(rp, lp, 5, 5.5)
    ps = PathwaySet([ReactionInterface("rxn", ReactionCriteria("contact", pairs, n_needed=2))])
    mob = MobilityTensor.from_radii(r_rec, r_lig)
    params = _params(n_traj, 60.0, seed, max_steps, overlap=True, hi=True)
    gn, gh = grid
    ra = _atoms5(rp, rq, rr)
    la = _atoms5(lp, lq, lr)
    specs = [GridSpec("elec", 0, (gn,) * 3, _cube_origin(gn, gh), gh, "float32", ra),
             GridSpec("born", 1, (gn,) * 3, _cube_origin(gn, gh), gh, "float32", ra),
             GridSpec("lig_born", 1, (gn,) * 3, _cube_origin(gn, gh), gh, "float32", la)]
    mp = MultipoleExpansion(rp, rq, DEBYE_015M, 78.0)
    kw = dict(alpha=ALPHA_DEFAULT, receptor_charge=float(rq.sum()), debye_length=DEBYE_015M,
              sdie=78.0, rec_positions=rp, rec_charges=rq, multipole_expansion=mp)
    return BDSystem("protein_protein", mol1, mol2, mob, ps, params, kw, specs,
                    {"lattice_spacing": (s1, s2), "geom_seeds": geom_seeds})


def build_grids(system: BDSystem, ctx) -> Dict[str, list]:
    """Evaluate every grid of ``system`` on the device; returns role -> [DXGrid]."""
    from ._lib import ptr
    out: Dict[str, list] = {"elec": [], "born": [], "lig_born": []}
    for gs in system.grid_specs:
        arr = np.zeros(gs.shape, dtype=np.float32 if gs.dtype == "float32" else np.float64)
        origin = np.ascontiguousarray(gs.origin, dtype=np.float64)
        sp = np.full(3, gs.spacing, dtype=np.float64)
        atoms = np.ascontiguousarray(gs.atoms, dtype=np.float64)
        rc = ctx.lib.pst_make_grid(ctx.h, gs.kind, *gs.shape, ptr(origin), ptr(sp), ptr(atoms),
                                   len(atoms), gs.debye, gs.eps_out, gs.eps_in,
                                   0 if gs.dtype == "float32" else 1, ptr(arr))
        ctx.check(rc, "pst_make_grid")
        out[gs.role].append(DXGrid(origin, np.diag(sp), arr))
    return out


def make_engine(system: BDSystem, grids: Dict[str, list], **kw):
    from .forces.gpu_batch_engine import GPUBatchForceEngine
    args = dict(system.engine_kwargs)
    args.update(kw)
    return GPUBatchForceEngine(grids["elec"], grids["born"], lig_born_grids=grids["lig_born"] or None,
                               **args)


def from_pqr(receptor_pqr, ligand_pqr, rxns_xml, *, b=None, n_traj=100_000, seed=1,
             grid=(193, 0.5), lig_grid=None, a_rec=None, a_lig=None, debye=DEBYE_015M,
             max_steps=1_000_000, hi=True, overlap=True, dtype="float32", name=None,
             alpha=ALPHA_DEFAULT, min_core_dt=0.2, r_escape=None, point_charges=False):
    """A batch system from the reference's input files: receptor and ligand
    PQR (pystarc/structures/pqr_io.py) and rxns.xml with 1-based atom indices
    (pystarc/pipeline/geometry.py:321-389).  Both molecules are centred on
    their centroids, as the pipeline centres them before the BD stage
    (prepare_bd_surface.py:260-295).  APBS is not part of this build: the
    receptor potential / Born grids and the ligand Born grid are the
    Debye-Hueckel and Born sums of the PQR charges and radii on cubic grids
    (``grid`` / ``lig_grid`` = (nodes, spacing)).  Hydrodynamic radii default
    to each molecule's bounding radius (farthest atom centre + its radius);
    b defaults to 1.1 x (a_rec + a_lig + 15 A); r_escape to 2 b as the
    pipeline passes it (geometry.py:291-292).  alpha is the Born desolvation
    coefficient (desolvation_alpha, input_parser.py:75-77) and min_core_dt the
    dt floor (minimum_core_dt), as an input.xml sets them.  point_charges:
    the potential grid holds the screened Coulomb field of point charges
    (ion-exclusion radius 0), the field of the analytic two-sphere solution
    (examples/two_charged_spheres/analytical.py:39-40)."""
    from .pipeline.geometry import pathway_set_from_rxns
    from .structures.pqr_io import parse_pqr
    mol1, mol2 = parse_pqr(receptor_pqr), parse_pqr(ligand_pqr)
    for mol in (mol1, mol2):
        c = mol.centroid()
        for a in mol.atoms:
            a.x, a.y, a.z = a.x - c[0], a.y - c[1], a.z - c[2]
    rp, rq, rr = mol1.positions_array(), mol1.charges_array(), mol1.radii_array()
    lp, lq, lr = mol2.positions_array(), mol2.charges_array(), mol2.radii_array()
    if a_rec is None:
        a_rec = float((np.linalg.norm(rp, axis=1) + rr).max())
    if a_lig is None:
        a_lig = float((np.linalg.norm(lp, axis=1) + lr).max())
    if b is None:
        b = 1.1 * (a_rec + a_lig + 15.0)
    ps = pathway_set_from_rxns(rxns_xml)
    mob = MobilityTensor.from_radii(a_rec, a_lig)
    params = _params(n_traj, float(b), seed, max_steps, overlap=overlap, hi=hi,
                     min_core_dt=min_core_dt)
    if r_escape is not None:
        params.r_escape = float(r_escape)
    gn, gh = grid
    ln, lh = lig_grid if lig_grid is not None else grid
    specs = [GridSpec("elec", 0, (gn,) * 3, _cube_origin(gn, gh), gh, dtype,
                      _atoms5(rp, rq, 0.0 * rr if point_charges else rr), debye),
             GridSpec("born", 1, (gn,) * 3, _cube_origin(gn, gh), gh, dtype, _atoms5(rp, rq, rr), debye),
             GridSpec("lig_born", 1, (ln,) * 3, _cube_origin(ln, lh), lh, dtype,
                      _atoms5(lp, lq, lr), debye)]
    mp = MultipoleExpansion(rp, rq, debye, 78.0)
    kw = dict(alpha=float(alpha), receptor_charge=float(rq.sum()), debye_length=debye, sdie=78.0,
              rec_positions=rp, rec_charges=rq, multipole_expansion=mp)
    return BDSystem(name or f"{mol1.name}-{mol2.name}", mol1, mol2, mob, ps, params, kw, specs,
                    {"a_rec": a_rec, "a_lig": a_lig, "files": [str(receptor_pqr), str(ligand_pqr),
                                                               str(rxns_xml)]})


def multi_ligand(n_traj=1_000_000, seed=5015, rec_seed=5005, lig_seed=5006, n_rec=4700,
                 n_lig=16, lig_atoms=(12, 64), rec_grid=(257, 0.5), lig_grid=(65, 0.4),
                 b=40.0, max_steps=1_000_000, rec_net=4.0, density=0.0826, rec_nn=2.05):
    """Config 5: a 4,700-atom receptor on 257^3 FP32 potential + Born grids
    (0.5 A) and n_lig ligands of U{12..64} atoms (charges N(0, 0.3) e, integer
    nets in [-2, 2]), each with its own 65^3 Born grid at 0.4 A and its own
    two-of-three contact criterion (4.0 A, docked pose); b = 40 A.  The
    receptor is a filled FCC ball (nearest neighbours rec_nn = 2.05 A, radius
    ~19 A) whose interior holds no empty sphere of the 1.5 A overlap radius,
    so ligand centroids stay outside it as they stay outside a real protein;
    ligands are jittered lattices at the atom density of the 1,700-atom
    globule (0.0826 / A^3).  Returns one BDSystem per ligand: the first
    carries the receptor grids, the others only their ligand grid (their
    engines share the receptor grids, ``make_engine(..., receptor_from=)``)."""
    rr_ = np.random.default_rng(rec_seed)
    rp = fcc_ball(n_rec, rec_nn, rr_)
    r_rec = float(np.linalg.norm(rp, axis=1).max())
    s1 = rec_nn
    rq = charges_with_net(n_rec, 0.35, rec_net, rr_)
    rrad = rr_.uniform(1.4, 2.0, n_rec)
    mol1 = Molecule.from_arrays(rp, rq, rrad, name="receptor")
    ra = _atoms5(rp, rq, rrad)
    gn, gh = rec_grid
    ln, lh = lig_grid
    mp = MultipoleExpansion(rp, rq, DEBYE_015M, 78.0)
    kw = dict(alpha=ALPHA_DEFAULT, receptor_charge=float(rq.sum()), debye_length=DEBYE_015M,
              sdie=78.0, rec_positions=rp, rec_charges=rq, multipole_expansion=mp)
    lr_ = np.random.default_rng(lig_seed)
    systems = []
    for k in range(n_lig):
        m = int(lr_.integers(lig_atoms[0], lig_atoms[1] + 1))
        net = float(lr_.integers(-2, 3))
        r_l = max((m / density * 3.0 / (4.0 * math.pi)) ** (1.0 / 3.0), 2.5)
        lp, _ = lattice_cloud(m, (r_l,) * 3, lr_)
        lq = charges_with_net(m, 0.3, net, lr_)
        lrad = lr_.uniform(1.4, 2.0, m)
        mol2 = Molecule.from_arrays(lp, lq, lrad, name=f"ligand{k}")
        pairs = _contact_pairs(rp, lp, 3, 4.0)
        ps = PathwaySet([ReactionInterface("rxn", ReactionCriteria("contact", pairs, n_needed=2))])
        mob = MobilityTensor.from_radii(r_rec, r_l)
        params = _params(n_traj, b, seed + k, max_steps, overlap=True)
        specs = [GridSpec("lig_born", 1, (ln,) * 3, _cube_origin(ln, lh), lh, "float32",
                          _atoms5(lp, lq, lrad))]
        if k == 0:
            specs = [GridSpec("elec", 0, (gn,) * 3, _cube_origin(gn, gh), gh, "float32", ra),
                     GridSpec("born", 1, (gn,) * 3, _cube_origin(gn, gh), gh, "float32", ra)] + specs
        systems.append(BDSystem(f"multi_ligand_{k}", mol1, mol2, mob, ps, params, dict(kw), specs,
                                {"ligand": k, "atoms": m, "net": net, "r_rec": r_rec,
                                 "r_lig": r_l, "lattice_spacing": s1}))
    return systems
