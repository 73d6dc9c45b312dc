"""Test helpers: small benchmark systems on CPU-evaluated grids, in the
form the oracle and the product engine take them."""
from __future__ import annotations

import numpy as np

from oracle import forces as OF
from oracle import grids as OG
from oracle.physics import EPS0  # noqa: F401
from pystarc_amd import systems as S
from pystarc_amd.forces.electrostatic.grid_force import DXGrid


def numpy_grids(system):
    """role -> list of (data, origin, spacing) evaluated with the oracle's NumPy generator."""
    out = {"elec": [], "born": [], "lig_born": []}
    for gs in system.grid_specs:
        data = OG.make_grid(gs.kind, gs.shape, gs.origin, gs.spacing, gs.atoms, gs.debye,
                            gs.eps_out, gs.eps_in)
        data = data.astype(np.float32 if gs.dtype == "float32" else np.float64)
        out[gs.role].append((data, np.asarray(gs.origin, float), gs.spacing))
    return out


def dx_grids(np_grids):
    return {k: [DXGrid(o, np.diag([h, h, h]), d) for (d, o, h) in v] for k, v in np_grids.items()}


def oracle_engine(system, np_grids, chunk=None):
    kw = system.engine_kwargs
    mp = kw.get("multipole_expansion")
    mpd = None if mp is None else {"dipole": mp.dipole, "quadrupole": mp.quadrupole}

    def prep(lst):
        # float32 grids are read as float32 values promoted to float64, like the kernel does
        return [OF.prepare_grid(d.astype(np.float64), o, (h, h, h)) for (d, o, h) in lst]

    return OF.build_engine(prep(np_grids["elec"]), prep(np_grids["born"]), kw.get("alpha", 0.0),
                           kw.get("receptor_charge", 0.0), kw.get("debye_length", 7.858),
                           kw.get("sdie", 78.0), prep(np_grids["lig_born"]) or None,
                           kw.get("rec_positions"), kw.get("rec_charges"), mpd, chunk=chunk)


def oracle_system(system, np_grids, chunk=None):
    mol1, mol2, mob = system.mol1, system.mol2, system.mob
    pairs = []
    nn = -1
    rxns = system.pathway_set.reactions
    sm = any(getattr(r, "state_before", None) is not None for r in rxns)
    sm_reactions = []
    for rxn in rxns:
        nn = rxn.criteria.n_needed
        plist = rxn.criteria.pairs[:1] if sm else rxn.criteria.pairs
        for p in plist:
            pairs.append((p.mol1_atom_index, p.mol2_atom_index, p.distance_cutoff))
        if sm and rxn.criteria.pairs:
            sm_reactions.append((rxn.state_before, rxn.state_after))
    return {
        "lig_body": mol2.positions_array() - mol2.centroid(),
        "lig_q": mol2.charges_array(),
        "rec_pos": mol1.positions_array(),
        "pairs": pairs, "n_needed": nn,
        "q_rec": mol1.total_charge(), "q_lig": mol2.total_charge(),
        "D_t": mob.relative_translational_diffusion(),
        "D_r": mob.relative_rotational_diffusion(),
        "a_rec": mob.radius1, "a_lig": mob.radius2,
        "engine": oracle_engine(system, np_grids, chunk),
        "debye": system.engine_kwargs.get("debye_length", 7.858),
        "sm_reactions": sm_reactions,
        "first_state": getattr(system.pathway_set, "first_state", None),
    }


def oracle_params(system, **over):
    p = system.params
    d = {"n_trajectories": p.n_trajectories, "seed": p.seed, "r_start": p.r_start,
         "r_escape": p.r_escape, "max_steps": p.max_steps,
         "hydrodynamic_interactions": p.hydrodynamic_interactions,
         "minimum_core_dt": p.minimum_core_dt, "max_dt": getattr(p, "max_dt", 0.0),
         "overlap_check": getattr(p, "_overlap_check", True),
         "gpu_force_batch": getattr(p, "gpu_force_batch", 0), "kT_scale": 1.0,
         "minimum_core_reaction_dt": getattr(p, "minimum_core_reaction_dt", 0.0) or 0.0}
    d.update(over)
    return d


def small_protein_ligand(n_traj=256, seed=7, n_rec=200, grid=(41, 0.5), lig_grid=(25, 0.4)):
    """A scaled-down protein-ligand system (same generator) for parity tests."""
    s = S.protein_ligand(n_traj=n_traj, seed=seed, n_rec=n_rec, rec_axes=(8.0, 7.0, 6.0),
                         rec_grid=grid, lig_grid=lig_grid, max_steps=100000)
    s.params.r_start = 14.0
    s.params.r_escape = 28.0
    return s


def small_protein_protein(n_traj=128, seed=9, n_rec=150, n_lig=120, grid=(41, 0.5)):
    s = S.protein_protein(n_traj=n_traj, seed=seed, n_rec=n_rec, n_lig=n_lig, r_rec=6.0,
                          r_lig=5.0, grid=grid, max_steps=100000)
    s.params.r_start = 16.0
    s.params.r_escape = 32.0
    return s


def two_spheres_small(n_traj=2000, seed=11111113, grid_n=97):
    return S.two_spheres(n_traj=n_traj, seed=seed, grid_n=grid_n)


def state_machine_protein_protein(n_traj=128, seed=9):
    """The small protein pair with three state-machine reactions on its first
    three contact pairs: start -(pair 0)-> mid -(pair 1)-> bound, and
    start -(pair 2)-> bound; 'bound' is terminal."""
    from pystarc_amd.pathways.reaction_interface import PathwaySet, ReactionInterface
    from pystarc_amd.structures.molecules import ReactionCriteria
    from pystarc_amd.structures.molecules import ContactPair
    s = small_protein_protein(n_traj=n_traj, seed=seed)
    pairs = s.pathway_set.reactions[0].criteria.pairs
    wide = lambda p, extra: ContactPair(p.mol1_atom_index, p.mol2_atom_index,
                                        p.distance_cutoff + extra)
    mk = lambda name, p, b, a: ReactionInterface(name, ReactionCriteria(name, [p], n_needed=1),
                                                 state_before=b, state_after=a)
    s.pathway_set = PathwaySet([mk("dock", wide(pairs[0], 6.0), "start", "mid"),
                                mk("lock", wide(pairs[1], 1.0), "mid", "bound"),
                                mk("direct", pairs[2], "start", "bound")], first_state="start")
    s.params.minimum_core_reaction_dt = 0.05
    return s
