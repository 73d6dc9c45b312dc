"""ctypes binding of libpystarc_hip.so (C ABI in include/pystarc_hip.h).

The library is the only compute path: when it is missing, or no HIP device
is visible, every operator raises RuntimeError instead of falling back to
anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpystarc_hip.so"
MAX_GRIDS = 4
MAX_PAIRS = 16

_lib = None


class StepParams(C.Structure):
    """Mirror of StepParams in csrc/pst_types.hpp (field order and types)."""
    _fields_ = [
        ("n_total", C.c_int64), ("n_slots", C.c_int64), ("seed", C.c_uint64),
        ("traj_base", C.c_int64),
        ("r_b", C.c_double), ("r_esc", C.c_double), ("return_prob", C.c_double),
        ("D_t", C.c_double), ("D_r", C.c_double),
        ("use_hi", C.c_int),
        ("hi_Df", C.c_double), ("hi_ainv", C.c_double), ("hi_a2", C.c_double),
        ("adt_debye", C.c_double), ("adt_V0", C.c_double),
        ("adt_has_force", C.c_int),
        ("min_core_dt", C.c_double), ("max_dt", C.c_double),
        ("use_overlap", C.c_int),
        ("overlap_thresh", C.c_double),
        ("n_pairs", C.c_int), ("n_needed", C.c_int), ("has_bb_deff", C.c_int),
        ("rec_gho", C.c_double * (MAX_PAIRS * 3)),
        ("lig_gho", C.c_double * (MAX_PAIRS * 3)),
        ("cutoff", C.c_double * MAX_PAIRS),
        ("bb_deff", C.c_double * MAX_PAIRS),
        ("max_steps", C.c_int64),
        ("hydro_mode", C.c_int),
        ("a_rec", C.c_double), ("a_lig", C.c_double), ("kT_over_eta", C.c_double),
        ("surrogate_r", C.c_double),
        ("stop_done", C.c_int64),
        ("sm_active", C.c_int),
        ("sm_before", C.c_int * MAX_PAIRS),
        ("sm_after", C.c_int * MAX_PAIRS),
        ("sm_terminal", C.c_int),
        ("min_core_rxn_dt", C.c_double),
        ("lmz_on", C.c_int), ("lmz_hi", C.c_int),
        ("lmz_trigger", C.c_double), ("lmz_b", C.c_double), ("lmz_q", C.c_double),
        ("lmz_b_cover", C.c_double), ("lmz_q_cover", C.c_double),
        ("lmz_V", C.c_double), ("lmz_L", C.c_double), ("lmz_return_q", C.c_double),
        ("lmz_Dt", C.c_double), ("lmz_dpre", C.c_double), ("lmz_ainv", C.c_double),
        ("lmz_a2", C.c_double), ("lmz_drot0", C.c_double), ("lmz_drot1", C.c_double),
    ]


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)

_SIGS = {
    "pst_abi_version": ([], C.c_int),
    "pst_struct_size": ([C.c_int], C.c_int),
    "pst_ctx_create": ([C.POINTER(_vp), C.c_int, _vp, C.c_int], C.c_int),
    "pst_ctx_destroy": ([_vp], C.c_int),
    "pst_last_error": ([_vp], C.c_char_p),
    "pst_add_grid": ([_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "pst_clear_grids": ([_vp], C.c_int),
    "pst_share_receptor_grids": ([_vp, _vp], C.c_int),
    "pst_group_run": ([_vp, C.c_int, C.c_longlong, C.c_int, _vp, _dp], C.c_int),
    "pst_set_farfield": ([_vp, C.c_double, C.c_double, C.c_double, _vp, _vp], C.c_int),
    "pst_set_alpha": ([_vp, C.c_double], C.c_int),
    "pst_set_receptor": ([_vp, C.c_int, _vp, _vp], C.c_int),
    "pst_set_ligand": ([_vp, C.c_int, _vp, _vp], C.c_int),
    "pst_set_torque_mode": ([_vp, C.c_int], C.c_int),
    "pst_set_force_group": ([_vp, C.c_int], C.c_int),
    "pst_forces_positions": ([_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                              C.c_int, C.c_int], C.c_int),
    "pst_forces_positions_noalloc": ([_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      C.c_int, C.c_int], C.c_int),
    "pst_forces_pose": ([_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int], C.c_int),
    "pst_forces_pose_group": ([_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int], C.c_int),
    "pst_time_forces_pose": ([_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp],
                             C.c_int),
    "pst_gather_ceiling": ([_vp, C.c_int, C.c_longlong, C.c_int, _vp], C.c_int),
    "pst_rpy12_eval": ([_vp, C.c_int, _vp, C.c_double, C.c_double, _vp, _vp], C.c_int),
    "pst_rpy12_move": ([_vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _vp, _vp, _vp,
                        _vp, _vp, _vp, _vp], C.c_int),
    "pst_sim_create": ([_vp, _vp, C.c_int, C.POINTER(_vp)], C.c_int),
    "pst_sim_destroy": ([_vp], C.c_int),
    "pst_sim_set_state": ([_vp, C.c_int, _vp, _vp], C.c_int),
    "pst_sim_get_state": ([_vp, _vp, _vp, _vp, _vp], C.c_int),
    "pst_replay_begin": ([_vp, _ip, _ip, _ip], C.c_int),
    "pst_replay_escape": ([_vp, _vp, C.c_int, _vp, C.c_int, _ip], C.c_int),
    "pst_replay_move": ([_vp, _vp, _vp, C.c_int, _ip, _ip], C.c_int),
    "pst_replay_bridge": ([_vp, _vp, C.c_int, _ip], C.c_int),
    "pst_replay_forces": ([_vp, _vp, C.c_int], C.c_int),
    "pst_prod_init": ([_vp, C.c_int], C.c_int),
    "pst_prod_run": ([_vp, C.c_longlong, C.c_int, _vp], C.c_int),
    "pst_prod_outcomes": ([_vp, _vp, _vp], C.c_int),
    "pst_prod_obs_sizes": ([_ip, _ip], C.c_int),
    "pst_prod_sm_counts": ([_vp, _vp], C.c_int),
    "pst_prod_observables": ([_vp, _vp, _vp], C.c_int),
    "pst_prod_autotune": ([_vp, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp], C.c_int),
    "pst_prod_set_launch": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "pst_set_gather_mode": ([_vp, C.c_int, C.c_int], C.c_int),
    "pst_last_force_eval": ([_vp, _ip, _ip], C.c_int),
    "pst_prod_time_forces": ([_vp, C.c_int, _vp], C.c_int),
    "pst_prod_last_ms_per_step": ([_vp, _dp], C.c_int),
    "pst_prod_counters": ([_vp, _vp], C.c_int),
    "pst_make_grid": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, C.c_double,
                       C.c_double, C.c_double, C.c_int, _vp], C.c_int),
    "pst_prod_state_bytes": ([_vp, C.POINTER(C.c_longlong)], C.c_int),
    "pst_prod_save_state": ([_vp, _vp, C.c_longlong], C.c_int),
    "pst_prod_load_state": ([_vp, _vp, C.c_longlong], C.c_int),
    "pst_prod_snapshot": ([_vp, C.c_longlong, _vp, _vp, _ip], C.c_int),
    "pst_prod_audit": ([_vp, _vp], C.c_int),
    "pst_mlp_param_count": ([], C.c_int),
    "pst_mlp_create": ([_vp, C.c_int, C.POINTER(_vp)], C.c_int),
    "pst_mlp_destroy": ([_vp], C.c_int),
    "pst_mlp_load": ([_vp, _vp], C.c_int),
    "pst_mlp_forward": ([_vp, C.c_int, _vp, _vp, C.c_int], C.c_int),
    "pst_mlp_time_infer": ([_vp, C.c_int, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "pst_mlp_backward": ([_vp, C.c_int, _vp, _vp], C.c_int),
    "pst_sim_set_surrogate": ([_vp, _vp, _vp], C.c_int),
    "pst_sim_reset_audit": ([_vp], C.c_int),
}


def exported_symbols():
    """Names every binding expects the shared library to export."""
    return sorted(_SIGS)


def load(path: os.PathLike | None = None):
    """Load and type the library; RuntimeError when it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -m pystarc_amd.build)")
    lib = C.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def ptr(a) -> int:
    """Raw address of a torch tensor (device) or a contiguous numpy array (host)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("numpy array must be C-contiguous")
        return a.ctypes.data
    return a.data_ptr()


class Context:
    """Owns one pst_ctx (device grids, ligand/receptor tables, stream)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        """stream None: torch's current stream on `device` (work stays ordered
        with torch); otherwise a raw hipStream_t handle."""
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("pystarc_amd needs a HIP device; none is visible")
        self.lib = load()
        h = C.c_void_p()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        rc = self.lib.pst_ctx_create(C.byref(h), device, C.c_void_p(stream or None), 1)
        self.h = h
        if rc:
            raise RuntimeError(f"pst_ctx_create failed: {self.error()}")

    def error(self) -> str:
        msg = self.lib.pst_last_error(self.h)
        return msg.decode() if msg else ""

    def check(self, rc: int, what: str):
        if rc:
            raise RuntimeError(f"{what} failed: {self.error()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.pst_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
