"""The C-ABI library loads and exports every entry point the header declares
(no compute: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from pystarc_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "pystarc_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pst_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_bindings():
    names = declared()
    assert names, "no declarations parsed"
    assert set(names) == set(_lib.exported_symbols())


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} not built (run python -m pystarc_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.pst_abi_version() == 1


def test_step_params_layout_matches_native():
    lib = _lib.load()
    assert lib.pst_struct_size(0) == ctypes.sizeof(_lib.StepParams)


def test_context_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    with pytest.raises(RuntimeError):
        _lib.Context(0)
