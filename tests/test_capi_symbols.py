"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/cffc_euler2d.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cffc_euler2d.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cffc_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cffc_create", "cffc_destroy", "cffc_dudt_stage", "cffc_update_stage", "cffc_cfl",
              "cffc_exchange_and_bcs", "cffc_residual_norms", "cffc_run_steps"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from cffc_amd import euler2d
    path = euler2d.library_path()
    if not os.path.exists(path):
        pytest.skip("libcffc_hip.so not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_device_fails_cleanly():
    """No HIP device in the build container: cffc_create must return
    CFFC_E_NODEVICE (-5) -- there is no CPU fallback."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from cffc_amd import abi, euler2d
    from cffc_amd import mesh as M
    m = M.cartesian(1, 1, 4, 4)
    with pytest.raises(euler2d.CffcError, match=r"\(-5\)"):
        euler2d.Euler2DQuadBlocks(m, abi.Scheme())
