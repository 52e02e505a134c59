"""The C-ABI library loads on the host (no GPU needed) and exports exactly
what include/diffcp.h declares; the ctypes table covers every declaration."""
import ctypes
import os
import re

import pytest

from diffcp_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diffcp.h")


def _declared():
    src = open(HEADER).read()
    return re.findall(r"^\s*(?:int|const char\*)\s+(dcp_\w+)\s*\(", src, re.M)


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) == len(set(names))
    for core in ("dcp_create", "dcp_destroy", "dcp_substep", "dcp_substep_adjoint", "dcp_loss",
                 "dcp_loss_adjoint", "dcp_param_grads", "dcp_rollout_grad", "dcp_last_error"):
        assert core in names


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    declared = set(_declared())
    assert declared == set(_abi.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="library not built")
def test_error_path_without_device_work():
    """A call with a null context fails with a message and touches no device."""
    lib = _abi.load()
    rc = lib.dcp_substep(None, 0, 0, 0, 1)
    assert rc < 0
    assert lib.dcp_last_error()
