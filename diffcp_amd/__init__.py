"""MI355X-native differentiable codimensional-MPM cloth substep.

The HIP library is opened on first use (``diffcp_amd._abi.load``); importing
the package loads nothing.  Library errors raise ``DiffcpError``.
"""
from ._abi import DiffcpError

__all__ = ["DiffcpError"]
