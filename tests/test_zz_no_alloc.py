"""dcp_polar_adjoint_noalloc (no device allocation, static scratch) against
the allocating dcp_polar_adjoint through the C-ABI: two calls in a row, the
second smaller, every precision and SVD-backward mode; the same kernel on the
same inputs, so the outputs are bit-identical.  Sizes beyond the static
scratch are rejected with an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    Ft = (np.eye(2) + 0.3 * rng.standard_normal((n, 2, 2))).astype(np.float32)
    gR = rng.standard_normal((n, 2, 2)).astype(np.float32)
    return Ft, gR


def test_noalloc_polar_adjoint_matches_the_allocating_entry():
    from diffcp_amd.engine import polar_adjoint
    for call, n in enumerate((16384, 3000)):
        Ft, gR = _inputs(n, call)
        for precision in ("mixed", "fp32"):
            for mode in ("reference", "stable"):
                R0, g0 = polar_adjoint(Ft, gR, precision=precision, svd_grad=mode)
                R1, g1 = polar_adjoint(Ft, gR, precision=precision, svd_grad=mode, no_alloc=True)
                assert np.array_equal(R0, R1) and np.array_equal(g0, g1), (call, n, precision, mode)
                assert np.isfinite(g1).all()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from diffcp_amd._abi import DiffcpError
    from diffcp_amd.engine import polar_adjoint
    Ft, gR = _inputs(16385, 7)
    with pytest.raises(DiffcpError, match="static scratch"):
        polar_adjoint(Ft, gR, no_alloc=True)
