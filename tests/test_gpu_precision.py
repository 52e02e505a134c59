"""Precision stress (the compress-then-fold trajectory of the precision-stress
workload on a 24x24 cloth, 64^3 grid): both precision modes against the full
FP64 oracle.

Measured on MI355X (the bars below have headroom over these):
  * positions after 120 substeps: 2.5e-6 relative in both modes, velocities
    2.4e-5 (mixed) and 2.7e-5 (all-FP32);
  * loss gradient w.r.t. E and nu over 80 substeps: 1.6e-4 relative in both
    modes and both SVD backwards.
Bars: positions 2e-5, velocities 5e-3 (the forward parity tolerances), and
gradients 1e-3 relative (the gradient bar for this path).
"""
import numpy as np
import pytest
import torch

from diffcp_amd import scene
from oracle import reduced as R
from oracle import sim_torch as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.fixture(scope="module")
def stress():
    return scene.precision_stress_instance(substeps=120)


def test_stress_scene_is_near_degenerate(stress):
    st, _ = O.rollout(stress, n_substeps=120)
    _, Rm = O.qr_gram_schmidt(st["qF"])
    sv = torch.linalg.svdvals(Rm[:, :2, :2])
    assert float(sv.min()) < 0.2 and float((sv < 1).double().mean()) > 0.3
    assert float((sv[:, 0] - sv[:, 1]).abs().min()) < 1e-2


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_stress_positions_vs_fp64(stress, precision):
    from diffcp_amd.engine import Engine
    N = 120
    states, _ = O.rollout(stress, n_substeps=N, keep_states=True)
    eng = Engine([stress], max_states=N + 2, precision=precision)
    eng.reset(0)
    for f in range(N):
        eng.substep(f, 0, f, f + 1)
    assert eng.status() & 3 == 0
    for s in (N // 2, N):
        gpu = eng.download_state(s)
        ref = R.reduce(states[s], states[s - 1])
        assert _rel(gpu["vertex_x"], ref["vx"].numpy()) < 2e-5, (s, precision)
        assert _rel(gpu["vertex_v"], ref["vv"].numpy()) < 5e-3, (s, precision)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
@pytest.mark.parametrize("svd_grad", ["reference", "stable"])
def test_stress_gradient_vs_fp64(stress, precision, svd_grad):
    from diffcp_amd.engine import Engine
    S = 80
    final, _ = O.rollout(stress, n_substeps=S)
    rng = np.random.default_rng(4)
    targets = final["vx"].numpy()[rng.choice(stress.n_vertices, 64, replace=False)] + rng.normal(0, 0.003, (64, 3))
    params = O.make_params(stress, requires_grad=True)
    lo, _ = O.tape_loss(stress, params, S, targets, svd_grad=svd_grad)
    lo.backward()
    ref = np.array([params["E"].grad.item(), params["nu"].grad.item()])
    eng = Engine([stress], max_states=Engine.slots_for(S, S), precision=precision, svd_grad=svd_grad)
    eng.set_targets([targets])
    eng.reset(0)
    loss, g, _ = eng.rollout_grad(S, 0, S, use_graph=False)
    assert abs(loss[0] - lo.item()) < 1e-4 * lo.item()
    err = np.abs(g[0, :2] - ref) / np.abs(ref)
    assert (err < 1e-3).all(), (precision, svd_grad, err)


def _near_degenerate(delta, n=2000, seed=0):
    """FP32 2x2 matrices U diag(s1, s2) V^T with s1^2 - s2^2 = delta, s2 in
    [0.3, 1.2], random rotations U and V (the regime of the 1/(si^2 - sj^2)
    terms of the SVD backward, simulation.py:71-94)."""
    rng = np.random.default_rng(seed)
    s2 = rng.uniform(0.3, 1.2, n)
    s1 = np.sqrt(s2 ** 2 + delta)
    a, b = rng.uniform(0, 2 * np.pi, (2, n))
    rot = lambda t: np.stack([np.stack([np.cos(t), -np.sin(t)], -1), np.stack([np.sin(t), np.cos(t)], -1)], -2)
    S = np.zeros((n, 2, 2))
    S[:, 0, 0], S[:, 1, 1] = s1, s2
    Ft = (rot(a) @ S @ rot(b).transpose(0, 2, 1)).astype(np.float32)
    gR = rng.standard_normal((n, 2, 2)).astype(np.float32)
    return Ft, gR


@pytest.mark.parametrize("delta", [1e-8, 1e-7, 1e-6, 1e-5])
def test_polar_adjoint_near_degenerate_stretch(delta):
    """The library's 2x2 polar adjoint against the oracle's literal SVD
    backward (the clamped reference formula) and the exact polar derivative,
    both evaluated in FP64 on the same FP32 inputs.  Bar: 1e-6 relative (to
    the largest entry) for the FP64 path; the all-FP32 path's error is
    printed and bounded at 1e-2."""
    from diffcp_amd.engine import polar_adjoint
    Ft, gR = _near_degenerate(delta)
    F64 = torch.as_tensor(Ft, dtype=torch.float64)
    g64 = torch.as_tensor(gR, dtype=torch.float64)
    ref = {"reference": O.reference_svd_backward(F64, g64).numpy()}
    Fg = F64.clone().requires_grad_(True)
    (O.polar_rotation_closed(Fg) * g64).sum().backward()
    ref["stable"] = Fg.grad.numpy()
    R_ref = O.polar_rotation_closed(F64).numpy()
    for mode in ("reference", "stable"):
        for precision in ("mixed", "fp32"):
            R, gF = polar_adjoint(Ft, gR, precision=precision, svd_grad=mode)
            err = _rel(gF, ref[mode])
            print(f"delta {delta:.0e} svd_grad {mode:9s} {precision:5s}: gradient rel err {err:.2e}, "
                  f"rotation rel err {_rel(R, R_ref):.2e}")
            assert _rel(R, R_ref) < 1e-6
            if precision == "mixed":
                assert err < 1e-6, (mode, err)
            else:
                assert err < 1e-2, (mode, err)
