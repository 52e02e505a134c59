"""The opt-in live element return mapping (Engine(return_mapping=True)):
F = Q return_map(R) with cf = the friction parameter (the mapping of
simulation.py:408-428, which the reference computes and discards).  Forward
states, single adjoint substeps and whole-window gradients, including
d loss / d friction, against the oracle run with the same mapping.

Tolerances as in test_gpu_parity.py: 2e-5 relative on positions and normals,
5e-3 on velocities and affine matrices, 2e-4 on substep state gradients,
2e-3 on parameter gradients as loss sensitivities (see _params_close; the
FP32 floor is 1e-5 of the largest sensitivity).
"""
import numpy as np
import pytest
import torch

from diffcp_amd import scene
from oracle import reduced as R
from oracle import sim_torch as O

pytestmark = pytest.mark.gpu
BAD_STATUS = 1 | 2 | 16
PARAMS = ("E", "nu", "ks", "kc", "damping", "friction")


def _engine(insts, **kw):
    from diffcp_amd.engine import Engine
    return Engine(insts, **kw)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _params_close(g, ref, theta, rtol=2e-3, floor=1e-5):
    """Parameter gradients g[j] against ref[j], compared as loss sensitivities
    theta_j g_j (theta_j = max(|parameter|, 1)): within rtol of each, plus
    floor times the largest sensitivity.  The floor is the FP32 limit: a
    component whose sensitivity is orders below the others (the shearing
    stiffness here, ~1e-7 of E's) differs by ~10 % between the FP32 and FP64
    oracles themselves."""
    th = np.maximum(np.abs(np.asarray(theta, np.float64)), 1.0)
    g, ref = np.asarray(g, np.float64) * th, np.asarray(ref, np.float64) * th
    bar = rtol * np.abs(ref) + floor * np.abs(ref).max()
    assert (np.abs(g - ref) <= bar).all(), (g / th, ref / th, np.abs(g - ref) / np.maximum(np.abs(ref), 1e-300))


def _scene():
    """The compress-and-fold trajectory on an 8x8 cloth: elements go into
    normal compression (R22 <= 1) where the friction cone acts."""
    inst = scene.precision_stress_instance(n=8, n_grid=32, substeps=40)
    inst.friction = 0.5
    return inst


def _gpu(inst, N, **kw):
    eng = _engine([inst], max_states=N + 2, return_mapping=True, **kw)
    eng.reset(0)
    for f in range(N):
        eng.substep(f, 0, f, f + 1)
    return eng


def test_return_mapping_forward_matches_oracle():
    inst = _scene()
    N = 30
    eng = _gpu(inst, N)
    assert eng.status() & BAD_STATUS == 0
    with torch.no_grad():
        on, _ = O.rollout(inst, n_substeps=N, keep_states=True, return_mapping=True)
        off, _ = O.rollout(inst, n_substeps=N, keep_states=True)
    for s in (1, 10, N):
        gpu = eng.download_state(s)
        ref = R.reduce(on[s], on[s - 1])
        for name, key in R.LIB_NAMES.items():
            tol = 2e-5 if name.endswith("_x") or name == "element_n" else 5e-3
            assert _rel(gpu[name], ref[key].numpy()) < tol, (s, name, _rel(gpu[name], ref[key].numpy()))
    # the mapping moves the normals far beyond that tolerance (1e-3 relative on this scene)
    assert _rel(R.reduce(off[N], off[N - 1])["qn"].numpy(), R.reduce(on[N], on[N - 1])["qn"].numpy()) > 5e-4


@pytest.mark.parametrize("svd_grad", ["stable", "reference"])
def test_return_mapping_adjoint_substep(svd_grad):
    inst = _scene()
    K = O.Consts.from_instance(inst)
    N = 24
    eng = _gpu(inst, N, svd_grad=svd_grad)
    _, static = O.initialize_state(inst)
    rng = np.random.default_rng(21)
    for f in (12, 20):
        st = eng.download_state(f)
        cot = {k: rng.standard_normal(v.shape).astype(np.float32) for k, v in st.items()}
        eng.upload_adjoint(0, **cot)
        eng.zero_param_grads()
        eng.substep_adjoint(f, 0, f, 0, 1)
        g = eng.download_adjoint(1)
        pg = eng.param_grads()[0]
        red = R.from_library(st)
        for v in red.values():
            v.requires_grad_(True)
        params = O.make_params(inst, requires_grad=True)
        out = R.reduced_substep(red, static, params, K, f, 0, svd_grad, return_mapping=True)
        tot = sum((out[R.LIB_NAMES[k]] * torch.as_tensor(cot[k], dtype=torch.float64)).sum() for k in cot)
        tot.backward()
        for name, key in R.LIB_NAMES.items():
            ref = red[key].grad
            ref = torch.zeros_like(red[key]) if ref is None else ref
            assert _rel(g[name], ref.numpy()) < 2e-4, (f, name, _rel(g[name], ref.numpy()))
        _params_close(pg, [params[pn].grad.item() for pn in PARAMS], inst.params())
        assert abs(params["friction"].grad.item()) > 0 and abs(pg[5]) > 0


def test_return_mapping_window_gradient_matches_oracle():
    """Whole window (40 substeps, loss on state 38) with the friction gradient."""
    from diffcp_amd.engine import Engine
    inst = _scene()
    S = 40
    with torch.no_grad():
        st, _ = O.rollout(inst, n_substeps=S - 2, return_mapping=True)
    targets = st["vx"].numpy() + np.array([0.002, -0.003, 0.001])
    eng = _engine([inst], max_states=Engine.slots_for(S, 10), return_mapping=True, svd_grad="stable")
    eng.set_targets([targets])
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, 10, use_graph=True)
    p = O.make_params(inst, requires_grad=True)
    ref_loss, _ = O.tape_loss(inst, p, S, targets, svd_grad="stable", return_mapping=True)
    ref_loss.backward()
    assert abs(loss[0] - ref_loss.item()) <= 1e-4 * ref_loss.item()
    _params_close(grads[0], [p[pn].grad.item() for pn in PARAMS], inst.params())
    assert abs(grads[0, 5]) > 0
    # without the mapping the friction gradient is exactly zero, as in the reference
    eng0 = _engine([inst], max_states=Engine.slots_for(S, 10), svd_grad="stable")
    eng0.set_targets([targets])
    eng0.reset(0)
    _, g0, _ = eng0.rollout_grad(S, 0, 10, use_graph=True)
    assert g0[0, 5] == 0.0
