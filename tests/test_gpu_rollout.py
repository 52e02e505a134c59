"""Whole-window tests of the HIP path through the C-ABI: gradients against
finite differences of the FP64 oracle, graph replay against eager launches,
checkpoint plans against each other, the trajectory gradient, event timing.

Tolerances:
  * gradient vs FP64 central FD (h = 1e-4 relative, committed fixture):
    1e-3 relative, the bar the gradient must meet;
  * graph replay vs eager and checkpoint plans vs each other: 1e-5 relative
    (the same kernels; only float-atomic summation order differs);
  * trajectory gradient vs the oracle's autograd: 2e-3 of its max norm.
"""
import os

import numpy as np
import pytest
import torch

from diffcp_amd import scene
from oracle import sim_torch as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "c1_grad_fd.npz")


def _engine(insts, **kw):
    from diffcp_amd.engine import Engine
    return Engine(insts, **kw)


def _slots(S, K):
    from diffcp_amd.engine import Engine
    return Engine.slots_for(S, K)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("svd_grad", ["reference", "stable"])
def test_c1_gradient_vs_fp64_finite_differences(svd_grad):
    ref = np.load(FIXTURE)
    w = scene.config_c1()
    S = int(ref["substeps"])
    eng = _engine(w.instances, max_states=_slots(S, S), svd_grad=svd_grad)
    eng.set_targets([ref["targets"]])
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=False)
    assert abs(loss[0] - float(ref["loss"])) < 1e-4 * abs(float(ref["loss"]))
    fd = ref["fd_0.0001"]
    for j, name in enumerate(("E", "nu")):
        err = abs(grads[0, j] - fd[j]) / abs(fd[j])
        assert err < 1e-3, (name, grads[0, j], fd[j], err)


def _c1_engine(S, K, **kw):
    w = scene.config_c1()
    inst = w.instances[0]
    inst.targets = scene.sample_targets(inst.vertex_x, 512, 3) + np.array([0.0, 0.004, 0.0])
    eng = _engine([inst], max_states=_slots(S, K), **kw)
    eng.reset(0)
    return eng


def test_graph_replay_matches_eager():
    S = 120
    eng = _c1_engine(S, 40)
    l0, g0, _ = eng.rollout_grad(S, 0, 40, use_graph=False)
    for rep in range(3):
        l1, g1, _ = eng.rollout_grad(S, 0, 40, use_graph=True)
        assert _rel(l1, l0) < 1e-5 and _rel(g1, g0) < 1e-5, rep
        # eager work and a synchronising read-back between replays
        eng.substep(0, 0, 0, 1)
        eng.download_state(1)
    assert eng.status() & 1 == 0


def test_checkpoint_plans_agree():
    S = 120
    res = []
    for K in (120, 40, 17):
        eng = _c1_engine(S, K)
        l, g, _ = eng.rollout_grad(S, 0, K, use_graph=True)
        res.append((l, g))
        eng.close()
    for l, g in res[1:]:
        assert _rel(l, res[0][0]) < 1e-5
        assert _rel(g[:, :5], res[0][1][:, :5]) < 1e-5


def test_grid_snapshots_agree_with_recompute():
    """Snapshots on (grids and element prep reused), off (recomputed) and
    bounded to 4 blocks per substep (every grid overflows: P2G reruns on the
    saved element prep) give the same window, with checkpoint recompute."""
    S, K = 120, 40
    res = []
    for snaps in (False, True, 4):
        eng = _c1_engine(S, K, grid_snapshots=snaps)
        l, g, _ = eng.rollout_grad(S, 0, K, use_graph=True)
        l2, g2, _ = eng.rollout_grad(S, 0, K, use_graph=True)
        assert _rel(l2, l) < 1e-5 and _rel(g2[:, :5], g[:, :5]) < 1e-5, snaps
        assert eng.status() & 19 == 0
        res.append((l, g))
        eng.close()
    for l, g in res[1:]:
        assert _rel(l, res[0][0]) < 1e-5
        assert _rel(g[:, :5], res[0][1][:, :5]) < 1e-5


def test_kernel_timing_events_eager_and_graph():
    S = 60
    eng = _c1_engine(S, S)
    eng.set_kernel_timing("p2g", 4 * S)
    eng.rollout_grad(S, 0, S, use_graph=False)
    ms_e, n_e = eng.kernel_timing()
    # the forward's P2G launches; the adjoint's, which exit at once when the
    # grid snapshot holds the substep, are not timed
    assert S - 1 <= n_e <= S and ms_e > 0
    eng.set_kernel_timing(None)
    ms_0, n_0 = eng.kernel_timing()
    assert n_0 == 0


@pytest.mark.parametrize("svd_grad", ["reference", "stable"])
def test_per_element_moduli_gradient_matches_oracle(svd_grad):
    """Gradients w.r.t. per-triangle (E, shearing, contact) moduli (seeded
    lognormal, sigma 0.3) against the oracle's FP64 autograd, 2e-3 of the
    largest component; nu stays a scalar."""
    inst = scene.tiny_instance(n=6, n_grid=16, substeps=40)
    rng = np.random.default_rng(11)
    base = np.array([inst.E, inst.shearing_stiffness, inst.contact_stiffness])
    inst.element_moduli = base[None, :] * np.exp(rng.normal(0.0, 0.3, size=(inst.n_elements, 3)))
    inst.targets = inst.vertex_x[rng.choice(inst.n_vertices, 16, replace=False)] + rng.normal(0, 0.01, (16, 3))
    S = 30
    eng = _engine([inst], max_states=_slots(S, S), svd_grad=svd_grad, per_element_moduli=True)
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=False)
    ge = eng.element_moduli_grads()

    params = O.make_params(inst, requires_grad=True)
    states, _ = O.rollout(inst, params, S, svd_grad=svd_grad, keep_states=True)
    lo, _ = O.compute_loss(states[S - 2]["vx"], torch.as_tensor(inst.targets))
    lo.backward()
    ref = params["elem"].grad.numpy()
    assert abs(loss[0] - lo.item()) < 1e-4 * lo.item()
    assert np.abs(ref).max() > 0
    for c in range(3):                     # E, shearing, contact columns
        scale = np.abs(ref[:, c]).max()
        if scale == 0:
            assert np.abs(ge[:, c]).max() == 0
            continue
        assert np.abs(ge[:, c] - ref[:, c]).max() < 2e-3 * scale, (c, np.abs(ge[:, c] - ref[:, c]).max(), scale)
    gnu = float(params["nu"].grad)
    assert abs(grads[0, 1] - gnu) < 2e-3 * abs(gnu)


def test_trajectory_gradient_matches_oracle():
    inst = scene.tiny_instance(n=5, n_grid=16, substeps=40)
    rng = np.random.default_rng(5)
    inst.targets = inst.vertex_x[rng.choice(inst.n_vertices, 12, replace=False)] + rng.normal(0, 0.01, (12, 3))
    S = 30
    eng = _engine([inst], max_states=_slots(S, S), svd_grad="stable", deterministic=True)
    eng.reset(0)
    loss, _, _ = eng.rollout_grad(S, 0, S, use_graph=False)
    gt = eng.trajectory_grads()[0]

    st, static = O.initialize_state(inst)
    static["traj"] = static["traj"].clone().requires_grad_(True)
    params = O.make_params(inst)
    states, _ = O.rollout(inst, params, S, svd_grad="stable", keep_states=True, state=st, static=static)
    lo, _ = O.compute_loss(states[S - 2]["vx"], torch.as_tensor(inst.targets))
    lo.backward()
    ref = static["traj"].grad.numpy()
    assert abs(loss[0] - float(lo)) < 1e-4 * float(lo)
    assert np.abs(ref).max() > 0
    assert np.abs(gt - ref).max() < 2e-3 * np.abs(ref).max(), (np.abs(gt - ref).max(), np.abs(ref).max())


# 16^3: tiles of 13 x 9 x 13 cells, cut in x only; 48^3: 33 x 22 x 33 cells,
# one x-plane alone above the cap, cut in x and y
@pytest.mark.parametrize("n_grid,spread_xz,y_lo,y_hi", [(16, 0.35, -0.1, 0.25), (48, 0.4, -0.15, 0.25)])
def test_overflowing_patch_tiles_through_sub_boxes(n_grid, spread_xz, y_lo, y_hi):
    """Patches whose tile exceeds the LDS cap are listed for the sub-box tile
    kernels (no status bit: nothing falls back to per-particle atomics).  An
    eager and a graph-replayed window: loss, (E, nu) and trajectory gradients
    match the oracle's FP64 autograd (loss 1e-4 relative; E and nu each 2e-3
    relative, the trajectory 2e-3 of its largest component) and, with
    deterministic scatter sums, each other bit for bit."""
    inst = scene.wide_grasp_instance(n_grid, spread_xz, y_lo, y_hi)
    gs = len(inst.grasper_x) // inst.trajectory.shape[1]
    base = np.trunc(inst.grasper_x[:gs] * inst.n_grid - 0.5)
    assert np.prod(base.max(0) - base.min(0) + 3) > 1024
    S = 30
    eng = _engine([inst], max_states=_slots(S, S), svd_grad="stable", deterministic=True)
    eng.reset(0)
    eng.status(clear=True)
    l1, g1, _ = eng.rollout_grad(S, 0, S, use_graph=False)
    t1 = eng.trajectory_grads()[0]
    assert eng.status(clear=True) & 23 == 0
    eng.reset(0)
    l2, g2, _ = eng.rollout_grad(S, 0, S, use_graph=True)
    t2 = eng.trajectory_grads()[0]
    assert eng.status() & 23 == 0

    st, static = O.initialize_state(inst)
    static["traj"] = static["traj"].clone().requires_grad_(True)
    params = O.make_params(inst, requires_grad=True)
    states, _ = O.rollout(inst, params, S, svd_grad="stable", keep_states=True, state=st, static=static)
    lo, _ = O.compute_loss(states[S - 2]["vx"], torch.as_tensor(inst.targets))
    lo.backward()
    ref_t = static["traj"].grad.numpy()
    ref_g = np.array([float(params["E"].grad), float(params["nu"].grad)])
    assert np.abs(ref_t).max() > 0
    for l, g, t in ((l1, g1, t1), (l2, g2, t2)):
        assert abs(l[0] - float(lo.detach())) < 1e-4 * float(lo.detach())
        assert np.all(np.abs(g[0, :2] - ref_g) < 2e-3 * np.abs(ref_g)), (g[0, :2], ref_g)
        assert np.abs(t - ref_t).max() < 2e-3 * np.abs(ref_t).max()
    assert np.array_equal(l2, l1) and np.array_equal(g2, g1) and np.array_equal(t2, t1)
