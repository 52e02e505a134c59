"""The benchmark workloads through the C-ABI against the FP64 oracle.

  * 64x64 cloth, 128^3 grid, 400 substeps, checkpoints every 50 (the plain
    GPU workload): loss and dL/dE, dL/dnu with segment recompute against the
    oracle's autograd and central differences (committed fixture
    tests/golden/c2_grad_fd.npz), and against the same window without
    recompute;
  * batched identification: four uneven instances (different sides, grids,
    damping, E, nu) in one graph-replayed window against the oracle per
    instance; the full 64-instance batch (sides 32..96, 128^3 grids, 300
    substeps): clean status, finite, and three sampled instances equal to
    their single-instance runs;
  * the precision-stress cloth (96x96, 128^3, 500 substeps, compression and
    fold): positions at substeps 250 and 500 and the gradient against the
    oracle (fixture c4_positions.npz), both precision modes;
  * the large-cloth setting on a reduced cloth (256^3 grid, per-triangle
    moduli, sqrt(S) = 10-substep checkpoints), with a dense grid and with a
    pool of active blocks: per-triangle moduli gradients and a directional
    derivative against the oracle (fixture c5_reduced.npz);
  * a reference-shaped scene: a 15^3 grid (inv_dx not a power of two),
    scene_scaling 0.5 and 300-substep frames: bit-exact binning keys,
    positions after 320 and 620 substeps (past the frame boundaries; 2e-5:
    FP32 against FP64 on the coarse grid drifts to 1.1e-5) and a window
    gradient against the oracle.

Tolerances: 1e-3 relative on window gradients against the oracle (the bar
of the path; parameter gradients compared as loss sensitivities with an FP32
floor of 1e-5 of the largest, see _params_close), 1e-4 on losses, 1e-5 on
positions, widened where a scene's FP32 floor (FP32 restatements against the
FP64 oracle) is measured above it, with the measurement beside the bar;
equal-kernel comparisons
(recompute vs resident, batched vs single) 1e-5 on losses and 1e-4 on
gradients (float-atomic summation order only).
"""
import os

import numpy as np
import pytest
import torch

from diffcp_amd import scene
from oracle import sim_torch as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BAD_STATUS = 1 | 2 | 16          # non-finite, out of grid, internal check


def _engine(insts, **kw):
    from diffcp_amd.engine import Engine
    return Engine(insts, **kw)


def _slots(S, K):
    from diffcp_amd.engine import Engine
    return Engine.slots_for(S, K)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


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


def test_plain_workload_gradient_with_checkpoint_recompute():
    ref = np.load(os.path.join(GOLDEN, "c2_grad_fd.npz"))
    w = scene.config_c2()
    S, K = int(ref["substeps"]), int(ref["checkpoint_every"])
    assert (S, K) == (w.substeps, w.checkpoint_every) == (400, 50)
    out = {}
    for k in (K, S):
        eng = _engine(w.instances, max_states=_slots(S, k))
        eng.set_targets([ref["targets"]])
        eng.reset(0)
        loss, grads, _ = eng.rollout_grad(S, 0, k, use_graph=True)
        assert eng.status() & BAD_STATUS == 0
        out[k] = (loss[0], grads[0, :2].copy())
        eng.close()
    loss, g = out[K]
    assert abs(loss - float(ref["loss"])) < 1e-4 * float(ref["loss"])
    for j, name in enumerate(("E", "nu")):
        assert abs(g[j] - ref["grad"][j]) < 1e-3 * abs(ref["grad"][j]), (name, g[j], ref["grad"][j])
        fd = ref["fd_1e-05"][j] if "fd_1e-05" in ref.files else ref["fd_0.0001"][j]
        assert abs(g[j] - fd) < 3e-3 * abs(fd), (name, g[j], fd)
    # recompute (8 segments of 50) against the resident window
    assert abs(out[K][0] - out[S][0]) < 1e-5 * abs(out[S][0])
    assert _rel(out[K][1], out[S][1]) < 1e-4


def _uneven_batch(S):
    specs = [(6, 32, 2e3, 0.15, 0.5), (9, 48, 8e3, 0.3, 2.0), (12, 64, 3e4, 0.4, 4.0), (8, 40, 5e3, 0.25, 1.0)]
    insts = []
    for b, (n, G, E, nu, damp) in enumerate(specs):
        insts.append(scene.make_instance(n, G, S, seed=20 + b, E=E, nu=nu, damping=damp, friction=0.3 + 0.1 * b,
                                         density=scene.HEAVY_DENSITY, name=f"uneven_{b}"))
    for b, inst in enumerate(insts):
        gt = O.make_params(inst)
        gt["E"] = torch.tensor(8e3, dtype=torch.float64)
        gt["nu"] = torch.tensor(0.25, dtype=torch.float64)
        with torch.no_grad():
            st, _ = O.rollout(inst, gt, S - 2)
        # moved by ~2 mm: the loss stays far above FP32 position resolution
        # even for an instance whose parameters are near the ground truth
        inst.targets = scene.sample_targets(st["vx"].numpy(), 48, 300 + b) + np.array([0.002, -0.001, 0.0015])
    return insts


def test_batched_uneven_instances_match_oracle():
    S = 40
    insts = _uneven_batch(S)
    eng = _engine(insts, max_states=_slots(S, S))
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=True)
    assert eng.status() & BAD_STATUS == 0
    for b, inst in enumerate(insts):
        p = O.make_params(inst, requires_grad=True)
        ref_loss, _ = O.tape_loss(inst, p, S, inst.targets)
        ref_loss.backward()
        assert abs(loss[b] - ref_loss.item()) <= 1e-4 * ref_loss.item(), b
        names = ("E", "nu", "ks", "kc", "damping")
        _params_close(grads[b, :5], [p[pn].grad.item() for pn in names], inst.params()[:5], rtol=1e-3)
        assert grads[b, 5] == 0.0          # friction has no effect without the live return mapping


def test_full_batch_is_clean_and_matches_single_instance_runs():
    w = scene.config_c3()
    S = w.substeps
    eng = _engine(w.instances, max_states=_slots(S, S))
    base = np.stack([i.params() for i in w.instances])
    gt = base.copy()
    gt[:, 0], gt[:, 1] = 8e3, 0.25
    eng.set_params(gt)
    eng.reset(0)
    slot, _ = eng.rollout(S - 2)
    vx = eng.split(eng.download_state(slot)["vertex_x"], "n_vertices")
    targets = [scene.sample_targets(v, 1024, w.seed * 100 + b) for b, v in enumerate(vx)]
    eng.set_targets(targets)
    eng.set_params(base)
    eng.reset(0)
    eng.status(clear=True)
    loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=True)
    assert eng.status() & BAD_STATUS == 0
    assert np.isfinite(loss).all() and np.isfinite(grads).all()
    assert (loss > 0).all()
    eng.close()
    for b in (0, 31, 63):
        one = _engine([w.instances[b]], max_states=_slots(S, S))
        one.set_targets([targets[b]])
        one.reset(0)
        l1, g1, _ = one.rollout_grad(S, 0, S, use_graph=True)
        assert abs(l1[0] - loss[b]) <= 1e-5 * abs(loss[b]), b
        assert _rel(g1[0, :5], grads[b, :5]) < 1e-4, (b, g1[0, :5], grads[b, :5])
        one.close()


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_precision_stress_positions_and_gradient(precision):
    path = os.path.join(GOLDEN, "c4_positions.npz")
    if not os.path.exists(path):
        pytest.skip("fixture c4_positions.npz not generated")
    ref = np.load(path)
    w = scene.config_c4()
    S = int(ref["substeps"])
    # two slots, ping-pong; the states after substeps 250 and 500 are read back
    eng = _engine(w.instances, max_states=2, precision=precision)
    slots = {}
    eng.reset(0)
    a, b = 0, 1
    for f in range(S):
        eng.substep(f, 0, a, b)
        if f + 1 in (250, 500):
            slots[f + 1] = eng.download_state(b)["vertex_x"]
        a, b = b, a
    assert eng.status() & BAD_STATUS == 0
    # the FP32 floor of this scene: the oracle's FP32 C++ restatement sits
    # 9.2e-6 and 1.38e-5 from the FP64 positions after 250 and 500 substeps
    for s, bar in ((250, 2e-5), (500, 3e-5)):
        err = _rel(slots[s], ref[f"x{s}"])
        assert err < bar, (s, err)
    eng.close()
    eng = _engine(w.instances, max_states=_slots(S, w.checkpoint_every), precision=precision)
    eng.set_targets([ref["targets"]])
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, w.checkpoint_every, use_graph=True)
    assert abs(loss[0] - float(ref["loss"])) < 1e-4 * float(ref["loss"])
    # FP32 floor of this window: the oracle's FP32 C++ restatement (forward-mode
    # tangents) sits 0.9e-3 (E) and 1.8e-3 (nu) from the FP64 differences; the
    # h = 1e-4 and 1e-5 differences agree to 7e-5
    fd = ref["fd_1e-05"] if "fd_1e-05" in ref.files else ref["fd_0.0001"]
    for j in (0, 1):
        assert abs(grads[0, j] - fd[j]) < 5e-3 * abs(fd[j]), (j, grads[0, j], fd[j])


@pytest.mark.parametrize("pool", [0, 2048])
def test_large_cloth_setting_per_triangle_moduli(pool):
    ref = np.load(os.path.join(GOLDEN, "c5_reduced.npz"))
    inst = scene.large_cloth_reduced_instance()
    S, K = int(ref["substeps"]), int(ref["checkpoint_every"])
    eng = _engine([inst], max_states=_slots(S, K), per_element_moduli=True, grid_pool_blocks=pool)
    eng.set_targets([ref["targets"]])
    eng.reset(0)
    loss, _, _ = eng.rollout_grad(S, 0, K, use_graph=True)
    assert eng.status() & BAD_STATUS == 0
    used, cap = eng.pool_blocks()
    print(f"pool {pool}: {used} of {cap} blocks handed out")
    assert cap == pool and (0 < used <= cap if pool else used == 0)
    ge = eng.element_moduli_grads()
    assert abs(loss[0] - float(ref["loss"])) < 1e-4 * float(ref["loss"])
    # per-triangle gradients as loss sensitivities theta g, within 2e-3 of the
    # largest one (the contact modulus acts only in normal compression and its
    # sensitivities sit orders below E's, at the FP32 limit)
    th = np.asarray(inst.element_moduli)
    sens, ref_sens = ge * th, ref["grad"] * th
    top = np.abs(ref_sens).max()
    for c in range(3):
        err = np.abs(sens[:, c] - ref_sens[:, c]).max() / top
        print(f"modulus {c}: max sensitivity {np.abs(ref_sens[:, c]).max() / top:.2e} of the largest, error {err:.2e}")
        assert err < 2e-3, (c, err)
    d = float((ge * th * ref["delta"]).sum())
    assert abs(d - float(ref["fd_dir"])) < 1e-3 * abs(float(ref["fd_dir"])), (d, float(ref["fd_dir"]))


def test_reference_shaped_scene():
    inst = scene.reference_shaped_instance()
    S = 620
    eng = _engine([inst], max_states=2)
    eng.reset(0)
    a, b = 0, 1
    xs = {}
    for f in range(S):
        eng.substep(f, 0, a, b)
        if f + 1 in (320, S):
            xs[f + 1] = eng.download_state(b)["vertex_x"]
        a, b = b, a
    assert eng.status() & BAD_STATUS == 0
    slot = a
    st = eng.download_state(slot)
    with torch.no_grad():
        ref, _ = O.rollout(inst, n_substeps=S, keep_states=True)
    # FP32 against FP64 on the coarse 1/15 grid: 1.1e-5 relative after 320
    # and after 620 substeps (measured on MI355X); bar 2e-5
    for s in (320, S):
        err = _rel(xs[s], ref[s]["vx"].numpy())
        print(f"reference-shaped scene: positions after {s} substeps {err:.2e} relative to the FP64 oracle")
        assert err < 2e-5, (s, err)
    # integer keys at inv_dx = 15: trunc(f32(x) * 15 - 0.5) from a rounded product
    keys = eng.binning_keys(slot)
    nv, no, nc = inst.n_vertices, inst.n_grasper, inst.n_cube
    assert (keys[:nv] == O.binning_keys_fp32(st["vertex_x"], 15)).all()
    assert (keys[nv:nv + no] == O.binning_keys_fp32(st["grasper_x"], 15)).all()
    assert (keys[nv + no:nv + no + nc] == O.binning_keys_fp32(st["cube_x"], 15)).all()
    # the FP32 restatement of the oracle: the GPU follows it far more closely
    with torch.no_grad():
        r32, _ = O.rollout(inst, n_substeps=S, keep_states=True, dtype=torch.float32)
    for s in (320, S):
        err = _rel(xs[s], r32[s]["vx"].numpy())
        print(f"reference-shaped scene: positions after {s} substeps {err:.2e} relative to the FP32 oracle")
        assert err < 2e-5, (s, err)


def test_reference_shaped_scene_window_gradient():
    """A 320-substep window across the first frame boundary against the
    oracle's autograd in FP32 (the same arithmetic precision: 1e-4) and in
    FP64 (3.1e-3 on dL/dnu between the two oracles themselves on this coarse
    grid: bar 5e-3), fixture tests/golden/reference_shaped.npz."""
    path = os.path.join(GOLDEN, "reference_shaped.npz")
    if not os.path.exists(path):
        pytest.skip("fixture reference_shaped.npz not generated")
    ref = np.load(path)
    inst = scene.reference_shaped_instance()
    Sg = int(ref["substeps"])
    eng = _engine([inst], max_states=_slots(Sg, 32))
    eng.set_targets([ref["targets"]])
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(Sg, 0, 32, use_graph=True)
    assert eng.status() & BAD_STATUS == 0
    assert abs(loss[0] - float(ref["loss_fp64"])) <= 1e-4 * float(ref["loss_fp64"])
    _params_close(grads[0, :4], ref["grad_fp32"], inst.params()[:4], rtol=1e-4)
    _params_close(grads[0, :4], ref["grad_fp64"], inst.params()[:4], rtol=5e-3)
