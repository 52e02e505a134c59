"""Ragged and degenerate batches through the C-ABI: one batch holding an
instance without cube particles, one without grasper particles and a
two-triangle cloth (classes with no particles, a patch of two elements),
each against the FP64 oracle; and an instance with no targets beside one
with targets (its loss and gradients are zero, the other's equal its
single-instance run).

Tolerances: positions 1e-5 relative, losses 1e-4, equal-kernel comparisons
1e-5.  Parameter gradients are compared as loss sensitivities against the
FP64 oracle with a bar of 2e-3 relative plus twice the FP32 oracle's own
deviation from the FP64 one for that component (plus 1e-5 of the largest
sensitivity): on these tiny scenes FP32 sets the floor of some components —
the cube-less instance's dL/dE differs by 6.4e-3 between the two oracles, and
the grasper-less cloth falls freely and does not stretch, so its elastic
gradients differ by 1-20 % (its damping gradient by 4e-5)."""
import numpy as np
import pytest

from diffcp_amd import scene
from oracle import sim_torch as O

pytestmark = pytest.mark.gpu
BAD_STATUS = 1 | 2 | 4 | 16
PARAMS = ("E", "nu", "ks", "kc", "damping")


def _engine(insts, S, **kw):
    from diffcp_amd.engine import Engine
    eng = Engine(insts, max_states=Engine.slots_for(S, S), **kw)
    eng.reset(0)
    return eng


def _targets(inst, seed, k=6):
    rng = np.random.default_rng(seed)
    idx = rng.choice(inst.n_vertices, min(k, inst.n_vertices), replace=False)
    return inst.vertex_x[idx] + rng.normal(0, 0.005, (len(idx), 3))


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _params_close(g, ref, ref32, theta, rtol=2e-3, floor=1e-5):
    th = np.maximum(np.abs(np.asarray(theta, np.float64)), 1.0)
    g, ref, ref32 = (np.asarray(a, np.float64) * th for a in (g, ref, ref32))
    bar = rtol * np.abs(ref) + 2 * np.abs(ref32 - ref) + floor * np.abs(ref).max()
    assert (np.abs(g - ref) <= bar).all(), (g / th, ref / th, ref32 / th)


def _oracle_grads(inst, S, dtype):
    import torch
    params = O.make_params(inst, dtype=dtype, requires_grad=True)
    lo, states = O.tape_loss(inst, params, S, inst.targets, dtype=dtype)
    lo.backward()
    return lo.item(), np.array([params[k].grad.item() for k in PARAMS]), states[S - 2]["vx"].detach().double().numpy()


def _ragged():
    a = scene.tiny_instance(n=6, n_grid=16, substeps=40, seed=5, n_cube=0)
    b = scene.tiny_instance(n=6, n_grid=16, substeps=40, seed=6, n_grasper_per_point=0)
    c = scene.tiny_instance(n=2, n_grid=16, substeps=40, seed=7)
    for i, inst in enumerate((a, b, c)):
        inst.targets = _targets(inst, 30 + i)
    assert a.n_cube == 0 and b.n_grasper == 0 and c.n_elements == 2
    return [a, b, c]


def test_ragged_batch_forward_and_gradients_match_oracle():
    import torch
    insts = _ragged()
    S = 20
    eng = _engine(insts, S)
    slot, _ = eng.rollout(S - 2)
    vx = eng.split(eng.download_state(slot)["vertex_x"], "n_vertices")
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=True)
    assert eng.status() & BAD_STATUS == 0
    for b, inst in enumerate(insts):
        lo, ref, x_ref = _oracle_grads(inst, S, torch.float64)
        _, ref32, _ = _oracle_grads(inst, S, torch.float32)
        assert _rel(vx[b], x_ref) < 1e-5, (b, _rel(vx[b], x_ref))
        assert abs(loss[b] - lo) < 1e-4 * lo, (b, loss[b], lo)
        _params_close(grads[b, :5], ref, ref32, inst.params()[:5])
    # the instance without graspers has no trajectory gradient; the others do
    tg = eng.trajectory_grads()
    assert np.abs(tg[1]).max() == 0 and np.abs(tg[0]).max() > 0


def test_instance_without_targets_beside_one_with_targets():
    a = scene.tiny_instance(n=6, n_grid=16, substeps=40, seed=8)
    b = scene.tiny_instance(n=5, n_grid=16, substeps=40, seed=9)
    a.targets = _targets(a, 40)
    S = 20
    eng = _engine([a, b], S)
    eng.set_targets([a.targets, np.zeros((0, 3))])
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=True)
    assert loss[1] == 0.0 and np.abs(grads[1]).max() == 0.0
    one = _engine([a], S)
    l1, g1, _ = one.rollout_grad(S, 0, S, use_graph=True)
    assert _rel(loss[0], l1[0]) < 1e-5 and _rel(grads[0, :5], g1[0, :5]) < 1e-5
