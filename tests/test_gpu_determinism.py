"""Bitwise run-to-run reproducibility through the C-ABI (dcp_options.deterministic).

The scatters (P2G and the G2P adjoint) add tile nodes into the grid with
atomics; in the deterministic form the sums are kept in two-word fixed point,
so they do not depend on the order in which workgroups arrive, and the
per-wave damping and trajectory-gradient partial sums are added the same way.  Every other
reduction runs in a fixed order (stable binning, in-order run and node sums,
per-(patch, wave) parameter-gradient slots summed on the host in order).  So
the same window gives the same bits:

  * an uneven batch (four instances, different sides, grids, damping, E,
    nu): loss, parameter gradients, trajectory gradients and the final
    states of two windows in one context and one window in a second context
    are equal bit for bit;
  * a scene whose patch tiles overflow (the sub-box scatter kernels) and a
    pooled grid give equal bits across two windows;
  * pooled and dense grids give the same bits (the same integers reach the
    same nodes), and so does the live return mapping across two windows;
  * the eager launch sequence and the graph replay of the same window agree
    to the last bit as well (the same kernels on the same inputs);
  * the default float-atomic form gives the same window within its
    summation-order noise (1e-5 on losses, 1e-4 on gradients, 1e-6 m on
    positions).
"""
import numpy as np
import pytest

from diffcp_amd import scene

pytestmark = pytest.mark.gpu


def _batch(S):
    insts = []
    for b, (n, G, damp, E, nu) in enumerate([(10, 32, 0.0, 5e3, 0.3), (14, 48, 2.0, 2e4, 0.2),
                                             (8, 32, 4.0, 1.5e3, 0.4), (12, 64, 1.0, 8e4, 0.35)]):
        inst = scene.make_instance(n, G, S + 2, seed=40 + b, side=0.2 * 32 / G * 2, lift=0.03, E=E, nu=nu,
                                   damping=damp, name=f"det{b}")
        inst.targets = scene.sample_targets(inst.vertex_x, 24, 7 + b) + np.array([0.003, 0.002, -0.001])
        insts.append(inst)
    return insts


def _window(eng, S, use_graph=True):
    eng.reset(0)
    loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=use_graph)
    traj = np.concatenate([t.ravel() for t in eng.trajectory_grads()])
    st = eng.download_state(eng.last_final_slot())
    return loss.copy(), grads.copy(), traj, st


def _same(a, b, what):
    la, ga, ta, sa = a
    lb, gb, tb, sb = b
    assert np.array_equal(la, lb), (what, "loss", la - lb)
    assert np.array_equal(ga, gb), (what, "grads", ga - gb)
    assert np.array_equal(ta, tb), (what, "trajectory grads")
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), (what, k, np.abs(sa[k] - sb[k]).max())


def test_same_window_same_bits():
    from diffcp_amd.engine import Engine
    S = 40
    insts = _batch(S)
    engs = [Engine(insts, max_states=Engine.slots_for(S, S), deterministic=True) for _ in range(2)]
    for e in engs:
        e.set_targets([i.targets for i in insts])
    first = _window(engs[0], S)
    assert engs[0].status() == 0
    assert np.isfinite(first[1]).all() and np.abs(first[1][:, :2]).max() > 0
    _same(first, _window(engs[0], S), "second window, same context")
    _same(first, _window(engs[1], S), "another context")
    _same(first, _window(engs[1], S, use_graph=False), "eager launches")
    for e in engs:
        e.close()
    # the float-atomic form: the same window within its summation-order noise
    fast = Engine(insts, max_states=Engine.slots_for(S, S))
    fast.set_targets([i.targets for i in insts])
    lf, gf, tf, sf = _window(fast, S)
    fast.close()
    assert np.abs(lf - first[0]).max() <= 1e-5 * np.abs(first[0]).max()
    top = np.abs(first[1][:, :5]).max(axis=1, keepdims=True)
    assert (np.abs(gf[:, :5] - first[1][:, :5]) <= 1e-4 * top).all(), (gf[:, :5], first[1][:, :5])
    assert np.abs(sf["vertex_x"] - first[3]["vertex_x"]).max() <= 1e-6


@pytest.mark.parametrize("pool", [0, 1024])
def test_overflowing_tiles_same_bits(pool):
    """Two grasper particles per grasp point moved to far corners of a 48^3
    grid: their patch tiles exceed the 1024-cell LDS tile and the sub-box
    kernels scatter them (fixed-point flushes as well)."""
    from diffcp_amd.engine import Engine
    S = 30
    inst = scene.wide_grasp_instance(48, 0.4, -0.15, 0.25)
    gs = len(inst.grasper_x) // inst.trajectory.shape[1]
    base = np.trunc(inst.grasper_x[:gs] * inst.n_grid - 0.5)
    assert np.prod(base.max(0) - base.min(0) + 3) > 1024
    eng = Engine([inst], max_states=Engine.slots_for(S, S), grid_pool_blocks=pool, deterministic=True)
    eng.set_targets([inst.targets])
    a = _window(eng, S)
    _same(a, _window(eng, S), "overflow tiles")
    assert eng.status() & 23 == 0
    eng.close()


def test_pooled_and_dense_grids_give_the_same_bits():
    """With fixed-point sums the pooled grid (four instances sharing one
    block pool) and the dense grids add the same integers into their nodes:
    the window gives the same bits either way."""
    from diffcp_amd.engine import Engine
    S = 30
    insts = _batch(S)
    outs = []
    for pool in (0, 2048):
        eng = Engine(insts, max_states=Engine.slots_for(S, S), grid_pool_blocks=pool, deterministic=True)
        eng.set_targets([i.targets for i in insts])
        outs.append(_window(eng, S))
        assert eng.status() == 0
        if pool:
            used, cap = eng.pool_blocks()
            assert 0 < used <= cap
        eng.close()
    _same(outs[0], outs[1], "pooled vs dense")


def test_return_mapping_same_bits():
    """The live return mapping (friction gradients) in the deterministic form."""
    from diffcp_amd.engine import Engine
    S = 30
    insts = _batch(S)[:2]
    for inst in insts:
        inst.friction = 0.4
    eng = Engine(insts, max_states=Engine.slots_for(S, S), return_mapping=True, deterministic=True)
    eng.set_targets([i.targets for i in insts])
    a = _window(eng, S)
    assert np.abs(a[1][:, 5]).max() > 0          # dL/dfriction is live
    _same(a, _window(eng, S), "return mapping")
    eng.close()
