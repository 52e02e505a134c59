"""Windowed fitting on the HIP path: the carry-over of compute_grad against the
substep-by-substep drop-in module (Tape windows with initial_*.from_numpy +
initialize_ti_field, parameter_fitting_functions.py:20-49) and against the
FP64 oracle window by window, and a short fit that must lower the loss.

The two paths give the same bits, both with deterministic scatter sums
(dcp_options.deterministic): the same kernels run on the same inputs.  (With float-atomic sums, whose order varies from run to run, a
last-bit change can move one of the 16 targets' nearest vertex on this
36-vertex cloth, since the loss takes the first minimum, which shifted the
nu gradient by about 1.2e-5 relative in either path.)"""
import numpy as np
import pytest

from diffcp_amd import scene

pytestmark = pytest.mark.gpu

WINDOW = 20


def _instance(E=5e3):
    inst = scene.tiny_instance(n=6, n_grid=16, substeps=2 * WINDOW, E=E)
    inst.frames_per_window = WINDOW // inst.steps_per_frame
    return inst


def _targets(inst, seed):
    rng = np.random.default_rng(seed)
    return inst.vertex_x[rng.choice(inst.n_vertices, 16, replace=False)] + rng.normal(0, 0.01, (16, 3))


def test_compute_grad_carry_over_matches_tape_windows():
    from diffcp_amd import simulation as sim
    from diffcp_amd.engine import Engine
    from diffcp_amd.fitting import WindowedFit

    inst = _instance()
    tg = [_targets(inst, 1), _targets(inst, 2)]
    eng = Engine([inst], max_states=Engine.slots_for(WINDOW, WINDOW), deterministic=True)
    fit = WindowedFit(eng, [[tg[0]], [tg[1]]], WINDOW)
    fit.initialize_objects()
    loss, grads = fit.compute_grad()

    s = sim.setup(inst, max_steps=WINDOW + 1, deterministic=True)
    sim.initialize_objects()
    losses, gE = [], []
    for g in range(2):
        sim.target_pos.from_numpy(tg[g])
        with sim.Tape(loss=sim.loss):
            sim.loss[None] = 0
            for f in range(sim.max_steps - 1):
                sim.forward(f, g)
            sim.compute_loss(f)
        losses.append(sim.loss[None])
        gE.append([sim.garment_E.grad[None], sim.garment_nu.grad[None]])
        sim.initial_vertex_x.from_numpy(sim.vertex_x[sim.max_steps - 1])
        sim.initial_operator_x.from_numpy(sim.operator_x[sim.max_steps - 1])
        sim.initial_cube_x.from_numpy(sim.cube_x[sim.max_steps - 1])
        sim.initialize_ti_field()
    ref = np.mean(gE, axis=0)
    assert loss[0] == np.mean(losses)
    assert np.array_equal(grads[0, :2], ref)
    s.engine.close()


def test_compute_grad_carry_over_matches_oracle():
    """Two instances (different E, nu, damping), two windows: window 1 starts
    from window 0's final positions with zero velocities and a fresh element
    state (initial_*_x.from_numpy + initialize_ti_field,
    parameter_fitting_functions.py:40-49) and its own grasp trajectory slice.
    The FP64 oracle repeats the same: autograd per window, the carried
    positions constants.  Bars: 1e-4 on the mean loss, 1e-3 on the mean
    E and nu gradients."""
    import torch

    from diffcp_amd.engine import Engine
    from diffcp_amd.fitting import WindowedFit
    from oracle import sim_torch as O

    insts = [_instance(E=5e3), _instance(E=9e3)]
    insts[1].nu, insts[1].damping = 0.35, 1.5
    tg = [[_targets(insts[0], 1), _targets(insts[1], 3)], [_targets(insts[0], 2), _targets(insts[1], 4)]]
    eng = Engine(insts, max_states=Engine.slots_for(WINDOW, WINDOW))
    fit = WindowedFit(eng, tg, WINDOW)
    fit.initialize_objects()
    loss, grads = fit.compute_grad()
    eng.close()
    for b, inst in enumerate(insts):
        cur = inst.copy()
        # particle volumes and masses stay those of the creation mesh
        # (initialization.py:113-131 runs once), not of the carried positions
        cur.area = inst.surface_area
        losses, gs = [], []
        for g in range(2):
            params = O.make_params(cur, requires_grad=True)
            l, states = O.tape_loss(cur, params, WINDOW, tg[g][b], g=g)
            l.backward()
            losses.append(l.item())
            gs.append([params["E"].grad.item(), params["nu"].grad.item()])
            last = states[WINDOW]
            cur = cur.copy()
            cur.vertex_x = last["vx"].detach().numpy().copy()
            cur.grasper_x = last["ox"].detach().numpy().copy()
            cur.cube_x = last["cx"].detach().numpy().copy()
        ref_l, ref_g = np.mean(losses), np.mean(gs, axis=0)
        assert abs(loss[b] - ref_l) <= 1e-4 * abs(ref_l), (b, loss[b], ref_l)
        # per component, with a floor of 1e-4 of the larger loss sensitivity
        # (theta dL/dtheta) for a mean that cancels between the windows
        theta = np.array([inst.E, inst.nu])
        floor = 1e-4 * np.abs(ref_g * theta).max() / theta
        for j in (0, 1):
            assert abs(grads[b, j] - ref_g[j]) <= 1e-3 * abs(ref_g[j]) + floor[j], (b, j, grads[b, j], ref_g[j])


def test_short_fit_moves_E_toward_the_ground_truth():
    """32x32 cloth, 200 substeps, one window: targets from a run at E = 8e3;
    candidates start at 6.5e3 and 9.5e3 with E bounded to [6e3, 1e4]."""
    from diffcp_amd.engine import Engine
    from diffcp_amd.fitting import WindowedFit

    w = scene.config_c1()
    S = w.substeps
    gt = w.instances[0].copy()
    gt.E = 8e3
    eng_gt = Engine([gt], max_states=3)
    eng_gt.reset(0)
    slot, _ = eng_gt.rollout(S - 2)
    tgt = scene.sample_targets(eng_gt.download_state(slot)["vertex_x"], 1024, 0)
    eng_gt.close()
    insts = [w.instances[0].copy(), w.instances[0].copy()]
    insts[0].E, insts[1].E = 6.5e3, 9.5e3
    eng = Engine(insts, max_states=Engine.slots_for(S, S))
    fit = WindowedFit(eng, [[tgt, tgt]], S)
    params = np.stack([i.params() for i in insts])
    res = fit.parameters_learning(params, epochs=5, bounds={"E": (6e3, 1e4), "nu": (0.28, 0.32)})
    for b in range(2):
        best = int(np.argmin([l[b] for l in res.losses]))
        assert res.losses[best][b] < res.losses[0][b]
        assert abs(res.history[best][b, 0] - 8e3) < abs(params[b, 0] - 8e3)
