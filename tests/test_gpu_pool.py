"""Grids from a pool of 4^3 blocks (dcp_options.grid_pool_blocks) through the
C-ABI: the same window as with dense grids, the pool's use reported and a
full pool flagged.

  * a small cloth, forward states and a graph-replayed window (loss and
    parameter gradients) with a pooled grid against the dense one: the same
    kernels on the same blocks, so only the float-atomic summation order of
    the tile flushes differs (1e-6 on positions, 1e-5 on the loss, 1e-4 on
    gradients);
  * the pool is handed out afresh at every reset and rollout, so a second
    window reports the same number of blocks;
  * a pool too small for the scene sets DCP_STATUS_POOL_FULL (its blocks
    share the trash slot: every access stays inside the pool) and the run
    completes.
"""
import numpy as np
import pytest

from diffcp_amd import scene

pytestmark = pytest.mark.gpu
POOL_FULL = 8


def _pair(**kw):
    from diffcp_amd.engine import Engine
    inst = scene.make_instance(12, 32, 60, seed=11, side=0.2, lift=0.03, name="pool")
    S = 40
    engs = [Engine([inst], max_states=Engine.slots_for(S, S), grid_pool_blocks=pool, **kw) for pool in (0, 512)]
    return inst, S, engs


def test_pooled_grid_matches_dense_grid():
    inst, S, (dense, pooled) = _pair()
    targets = scene.sample_targets(inst.vertex_x, 64, 3) + np.array([0.0, 0.01, 0.0])
    out = []
    for eng in (dense, pooled):
        eng.set_targets([targets])
        eng.reset(0)
        slot, _ = eng.rollout(S - 2)
        st = eng.download_state(slot)
        eng.reset(0)
        loss, grads, _ = eng.rollout_grad(S, 0, S, use_graph=True)
        out.append((st, loss, grads, eng.status()))
    (sd, ld, gd, std), (sp, lp, gp, stp) = out
    assert std == 0 and stp == 0
    for k in ("vertex_x", "element_C"):
        rel = np.abs(sp[k] - sd[k]).max() / max(np.abs(sd[k]).max(), 1e-30)
        assert rel < 1e-6, (k, rel)
    assert abs(lp[0] - ld[0]) <= 1e-5 * abs(ld[0])
    top = np.abs(gd[0, :4]).max()
    assert np.abs(gp[0, :4] - gd[0, :4]).max() <= 1e-4 * top, (gp[0], gd[0])
    used, cap = pooled.pool_blocks()
    assert cap == 512 and 0 < used <= cap
    # handed out afresh: a second window reports the same count
    pooled.reset(0)
    pooled.rollout_grad(S, 0, S, use_graph=True)
    assert pooled.pool_blocks() == (used, cap)
    assert dense.pool_blocks() == (0, 0)
    for eng in (dense, pooled):
        eng.close()


def test_full_pool_is_reported():
    from diffcp_amd.engine import Engine
    inst = scene.make_instance(12, 32, 60, seed=11, side=0.2, lift=0.03, name="pool")
    eng = Engine([inst], max_states=3, grid_pool_blocks=4)
    eng.reset(0)
    eng.status(clear=True)
    eng.rollout(10)
    assert eng.status() & POOL_FULL
    used, cap = eng.pool_blocks()
    assert cap == 4 and used > cap
    eng.close()
