"""GPU: forward-only animation with rest-shape re-bases (dcp_rebase_elements)
against the oracle's reset_quadrature_particles loop (FP64).  Tolerances as
the forward parity tests: 2e-5 on positions, 5e-3 on v and C (relative to
each field's largest magnitude)."""
import numpy as np
import pytest

from diffcp_amd import scene
from oracle import reduced as R
from oracle import sim_torch as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_animation_with_rebases_matches_oracle(tmp_path):
    from diffcp_amd.engine import Engine
    from diffcp_amd.state_estimation import animate

    inst = scene.tiny_instance(n=5, n_grid=16, substeps=20)
    per, chunks = 4, 3
    eng = Engine([inst], max_states=2)
    eng.reset(0)
    seen = []
    final = animate(eng, animation_time=chunks * per * inst.dt, ifs=per * inst.dt, n_outputs=2,
                    out_dir=str(tmp_path), on_output=lambda i, c, vx: seen.append((i, c)))
    assert eng.status() & 3 == 0
    gpu = eng.download_state(final)

    st, static = O.initialize_state(inst)
    params = O.make_params(inst)
    K = O.Consts.from_instance(inst)
    prev = st
    for c in range(chunks):
        for f in range(c * per, (c + 1) * per):
            prev, st = st, O.substep(st, static, params, K, f)
        st, static = O.reset_quadrature_particles(st, static)
    ref = R.reduce(st, prev)
    for name in ("vertex_x", "grasper_x", "cube_x"):
        assert _rel(gpu[name], ref[R.LIB_NAMES[name]].numpy()) < 2e-5, name
    for name in ("vertex_v", "vertex_C", "element_C"):
        assert _rel(gpu[name], ref[R.LIB_NAMES[name]].numpy()) < 5e-3, name
    # outputs after chunks int(3 i / 2 - 2) for i = 0..2 (truncation toward 0):
    # i = 1 after chunk 0, i = 2 after chunk 1
    assert seen == [(1, 0), (2, 1)]
    assert (tmp_path / "0" / "1.ply").exists() and (tmp_path / "0" / "2.ply").exists()


def test_rollout_refuses_pending_rebase():
    from diffcp_amd._abi import DiffcpError
    from diffcp_amd.engine import Engine

    inst = scene.tiny_instance(n=5, n_grid=16, substeps=8)
    inst.targets = inst.vertex_x[:4].copy()
    eng = Engine([inst], max_states=Engine.slots_for(6, 6))
    eng.reset(0)
    eng.rebase_elements(0)
    with pytest.raises(DiffcpError):
        eng.rollout_grad(6, 0, 6)
    eng.reset(0)                               # clears the pending re-base
    loss, grads, _ = eng.rollout_grad(6, 0, 6)
    assert np.isfinite(loss).all()
