"""Forward-only state estimation pieces that need no GPU: PLY round trips, the
oracle's reset_quadrature_particles, and the reference's output schedule."""
import os

import numpy as np
import torch

from diffcp_amd import scene
from diffcp_amd.scene_io import read_ply, write_ply
from diffcp_amd.state_estimation import output_chunks
from oracle import sim_torch as O


def test_ply_binary_round_trip(tmp_path):
    inst = scene.tiny_instance(n=4, n_grid=16, substeps=4)
    p = str(tmp_path / "m.ply")
    write_ply(p, inst.vertex_x, inst.triangles)
    v, t = read_ply(p)
    assert np.array_equal(v, inst.vertex_x) and np.array_equal(t, inst.triangles)


def test_ply_ascii_reader(tmp_path):
    p = tmp_path / "a.ply"
    p.write_text("ply\nformat ascii 1.0\ncomment hand written\nelement vertex 3\nproperty float x\n"
                 "property float y\nproperty float z\nproperty uchar red\nelement face 1\n"
                 "property list uchar int vertex_indices\nend_header\n"
                 "0 0 0 1\n1 0 0 2\n0 1 0.5 3\n3 0 2 1\n")
    v, t = read_ply(str(p))
    assert v.shape == (3, 3) and v[2, 2] == 0.5
    assert t.tolist() == [[0, 2, 1]]


def test_reset_at_rest_is_identity():
    """Re-basing the initial state reproduces initialize_ti_field's d and Dinv."""
    inst = scene.tiny_instance(n=5, n_grid=16, substeps=4)
    st, static = O.initialize_state(inst)
    st2, static2 = O.reset_quadrature_particles(st, static)
    assert torch.allclose(st2["qd"], st["qd"], rtol=0, atol=0)
    assert torch.allclose(static2["Dinv"], static["Dinv"], rtol=0, atol=0)
    assert torch.equal(st2["qF"], st["qF"]) and torch.equal(st2["qv"], st["qv"])


def test_reset_after_motion_rebases_rest_shape():
    """After some substeps, the re-based Dinv maps the current edges to the
    rest-frame identity: d(x) @ Dinv = I, while F, v and C are kept."""
    inst = scene.tiny_instance(n=5, n_grid=16, substeps=12)
    st, static = O.rollout(inst, n_substeps=6)
    st2, static2 = O.reset_quadrature_particles(st, static)
    eye = torch.eye(3, dtype=st2["qd"].dtype).expand_as(st2["qd"])
    assert torch.allclose(st2["qd"] @ static2["Dinv"], eye, atol=1e-10)
    assert not torch.allclose(static2["Dinv"], static["Dinv"])
    for k in ("qF", "qv", "qC", "vx", "vv"):
        assert torch.equal(st2[k], st[k]), k
    tri = static["tri"]
    cen = (st["vx"][tri[:, 0]] + st["vx"][tri[:, 1]] + st["vx"][tri[:, 2]]) / 3
    assert torch.allclose(st2["qx"], cen)


def test_output_schedule_matches_reference_formula():
    n_chunks, n_out = 1941 * 30, 35            # the reference animation: 58.2 s of 1 ms chunks
    outs = output_chunks(n_chunks, n_out)
    got = sorted(i for lst in outs for i in lst)
    want = [i for i in range(n_out + 1) if 0 <= int(n_chunks * i / n_out - 2) < n_chunks]
    assert got == want and len(got) == n_out    # i = 0 maps to chunk -2: never written
    assert outs[int(n_chunks * 7 / n_out - 2)] == [7]
