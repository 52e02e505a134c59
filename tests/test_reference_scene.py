"""The reference's real-data scene built from its own files (its garment PLY,
cube mark, grasp trajectory and target point clouds), checked against the
quantities initialization.py derives from them, and run for a few substeps
through the FP64 oracle.  Runs where the reference checkout is present (its
data are not copied into this repository); skipped elsewhere."""
import os

import numpy as np
import pytest
import torch

REF_SRC = os.path.join(os.environ.get("DIFFCP_REFERENCE", "/root/reference"), "src")
pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "simulation_arguments.json")),
                                reason="reference checkout not present")


@pytest.fixture(scope="module")
def ref():
    from diffcp_amd.reference_scene import load_reference_scene
    return load_reference_scene(REF_SRC, seed=0)


def test_counts_and_windows(ref):
    i = ref.instance
    assert (i.n_vertices, i.n_elements) == (427, 777)          # cloth_mesh.ply header
    assert (i.n_cube, i.n_grasper, i.n_grasp_points) == (150, 250, 2)
    assert i.n_actions == 1942                                  # end_frame 9000 clipped to the data
    assert (i.steps_per_frame, i.frames_per_window, ref.max_steps) == (300, 100, 30001)
    assert len(ref.targets) == 8 and ref.targets[0].shape == (31, 3)
    assert i.n_grid == 15 and i.dt == 1e-4


def test_derived_constants(ref):
    i = ref.instance
    s = 0.5
    assert i.garment_mass == pytest.approx(0.5 * s ** 3)
    assert i.quadrature_p_mass == pytest.approx(i.garment_mass / 427)                  # :129
    assert i.vertex_p_mass == pytest.approx(i.surface_area * i.dx / 427 * i.garment_mass / (i.surface_area * i.dx))
    assert i.cube_p_mass == pytest.approx(999 / 150) and i.operator_p_mass == pytest.approx(99 / 250)
    # scene placement: x-z centre of garment + trajectory at 0.5 (initialization.py:171-178)
    pts = np.concatenate((i.vertex_x, i.trajectory.reshape(-1, 3)))
    c = (pts.max(0) + pts.min(0)) / 2
    assert abs(c[0] - 0.5) < 1e-12 and abs(c[2] - 0.5) < 1e-12
    # grasper particles: one offset set per grasp point, around the first frame's points
    g = i.grasper_x.reshape(2, 125, 3)
    assert np.allclose(g[0] - i.trajectory[0, 0], g[1] - i.trajectory[0, 1])
    assert np.abs(g[0] - i.trajectory[0, 0]).max() <= 0.03 * 0.5 / 2 + 1e-12


def test_stencils_inside_grid_and_oracle_runs(ref):
    from diffcp_amd.reference_scene import stencils_inside_grid
    from oracle import sim_torch as O
    i = ref.instance
    assert stencils_inside_grid(i)
    st, _ = O.rollout(i, n_substeps=5)
    assert bool(torch.isfinite(st["vx"]).all())
    assert float((st["vx"] - torch.as_tensor(i.vertex_x)).abs().max()) < 1e-3
    loss, _ = O.compute_loss(st["vx"], torch.as_tensor(ref.targets[0]))
    assert np.isfinite(float(loss)) and float(loss) > 0
