"""The reference's real-data scene: its garment mesh, rigid cube, grasp
trajectory and target point clouds, built into an ``Instance`` the way
src/initialization.py:27-182 builds its module globals.

The data stay where the reference keeps them: ``load_reference_scene`` reads
``simulation_arguments.json`` from the reference's ``src`` directory and the
files under its ``data_dir`` (``cloth_mesh.ply``,
``rigid_cube_top_face_center_position.npy``, ``grasp_traj_array.npy``,
``target_pcd_sample.npy``); nothing is copied.  ``.npy`` files are read with
``allow_pickle=False``.

The reference draws the cube's and the grasper's particle offsets from the
unseeded ``np.random`` / ``random`` generators (initialization.py:84,
163-169); here they come from a seeded generator, in the same arrangement.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import Dict, List, Optional

import numpy as np

from .scene import Instance
from .scene_io import read_ply


@dataclasses.dataclass
class ReferenceScene:
    instance: Instance
    targets: List[np.ndarray]        # per gradient window g: [n_points, 3] (initialization.py:181)
    max_steps: int                   # substeps of a window + 1 (initialization.py:190)
    lower_bounds: Dict[str, float]   # constitutive_parameter_lower_bounds
    upper_bounds: Dict[str, float]
    arguments: Dict


def load_reference_scene(src_dir: str, seed: int = 0, arguments: Optional[Dict] = None) -> ReferenceScene:
    """Scene of simulation_arguments.json + data files (initialization.py:27-182)."""
    if arguments is None:
        with open(os.path.join(src_dir, "simulation_arguments.json")) as f:
            arguments = json.load(f)
    a = arguments
    dt, n_grid, s = float(a["dt"]), int(a["n_grid"]), float(a["scene_scaling"])
    op, cube, garment = a["operator_parameters"], a["rigid_cube_parameters"], a["garment_parameters"]
    dx = 1.0 / n_grid
    data_dir = src_dir + a["data_dir"]                                   # :71
    load = lambda name: np.load(os.path.join(data_dir, name), allow_pickle=False)
    targets = np.asarray(load("target_pcd_sample.npy"), np.float64)     # [windows, points, 3]
    mark = np.asarray(load("rigid_cube_top_face_center_position.npy"), np.float64)
    grasp = np.asarray(load("grasp_traj_array.npy"), np.float64)        # [frames, grasp points, 3]
    rng = np.random.default_rng(seed)

    # rigid cube particles below the marked top-face centre (:84-88)
    L = float(cube["rigid_cube_side_length"])
    n_cube = int(cube["n_rigid_cube_particles"])
    cube_x = rng.random((n_cube, 3)) * L - L / 2
    cube_x[:, 0] += mark[0]
    cube_x[:, 1] += mark[1] - L / 2
    cube_x[:, 2] += mark[2]

    # frame window (:91-100)
    n_actions = grasp.shape[0]
    initial_frame, end_frame = int(a["initial_frame"]), int(a["end_frame"])
    if end_frame >= n_actions or end_frame == -1:
        end_frame = n_actions - 1
    if initial_frame > end_frame:
        initial_frame = end_frame - 1
    grasp = grasp[initial_frame:end_frame + 1]
    n_grasp = grasp.shape[1]

    # scaling (:106-111) and derived counts (:113-116)
    verts, tris = read_ply(os.path.join(data_dir, "cloth_mesh.ply"))
    verts = verts * s
    grasp = grasp * s
    cube_x = cube_x * s
    targets = targets * s
    n_op = int(int(op["n_operator_particles"]) / n_grasp) * n_grasp
    edge = float(op["operator_edge_length"]) * s                        # :119

    # grasper particle offsets, one set shared by every grasp point (:163-169)
    group = n_op // n_grasp
    offsets = (rng.random((group, 3)) - 0.5) * edge

    # scene placement: bottom centre of garment + trajectory at (0.5, 2 dx, 0.5) (:171-182)
    scene_pts = np.concatenate((verts, grasp.reshape(-1, 3)))
    centre = (scene_pts.max(axis=0) - scene_pts.min(axis=0)) / 2 + scene_pts.min(axis=0)
    shift = np.array([0.5, 2 * dx, 0.5]) - np.array([centre[0], 0.0, centre[2]])
    verts = verts + shift
    grasp = grasp + shift
    cube_x = cube_x + shift
    targets = targets + shift

    # grasper particles around the first frame's grasp points (simulation.py:433-437)
    grasper_x = np.concatenate([offsets + grasp[0, g] for g in range(n_grasp)])

    init = garment["initial_constitutive_parameters"]
    interval = float(a["operator_update_interval"])
    inst = Instance(
        n_grid=n_grid, dt=dt, vertex_x=verts, triangles=tris, cube_x=cube_x, grasper_x=grasper_x,
        trajectory=grasp,
        # frame f * dt / interval in integers (exact frame boundaries)
        steps_per_frame=int(round(interval / dt)),
        frames_per_window=int(float(a["grad_update_interval"]) / interval),
        garment_mass=float(garment["garment_mass"]) * s ** 3,                  # :69
        cube_mass=float(cube["cube_mass"]), operator_mass=float(op["operator_mass"]),
        scene_scaling=s, E=float(init["E"]), nu=float(init["nu"]),
        shearing_stiffness=float(init["shearing_stiffness"]),
        contact_stiffness=float(init["contact_stiffness"]), name="reference",
    )
    return ReferenceScene(
        instance=inst, targets=[targets[g] for g in range(targets.shape[0])],
        max_steps=int(float(a["grad_update_interval"]) / dt) + 1,
        lower_bounds=dict(garment["constitutive_parameter_lower_bounds"]),
        upper_bounds=dict(garment["constitutive_parameter_upper_bounds"]), arguments=a)


def stencils_inside_grid(inst: Instance) -> bool:
    """Every particle's (and every grasp point's) 3x3x3 stencil inside the grid:
    base cell trunc(x inv_dx - 0.5) in [0, n_grid - 3] (simulation.py:140)."""
    pts = np.concatenate((inst.vertex_x, inst.cube_x, inst.grasper_x, inst.trajectory.reshape(-1, 3)))
    base = np.trunc(pts * inst.inv_dx - 0.5)
    return bool(((base >= 0) & (base <= inst.n_grid - 3)).all())
