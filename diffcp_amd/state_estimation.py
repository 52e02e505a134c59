"""Forward-only state estimation on the batched library.

Restates the reference's no-grad animation (``animation`` in
src/simulation_without_grad.py:488-520, driven by src/state_estimation_main.py):
substeps in chunks of ``IFS = 1 ms``; after every chunk the element rest shapes
are re-derived from the current vertex positions (``reset_quadrature_particles``,
:472-486 -> ``dcp_rebase_elements``); at the chunks the reference picks,
``n_pose_estimation_output + 1`` garment meshes are written as PLY files.

The substep index runs on from 0 through the whole animation (the grasp
trajectory frame is ``f * dt / operator_update_interval``) with window 0.
"""
from __future__ import annotations

import os
from typing import Callable, List, Optional

import numpy as np

from .engine import Engine
from .scene_io import write_ply

IFS = 1e-3   # simulation_without_grad.py:503


def output_chunks(n_chunks: int, n_outputs: int) -> List[List[int]]:
    """For each chunk, the output indices i the reference writes after it:
    chunk == int(n_chunks * i / n_outputs - 2) (simulation_without_grad.py:512-514)."""
    out: List[List[int]] = [[] for _ in range(n_chunks)]
    for i in range(n_outputs + 1):
        c = int(n_chunks * i / n_outputs - 2)
        if 0 <= c < n_chunks:
            out[c].append(i)
    return out


def animate(engine: Engine, animation_time: float, ifs: float = IFS, g: int = 0, n_outputs: int = 0,
            out_dir: Optional[str] = None,
            on_output: Optional[Callable[[int, int, List[np.ndarray]], None]] = None) -> int:
    """Run the animation from the state in slot 0 (``Engine.reset``) and return
    the slot holding the final state.  Needs ``max_states >= 2``.

    ``out_dir``: write ``<out_dir>/<b>/<i>.ply`` (instance b, output i) with the
    instance's triangles; ``on_output(i, chunk, vertex_x_per_instance)`` is
    called for each output as well.
    """
    dts = {float(i.dt) for i in engine.instances}
    if len(dts) != 1:
        raise ValueError("instances of one animation must share dt")
    dt = dts.pop()
    per = int(ifs / dt)
    n_chunks = int(animation_time / ifs)
    if per < 1 or n_chunks < 1:
        raise ValueError("animation shorter than one chunk")
    outs = output_chunks(n_chunks, n_outputs) if (n_outputs and (out_dir or on_output)) else None
    a, b = 0, 1
    for count in range(n_chunks):
        for f in range(count * per, (count + 1) * per):
            engine.substep(f, g, a, b)
            a, b = b, a
        engine.rebase_elements(a)
        if outs and outs[count]:
            vx = engine.split(engine.download_state(a)["vertex_x"], "n_vertices")
            for i in outs[count]:
                if out_dir:
                    for bi, (inst, v) in enumerate(zip(engine.instances, vx)):
                        write_ply(os.path.join(out_dir, str(bi), f"{i}.ply"), v, inst.triangles)
                if on_output:
                    on_output(i, count, vx)
    return a
