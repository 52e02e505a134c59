"""Drop-in for the reference's simulation module on the HIP library.

Mirrors the names a diffcp caller uses (src/simulation.py,
src/taichi_variable_definition.py, src/parameter_fitting_functions.py:3-58):

    from diffcp_amd import simulation as sim
    sim.setup(instance, max_steps=S + 1)         # replaces the module globals of initialization.py
    sim.initialize_objects()
    sim.garment_E[None] = 5e3
    sim.target_pos.from_numpy(targets)
    with sim.Tape(loss=sim.loss):                # replaces ti.ad.Tape
        sim.loss[None] = 0
        for f in range(sim.max_steps - 1):
            sim.forward(f, g)
        sim.compute_loss(f)
    sim.garment_E.grad[None], sim.vertex_x.to_numpy()[sim.max_steps - 1]

Every state index f is a device state slot (all ``max_steps`` states are
kept, as the reference's time-indexed fields are).  ``forward(f, g)`` is one
substep f -> f+1 (simulation.py:480-488); ``backward(f, g)`` is its adjoint
(simulation.py:490-506) and is what ``Tape`` replays in reverse.  The module
functions act on the simulation made by the last ``setup`` call; several
scenes can coexist as ``Simulation`` objects.

Errors raise ``diffcp_amd.DiffcpError`` (the reference has no error path).
``clear_distances`` is a no-op: the loss kernel keeps no distance matrix.
"""
from __future__ import annotations

from typing import List, Optional, Tuple

import numpy as np

from . import scene as _scene
from .engine import Engine

PARAM_FIELDS = ("garment_E", "garment_nu", "shearing_stiffness", "contact_stiffness", "damping", "friction")


class _Grad:
    def __init__(self, owner: "_Scalar"):
        self._o = owner

    def __getitem__(self, key):
        return self._o._grad

    def __setitem__(self, key, value):
        self._o._grad = float(value)


class _Scalar:
    """A 0-d field: ``x[None]`` reads, ``x[None] = v`` writes, ``x.grad[None]``."""

    def __init__(self, sim: "Simulation", index: Optional[int], value: float = 0.0):
        self._sim, self._index = sim, index
        self._value = float(value)
        self._grad = 0.0
        self.grad = _Grad(self)

    def __getitem__(self, key):
        return self._value

    def __setitem__(self, key, value):
        self._value = float(value)
        if self._index is not None:
            self._sim._params_dirty = True

    def __float__(self):
        return self._value


class _StateField:
    """A time-indexed particle field (vertex_x, operator_x, cube_x, ...):
    ``to_numpy()`` gives [max_steps, n, 3] (or [max_steps, n, 3, 3] for C)."""

    def __init__(self, sim: "Simulation", lib_name: str):
        self._sim, self._name = sim, lib_name

    def __getitem__(self, f: int) -> np.ndarray:
        return self._sim._download(f)[self._name]

    def to_numpy(self) -> np.ndarray:
        return np.stack([self[f] for f in range(self._sim.max_steps)])


class _HostField:
    """A host-side input field (initial_vertex_x, target_pos, ...)."""

    def __init__(self, sim: "Simulation", on_set=None, value=None):
        self._sim, self._on_set = sim, on_set
        self._value = None if value is None else np.array(value)

    def from_numpy(self, arr) -> None:
        self._value = np.array(arr)
        if self._on_set is not None:
            self._on_set(self._value)

    def to_numpy(self) -> np.ndarray:
        return None if self._value is None else self._value.copy()


class _TrajField(_HostField):
    """grasp_tracking_seq: ``from_numpy`` uploads, ``.grad`` is the gradient
    (gradient_descent_for_traj_generation.py:24)."""

    class _G:
        def __init__(self, sim):
            self._sim = sim

        def to_numpy(self) -> np.ndarray:
            return self._sim.engine.trajectory_grads()[0]

    def __init__(self, sim, value):
        super().__init__(sim, on_set=lambda v: sim.engine.set_trajectory([v]), value=value)
        self.grad = _TrajField._G(sim)


class Tape:
    """Records ``forward`` / ``compute_loss`` calls of one simulation and, on
    exit, runs the adjoint in reverse with d loss / d loss = 1 (what
    ``ti.ad.Tape(loss=loss)`` does), leaving the parameter gradients in the
    fields' ``.grad[None]``."""

    def __init__(self, loss: _Scalar):
        self.loss = loss
        self.sim = loss._sim

    def __enter__(self):
        sim = self.sim
        sim._tape = []
        sim.engine.zero_param_grads()
        for name in PARAM_FIELDS:
            getattr(sim, name)._grad = 0.0
        return self

    def __exit__(self, exc_type, exc, tb):
        sim = self.sim
        ops, sim._tape = sim._tape, None
        if exc_type is not None:
            return False
        sim._replay_adjoint(ops)
        grads = sim.engine.param_grads()[0]
        for j, name in enumerate(PARAM_FIELDS):
            getattr(sim, name)._grad = float(grads[j])
        return False


class Simulation:
    """One cloth scene on the HIP library with the reference's field names."""

    def __init__(self, instance: "_scene.Instance", max_steps: Optional[int] = None, **engine_options):
        self.instance = instance
        # default: one window of the grasp trajectory (frames_per_window frames)
        self.max_steps = int(max_steps if max_steps is not None
                             else instance.frames_per_window * instance.steps_per_frame + 1)
        if self.max_steps < 3:
            raise ValueError("max_steps must be >= 3")
        self.engine = Engine([instance], max_states=self.max_steps, **engine_options)
        self._tape: Optional[List[Tuple]] = None
        self._params_dirty = True
        self._cache = {}
        p = instance.params()
        for j, name in enumerate(PARAM_FIELDS):
            setattr(self, name, _Scalar(self, j, p[j]))
        self.loss = _Scalar(self, None)
        self.vertex_x = _StateField(self, "vertex_x")
        self.vertex_v = _StateField(self, "vertex_v")
        self.vertex_C = _StateField(self, "vertex_C")
        self.operator_x = _StateField(self, "grasper_x")
        self.operator_v = _StateField(self, "grasper_v")
        self.operator_C = _StateField(self, "grasper_C")
        self.cube_x = _StateField(self, "cube_x")
        self.cube_v = _StateField(self, "cube_v")
        self.cube_C = _StateField(self, "cube_C")
        self.initial_vertex_x = _HostField(self, value=instance.vertex_x)
        self.initial_operator_x = _HostField(self, value=instance.grasper_x)
        self.initial_cube_x = _HostField(self, value=instance.cube_x)
        self.lagrangian_mesh_element = _HostField(self, on_set=self._check_topology, value=instance.triangles)
        self.target_pos = _HostField(self, on_set=lambda t: self.engine.set_targets([t]))
        self.grasp_tracking_seq = _TrajField(self, instance.trajectory)
        self.n_vertex_particles = instance.n_vertices
        self.n_quadrature_particles = instance.n_elements
        self.n_operator_particles = instance.n_grasper
        self.n_cube_particles = instance.n_cube
        self.n_grid = instance.n_grid
        self.dt = instance.dt

    # ---------------------------------------------------------------- helpers
    def _check_topology(self, tri) -> None:
        if not np.array_equal(np.asarray(tri, np.int64), np.asarray(self.instance.triangles, np.int64)):
            raise ValueError("lagrangian_mesh_element differs from the mesh the simulation was built with")

    def _sync_params(self) -> None:
        if self._params_dirty:
            self.engine.set_params(np.array([[getattr(self, n)[None] for n in PARAM_FIELDS]]))
            self._params_dirty = False

    def _download(self, f: int):
        if not 0 <= f < self.max_steps:
            raise IndexError(f"state index {f} outside [0, {self.max_steps})")
        if f not in self._cache:
            self._cache[f] = self.engine.download_state(f)
        return self._cache[f]

    # -------------------------------------------------------- reference names
    def initialize_ti_field(self) -> None:
        """simulation.py:447-477: state 0 from the initial_* fields (rest shape
        re-based on initial_vertex_x, F = I, v = C = 0)."""
        self._cache.clear()
        self.engine.reset(0, vertex_x=[self.initial_vertex_x.to_numpy()],
                          grasper_x=[self.initial_operator_x.to_numpy()],
                          cube_x=[self.initial_cube_x.to_numpy()])

    def initialize_objects(self) -> None:
        """simulation.py:432-445: the scene's creation data into the initial
        fields and the trajectory, then initialize_ti_field()."""
        inst = self.instance
        self.grasp_tracking_seq.from_numpy(inst.trajectory)
        self.initial_operator_x.from_numpy(inst.grasper_x)
        self.initial_vertex_x.from_numpy(inst.vertex_x)
        self.lagrangian_mesh_element.from_numpy(inst.triangles)
        self.initial_cube_x.from_numpy(inst.cube_x)
        self.initialize_ti_field()

    def forward(self, s: int, grad_update: int = 0) -> None:
        """One substep, state s -> s+1 (simulation.py:480-488)."""
        if not 0 <= s < self.max_steps - 1:
            raise IndexError(f"substep {s} outside [0, {self.max_steps - 1})")
        self._sync_params()
        self._cache.pop(s + 1, None)
        self.engine.substep(s, grad_update, s, s + 1)
        if self._tape is not None:
            self._tape.append(("forward", s, grad_update))

    def backward(self, s: int, grad_update: int = 0, gslot_in: int = 0, gslot_out: int = 1) -> None:
        """Adjoint of substep s (simulation.py:490-506): d loss / d state s+1 in
        adjoint slot gslot_in -> d loss / d state s in gslot_out, parameter
        gradients accumulated."""
        self._sync_params()
        self.engine.substep_adjoint(s, grad_update, s, gslot_in, gslot_out)

    def clear_distances(self) -> None:
        """simulation.py:3-7 (nothing to clear: no distance matrix is stored)."""

    def compute_loss(self, f: int) -> float:
        """parameter_fitting_functions.py:3-15: loss[None] += mean over targets
        of the min over vertices of |x - t|^2 / 3, on vertex_x[f - 1]."""
        if self.target_pos.to_numpy() is None:
            raise RuntimeError("target_pos is not set")
        val = float(self.engine.loss(f - 1)[0])
        self.loss[None] = self.loss[None] + val
        if self._tape is not None:
            self._tape.append(("loss", f - 1))
        return val

    def _replay_adjoint(self, ops) -> None:
        fwd = [op for op in ops if op[0] == "forward"]
        loss_states = [op[1] for op in ops if op[0] == "loss"]
        if not fwd and not loss_states:
            return
        steps = [op[1] for op in fwd]
        if steps != list(range(len(steps))):
            raise RuntimeError("Tape expects forward(0), forward(1), ... in order")
        T = len(steps)
        cur, other = 0, 1
        self.engine.zero_adjoint(cur)
        for t in loss_states:
            if t == T:
                self.engine.loss_adjoint(t, cur, seed=1.0)
        for s in reversed(range(T)):
            self.backward(s, fwd[s][2], cur, other)
            cur, other = other, cur
            for t in loss_states:
                if t == s:
                    self.engine.loss_adjoint(t, cur, seed=1.0)


_current: Optional[Simulation] = None


def setup(instance: "_scene.Instance", max_steps: Optional[int] = None, **engine_options) -> Simulation:
    """Build the scene the module-level names act on (the role of the
    reference's initialization.py globals)."""
    global _current
    if _current is not None:
        _current.engine.close()
    _current = Simulation(instance, max_steps, **engine_options)
    return _current


def current() -> Simulation:
    if _current is None:
        raise RuntimeError("call diffcp_amd.simulation.setup(instance, max_steps) first")
    return _current


_FUNCS = ("initialize_objects", "initialize_ti_field", "forward", "backward", "clear_distances", "compute_loss")
_FIELDS = PARAM_FIELDS + ("loss", "vertex_x", "vertex_v", "vertex_C", "operator_x", "operator_v", "operator_C",
                          "cube_x", "cube_v", "cube_C", "initial_vertex_x", "initial_operator_x", "initial_cube_x",
                          "lagrangian_mesh_element", "target_pos", "grasp_tracking_seq", "max_steps",
                          "n_vertex_particles", "n_quadrature_particles", "n_operator_particles",
                          "n_cube_particles", "n_grid", "dt")
__all__ = ["Simulation", "Tape", "setup", "current"] + list(_FUNCS) + list(_FIELDS)


def __getattr__(name):
    if name in _FUNCS or name in _FIELDS:
        return getattr(current(), name)
    raise AttributeError(name)
