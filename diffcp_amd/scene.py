"""Synthetic cloth-manipulation scenes with the reference's mass/volume rules.

A scene instance is one garment (vertex particles + one element particle per
triangle), a two-point grasper, a heavy rigid cube and a grasp trajectory, on
its own uniform MPM grid.  The derived constants follow
``src/initialization.py:113-131`` (masses, volumes, density) with
``scene_scaling = 1`` and ``bound``/``gravity`` from ``initialization.py:21-23``.

Synthetic layout (all lengths in metres, domain is the unit cube):
  * cloth: n x n vertex square of side ``side`` in the x-z plane, centred at
    x = z = 0.5, one cell above the cube top; every quad split by the same
    diagonal so that every triangle normal points +y;
  * cube: side 0.04, centred under the cloth, bottom at y = 2 dx (the
    reference puts the scene bottom there, initialization.py:178);
  * grasper: ``n_grasper_per_point`` particles per grasp point, uniform in a
    0.015 m cube around two adjacent cloth corners, same offsets for every
    grasp point (initialization.py:163-169, simulation.py:433-437);
  * trajectory: seeded smooth Hermite cubic (zero end velocities) lifting the
    corners by ``lift`` plus a small lateral jitter, sampled every
    ``steps_per_frame`` substeps.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

GRAVITY = -9.81          # initialization.py:23
BOUND = 3                # initialization.py:22
N_PARAMS = 6             # E, nu, shearing, contact, damping, friction
PARAM_NAMES = ("E", "nu", "shearing_stiffness", "contact_stiffness", "damping", "friction")


@dataclasses.dataclass
class Instance:
    """One cloth scene; arrays are float64 host copies (uploaded as FP32)."""
    n_grid: int
    dt: float
    vertex_x: np.ndarray          # [Nv, 3]
    triangles: np.ndarray         # [Ne, 3] int32
    cube_x: np.ndarray            # [Nc, 3]
    grasper_x: np.ndarray         # [No, 3]
    trajectory: np.ndarray        # [n_actions, n_grasp, 3]
    steps_per_frame: int          # operator_update_interval / dt
    frames_per_window: int        # int(grad_update_interval / operator_update_interval)
    garment_mass: float
    cube_mass: float = 999.0      # simulation_arguments.json:18
    operator_mass: float = 99.0   # simulation_arguments.json:14
    scene_scaling: float = 1.0
    E: float = 5e3
    nu: float = 0.3
    shearing_stiffness: float = 100.0   # simulation_arguments.json:38-43 initial values
    contact_stiffness: float = 100.0
    damping: float = 0.0
    friction: float = 0.0
    element_moduli: Optional[np.ndarray] = None   # [Ne, 3] per-triangle (E, shear, contact)
    targets: Optional[np.ndarray] = None          # [Nt, 3]
    name: str = ""
    area: Optional[float] = None                  # surface area override (None: from the mesh)

    # ---- derived constants (initialization.py:79, 113-131) ----
    @property
    def dx(self) -> float:
        return 1.0 / self.n_grid

    @property
    def inv_dx(self) -> float:
        return float(self.n_grid)

    @property
    def n_vertices(self) -> int:
        return int(self.vertex_x.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.triangles.shape[0])

    @property
    def n_cube(self) -> int:
        return int(self.cube_x.shape[0])

    @property
    def n_grasper(self) -> int:
        return int(self.grasper_x.shape[0])

    @property
    def n_grasp_points(self) -> int:
        return int(self.trajectory.shape[1])

    @property
    def n_actions(self) -> int:
        return int(self.trajectory.shape[0])

    @property
    def grasper_group_size(self) -> int:
        # initialization.py:115,164
        return self.n_grasper // self.n_grasp_points

    @property
    def surface_area(self) -> float:
        if self.area is not None:
            return float(self.area)
        x = self.vertex_x
        t = self.triangles
        e1 = x[t[:, 1]] - x[t[:, 0]]
        e2 = x[t[:, 2]] - x[t[:, 0]]
        return float(0.5 * np.linalg.norm(np.cross(e1, e2), axis=1).sum())

    @property
    def garment_rho(self) -> float:
        return self.garment_mass / (self.surface_area * self.dx)          # :118

    @property
    def quadrature_p_vol(self) -> float:
        if self.n_elements == 0:
            return 0.0
        return 3.0 * self.surface_area * self.dx / self.n_elements        # :127

    @property
    def vertex_p_vol(self) -> float:
        return self.surface_area * self.dx / self.n_vertices              # :128

    @property
    def quadrature_p_mass(self) -> float:
        return self.garment_mass / self.n_vertices                        # :129

    @property
    def vertex_p_mass(self) -> float:
        return self.vertex_p_vol * self.garment_rho                       # :130

    @property
    def cube_p_mass(self) -> float:
        return self.cube_mass / self.n_cube if self.n_cube else 0.0       # :123

    @property
    def operator_p_mass(self) -> float:
        return self.operator_mass / self.n_grasper if self.n_grasper else 0.0   # :125

    @property
    def gravity(self) -> float:
        return self.scene_scaling * GRAVITY                               # simulation.py:182

    def params(self) -> np.ndarray:
        return np.array([self.E, self.nu, self.shearing_stiffness, self.contact_stiffness,
                         self.damping, self.friction], dtype=np.float64)

    def set_params(self, p) -> None:
        (self.E, self.nu, self.shearing_stiffness, self.contact_stiffness,
         self.damping, self.friction) = [float(v) for v in p]

    def copy(self) -> "Instance":
        return dataclasses.replace(self)

    @property
    def n_cloth_particles(self) -> int:
        return self.n_vertices + self.n_elements


# ----------------------------------------------------------------------------
# builders
# ----------------------------------------------------------------------------

def square_cloth(n: int, side: float, cx: float, cz: float, y: float):
    """n x n vertex square in the x-z plane; triangles with +y normals."""
    h = side / (n - 1)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = cx - side / 2 + i * h
    z = cz - side / 2 + j * h
    verts = np.stack([x, np.full_like(x, y, dtype=np.float64), z], -1).reshape(-1, 3)
    ii, jj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (ii * n + jj).ravel()
    v01 = v00 + 1
    v10 = v00 + n
    v11 = v10 + 1
    ta = np.stack([v00, v01, v11], 1)
    tb = np.stack([v00, v11, v10], 1)
    tris = np.stack([ta, tb], 1).reshape(-1, 3).astype(np.int32)
    return verts.astype(np.float64), tris


def hermite_lift(p0: np.ndarray, delta: np.ndarray, n_frames: int, frame_dt: float,
                 total_time: float) -> np.ndarray:
    """Smooth cubic p0 + delta * (3 tau^2 - 2 tau^3), zero end velocities."""
    t = np.arange(n_frames) * frame_dt
    tau = np.clip(t / total_time, 0.0, 1.0)
    s = 3 * tau ** 2 - 2 * tau ** 3
    return p0[None, :] + s[:, None] * delta[None, :]


def make_instance(n: int, n_grid: int, substeps: int, *, seed: int, dt: float = 1e-4,
                  side: float = 0.2, density: float = 0.2, lift: float = 0.1,
                  jitter: float = 0.005, cube_side: float = 0.04,
                  n_cube: int = 150, n_grasper_per_point: int = 125,
                  steps_per_frame: int = 10, E: float = 5e3, nu: float = 0.3,
                  shearing: float = 100.0, contact: float = 100.0,
                  damping: float = 0.0, friction: float = 0.0,
                  compress: float = 0.0, fold: bool = False,
                  name: str = "") -> Instance:
    rng = np.random.default_rng(seed)
    dx = 1.0 / n_grid
    y0 = 2 * dx
    cloth_y = y0 + cube_side + dx
    verts, tris = square_cloth(n, side, 0.5, 0.5, cloth_y)
    cube = rng.random((n_cube, 3)) * cube_side
    cube[:, 0] += 0.5 - cube_side / 2
    cube[:, 1] += y0
    cube[:, 2] += 0.5 - cube_side / 2
    edge = 0.015
    offs = (rng.random((n_grasper_per_point, 3)) - 0.5) * edge
    corners = np.stack([verts[0], verts[n - 1]])            # (i=0,j=0) and (i=0,j=n-1)
    grasper = np.concatenate([offs + c for c in corners], 0)
    n_frames = substeps // steps_per_frame + 2
    frame_dt = steps_per_frame * dt
    total = substeps * dt
    trajs = []
    for c in corners:
        jit = (rng.random(3) - 0.5) * 2 * jitter
        jit[1] = 0.0
        if fold:
            # compress in-plane towards +x, then lift over the cube and lay down
            t = np.arange(n_frames) * frame_dt
            tau = np.clip(t / total, 0.0, 1.0)
            half = np.clip(tau * 2, 0, 1)
            s1 = 3 * half ** 2 - 2 * half ** 3
            rest = np.clip(tau * 2 - 1, 0, 1)
            s2 = 3 * rest ** 2 - 2 * rest ** 3
            p = np.repeat(c[None], n_frames, 0).astype(np.float64)
            p[:, 0] += compress * side * s1 + (side * 0.45) * s2
            p[:, 1] += lift * np.sin(np.pi * s2)
            p += jit[None] * tau[:, None]
            trajs.append(p)
        else:
            delta = np.array([0.0, lift, 0.0]) + jit
            delta[0] += compress * side
            trajs.append(hermite_lift(c, delta, n_frames, frame_dt, total))
    traj = np.stack(trajs, 1)
    area = side * side
    inst = Instance(n_grid=n_grid, dt=dt, vertex_x=verts, triangles=tris, cube_x=cube,
                    grasper_x=grasper, trajectory=traj, steps_per_frame=steps_per_frame,
                    frames_per_window=max(1, substeps // steps_per_frame),
                    garment_mass=density * area, E=E, nu=nu, shearing_stiffness=shearing,
                    contact_stiffness=contact, damping=damping, friction=friction, name=name)
    return inst


def sample_targets(vertex_x: np.ndarray, n_targets: int, seed: int) -> np.ndarray:
    """Seeded sample without replacement of vertex positions (the loss state)."""
    rng = np.random.default_rng(seed + 7919)
    nv = vertex_x.shape[0]
    idx = rng.choice(nv, size=min(n_targets, nv), replace=False)
    return np.asarray(vertex_x[idx], dtype=np.float64)


# ----------------------------------------------------------------------------
# the five workload configurations (C1..C5)
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class Workload:
    name: str
    instances: List[Instance]
    substeps: int
    checkpoint_every: int
    grad_params: tuple
    seed: int
    per_element_moduli: bool = False
    n_targets: int = 1024
    gt_params: Optional[tuple] = None     # ground-truth (E, nu) for target generation
    grid_pool_blocks: int = 0             # > 0: grids share a pool of this many 4^3 blocks

    @property
    def particles(self) -> int:
        return sum(i.n_cloth_particles for i in self.instances)


def config_c1(seed: int = 0) -> Workload:
    """32x32-vertex cloth, grid 64^3, 200 substeps, E=5e3, nu=0.3, friction 0.5."""
    inst = make_instance(32, 64, 200, seed=seed, E=5e3, nu=0.3, friction=0.5, name="c1")
    return Workload("c1_plain_cpu", [inst], 200, 200, ("E", "nu"), seed, gt_params=(8e3, 0.25))


# Areal density of the C2-C5 cloths (unstated for them; C1 states 0.2 kg/m^2).
# With E up to 1e5 and nu up to 0.45 on a 1/128 grid at dt = 1e-4, 1.0 kg/m^2
# keeps the elastic CFL number c dt / dx below 0.5 (0.2 kg/m^2 reaches ~0.9
# and some seeded candidates diverge).
HEAVY_DENSITY = 1.0


def config_c2(seed: int = 1) -> Workload:
    """64x64-vertex cloth, grid 128^3, 400 substeps, checkpoints every 50."""
    rng = np.random.default_rng(seed)
    E = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
    nu = float(rng.uniform(0.1, 0.45))
    inst = make_instance(64, 128, 400, seed=seed, E=E, nu=nu, density=HEAVY_DENSITY, name="c2")
    return Workload("c2_plain_gpu", [inst], 400, 50, ("E", "nu"), seed, gt_params=(8e3, 0.25))


def config_c3(seed: int = 2, n_instances: int = 64, substeps: int = 300) -> Workload:
    """64 instances, side 32..96 vertices, own 128^3 grids, per-instance E/nu/damping/friction."""
    rng = np.random.default_rng(seed)
    insts = []
    for b in range(n_instances):
        n = int(rng.integers(32, 97))
        E = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        nu = float(rng.uniform(0.1, 0.45))
        damping = float(rng.uniform(0.0, 5.0))
        friction = float(rng.uniform(0.1, 1.0))
        insts.append(make_instance(n, 128, substeps, seed=seed * 1000 + b, E=E, nu=nu, density=HEAVY_DENSITY,
                                   damping=damping, friction=friction, name=f"c3_{b}"))
    return Workload("c3_batched", insts, substeps, substeps,
                    ("E", "nu", "damping", "friction"), seed, gt_params=(8e3, 0.25))


def config_c4(seed: int = 3) -> Workload:
    """96x96-vertex cloth compressed 40% in-plane then folded over the cube, 500 substeps."""
    inst = make_instance(96, 128, 500, seed=seed, E=5e3, nu=0.3, compress=0.4, fold=True, density=HEAVY_DENSITY,
                         lift=0.08, name="c4")
    return Workload("c4_precision", [inst], 500, 50, ("E", "nu"), seed, gt_params=(8e3, 0.25))


# Grid block pool of the large cloth (only active blocks allocated): 4^3
# blocks, against 64^3 = 262,144 for a dense 256^3 grid.
C5_POOL_BLOCKS = 4096


def config_c5(seed: int = 4, substeps: int = 1000) -> Workload:
    """256x256-vertex cloth on 256^3, per-triangle lognormal moduli, sqrt(S) checkpoints."""
    inst = make_instance(256, 256, substeps, seed=seed, E=5e3, nu=0.3, density=HEAVY_DENSITY, name="c5")
    rng = np.random.default_rng(seed + 17)
    base = np.array([inst.E, inst.shearing_stiffness, inst.contact_stiffness])
    inst.element_moduli = base[None, :] * np.exp(rng.normal(0.0, 0.3, size=(inst.n_elements, 3)))
    k = max(1, int(round(math.sqrt(substeps))))
    return Workload("c5_large", [inst], substeps, k, ("element_moduli",), seed,
                    per_element_moduli=True, gt_params=(8e3, 0.25), grid_pool_blocks=C5_POOL_BLOCKS)


def precision_stress_instance(n: int = 24, n_grid: int = 64, substeps: int = 120, seed: int = 3) -> Instance:
    """The precision-stress trajectory of config_c4 on a smaller cloth: 40 %
    in-plane compression, then a fold over the cube (in-plane singular values
    down to ~0.1 with near-equal pairs)."""
    return make_instance(n, n_grid, substeps + 2, seed=seed, E=5e3, nu=0.3, compress=0.4, fold=True,
                         density=HEAVY_DENSITY, lift=0.08, name="c4_small")


def large_cloth_reduced_instance(seed: int = 4, substeps: int = 100) -> Instance:
    """The large-cloth workload's setting (256^3 grid, per-triangle lognormal
    (E, shearing, contact) with sigma 0.3, areal density 1.0) on a 40x40 cloth
    of side 0.05 m (vertex spacing 0.33 dx) over a 0.02 m cube."""
    inst = make_instance(40, 256, substeps, seed=seed, side=0.05, cube_side=0.02, density=HEAVY_DENSITY,
                         lift=0.02, name="c5_reduced")
    rng = np.random.default_rng(seed + 17)
    base = np.array([inst.E, inst.shearing_stiffness, inst.contact_stiffness])
    inst.element_moduli = base[None, :] * np.exp(rng.normal(0.0, 0.3, size=(inst.n_elements, 3)))
    return inst


def reference_shaped_instance(substeps: int = 620, seed: int = 5) -> Instance:
    """A synthetic scene shaped like the reference's own (simulation_arguments.json,
    initialization.py:79-131): a 15^3 grid (dx = 1/15, not a power of two),
    scene_scaling 0.5 (gravity halved), frames of 300 substeps
    (operator_update_interval 0.03 s at dt = 1e-4), 150 cube and 2 x 125
    grasper particles; a 15x15 cloth of side 0.4 m."""
    inst = make_instance(15, 15, substeps, seed=seed, side=0.4, cube_side=0.15, density=HEAVY_DENSITY,
                         lift=0.05, steps_per_frame=300, name="reference_shaped")
    inst.scene_scaling = 0.5
    return inst


CONFIGS = {"c1": config_c1, "c2": config_c2, "c3": config_c3, "c4": config_c4, "c5": config_c5}


def tiny_instance(n: int = 4, n_grid: int = 16, substeps: int = 8, seed: int = 0, **kw) -> Instance:
    """A few-vertex scene for oracle/kernel unit tests (fits a 16^3 grid)."""
    kw.setdefault("side", 0.25)
    kw.setdefault("lift", 0.02)
    kw.setdefault("n_cube", 20)
    kw.setdefault("n_grasper_per_point", 8)
    kw.setdefault("cube_side", 0.1)
    kw.setdefault("steps_per_frame", 2)
    return make_instance(n, n_grid, substeps, seed=seed, **kw)


def wide_grasp_instance(n_grid, spread_xz, y_lo, y_hi):
    """Tiny scene whose two grasp points each have two particles moved to far
    corners (x, z -+spread_xz, y y_lo / y_hi): each grasp point's patch tile
    exceeds the 1024-cell LDS tile, while its other six particles keep holding
    the cloth corner."""
    inst = tiny_instance(n=5, n_grid=n_grid, substeps=40)
    ng = inst.trajectory.shape[1]
    gs = len(inst.grasper_x) // ng
    gx = inst.grasper_x.copy()
    for g in range(ng):
        c = gx[g * gs:(g + 1) * gs].mean(0)
        gx[g * gs] = c + np.array([-spread_xz, y_lo, -spread_xz])
        gx[g * gs + 1] = c + np.array([spread_xz, y_hi, spread_xz])
    gx[:, [0, 2]] = np.clip(gx[:, [0, 2]], 0.15, 0.85)
    inst.grasper_x = gx
    rng = np.random.default_rng(7)
    inst.targets = inst.vertex_x[rng.choice(inst.n_vertices, 12, replace=False)] + rng.normal(0, 0.01, (12, 3))
    return inst
