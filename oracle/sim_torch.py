"""CPU restatement of DiffCP's differentiable codim-MPM cloth substep.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  Nothing in ``diffcp_amd/`` imports or calls it, and the product path never
routes through it.

It follows the reference op for op (same stencil order, the element scatter
into its own stencil plus the three vertex stencils, the dead return mapping
left out because it never writes back, the d <- F @ inverse(Dinv) round trip,
the loss read two substeps before the end).  Every function cites the
reference file:line it restates.  Arithmetic runs in float64 by default
(``dtype=torch.float32`` gives the FP32 variant).  Gradients come from torch
autograd through this restatement, which is what Taichi's tape computes
(source-transform AD of the same kernels), with the reference's hand-written
2x2 SVD backward (simulation.py:44-94) reproduced literally in a custom
autograd Function (``svd_grad="reference"``) or replaced by the exact polar
derivative (``svd_grad="stable"``).

The grid is stored sparsely (only cells some stencil touches, found with
``torch.unique``) so that autograd over hundreds of substeps fits in memory;
the values are those of the reference's dense grid at those cells, and
untouched cells of the dense grid hold v = 0/(0+1e-9) = 0 which no particle
reads.

Pins (tests/test_oracle.py, tests/test_oracle_pins.py): closed forms and
invariants of the B-spline weights, Gram-Schmidt QR vs numpy, the 2x2 polar vs
numpy SVD, the literal SVD backward vs finite differences away from
degeneracy and its exact zero at sigma1 = sigma2; hand-derived closed forms
of ``element_stress`` (upper-triangular F with shear and normal compression,
non-symmetric in-plane F against numpy's SVD and inverse, frame
indifference); the element force vectors and the normal impulse isolated on
a one-triangle scene (per-stencil sums, first and second moments);
``grid_op`` wall by wall (both sides of every wall, the floor, corners,
damping); ``g2p_gather`` on affine velocity fields; ``damp_element_C``;
``update_elements`` (rest, translation, third-order rigid rotation, round
trip); ``grasper_velocity`` at frame and window boundaries and past the last
frame; P2G mass/momentum conservation, free fall, the brute-force loss, and
central finite differences of whole rollouts in float64.  A mutation check
plants a sign, index or transpose error in each of these functions in memory
and shows that its pin rejects it.

Parity unpinned (Taichi is not installed, no import possible): Taichi's FP32
codegen order (FMA contraction, fast-math division, including the FP32
frame index int(f*dt/interval) of simulation.py:160, replaced here by the
exact integer f // steps_per_frame), ti.svd's internal small-off-diagonal
branch, float-atomic summation order.
"""
from __future__ import annotations

import dataclasses
import itertools
from typing import Dict, Optional

import numpy as np
import torch

OFFSETS = torch.tensor(list(itertools.product(range(3), range(3), range(3))), dtype=torch.int64)  # [27,3]
STATE_KEYS = ("vx", "vv", "vC", "ox", "ov", "oC", "cx", "cv", "cC", "qx", "qv", "qC", "qF", "qd")


@dataclasses.dataclass
class Consts:
    """Per-instance constants (initialization.py:21-23,79,113-131)."""
    G: int
    dt: float
    dx: float
    inv_dx: float
    bound: int
    gravity: float
    vertex_p_mass: float
    quadrature_p_mass: float
    quadrature_p_vol: float
    cube_p_mass: float
    operator_p_mass: float
    steps_per_frame: int
    frames_per_window: int
    group_size: int
    n_actions: int

    @staticmethod
    def from_instance(inst) -> "Consts":
        from diffcp_amd.scene import BOUND
        return Consts(G=inst.n_grid, dt=inst.dt, dx=inst.dx, inv_dx=inst.inv_dx, bound=BOUND,
                      gravity=inst.gravity, vertex_p_mass=inst.vertex_p_mass,
                      quadrature_p_mass=inst.quadrature_p_mass,
                      quadrature_p_vol=inst.quadrature_p_vol, cube_p_mass=inst.cube_p_mass,
                      operator_p_mass=inst.operator_p_mass,
                      steps_per_frame=inst.steps_per_frame,
                      frames_per_window=inst.frames_per_window,
                      group_size=inst.grasper_group_size, n_actions=inst.n_actions)


# ----------------------------------------------------------------------------
# small matrix helpers (Taichi Matrix semantics)
# ----------------------------------------------------------------------------

def tnorm(v: torch.Tensor, eps: float) -> torch.Tensor:
    """Taichi ``Vector.norm(eps)`` = sqrt(v.v + eps)."""
    return torch.sqrt((v * v).sum(-1) + eps)


def inv3(m: torch.Tensor) -> torch.Tensor:
    """General 3x3 inverse by cofactors (Taichi ``Matrix.inverse``)."""
    a = m
    c00 = a[..., 1, 1] * a[..., 2, 2] - a[..., 1, 2] * a[..., 2, 1]
    c01 = a[..., 1, 2] * a[..., 2, 0] - a[..., 1, 0] * a[..., 2, 2]
    c02 = a[..., 1, 0] * a[..., 2, 1] - a[..., 1, 1] * a[..., 2, 0]
    det = a[..., 0, 0] * c00 + a[..., 0, 1] * c01 + a[..., 0, 2] * c02
    inv_det = 1.0 / det
    r0 = torch.stack([c00, a[..., 0, 2] * a[..., 2, 1] - a[..., 0, 1] * a[..., 2, 2],
                      a[..., 0, 1] * a[..., 1, 2] - a[..., 0, 2] * a[..., 1, 1]], -1)
    r1 = torch.stack([c01, a[..., 0, 0] * a[..., 2, 2] - a[..., 0, 2] * a[..., 2, 0],
                      a[..., 0, 2] * a[..., 1, 0] - a[..., 0, 0] * a[..., 1, 2]], -1)
    r2 = torch.stack([c02, a[..., 0, 1] * a[..., 2, 0] - a[..., 0, 0] * a[..., 2, 1],
                      a[..., 0, 0] * a[..., 1, 1] - a[..., 0, 1] * a[..., 1, 0]], -1)
    return torch.stack([r0, r1, r2], -2) * inv_det[..., None, None]


def inv2(m: torch.Tensor) -> torch.Tensor:
    det = m[..., 0, 0] * m[..., 1, 1] - m[..., 0, 1] * m[..., 1, 0]
    r0 = torch.stack([m[..., 1, 1], -m[..., 0, 1]], -1)
    r1 = torch.stack([-m[..., 1, 0], m[..., 0, 0]], -1)
    return torch.stack([r0, r1], -2) / det[..., None, None]


def cols(c1, c2, c3) -> torch.Tensor:
    return torch.stack([c1, c2, c3], -1)


def qr_gram_schmidt(F: torch.Tensor):
    """QR_decomposition_3d, simulation.py:123-135 (eps 1e-9, R = Q.inverse() @ F)."""
    c1, c2, c3 = F[..., :, 0], F[..., :, 1], F[..., :, 2]
    q1 = c1 / tnorm(c1, 1e-9)[..., None]
    h2 = c2 - (c2 * q1).sum(-1, keepdim=True) * q1
    q2 = h2 / tnorm(h2, 1e-9)[..., None]
    h3 = c3 - (c3 * q1).sum(-1, keepdim=True) * q1 - (c3 * q2).sum(-1, keepdim=True) * q2
    q3 = h3 / tnorm(h3, 1e-9)[..., None]
    Q = cols(q1, q2, q3)
    R = inv3(Q) @ F
    return Q, R


def bspline(x: torch.Tensor, inv_dx: float):
    """Quadratic B-spline base/fx/weights, simulation.py:140-142 and :196-197.

    ``base`` is ``ti.cast(x*inv_dx - 0.5, i32)`` (truncation toward zero).
    Returns base [N,3] int64, fx [N,3], w [N,3(node),3(dim)], dw [N,3,3] (d w / d fx).
    """
    base = (x * inv_dx - 0.5).detach().trunc().to(torch.int64)
    fx = x * inv_dx - base.to(x.dtype)
    w = torch.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1) ** 2, 0.5 * (fx - 0.5) ** 2], 1)
    dw = torch.stack([fx - 1.5, 2 * (1.0 - fx), fx - 0.5], 1)
    return base, fx, w, dw


def stencil_weights(w: torch.Tensor) -> torch.Tensor:
    """[N,27] products w[i][0]*w[j][1]*w[k][2] in the reference's i,j,k loop order."""
    o = OFFSETS
    return w[:, o[:, 0], 0] * w[:, o[:, 1], 1] * w[:, o[:, 2], 2]


def stencil_keys(base: torch.Tensor, G: int) -> torch.Tensor:
    cell = base[:, None, :] + OFFSETS[None]
    return (cell[..., 0] * G + cell[..., 1]) * G + cell[..., 2]          # [N,27]


def binning_keys_fp32(x: np.ndarray, n_grid: int) -> np.ndarray:
    """Integer base cells trunc(f32(x) * inv_dx - 0.5) computed in FP32 (simulation.py:140)."""
    xf = np.asarray(x, dtype=np.float32)
    return np.trunc(xf * np.float32(n_grid) - np.float32(0.5)).astype(np.int32)


# ----------------------------------------------------------------------------
# 2x2 polar rotation with the reference's SVD backward
# ----------------------------------------------------------------------------

def _clamp(a: torch.Tensor) -> torch.Tensor:
    """simulation.py:44-50: keep the sign (a >= 0 counts as positive), |a| >= 1e-6."""
    return torch.where(a >= 0, torch.clamp(a, min=1e-6), torch.clamp(a, max=-1e-6))


def rotation_variant_svd(Ft: torch.Tensor):
    """SVD with det(U) = det(V) = +1 (the last singular value carries the sign)."""
    U, S, Vh = torch.linalg.svd(Ft)
    V = Vh.transpose(-1, -2).clone()
    U = U.clone()
    S = S.clone()
    fu = torch.linalg.det(U) < 0
    U[fu, :, 1] *= -1
    S[fu, 1] *= -1
    fv = torch.linalg.det(V) < 0
    V[fv, :, 1] *= -1
    S[fv, 1] *= -1
    return U, S, V


def reference_svd_backward(Ft: torch.Tensor, gR: torch.Tensor) -> torch.Tensor:
    """Literal backward_inplane_svd (simulation.py:71-88) for R = U V^T.

    gU = gR V and gV = gR^T U (the adjoint of ``inplane_U @ inplane_V.transpose()``
    at simulation.py:235); the sigma gradient is never written (zero).
    """
    U, S, V = rotation_variant_svd(Ft)
    gU = gR @ V
    gV = gR.transpose(-1, -2) @ U
    sig = torch.diag_embed(S)
    s = S ** 2
    Fm = torch.zeros_like(Ft)
    Fm[:, 0, 1] = 1.0 / _clamp(s[:, 1] - s[:, 0])
    Fm[:, 1, 0] = 1.0 / _clamp(s[:, 0] - s[:, 1])
    ut, vt = U.transpose(-1, -2), V.transpose(-1, -2)
    u_term = U @ ((Fm * (ut @ gU - gU.transpose(-1, -2) @ U)) @ sig) @ vt
    v_term = U @ (sig @ ((Fm * (vt @ gV - gV.transpose(-1, -2) @ V)) @ vt))
    return u_term + v_term


def polar_rotation_closed(Ft: torch.Tensor) -> torch.Tensor:
    x = Ft[..., 0, 0] + Ft[..., 1, 1]
    y = Ft[..., 1, 0] - Ft[..., 0, 1]
    r = torch.sqrt(x * x + y * y)
    c, s = x / r, y / r
    return torch.stack([torch.stack([c, -s], -1), torch.stack([s, c], -1)], -2)


class _PolarReference(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Ft):
        ctx.save_for_backward(Ft)
        return polar_rotation_closed(Ft)

    @staticmethod
    def backward(ctx, gR):
        (Ft,) = ctx.saved_tensors
        with torch.no_grad():
            g = reference_svd_backward(Ft.to(torch.float64), gR.to(torch.float64)).to(Ft.dtype)
        return g


def polar_rotation(Ft: torch.Tensor, svd_grad: str = "reference") -> torch.Tensor:
    """U V^T of ``ti.svd`` (simulation.py:39-42, used at :235)."""
    if svd_grad == "reference":
        return _PolarReference.apply(Ft)
    if svd_grad == "stable":
        return polar_rotation_closed(Ft)
    raise ValueError(svd_grad)


# ----------------------------------------------------------------------------
# state init (simulation.py:432-477)
# ----------------------------------------------------------------------------

def initialize_state(inst, dtype=torch.float64) -> Dict[str, torch.Tensor]:
    """initialize_ti_field, simulation.py:447-477 (plus the static Dinv)."""
    T = lambda a: torch.as_tensor(np.asarray(a), dtype=dtype)
    vx = T(inst.vertex_x)
    tri = torch.as_tensor(np.asarray(inst.triangles), dtype=torch.int64)
    x0, x1, x2 = vx[tri[:, 0]], vx[tri[:, 1]], vx[tri[:, 2]]
    d1 = x1 - x0
    d2 = x2 - x0
    n1 = d1 / tnorm(d1, 1e-6)[:, None]
    n2 = d2 / tnorm(d2, 1e-6)[:, None]
    c3 = torch.linalg.cross(n1, n2)
    nc3 = c3 / tnorm(c3, 1e-6)[:, None]
    initial_d = cols(d1, d2, nc3)
    Ne = tri.shape[0]
    z3 = lambda n: torch.zeros(n, 3, dtype=dtype)
    z33 = lambda n: torch.zeros(n, 3, 3, dtype=dtype)
    st = {
        "vx": vx, "vv": z3(vx.shape[0]), "vC": z33(vx.shape[0]),
        "ox": T(inst.grasper_x), "ov": z3(inst.n_grasper), "oC": z33(inst.n_grasper),
        "cx": T(inst.cube_x), "cv": z3(inst.n_cube), "cC": z33(inst.n_cube),
        "qx": (x0 + x1 + x2) / 3, "qv": z3(Ne), "qC": z33(Ne),
        "qF": torch.eye(3, dtype=dtype).expand(Ne, 3, 3).clone(), "qd": initial_d,
    }
    static = {"tri": tri, "Dinv": inv3(initial_d), "traj": T(inst.trajectory)}
    return st, static


def reset_quadrature_particles(st: Dict[str, torch.Tensor], static):
    """reset_quadrature_particles, simulation_without_grad.py:472-486: element
    centroids, d and Dinv re-derived from the current vertex positions (the
    same normalised-normal construction as initialize_ti_field); F, v and C are
    kept.  Returns (new state, new static)."""
    vx, tri = st["vx"], static["tri"]
    x0, x1, x2 = vx[tri[:, 0]], vx[tri[:, 1]], vx[tri[:, 2]]
    d1, d2 = x1 - x0, x2 - x0
    n1 = d1 / tnorm(d1, 1e-6)[:, None]
    n2 = d2 / tnorm(d2, 1e-6)[:, None]
    c3 = torch.linalg.cross(n1, n2)
    d = cols(d1, d2, c3 / tnorm(c3, 1e-6)[:, None])
    new = dict(st)
    new["qx"] = (x0 + x1 + x2) / 3
    new["qd"] = d
    new_static = dict(static)
    new_static["Dinv"] = inv3(d)
    return new, new_static


def make_params(inst, dtype=torch.float64, requires_grad=False) -> Dict[str, torch.Tensor]:
    p = {k: torch.tensor(float(getattr(inst, a)), dtype=dtype, requires_grad=requires_grad)
         for k, a in (("E", "E"), ("nu", "nu"), ("ks", "shearing_stiffness"),
                      ("kc", "contact_stiffness"), ("damping", "damping"),
                      ("friction", "friction"))}
    if inst.element_moduli is not None:
        p["elem"] = torch.tensor(np.asarray(inst.element_moduli), dtype=dtype,
                                 requires_grad=requires_grad)
    return p


# ----------------------------------------------------------------------------
# element stress (simulation.py:214-248)
# ----------------------------------------------------------------------------

def element_stress(F: torch.Tensor, params: Dict[str, torch.Tensor], svd_grad: str):
    """quadrature_P for every element.  Returns (P, Q, R, Ft)."""
    Q, R = qr_gram_schmidt(F)                           # :214 (and :27)
    Ft = R[:, :2, :2]                                   # :28
    if "elem" in params:
        E = params["elem"][:, 0]
        ks = params["elem"][:, 1]
        kc = params["elem"][:, 2]
    else:
        E, ks, kc = params["E"], params["ks"], params["kc"]
    nu = params["nu"]
    R02, R12, R22 = R[:, 0, 2], R[:, 1, 2], R[:, 2, 2]
    a00 = ks * R02 * R02                                # :216-217 (the E term is overwritten)
    a01 = ks * R02 * R12                                # :218
    a02 = ks * R02 * R22                                # :219
    a11 = ks * R12 * R12                                # :220
    a12 = ks * R12 * R22                                # :221
    a22 = ((R22 <= 1).to(F.dtype) * (-kc) * ((1 - R22) ** 2)) * R22   # :222
    A = torch.stack([torch.stack([a00, a01, a02], -1),
                     torch.stack([a01, a11, a12], -1),
                     torch.stack([a02, a12, a22 * torch.ones_like(a00)], -1)], -2)
    mu = E / (2 * (1 + nu))                             # :230
    la = E * nu / ((1 + nu) * (1 - 2 * nu))             # :230-232
    J = Ft[:, 0, 0] * Ft[:, 1, 1] - Ft[:, 0, 1] * Ft[:, 1, 0]     # :234
    Rr = polar_rotation(Ft, svd_grad)                   # :235
    mu_ = mu[:, None, None] if torch.is_tensor(mu) and mu.dim() == 1 else mu
    la_ = la[:, None, None] if torch.is_tensor(la) and la.dim() == 1 else la
    P_fc = 2 * mu_ * (Ft - Rr) + la_ * ((J - 1) * J)[:, None, None] * inv2(Ft.transpose(-1, -2))  # :236-237
    Pfc3 = torch.cat([torch.cat([P_fc, torch.zeros_like(P_fc[:, :, :1])], 2),
                      torch.zeros_like(F[:, :1, :])], 1)
    P = Q @ Pfc3 + Q @ A @ inv3(R).transpose(-1, -2)   # :238-240
    return P, Q, R, Ft



# ----------------------------------------------------------------------------
# grid update, G2P gather and element update (simulation.py:274-430)
# ----------------------------------------------------------------------------

def grid_op(ijk: torch.Tensor, grid_m: torch.Tensor, grid_v: torch.Tensor, G: int, bound: int,
            damping, dt: float) -> torch.Tensor:
    """grid_op, simulation.py:274-302, on the cells ``ijk`` [n,3] (int): v = mv/(m+1e-9),
    the damping extension v *= (1 - damping dt) (identity at 0), then the wall
    conditions in the reference's order: x low, x high, z low, z high (all
    components zeroed when moving outward), the floor j < bound (only v_y
    zeroed when v_y < 0), y high (all components).  Low walls are i < bound,
    high walls i > G - bound."""
    gi, gj, gk = ijk[:, 0], ijk[:, 1], ijk[:, 2]
    inv_m = 1 / (grid_m + 1e-9)
    v_out = inv_m[:, None] * grid_v
    v_out = v_out * (1 - damping * dt)
    b = bound
    zero3 = torch.zeros_like(v_out)

    def zero_all(cond):
        return torch.where(cond[:, None], zero3, v_out)
    v_out = zero_all((gi < b) & (v_out[:, 0] < 0))                                # :279-282
    v_out = zero_all((gi > G - b) & (v_out[:, 0] > 0))                            # :283-286
    v_out = zero_all((gk < b) & (v_out[:, 2] < 0))                                # :287-290
    v_out = zero_all((gk > G - b) & (v_out[:, 2] > 0))                            # :291-294
    c5 = (gj < b) & (v_out[:, 1] < 0)                                             # :295-296
    v_out = torch.stack([v_out[:, 0], torch.where(c5, torch.zeros_like(v_out[:, 1]), v_out[:, 1]),
                         v_out[:, 2]], -1)
    v_out = zero_all((gj > G - b) & (v_out[:, 1] > 0))                            # :297-300
    return v_out


def g2p_gather(x: torch.Tensor, gv: torch.Tensor, inv_dx: float):
    """APIC gather of g2p, simulation.py:304-380: new_v = sum w g and
    new_C = 4 inv_dx sum w g (i - fx)^T over the 27 stencil nodes, whose grid
    velocities ``gv`` [N,27,3] are in the reference's i,j,k loop order."""
    base, fx, w, _ = bspline(x, inv_dx)
    W = stencil_weights(w)
    dp = OFFSETS[None].to(x.dtype) - fx[:, None, :]
    new_v = (W[..., None] * gv).sum(1)
    new_C = ((4 * W[..., None, None] * (gv[..., :, None] * dp[..., None, :])) * inv_dx).sum(1)
    return new_v, new_C


def damp_element_C(C: torch.Tensor) -> torch.Tensor:
    """Element affine matrix kept as skew(C) + (1 - 0.95) sym(C), simulation.py:381-383."""
    C_skew = 0.5 * (C - C.transpose(-1, -2))
    C_sym = 0.5 * (C + C.transpose(-1, -2))
    return C_skew + (1 - 0.95) * C_sym


def return_map(R: torch.Tensor, ks, kc, cf) -> torch.Tensor:
    """The return mapping of update_particle_state, simulation.py:408-428, on
    the R of the QR of the new F (only the third column changes):
      ks == 0:          R02 = R12 = 1e-7, R22 = min(R22, 1)
      else R22 > 1:     R02 = R12 = 1e-7, R22 = 1
      else R22 < 0:     R02 = R12 = 1e-7, R22 = max(R22, -1)
      else R22 <= 1:    with shear = R02^2 + R12^2 and k = ks / kc, if
                        k^2 shear - (cf (R22-1)^2)^2 > 0 the pair (R02, R12)
                        is scaled by cf (R22-1)^2 / (k sqrt(shear)).
    The reference computes it with cf = 0.1 and never writes it back (so it
    changes nothing there); ``update_elements`` applies it only when asked,
    with cf = the friction parameter."""
    R02, R12, R22 = R[:, 0, 2], R[:, 1, 2], R[:, 2, 2]
    one = torch.ones_like(R22)
    ks, kc, cf = ks * one, kc * one, cf * one
    b1 = ks == 0
    b2 = ~b1 & (R22 > 1)
    b3 = ~b1 & ~b2 & (R22 < 0)
    b4 = ~b1 & ~b2 & ~b3 & (R22 <= 1)
    shear = R02 ** 2 + R12 ** 2
    k = ks / kc
    act = b4 & ((k ** 2) * shear - ((cf * (R22 - 1) ** 2) ** 2) > 0)
    safe = torch.where(act, shear, one)
    scale = cf * ((R22 - 1) ** 2) / (k * torch.sqrt(safe))
    const = b1 | b2 | b3
    tiny = torch.full_like(R22, 1e-7)
    n02 = torch.where(const, tiny, torch.where(act, R02 * scale, R02))
    n12 = torch.where(const, tiny, torch.where(act, R12 * scale, R12))
    n22 = torch.where(b1, torch.minimum(R22, one),
                      torch.where(b2, one, torch.where(b3, torch.maximum(R22, -one), R22)))
    col3 = torch.stack([n02, n12, n22], -1)
    return torch.cat([R[:, :, :2], col3[:, :, None]], 2)


def update_elements(vx: torch.Tensor, vv: torch.Tensor, C: torch.Tensor, d: torch.Tensor,
                    Dinv: torch.Tensor, tri: torch.Tensor, dt: float, rm=None) -> Dict[str, torch.Tensor]:
    """update_particle_state, simulation.py:386-430, from the new vertex state
    (vx, vv), the new element affine matrix C and the old d: element v and x
    are the vertex means; d = [x2-x1 | x3-x1 | col3 + (dt C + dt^2/2 C C) col3]
    with col3 the old d's normal column; F = d Dinv; the stored d is the
    round trip F inverse(Dinv).  The return mapping of :408-428 changes only a
    local R that is never written back, so by default it is left out; with
    rm = (ks, kc, cf) it is applied (F = Q return_map(R)) before F and d are
    stored."""
    x1, x2, x3 = vx[tri[:, 0]], vx[tri[:, 1]], vx[tri[:, 2]]
    v1, v2, v3 = vv[tri[:, 0]], vv[tri[:, 1]], vv[tri[:, 2]]
    out = {"qv": (1 / 3) * (v1 + v2 + v3),                                        # :392
           "qx": (1 / 3) * (x1 + x2 + x3)}                                        # :393
    col3 = d[:, :, 2]
    M = dt * C + 0.5 * dt * dt * C @ C
    col3n = col3 + (M @ col3[..., None])[..., 0]                                  # :397-400
    new_d = cols(x2 - x1, x3 - x1, col3n)
    new_F = new_d @ Dinv                                                          # :407
    if rm is not None:
        Q, R = qr_gram_schmidt(new_F)                                             # :408
        new_F = Q @ return_map(R, *rm)
    out["qd"] = new_F @ inv3(Dinv)                                                # :429
    out["qF"] = new_F                                                             # :430
    return out

# ----------------------------------------------------------------------------
# one substep
# ----------------------------------------------------------------------------

def grasper_velocity(static, K: Consts, f: int, g: int, n_op: int, dtype):
    """simulation.py:160-165 with the frame index taken as an integer.

    frame = f // steps_per_frame + frames_per_window * g (the reference computes
    int(f*dt/interval) in FP32; with interval = k*dt the integer form is exact).
    """
    frame = f // K.steps_per_frame + K.frames_per_window * g
    interval = K.steps_per_frame * K.dt
    gid = torch.arange(n_op) // K.group_size
    traj = static["traj"]
    if frame + 1 < K.n_actions:
        v = (traj[frame + 1, gid] - traj[frame, gid]) / interval
    else:
        v = torch.zeros(n_op, 3, dtype=dtype)
    return v


def substep(st: Dict[str, torch.Tensor], static, params, K: Consts, f: int, g: int = 0,
            svd_grad: str = "reference", return_grid: bool = False, return_mapping: bool = False):
    """forward(s, grad_update), simulation.py:480-488 -> new state dict.
    ``return_mapping``: apply the element return mapping with cf = friction
    (off in the reference, where it is computed and discarded)."""
    dtype = st["vx"].dtype
    G, dt, dx, inv_dx = K.G, K.dt, K.dx, K.inv_dx
    tri = static["tri"]
    Dinv = static["Dinv"]

    # compute_inplane_F_tmp + inplane_svd + element stress (:24-28, :39-42, :214-240)
    P, Q, R, Ft = element_stress(st["qF"], params, svd_grad)

    # ---- p2g (:137-271): collect every scatter as (keys, momentum, mass) ----
    keys, mom, mass = [], [], []

    def point_class(x, C, pmass, vel):
        base, fx, w, _ = bspline(x, inv_dx)
        W = stencil_weights(w)                                   # [N,27]
        dpos = (OFFSETS[None].to(dtype) - fx[:, None, :]) * dx  # [N,27,3]
        affine = pmass * C
        adp = torch.einsum("nij,nkj->nki", affine, dpos)         # affine @ dpos
        if vel is None:
            m = W[..., None] * adp
        else:
            m = W[..., None] * (pmass * vel[:, None, :] + adp)
        keys.append(stencil_keys(base, G))
        mom.append(m)
        mass.append(W * pmass)
        return base, fx, w

    point_class(st["cx"], st["cC"], K.cube_p_mass, None)                           # :139-152
    vop = grasper_velocity(static, K, f, g, st["ox"].shape[0], dtype)
    point_class(st["ox"], st["oC"], K.operator_p_mass, vop)                        # :154-174
    grav = torch.tensor([0.0, dt * K.gravity, 0.0], dtype=dtype)
    point_class(st["vx"], st["vC"], K.vertex_p_mass, st["vv"] + grav)              # :176-191

    # element particles (:193-271)
    base, fx, w, dwr = bspline(st["qx"], inv_dx)
    xjw = dwr * inv_dx                                                             # :197
    W = stencil_weights(w)
    o = OFFSETS
    gradw = torch.stack([xjw[:, o[:, 0], 0] * w[:, o[:, 1], 1] * w[:, o[:, 2], 2],
                         w[:, o[:, 0], 0] * xjw[:, o[:, 1], 1] * w[:, o[:, 2], 2],
                         w[:, o[:, 0], 0] * w[:, o[:, 1], 1] * xjw[:, o[:, 2], 2]], -1)  # [N,27,3]
    dpos = (o[None].to(dtype) - fx[:, None, :]) * dx
    qm = K.quadrature_p_mass
    affine = st["qC"] * qm
    adp = torch.einsum("nij,nkj->nki", affine, dpos)
    vol = K.quadrature_p_vol
    d = st["qd"]
    dcol3 = d[:, :, 2]
    Pcol3 = P[:, :, 2]
    own = W[..., None] * (qm * st["qv"][:, None, :] + adp)                          # :262
    normal = ((-1) * vol * Pcol3[:, None, :] * (gradw * dcol3[:, None, :]).sum(-1, keepdim=True)) * dt  # :264-265
    keys.append(stencil_keys(base, G))
    mom.append(own)
    mass.append(W * qm)
    keys.append(stencil_keys(base, G))
    mom.append(normal)
    mass.append(torch.zeros_like(W))
    raw1 = torch.stack([Dinv[:, 0, 0], d[:, 0, 1], d[:, 0, 2]], -1)                # :247
    raw2 = torch.stack([Dinv[:, 1, 0], d[:, 1, 1], d[:, 1, 2]], -1)                # :248
    f1 = (P @ (raw1 + raw2)[..., None])[..., 0]
    f2 = (P @ raw1[..., None])[..., 0]
    f3 = (P @ raw2[..., None])[..., 0]
    for vi, fv, sgn in ((0, f1, 1.0), (1, f2, -1.0), (2, f3, -1.0)):              # :266-271
        vb, _, vw, _ = bspline(st["vx"][tri[:, vi]], inv_dx)
        vW = stencil_weights(vw)
        keys.append(stencil_keys(vb, G))
        if sgn > 0:
            mom.append((vol * vW[..., None] * fv[:, None, :]) * dt)
        else:
            mom.append(((-1) * vol * vW[..., None] * fv[:, None, :]) * dt)
        mass.append(torch.zeros_like(vW))

    allk = torch.cat([k.reshape(-1) for k in keys])
    uniq, inv = torch.unique(allk, return_inverse=True)
    nC = uniq.shape[0]
    grid_v = torch.zeros(nC, 3, dtype=dtype).index_add(0, inv, torch.cat([m.reshape(-1, 3) for m in mom]))
    grid_m = torch.zeros(nC, dtype=dtype).index_add(0, inv, torch.cat([m.reshape(-1) for m in mass]))

    # ---- grid_op (:274-302) ----
    gi = uniq // (G * G)
    gj = (uniq // G) % G
    gk = uniq % G
    v_out = grid_op(torch.stack([gi, gj, gk], -1), grid_m, grid_v, G, K.bound, params["damping"], dt)

    # ---- g2p (:304-383) ----
    sizes = [k.shape[0] for k in keys]
    inv_split = torch.split(inv, [s * 27 for s in sizes])

    def g2p_point(x, which):
        return g2p_gather(x, v_out[inv_split[which]].reshape(x.shape[0], 27, 3), inv_dx)

    new = {}
    cv, cC = g2p_point(st["cx"], 0)
    new["cv"], new["cx"], new["cC"] = cv, st["cx"] + dt * cv, cC
    ov, oC = g2p_point(st["ox"], 1)
    new["ov"], new["ox"], new["oC"] = ov, st["ox"] + dt * ov, oC
    vv, vC = g2p_point(st["vx"], 2)
    new["vv"], new["vx"], new["vC"] = vv, st["vx"] + dt * vv, vC
    _, qCn = g2p_point(st["qx"], 3)
    qC1 = damp_element_C(qCn)                                                     # :381-383

    # ---- update_particle_state (:386-430) ----
    rm = None
    if return_mapping:
        if "elem" in params:
            rm = (params["elem"][:, 1], params["elem"][:, 2], params["friction"])
        else:
            rm = (params["ks"], params["kc"], params["friction"])
    new.update(update_elements(new["vx"], new["vv"], qC1, d, Dinv, tri, dt, rm))
    new["qC"] = qC1
    if return_grid:
        return new, {"keys": uniq, "m": grid_m, "mv": grid_v, "v_out": v_out, "P": P}
    return new


# ----------------------------------------------------------------------------
# loss (parameter_fitting_functions.py:3-15) and the tape driver (:17-58)
# ----------------------------------------------------------------------------

def compute_loss(vertex_x: torch.Tensor, targets: torch.Tensor):
    """Mean over targets of min over vertices of |x - t|^2 / 3 (first argmin)."""
    dist = (vertex_x[:, None, :] - targets[None, :, :]) ** 2
    dn = (dist[..., 0] + dist[..., 1] + dist[..., 2]) / 3
    idx = torch.argmin(dn, dim=0)                     # first minimum, strict <
    mins = dn.gather(0, idx[None])[0]
    nt = targets.shape[0]
    loss = (mins / nt).sum()
    return loss, idx


def rollout(inst, params=None, n_substeps: Optional[int] = None, g: int = 0,
            dtype=torch.float64, svd_grad: str = "reference", keep_states: bool = False,
            state=None, static=None, return_mapping: bool = False):
    """Run ``n_substeps`` forward substeps from the initial state.

    Returns (states_list_or_final, static).  The loss of the reference tape is on
    state ``n_substeps - 2`` (parameter_fitting_functions.py:27-29 reads
    vertex_x[f-1] with f = max_steps-2).
    """
    K = Consts.from_instance(inst)
    if params is None:
        params = make_params(inst, dtype)
    if state is None:
        state, static = initialize_state(inst, dtype)
    S = n_substeps
    states = [state] if keep_states else None
    st = state
    for f in range(S):
        st = substep(st, static, params, K, f, g, svd_grad, return_mapping=return_mapping)
        if keep_states:
            states.append(st)
    return (states if keep_states else st), static


def tape_loss(inst, params, n_substeps: int, targets, g: int = 0, dtype=torch.float64,
              svd_grad: str = "reference", return_mapping: bool = False):
    """The loss the reference's Tape records: compute_loss on state S-2."""
    states, _ = rollout(inst, params, n_substeps, g, dtype, svd_grad, keep_states=True,
                        return_mapping=return_mapping)
    tgt = torch.as_tensor(np.asarray(targets), dtype=dtype)
    loss, idx = compute_loss(states[n_substeps - 2]["vx"], tgt)
    return loss, states
