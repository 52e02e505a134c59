"""Known-answer pins of the oracle's element stress, element force vectors,
grid update, G2P gather, element update and grasper velocity, and a mutation
check that each pin fails on a plausible mistake planted in the oracle.

Every expected value below comes from the mathematics (hand-derived closed
forms, numpy's SVD and inverse, exact reproduction properties of the
quadratic B-spline, convergence orders) or from literal numbers, never from
the oracle's own formulas.  The reference tests none of these functions and
ships no vectors for them.

Reference lines pinned: simulation.py:160-165 (grasper velocity), :214-248
(stress, force vectors), :262-271 (element scatter), :274-302 (grid update),
:304-383 (G2P, element C damping), :386-430 (element update).

The mutation check (``test_pin_catches_mutation``) rewrites one expression of
oracle/sim_torch.py in memory (a dropped term, a sign, an index, a
transpose), loads the result as a separate module and asserts that the
named pin rejects it.  Nothing is written to disk.
"""
import math
import sys
import types

import numpy as np
import pytest
import torch

from oracle import sim_torch as O_REAL

torch.set_default_dtype(torch.float64)

E_, NU_, KS_, KC_ = 5e3, 0.3, 100.0, 80.0


def _t(v):
    return torch.tensor(float(v), dtype=torch.float64)


def _params(E=E_, nu=NU_, ks=KS_, kc=KC_, damping=0.0):
    return {"E": _t(E), "nu": _t(nu), "ks": _t(ks), "kc": _t(kc), "damping": _t(damping), "friction": _t(0.0)}


def _lame(E, nu):
    # textbook conversion of Young's modulus and Poisson ratio
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def _close(a, b, rtol, what=""):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.abs(b).max()), 1e-300)
    err = float(np.abs(a - b).max()) / scale
    assert err <= rtol, f"{what}: relative error {err:.3e} > {rtol:.1e}"


# ----------------------------------------------------------------------------
# element stress (simulation.py:214-240)
# ----------------------------------------------------------------------------
# F = [[s, 0, g], [0, t, b], [0, 0, r]] is upper triangular with a positive
# diagonal, so its QR has Q = I and R = F (up to the 1e-9 in the column norms).
# With Ft = diag(s, t) the polar rotation is I, J = s t, inverse(Ft^T) =
# diag(1/s, 1/t), and by hand
#   P = [[2mu(s-1) + la(st-1)t, 0,                    ks g],
#        [0,                    2mu(t-1) + la(st-1)s, ks b],
#        [g (ks r + c) / s,     b (ks r + c) / t,     -c  ]],   c = kc (1-r)^2 if r <= 1 else 0.
UPPER_CASES = [(1.3, 0.8, 0.2, -0.15, 0.7), (0.9, 1.2, -0.3, 0.25, 1.2), (1.1, 0.75, 0.0, 0.4, 0.5)]


def _upper_expected(s, t, g, b, r):
    mu, la = _lame(E_, NU_)
    c = KC_ * (1 - r) ** 2 if r <= 1 else 0.0
    return np.array([[2 * mu * (s - 1) + la * (s * t - 1) * t, 0.0, KS_ * g],
                     [0.0, 2 * mu * (t - 1) + la * (s * t - 1) * s, KS_ * b],
                     [g * (KS_ * r + c) / s, b * (KS_ * r + c) / t, -c]])


def pin_stress_upper(O):
    F = torch.tensor([[[s, 0, g], [0, t, b], [0, 0, r]] for s, t, g, b, r in UPPER_CASES])
    for mode in ("stable", "reference"):
        P = O.element_stress(F, _params(), mode)[0].detach().numpy()
        for n, case in enumerate(UPPER_CASES):
            _close(P[n], _upper_expected(*case), 1e-7, f"upper-triangular F {case} ({mode})")


def pin_stress_inplane(O):
    """Non-symmetric in-plane F = [[a, b, 0], [0, d, 0], [0, 0, 1]]: P = [P_fc 0; 0 0]
    with P_fc = 2mu(Ft - U V^T) + la (J-1) J inverse(Ft)^T from numpy's SVD and
    inverse (R22 = sqrt(1 + 1e-9) > 1 switches the contact term off, R02 = R12 = 0)."""
    mu, la = _lame(E_, NU_)
    cases = [(1.2, 0.35, 0.85), (0.7, -0.5, 1.1), (0.95, 0.6, 0.9)]
    F = torch.tensor([[[a, b, 0], [0, d, 0], [0, 0, 1.0]] for a, b, d in cases])
    P = O.element_stress(F, _params(), "stable")[0].detach().numpy()
    for n, (a, b, d) in enumerate(cases):
        Ft = np.array([[a, b], [0.0, d]])
        U, _, Vh = np.linalg.svd(Ft)
        J = np.linalg.det(Ft)
        Pfc = 2 * mu * (Ft - U @ Vh) + la * (J - 1) * J * np.linalg.inv(Ft).T
        exp = np.zeros((3, 3))
        exp[:2, :2] = Pfc
        _close(P[n], exp, 1e-7, f"in-plane F {(a, b, d)}")


def pin_stress_rotation(O):
    """Frame indifference: P(Rot F) = Rot P(F) for a rotation Rot (QR gives Q = Rot)."""
    rng = np.random.default_rng(5)
    F1 = torch.tensor([[[s, 0.1, g], [0, t, b], [0, 0, r]] for s, t, g, b, r in UPPER_CASES])
    rots = []
    for _ in UPPER_CASES:
        q, rr = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(rr))[None, :]
        if np.linalg.det(q) < 0:
            q[:, 0] *= -1
        rots.append(q)
    Rot = torch.tensor(np.stack(rots))
    P1 = O.element_stress(F1, _params(), "stable")[0]
    P2 = O.element_stress(Rot @ F1, _params(), "stable")[0]
    _close((P2).detach().numpy(), (Rot @ P1).detach().numpy(), 1e-9, "rotated F")


# ----------------------------------------------------------------------------
# element force vectors and normal impulse (simulation.py:243-248, 262-271)
# ----------------------------------------------------------------------------
# One triangle on a 32^3 grid with every stencil apart from the others.  The
# rest shape and current d are literal matrices of primes, so that raw1 =
# (Dinv00, d01, d02) = (2, 31, 37) and raw2 = (Dinv10, d11, d12) = (7, 43, 47).
# With v = C = 0 the 27 cells of vertex k's stencil hold (sum of weights = 1)
# m_v (0, g dt, 0) plus vol dt (P(raw1 + raw2), -P raw1, -P raw2)[k]; the
# element's own stencil holds zero total momentum (the weight gradients sum to
# zero) whose first moment about the element, sum_i mv_i (x_i - x_e)^T, is
# -vol dt P[:, 2] d[:, 2]^T (sum_i grad w_i (x_i - x_e)^T = I exactly).
FORCE_DINV = [[2.0, 3, 5], [7, 11, 13], [17, 19, 23]]
FORCE_D = [[29.0, 31, 37], [41, 43, 47], [53, 59, 61]]
RAW1, RAW2, DCOL3 = np.array([2.0, 31, 37]), np.array([7.0, 43, 47]), np.array([37.0, 47, 61])


def _force_scene(O):
    G = 32
    dx, dt = 1.0 / G, 1e-4
    K = O.Consts(G=G, dt=dt, dx=dx, inv_dx=float(G), bound=3, gravity=-9.81, vertex_p_mass=0.01,
                 quadrature_p_mass=0.02, quadrature_p_vol=3e-4, cube_p_mass=2.0, operator_p_mass=1.5,
                 steps_per_frame=10, frames_per_window=1, group_size=1, n_actions=4)
    z3 = lambda n: torch.zeros(n, 3)
    z33 = lambda n: torch.zeros(n, 3, 3)
    st = {"vx": torch.tensor([[8.7, 16.2, 8.3], [16.4, 16.6, 8.9], [8.2, 16.1, 16.8]]) * dx,
          "vv": z3(3), "vC": z33(3),
          "ox": torch.tensor([[24.5, 24.5, 24.5], [25.0, 24.6, 24.4]]) * dx, "ov": z3(2), "oC": z33(2),
          "cx": torch.tensor([[24.5, 6.5, 24.5]]) * dx, "cv": z3(1), "cC": z33(1),
          "qx": torch.tensor([[11.6, 16.4, 11.7]]) * dx, "qv": z3(1), "qC": z33(1),
          "qF": torch.tensor([[[1.3, 0, 0.2], [0, 0.8, -0.15], [0, 0, 0.7]]]),
          "qd": torch.tensor([FORCE_D])}
    static = {"tri": torch.tensor([[0, 1, 2]]), "Dinv": torch.tensor([FORCE_DINV]), "traj": torch.zeros(4, 2, 3)}
    return K, st, static


def _stencil_cells(x, G):
    base = np.trunc(np.asarray(x) * G - 0.5).astype(np.int64)
    off = np.array([(i, j, k) for i in range(3) for j in range(3) for k in range(3)])
    return base[None, :] + off


def pin_forces(O):
    K, st, static = _force_scene(O)
    _, grid = O.substep(st, static, _params(), K, 0, 0, "stable", return_grid=True)
    keys = grid["keys"].numpy()
    mv = grid["mv"].detach().numpy()
    m = grid["m"].detach().numpy()
    P = grid["P"][0].detach().numpy()
    G, vol, dt = K.G, K.quadrature_p_vol, K.dt

    def cell_rows(cells):
        k = (cells[:, 0] * G + cells[:, 1]) * G + cells[:, 2]
        idx = np.searchsorted(keys, k)
        assert (keys[idx] == k).all()
        return idx

    fk = [P @ (RAW1 + RAW2), -(P @ RAW1), -(P @ RAW2)]
    for v in range(3):
        rows = cell_rows(_stencil_cells(st["vx"][v].numpy(), G))
        expect = np.array([0.0, K.vertex_p_mass * dt * K.gravity, 0.0]) + vol * dt * fk[v]
        _close(mv[rows].sum(0), expect, 1e-9, f"vertex {v} stencil momentum")
    cells = _stencil_cells(st["qx"][0].numpy(), G)
    rows = cell_rows(cells)
    assert abs(m[rows].sum() - K.quadrature_p_mass) < 1e-12
    assert np.abs(mv[rows].sum(0)).max() < 1e-9 * np.abs(vol * dt * P[:, 2]).max() * 100
    rel = cells * K.dx - st["qx"][0].numpy()[None, :]
    moment = mv[rows].T @ rel
    _close(moment, -vol * dt * np.outer(P[:, 2], DCOL3), 1e-9, "element normal-impulse moment")
    # second moments vanish: sum_i w_i (x_i - x_e)_a (x_i - x_e)_b = dx^2/4 delta_ab
    # does not depend on x_e, so its gradient sum_i grad w_i (x_i - x_e)_a (x_i - x_e)_b = 0
    second = np.einsum("nc,na,nb->cab", mv[rows], rel, rel)
    scale = vol * dt * np.abs(P[:, 2]).max() * np.abs(DCOL3).max() * K.dx
    assert np.abs(second).max() < 1e-9 * scale, np.abs(second).max() / scale


# ----------------------------------------------------------------------------
# grid update (simulation.py:274-302)
# ----------------------------------------------------------------------------
GRID_G = 32
# (cell, velocity, expected velocity after the wall conditions); low walls
# i < 3, high walls i > G - 3 = 29, the floor (j < 3) zeroes only v_y
GRID_CASES = [
    ((2, 10, 10), (-1.0, 2.0, 3.0), (0.0, 0.0, 0.0)),
    ((3, 10, 10), (-1.0, 2.0, 3.0), (-1.0, 2.0, 3.0)),
    ((2, 10, 10), (1.0, 2.0, 3.0), (1.0, 2.0, 3.0)),
    ((29, 10, 10), (1.0, 2.0, 3.0), (1.0, 2.0, 3.0)),
    ((30, 10, 10), (1.0, 2.0, 3.0), (0.0, 0.0, 0.0)),
    ((30, 10, 10), (-1.0, 2.0, 3.0), (-1.0, 2.0, 3.0)),
    ((10, 10, 2), (1.0, 2.0, -3.0), (0.0, 0.0, 0.0)),
    ((10, 10, 3), (1.0, 2.0, -3.0), (1.0, 2.0, -3.0)),
    ((10, 10, 30), (1.0, 2.0, 3.0), (0.0, 0.0, 0.0)),
    ((10, 10, 29), (1.0, 2.0, 3.0), (1.0, 2.0, 3.0)),
    ((10, 2, 10), (1.0, -2.0, 3.0), (1.0, 0.0, 3.0)),
    ((10, 2, 10), (1.0, 2.0, 3.0), (1.0, 2.0, 3.0)),
    ((10, 3, 10), (1.0, -2.0, 3.0), (1.0, -2.0, 3.0)),
    ((10, 30, 10), (1.0, 2.0, 3.0), (0.0, 0.0, 0.0)),
    ((10, 30, 10), (1.0, -2.0, 3.0), (1.0, -2.0, 3.0)),
    ((2, 2, 10), (-1.0, -2.0, 3.0), (0.0, 0.0, 0.0)),
    ((2, 2, 10), (1.0, -2.0, 3.0), (1.0, 0.0, 3.0)),
    ((10, 2, 30), (1.0, -2.0, 3.0), (0.0, 0.0, 0.0)),
    ((15, 15, 15), (0.3, -0.7, 0.2), (0.3, -0.7, 0.2)),
]


def pin_grid_op(O):
    ijk = torch.tensor([c for c, _, _ in GRID_CASES])
    v = torch.tensor([u for _, u, _ in GRID_CASES])
    m = torch.full((len(GRID_CASES),), 2.0)
    out = O.grid_op(ijk, m, m[:, None] * v, GRID_G, 3, _t(0.0), 1e-4).numpy()
    exp = np.array([e for _, _, e in GRID_CASES]) * (2.0 / (2.0 + 1e-9))
    _close(out, exp, 1e-14, "wall conditions")
    # damping extension v (1 - damping dt) and the 1/(m + 1e-9) normalisation
    out = O.grid_op(torch.tensor([[15, 15, 15]]), torch.tensor([0.5]), torch.tensor([[0.5, -1.0, 1.5]]),
                    GRID_G, 3, _t(3.0), 1e-4).numpy()
    _close(out[0], np.array([1.0, -2.0, 3.0]) * (0.5 / (0.5 + 1e-9)) * (1 - 3.0 * 1e-4), 1e-14, "damping")


# ----------------------------------------------------------------------------
# G2P gather (simulation.py:304-380) and element C damping (:381-383)
# ----------------------------------------------------------------------------
def pin_g2p(O):
    """APIC with quadratic B-splines reproduces an affine grid velocity field
    v(x) = v0 + M x exactly: the particle gets v0 + M x_p and C = M."""
    G = 32
    rng = np.random.default_rng(7)
    x = torch.tensor(rng.random((30, 3)) * 0.6 + 0.2)
    M = torch.tensor(rng.standard_normal((3, 3)))
    v0 = torch.tensor([0.1, -0.2, 0.3])
    cells = np.stack([_stencil_cells(xp, G) for xp in x.numpy()])     # [N,27,3], i,j,k order
    gv = v0 + torch.tensor(cells / G) @ M.T
    v, C = O.g2p_gather(x, gv, float(G))
    _close(v.numpy(), (v0 + x @ M.T).numpy(), 1e-12, "affine velocity")
    _close(C.numpy(), M.expand(30, 3, 3).numpy(), 1e-11, "affine C")


def pin_damp(O):
    rng = np.random.default_rng(8)
    C = torch.tensor(rng.standard_normal((10, 3, 3)))
    skew = 0.5 * (C - C.transpose(-1, -2))
    sym = 0.5 * (C + C.transpose(-1, -2))
    out = O.damp_element_C(C)
    _close(out.numpy(), (skew + 0.05 * sym).numpy(), 1e-14, "element C damping")
    _close(O.damp_element_C(skew).numpy(), skew.numpy(), 1e-15, "skew C kept")


# ----------------------------------------------------------------------------
# element update (simulation.py:386-430)
# ----------------------------------------------------------------------------
def _rodrigues(axis, theta):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    Kx = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(theta) * Kx + (1 - math.cos(theta)) * Kx @ Kx, Kx


def pin_update(O):
    # literal centroid and mean velocity
    vx = torch.tensor([[0.0, 0, 0], [3, 0, 0], [0, 6, 0]])
    vv = torch.tensor([[1.0, 2, 3], [4, 5, 6], [7, 8, 9]])
    tri = torch.tensor([[0, 1, 2]])
    d0 = torch.tensor([[[3.0, 0, 0], [0, 6, 0], [0, 0, 1]]])
    d0[0, :, 2] = torch.tensor([0.0, 0.0, 1.0])
    Dinv = torch.linalg.inv(d0)
    out = O.update_elements(vx, vv, torch.zeros(1, 3, 3), d0, Dinv, tri, 1e-4)
    _close(out["qx"][0].numpy(), [1.0, 2.0, 0.0], 1e-15, "centroid")
    _close(out["qv"][0].numpy(), [4.0, 5.0, 6.0], 1e-15, "mean velocity")
    _close(out["qF"][0].numpy(), np.eye(3), 1e-14, "F at rest")
    _close(out["qd"][0].numpy(), d0[0].numpy(), 1e-14, "d round trip at rest")
    # a translated rest shape keeps F = I
    out = O.update_elements(vx + torch.tensor([0.3, -0.1, 0.2]), vv, torch.zeros(1, 3, 3), d0, Dinv, tri, 1e-4)
    _close(out["qF"][0].numpy(), np.eye(3), 1e-14, "F after a translation")
    # rigid rotation at rate w about an axis: vertices rotated by theta = |w| dt,
    # C = the rate's skew matrix; the normal column is advected by the second-
    # order Taylor map, so F - Rot = O(theta^3): halving theta divides it by 8
    rng = np.random.default_rng(9)
    e1, e2 = rng.standard_normal(3) * 0.01, rng.standard_normal(3) * 0.01
    n0 = np.cross(e1, e2)
    n0 /= np.linalg.norm(n0)
    x0 = np.array([0.5, 0.5, 0.5])
    d_rest = torch.tensor(np.stack([e1, e2, n0], 1))[None]
    Dinv = torch.linalg.inv(d_rest)
    axis = [0.3, 1.0, -0.4]
    dt = 1e-4
    errs = []
    for theta in (0.08, 0.04):
        Rot, Kx = _rodrigues(axis, theta)
        verts = torch.tensor(np.stack([x0, x0 + Rot @ e1, x0 + Rot @ e2]))
        C = torch.tensor(Kx * (theta / dt))[None]
        out = O.update_elements(verts, torch.zeros(3, 3), C, d_rest, Dinv, tri, dt)
        errs.append(np.abs(out["qF"][0].numpy() - Rot).max())
        _close(out["qd"][0].numpy(), (out["qF"][0] @ d_rest[0]).numpy(), 1e-12, "d = F inverse(Dinv)")
        assert abs(np.linalg.norm(out["qd"][0, :, 2].numpy()) - 1) < theta ** 3
    assert 6.0 < errs[0] / errs[1] < 10.0, errs


# ----------------------------------------------------------------------------
# grasper velocity (simulation.py:160-165)
# ----------------------------------------------------------------------------
def pin_grasper(O):
    """Frames of 5 substeps, 3 frames per window, 2 grasp points of 2 particles,
    6 trajectory frames traj[k, p] = (k^2, 10 p + k, -k p): the velocity is
    (traj[frame + 1] - traj[frame]) / (5 dt), frame = s // 5 + 3 g, zero
    once frame + 1 reaches the end."""
    dt = 1e-4
    K = O.Consts(G=16, dt=dt, dx=1 / 16, inv_dx=16.0, bound=3, gravity=-9.81, vertex_p_mass=1.0,
                 quadrature_p_mass=1.0, quadrature_p_vol=1.0, cube_p_mass=1.0, operator_p_mass=1.0,
                 steps_per_frame=5, frames_per_window=3, group_size=2, n_actions=6)
    traj = torch.tensor([[[k * k, 10 * p + k, -k * p] for p in range(2)] for k in range(6)], dtype=torch.float64)
    static = {"traj": traj}
    iv = 5 * dt
    # (s, g) -> frame
    for s, g, frame in ((4, 0, 0), (9, 0, 1), (10, 0, 2), (0, 1, 3), (14, 1, 5), (9, 1, 4), (10, 1, 5)):
        v = O.grasper_velocity(static, K, s, g, 4, torch.float64).numpy()
        for q in range(4):
            p = q // 2
            if frame + 1 < 6:
                exp = np.array([(frame + 1) ** 2 - frame ** 2, 1.0, -p]) / iv
            else:
                exp = np.zeros(3)
            assert np.abs(v[q] - exp).max() <= 1e-9 * max(1.0, np.abs(exp).max()), (s, g, q, v[q], exp)


# ----------------------------------------------------------------------------
# the optional live return mapping (simulation.py:408-428)
# ----------------------------------------------------------------------------
def pin_return_map(O):
    """Branch by branch on literal R's (ks = 100 or 0, kc = 80, cf = 0.5): a
    compressed element with shear (R02, R12) = (0.3, 0.4), R22 = 0.6 has
    k = 1.25, k^2 |u|^2 = 0.390625 > (cf (R22-1)^2)^2 = 0.0064, so the shear
    is scaled by 0.08 / (1.25 * 0.5) = 0.128 to (0.0384, 0.0512)."""
    base = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.0], [0.0, 0.0, 0.0]])
    cases = [  # (ks, R02, R12, R22) -> mapped (R02, R12, R22)
        (100.0, 0.3, 0.4, 0.6, (0.0384, 0.0512, 0.6)),
        (100.0, 0.01, 0.0, 0.6, (0.01, 0.0, 0.6)),
        (100.0, 0.3, 0.4, 1.3, (1e-7, 1e-7, 1.0)),
        (100.0, 0.3, 0.4, 1.0, (0.0, 0.0, 1.0)),       # R22 = 1: the cone has zero width
        (100.0, 0.3, 0.4, -0.4, (1e-7, 1e-7, -0.4)),
        (100.0, 0.3, 0.4, -1.5, (1e-7, 1e-7, -1.0)),
        (0.0, 0.3, 0.4, 0.5, (1e-7, 1e-7, 0.5)),
        (0.0, 0.3, 0.4, 1.5, (1e-7, 1e-7, 1.0)),
    ]
    for ks, r02, r12, r22, exp in cases:
        R = torch.tensor(base)[None].clone()
        R[0, 0, 2], R[0, 1, 2], R[0, 2, 2] = r02, r12, r22
        out = O.return_map(R, _t(ks), _t(80.0), _t(0.5))[0].numpy()
        want = base.copy()
        want[0, 2], want[1, 2], want[2, 2] = exp
        assert np.abs(out - want).max() < 1e-12, (ks, r02, r12, r22, out[:, 2])


PINS = {"stress_upper": pin_stress_upper, "return_map": pin_return_map, "stress_inplane": pin_stress_inplane,
        "stress_rotation": pin_stress_rotation, "forces": pin_forces, "grid_op": pin_grid_op,
        "g2p": pin_g2p, "damp": pin_damp, "update": pin_update, "grasper": pin_grasper}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name](O_REAL)


# (expression in oracle/sim_torch.py, planted mistake, pin that must reject it)
MUTATIONS = [
    ("a01 = ks * R02 * R12", "a01 = ks * R02 * R02", "stress_upper"),
    ("a12 = ks * R12 * R22", "a12 = ks * R02 * R22", "stress_upper"),
    ("a02 = ks * R02 * R22", "a02 = ks * R12 * R22", "stress_upper"),
    ("(-kc) * ((1 - R22) ** 2)", "(kc) * ((1 - R22) ** 2)", "stress_upper"),
    ("a22 = ((R22 <= 1)", "a22 = ((R22 >= 1)", "stress_upper"),
    ("a00 = ks * R02 * R02 ", "a00 = ks * R02 * R02 + E * R[:, 0, 0] * (R[:, 0, 0] - 1) ", "stress_upper"),
    ("la = E * nu / ((1 + nu) * (1 - 2 * nu))", "la = E * nu / ((1 + nu) * (1 - nu))", "stress_upper"),
    ("mu = E / (2 * (1 + nu))", "mu = E / (1 + nu)", "stress_upper"),
    ("E, ks, kc = params[\"E\"], params[\"ks\"], params[\"kc\"]",
     "E, kc, ks = params[\"E\"], params[\"ks\"], params[\"kc\"]", "stress_upper"),
    ("Q @ A @ inv3(R).transpose(-1, -2)", "Q @ A @ inv3(R)", "stress_upper"),
    ("inv2(Ft.transpose(-1, -2))", "inv2(Ft)", "stress_inplane"),
    ("2 * mu_ * (Ft - Rr)", "2 * mu_ * (Ft - Rr.transpose(-1, -2))", "stress_inplane"),
    ("P = Q @ Pfc3 + ", "P = Pfc3 @ Q + ", "stress_rotation"),
    ("raw1 = torch.stack([Dinv[:, 0, 0], d[:, 0, 1], d[:, 0, 2]], -1)",
     "raw1 = torch.stack([d[:, 0, 0], d[:, 0, 1], d[:, 0, 2]], -1)", "forces"),
    ("raw2 = torch.stack([Dinv[:, 1, 0], d[:, 1, 1], d[:, 1, 2]], -1)",
     "raw2 = torch.stack([Dinv[:, 0, 1], d[:, 1, 1], d[:, 1, 2]], -1)", "forces"),
    ("(1, f2, -1.0), (2, f3, -1.0)", "(1, f3, -1.0), (2, f2, -1.0)", "forces"),
    ("(0, f1, 1.0)", "(0, f1, -1.0)", "forces"),
    ("Pcol3 = P[:, :, 2]", "Pcol3 = P[:, 2, :]", "forces"),
    ("dcol3 = d[:, :, 2]", "dcol3 = d[:, 2, :]", "forces"),
    ("w[:, o[:, 0], 0] * xjw[:, o[:, 1], 1] * w[:, o[:, 2], 2]",
     "w[:, o[:, 0], 0] * xjw[:, o[:, 1], 0] * w[:, o[:, 2], 2]", "forces"),
    ("(gi > G - b)", "(gi >= G - b)", "grid_op"),
    ("(gi < b)", "(gi <= b)", "grid_op"),
    ("(gk < b) & (v_out[:, 2] < 0)", "(gk < b) & (v_out[:, 2] > 0)", "grid_op"),
    ("(gj > G - b) & (v_out[:, 1] > 0)", "(gj > G - b) & (v_out[:, 0] > 0)", "grid_op"),
    ("torch.where(c5, torch.zeros_like(v_out[:, 1]), v_out[:, 1])", "v_out[:, 1]", "grid_op"),
    ("v_out = v_out * (1 - damping * dt)", "v_out = v_out * (1 + damping * dt)", "grid_op"),
    ("(gv[..., :, None] * dp[..., None, :])", "(dp[..., :, None] * gv[..., None, :])", "g2p"),
    ("return C_skew + (1 - 0.95) * C_sym", "return C_sym + (1 - 0.95) * C_skew", "damp"),
    ("col3n = col3 + (M @ col3[..., None])[..., 0]",
     "col3n = col3 + (M.transpose(-1, -2) @ col3[..., None])[..., 0]", "update"),
    ("M = dt * C + 0.5 * dt * dt * C @ C", "M = dt * C", "update"),
    ("new_d = cols(x2 - x1, x3 - x1, col3n)", "new_d = cols(x3 - x1, x2 - x1, col3n)", "update"),
    ("\"qv\": (1 / 3) * (v1 + v2 + v3)", "\"qv\": (1 / 3) * (v1 + v2 + v2)", "update"),
    ("out[\"qd\"] = new_F @ inv3(Dinv)", "out[\"qd\"] = new_F @ Dinv", "update"),
    ("b2 = ~b1 & (R22 > 1)", "b2 = ~b1 & (R22 >= 1)", "return_map"),
    ("scale = cf * ((R22 - 1) ** 2) / (k * torch.sqrt(safe))", "scale = cf * ((R22 - 1) ** 2) / (k * safe)",
     "return_map"),
    ("torch.where(b3, torch.maximum(R22, -one), R22)", "torch.where(b3, R22, R22)", "return_map"),
    ("    k = ks / kc\n    act", "    k = kc / ks\n    act", "return_map"),
    ("frame = f // K.steps_per_frame + K.frames_per_window * g",
     "frame = (f + 1) // K.steps_per_frame + K.frames_per_window * g", "grasper"),
    ("if frame + 1 < K.n_actions:", "if frame < K.n_actions:", "grasper"),
    ("gid = torch.arange(n_op) // K.group_size", "gid = torch.arange(n_op) % K.group_size", "grasper"),
    ("(traj[frame + 1, gid] - traj[frame, gid]) / interval",
     "(traj[frame + 1, gid] - traj[frame, gid]) / K.dt", "grasper"),
]


def _mutant(old, new):
    src = open(O_REAL.__file__).read()
    assert src.count(old) == 1, f"expression not found exactly once: {old!r}"
    name = "oracle_mutant"
    mod = types.ModuleType(name)
    mod.__file__ = O_REAL.__file__
    sys.modules[name] = mod          # dataclasses look their module up while the class is built
    try:
        exec(compile(src.replace(old, new), O_REAL.__file__, "exec"), mod.__dict__)
    finally:
        del sys.modules[name]
    return mod


@pytest.mark.parametrize("old,new,pin", MUTATIONS, ids=[f"{p}:{i}" for i, (_, _, p) in enumerate(MUTATIONS)])
def test_pin_catches_mutation(old, new, pin):
    mod = _mutant(old, new)
    with pytest.raises(Exception):
        PINS[pin](mod)
