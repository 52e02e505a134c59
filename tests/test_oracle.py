"""Pins of the CPU oracle (oracle/sim_torch.py) against closed forms, invariants,
library routines and central finite differences.

The reference ships no tests, golden vectors or fixtures for the substep
(its data/ files are inputs only), so every oracle function is pinned here by
mathematics chosen so that a dropped term, a sign or index error or a
transposed operand fails a check.
"""
import itertools

import numpy as np
import pytest
import torch

from diffcp_amd import scene
from oracle import reduced as R
from oracle import sim_torch as O

torch.set_default_dtype(torch.float64)


def test_bspline_partition_of_unity_and_first_moment():
    x = torch.rand(500, 3) * 0.8 + 0.1
    base, fx, w, dw = O.bspline(x, 64.0)
    assert torch.allclose(w.sum(1), torch.ones(500, 3), atol=1e-14)
    nodes = torch.arange(3.0)[None, :, None]
    # sum_i w_i (i - fx) = 0 (linear reproduction)
    assert torch.allclose((w * (nodes - fx[:, None, :])).sum(1), torch.zeros(500, 3), atol=1e-13)
    # sum_i w_i (i - fx)^2 = 1/4 (quadratic B-spline second moment)
    assert torch.allclose((w * (nodes - fx[:, None, :]) ** 2).sum(1), torch.full((500, 3), 0.25), atol=1e-13)
    # fx in [0.5, 1.5): the stencil is centred
    assert (fx >= 0.5).all() and (fx < 1.5).all()


def test_bspline_derivative_matches_fd():
    x = torch.rand(100, 3) * 0.5 + 0.2
    inv_dx = 32.0
    _, fx, w, dw = O.bspline(x, inv_dx)
    h = 1e-7
    _, _, wp, _ = O.bspline(x + h / inv_dx, inv_dx)
    _, _, wm, _ = O.bspline(x - h / inv_dx, inv_dx)
    fd = (wp - wm) / (2 * h)
    ok = (O.bspline(x + h / inv_dx, inv_dx)[0] == O.bspline(x - h / inv_dx, inv_dx)[0]).all(-1)
    assert torch.allclose(fd[ok], dw[ok], atol=1e-6)


def test_binning_keys_truncate_toward_zero_in_fp32():
    x = np.array([[0.5 / 64, 1.0 / 64, 63.49 / 64], [0.1, 0.2, 0.3], [-0.001, 0.0, 0.004]], np.float64)
    k = O.binning_keys_fp32(x, 64)
    # trunc(f32(x)*64 - 0.5): 0, 0, 62 / 5, 12, 18 / trunc(-0.564)=0, trunc(-0.5)=0, trunc(-0.244)=0
    assert k.tolist() == [[0, 0, 62], [5, 12, 18], [0, 0, 0]]


def test_qr_matches_numpy_and_reconstructs():
    torch.manual_seed(0)
    F = torch.eye(3) + 0.3 * torch.randn(200, 3, 3)
    Q, R = O.qr_gram_schmidt(F)
    assert torch.allclose(Q @ R, F, atol=1e-12)
    # R upper triangular up to the eps term
    assert R[:, 1, 0].abs().max() < 1e-6 and R[:, 2, 0].abs().max() < 1e-6 and R[:, 2, 1].abs().max() < 1e-6
    Qn, Rn = np.linalg.qr(F.numpy())
    s = np.sign(np.diagonal(Rn, axis1=1, axis2=2))
    Qn = Qn * s[:, None, :]
    # eps = 1e-9 in the norm shrinks a column by ~eps / (2 |h|^2) relative
    assert np.abs(Q.numpy() - Qn).max() < 1e-5


def test_inv3_matches_linalg():
    A = torch.randn(100, 3, 3) + 3 * torch.eye(3)
    assert torch.allclose(O.inv3(A), torch.linalg.inv(A), atol=1e-12)


def test_polar_matches_numpy_svd():
    torch.manual_seed(1)
    Ft = torch.randn(300, 2, 2)
    Rr = O.polar_rotation_closed(Ft)
    U, S, Vh = np.linalg.svd(Ft.numpy())
    # polar rotation of the rotation-variant SVD: det(U V^T) = +1 and R^T F symmetric with trace >= 0
    assert torch.allclose(torch.linalg.det(Rr), torch.ones(300), atol=1e-12)
    Sym = Rr.transpose(-1, -2) @ Ft
    assert torch.allclose(Sym, Sym.transpose(-1, -2), atol=1e-12)
    assert (Sym.diagonal(dim1=1, dim2=2).sum(-1) >= 0).all()
    pos = torch.linalg.det(Ft) > 0
    ref = torch.as_tensor(U @ Vh)
    assert torch.allclose(Rr[pos], ref[pos], atol=1e-10)


def test_reference_svd_backward_equals_polar_derivative_away_from_degeneracy():
    torch.manual_seed(2)
    Ft = torch.eye(2) + 0.2 * torch.randn(200, 2, 2)
    gR = torch.randn(200, 2, 2)
    g_ref = O.reference_svd_backward(Ft, gR)
    Fg = Ft.clone().requires_grad_(True)
    (O.polar_rotation_closed(Fg) * gR).sum().backward()
    assert torch.allclose(g_ref, Fg.grad, atol=1e-9)
    # central FD of <gR, R(F)>
    h = 1e-6
    fd = torch.zeros_like(Ft)
    for i, j in itertools.product(range(2), range(2)):
        E = torch.zeros(2, 2)
        E[i, j] = h
        fd[:, i, j] = ((O.polar_rotation_closed(Ft + E) - O.polar_rotation_closed(Ft - E)) * gR).sum((1, 2)) / (2 * h)
    assert torch.allclose(fd, g_ref, atol=1e-7)


def test_reference_svd_backward_is_exactly_zero_at_degeneracy():
    Ft = torch.eye(2).expand(10, 2, 2) * torch.linspace(0.5, 2.0, 10)[:, None, None]
    gR = torch.randn(10, 2, 2)
    g = O.reference_svd_backward(Ft, gR)
    assert g.abs().max() < 1e-9
    # the exact derivative is not zero there
    Fg = Ft.clone().requires_grad_(True)
    (O.polar_rotation_closed(Fg) * gR).sum().backward()
    assert Fg.grad.abs().max() > 1e-3


def test_clamped_scale_closed_form():
    """reference backward = exact * min(1, |s1^2 - s0^2| / 1e-6) (used by the HIP kernels)."""
    torch.manual_seed(3)
    for scale in (1e-3, 1e-6, 1e-7, 1e-8):
        Ft = torch.eye(2) + scale * torch.randn(400, 2, 2)
        gR = torch.randn(400, 2, 2)
        lit = O.reference_svd_backward(Ft, gR)
        Fg = Ft.clone().requires_grad_(True)
        (O.polar_rotation_closed(Fg) * gR).sum().backward()
        a, b, c, d = Ft[:, 0, 0], Ft[:, 0, 1], Ft[:, 1, 0], Ft[:, 1, 1]
        r = torch.sqrt((a + d) ** 2 + (c - b) ** 2)
        delta = 2 * r * torch.sqrt(((a - d) / 2) ** 2 + ((c + b) / 2) ** 2)
        sv = torch.linalg.svdvals(Ft)
        assert torch.allclose(delta, sv[:, 0] ** 2 - sv[:, 1] ** 2, atol=1e-12)
        closed = Fg.grad * torch.clamp(delta / 1e-6, max=1.0)[:, None, None]
        assert (lit - closed).abs().max() < 1e-8 * max(1.0, closed.abs().max().item())


def test_rest_state_stress_vanishes():
    inst = scene.tiny_instance()
    st, static = O.initialize_state(inst)
    params = O.make_params(inst)
    P, Q, Rm, Ft = O.element_stress(st["qF"], params, "stable")
    # F = I: in-plane P_fc = 2mu(Ft - R) + la (J-1) J Ft^-T with Ft = sqrt(1+1e-9) I -> O(1e-9 E)
    assert P.abs().max() < 1e-4 * inst.E * 1e-3
    # R22 = sqrt(1+1e-9) > 1: the contact term is off, shear terms are 0
    assert (Rm[:, 2, 2] > 1).all()


def test_stress_is_gradient_of_energy_in_plane():
    """The oracle's in-plane stress is d/dFt of psi = mu |Ft - R|^2 + la/2 (J-1)^2
    (fixed-corotated, simulation.py:236-237), with R the polar rotation from
    torch's SVD: for an upper-triangular F with a positive diagonal QR gives
    Q = I and Ft = F[:2, :2], and the contact/shear terms vanish (R02 = R12 = 0,
    R22 > 1)."""
    torch.manual_seed(4)
    n = 50
    F = torch.eye(3).repeat(n, 1, 1)
    F[:, 0, 0] = 0.7 + 0.6 * torch.rand(n)
    F[:, 1, 1] = 0.7 + 0.6 * torch.rand(n)
    F[:, 0, 1] = 0.4 * torch.randn(n)
    E, nu = 3e3, 0.35
    mu, la = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    params = {"E": torch.tensor(E), "nu": torch.tensor(nu), "ks": torch.tensor(50.0), "kc": torch.tensor(70.0)}
    P = O.element_stress(F, params, "stable")[0]
    Ft = F[:, :2, :2].clone().requires_grad_(True)
    U, S, Vh = torch.linalg.svd(Ft)
    R = U @ Vh
    J = torch.linalg.det(Ft)
    psi = (mu * ((Ft - R) ** 2).sum((1, 2)) + 0.5 * la * (J - 1) ** 2).sum()
    (g,) = torch.autograd.grad(psi, Ft)
    assert torch.allclose(P[:, :2, :2], g, atol=1e-6 * E)
    assert P[:, 2, :].abs().max() < 1e-6 * E and P[:, :2, 2].abs().max() < 1e-6 * E


def test_initial_state_literal():
    inst = scene.tiny_instance()
    st, static = O.initialize_state(inst)
    d = st["qd"]
    # third column: unit normal times (1 - ~eps) and +y for the synthetic mesh
    n = d[:, :, 2]
    assert torch.allclose(n[:, 1], torch.ones_like(n[:, 1]), atol=1e-4)
    assert torch.allclose(static["Dinv"] @ d, torch.eye(3).expand_as(d), atol=1e-10)
    assert torch.allclose(st["qx"], st["vx"][static["tri"]].mean(1), atol=1e-14)


def test_p2g_conserves_mass_and_momentum():
    inst = scene.tiny_instance(n=5, substeps=6)
    K = O.Consts.from_instance(inst)
    states, static = O.rollout(inst, n_substeps=3, keep_states=True)
    st = states[3]
    params = O.make_params(inst)
    _, grid = O.substep(st, static, params, K, 3, 0, "stable", return_grid=True)
    total_m = (K.cube_p_mass * inst.n_cube + K.operator_p_mass * inst.n_grasper
               + K.vertex_p_mass * inst.n_vertices + K.quadrature_p_mass * inst.n_elements)
    assert abs(grid["m"].sum().item() - total_m) < 1e-9 * total_m
    # momentum: affine terms and internal forces integrate to zero
    vop = O.grasper_velocity(static, K, 3, 0, inst.n_grasper, torch.float64)
    grav = torch.tensor([0.0, K.dt * K.gravity, 0.0])
    mom = (K.operator_p_mass * vop.sum(0) + K.vertex_p_mass * (st["vv"] + grav).sum(0)
           + K.quadrature_p_mass * st["qv"].sum(0))
    assert torch.allclose(grid["mv"].sum(0), mom, atol=1e-9 * (1 + mom.abs().max().item()))


def test_g2p_reproduces_rigid_motion():
    """The oracle's APIC gather transfers a rigid translation + rotation exactly (v and C)."""
    G = 32
    x = torch.rand(40, 3) * 0.5 + 0.25
    base, _, _, _ = O.bspline(x, G)
    cell = (base[:, None, :] + O.OFFSETS[None]).double() / G
    omega = torch.tensor([0.3, -1.0, 0.7])
    v0 = torch.tensor([0.1, 0.2, -0.3])
    Om = torch.tensor([[0, -omega[2], omega[1]], [omega[2], 0, -omega[0]], [-omega[1], omega[0], 0]])
    vp, Cp = O.g2p_gather(x, v0 + cell @ Om.T, float(G))
    assert torch.allclose(vp, v0 + x @ Om.T, atol=1e-12)
    assert torch.allclose(Cp, Om.expand_as(Cp), atol=1e-12)


def test_free_fall_closed_form():
    """Vertex particles far from walls, cube and grasper fall with v = s * dt * g.

    Only vertex particles receive gravity (simulation.py:181-182); element
    particles carry mass without it, so the closed form holds for a cloth
    without triangles.
    """
    # a heavy cloth makes the 1/(m + 1e-9) normalisation (simulation.py:277) negligible
    inst = scene.tiny_instance(n=4, n_grid=32, substeps=6, side=0.1, density=2000.0)
    inst.area = inst.surface_area
    inst.triangles = inst.triangles[:0]
    inst.vertex_x[:, 1] += 0.4
    inst.grasper_x[:, 1] += 0.7   # grasper above, outside the cloth stencils
    inst.grasper_x[:, 0] = 0.85
    inst.trajectory[:] = inst.trajectory[:1]
    states, _ = O.rollout(inst, n_substeps=5, keep_states=True)
    for s in (1, 3, 5):
        vy = states[s]["vv"][:, 1]
        assert torch.allclose(vy, torch.full_like(vy, s * inst.dt * inst.gravity), rtol=1e-7, atol=1e-12)


def test_loss_brute_force_and_first_argmin():
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(0)
    x = rng.random((50, 3))
    t = rng.random((20, 3))
    t[3] = x[7]
    x[9] = x[7]          # a tie: the first index (7) wins
    loss, idx = O.compute_loss(torch.as_tensor(x), torch.as_tensor(t))
    d = cdist(x, t) ** 2 / 3
    assert abs(loss.item() - d.min(0).mean()) < 1e-14
    assert idx[3].item() == 7
    assert (idx.numpy() == d.argmin(0)).all()


def test_rollout_gradient_matches_central_fd():
    inst = scene.tiny_instance(n=4, substeps=8)
    tgt = inst.vertex_x + np.array([0.003, 0.01, -0.002])
    for mode in ("stable", "reference"):
        p = O.make_params(inst, requires_grad=True)
        loss, _ = O.tape_loss(inst, p, 8, tgt, svd_grad=mode)
        loss.backward()
        for name, attr in (("E", "E"), ("nu", "nu"), ("ks", "shearing_stiffness")):
            h = 1e-4 * abs(getattr(inst, attr))
            vals = []
            for sgn in (1, -1):
                q = O.make_params(inst)
                q[name] = q[name] + sgn * h
                with torch.no_grad():
                    vals.append(O.tape_loss(inst, q, 8, tgt, svd_grad=mode)[0].item())
            fd = (vals[0] - vals[1]) / (2 * h)
            ad = p[name].grad.item()
            assert abs(fd - ad) <= 1e-5 * abs(fd) + 1e-18, (mode, name, fd, ad)


def test_reduced_bridge_is_identity_on_rollouts():
    inst = scene.tiny_instance(n=4, substeps=6)
    K = O.Consts.from_instance(inst)
    states, static = O.rollout(inst, n_substeps=4, keep_states=True)
    params = O.make_params(inst)
    red = R.initial_reduced(states[0])
    for f in range(4):
        red = R.reduced_substep(red, static, params, K, f)
        ref = R.reduce(states[f + 1], states[f])
        for k in R.RED_KEYS:
            assert torch.allclose(red[k], ref[k], atol=1e-12), (f, k)
