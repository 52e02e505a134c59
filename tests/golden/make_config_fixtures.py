"""Writes the FP64-oracle fixtures of the GPU workload tests:

  c2_grad_fd.npz      64x64 cloth, 128^3 grid, 400 substeps: targets from a
                      ground-truth run (E = 8e3, nu = 0.25), the loss, its
                      gradient w.r.t. E and nu by autograd (reference SVD
                      backward) and by central differences at h = 1e-3 and
                      1e-4 relative;
  c4_positions.npz    96x96 cloth compressed and folded, 128^3 grid, 500
                      substeps: vertex positions at substeps 250 and 500 and
                      central differences (h = 1e-4 and 1e-5) w.r.t. E and
                      nu of a loss on state 498;
  reference_shaped.npz  the 15^3, scene_scaling 0.5, 300-substep-frame
                      scene: a 320-substep window's loss and gradient by
                      autograd of the oracle in FP64 and in FP32;
  c5_reduced.npz      40x40 cloth on the 256^3 grid with per-triangle
                      lognormal moduli, 100 substeps: the gradient w.r.t.
                      every triangle's (E, shearing, contact) by autograd and
                      the central difference of the loss along a seeded
                      direction.

Autograd runs through torch.utils.checkpoint segments (the rollout is
recomputed segment by segment in the backward pass), so that memory stays
bounded.  Usage, from the repository root:
    python tests/golden/make_config_fixtures.py c2|c2fd|c4|c4fd|c5|refshape
(c2: about 10 minutes on 8 cores; c2fd, which adds h = 1e-5 differences to
c2's fixture: 5; c4: about 30; c4fd, the same for c4: 30; c5: about 1;
refshape: about 1)."""
import os
import sys
import time

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from diffcp_amd import scene  # noqa: E402
from oracle import sim_torch as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
KEYS = O.STATE_KEYS


def rollout_segments(inst, params, n, seg, svd_grad="reference", keep=()):
    """State after n substeps, autograd through checkpointed segments; also
    returns detached copies of the vertex positions at the substeps in `keep`."""
    K = O.Consts.from_instance(inst)
    st, static = O.initialize_state(inst)
    kept = {}
    f = 0
    pnames = sorted(params)
    while f < n:
        m = min(seg, n - f)

        def run(f0, m_, *flat):
            s = dict(zip(KEYS, flat[:len(KEYS)]))
            p = dict(zip(pnames, flat[len(KEYS):]))
            for t in range(f0, f0 + m_):
                s = O.substep(s, static, p, K, t, 0, svd_grad)
            return tuple(s[k] for k in KEYS)

        flat = tuple(st[k] for k in KEYS) + tuple(params[k] for k in pnames)
        out = checkpoint(run, f, m, *flat, use_reentrant=False)
        st = dict(zip(KEYS, out))
        f += m
        if f in keep:
            kept[f] = st["vx"].detach().numpy().copy()
    return st, kept


def loss_at(inst, params, n_loss, targets, seg, svd_grad="reference"):
    st, _ = rollout_segments(inst, params, n_loss, seg, svd_grad)
    loss, _ = O.compute_loss(st["vx"], torch.as_tensor(targets))
    return loss


def gt_targets(inst, n_loss, seed, E=8e3, nu=0.25, elem_scale=None):
    p = O.make_params(inst)
    p["E"] = torch.tensor(E, dtype=torch.float64)
    p["nu"] = torch.tensor(nu, dtype=torch.float64)
    if elem_scale is not None:
        p["elem"] = p["elem"] * torch.tensor(elem_scale)
    with torch.no_grad():
        st, _ = O.rollout(inst, p, n_loss)
    return scene.sample_targets(st["vx"].numpy(), 1024, seed)


def grad_and_fd(inst, S, K, targets, names=("E", "nu"), rels=(1e-3, 1e-4)):
    n_loss = S - 2
    p = O.make_params(inst, requires_grad=True)
    loss = loss_at(inst, p, n_loss, targets, K)
    loss.backward()
    out = {"loss": np.float64(loss.detach()), "grad": np.array([float(p[k].grad) for k in names])}
    for rel in rels:
        fd = []
        for name in names:
            th = float(getattr(inst, name))
            h = rel * abs(th)
            vals = []
            for sgn in (1.0, -1.0):
                q = O.make_params(inst)
                q[name] = torch.tensor(th + sgn * h, dtype=torch.float64)
                with torch.no_grad():
                    vals.append(float(loss_at(inst, q, n_loss, targets, K)))
            fd.append((vals[0] - vals[1]) / (2 * h))
        out[f"fd_{rel:g}"] = np.array(fd)
    return out


def make_c2():
    w = scene.config_c2()
    inst = w.instances[0]
    S, K = w.substeps, w.checkpoint_every
    targets = gt_targets(inst, S - 2, w.seed * 100)
    out = grad_and_fd(inst, S, K, targets)
    out.update(targets=targets, substeps=np.int64(S), checkpoint_every=np.int64(K),
               E=np.float64(inst.E), nu=np.float64(inst.nu))
    np.savez(os.path.join(HERE, "c2_grad_fd.npz"), **out)
    return out


def fd_1e5(inst, S, K, targets):
    """Central differences at h = 1e-5 relative of the loss on state S - 2."""
    fd = []
    for name in ("E", "nu"):
        th = float(getattr(inst, name))
        h = 1e-5 * abs(th)
        vals = []
        for sgn in (1.0, -1.0):
            q = O.make_params(inst)
            q[name] = torch.tensor(th + sgn * h, dtype=torch.float64)
            with torch.no_grad():
                vals.append(float(loss_at(inst, q, S - 2, targets, K)))
        fd.append((vals[0] - vals[1]) / (2 * h))
    return np.array(fd)


def refine_c2():
    """Adds central differences at h = 1e-5 relative to c2_grad_fd.npz (the
    FD truncation error at 1e-3 and 1e-4 is visible for nu)."""
    path = os.path.join(HERE, "c2_grad_fd.npz")
    out = dict(np.load(path))
    inst = scene.config_c2().instances[0]
    out["fd_1e-05"] = fd_1e5(inst, int(out["substeps"]), int(out["checkpoint_every"]), out["targets"])
    np.savez(path, **out)
    return {k: v for k, v in out.items() if k.startswith("fd") or k == "grad"}


def refine_c4():
    """Adds central differences at h = 1e-5 relative to c4_positions.npz."""
    path = os.path.join(HERE, "c4_positions.npz")
    out = dict(np.load(path))
    w = scene.config_c4()
    out["fd_1e-05"] = fd_1e5(w.instances[0], int(out["substeps"]), w.checkpoint_every, out["targets"])
    np.savez_compressed(path, **out)
    return {k: v for k, v in out.items() if k.startswith("fd") or k == "loss"}


def make_c4():
    """Positions from one forward run; the gradient by central differences of
    forward runs (h = 1e-4 relative; autograd through 500 substeps of this
    cloth exceeds the memory of the build machine even in checkpoint
    segments)."""
    w = scene.config_c4()
    inst = w.instances[0]
    S, K = w.substeps, w.checkpoint_every
    with torch.no_grad():
        _, kept = rollout_segments(inst, O.make_params(inst), S, K, keep=(250, 500))
    targets = gt_targets(inst, S - 2, w.seed * 100) + np.array([0.002, -0.001, 0.0015])
    with torch.no_grad():
        loss = float(loss_at(inst, O.make_params(inst), S - 2, targets, K))
        fd = []
        for name in ("E", "nu"):
            th = float(getattr(inst, name))
            h = 1e-4 * abs(th)
            vals = []
            for sgn in (1.0, -1.0):
                q = O.make_params(inst)
                q[name] = torch.tensor(th + sgn * h, dtype=torch.float64)
                vals.append(float(loss_at(inst, q, S - 2, targets, K)))
            fd.append((vals[0] - vals[1]) / (2 * h))
    out = {"x250": kept[250].astype(np.float32), "x500": kept[500].astype(np.float32), "targets": targets,
           "substeps": np.int64(S), "loss": np.float64(loss), "fd_0.0001": np.array(fd)}
    np.savez_compressed(os.path.join(HERE, "c4_positions.npz"), **out)
    return {k: v for k, v in out.items() if k not in ("x250", "x500", "targets")}


def make_c5():
    inst = scene.large_cloth_reduced_instance()
    S, K = 100, 10
    # ground-truth samples moved by ~1 mm, so that the loss is far above the
    # FP32 position resolution of a 1/256 grid
    targets = gt_targets(inst, S - 2, 400, elem_scale=np.array([1.6, 1.0, 1.0])) + np.array([0.001, 0.0005, -0.001])
    p = O.make_params(inst, requires_grad=True)
    loss = loss_at(inst, p, S - 2, targets, K)
    loss.backward()
    grad = p["elem"].grad.numpy().copy()
    rng = np.random.default_rng(44)
    delta = rng.standard_normal(grad.shape)
    delta /= np.abs(delta).max()
    th = np.asarray(inst.element_moduli)
    h = 1e-4
    vals = []
    for sgn in (1.0, -1.0):
        q = O.make_params(inst)
        q["elem"] = torch.tensor(th * (1 + sgn * h * delta))
        with torch.no_grad():
            vals.append(float(loss_at(inst, q, S - 2, targets, K)))
    fd = (vals[0] - vals[1]) / (2 * h)          # d/de L(theta (1 + e delta)) at e = 0
    out = {"targets": targets, "substeps": np.int64(S), "checkpoint_every": np.int64(K),
           "loss": np.float64(loss.detach()), "grad": grad, "delta": delta, "fd_dir": np.float64(fd),
           "ad_dir": np.float64((grad * th * delta).sum())}
    np.savez_compressed(os.path.join(HERE, "c5_reduced.npz"), **out)
    return {k: v for k, v in out.items() if k in ("loss", "fd_dir", "ad_dir")}


def make_refshape():
    """The reference-shaped scene (15^3 grid, scene_scaling 0.5, 300-substep
    frames): targets 4-6 mm from the state of substep 318, the loss of a
    320-substep window and its gradient w.r.t. E, nu, shearing and contact
    stiffness by autograd of the oracle in FP64 and in FP32."""
    inst = scene.reference_shaped_instance()
    S = 320
    with torch.no_grad():
        st, _ = O.rollout(inst, n_substeps=S - 2)
    targets = st["vx"].numpy() + np.array([0.004, 0.006, -0.003])
    out = {"targets": targets, "substeps": np.int64(S)}
    for tag, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        p = O.make_params(inst, dtype=dt, requires_grad=True)
        loss, _ = O.tape_loss(inst, p, S, targets, dtype=dt)
        loss.backward()
        out[f"loss_{tag}"] = np.float64(loss.detach())
        out[f"grad_{tag}"] = np.array([float(p[k].grad) for k in ("E", "nu", "ks", "kc")])
    np.savez(os.path.join(HERE, "reference_shaped.npz"), **out)
    return {k: v for k, v in out.items() if k != "targets"}


if __name__ == "__main__":
    t = time.time()
    which = sys.argv[1]
    res = {"c2": make_c2, "c2fd": refine_c2, "c4": make_c4, "c4fd": refine_c4, "c5": make_c5,
           "refshape": make_refshape}[which]()
    print(which, {k: v for k, v in res.items() if k != "targets"}, f"{time.time() - t:.0f} s", flush=True)
