"""C-ABI arrays in device memory (torch CUDA tensors' data_ptr): the state
slot through the device export/import kernels, targets, parameters and
per-element moduli read from device memory, losses and gradients written to
it, and the torch.autograd.Function on CUDA tensors.  Each is compared with
the same call on host arrays: bit-exact where the same kernels run on the
same inputs once (state transfer, parameters), 1e-5 relative on window
results (float-atomic summation order differs between runs)."""
import numpy as np
import pytest
import torch

from diffcp_amd import scene
from diffcp_amd._abi import DiffcpError

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _insts():
    a = scene.tiny_instance(n=6, n_grid=16, substeps=40, seed=1)
    b = scene.tiny_instance(n=5, n_grid=16, substeps=40, seed=2)
    b.E = 9e3
    for i, inst in enumerate((a, b)):
        rng = np.random.default_rng(10 + i)
        inst.targets = inst.vertex_x[rng.choice(inst.n_vertices, 12, replace=False)] + rng.normal(0, 0.01, (12, 3))
    return [a, b]


def _engine(insts, S=24, **kw):
    from diffcp_amd.engine import Engine
    eng = Engine(insts, max_states=Engine.slots_for(S, S), **kw)
    eng.reset(0)
    return eng


def test_state_download_and_upload_through_device_memory():
    eng = _engine(_insts())
    for f in range(5):
        eng.substep(f, 0, f % 2, 1 - f % 2)
    host = eng.download_state(1)
    out = {k: torch.full(v.shape, -7.0, dtype=torch.float32, device=DEV) for k, v in host.items()}
    eng.download_state(1, out=out)
    for k in host:
        assert np.array_equal(out[k].cpu().numpy(), host[k]), k
    # a subset downloads only those arrays
    sub = {"vertex_x": torch.zeros(host["vertex_x"].shape, dtype=torch.float32, device=DEV)}
    eng.download_state(1, out=sub)
    assert np.array_equal(sub["vertex_x"].cpu().numpy(), host["vertex_x"])
    # upload from device tensors (element C left out: zero) and read back on the host
    up = {k: (t * 0.5 + 0.25).contiguous() for k, t in out.items() if k != "element_C"}
    eng.upload_state(0, **up)
    back = eng.download_state(0)
    for k in host:
        want = np.zeros_like(host[k]) if k == "element_C" else up[k].cpu().numpy()
        assert np.array_equal(back[k], want), k


def test_mixed_host_and_device_state_arrays_are_refused():
    eng = _engine(_insts())
    host = eng.download_state(0)
    mixed = {"vertex_x": torch.zeros(host["vertex_x"].shape, dtype=torch.float32, device=DEV),
             "vertex_v": torch.zeros(host["vertex_v"].shape, dtype=torch.float32)}
    with pytest.raises(DiffcpError, match="all host or all device"):
        eng.download_state(0, out=mixed)


def test_window_inputs_and_outputs_in_device_memory():
    insts = _insts()
    S = 24
    eng = _engine(insts, S)
    l_h, g_h, _ = eng.rollout_grad(S, 0, S, use_graph=True)

    # parameters and targets from device tensors, results into device tensors
    params = torch.tensor(np.stack([i.params() for i in insts]), dtype=torch.float64, device=DEV)
    eng.set_params(params)
    tg = torch.tensor(np.concatenate([i.targets for i in insts]), dtype=torch.float64, device=DEV)
    eng.set_targets(tg, counts=[len(i.targets) for i in insts])
    eng.reset(0)
    loss = torch.zeros(2, dtype=torch.float64, device=DEV)
    grads = torch.zeros(2, 6, dtype=torch.float64, device=DEV)
    eng.rollout_grad(S, 0, S, use_graph=True, loss_out=loss, grads_out=grads)
    assert _rel(loss.cpu().numpy(), l_h) < 1e-5
    assert _rel(grads.cpu().numpy(), g_h) < 1e-5
    pg = torch.zeros(2, 6, dtype=torch.float64, device=DEV)
    eng.param_grads(out=pg)
    assert _rel(pg.cpu().numpy(), grads.cpu().numpy()) == 0.0


def test_reset_positions_from_device_memory():
    insts = _insts()
    eng = _engine(insts)
    vx = [i.vertex_x + 0.01 for i in insts]
    eng.reset(0, vertex_x=vx)
    want = eng.download_state(0)
    eng2 = _engine(insts)
    eng2.reset(0, vertex_x=torch.tensor(np.concatenate(vx), dtype=torch.float64, device=DEV))
    got = eng2.download_state(0)
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_element_moduli_in_device_memory():
    insts = _insts()
    rng = np.random.default_rng(3)
    for i in insts:
        base = np.array([i.E, i.shearing_stiffness, i.contact_stiffness])
        i.element_moduli = base[None, :] * np.exp(rng.normal(0.0, 0.3, size=(i.n_elements, 3)))
    S = 24
    eng = _engine(insts, S, per_element_moduli=True)
    eng.rollout_grad(S, 0, S, use_graph=False)
    gm_h = eng.element_moduli_grads()
    m = torch.tensor(np.concatenate([i.element_moduli for i in insts]), dtype=torch.float64, device=DEV)
    eng.set_element_moduli(m)
    eng.reset(0)
    eng.rollout_grad(S, 0, S, use_graph=False)
    gm = torch.zeros(gm_h.shape, dtype=torch.float64, device=DEV)
    eng.element_moduli_grads(out=gm)
    assert _rel(gm.cpu().numpy(), gm_h) < 1e-5


def test_autograd_function_on_device_tensors():
    from diffcp_amd.autograd import rollout_loss
    insts = _insts()
    S = 24
    eng = _engine(insts, S)
    p_host = torch.tensor(np.stack([i.params() for i in insts]), dtype=torch.float64, requires_grad=True)
    l_host = rollout_loss(eng, p_host, S)
    (l_host * torch.tensor([1.0, 3.0], dtype=torch.float64)).sum().backward()
    eng.reset(0)
    p_dev = p_host.detach().to(DEV).requires_grad_(True)
    l_dev = rollout_loss(eng, p_dev, S)
    assert l_dev.device.type == "cuda"
    (l_dev * torch.tensor([1.0, 3.0], dtype=torch.float64, device=DEV)).sum().backward()
    assert p_dev.grad.device.type == "cuda"
    assert _rel(l_dev.detach().cpu().numpy(), l_host.detach().numpy()) < 1e-5
    assert _rel(p_dev.grad.cpu().numpy(), p_host.grad.numpy()) < 1e-5
