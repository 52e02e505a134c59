"""torch.autograd.Function over a whole window of the batched substep.

Replaces the reference's tape window (``with ti.ad.Tape(loss=loss): for f in
range(max_steps-1): forward(f, g); compute_loss(f)``,
src/parameter_fitting_functions.py:25-29) for torch callers: the loss of
every instance is a differentiable function of the per-instance parameter
tensor (and, optionally, of per-element moduli).

The library runs forward and adjoint of the window in one call
(``dcp_rollout_grad``: checkpoints every ``checkpoint_every`` substeps,
recompute in the backward, hipGraph-captured segments).  Instances are
independent and the adjoint is linear in the loss seed, so
d(sum_b w_b L_b)/d theta_b = w_b dL_b/d theta_b: ``backward`` scales the
saved unit-seed gradients by ``grad_output``.  The window's initial state
must already be in state slot 0 (``Engine.reset``).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .engine import Engine


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params: torch.Tensor, element_moduli: Optional[torch.Tensor], engine: Engine,
                n_substeps: int, g: int, checkpoint_every: int, use_graph: bool):
        # the library reads and writes the tensors' memory directly (host or
        # device): no numpy copies
        if tuple(params.shape) != (engine.B, 6):
            raise ValueError(f"params must be [{engine.B}, 6] (E, nu, shearing, contact, damping, friction)")
        engine.set_params(params.detach().to(torch.float64).contiguous())
        if element_moduli is not None:
            engine.set_element_moduli(element_moduli.detach().to(torch.float64).contiguous())
        loss = torch.empty(engine.B, dtype=torch.float64, device=params.device)
        grads = torch.empty(engine.B, 6, dtype=torch.float64, device=params.device)
        engine.rollout_grad(n_substeps, g, checkpoint_every, use_graph, loss_out=loss, grads_out=grads)
        # the reference's caller sanitises with nan_to_num (parameter_fitting_functions.py:54-57)
        grads = torch.nan_to_num(grads)
        ctx.has_moduli = element_moduli is not None
        saved = [grads]
        if element_moduli is not None:
            ne = sum(i.n_elements for i in engine.instances)
            gm = torch.empty(ne, 3, dtype=torch.float64, device=element_moduli.device)
            saved.append(engine.element_moduli_grads(out=gm))
        ctx.save_for_backward(*saved)
        ctx.elem_inst = (np.repeat(np.arange(engine.B), [i.n_elements for i in engine.instances])
                         if element_moduli is not None else None)
        return loss

    @staticmethod
    def backward(ctx, grad_loss: torch.Tensor):
        saved = ctx.saved_tensors
        w = grad_loss.to(torch.float64)
        g_params = w[:, None] * saved[0]
        g_moduli = None
        if ctx.has_moduli:
            idx = torch.as_tensor(ctx.elem_inst, device=saved[1].device)
            g_moduli = w.to(saved[1].device)[idx][:, None] * saved[1]
        return g_params, g_moduli, None, None, None, None, None


def rollout_loss(engine: Engine, params: torch.Tensor, n_substeps: int, *, g: int = 0,
                 checkpoint_every: int = 0, use_graph: bool = True,
                 element_moduli: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Per-instance point-cloud loss of one window, differentiable in
    ``params`` ([B, 6] float64: E, nu, shearing, contact, damping, friction)
    and, with per-element moduli enabled, in ``element_moduli`` ([sum NE, 3])."""
    return RolloutLoss.apply(params, element_moduli, engine, n_substeps, g, checkpoint_every, use_graph)
