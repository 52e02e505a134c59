"""Windowed parameter identification over a batch of cloth instances.

Restates the reference's fitting loop (src/parameter_fitting_functions.py)
on the batched library, every instance being an independent candidate:

* ``compute_grad`` (parameter_fitting_functions.py:17-58): for each gradient
  window g the loss of that window's targets and its parameter gradients come
  from one ``dcp_rollout_grad`` call (forward + adjoint, the tape window of
  :25-29); the final state becomes the next window's initial state with
  v = C = 0, F = I and the rest shape re-based (:40-49); losses and gradients
  are averaged over the windows after ``nan_to_num`` (:54-57).
* ``single_parameter_fitting`` (:64-92): the adaptive-magnitude update of one
  bounded parameter.
* ``parameters_learning`` (:94-192): epochs of compute_grad + updates of E and
  nu and/or the contact and shearing stiffnesses, keeping the best parameters
  seen (by loss; the reference also replaces them at epoch 1 unconditionally).

Gradient order follows the reference's ``grad_result_list``:
(E, nu, contact_stiffness, shearing_stiffness).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .engine import Engine

# parameter columns of Engine.set_params / param_grads
P_E, P_NU, P_SHEAR, P_CONTACT = 0, 1, 2, 3

# simulation_arguments.json "constitutive_parameter_{lower,upper}_bounds"
REFERENCE_BOUNDS = {"E": (10.0, 1000.0), "nu": (0.05, 0.45),
                    "contact_stiffness": (10.0, 5000.0), "shearing_stiffness": (1.0, 1000.0)}


def magnitude_level(a: float) -> int:
    """floor(log10(|a + 1e-19|)) (parameter_fitting_functions.py:60-63)."""
    return math.floor(math.log10(abs(a + 1e-19)))


def single_parameter_fitting(parameter: float, grad: float, lower: float, upper: float, lr: float,
                             epochs: int, epoch_index: int) -> Tuple[float, float]:
    """One bounded update with a learning rate rescaled so that the step has
    the order of magnitude of the remaining distance to the bound the gradient
    points at, spread over the remaining epochs (parameter_fitting_functions.py:64-92)."""
    period = epochs - epoch_index
    if not abs(grad) > 0:
        return parameter, lr
    delta = lr * grad
    target = upper if grad < 0 else lower
    new_mag = magnitude_level((target - parameter) / period)
    if epoch_index < 1:
        lr *= 10.0 ** (new_mag - magnitude_level(delta))
        delta = lr * grad
    update = parameter - delta
    if lower < update < upper:
        parameter = update
        lr *= 10.0 ** (new_mag - magnitude_level(delta))
    else:
        if update < lower:
            new_mag = magnitude_level((upper - parameter) / period)
        if update > upper:
            new_mag = magnitude_level((lower - parameter) / period)
        lr *= 10.0 ** (new_mag - magnitude_level(delta))
        update = parameter - lr * grad
        while update > upper or update < lower:
            lr *= 0.5
            update = parameter - lr * grad
        parameter = update
    return parameter, lr


@dataclasses.dataclass
class FitResult:
    params: np.ndarray                 # [B, 6] best parameters (by loss)
    losses: List[np.ndarray]           # per epoch [B]
    grads: List[np.ndarray]            # per epoch [B, 4]
    history: List[np.ndarray]          # per epoch [B, 6] parameters used


class WindowedFit:
    """The fitting loop of the reference on an Engine of B instances.

    ``targets[g][b]`` is the target point cloud of window g for instance b.
    Every window runs ``substeps`` substeps with checkpoints every
    ``checkpoint_every``; the engine needs ``Engine.slots_for(substeps,
    checkpoint_every)`` state slots.
    """

    def __init__(self, engine: Engine, targets: Sequence[Sequence[np.ndarray]], substeps: int,
                 checkpoint_every: int = 0, use_graph: bool = True):
        self.engine = engine
        self.targets = [list(t) for t in targets]
        if any(len(t) != engine.B for t in self.targets):
            raise ValueError("targets[g] must hold one point cloud per instance")
        self.S = substeps
        self.K = checkpoint_every if checkpoint_every > 0 else substeps
        self.use_graph = use_graph

    def initialize_objects(self) -> None:
        """State 0 from the instances' creation data (simulation.py:432-445)."""
        self.engine.reset(0)

    def compute_grad(self) -> Tuple[np.ndarray, np.ndarray]:
        """Mean loss [B] and mean gradients [B, 4] (E, nu, contact, shearing) over the windows."""
        eng = self.engine
        losses, grads = [], []
        for g, tg in enumerate(self.targets):
            eng.set_targets(tg)
            loss, gr, _ = eng.rollout_grad(self.S, g, self.K, use_graph=self.use_graph)
            losses.append(loss)
            grads.append(gr[:, [P_E, P_NU, P_CONTACT, P_SHEAR]])
            # carry-over: final positions become the next window's rest state
            st = eng.download_state(eng.last_final_slot())
            eng.reset(0, vertex_x=eng.split(st["vertex_x"], "n_vertices"),
                      grasper_x=eng.split(st["grasper_x"], "n_grasper"),
                      cube_x=eng.split(st["cube_x"], "n_cube"))
        loss_mean = np.mean(np.stack(losses), axis=0)
        grad_mean = np.mean(np.nan_to_num(np.stack(grads)), axis=0)
        # the last window's raw dL/dE, before nan_to_num: what the reference's
        # NaN stop reads (garment_E.grad[None], parameter_fitting_functions.py:148-150)
        self.last_window_E_grad = grads[-1][:, 0].copy()
        return loss_mean, grad_mean

    def parameters_learning(self, params: np.ndarray, epochs: int, elasticity: bool = True,
                            contact: bool = False, bounds: Optional[dict] = None) -> FitResult:
        """Epochs of compute_grad + bounded updates, per instance independently."""
        bounds = dict(REFERENCE_BOUNDS, **(bounds or {}))
        params = np.array(params, np.float64, copy=True)
        B = self.engine.B
        lr = np.ones((B, 4))
        best = params.copy()
        best_loss = np.full(B, 1e9)
        alive = np.ones(B, bool)         # a NaN gradient stops a candidate (the reference stops the loop)
        res = FitResult(best, [], [], [])
        cols = ((P_E, "E", 0), (P_NU, "nu", 1), (P_CONTACT, "contact_stiffness", 2),
                (P_SHEAR, "shearing_stiffness", 3))
        for i in range(epochs):
            self.engine.set_params(params)
            self.initialize_objects()
            loss, grads = self.compute_grad()
            res.losses.append(loss)
            res.grads.append(grads)
            res.history.append(params.copy())
            take = (best_loss >= loss) | (i == 1)
            best[take] = params[take]
            best_loss[take] = loss[take]
            for b in range(B):
                if not (abs(self.last_window_E_grad[b]) >= 0):
                    alive[b] = False
                if not alive[b]:
                    continue
                for col, name, gi in cols:
                    learn = elasticity if name in ("E", "nu") else contact
                    if not learn:
                        continue
                    lo, hi = bounds[name]
                    params[b, col], lr[b, gi] = single_parameter_fitting(params[b, col], grads[b, gi], lo, hi,
                                                                         lr[b, gi], epochs, i)
        res.params = best
        return res
