"""Bridge between the reference's state fields and the HIP library's state.

TEST INFRASTRUCTURE ONLY (the checker; see sim_torch.py).

The reference stores, per element, x, v, C, F and d at every substep
(taichi_variable_definition.py:16-34).  The HIP library stores only C and the
third column ``n`` of the previous substep's d, and rebuilds F and d from the
vertex positions, exactly as update_particle_state does
(simulation.py:392-430):

  F[f] = [x2-x1 | x3-x1 | n + (dt C + 0.5 dt^2 C C) n] @ Dinv,  d[f] = F[f] @ inverse(Dinv)

with x = vertex_x[f], C = quadrature_C[f], n = d[f-1][:, 2]; at the initial
state F = I and d = the rest shape (simulation.py:468-476).  ``expand`` maps the
library's state to the reference's and ``reduce`` maps back, both
differentiable, so that autograd through expand -> substep -> reduce is the
reference tape's vector-Jacobian product in the library's coordinates.
"""
from __future__ import annotations

from typing import Dict

import numpy as np
import torch

from . import sim_torch as O

POINT_KEYS = ("vx", "vv", "vC", "ox", "ov", "oC", "cx", "cv", "cC")
RED_KEYS = POINT_KEYS + ("qC", "qn")

# library array name -> reduced key
LIB_NAMES = {"vertex_x": "vx", "vertex_v": "vv", "vertex_C": "vC", "grasper_x": "ox",
             "grasper_v": "ov", "grasper_C": "oC", "cube_x": "cx", "cube_v": "cv", "cube_C": "cC",
             "element_C": "qC", "element_n": "qn"}


def expand(red: Dict[str, torch.Tensor], static, K: O.Consts, fresh: bool, params=None) -> Dict[str, torch.Tensor]:
    """Reference state from the library's; with ``params`` (live return
    mapping) F is the return-mapped F = Q return_map(R) of the rebuilt one, as
    update_particle_state would have stored it."""
    tri, Dinv = static["tri"], static["Dinv"]
    vx, vv = red["vx"], red["vv"]
    x1, x2, x3 = vx[tri[:, 0]], vx[tri[:, 1]], vx[tri[:, 2]]
    v1, v2, v3 = vv[tri[:, 0]], vv[tri[:, 1]], vv[tri[:, 2]]
    n = red["qn"]
    C = red["qC"]
    if fresh:
        qx = (x1 + x2 + x3) / 3
        qv = torch.zeros_like(qx)
        d = O.cols(x2 - x1, x3 - x1, n)
        F = torch.eye(3, dtype=vx.dtype).expand_as(d).clone()
    else:
        qx = (1 / 3) * (x1 + x2 + x3)
        qv = (1 / 3) * (v1 + v2 + v3)
        M = K.dt * C + 0.5 * K.dt * K.dt * C @ C
        col3 = n + (M @ n[..., None])[..., 0]
        dn = O.cols(x2 - x1, x3 - x1, col3)
        F = dn @ Dinv
        if params is not None:
            if "elem" in params:
                ks, kc = params["elem"][:, 1], params["elem"][:, 2]
            else:
                ks, kc = params["ks"], params["kc"]
            Q, R = O.qr_gram_schmidt(F)
            F = Q @ O.return_map(R, ks, kc, params["friction"])
        d = F @ O.inv3(Dinv)
    st = {k: red[k] for k in POINT_KEYS}
    st.update(qx=qx, qv=qv, qC=C, qF=F, qd=d)
    return st


def reduce(ref_next: Dict[str, torch.Tensor], ref_cur: Dict[str, torch.Tensor]) -> Dict[str, torch.Tensor]:
    out = {k: ref_next[k] for k in POINT_KEYS}
    out["qC"] = ref_next["qC"]
    out["qn"] = ref_cur["qd"][:, :, 2]
    return out


def initial_reduced(st0: Dict[str, torch.Tensor]) -> Dict[str, torch.Tensor]:
    out = {k: st0[k] for k in POINT_KEYS}
    out["qC"] = st0["qC"]
    out["qn"] = st0["qd"][:, :, 2]
    return out


def reduced_substep(red, static, params, K, f, g=0, svd_grad="reference", return_mapping=False):
    cur = expand(red, static, K, fresh=(f == 0), params=params if return_mapping else None)
    nxt = O.substep(cur, static, params, K, f, g, svd_grad, return_mapping=return_mapping)
    return reduce(nxt, cur)


def from_library(arrs: Dict[str, np.ndarray], dtype=torch.float64) -> Dict[str, torch.Tensor]:
    """Library download (natural order, one instance) -> reduced torch state."""
    return {LIB_NAMES[k]: torch.as_tensor(np.asarray(v), dtype=dtype) for k, v in arrs.items() if k in LIB_NAMES}


def to_library(red: Dict[str, torch.Tensor]) -> Dict[str, np.ndarray]:
    inv = {v: k for k, v in LIB_NAMES.items()}
    return {inv[k]: red[k].detach().cpu().numpy().astype(np.float32) for k in RED_KEYS}
