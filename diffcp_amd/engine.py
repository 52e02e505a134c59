"""Python handle on one device context of the HIP library (a batch of cloth instances)."""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _abi
from .scene import Instance, N_PARAMS

PRECISION = {"mixed": 0, "fp32": 1}
SVD_GRAD = {"reference": 0, "stable": 1}


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_abi.c_dbl_p)


def _fp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_abi.c_flt_p)


def _is_tensor(x) -> bool:
    return type(x).__module__.split(".")[0] == "torch"


def _tensor_arg(t, dtype: str, shape, name: str):
    """A torch tensor (host or device memory of the engine's GPU) as a C-ABI
    array argument: dtype, contiguity and shape are checked here, the memory
    kind by the library.  A device tensor's pending work on torch's current
    stream is waited for first: the library works on its own stream, and it
    has finished with the tensor when the call returns."""
    import torch
    want = {"float32": torch.float32, "float64": torch.float64, "int32": torch.int32}[dtype]
    if t.dtype != want:
        raise TypeError(f"{name}: {dtype} tensor expected, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(shape)} expected, got {tuple(t.shape)}")
    if t.device.type not in ("cpu", "cuda"):
        raise ValueError(f"{name}: tensor on {t.device} (cpu or the engine's GPU expected)")
    if t.device.type == "cuda":
        torch.cuda.current_stream(t.device).synchronize()
    ptr_type = {"float32": _abi.c_flt_p, "float64": _abi.c_dbl_p, "int32": _abi.c_int_p}[dtype]
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ptr_type)


def _snap_option(v) -> int:
    """dcp_options.grid_snapshots from True (on), False (off) or a block bound per substep."""
    if isinstance(v, bool):
        return 0 if v else -1
    if int(v) <= 0:
        raise ValueError("grid_snapshots: True, False or a positive block bound")
    return int(v)


class Engine:
    """Owns a ``dcp_ctx``: all device buffers of B instances.

    Slots: ``max_states`` state slots (0 is where :meth:`reset` writes) and 4
    adjoint slots.
    """

    def __init__(self, instances: Sequence[Instance], *, max_states: int = 2,
                 precision: str = "mixed", svd_grad: str = "reference",
                 per_element_moduli: bool = False, patch_size: int = 0, scatter_wgs: int = 0,
                 grid_wgs: int = 0, bin_cells: int = 0, grid_snapshots: bool = True,
                 return_mapping: bool = False, grid_pool_blocks: int = 0, deterministic: bool = False,
                 stream: Optional[int] = None):
        self.lib = _abi.load()
        self.instances = list(instances)
        self.B = len(self.instances)
        self._keep = []
        descs = (_abi.InstanceDesc * self.B)()
        for b, inst in enumerate(self.instances):
            vx = np.ascontiguousarray(inst.vertex_x, dtype=np.float64)
            tri = np.ascontiguousarray(inst.triangles, dtype=np.int32)
            cx = np.ascontiguousarray(inst.cube_x, dtype=np.float64).reshape(-1, 3)
            ox = np.ascontiguousarray(inst.grasper_x, dtype=np.float64).reshape(-1, 3)
            tr = np.ascontiguousarray(inst.trajectory, dtype=np.float64)
            self._keep += [vx, tri, cx, ox, tr]
            d = descs[b]
            d.n_grid = inst.n_grid
            d.n_vertices = inst.n_vertices
            d.n_elements = inst.n_elements
            d.n_cube = inst.n_cube
            d.n_grasper = inst.n_grasper
            d.n_grasp_points = inst.n_grasp_points
            d.n_actions = inst.n_actions
            d.steps_per_frame = inst.steps_per_frame
            d.frames_per_window = inst.frames_per_window
            d.dt = inst.dt
            d.gravity = inst.gravity
            d.vertex_p_mass = inst.vertex_p_mass
            d.quadrature_p_mass = inst.quadrature_p_mass
            d.quadrature_p_vol = inst.quadrature_p_vol
            d.cube_p_mass = inst.cube_p_mass
            d.operator_p_mass = inst.operator_p_mass
            d.vertex_x = _dp(vx)
            d.triangles = tri.ctypes.data_as(_abi.c_int_p)
            d.cube_x = _dp(cx)
            d.grasper_x = _dp(ox)
            d.trajectory = _dp(tr)
        opt = _abi.Options(max_states=max_states, precision=PRECISION[precision],
                           svd_grad=SVD_GRAD[svd_grad], per_element_moduli=int(per_element_moduli),
                           patch_size=patch_size, scatter_wgs=scatter_wgs, grid_wgs=grid_wgs,
                           bin_cells=bin_cells, grid_snapshots=_snap_option(grid_snapshots),
                           return_mapping=int(return_mapping), grid_pool_blocks=int(grid_pool_blocks),
                           deterministic=int(deterministic))
        ctx = ctypes.c_void_p()
        _abi.check(self.lib.dcp_create(descs, self.B, ctypes.byref(opt), stream, ctypes.byref(ctx)))
        self.ctx = ctx
        self.max_states = max_states
        self.per_element_moduli = per_element_moduli
        sz = (ctypes.c_int64 * 5)()
        _abi.check(self.lib.dcp_sizes(self.ctx, sz))
        self.n_points, self.n_elements, self.slot_floats = int(sz[1]), int(sz[2]), int(sz[3])
        self.set_params(np.stack([i.params() for i in self.instances]))
        if per_element_moduli:
            self.set_element_moduli(np.concatenate([i.element_moduli for i in self.instances]))
        if all(i.targets is not None for i in self.instances):
            self.set_targets([i.targets for i in self.instances])

    # ------------------------------------------------------------------ setup
    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.dcp_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params) -> None:
        """[B, 6] float64 (E, nu, shearing, contact, damping, friction): numpy,
        or a torch tensor in host or device memory."""
        if _is_tensor(params):
            _abi.check(self.lib.dcp_set_params(self.ctx, _tensor_arg(params, "float64", (self.B, N_PARAMS), "params")))
            self.params = params.detach().cpu().numpy().copy()
            return
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(self.B, N_PARAMS))
        _abi.check(self.lib.dcp_set_params(self.ctx, _dp(p)))
        self.params = p.copy()

    def set_element_moduli(self, moduli) -> None:
        if _is_tensor(moduli):
            n = sum(i.n_elements for i in self.instances)
            _abi.check(self.lib.dcp_set_element_moduli(self.ctx, _tensor_arg(moduli, "float64", (n, 3), "moduli")))
            return
        m = np.ascontiguousarray(np.asarray(moduli, dtype=np.float64).reshape(-1, 3))
        _abi.check(self.lib.dcp_set_element_moduli(self.ctx, _dp(m)))

    def set_targets(self, targets, counts: Optional[Sequence[int]] = None) -> None:
        """One point cloud per instance (a list of [n_b, 3] arrays), or all of
        them concatenated in one [sum n_b, 3] float64 torch tensor (host or
        device memory) with ``counts``."""
        if _is_tensor(targets):
            if counts is None or len(counts) != self.B:
                raise ValueError("a concatenated target tensor needs one count per instance")
            c = np.ascontiguousarray(np.asarray(counts, dtype=np.int32))
            _abi.check(self.lib.dcp_set_targets(self.ctx, _tensor_arg(targets, "float64", (int(c.sum()), 3), "targets"),
                                                c.ctypes.data_as(_abi.c_int_p)))
            return
        counts = np.array([len(t) for t in targets], dtype=np.int32)
        t = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3)
                                                 for x in targets]))
        _abi.check(self.lib.dcp_set_targets(self.ctx, _dp(t), counts.ctypes.data_as(_abi.c_int_p)))

    def reset(self, slot: int = 0, vertex_x=None, grasper_x=None, cube_x=None) -> None:
        """Rest state in ``slot`` (initialize_ti_field) from the creation
        positions, or from given ones: one array per instance, or all
        instances' concatenated in one [n, 3] float64 torch tensor (host or
        device memory)."""
        def arg(arrs, count):
            if arrs is None:
                return None
            if _is_tensor(arrs):
                n = sum(getattr(i, count) for i in self.instances)
                return _tensor_arg(arrs, "float64", (n, 3), count)
            a = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).reshape(-1, 3) for x in arrs]))
            keep.append(a)
            return _dp(a)
        keep = []
        _abi.check(self.lib.dcp_reset(self.ctx, slot, arg(vertex_x, "n_vertices"), arg(grasper_x, "n_grasper"),
                                      arg(cube_x, "n_cube")))

    def device_bytes(self) -> int:
        """Device memory held by the context (all buffers, snapshots included)."""
        out = ctypes.c_int64()
        _abi.check(self.lib.dcp_device_bytes(self.ctx, ctypes.byref(out)))
        return int(out.value)

    def pool_blocks(self) -> tuple:
        """(slots handed out since the last reset, capacity) of the grid block
        pool; (0, 0) for dense grids."""
        used, cap = ctypes.c_int(), ctypes.c_int()
        _abi.check(self.lib.dcp_pool_blocks(self.ctx, ctypes.byref(used), ctypes.byref(cap)))
        return int(used.value), int(cap.value)

    def rebase_elements(self, slot: int) -> None:
        """Element rest shapes from the vertex positions in ``slot`` (reset_quadrature_particles)."""
        _abi.check(self.lib.dcp_rebase_elements(self.ctx, slot))

    # ------------------------------------------------------------------ steps
    def substep(self, s: int, g: int, slot_in: int, slot_out: int) -> None:
        _abi.check(self.lib.dcp_substep(self.ctx, s, g, slot_in, slot_out))

    def substep_adjoint(self, s: int, g: int, slot_in: int, gslot_in: int, gslot_out: int) -> None:
        _abi.check(self.lib.dcp_substep_adjoint(self.ctx, s, g, slot_in, gslot_in, gslot_out))

    def zero_adjoint(self, gslot: int) -> None:
        _abi.check(self.lib.dcp_zero_adjoint(self.ctx, gslot))

    def zero_param_grads(self) -> None:
        _abi.check(self.lib.dcp_zero_param_grads(self.ctx))

    def loss(self, slot: int) -> np.ndarray:
        out = np.zeros(self.B, np.float64)
        _abi.check(self.lib.dcp_loss(self.ctx, slot, _dp(out)))
        return out

    def loss_adjoint(self, slot: int, gslot: int, seed=None) -> None:
        sd = None if seed is None else np.ascontiguousarray(np.broadcast_to(np.asarray(seed, np.float64), (self.B,)))
        _abi.check(self.lib.dcp_loss_adjoint(self.ctx, slot, gslot, _dp(sd)))

    def param_grads(self, out=None):
        """[B, 6] float64; ``out``: a torch tensor to fill (host or device memory)."""
        if out is not None:
            _abi.check(self.lib.dcp_param_grads(self.ctx, _tensor_arg(out, "float64", (self.B, N_PARAMS), "out")))
            return out
        out = np.zeros((self.B, N_PARAMS), np.float64)
        _abi.check(self.lib.dcp_param_grads(self.ctx, _dp(out)))
        return out

    def element_moduli_grads(self, out=None):
        """[sum NE, 3] float64; ``out``: a torch tensor to fill (host or device memory)."""
        n = sum(i.n_elements for i in self.instances)
        if out is not None:
            _abi.check(self.lib.dcp_element_moduli_grads(self.ctx, _tensor_arg(out, "float64", (n, 3), "out")))
            return out
        out = np.zeros((n, 3), np.float64)
        _abi.check(self.lib.dcp_element_moduli_grads(self.ctx, _dp(out)))
        return out

    def set_trajectory(self, trajectories: List[np.ndarray]) -> None:
        """Replace every instance's grasp trajectory ([n_actions, n_grasp_points, 3] each)."""
        t = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).reshape(-1) for x in trajectories]))
        if t.size != sum(i.trajectory.size for i in self.instances):
            raise ValueError("trajectory sizes differ from the ones the engine was created with")
        _abi.check(self.lib.dcp_set_trajectory(self.ctx, _dp(t)))

    def trajectory_grads(self) -> List[np.ndarray]:
        n = sum(i.trajectory.size for i in self.instances)
        out = np.zeros(n, np.float64)
        _abi.check(self.lib.dcp_trajectory_grads(self.ctx, _dp(out)))
        res, o = [], 0
        for i in self.instances:
            res.append(out[o:o + i.trajectory.size].reshape(i.trajectory.shape))
            o += i.trajectory.size
        return res

    def status(self, clear: bool = False) -> int:
        v = ctypes.c_uint32()
        _abi.check(self.lib.dcp_status(self.ctx, ctypes.byref(v), int(clear)))
        return int(v.value)

    # ------------------------------------------------------------------ I/O
    STATE_NAMES = ("vertex_x", "vertex_v", "vertex_C", "grasper_x", "grasper_v", "grasper_C",
                   "cube_x", "cube_v", "cube_C", "element_C", "element_n")

    def _alloc_state(self):
        nv = sum(i.n_vertices for i in self.instances)
        ne = sum(i.n_elements for i in self.instances)
        no = sum(i.n_grasper for i in self.instances)
        nc = sum(i.n_cube for i in self.instances)
        shapes = {"vertex": nv, "grasper": no, "cube": nc}
        out = {}
        for k in self.STATE_NAMES:
            cls, f = k.split("_")
            if cls == "element":
                out[k] = np.zeros((ne, 3, 3) if f == "C" else (ne, 3), np.float32)
            else:
                out[k] = np.zeros((shapes[cls], 3, 3) if f == "C" else (shapes[cls], 3), np.float32)
        return out

    def state_shapes(self) -> Dict[str, tuple]:
        return {k: v.shape for k, v in self._alloc_state().items()}

    def _tensor_arrays(self, arrs):
        shapes = self.state_shapes()
        return [None if arrs.get(k) is None else _tensor_arg(arrs[k], "float32", shapes[k], k)
                for k in self.STATE_NAMES]

    def download_state(self, slot: int, out=None) -> Dict[str, np.ndarray]:
        """The state in ``slot`` in natural particle order.  ``out``: a dict of
        float32 torch tensors (shapes ``state_shapes()``, any subset, all in
        host or all in device memory) filled in place."""
        if out is not None:
            _abi.check(self.lib.dcp_download_state(self.ctx, slot, *self._tensor_arrays(out)))
            return out
        st = self._alloc_state()
        _abi.check(self.lib.dcp_download_state(self.ctx, slot, *[_fp(st[k]) for k in self.STATE_NAMES]))
        return st

    def download_adjoint(self, gslot: int) -> Dict[str, np.ndarray]:
        st = self._alloc_state()
        _abi.check(self.lib.dcp_download_adjoint(self.ctx, gslot, *[_fp(st[k]) for k in self.STATE_NAMES]))
        return st

    def _conv(self, arrs):
        return {k: np.ascontiguousarray(np.asarray(arrs[k], np.float32))
                for k in self.STATE_NAMES if arrs.get(k) is not None}

    def upload_state(self, slot: int, **arrs) -> None:
        """Arrays named as ``STATE_NAMES`` (missing ones are zero): numpy, or
        float32 torch tensors all in host or all in device memory."""
        if any(_is_tensor(v) for v in arrs.values() if v is not None):
            _abi.check(self.lib.dcp_upload_state(self.ctx, slot, *self._tensor_arrays(arrs)))
            return
        conv = self._conv(arrs)
        _abi.check(self.lib.dcp_upload_state(self.ctx, slot, *[_fp(conv.get(k)) for k in self.STATE_NAMES]))

    def upload_adjoint(self, gslot: int, **arrs) -> None:
        conv = self._conv(arrs)
        _abi.check(self.lib.dcp_upload_adjoint(self.ctx, gslot, *[_fp(conv.get(k)) for k in self.STATE_NAMES]))

    def split(self, arr: np.ndarray, count_attr: str) -> List[np.ndarray]:
        out, o = [], 0
        for i in self.instances:
            n = getattr(i, count_attr)
            out.append(arr[o:o + n])
            o += n
        return out

    def binning_keys(self, slot: int) -> np.ndarray:
        n = sum(i.n_vertices + i.n_grasper + i.n_cube + i.n_elements for i in self.instances)
        out = np.zeros((n, 3), np.int32)
        _abi.check(self.lib.dcp_binning_keys(self.ctx, slot, out.ctypes.data_as(_abi.c_int_p)))
        return out

    PATCH_ORDER_COLUMNS = ("patch", "lane", "instance", "class", "natural", "b0", "b1", "b2", "position")

    def patch_order(self, slot: int) -> np.ndarray:
        """The scatters' per-patch binning of ``slot``: one row per particle,
        columns ``PATCH_ORDER_COLUMNS`` (class 0..3 = vertex, grasper, cube, element)."""
        n = self.n_points + self.n_elements
        out = np.zeros((n, 9), np.int32)
        _abi.check(self.lib.dcp_patch_order(self.ctx, slot, out.ctypes.data_as(_abi.c_int_p)))
        return out

    # ------------------------------------------------------------------ rollouts
    def rollout_grad(self, n_substeps: int, g: int = 0, checkpoint_every: int = 0,
                     use_graph: bool = True, loss_out=None, grads_out=None):
        """Forward + adjoint of a whole window from the reset state in slot 0.

        Returns (loss[B], grads[B, 6], device milliseconds).  ``loss_out`` /
        ``grads_out``: float64 torch tensors ([B], [B, 6], host or device
        memory) filled in place and returned instead of numpy arrays.
        """
        if loss_out is not None or grads_out is not None:
            if loss_out is None or grads_out is None:
                raise ValueError("loss_out and grads_out go together")
            lp = _tensor_arg(loss_out, "float64", (self.B,), "loss_out")
            gp = _tensor_arg(grads_out, "float64", (self.B, N_PARAMS), "grads_out")
            _abi.check(self.lib.dcp_rollout_grad(self.ctx, n_substeps, g, checkpoint_every, int(use_graph), lp, gp))
            ms = ctypes.c_double()
            _abi.check(self.lib.dcp_last_rollout_ms(self.ctx, ctypes.byref(ms)))
            return loss_out, grads_out, ms.value
        loss = np.zeros(self.B, np.float64)
        grads = np.zeros((self.B, N_PARAMS), np.float64)
        _abi.check(self.lib.dcp_rollout_grad(self.ctx, n_substeps, g, checkpoint_every,
                                             int(use_graph), _dp(loss), _dp(grads)))
        ms = ctypes.c_double()
        _abi.check(self.lib.dcp_last_rollout_ms(self.ctx, ctypes.byref(ms)))
        return loss, grads, ms.value

    def last_final_slot(self) -> int:
        """State slot of the final state of the last rollout."""
        s = ctypes.c_int32()
        _abi.check(self.lib.dcp_last_rollout_final_slot(self.ctx, ctypes.byref(s)))
        return s.value

    def synchronize(self) -> None:
        """Wait for all work queued on the engine's stream."""
        _abi.check(self.lib.dcp_synchronize(self.ctx))

    KERNELS = {"prep": 1, "p2g": 2, "gridop": 3, "g2p": 4, "g2p_adj": 5, "gridop_adj": 6,
               "p2g_adj": 7, "prep_adj": 8, "vertex_final": 9}

    def set_kernel_timing(self, kernel: Optional[str], max_launches: int = 4096) -> None:
        """Time every launch of one kernel kind with HIP events (None: off)."""
        kid = 0 if kernel is None else self.KERNELS[kernel]
        _abi.check(self.lib.dcp_set_kernel_timing(self.ctx, kid, max_launches if kid else 0))

    def kernel_timing(self):
        """(total ms, launches) recorded by set_kernel_timing."""
        ms = ctypes.c_double()
        n = ctypes.c_int32()
        _abi.check(self.lib.dcp_kernel_timing(self.ctx, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def rollout(self, n_substeps: int, g: int = 0, use_graph: bool = False):
        slot = ctypes.c_int32()
        _abi.check(self.lib.dcp_rollout(self.ctx, n_substeps, g, int(use_graph), ctypes.byref(slot)))
        ms = ctypes.c_double()
        _abi.check(self.lib.dcp_last_rollout_ms(self.ctx, ctypes.byref(ms)))
        return slot.value, ms.value

    @staticmethod
    def slots_for(n_substeps: int, checkpoint_every: int) -> int:
        """State slots dcp_rollout_grad needs for a plan."""
        K = checkpoint_every if 0 < checkpoint_every <= n_substeps else n_substeps
        nseg = (n_substeps + K - 1) // K
        return nseg + 1 + K + 1


def polar_adjoint(Ft: np.ndarray, gR: np.ndarray, precision: str = "mixed", svd_grad: str = "reference",
                  no_alloc: bool = False):
    """The library's 2x2 polar rotation R and adjoint gF = d<gR, R>/dFt for
    arrays of shape [n, 2, 2] (FP32; ``precision`` "mixed" computes in FP64).
    ``no_alloc``: the entry point that allocates no device memory (at most
    16384 matrices per call)."""
    lib = _abi.load()
    F = np.ascontiguousarray(np.asarray(Ft, np.float32).reshape(-1, 2, 2))
    g = np.ascontiguousarray(np.asarray(gR, np.float32).reshape(-1, 2, 2))
    if F.shape != g.shape:
        raise ValueError("Ft and gR must have the same shape")
    gF = np.zeros_like(F)
    R = np.zeros_like(F)
    fn = lib.dcp_polar_adjoint_noalloc if no_alloc else lib.dcp_polar_adjoint
    _abi.check(fn(_fp(F), _fp(g), F.shape[0], PRECISION[precision], SVD_GRAD[svd_grad], _fp(gF), _fp(R)))
    return R, gF
