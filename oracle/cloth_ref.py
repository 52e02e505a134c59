"""ctypes binding of the oracle's FP32 C++/OpenMP restatement (cloth_ref.cpp).

TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg,
never by the product package.  ``build()`` runs oracle/Makefile.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "bin", "libcloth_ref.so")

_i = ctypes.c_int32
_dp = ctypes.POINTER(ctypes.c_double)


class Scene(ctypes.Structure):
    _fields_ = [("n_grid", _i), ("n_vertices", _i), ("n_elements", _i), ("n_cube", _i), ("n_grasper", _i),
                ("n_grasp_points", _i), ("n_actions", _i), ("steps_per_frame", _i), ("frames_per_window", _i),
                ("reserved", _i), ("dt", ctypes.c_double), ("gravity", ctypes.c_double),
                ("vertex_p_mass", ctypes.c_double), ("quadrature_p_mass", ctypes.c_double),
                ("quadrature_p_vol", ctypes.c_double), ("cube_p_mass", ctypes.c_double),
                ("operator_p_mass", ctypes.c_double), ("damping", ctypes.c_double), ("vertex_x", _dp), ("triangles", ctypes.POINTER(_i)),
                ("cube_x", _dp), ("grasper_x", _dp), ("trajectory", _dp)]


def build() -> str:
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle/Makefile failed:\n" + r.stderr[-4000:])
    return LIB


_LIB = None


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.cref_loss_grad.restype = _i
        lib.cref_loss_grad.argtypes = [ctypes.POINTER(Scene), _dp, _dp, _i, _i, _i, _dp, _dp, _dp,
                                       ctypes.POINTER(ctypes.c_float)]
        lib.cref_max_threads.restype = _i
        _LIB = lib
    return _LIB


def max_threads() -> int:
    return int(load().cref_max_threads())


def loss_grad(inst, targets, n_substeps: int, threads: int = 0):
    """(loss, [dL/dE, dL/dnu], seconds, final vertex positions) of one window:
    the loss on the state after n_substeps - 2 substeps, FP32, dual numbers."""
    lib = load()
    keep = []

    def arr(a, dt=np.float64):
        a = np.ascontiguousarray(np.asarray(a, dtype=dt))
        keep.append(a)
        return a.ctypes.data_as(_dp if dt == np.float64 else ctypes.POINTER(_i))

    sc = Scene(n_grid=inst.n_grid, n_vertices=inst.n_vertices, n_elements=inst.n_elements, n_cube=inst.n_cube,
               n_grasper=inst.n_grasper, n_grasp_points=inst.n_grasp_points, n_actions=inst.n_actions,
               steps_per_frame=inst.steps_per_frame, frames_per_window=inst.frames_per_window, reserved=0,
               dt=inst.dt, gravity=inst.gravity, vertex_p_mass=inst.vertex_p_mass,
               quadrature_p_mass=inst.quadrature_p_mass, quadrature_p_vol=inst.quadrature_p_vol,
               cube_p_mass=inst.cube_p_mass, operator_p_mass=inst.operator_p_mass, damping=inst.damping,
               vertex_x=arr(inst.vertex_x), triangles=arr(inst.triangles, np.int32), cube_x=arr(inst.cube_x),
               grasper_x=arr(inst.grasper_x), trajectory=arr(inst.trajectory))
    params = np.array([inst.E, inst.nu, inst.shearing_stiffness, inst.contact_stiffness], np.float64)
    tg = np.ascontiguousarray(np.asarray(targets, np.float64).reshape(-1, 3))
    loss = ctypes.c_double()
    grad = np.zeros(2, np.float64)
    secs = ctypes.c_double()
    xf = np.zeros((inst.n_vertices, 3), np.float32)
    rc = lib.cref_loss_grad(ctypes.byref(sc), params.ctypes.data_as(_dp), tg.ctypes.data_as(_dp), tg.shape[0],
                            n_substeps, threads, ctypes.byref(loss), grad.ctypes.data_as(_dp), ctypes.byref(secs),
                            xf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc == -2:
        raise RuntimeError("cloth_ref: a stencil left the grid")
    if rc != 0:
        raise RuntimeError(f"cloth_ref: error {rc}")
    return loss.value, grad, secs.value, xf
