#!/usr/bin/env python3
"""Throughput of the batched differentiable cloth substep on one MI355X.

One step = one forward + adjoint pass over a whole window of the batched
parameter-identification workload: 64 cloth instances (sides 32..96
vertices, own 128^3 grids, seeded E/nu/damping/friction and grasp
trajectories), 300 substeps, point-cloud loss on 1024 targets per instance,
gradients of the loss w.r.t. every instance's parameters.  The substep loop
runs as captured hipGraphs.  Counted particles are cloth vertices plus
element (quadrature) particles; grasper and cube particles are processed but
not counted.

Also reported on the same JSON line:
  * the plain path (one 64x64 cloth on a 128^3 grid, 400 substeps, checkpoints
    every 50, one launch per kernel, no graph) and the batched/plain ratio;
  * the GPU gradient w.r.t. E and nu vs central finite differences of the
    FP64 oracle on the 32x32 cloth (committed fixture, h = 1e-4 |theta|);
  * ``roofline``: the G2P adjoint (the kernel with the largest share of the
    window): its algorithmic bytes per launch over its average launch
    duration, measured with HIP events recorded on the library's stream
    around every G2P-adjoint launch of the timed region, against the 8 TB/s
    HBM peak; ``traffic`` comes from the committed rocprofv3 counter summary
    when present;
  * ``large``: the large single cloth (256x256 vertices, 256^3 grid, 1000
    substeps, sqrt(S) checkpoints, per-triangle moduli gradients): window
    time, throughput, G2P-adjoint HBM rate and the device memory held;
  * ``deterministic``: the batched window again with fixed-point scatter
    sums (bitwise reproducible windows), 2 timed steps;
  * ``cpu_baseline``: the FP32 C++/OpenMP restatement in ``oracle/`` timed
    on the host on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GT_PARAMS = (8e3, 0.25)
N_TARGETS = 1024


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def set_ground_truth_targets(eng, w, n_loss_state: int) -> None:
    """Targets = seeded samples of the vertex positions at the loss state of a
    forward run with E = 8e3, nu = 0.25 (the other parameters unchanged)."""
    from diffcp_amd import scene
    base = np.stack([i.params() for i in w.instances])
    gt = base.copy()
    gt[:, 0], gt[:, 1] = GT_PARAMS
    eng.set_params(gt)
    eng.reset(0)
    slot, _ = eng.rollout(n_loss_state)
    vx = eng.split(eng.download_state(slot)["vertex_x"], "n_vertices")
    eng.set_targets([scene.sample_targets(v, N_TARGETS, w.seed * 100 + b) for b, v in enumerate(vx)])
    eng.set_params(base)
    eng.reset(0)


def distinct_active_cells(eng, slot: int) -> int:
    """Distinct grid cells the quadratic stencils of all particles touch."""
    b = eng.binning_keys(slot).astype(np.int64)     # [n][3] base cells, instances concatenated
    counts = [i.n_vertices + i.n_grasper + i.n_cube + i.n_elements for i in eng.instances]
    inst = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    cells = []
    for di in range(3):
        for dj in range(3):
            for dk in range(3):
                c = ((inst * 1024 + b[:, 0] + di) * 1024 + b[:, 1] + dj) * 1024 + b[:, 2] + dk
                cells.append(c)
    return int(np.unique(np.concatenate(cells)).size)


ROOFLINE_KERNEL = "g2p_adj"     # the kernel with the largest share of the window


def g2p_adj_algorithmic_bytes(w, active_cells: int) -> float:
    """Compulsory HBM bytes of one G2P-adjoint launch: per vertex, grasper and
    cube particle its position and the incoming dL/d(x, v, C) read and dL/dx
    written (84 B); per element its position and dL/dC read and dL/dx written
    (60 B); per distinct active cell grid_v_out read and dL/d(m v) written
    (12 + 12 B)."""
    npt = sum(i.n_vertices + i.n_grasper + i.n_cube for i in w.instances)
    ne = sum(i.n_elements for i in w.instances)
    return 84.0 * npt + 60.0 * ne + 24.0 * active_cells


def load_traffic():
    """HBM bytes per G2P-adjoint launch from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "k_g2p_adj_c3_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("bytes_per_dispatch")


def bench_batched(steps: int, warmup: int, use_graph: bool = True, deterministic: bool = False, w=None):
    from diffcp_amd import scene
    from diffcp_amd.engine import Engine

    w = w or scene.config_c3()
    S, K = w.substeps, w.checkpoint_every
    eng = Engine(w.instances, max_states=Engine.slots_for(S, K), deterministic=deterministic)
    set_ground_truth_targets(eng, w, S - 2)
    # active cells at the initial state and at the loss state (mean)
    slot, _ = eng.rollout(S - 2)
    cells = 0.5 * (distinct_active_cells(eng, 0) + distinct_active_cells(eng, slot))
    eng.reset(0)
    # per-launch HIP events around the G2P adjoint, captured into the rollout graphs.  torch
    # is kept out of this process until the GPU legs are done: with torch's
    # HIP runtime initialised first, capturing an event-record node fails
    # (hipErrorInvalidValue), while the same capture succeeds without it.
    eng.set_kernel_timing(ROOFLINE_KERNEL, max_launches=4 * S + 16)
    for _ in range(warmup):
        eng.rollout_grad(S, 0, K, use_graph=use_graph)
    eng.status(clear=True)
    eng.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(steps):
        loss, grads, ms = eng.rollout_grad(S, 0, K, use_graph=use_graph)
        dev_ms += ms
    eng.synchronize()
    wall = time.perf_counter() - t0
    k_ms, k_n = eng.kernel_timing()
    status = eng.status()
    finite = bool(np.isfinite(loss).all() and np.isfinite(grads).all())
    eng.close()
    return dict(w=w, S=S, K=K, wall=wall, dev_ms=dev_ms, cells=cells, k_ms=k_ms, k_n=k_n,
                status=status, finite=finite)


def bench_plain(reps: int = 2):
    from diffcp_amd import scene
    from diffcp_amd.engine import Engine

    w = scene.config_c2()
    S, K = w.substeps, w.checkpoint_every
    eng = Engine(w.instances, max_states=Engine.slots_for(S, K))
    set_ground_truth_targets(eng, w, S - 2)
    eng.rollout_grad(S, 0, K, use_graph=False)
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, grads, _ = eng.rollout_grad(S, 0, K, use_graph=False)
    wall = (time.perf_counter() - t0) / reps
    rate = w.particles * S / wall
    eng.close()
    return dict(rate=rate, ms=wall * 1e3, particles=w.particles, S=S)


def device_free_bytes():
    """Free device memory reported by the HIP runtime (hipMemGetInfo), or None."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return None
    return int(free.value)


def bench_large(reps: int = 2):
    """The large single cloth: 256x256 vertices (130,050 triangles) on a 256^3
    grid, 1000 substeps fwd+adj with sqrt(S) = 32-substep checkpoints, gradient
    w.r.t. per-triangle (E, shearing, contact) moduli (seeded lognormal).
    Targets: the loss-state vertices of a run with every E scaled by 1.6.
    Reports the window time, the G2P-adjoint launch average and HBM rate, and
    the device memory the context holds (state slots, grids, snapshots) and
    the drop in free device memory from before the context was built to the
    end of the last window; the grid is a pool of active 4^3 blocks."""
    from diffcp_amd import scene
    from diffcp_amd.engine import Engine

    w = scene.config_c5()
    S, K = w.substeps, w.checkpoint_every
    inst = w.instances[0]
    free0 = device_free_bytes()
    eng = Engine(w.instances, max_states=Engine.slots_for(S, K), per_element_moduli=True,
                 grid_pool_blocks=w.grid_pool_blocks)
    gt = inst.element_moduli.copy()
    gt[:, 0] *= 1.6
    eng.set_element_moduli(gt)
    eng.reset(0)
    slot, _ = eng.rollout(S - 2)
    vx = eng.download_state(slot)["vertex_x"]
    eng.set_targets([scene.sample_targets(vx, N_TARGETS, w.seed * 100)])
    cells = 0.5 * (distinct_active_cells(eng, 0) + distinct_active_cells(eng, slot))
    eng.set_element_moduli(inst.element_moduli)
    eng.reset(0)
    eng.set_kernel_timing(ROOFLINE_KERNEL, max_launches=4 * S + 16)
    eng.rollout_grad(S, 0, K, use_graph=True)
    eng.synchronize()
    # the status bits of the timed windows only
    eng.status(clear=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, grads, _ = eng.rollout_grad(S, 0, K, use_graph=True)
    eng.synchronize()
    wall = (time.perf_counter() - t0) / reps
    k_ms, k_n = eng.kernel_timing()
    ge = eng.element_moduli_grads()
    dev = eng.device_bytes()
    used_dev = free0 - device_free_bytes() if free0 is not None else None
    pool_used, pool_cap = eng.pool_blocks()
    status = eng.status()
    eng.close()
    avg_ms = k_ms / max(k_n, 1)
    alg = g2p_adj_algorithmic_bytes(w, int(cells))
    return {"workload": "256x256-vertex cloth (130,050 triangles), 256^3 grid (pool of active 4^3 blocks), "
                        "1000 substeps fwd+adj, 32-substep checkpoints, per-triangle moduli gradients",
            "particles": w.particles, "substeps": S, "checkpoint_every": K, "ms_per_window": wall * 1e3,
            "value": w.particles * S / wall, "unit": "particle-substeps/s",
            "device_bytes": dev, "device_bytes_measured": used_dev,
            "grid_pool_blocks": {"used": pool_used, "capacity": pool_cap,
                                 "dense_equivalent": (inst.n_grid // 4) ** 3},
            "g2p_adj_avg_launch_us": avg_ms * 1e3,
            "g2p_adj_achieved_gbs": alg / (avg_ms * 1e-3) / 1e9 if k_n else None,
            "finite": bool(np.isfinite(loss).all() and np.isfinite(ge).all()), "status": status}


def grad_vs_fd():
    """Gradient of the 32x32 cloth's loss (64^3 grid, 200 substeps) w.r.t. E
    and nu on the GPU vs central finite differences of the FP64 oracle
    (h = 1e-4 relative; committed fixture tests/golden/c1_grad_fd.npz, same
    targets).  At h = 1e-3 the FD's own truncation error on this loss is
    ~2e-3 (E) and ~8e-3 (nu) — the fixture holds both."""
    from diffcp_amd import scene
    from diffcp_amd.engine import Engine

    path = os.path.join(ROOT, "tests", "golden", "c1_grad_fd.npz")
    if not os.path.exists(path):
        return None
    ref = np.load(path)
    w = scene.config_c1()
    S = int(ref["substeps"])
    out = {}
    for mode in ("reference", "stable"):
        eng = Engine(w.instances, max_states=Engine.slots_for(S, S), svd_grad=mode)
        eng.set_targets([ref["targets"]])
        eng.reset(0)
        _, grads, _ = eng.rollout_grad(S, 0, S, use_graph=False)
        eng.close()
        fd = ref["fd_0.0001"]
        out[mode] = {"E": abs(grads[0, 0] - fd[0]) / abs(fd[0]), "nu": abs(grads[0, 1] - fd[1]) / abs(fd[1])}
    return out


def cpu_baseline(target_seconds: float = 15.0):
    """The oracle's FP32 C++/OpenMP restatement (oracle/cloth_ref.cpp): the
    same substep with forward-mode dual numbers carrying dx/dE and dx/dnu (one
    pass gives the loss and both gradients).  Whole windows of the batched
    workload's instances, one after the other from instance 0, until
    ~target_seconds have run on all the host threads OpenMP is given; then
    part of instance 0's window on one thread for a quarter of that."""
    from diffcp_amd import scene
    from oracle import cloth_ref as CR

    CR.build()
    w = scene.config_c3()
    threads = CR.max_threads()
    done, secs, n_inst = 0, 0.0, 0
    for inst in w.instances:
        targets = scene.sample_targets(inst.vertex_x, N_TARGETS, 0)
        _, _, t, _ = CR.loss_grad(inst, targets, w.substeps, threads=threads)
        done += inst.n_cloth_particles * w.substeps
        secs += t
        n_inst += 1
        if secs >= target_seconds:
            break
    inst = w.instances[0]
    targets = scene.sample_targets(inst.vertex_x, N_TARGETS, 0)
    _, _, t10, _ = CR.loss_grad(inst, targets, 10, threads=1)
    n1 = int(min(w.substeps, max(10, round(target_seconds / 4 / max(t10 / 10, 1e-6)))))
    _, _, secs1, _ = CR.loss_grad(inst, targets, n1, threads=1)
    return {"value": done / secs, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(),
            "single_thread_value": inst.n_cloth_particles * n1 / secs1,
            "sample": (f"instances 0..{n_inst - 1} of the batched workload ({w.substeps}-substep windows, "
                       f"{w.instances[0].n_grid}^3 grids), forward with dE/dnu dual numbers, FP32 C++/OpenMP "
                       f"restatement, {threads} threads, {secs:.1f} s; {n1} substeps of instance 0 on 1 thread, "
                       f"{secs1:.1f} s")}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the large single-cloth leg")
    ap.add_argument("--no-det", action="store_true", help="skip the deterministic-sums batched leg")
    ap.add_argument("--eager", action="store_true",
                    help="launch kernels one by one instead of replaying graphs (for counter profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 512 MiB device-to-device copy 5 times (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4}))
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the batched workload fills one MI355X; "
              f"--gpus {args.gpus} is not supported)", file=sys.stderr)
        return 2

    from diffcp_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()

    log("batched workload: building 64 instances")
    b = bench_batched(args.steps, args.warmup, use_graph=not args.eager)
    w, S = b["w"], b["S"]
    value = w.particles * S * args.steps / b["wall"]
    log(f"batched: {value / 1e9:.4f} G particle-substeps/s, {b['wall'] / args.steps * 1e3:.1f} ms/step")

    avg_ms = b["k_ms"] / max(b["k_n"], 1)
    alg = g2p_adj_algorithmic_bytes(w, int(b["cells"]))
    achieved = alg / (avg_ms * 1e-3) / 1e9 if b["k_n"] else None
    traffic = load_traffic()
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                "kernel": "k_" + ROOFLINE_KERNEL, "launches_timed": b["k_n"], "avg_launch_us": avg_ms * 1e3,
                "algorithmic_bytes_per_launch": alg}

    out = {
        "metric": "particle-substeps/s fwd+adjoint, batched vs plain, 1 GPU; grad rel err vs FD",
        "value": value, "unit": "particle-substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": b["wall"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded cloths, trajectories and ground-truth targets)",
        "config": {"workload": "batched parameter identification: 64 instances, sides 32..96 vertices, "
                               "128^3 grids, 300 substeps fwd+adj, hipGraph loop",
                   "instances": len(w.instances), "particles": w.particles, "substeps": S,
                   "checkpoint_every": b["K"], "precision": "fp32 bulk, fp64 2x2 polar + accumulation"},
        "roofline": roofline,
        "status": b["status"], "finite": b["finite"],
    }
    if not args.no_det:
        log("batched workload with deterministic (fixed-point) scatter sums")
        d = bench_batched(2, 1, use_graph=not args.eager, deterministic=True, w=w)
        out["deterministic"] = {"value": w.particles * S * 2 / d["wall"], "ms_per_step": d["wall"] / 2 * 1e3,
                                "g2p_adj_avg_launch_us": d["k_ms"] / max(d["k_n"], 1) * 1e3,
                                "status": d["status"], "finite": d["finite"],
                                "note": "dcp_options.deterministic = 1: bitwise reproducible windows"}
    if not args.no_plain:
        log("plain workload (single instance, per-kernel launches)")
        p = bench_plain()
        out["plain"] = {"value": p["rate"], "ms_per_rollout": p["ms"], "particles": p["particles"],
                        "substeps": p["S"]}
        out["batched_over_plain"] = value / p["rate"]
        g = grad_vs_fd()
        if g is not None:
            out["grad_rel_err_fd"] = g["reference"]
            out["grad_rel_err_fd_stable_svd_grad"] = g["stable"]
            out["grad_fd_reference"] = ("FP64 oracle central FD, h = 1e-4 relative, 32x32 cloth, 64^3 grid, "
                                        "200 substeps (tests/golden/c1_grad_fd.npz)")
    if not args.no_large:
        log("large single cloth (256x256, 256^3 grid, 1000 substeps)")
        out["large"] = bench_large()
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle restatement on the host)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
