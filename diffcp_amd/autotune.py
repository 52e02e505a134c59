"""Autotuner for the launch and patch configuration of a workload.

The knobs (all in ``dcp_options``):

* ``bin_cells``  — side of the spatial bins patches are cut from; it bounds how
  far a patch's particles spread and therefore its LDS tile;
* ``patch_size`` — particles per patch = particles per 256-lane workgroup of
  the per-patch kernels (the particles-per-wave knob);
* ``scatter_wgs`` — persistent workgroups of the per-patch kernels;
* ``grid_wgs`` — persistent workgroups of the grid update and its adjoint.

``tune`` times short forward + adjoint windows of the caller's own instances
(device time from ``dcp_rollout_grad``, graphs on) and walks the knobs one at
a time (coordinate descent from the defaults), keeping the fastest value of
each.  Results are cached in a JSON file (``$DIFFCP_TUNE_CACHE``, default
``~/.cache/diffcp_amd/autotune.json``, one per machine) keyed by a signature
of the workload (instance count, grid sizes, particle counts) and the window
length.
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Dict, Optional, Sequence

import numpy as np

from .engine import Engine
from .scene import Instance, sample_targets

DEFAULTS = {"bin_cells": 6, "patch_size": 256, "scatter_wgs": 0, "grid_wgs": 0}
CANDIDATES = {
    "bin_cells": (4, 5, 6, 8, 10),
    "patch_size": (128, 192, 256),
    "scatter_wgs": (0, 2048, 8192),
    "grid_wgs": (0, 512, 2048),
}
ORDER = ("bin_cells", "patch_size", "scatter_wgs", "grid_wgs")


def signature(instances: Sequence[Instance]) -> str:
    parts = [f"{len(instances)}"] + [f"{i.n_grid}:{i.n_vertices}:{i.n_elements}:{i.n_cube}:{i.n_grasper}"
                                     for i in instances]
    return hashlib.sha1(";".join(parts).encode()).hexdigest()[:16]


def _cache_path() -> str:
    base = os.environ.get("DIFFCP_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "diffcp_amd", "autotune.json")
    return base


def _load_cache() -> Dict[str, Dict]:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: Dict[str, Dict]) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass


def measure(instances: Sequence[Instance], config: Dict[str, int], substeps: int = 24, reps: int = 2,
            **engine_kw) -> float:
    """Device ms of one forward + adjoint window of `substeps` substeps (best of `reps`)."""
    insts = list(instances)
    for inst in insts:
        if inst.targets is None:
            inst.targets = sample_targets(inst.vertex_x, 256, 0)
    eng = Engine(insts, max_states=Engine.slots_for(substeps, substeps), **config, **engine_kw)
    try:
        eng.reset(0)
        eng.rollout_grad(substeps, 0, substeps, use_graph=True)   # capture
        best = np.inf
        for _ in range(reps):
            _, _, ms = eng.rollout_grad(substeps, 0, substeps, use_graph=True)
            best = min(best, ms)
        return float(best)
    finally:
        eng.close()


def tune(instances: Sequence[Instance], substeps: int = 24, use_cache: bool = True, verbose: bool = False,
         candidates: Optional[Dict[str, Sequence[int]]] = None, **engine_kw) -> Dict[str, int]:
    """Fastest configuration found for these instances; pass it to ``Engine(**config)``."""
    key = signature(instances) + f":{substeps}"
    cache = _load_cache() if use_cache else {}
    if key in cache:
        return dict(cache[key]["config"])
    cands = dict(CANDIDATES)
    if candidates:
        cands.update(candidates)
    best = dict(DEFAULTS)
    best_ms = measure(instances, best, substeps, **engine_kw)
    trials = [(dict(best), best_ms)]
    if verbose:
        print(f"[autotune] {best} -> {best_ms:.2f} ms (defaults)", flush=True)
    for knob in ORDER:
        for value in cands[knob]:
            if value == best[knob]:
                continue
            cfg = dict(best)
            cfg[knob] = value
            ms = measure(instances, cfg, substeps, **engine_kw)
            trials.append((cfg, ms))
            if verbose:
                print(f"[autotune] {cfg} -> {ms:.2f} ms", flush=True)
            if ms < best_ms:
                best, best_ms = cfg, ms
    if use_cache:
        cache[key] = {"config": best, "ms": best_ms, "trials": [[c, m] for c, m in trials]}
        _store_cache(cache)
    return best
