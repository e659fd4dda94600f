"""Launch-shape autotuner of the CRKSPH passes.

Every pair pass runs with G lanes per particle row (crk_set_group_size), and
the passes that stage LDS tiles run one block of 256, 512 or 1024 threads per
tile (crk_set_block_size).  The best shape depends on the pass (register
pressure and LDS per block, hence occupancy; number of compensated
accumulators to combine; edge-array traffic), the dimension and the
neighbour count, so it is measured: for each pass in turn, every candidate
(G, threads) is timed over full RK2 steps of the caller's workload on a
scratch copy of the state (HIP events on the launch stream), and the fastest
is kept.

A tuning is a dict {pass: G, "<pass>.threads": threads}.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional

import torch

from . import abi
from .runtime import runtime
from .timing import PassTimer

# pass -> timing scope that contains it
SCOPES = {
    "owen": "owen", "support": "support", "moments": "factors_eos", "velgrad": "velocity_gradient",
    "switch": "switch", "accel": "accel_dudt", "balance": "balance", "compat": "compatible_energy",
}


# passes that run on LDS tiles (their block size is a knob too)
TILED = ("owen", "support", "moments", "velgrad", "accel", "compat")
THREADS = (256, 512, 1024)


def set_groups(device, groups: Dict[str, int]):
    rt = runtime(device)
    for name, g in groups.items():
        if name.endswith(".threads"):
            abi.call("crk_set_block_size", rt.ctx, abi.PASSES.index(name[:-len(".threads")]), int(g))
        else:
            abi.call("crk_set_group_size", rt.ctx, abi.PASSES.index(name), int(g))
    rt.groups = dict(getattr(rt, "groups", {}), **groups)


def _copy(system):
    s = system.initializeNewState()
    s.adjacency = system.adjacency
    return s


def autotune(system, config, scheme, dt: float, candidates: Iterable[int] = (1, 2, 4, 8, 16, 32),
             reps: int = 2, passes: Optional[Iterable[str]] = None, verbose: bool = False) -> Dict[str, int]:
    from .integrator import getIntegrator
    from .scheme import crkSPH_step
    dev = system.state.positions.device
    rt = runtime(dev)
    integ = getIntegrator()
    current = dict(getattr(rt, "groups", {}))
    defaults = {"owen": 4, "support": 32, "moments": 2, "velgrad": 8, "switch": 8, "accel": 16, "balance": 32,
                "compat": 32, "owen.threads": 256, "support.threads": 256, "moments.threads": 256,
                "velgrad.threads": 256, "accel.threads": 1024, "compat.threads": 256}
    for k, v in defaults.items():
        current.setdefault(k, v)
    set_groups(dev, current)
    # one untimed step so that the adjacency exists and memory is allocated
    base = _copy(system)
    base = integ.function(base, crkSPH_step, dt, config, scheme).state
    torch.cuda.synchronize()
    names = list(passes) if passes is not None else list(SCOPES)
    separate_switch = (scheme.viscositySwitchParams.scheme.name == "CullenDehnen2010" and
                       (not rt.fuse_switch or bool(getattr(scheme, "alphaWeightedViscosity", False))))
    if not separate_switch and "switch" in names:
        names.remove("switch")
    report = {}
    for name in names:
        tkey = name + ".threads"
        shapes = [(g, tb) for g in candidates for tb in (THREADS if name in TILED else (None,))]
        best, best_t = (current[name], current.get(tkey)), None
        for g, tb in shapes:
            shape = {name: g} if tb is None else {name: g, tkey: tb}
            set_groups(dev, shape)
            s = _copy(base)
            integ.function(s, crkSPH_step, dt, config, scheme)  # warm this shape
            with PassTimer() as pt:
                for _ in range(reps):
                    s = integ.function(_copy(base), crkSPH_step, dt, config, scheme).state
            summ = pt.summary()
            if SCOPES[name] not in summ:  # the pass does not run in this configuration
                break
            t = summ[SCOPES[name]][1]
            report.setdefault(name, {})[g if tb is None else f"{g}/{tb}"] = t
            if best_t is None or t < best_t:
                best, best_t = (g, tb), t
        current[name] = best[0]
        shape = {name: best[0]}
        if best[1] is not None:
            current[tkey] = best[1]
            shape[tkey] = best[1]
        set_groups(dev, shape)
        if verbose and name in report:
            print(f"autotune {name}: " + ", ".join(f"{k}:{t:.3f}ms" for k, t in report[name].items()) + f" -> {best}")
    rt.tuning_report = report
    return current


def save(groups: Dict[str, int], path: str):
    with open(path, "w") as f:
        json.dump(groups, f, indent=1)


def load(path: str) -> Optional[Dict[str, int]]:
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return {k: int(v) for k, v in json.load(f).items()}
