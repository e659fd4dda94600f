"""Throughput of the MI355X CRKSPH step on the 2D periodic shear layer.

Workload: 1024 x 1024 = 1,048,576 particles, [0,1]^2 periodic, Wendland C2
with n_H = 32, Owen adaptive h, Cullen-Dehnen switch, Heun RK2 with the
compatible-energy finalize, FP32 storage.  One "step" is one full RK2 step
(two derivative evaluations + finalize) of every particle.

Prints one JSON line (see the keys below).  ``roofline`` is for the dominant
pass, timed with HIP events on the launch stream; ``cpu_baseline`` times the
C++/OpenMP CPU oracle on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# lane-group sizes chosen by the autotuner for the 1M shear layer on MI355X
DEFAULT_TUNING = os.path.join(ROOT, "warpsph_amd", "tuning", "mi355x_shear2d_1M.json")


# Algorithmic bytes per launch, from the reference dataflow (each per-particle
# array of a reference pass read once and written once; neighbour gathers
# cache resident): per-particle bytes of the reference passes that one launch
# performs, 2D FP32, plus bytes per pair entry the launch processes (ap, av
# written: 8 + 8 B).  Reference passes: Owen 16; CRK volume 16, v_sig 28;
# moments + solve 52, CRK density 32, EOS 20; grad v 76, CD divergence 32;
# R 28, alpha update 32; accel 100, dudt 96; balance 32; finalize 28.
REF_BYTES_2D = {
    "owen": (16, 0),
    "support": (16 + 28, 0),
    "factors_eos": (52 + 32 + 20, 0),
    "velocity_gradient": (76 + 32, 0),
    "switch": (28 + 32, 0),
    "accel_dudt": (100 + 96, 16),
    "accel_dudt+switch": (100 + 96 + 28 + 32, 16),
    "balance": (32, 0),
    "compatible_energy": (28, 0),
}


def step_bytes(N: int, E: float, r: float) -> float:
    """B_alg(step) = N (2*632 + 52 + 164 r) + E (2*60 + 4 r): the reference
    dataflow of one RK2 step, 2D FP32, r = adjacency rebuilds per step."""
    return N * (2 * 632 + 52 + 164 * r) + E * (2 * 60 + 4 * r)


def measured_traffic(*kernel_prefixes: str):
    """HBM bytes per launch of an entry point (the sum over its kernels) from
    the committed counter summary (FETCH_SIZE / WRITE_SIZE scaled by the copy
    calibration), or None."""
    path = os.path.join(ROOT, "profiles", "crksph_step_shear2d_1M_hbm_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    total, found = 0.0, 0
    for prefix in kernel_prefixes:
        for name, v in data.get("kernels", {}).items():
            if name.startswith(prefix):
                total += v["traffic_bytes"]
                found += 1
                break
    return total if found == len(kernel_prefixes) else None


def run_gpu(nx, steps, warmup, timer=True, tuning_file=None, retune=False):
    import torch
    from warpsph_amd import sample
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.scheme import crkSPH_step
    from warpsph_amd.timing import PassTimer
    system, config, scheme = sample.shear_layer_2d(nx=nx, device="cuda:0")
    dt = sample.cfl_dt(system.state, config, scheme) * (2.0 / 3.0)
    integ = getIntegrator()
    groups = None
    if tuning_file and os.path.exists(tuning_file) and not retune:
        from warpsph_amd.autotune import load, set_groups
        groups = load(tuning_file)
        set_groups(system.state.positions.device, groups)
    else:
        from warpsph_amd.autotune import autotune, save
        groups = autotune(system, config, scheme, dt)
        if tuning_file:
            save(groups, tuning_file)
    s = system
    for _ in range(warmup):
        s = integ.function(s, crkSPH_step, dt, config, scheme).state
    torch.cuda.synchronize()
    # Move the objects that exist now (torch's modules, the system) out of the
    # collector's view: a full collection over them would stall the host for
    # tens of ms in the middle of the timed steps (as timeit avoids by
    # disabling the collector; here it stays enabled for new objects).
    gc.collect()
    gc.freeze()
    edges, pairs = [], []
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    lists = []  # the adjacency of each stage (a new object = a rebuild)
    t0 = time.perf_counter()
    marks[0].record()
    for it in range(steps):
        r = integ.function(s, crkSPH_step, dt, config, scheme)
        s = r.state
        marks[it + 1].record()
        edges.append(sum(st.aux[0].n_edges for st in r.stages) / len(r.stages))
        lists += [st.aux[0] for st in r.stages]
        last = r
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the per-pass times come from further steps with HIP events around each
    # pass (the events' host cost stays out of the timed steps above)
    pt = PassTimer()
    if timer:
        with pt:
            for _ in range(min(steps, 10)):
                s = integ.function(s, crkSPH_step, dt, config, scheme).state
        torch.cuda.synchronize()
    # in-support pairs per stage of the last step (read after the timed region)
    P_mean = sum(st.aux[1].pairs.n_pairs for st in last.stages) / len(last.stages)
    seq = [marks[i].elapsed_time(marks[i + 1]) for i in range(steps)]
    step_ms = sorted(seq)
    spread = {"min": step_ms[0], "median": step_ms[len(step_ms) // 2], "max": step_ms[-1]}
    if os.environ.get("CRK_BENCH_STEP_TIMES"):
        spread["all"] = [round(t, 3) for t in seq]
    spread["rebuilds"] = sum(1 for a, b in zip(lists, lists[1:]) if b is not a)
    return (s, (t1 - t0), pt.summary() if timer else {}, float(sum(edges) / max(len(edges), 1)), float(P_mean), dt,
            groups, spread)


def host_description(threads):
    """The CPU the baseline ran on: nproc, the lscpu model name, the OpenMP
    settings and the threads the oracle used."""
    import platform
    import subprocess
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "threads_used": threads}


def cpu_baseline(nx, nsteps, warm):
    """C++/OpenMP CPU oracle (FP32, Kahan) on an nx^2 sample; returns the
    rate and the oracle's final density for the accuracy check."""
    import numpy as np
    from oracle import oracle as O
    from warpsph_amd import sample
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = sample.shear_layer_2d(nx=nx, device="cpu")
    sim = O.OracleSim(toCrkConfig(config, scheme), system.state.positions.shape[0], 32)
    for k, name in O.STATE_MAP.items():
        sim.set(k, getattr(system.state, name).double().numpy())
    dt = sim.cfl_dt() * (2.0 / 3.0)
    sim.step(dt, warm)
    t0 = time.perf_counter()
    done = 0
    while done < nsteps:  # in chunks, with a progress line for long samples
        k = min(10, nsteps - done)
        sim.step(dt, k)
        done += k
        print(f"cpu baseline: {done}/{nsteps} steps, {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    n = system.state.positions.shape[0]
    return n * nsteps / (t1 - t0), sim.get("rho"), dt, O.num_threads(), (t1 - t0)


def blast3d(args):
    """3D point blast (128^3 jittered lattice, Wendland C4, n_H = 64): step
    throughput on the GPU, plus the precision statement at a reduced size:
    FP32 GPU (Kahan pair sums, FP64 moment solve and reductions) and the FP32
    oracle, both against the FP64-everywhere oracle from the same state."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from warpsph_amd import sample
    from warpsph_amd.autotune import autotune, load, save, set_groups
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.scheme import crkSPH_step
    from warpsph_amd.timing import PassTimer
    torch.cuda.set_device(0)
    integ = getIntegrator()
    nx = args.nx if args.nx != 1024 else 128
    system, config, scheme = sample.blast_3d(nx=nx, device="cuda:0")
    dt = sample.cfl_dt(system.state, config, scheme)
    tf = os.path.join(ROOT, "warpsph_amd", "tuning", f"mi355x_blast3d_{nx}.json")
    if os.path.exists(tf) and not args.retune:
        groups = load(tf)
        set_groups(system.state.positions.device, groups)
    else:
        groups = autotune(system, config, scheme, dt)
        save(groups, tf)
    s = system
    for _ in range(args.warmup):
        s = integ.function(s, crkSPH_step, dt, config, scheme).state
    torch.cuda.synchronize()
    gc.collect()
    gc.freeze()
    pt = PassTimer()
    t0 = time.perf_counter()
    with pt:
        for _ in range(args.steps):
            s = integ.function(s, crkSPH_step, dt, config, scheme).state
        torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    n = nx ** 3
    summary = pt.summary()
    # precision statement at a reduced size
    ne, ksteps = 20, 10
    sysg, cfg_e, sch_e = sample.blast_3d(nx=ne, device="cuda:0")
    cpu = {k: getattr(sysg.state, v).detach().cpu().double().numpy() for k, v in O.STATE_MAP.items()}
    from warpsph_amd.configurations import toCrkConfig
    sims = {}
    for prec in (32, 64):
        sim = O.OracleSim(toCrkConfig(cfg_e, sch_e), ne ** 3, prec)
        for k, v in cpu.items():
            sim.set(k, v)
        sims[prec] = sim
    dte = sims[64].cfl_dt()
    g = sysg
    for _ in range(ksteps):
        g = integ.function(g, crkSPH_step, dte, cfg_e, sch_e).state
    torch.cuda.synchronize()
    for sim in sims.values():
        sim.step(dte, ksteps)
    ref = {k: sims[64].get(k) for k in ("rho", "u", "v")}
    gpu = {"rho": g.state.densities.double().cpu().numpy(), "u": g.state.internalEnergies.double().cpu().numpy(),
           "v": g.state.velocities.double().cpu().numpy()}

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    err = {"sample": f"{ne}^3 = {ne ** 3} particles, {ksteps} RK2 steps, dt {dte:.3e}",
           "gpu_fp32_vs_fp64_oracle": {k: rel(gpu[k].reshape(ref[k].shape), ref[k]) for k in ref},
           "oracle_fp32_vs_fp64_oracle": {k: rel(sims[32].get(k), ref[k]) for k in ref},
           "density_l1_gpu_vs_fp64": float(np.mean(np.abs(gpu["rho"] - ref["rho"])) / np.mean(np.abs(ref["rho"])))}
    line = {
        "metric": "particle-updates/s (3D CRKSPH point blast, 128^3)", "value": n * args.steps / secs,
        "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3D periodic point blast, {nx}^3 = {n} particles, Wendland C4 n_H=64, E=1, "
                               "Owen h, Cullen-Dehnen switch, Heun RK2 + compatible energy", "dt": dt},
        "precision": err,
        "passes": {k: {"calls": v[0], "mean_ms": round(v[2], 4)} for k, v in summary.items()},
        "launch_groups": groups,
    }
    print(json.dumps(line))


def ensemble(args):
    """Batched ensemble: ``--runs`` independent 2D two-state runs (sizes
    log-uniform in [4096, 65536], the data generator's seeds) stepped together,
    every run with its own CFL dt.  Reports whole-ensemble particle-updates/s."""
    import numpy as np
    import torch
    from warpsph_amd import ensemble as ens_mod
    from warpsph_amd.autotune import load, set_groups
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.scheme import crkSPH_step
    from warpsph_amd.timing import PassTimer
    torch.cuda.set_device(0)
    integ = getIntegrator()
    t0 = time.perf_counter()
    e, config, scheme = ens_mod.sample_ensemble(n_runs=args.runs, nmin=args.run_min, nmax=args.run_max,
                                                seed=args.seed, device="cuda:0")
    gen_s = time.perf_counter() - t0
    # the ensemble's own launch shapes (autotuned on the 256-run workload)
    # unless a tuning file is given
    tf = args.tuning_file
    if tf == DEFAULT_TUNING:
        tf = os.path.join(ROOT, "warpsph_amd", "tuning", "mi355x_ensemble_256.json")
    if os.path.exists(tf):
        groups = load(tf)
        set_groups(e.state.positions.device, groups)
    else:
        groups = None
    n = e.runs.n
    summary, graph_stats = {}, None
    if args.graph:
        from warpsph_amd.graph import GraphStepper
        gs = GraphStepper(e, config, scheme, integ)
        for _ in range(args.warmup):
            gs.step()
        torch.cuda.synchronize()
        gc.collect()
        gc.freeze()
        r0, c0, d0, q0 = gs.replays, gs.captures, gs.redos, gs.eager_steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            gs.step()
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        graph_stats = {"replays": gs.replays - r0, "captures": gs.captures - c0, "eager_redos": gs.redos - d0,
                       "eager_steps": gs.eager_steps - q0}
        e.t = gs.times()
    else:
        prev = None
        for _ in range(args.warmup):
            prev = ens_mod.cfl_dt(e, config, scheme, prev)
            e = integ.function(e, crkSPH_step, prev, config, scheme).state
        torch.cuda.synchronize()
        gc.collect()
        gc.freeze()
        pt = PassTimer()
        t0 = time.perf_counter()
        with pt:
            for _ in range(args.steps):
                prev = ens_mod.cfl_dt(e, config, scheme, prev)
                e = integ.function(e, crkSPH_step, prev, config, scheme).state
            torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        summary = pt.summary()
    counts = e.runs.counts
    line = {
        "metric": "particle-updates/s (2D CRKSPH batched ensemble)", "value": n * args.steps / secs,
        "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.runs} independent 2D two-state runs (random line split, rho, P ~ U[0.1,1], "
                               f"v ~ U[-0.5,0.5]^2, gamma 5/3), N log-uniform in [{args.run_min}, {args.run_max}], "
                               "Wendland C2 n_H=32, Owen h, Cullen-Dehnen switch, Heun RK2 + compatible energy, "
                               "per-run CFL dt" + (", steps replayed as HIP graphs" if args.graph else ""),
                   "particles": n, "runs": args.runs, "run_min": int(counts.min()), "run_max": int(counts.max()),
                   "seed": args.seed},
        "graph": graph_stats,
        "generate_s": gen_s,
        "t_run_range": [float(np.min(e.t)), float(np.max(e.t))],
        "passes": {k: {"calls": v[0], "mean_ms": round(v[2], 4)} for k, v in summary.items()},
        "launch_groups": groups,
    }
    print(json.dumps(line))


def learned(args):
    """Learned viscosity closure: training data from coarse / 4x-resolution
    pairs of the two-state generator, training on the GPU (BF16 MFMA kernels,
    Adam), then a config-2-shape shear layer (--nx^2, alpha-weighted Q): the
    coarse run with the learned switch and with the Cullen-Dehnen switch
    against the Cullen-Dehnen run at 4x the particle count (L1 of the SPH
    densities averaged per cell of a 256^2 grid at --t-end), with wall times."""
    import numpy as np
    import torch
    from warpsph_amd import learned as Lm
    from warpsph_amd import sample
    from warpsph_amd.enumTypes import ViscositySwitch
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    feats, targs = [], []
    for k in range(args.pairs):
        f, t = Lm.pair_samples(1000 + k, nx=args.train_nx, device="cuda:0")
        feats.append(f)
        targs.append(t)
        print(f"[learned] pair {k}: {f.shape[0]} samples, target mean {float(t.mean()):.3f}", flush=True)
    X, T = torch.cat(feats), torch.cat(targs)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    model, hist = Lm.fit(X, T, epochs=args.epochs, batch=16384, lr=2e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    model.save(args.model_out or os.path.join(ROOT, "warpsph_amd", "tuning", "learned_switch.npz"))
    print(f"[learned] trained: loss {hist[0]:.5f} -> {hist[-1]:.5f} ({train_s:.1f} s)", flush=True)

    def run(nx, switch, mdl=None):
        system, config, scheme = sample.shear_layer_2d(nx=nx, device="cuda:0")
        scheme.alphaWeightedViscosity = True
        scheme.viscositySwitchParams.scheme = switch
        scheme.viscositySwitchParams.model = mdl
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        s = Lm.advance(system, config, scheme, args.t_end)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        rho, _ = Lm.density_field(s.state, 256)
        fb = None
        if mdl is not None:  # fraction of particles on the Cullen-Dehnen fallback in the final state
            f = Lm.stage_features(s, config, scheme)
            lo = torch.as_tensor(mdl.lo, device=f.device)
            hi = torch.as_tensor(mdl.hi, device=f.device)
            fb = float((~((f >= lo) & (f <= hi)).all(dim=1)).float().mean())
        return rho, wall, fb

    rho_ref, wall_ref, _ = run(2 * args.nx, ViscositySwitch.CullenDehnen2010)
    print(f"[learned] 4x run: {wall_ref:.1f} s", flush=True)
    rho_cd, wall_cd, _ = run(args.nx, ViscositySwitch.CullenDehnen2010)
    rho_ml, wall_ml, fallback = run(args.nx, ViscositySwitch.Learned, model)

    def l1(a, b):
        return float((a - b).abs().mean() / b.abs().mean())

    hi_ref = None
    if args.hi_ref:  # a 16x-resolution Cullen-Dehnen run as the reference of all three
        rho_hi, wall_hi, _ = run(4 * args.nx, ViscositySwitch.CullenDehnen2010)
        print(f"[learned] 16x run: {wall_hi:.1f} s", flush=True)
        hi_ref = {"l1_learned_vs_16x": l1(rho_ml, rho_hi), "l1_cullen_dehnen_vs_16x": l1(rho_cd, rho_hi),
                  "l1_cullen_dehnen_4x_vs_16x": l1(rho_ref, rho_hi), "wall_s_16x": wall_hi}

    # the fraction of particles that fell back to Cullen-Dehnen in the evaluation's last state
    line = {
        "metric": "density L1 vs 4x-resolution run, learned vs Cullen-Dehnen switch (2D shear layer)",
        "value": l1(rho_ml, rho_ref), "unit": "relative L1", "higher_is_better": False, "n_gpus": 1,
        "dtype": "bf16 MLP (fp32 accumulate), f32 solver", "data": "synthetic",
        "config": {"workload": f"shear layer {args.nx}^2 (alpha-weighted Q) to t = {args.t_end}, reference "
                               f"{2 * args.nx}^2; training: {args.pairs} two-state pairs {args.train_nx}^2 / "
                               f"{2 * args.train_nx}^2, {int(X.shape[0])} samples, {args.epochs} epochs"},
        "l1_learned_vs_4x": l1(rho_ml, rho_ref), "l1_cullen_dehnen_vs_4x": l1(rho_cd, rho_ref),
        "wall_s": {"learned": wall_ml, "cullen_dehnen": wall_cd, "cullen_dehnen_4x": wall_ref},
        "speedup_learned_vs_4x": wall_ref / wall_ml, "fallback_fraction_final": fallback,
        "against_16x": hi_ref,
        "training": {"generate_s": gen_s, "train_s": train_s, "loss_first": hist[0], "loss_last": hist[-1]},
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--cpu-nx", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tuning-file", default=DEFAULT_TUNING,
                    help="load (or, with --retune or when absent, save after tuning) the lane-group sizes")
    ap.add_argument("--retune", action="store_true")
    ap.add_argument("--workload", default="shear2d", choices=["shear2d", "blast3d", "ensemble", "learned"],
                    help="shear2d: the 1M-particle 2D headline; blast3d: the 3D point blast with the FP64 precision "
                         "statement; ensemble: the batched two-state runs (--runs, --seed)")
    ap.add_argument("--runs", type=int, default=256)
    ap.add_argument("--run-min", type=int, default=4096, help="ensemble: smallest run")
    ap.add_argument("--run-max", type=int, default=65536, help="ensemble: largest run")
    ap.add_argument("--graph", action="store_true", help="ensemble: replay the steps as HIP graphs")
    ap.add_argument("--pairs", type=int, default=6, help="learned: training pairs")
    ap.add_argument("--train-nx", type=int, default=256, help="learned: coarse side of a training pair")
    ap.add_argument("--epochs", type=int, default=30, help="learned: training epochs")
    ap.add_argument("--t-end", type=float, default=0.4, help="learned: evaluation end time")
    ap.add_argument("--hi-ref", action="store_true", help="learned: also run 16x the particles as the reference")
    ap.add_argument("--model-out", default=None, help="learned: where to save the trained network (.npz)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 256 MiB buffer 4 times (calibrates the HBM traffic counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": a.numel() * 4, "copies": 4}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "the CRKSPH step runs on one GPU only; --gpus N with N > 1 is not "
                                   "supported (domain decomposition over GPUs is not built)"}))
        sys.exit(2)

    if args.workload == "blast3d":
        blast3d(args)
        return
    if args.workload == "ensemble":
        ensemble(args)
        return
    if args.workload == "learned":
        learned(args)
        return

    import numpy as np
    import torch
    torch.cuda.set_device(0)

    s, secs, summary, E_mean, P_mean, dt, groups, spread = run_gpu(args.nx, args.steps, args.warmup, tuning_file=args.tuning_file,
                                                   retune=args.retune)
    n = args.nx * args.nx
    value = n * args.steps / secs

    # dominant pass by total time
    dom = max(summary.items(), key=lambda kv: kv[1][1]) if summary else None
    roofline = None
    from warpsph_amd.runtime import runtime
    fused = bool(getattr(runtime(s.state.positions.device), "fuse_switch", False))
    if dom is not None:
        name, (calls, total_ms, mean_ms) = dom
        key = name + ("+switch" if name == "accel_dudt" and fused else "")
        pb = REF_BYTES_2D.get(key)
        if pb is not None:
            bytes_launch = pb[0] * n + pb[1] * P_mean
            achieved = bytes_launch / (mean_ms * 1e-3)
            kernel = {"accel_dudt": "crk::k_accel_dudt", "support": "crk::k_support", "factors_eos": "crk::k_moments",
                      "compatible_energy": "crk::k_compat"}.get(name, "crk::k_" + name)
            # the acceleration entry point launches the record pack and the sweep
            kernels = (kernel, "crk::k_pack_rec") if name == "accel_dudt" else (kernel,)
            roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK, "traffic": measured_traffic(*kernels),
                        "kernel": " + ".join(kernels),
                        "algorithmic_bytes_per_launch": bytes_launch, "mean_launch_ms": mean_ms,
                        "bytes_per_particle": pb[0], "bytes_per_pair": pb[1], "pairs_per_launch": P_mean}
    rebuilds = spread.pop("rebuilds") / max(args.steps, 1)  # adjacency rebuilds per timed step
    sb = step_bytes(n, E_mean, rebuilds)
    step_roofline = {"bytes_per_step": sb, "edges": E_mean, "rebuilds_per_step": rebuilds,
                     "achieved": sb / (secs / args.steps) / 1e9, "unit": "GB/s",
                     "frac": sb / (secs / args.steps) / HBM_PEAK}

    cpu = None
    l1 = None
    if not args.no_cpu_baseline:
        rate, rho_orc, dt_c, threads, wall = cpu_baseline(args.cpu_nx, args.cpu_steps, 1)
        cpu = {"value": rate, "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "sample": f"C++/OpenMP FP32 CPU oracle, {args.cpu_nx}x{args.cpu_nx} shear layer, "
                         f"{args.cpu_steps} RK2 steps after 1 warm-up step ({wall:.1f} s)",
               "label": "build's C++ CPU oracle, not the reference Warp CPU path",
               "host": host_description(threads)}
        # density L1 of the GPU against the oracle on the same sample and steps
        from warpsph_amd import sample
        from warpsph_amd.integrator import getIntegrator
        from warpsph_amd.scheme import crkSPH_step
        sysg, config, scheme = sample.shear_layer_2d(nx=args.cpu_nx, device="cuda:0")
        g = sysg
        for _ in range(1 + args.cpu_steps):
            g = getIntegrator().function(g, crkSPH_step, dt_c, config, scheme).state
        rho_g = g.state.densities.double().cpu().numpy()
        l1 = float(np.mean(np.abs(rho_g - rho_orc)) / np.mean(np.abs(rho_orc)))

    passes = {k: {"calls": v[0], "total_ms": round(v[1], 4), "mean_ms": round(v[2], 4)} for k, v in summary.items()}
    line = {
        "metric": "particle-updates/s (2D CRKSPH, 1M) + density L1 vs ref; learned vs plain wall@err",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"2D periodic shear layer, {args.nx}x{args.nx} = {n} particles, Wendland C2 "
                               f"n_H=32, Owen h, Cullen-Dehnen switch, Heun RK2 + compatible energy",
                   "particles": n, "edges_per_stage": E_mean, "pairs_per_stage": P_mean, "dt": dt},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "cpu_baseline": cpu,
        "density_l1_vs_oracle": None if l1 is None else {"value": l1, "sample": f"{args.cpu_nx}^2, {1 + args.cpu_steps} steps"},
        "passes": passes,
        "launch_groups": groups,
        "step_ms": spread,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
