"""The CRKSPH derivative evaluation and the scheme builder.

``crkSPH_step`` has the signature and side effects of the reference's
``crkSPH_step`` (src/warpSPH/schemes/crkSPH.py:19-366); every pass runs in the
gfx950 HIP library through the C-ABI of include/crksph.h.  There is no CPU
path: without the library the call raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Optional

import numpy as np
import torch

from . import abi, boundary, timing
from .configurations import (CRKSPHConfig, SimulationConfig, crkSPHConfigToDict, dictToCRKSPHConfig,
                             toCrkConfig)
from .enumTypes import AdaptiveSupportScheme, CompressibleSPHScheme, ViscositySwitch
from .runtime import begin_checked, begin_speculative, ensure_adjacency, rebuild_adjacency, runtime
from .systems import CompressibleSystemUpdate, CompSPHState, CompSPHSystem, StagePairs

# dt argument of a step whose per-particle dt is already on the device
# (bound with the run table; graph replay)
DEVICE_DT = "device"


def _require(st: CompSPHState, names):
    for k in names:
        if getattr(st, k) is None:
            raise ValueError(f"{k} must be provided as a property of the particle state.")


def _fill_defaults(st: CompSPHState):
    z = st.supports
    for k in ("densities", "entropies", "pressures", "soundspeeds"):
        if getattr(st, k) is None:
            setattr(st, k, torch.zeros_like(z))
    if st.divergence is None:
        st.divergence = torch.zeros_like(z)
    if st.alpha0s is None:
        st.alpha0s = torch.ones_like(z)
    if st.alphas is None:
        st.alphas = torch.ones_like(z)


_FALSE = {}


def _no_passive(n: int, device) -> torch.Tensor:
    """The update's passive mask (no passive particles): an n-element view of
    one cached False per device (no fill per stage; writes to it raise)."""
    key = str(device)
    z = _FALSE.get(key)
    if z is None:
        z = torch.zeros(1, dtype=torch.bool, device=device)
        if not torch.cuda.is_current_stream_capturing():  # (graph pool memory is not kept)
            _FALSE[key] = z
    return z.expand(n)


def crkSPH_step(system: CompSPHSystem, dt: float, config: SimulationConfig, compParams: CRKSPHConfig,
                verbose: bool = False):
    """One CRKSPH derivative evaluation (crkSPH.py:19-366).

    Side effects on ``system.state``: supports, densities, entropies,
    pressures, soundspeeds, alphas (alpha0s when evolveAlpha0), divergence,
    pairs (read as ap_ij, av_ij, f_ij); ``system.adjacency`` is the adjacency used.
    Returns ``(update, adjacency, state)``.

    For an ensemble (``system.runs`` set, warpsph_amd.ensemble) ``dt`` is the
    per-run dt (or one value for every run); each run is evaluated as it
    would be alone.
    """
    runs = getattr(system, "runs", None)
    bcs = list(getattr(compParams, "boundaryConditions", None) or [])  # (bcs.py hooks, below)
    if bcs and runs is not None:
        raise NotImplementedError("boundary conditions are supported for single runs (not for ensembles)")
    if runs is not None:
        if isinstance(dt, str) and dt == DEVICE_DT:
            dt = 1.0  # each particle's dt is already on the device (the library reads it)
        else:
            runs.set_dt(dt.detach().cpu().numpy() if isinstance(dt, torch.Tensor) else dt)
            dt = float(np.max(runs.dt))  # (the library reads each particle's dt)
    else:
        dt = float(dt.detach().cpu().item()) if isinstance(dt, torch.Tensor) else float(dt)
    st = system.state
    _require(st, ("positions", "velocities", "supports", "masses", "internalEnergies"))
    _fill_defaults(st)
    n = st.numParticles
    dim = int(config.domain.dim)
    if st.positions.dim() != 2 or st.positions.shape[1] != dim:
        raise ValueError(f"positions must be [N, {dim}]")
    if n == 0:
        raise ValueError("the state holds no particles")
    cfg = toCrkConfig(config, compParams)
    rt = runtime(st.positions.device)
    if bcs and rt.deferred:
        raise NotImplementedError("boundary conditions are supported for steps run eagerly (not graph-replayed)")
    s = rt.stream()
    rt.bind_runs(runs)
    if runs is not None and runs.n != n:
        raise ValueError(f"the ensemble's runs hold {runs.n} particles, the state {n}")
    cst = st.cabi()
    adj = system.adjacency

    scheme_h = compParams.adaptiveSupportScheme
    owen = scheme_h in (AdaptiveSupportScheme.Owen, AdaptiveSupportScheme.Monaghan)  # an adaptive-h pass runs
    fused_support = False
    owen_checked = False
    spec = None  # (adjacency, pairs, post-update test failed) of a speculative stage

    # the interacting pairs of this stage (compacted by the support sweep,
    # which also computes the apparent volume) and their pair data; set up
    # for the list in hand before the host waits on its test (and set up
    # again if the list is rebuilt)
    def stage_pairs(a):
        p = StagePairs(a, dim, st.masses)
        if runs is not None:
            p.runs, p.dtp = runs, runs.dtp
        return p

    if owen:
        # evaluateOptimalSupport (optimalSupport.py:12-43): Owen,
        # optimalSupportOwen.py:35-127, or Monaghan, optimalSupportMonaghan.py:20-108
        monaghan = scheme_h == AdaptiveSupportScheme.Monaghan
        if not monaghan:
            rt.ensure_lut(cfg.kernel, dim, cfg.targetNeighbors)
        update = "crk_monaghan_iteration" if monaghan else "crk_owen_iteration"
        hmm = rt.scratch[0:2] if runs is None else runs.hmm
        ratio = rt.scratch[2:3]
        iters = int(compParams.adaptiveSupportIterations)
        # a single Owen iteration also computes V and v_sig with the new
        # supports (crk_owen_support); the support sweep then folds into the moments
        fused_support = iters == 1 and rt.fuse_support and not monaghan
        if iters > 1 and rt.deferred:
            raise ValueError("deferred (graph) steps support one adaptive-support iteration")
        if iters == 1:
            # the stage's gather carries the Verlet test of the prior list and
            # the support range hMin, hMax (buildVerletList with priorNeighborhood)
            def owen_pass():
                with timing.scope("owen"):
                    abi.call("crk_owen_support" if fused_support else update, rt.ctx, C.byref(cfg),
                             C.byref(cst), abi.ptr(hmm), None, s)
            if rt.speculate and not rt.deferred and not fused_support:
                spec = begin_speculative(cfg, adj, st, cst, hmm, runs, owen_pass, stage_pairs, rt.nsing)
                if spec is not None and spec[0] is None:
                    # the list failed its test: rebuilt at once (no second test
                    # of it; an ensemble rebuilds the runs that failed it)
                    spec, adj = None, rebuild_adjacency(cfg, st.positions, st.supports, runs, adj, 0)
            if spec is None:
                adj = begin_checked(cfg, adj, st, cst, hmm, runs, then=owen_pass)
            else:
                adj = spec[0]
            owen_checked = True
        else:
            abi.call("crk_minmax", rt.ctx, n, abi.ptr(st.supports), abi.ptr(hmm), s)
            for it in range(iters):
                adj, _ = ensure_adjacency(cfg, adj, st.positions, st.supports, runs)
                A = adj.cabi()
                abi.call("crk_stage_begin", rt.ctx, C.byref(cfg), C.byref(A), C.byref(cst), s)
                with timing.scope("owen"):
                    abi.call(update, rt.ctx, C.byref(cfg), C.byref(cst), abi.ptr(hmm), abi.ptr(ratio), s)
                if it + 1 == iters:
                    break
                # hMin, hMax of the new supports bound the next iteration
                # (optimalSupportOwen.py:103-104, optimalSupportMonaghan.py:74-75)
                abi.call("crk_minmax", rt.ctx, n, abi.ptr(st.supports), abi.ptr(hmm), s)
                if float(ratio.item()) < compParams.adaptiveSupportThreshold:
                    break
        # The reference's final gather-density pass (optimalSupportOwen.py:116-125,
        # optimalSupportMonaghan.py:96-107) is overwritten by computeCRKFactors
        # (crkSPH.py:61) before any use; it is not evaluated.

    support_done = False
    # buildVerletList(SuperSymmetric, verletScale, prior), crkSPH.py:54-58
    if spec is not None:
        # the speculative stage's pairs and support sweep (valid unless the
        # updated supports failed the list's test)
        pairs = spec[1]
        rebuilt = spec[2]
        support_done = not rebuilt
        if rebuilt:
            adj = rebuild_adjacency(cfg, st.positions, st.supports, runs, adj, 1)
    elif owen_checked:
        pairs = stage_pairs(adj) if adj is not None else None
        rt.nsing.zero_()
        # positions are unchanged since the checked gather: only the supports moved
        if rt.deferred:
            abi.call("crk_owen_valid", rt.ctx, C.byref(cfg), None, s)  # into the device flag
            rebuilt = False
        else:
            ok = C.c_int32()
            abi.call("crk_owen_valid", rt.ctx, C.byref(cfg), C.byref(ok), s)
            rebuilt = not ok.value
        if rebuilt:
            adj = rebuild_adjacency(cfg, st.positions, st.supports, runs, adj, 1)
    else:
        pairs = None
        rt.nsing.zero_()
        adj, rebuilt = ensure_adjacency(cfg, adj, st.positions, st.supports, runs)
    system.adjacency = adj
    A = adj.cabi()
    if rebuilt or not owen:
        # (with Owen and a reused adjacency the sorted copies are current:
        # the Owen pass updated the sorted supports in place)
        abi.call("crk_stage_begin", rt.ctx, C.byref(cfg), C.byref(A), C.byref(cst), s)
    if pairs is None or rebuilt or pairs.adjacency is not adj:
        pairs = stage_pairs(adj)
    st.pairs = pairs
    if fused_support and not rebuilt:
        pairs.loc = None  # the moments pass compacts the pairs by their sorted ids only
    ed = st.pairs.cabi()
    if fused_support and not rebuilt:
        # computeCRKFactors + idealGasEOS (crkSPH.py:61, 95-101) with the
        # support sweep's compaction in the same pass
        with timing.scope("factors_eos"):
            abi.call("crk_pairs_factors_eos", rt.ctx, C.byref(cfg), C.byref(cst), C.byref(ed), abi.ptr(rt.nsing), s)
    else:
        if not support_done:
            with timing.scope("support"):
                abi.call("crk_support", rt.ctx, C.byref(cfg), C.byref(ed), s)
        # computeCRKFactors + idealGasEOS (crkSPH.py:61, 95-101)
        with timing.scope("factors_eos"):
            abi.call("crk_factors_eos", rt.ctx, C.byref(cfg), C.byref(cst), abi.ptr(rt.nsing), s)
    t_now = system.t
    if bcs:
        # enforceDirichlet between the CRK factors and the equation of state
        # (crkSPH.py:94); the library re-reads the state and redoes the EOS
        if boundary.enforceDirichlet(system, t_now, dt, config, compParams):
            abi.call("crk_state_refresh", rt.ctx, C.byref(cfg), C.byref(cst), s)
    update = CompressibleSystemUpdate(
        dxdt=st.velocities,  # dx/dt = v (the stage's velocity tensor, read only)
        dvdt=torch.empty_like(st.velocities),
        dudt=torch.empty_like(st.internalEnergies),
        dEdt=torch.empty_like(st.internalEnergies),
        drhodt=torch.empty_like(st.internalEnergies),
        passive=_no_passive(n, st.positions.device))
    cu = update.cabi()
    # velocity gradient (crkSPH.py:158-172)
    with timing.scope("velocity_gradient"):
        abi.call("crk_velocity_gradient", rt.ctx, C.byref(cfg), C.byref(cu), s)
    # viscosity switch (crkSPH.py:195-200).  Without alpha-weighted Q the
    # accelerations do not read alpha, and the Cullen-Dehnen sum is evaluated
    # in the acceleration sweep (crk_switch_accel_dudt).
    sw = compParams.viscositySwitchParams.scheme
    fused = sw == ViscositySwitch.CullenDehnen2010 and not cfg.alphaWeightedQ and rt.fuse_switch
    if sw == ViscositySwitch.CullenDehnen2010 and not fused:
        with timing.scope("switch"):
            abi.call("crk_switch", rt.ctx, C.byref(cfg), C.byref(cst), float(dt), s)
    elif sw == ViscositySwitch.Learned:
        # the MLP closure with the Cullen-Dehnen switch as its fallback
        # (warpsph_amd/learned.py); alpha enters the viscosity with alpha-weighted Q
        model = compParams.viscositySwitchParams.model
        if model is None:
            raise ValueError("ViscositySwitch.Learned needs viscositySwitchParams.model (a trained LearnedSwitch)")
        with timing.scope("switch"):
            abi.call("crk_learned_switch", rt.ctx, C.byref(cfg), C.byref(cst), float(dt), C.byref(model.cabi()), s)
    # accel + dudt (crkSPH.py:241-284), divergence and dEdt (:324-325)
    with timing.scope("accel_dudt"):
        if fused:
            abi.call("crk_switch_accel_dudt", rt.ctx, C.byref(cfg), C.byref(cst), C.byref(cu), C.byref(ed),
                     float(dt), s)
        else:
            abi.call("crk_accel_dudt", rt.ctx, C.byref(cfg), C.byref(cst), C.byref(cu), C.byref(ed), s)
    if bcs:
        # forcing (crkSPH.py:327-328, with the reference's argument order: the
        # callbacks get t and dt exchanged) and enforceUpdates (:339); the
        # energy partition re-reads dv/dt
        forcing = boundary.computeForcing(system, t_now, dt, config, compParams)
        update.dvdt += forcing / st.masses.view(-1, 1)
        boundary.enforceUpdates(update, system, dt, t_now, config, compParams)
        abi.call("crk_update_refresh", rt.ctx, C.byref(update.cabi()), s)
    # energy partition with v_half = v + dt/2 dvdt (crkSPH.py:341-364): the
    # per-particle records; f_ij is evaluated where it is used
    st.pairs.dt = float(dt)
    with timing.scope("balance"):
        abi.call("crk_balance", rt.ctx, C.byref(cfg), float(dt), C.byref(ed), s)
    return update, adj, st


def buildScheme(schemeName):
    """schemes/builder.py:11-36, CRKSPH branch (:22-25)."""
    if (isinstance(schemeName, str) and schemeName == 'crkSPH') or \
            (isinstance(schemeName, CompressibleSPHScheme) and schemeName == CompressibleSPHScheme.CRKSPH):
        return (CompSPHSystem, CompSPHState, CRKSPHConfig, CompressibleSystemUpdate, crkSPH_step,
                crkSPHConfigToDict, dictToCRKSPHConfig)
    raise ValueError(f"Scheme {schemeName} not recognized.")
