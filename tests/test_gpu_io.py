"""Export after a GPU step and import (src/warpSPH/io.py:63-270), called the
way examples/compressible/01-sod-shock-tube-1d.py:236-248 calls them: the
adjacency comes back in the reference's user-index form with the CSR the
passes read, the stages with their pair data, and a step from the imported
system equals the step from the original one bit for bit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_export_import_after_a_step(tmp_path):
    from warpsph_amd import io as wio
    from warpsph_amd import sample
    from warpsph_amd.enumTypes import CompressibleSPHScheme
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.runtime import runtime
    from warpsph_amd.scheme import crkSPH_step
    from warpsph_amd.systems import AdjacencyList
    print("[gpu-case] export / import after a step", flush=True)
    rt = runtime("cuda:0")
    keep = rt.tile_cap
    rt.tile_cap = None  # an imported list carries no LDS tiles; compare like with like
    try:
        scheme = CompressibleSPHScheme.CRKSPH
        system, config, schemeConfig = sample.shear_layer_2d(nx=24, device="cuda")
        integrator = getIntegrator()
        dt = sample.cfl_dt(system.state, config, schemeConfig)
        res = integrator.function(system, crkSPH_step, dt, config, schemeConfig)
        s = res.state
        ke = 0.5 * (torch.linalg.norm(s.state.velocities, dim=-1) ** 2 * s.state.masses).sum()
        path = wio.exportSimulationSystem(str(tmp_path), f'state_{1:04d}', scheme, s, exportAdjacency=False,
                                          stages=res.stages, exportStagesAdjacency=True,
                                          extraData={'kineticEnergy': ke, 'frame_num': 1})
        back, stages, schemeEnum, extra = wio.importSimulationSystem(path, 'cuda', torch.float32)
        assert schemeEnum is scheme and abs(back.t - s.t) < 1e-12
        assert torch.equal(extra['kineticEnergy'], ke) and extra['frame_num'] == 1
        for k in ('positions', 'velocities', 'internalEnergies', 'supports', 'alphas', 'divergence'):
            assert torch.equal(getattr(back.state, k), getattr(s.state, k)), k
        a0, a1 = s.adjacency, back.adjacency
        assert isinstance(a1, AdjacencyList)
        for k in ('i', 'j', 'numNeighbors', 'edgeOffsets', 'perm', 'row_offsets', 'nbr', 'xb', 'hb'):
            assert torch.equal(getattr(a1, k), getattr(a0, k)), k
        # both stages, with the pair data over each stage's edges
        assert len(stages) == 2
        for st0, st1 in zip(res.stages, stages):
            assert torch.equal(st1.aux[0].j, st0.aux[0].j)
            for x0, x1 in zip(st0.aux[1].pairs.full(), st1.aux[1].pairs.full()):
                assert torch.equal(x0, x1)
            for f in ('dvdt', 'dudt', 'drhodt', 'dEdt'):
                assert torch.equal(getattr(st1.update, f), getattr(st0.update, f)), f
        assert stages[1].aux[1].ap_ij.shape[0] == stages[1].aux[0].j.shape[0]
        # the next step from the imported system (its adjacency reused or
        # rebuilt by the same Verlet test) equals the next step from the original
        r_a = integrator.function(s, crkSPH_step, dt, config, schemeConfig)
        r_b = integrator.function(back, crkSPH_step, dt, config, schemeConfig)
        torch.cuda.synchronize()
        for k in ('positions', 'velocities', 'internalEnergies', 'densities', 'supports', 'alphas'):
            va, vb = getattr(r_a.state.state, k), getattr(r_b.state.state, k)
            assert torch.equal(va, vb), (k, float((va - vb).abs().max()))
        assert np.array_equal(r_a.state.adjacency.j.cpu().numpy(), r_b.state.adjacency.j.cpu().numpy())
    finally:
        rt.tile_cap = keep
