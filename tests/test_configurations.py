"""buildScheme and the configuration objects (no GPU).

Defaults are the values the reference's dataclasses state:
configurations/crkSPH.py:17-68, compressibleConfig.py:23-40,
moduleConfigurations/viscositySwitchParameters.py:12-24."""
import math

import pytest


def test_build_scheme_returns_the_reference_tuple():
    from warpsph_amd import buildScheme
    from warpsph_amd.configurations import CRKSPHConfig, crkSPHConfigToDict, dictToCRKSPHConfig
    from warpsph_amd.enumTypes import CompressibleSPHScheme
    from warpsph_amd.scheme import crkSPH_step
    from warpsph_amd.systems import CompressibleSystemUpdate, CompSPHState, CompSPHSystem
    expect = (CompSPHSystem, CompSPHState, CRKSPHConfig, CompressibleSystemUpdate, crkSPH_step, crkSPHConfigToDict,
              dictToCRKSPHConfig)
    assert buildScheme(CompressibleSPHScheme.CRKSPH) == expect
    assert buildScheme("crkSPH") == expect
    with pytest.raises(ValueError):
        buildScheme("not-a-scheme")


def test_crksph_defaults():
    from warpsph_amd.configurations import CRKSPHConfig
    from warpsph_amd.enumTypes import AdaptiveSupportScheme, EnergyScheme, ViscositySwitch
    c = CRKSPHConfig()
    assert c.gamma == 1.4 and c.backgroundPressure == 0.0 and c.rho0 == 1.0
    assert c.adaptiveSupportScheme == AdaptiveSupportScheme.Monaghan
    assert c.adaptiveSupportIterations == 1 and c.adaptiveSupportThreshold == 1e-3
    assert c.energyScheme == EnergyScheme.CRK and c.compatibleEnergy is True
    assert c.schemeName == "CRKSPH" and c.boundaryConditions == []
    v = c.crkViscosityParams
    assert v.eta_fold == pytest.approx(0.2) and v.eta_crit == pytest.approx(0.3333333)
    assert v.enableCRKLimiter and v.enableVanLeerLimiter and not v.forceVanLeerOff and not v.forceVanLeerOn
    d = c.diffusionParams
    assert (d.c_s, d.C_l, d.C_q, d.Cu_l, d.Cu_q, d.K) == (1, 1, 1, 1, 1, 1.0)
    assert d.thermalConductivity == 0.5 and not d.scaleBeta and d.monaghanSwitch and d.correctXi
    s = c.viscositySwitchParams
    assert s.scheme == ViscositySwitch.NoneSwitch and s.limitXi is True  # the later of its two declarations
    assert (s.alpha_min, s.alpha_max, s.beta_c, s.beta_d, s.beta_xi) == (0.02, 2.0, 0.7, 0.05, 2.0)


def test_crksph_config_dict_round_trip():
    from warpsph_amd.configurations import CRKSPHConfig, crkSPHConfigToDict, dictToCRKSPHConfig
    from warpsph_amd.enumTypes import AdaptiveSupportScheme, ViscositySwitch
    c = CRKSPHConfig(gamma=5.0 / 3.0, adaptiveSupportScheme=AdaptiveSupportScheme.Owen, adaptiveSupportIterations=2)
    c.viscositySwitchParams.scheme = ViscositySwitch.CullenDehnen2010
    c.viscositySwitchParams.alpha_max = 1.5
    c.crkViscosityParams.eta_fold = 0.25
    c.crkViscosityParams.forceVanLeerOn = True
    c.compatibleEnergy = False
    d = crkSPHConfigToDict(c)
    assert d["energyScheme"] == "CRK" and d["adaptiveSupportScheme"] == "Owen"
    r = dictToCRKSPHConfig(d)
    assert crkSPHConfigToDict(r) == d
    assert r.gamma == c.gamma and r.adaptiveSupportScheme == AdaptiveSupportScheme.Owen
    assert r.viscositySwitchParams.scheme == ViscositySwitch.CullenDehnen2010
    assert r.viscositySwitchParams.alpha_max == 1.5 and r.crkViscosityParams.eta_fold == 0.25
    assert r.crkViscosityParams.forceVanLeerOn and r.compatibleEnergy is False


def test_simulation_config_kernel_support():
    """buildConfig: the support radius holding n_H neighbours (utils/support.py:18-44)."""
    from warpsph_amd import sample
    from warpsph_amd.configurations import buildConfig
    from warpsph_amd.utils import volumeToSupportHelper
    dom = sample._domain([0.0, 0.0], [1.0, 1.0], [True, True], "cpu")
    import torch
    config, _ = buildConfig(device=torch.device("cpu"), domain=dom, dim=2, targetNeighbors=32.0)
    assert config.targetNeighbors == 32.0 and config.verletScale == pytest.approx(math.sqrt(2.0))
    dx = 1.0 / 64
    h = float(volumeToSupportHelper(dx ** 2, 32.0, 2))
    assert math.pi * h * h / dx ** 2 == pytest.approx(32.0, rel=1e-6)
