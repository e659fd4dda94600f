"""Host-side contracts that need no GPU: bench.py's command line, refusals of unsupported options."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    line = [x for x in r.stdout.splitlines() if x.startswith("{")][-1]
    assert "one GPU" in json.loads(line)["error"]


def test_boundary_conditions_are_refused_not_ignored():
    import pytest
    from warpsph_amd import sample
    from warpsph_amd.scheme import crkSPH_step
    system, config, scheme = sample.shear_layer_2d(nx=8, device="cpu")
    scheme.boundaryConditions = [object()]
    with pytest.raises(NotImplementedError):
        crkSPH_step(system, 1e-4, config, scheme)
