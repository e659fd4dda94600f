"""crk_conservation_noalloc (fixed device scratch, no allocation) against
crk_conservation, through the C-ABI: two calls in a row on one context, the
second with fewer particles, each bit-identical to the allocating entry point."""
import ctypes as C

import numpy as np
import pytest
import torch

from warpsph_amd import abi
from warpsph_amd.runtime import runtime


def _inputs(n, dim, seed):
    g = torch.Generator().manual_seed(seed)
    m = (0.5 + torch.rand(n, generator=g)).float()
    v = (torch.rand(n, dim, generator=g) - 0.5).float()
    u = (1.0 + torch.rand(n, generator=g)).float()
    dev = torch.device("cuda:0")
    return m.to(dev), v.contiguous().to(dev), u.to(dev)


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_noalloc_conservation_matches_allocating_entry_point(dim):
    rt = runtime("cuda:0")
    s = rt.stream()
    out_dev = torch.full((dim + 3,), float("nan"), dtype=torch.float64, device="cuda:0")
    for call, n in enumerate((300_001, 4_099)):
        print(f"noalloc conservation: dim {dim}, call {call}, n {n}", flush=True)
        m, v, u = _inputs(n, dim, seed=17 + call)
        ref = np.empty(dim + 3, dtype=np.float64)
        abi.call("crk_conservation", rt.ctx, dim, n, abi.ptr(m), abi.ptr(v), abi.ptr(u),
                 C.c_void_p(ref.ctypes.data), s)
        abi.call("crk_conservation_noalloc", rt.ctx, dim, n, abi.ptr(m), abi.ptr(v), abi.ptr(u),
                 abi.ptr(out_dev), s)
        torch.cuda.synchronize()
        got = out_dev.cpu().numpy()
        assert np.array_equal(got, ref), (call, got, ref)
        # and the sums themselves against torch's FP64 sums
        md, vd, ud = m.double(), v.double(), u.double()
        want = [md.sum().item()] + (md[:, None] * vd).sum(0).tolist() + \
               [(md * ud).sum().item(), (0.5 * md * (vd * vd).sum(1)).sum().item()]
        assert np.allclose(got, want, rtol=1e-12, atol=0.0)


@pytest.mark.gpu
def test_noalloc_conservation_rejects_oversized_input():
    rt = runtime("cuda:0")
    t = torch.zeros(1, device="cuda:0")
    out = torch.zeros(5, dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError, match="capacity"):
        abi.call("crk_conservation_noalloc", rt.ctx, 2, (1 << 23) + 1, abi.ptr(t), abi.ptr(t), abi.ptr(t),
                 abi.ptr(out), rt.stream())
