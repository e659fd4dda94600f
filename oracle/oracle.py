"""Python harness of the CPU oracle (test infrastructure only).

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg.  Wraps oracle/liboracle.so (built by oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from warpsph_amd.abi import CrkConfig, CrkGrid

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FIELDS = ["x", "v", "h", "m", "rho", "u", "s", "P", "c", "div", "a0", "alpha",
          "V", "A", "B", "gA", "gB", "J", "divcd", "vsig", "Rcd", "psiH",
          "dvdt", "dudt", "drhodt", "dEdt"]
FIELD_ID = {k: i for i, k in enumerate(FIELDS)}
# components per particle (d = dimension)
_COMPS = {"x": 1, "v": 1, "B": 1, "gA": 1, "dvdt": 1, "gB": 2, "J": 2}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_kernel_constants.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_kernel_eval.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P]
        L.orc_owen_lut.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P]
        L.orc_lut_interp.argtypes = [C.c_int, P, P, C.c_int, P, P]
        L.orc_grid_plan.argtypes = [C.POINTER(CrkConfig), C.c_float, P, P, C.POINTER(CrkGrid)]
        L.orc_morton.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_morton.restype = C.c_uint32
        L.orc_sim_create.argtypes = [C.POINTER(CrkConfig), C.c_int, C.c_int]
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_destroy.argtypes = [P]
        L.orc_sim_destroy.restype = None
        L.orc_sim_set.argtypes = [P, C.c_int, P]
        L.orc_sim_get.argtypes = [P, C.c_int, P]
        L.orc_sim_evaluate.argtypes = [P, C.c_double]
        L.orc_sim_evaluate_phase.argtypes = [P, C.c_int, C.c_double]
        L.orc_sim_step.argtypes = [P, C.c_double, C.c_int]
        L.orc_sim_edges.argtypes = [P, C.c_int]
        L.orc_sim_edges.restype = C.c_int64
        L.orc_sim_adjacency.argtypes = [P, C.c_int, C.POINTER(CrkGrid), P, P, P, P, P, P]
        L.orc_sim_edge_field.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_sim_cfl_dt.argtypes = [P, C.c_double]
        L.orc_sim_cfl_dt.restype = C.c_double
        L.orc_sim_stats.argtypes = [P, P]
        L.orc_num_threads.restype = C.c_int
        L.orc_sim_build_adjacency.argtypes = [P]
        L.orc_limiter_phi.argtypes = [C.POINTER(CrkConfig), P, C.c_double, C.c_double, P, P, P]
        L.orc_energy_fraction.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_energy_fraction.restype = C.c_double
        L.orc_owen_new_support.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_owen_new_support.restype = C.c_double
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def kernel_constants(kernel: int, dim: int):
    c, s = C.c_double(), C.c_double()
    assert lib().orc_kernel_constants(kernel, dim, C.byref(c), C.byref(s)) == 0
    return c.value, s.value


def kernel_xi(kernel: int, dim: int) -> float:
    """xi of the CFL step and the Cullen-Dehnen time."""
    x = C.c_double()
    L = lib()
    L.orc_kernel_xi.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    assert L.orc_kernel_xi(kernel, dim, C.byref(x)) == 0
    return x.value


def kernel_eval(kernel: int, dim: int, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    k = np.empty_like(q)
    dk = np.empty_like(q)
    assert lib().orc_kernel_eval(kernel, dim, q.size, _p(q), _p(k), _p(dk)) == 0
    return k, dk


def owen_lut(kernel: int, dim: int, nlut=4096, nmin=2.0, nmax=6.0):
    nh = np.empty(nlut)
    ps = np.empty(nlut)
    assert lib().orc_owen_lut(kernel, dim, nlut, nmin, nmax, _p(nh), _p(ps)) == 0
    return nh, ps


def lut_interp(xs, ys, x):
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    assert lib().orc_lut_interp(xs.size, _p(xs), _p(ys), x.size, _p(x), _p(out)) == 0
    return out


def grid_plan(cfg: CrkConfig, hmax: float, lo=(0, 0, 0), hi=(0, 0, 0)):
    g = CrkGrid()
    lo = np.asarray(lo, dtype=np.float32)
    hi = np.asarray(hi, dtype=np.float32)
    st = lib().orc_grid_plan(C.byref(cfg), hmax, _p(lo), _p(hi), C.byref(g))
    if st != 0:
        raise ValueError(f"grid plan failed ({st})")
    return g


def morton(dim, cx, cy=0, cz=0):
    return int(lib().orc_morton(dim, cx, cy, cz))


def limiter_phi(cfg: CrkConfig, dx, hi, hj, Ji, Jj) -> float:
    """phi_ij of the CRK limiter (limiter.py:11-127) for one pair (FP64)."""
    a = [np.ascontiguousarray(np.asarray(t, dtype=np.float64).ravel()) for t in (dx, Ji, Jj)]
    out = C.c_double()
    assert lib().orc_limiter_phi(C.byref(cfg), _p(a[0]), float(hi), float(hj), _p(a[1]), _p(a[2]),
                                 C.byref(out)) == 0
    return out.value


def energy_fraction(dE, si, sj) -> float:
    return float(lib().orc_energy_fraction(float(dE), float(si), float(sj)))


def owen_new_support(nh_target, nh, h) -> float:
    return float(lib().orc_owen_new_support(float(nh_target), float(nh), float(h)))


def num_threads():
    return int(lib().orc_num_threads())


class OracleSim:
    """One CRKSPH simulation held by the oracle (user particle order)."""

    def __init__(self, cfg: CrkConfig, n: int, precision: int = 32):
        self.cfg = cfg
        self.n = int(n)
        self.dim = int(cfg.domain.dim)
        self.precision = precision
        self._destroy = lib().orc_sim_destroy
        self._h = lib().orc_sim_create(C.byref(cfg), precision, self.n)
        if not self._h:
            raise ValueError("orc_sim_create failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def build_adjacency(self):
        st = lib().orc_sim_build_adjacency(self._h)
        if st != 0:
            raise ValueError(f"oracle adjacency build failed ({st})")

    def _shape(self, name):
        k = _COMPS.get(name, 0)
        if k == 0:
            return (self.n,)
        if k == 1:
            return (self.n, self.dim)
        return (self.n, self.dim, self.dim)

    def set(self, name, value):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), self._shape(name)))
        st = lib().orc_sim_set(self._h, FIELD_ID[name], _p(a))
        assert st == 0, st

    def get(self, name):
        out = np.empty(self._shape(name), dtype=np.float64)
        st = lib().orc_sim_get(self._h, FIELD_ID[name], _p(out))
        assert st == 0, st
        return out

    def load_state(self, **fields):
        for k, v in fields.items():
            self.set(k, v)

    def evaluate(self, dt: float):
        st = lib().orc_sim_evaluate(self._h, float(dt))
        if st != 0:
            raise ValueError(f"oracle evaluate failed ({st})")

    def evaluate_phase(self, phase: int, dt: float):
        """One phase of evaluate (0: support, adjacency, CRK factors; 1: EOS
        to accel/dudt; 2: energy partition), for the boundary hooks between."""
        st = lib().orc_sim_evaluate_phase(self._h, int(phase), float(dt))
        if st != 0:
            raise ValueError(f"oracle evaluate phase {phase} failed ({st})")

    def step(self, dt: float, nsteps: int = 1):
        st = lib().orc_sim_step(self._h, float(dt), int(nsteps))
        if st != 0:
            raise ValueError(f"oracle step failed ({st})")

    def cfl_dt(self, prev: float = -1.0) -> float:
        return float(lib().orc_sim_cfl_dt(self._h, float(prev)))

    def stats(self):
        out = np.zeros(2, dtype=np.int64)
        lib().orc_sim_stats(self._h, _p(out))
        return {"builds": int(out[0]), "singular": int(out[1])}

    def adjacency(self, stage: int = -1):
        """Adjacency of a stage (-1: current).  Rows in user order."""
        E = int(lib().orc_sim_edges(self._h, stage))
        if E < 0:
            raise ValueError("no adjacency")
        g = CrkGrid()
        assert lib().orc_sim_adjacency(self._h, stage, C.byref(g), None, None, None, None, None, None) == 0
        n = self.n
        keys = np.empty(n, np.uint32)
        perm = np.empty(n, np.int32)
        cs = np.empty(g.table_size, np.int32)
        ce = np.empty(g.table_size, np.int32)
        off = np.empty(n + 1, np.int64)
        nbr = np.empty(E, np.int32)
        assert lib().orc_sim_adjacency(self._h, stage, C.byref(g), _p(keys), _p(perm), _p(cs), _p(ce),
                                       _p(off), _p(nbr)) == 0
        return {"grid": g, "keys": keys, "perm": perm, "cell_start": cs, "cell_end": ce,
                "offsets": off, "nbr": nbr, "n_edges": E}

    def edge_field(self, stage: int, name: str):
        E = int(lib().orc_sim_edges(self._h, stage))
        comps = self.dim if name in ("ap", "av") else 1
        out = np.empty((E, comps) if comps > 1 else (E,), np.float64)
        fid = {"ap": 0, "av": 1, "f": 2}[name]
        assert lib().orc_sim_edge_field(self._h, stage, fid, _p(out)) == 0
        return out


STATE_MAP = {"x": "positions", "v": "velocities", "h": "supports", "m": "masses", "rho": "densities",
             "u": "internalEnergies", "s": "entropies", "P": "pressures", "c": "soundspeeds",
             "div": "divergence", "a0": "alpha0s", "alpha": "alphas"}


def from_system(system, config, scheme, precision: int = 32) -> OracleSim:
    """Oracle simulation holding a copy of a CompSPHSystem's state."""
    from warpsph_amd.configurations import toCrkConfig
    cfg = toCrkConfig(config, scheme)
    st = system.state
    sim = OracleSim(cfg, st.positions.shape[0], precision)
    for k, name in STATE_MAP.items():
        t = getattr(st, name)
        if t is not None:
            sim.set(k, t.detach().cpu().double().numpy())
    return sim
