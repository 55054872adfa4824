"""TEST INFRASTRUCTURE: the CPU oracle for the explicit update path.

ctypes bindings to liboracle.so (built from mmsp_oracle.c by `make -C oracle`)
plus numpy-friendly wrappers. Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py import this package, and only as the checker.
Reference citations for each routine are in mmsp_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

MIRROR, NEUMANN, PERIODIC, PARALLEL, DIRICHLET = 0, 1, 2, 3, 4
CUBE_POW, CUBE_CR = 0, 1
MU_POLY, MU_DW = 0, 1
E_POLY, E_DW, E_GG = 0, 1, 2


class OrcGrid(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("n", C.c_int * 3),
        ("b0", C.c_int * 3),
        ("b1", C.c_int * 3),
        ("dx", C.c_double * 3),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        G = P(OrcGrid)
        dp, fp = P(C.c_double), P(C.c_float)
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_fill_uniform.argtypes = [dp, C.c_int64, C.c_uint64, C.c_double, C.c_double]
        L.orc_check_boundary.restype = C.c_int
        L.orc_check_boundary.argtypes = [C.c_int] * 5
        L.orc_cube_cr.restype = C.c_double
        L.orc_cube_cr.argtypes = [C.c_double]
        L.orc_laplacian_f64.restype = C.c_double
        L.orc_laplacian_f64.argtypes = [G, dp, C.c_int64]
        L.orc_laplacian_f32ref.restype = C.c_float
        L.orc_laplacian_f32ref.argtypes = [G, fp, C.c_int64]
        L.orc_laplacian_f32.restype = C.c_float
        L.orc_laplacian_f32.argtypes = [G, fp, C.c_int64]
        L.orc_ac_steps_f64.argtypes = [G, dp, dp, dp, C.c_int, C.c_int]
        L.orc_ac_steps_f32ref.argtypes = [G, fp, fp, dp, C.c_int, C.c_int]
        L.orc_ac_steps_f32.argtypes = [G, fp, fp, dp, C.c_int]
        L.orc_ch_steps_f64.argtypes = [G, dp, dp, dp, dp, C.c_int, C.c_int, C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        u32p = P(C.c_uint32)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_ln_unit.restype = C.c_float
        L.orc_ln_unit.argtypes = [C.c_float]
        L.orc_cos_sin_2pi.argtypes = [C.c_float, P(C.c_float), P(C.c_float)]
        L.orc_normal.restype = C.c_float
        L.orc_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_ac_noise_steps_f64.argtypes = [G, dp, dp, dp, C.c_int, C.c_uint64, C.c_int]
        L.orc_ac_noise_steps_f32ref.argtypes = [G, fp, fp, dp, C.c_int, C.c_uint64, C.c_int]
        L.orc_ac_noise_steps_f32.argtypes = [G, fp, fp, dp, C.c_int, C.c_uint64]
        L.orc_ch_noise_steps_f64.argtypes = [G, dp, dp, dp, dp, C.c_int, C.c_uint64, C.c_int]
        L.orc_ch_noise_steps_f32ref.argtypes = [G, fp, fp, fp, dp, C.c_int, C.c_uint64, C.c_int]
        L.orc_ch_noise_steps_f32.argtypes = [G, fp, fp, fp, dp, C.c_int, C.c_uint64]
        L.orc_zener_steps_f64.argtypes = [G, C.c_int, dp, dp, C.c_double, C.c_int, C.c_int]
        L.orc_zener_steps_f32.argtypes = [G, C.c_int, fp, fp, C.c_double, C.c_int]
        L.orc_ostwald_steps_f64.argtypes = [G, C.c_int, dp, dp, dp, dp, C.c_int, C.c_int]
        L.orc_ostwald_steps_f32.argtypes = [G, C.c_int, fp, fp, fp, dp, C.c_int]
        L.orc_gga_steps_f64.argtypes = [G, C.c_int, dp, dp, dp, C.c_int]
        L.orc_gga_steps_f32.argtypes = [G, C.c_int, fp, fp, dp, C.c_int]
        L.orc_ch_steps_f32ref.argtypes = [G, fp, fp, fp, dp, C.c_int, C.c_int, C.c_int]
        L.orc_ch_steps_f32.argtypes = [G, fp, fp, fp, dp, C.c_int, C.c_int]
        L.orc_gg_steps_f64.argtypes = [G, C.c_int, dp, dp, C.c_double, C.c_int, C.c_int]
        L.orc_gg_steps_f32ref.argtypes = [G, C.c_int, fp, fp, C.c_double, C.c_int, C.c_int]
        L.orc_gg_steps_f32.argtypes = [G, C.c_int, fp, fp, C.c_double, C.c_int]
        L.orc_mass_f64.argtypes = [dp, C.c_int64, C.c_int, dp]
        L.orc_mass_f32.argtypes = [fp, C.c_int64, C.c_int, dp]
        L.orc_energy_f64.restype = C.c_double
        L.orc_energy_f64.argtypes = [G, C.c_int, dp, C.c_int, dp]
        L.orc_energy_f32.restype = C.c_double
        L.orc_energy_f32.argtypes = [G, C.c_int, fp, C.c_int, dp]
        _lib = L
    return _lib


def grid_desc(shape, bc=None, dx=None) -> OrcGrid:
    dim = len(shape)
    g = OrcGrid()
    g.dim = dim
    for d in range(3):
        g.n[d] = int(shape[d]) if d < dim else 1
        lo, hi = (bc[d] if bc is not None and d < dim else (PERIODIC, PERIODIC))
        g.b0[d], g.b1[d] = lo, hi
        g.dx[d] = float(dx[d]) if dx is not None and d < dim else 1.0
    return g


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _params(p):
    return np.ascontiguousarray(np.asarray(p, dtype=np.float64))


# ---- seeded inputs ---------------------------------------------------------
def uniform(n: int, seed: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty(int(n), dtype=np.float64)
    lib().orc_fill_uniform(_p(out, C.c_double), int(n), seed, lo, hi)
    return out


def check_boundary(x, x0, x1, b0, b1) -> int:
    return lib().orc_check_boundary(x, x0, x1, b0, b1)


def cube_cr(x: float) -> float:
    return lib().orc_cube_cr(x)


def laplacian(a: np.ndarray, i: int, bc=None, dx=None, form="f64") -> float:
    g = grid_desc(a.shape, bc, dx)
    flat = np.ascontiguousarray(a).reshape(-1)
    if form == "f64":
        return lib().orc_laplacian_f64(C.byref(g), _p(flat, C.c_double), i)
    if form == "f32ref":
        return lib().orc_laplacian_f32ref(C.byref(g), _p(flat, C.c_float), i)
    return lib().orc_laplacian_f32(C.byref(g), _p(flat, C.c_float), i)


# ---- steppers: return a new array after `steps` explicit steps ---------------
def ac(a, params, steps, form="f64", cube=CUBE_CR, bc=None, dx=None):
    g = grid_desc(a.shape, bc, dx)
    p = _params(params)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y = np.empty_like(x)
        lib().orc_ac_steps_f64(C.byref(g), _p(x, C.c_double), _p(y, C.c_double), _p(p, C.c_double), steps, cube)
    elif form == "f32ref":
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_ac_steps_f32ref(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(p, C.c_double), steps, cube)
    else:
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_ac_steps_f32(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(p, C.c_double), steps)
    return x


def ch(a, params, steps, mu=MU_DW, form="f64", cube=CUBE_CR, bc=None, dx=None):
    g = grid_desc(a.shape, bc, dx)
    p = _params(params)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y, m = np.empty_like(x), np.empty_like(x)
        lib().orc_ch_steps_f64(C.byref(g), _p(x, C.c_double), _p(y, C.c_double), _p(m, C.c_double),
                               _p(p, C.c_double), mu, steps, cube)
    elif form == "f32ref":
        x = np.array(a, dtype=np.float32, order="C")
        y, m = np.empty_like(x), np.empty_like(x)
        lib().orc_ch_steps_f32ref(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(m, C.c_float),
                                  _p(p, C.c_double), mu, steps, cube)
    else:
        x = np.array(a, dtype=np.float32, order="C")
        y, m = np.empty_like(x), np.empty_like(x)
        lib().orc_ch_steps_f32(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(m, C.c_float),
                               _p(p, C.c_double), mu, steps)
    return x


# ---- stochastic forms: p = {r, u, K, M, dt, sigma, seed} ----------------------
def set_threads(n: int) -> int:
    """Threads for the node loops of the scalar steppers (Allen-Cahn,
    Cahn-Hilliard and their stochastic forms) on grids of 2^16 nodes or more;
    n <= 0 keeps the current count. Returns the count in use. The results do
    not depend on it."""
    return int(lib().orc_set_threads(int(n)))


def philox4x32_10(counter, key):
    c = np.ascontiguousarray(np.asarray(counter, dtype=np.uint32))
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return [int(v) for v in out]


def ln_unit(u: float) -> float:
    return lib().orc_ln_unit(u)


def cos_sin_2pi(t: float):
    c, s = C.c_float(), C.c_float()
    lib().orc_cos_sin_2pi(t, C.byref(c), C.byref(s))
    return c.value, s.value


def normal(seed: int, step: int, x: int, r: int = 0) -> float:
    """The deviate of the node with axis-0 index x and index r within its axis-0 plane."""
    return lib().orc_normal(seed, step, x, r)


def ac_noise(a, params, steps, form="f64", step0=0, cube=CUBE_CR, bc=None, dx=None):
    """Model A from the state `a`; the noise of step k is drawn for step number step0 + k."""
    g = grid_desc(a.shape, bc, dx)
    p = _params(params)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y = np.empty_like(x)
        lib().orc_ac_noise_steps_f64(C.byref(g), _p(x, C.c_double), _p(y, C.c_double), _p(p, C.c_double), steps,
                                     step0, cube)
    elif form == "f32ref":
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_ac_noise_steps_f32ref(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(p, C.c_double), steps,
                                        step0, cube)
    else:
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_ac_noise_steps_f32(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(p, C.c_double), steps,
                                     step0)
    return x


def ch_noise(a, params, steps, form="f64", step0=0, cube=CUBE_CR, bc=None, dx=None):
    """Model B from the state `a` (spinodal: r = u = K = 1, M = 1/(2 dV))."""
    g = grid_desc(a.shape, bc, dx)
    p = _params(params)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y, m = np.empty_like(x), np.empty_like(x)
        lib().orc_ch_noise_steps_f64(C.byref(g), _p(x, C.c_double), _p(y, C.c_double), _p(m, C.c_double),
                                     _p(p, C.c_double), steps, step0, cube)
    elif form == "f32ref":
        x = np.array(a, dtype=np.float32, order="C")
        y, m = np.empty_like(x), np.empty_like(x)
        lib().orc_ch_noise_steps_f32ref(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(m, C.c_float),
                                        _p(p, C.c_double), steps, step0, cube)
    else:
        x = np.array(a, dtype=np.float32, order="C")
        y, m = np.empty_like(x), np.empty_like(x)
        lib().orc_ch_noise_steps_f32(C.byref(g), _p(x, C.c_float), _p(y, C.c_float), _p(m, C.c_float),
                                     _p(p, C.c_double), steps, step0)
    return x


def zener(a, dt, steps, form="f64", cube=CUBE_CR, bc=None, dx=None):
    """Zener pinning; `a` node-major (extents + (Q,)), field 0 the particles."""
    shape, q = a.shape[:-1], a.shape[-1]
    g = grid_desc(shape, bc, dx)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y = np.empty_like(x)
        lib().orc_zener_steps_f64(C.byref(g), q, _p(x, C.c_double), _p(y, C.c_double), dt, steps, cube)
    elif form == "f32":
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_zener_steps_f32(C.byref(g), q, _p(x, C.c_float), _p(y, C.c_float), dt, steps)
    else:
        raise ValueError("Zener pinning: forms f64 and f32")
    return x


def ostwald(a, params, steps, form="f64", cube=CUBE_CR, bc=None, dx=None):
    """Ostwald ripening; `a` node-major, field 0 the concentration;
    params = {dt, L, D, Calpha, Cbeta, A, delta, epsilon, kappa}."""
    shape, q = a.shape[:-1], a.shape[-1]
    g = grid_desc(shape, bc, dx)
    p = _params(params)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y, w = np.empty_like(x), np.empty(int(np.prod(shape)), dtype=np.float64)
        lib().orc_ostwald_steps_f64(C.byref(g), q, _p(x, C.c_double), _p(y, C.c_double), _p(w, C.c_double),
                                    _p(p, C.c_double), steps, cube)
    elif form == "f32":
        x = np.array(a, dtype=np.float32, order="C")
        y, w = np.empty_like(x), np.empty(int(np.prod(shape)), dtype=np.float32)
        lib().orc_ostwald_steps_f32(C.byref(g), q, _p(x, C.c_float), _p(y, C.c_float), _p(w, C.c_float),
                                    _p(p, C.c_double), steps)
    else:
        raise ValueError("Ostwald ripening: forms f64 and f32")
    return x


def gg_aniso(a, params, steps, form="f64", bc=None, dx=None):
    """Multi-phase grain growth; `a` node-major; params = {dt, width, gamma, mu}."""
    shape, q = a.shape[:-1], a.shape[-1]
    g = grid_desc(shape, bc, dx)
    p = _params(params)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y = np.empty_like(x)
        lib().orc_gga_steps_f64(C.byref(g), q, _p(x, C.c_double), _p(y, C.c_double), _p(p, C.c_double), steps)
    elif form == "f32":
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_gga_steps_f32(C.byref(g), q, _p(x, C.c_float), _p(y, C.c_float), _p(p, C.c_double), steps)
    else:
        raise ValueError("multi-phase grain growth: forms f64 and f32")
    return x


def gg(a, dt, steps, form="f64", cube=CUBE_CR, bc=None, dx=None):
    """`a` is node-major: shape extents + (Q,)."""
    shape = a.shape[:-1]
    q = a.shape[-1]
    g = grid_desc(shape, bc, dx)
    if form == "f64":
        x = np.array(a, dtype=np.float64, order="C")
        y = np.empty_like(x)
        lib().orc_gg_steps_f64(C.byref(g), q, _p(x, C.c_double), _p(y, C.c_double), dt, steps, cube)
    elif form == "f32ref":
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_gg_steps_f32ref(C.byref(g), q, _p(x, C.c_float), _p(y, C.c_float), dt, steps, cube)
    else:
        x = np.array(a, dtype=np.float32, order="C")
        y = np.empty_like(x)
        lib().orc_gg_steps_f32(C.byref(g), q, _p(x, C.c_float), _p(y, C.c_float), dt, steps)
    return x


# ---- reductions --------------------------------------------------------------
def mass(a, vector=False) -> np.ndarray:
    q = a.shape[-1] if vector else 1
    n = a.size // q
    out = np.zeros(q, dtype=np.float64)
    x = np.ascontiguousarray(a)
    if x.dtype == np.float64:
        lib().orc_mass_f64(_p(x, C.c_double), n, q, _p(out, C.c_double))
    else:
        x = x.astype(np.float32, copy=False)
        lib().orc_mass_f32(_p(x, C.c_float), n, q, _p(out, C.c_double))
    return out


def energy(a, kind, params, vector=False, bc=None, dx=None) -> float:
    shape = a.shape[:-1] if vector else a.shape
    q = a.shape[-1] if vector else 1
    g = grid_desc(shape, bc, dx)
    p = _params(params)
    x = np.ascontiguousarray(a)
    if x.dtype == np.float64:
        return lib().orc_energy_f64(C.byref(g), q, _p(x, C.c_double), kind, _p(p, C.c_double))
    x = x.astype(np.float32, copy=False)
    return lib().orc_energy_f32(C.byref(g), q, _p(x, C.c_float), kind, _p(p, C.c_double))


# ---- the reference itself (compiled here from its own sources) ---------------
def ref_available() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, b)) for b in ("ref_ac", "ref_ch", "ref_gg"))


def ref_run(model: str, inp: str, out: str, steps: int, variant="stock", params=(), timeout=None):
    """Run oracle/_ref/ref_<model>; returns update() seconds reported by it."""
    exe = os.path.join(REF_DIR, "ref_" + model)
    cmd = [exe, "run", inp, out, str(steps), variant] + ["%.17g" % v for v in params]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("reference harness failed: %s\n%s" % (" ".join(cmd), r.stderr))
    for line in r.stdout.splitlines():
        if line.startswith("REFTIME"):
            return float(line.split()[1])
    raise RuntimeError("reference harness printed no REFTIME line")
