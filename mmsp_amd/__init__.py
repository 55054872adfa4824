"""mmsp_amd -- MI355X-native explicit finite-difference update path of MMSP.

Python bindings (ctypes) over the C ABI in include/mmsp_amd.h, implemented by
libmmsp_amd.so (HIP kernels for gfx950 + host C++). There is no CPU fallback:
importing this package raises if the library has not been built, and every
compute call goes to the GPU.

    import mmsp_amd as mx
    g = mx.Grid((512, 512), precision="f64")
    g.upload(phi0)
    g.step("ac", [1, 1, 1, 1, 0.01], 1000)   # allen-cahn.cpp:58-71
    phi = g.download()
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmmsp_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mmsp_amd.h")

MIRROR, NEUMANN, PERIODIC, PARALLEL, DIRICHLET = 0, 1, 2, 3, 4
F64, F32, F32REF = 0, 1, 2
AC, CH, CH_DW, GG, AC_NOISE, CH_NOISE, ZENER, OSTWALD, GG_ANISO = 0, 1, 2, 3, 4, 5, 6, 7, 8
NODE_MAJOR, FIELD_MAJOR = 0, 1
MASS, ENERGY = 0, 1

PRECISIONS = {"f64": F64, "f32": F32, "f32ref": F32REF}
MODELS = {"ac": AC, "ch": CH, "ch_dw": CH_DW, "gg": GG, "ac_noise": AC_NOISE, "ch_noise": CH_NOISE, "zener": ZENER,
          "ostwald": OSTWALD, "gg_aniso": GG_ANISO}
NPARAMS = {AC: 5, CH: 5, CH_DW: 3, GG: 1, AC_NOISE: 7, CH_NOISE: 7, ZENER: 1, OSTWALD: 9, GG_ANISO: 4}


class MmspxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_strerror(code), code, msg))
        self.code = code


class Desc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("g0", C.c_int * 3),
        ("n", C.c_int * 3),
        ("fields", C.c_int),
        ("vector", C.c_int),
        ("precision", C.c_int),
        ("b0", C.c_int * 3),
        ("b1", C.c_int * 3),
        ("dx", C.c_double * 3),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_size_t),
        ("max_slots", C.c_int),
        ("chunk_steps", C.c_int),
        ("model", C.c_int),
        ("precision", C.c_int),
        ("fp64_above_steps", C.c_int),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("id", C.c_int64),
        ("status", C.c_int),
        ("steps_done", C.c_int),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
        ("mass0", C.c_double),
        ("mass1", C.c_double),
        ("precision", C.c_int),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "mmsp_amd: %s is missing; build it with `python -m mmsp_amd.build` "
            "(there is no CPU fallback)" % LIB_PATH
        )
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp, dp, i64p = C.c_void_p, P(C.c_double), P(C.c_int64)
    sig = {
        "mmspx_strerror": (C.c_char_p, [C.c_int]),
        "mmspx_last_error": (C.c_char_p, []),
        "mmspx_abi_version": (C.c_int, []),
        "mmspx_create": (C.c_int, [P(Desc), P(vp)]),
        "mmspx_destroy": (None, [vp]),
        "mmspx_get_desc": (C.c_int, [vp, P(Desc)]),
        "mmspx_footprint": (C.c_size_t, [P(Desc), C.c_int]),
        "mmspx_upload": (C.c_int, [vp, vp, C.c_int]),
        "mmspx_download": (C.c_int, [vp, vp, C.c_int]),
        "mmspx_step": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_int]),
        "mmspx_step_noalloc": (C.c_int, [P(Desc), vp, vp, C.c_int, dp, C.c_int, C.c_int, C.c_int64, vp,
                                         P(C.c_int)]),
        "mmspx_sync": (C.c_int, [vp]),
        "mmspx_step_timed": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_int, dp, i64p]),
        "mmspx_reduce": (C.c_int, [vp, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int]),
        "mmspx_write_dat": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "mmspx_digest": (C.c_int, [vp, P(C.c_uint64)]),
        "mmspx_host_digest": (C.c_int, [vp, C.c_int64, C.c_int, P(C.c_uint64)]),
        "mmspx_read_dat": (C.c_int, [C.c_char_p, C.c_int, P(vp)]),
        "mmspx_peek_dat": (C.c_int, [C.c_char_p, P(Desc)]),
        "mmspx_checkpoint_begin": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, P(vp)]),
        "mmspx_checkpoint_wait": (C.c_int, [vp, dp]),
        "mmspx_write_dat_host": (C.c_int, [P(Desc), vp, C.c_char_p, C.c_int, C.c_int]),
        "mmspx_fill_uniform": (None, [dp, C.c_int64, C.c_uint64, C.c_double, C.c_double]),
        "mmspx_set_variant": (C.c_int, [vp, C.c_int]),
        "mmspx_set_tile": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "mmspx_get_tile": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "mmspx_set_tile_options": (C.c_int, [vp, C.c_int]),
        "mmspx_autotune": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_int, C.c_char_p, C.c_size_t]),
        "mmspx_kernel_time": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_int, dp, dp]),
        "mmspx_copy_bandwidth": (C.c_int, [C.c_size_t, C.c_int, dp]),
        "mmspx_batch_create": (C.c_int, [P(Desc), C.c_int, C.c_int, P(vp)]),
        "mmspx_batch_destroy": (None, [vp]),
        "mmspx_batch_upload": (C.c_int, [vp, C.c_int, vp, C.c_int]),
        "mmspx_batch_download": (C.c_int, [vp, C.c_int, vp, C.c_int]),
        "mmspx_batch_set_params": (C.c_int, [vp, dp, C.c_int]),
        "mmspx_batch_step": (C.c_int, [vp, C.c_int, C.c_int]),
        "mmspx_batch_step_timed": (C.c_int, [vp, C.c_int, C.c_int, dp]),
        "mmspx_batch_reduce": (C.c_int, [vp, C.c_int, C.c_int, dp, C.c_int]),
        "mmspx_batch_autotune": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_size_t]),
        "mmspx_batch_get_tile": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "mmspx_batch_set_tile": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "mmspx_server_create": (C.c_int, [P(ServerConfig), P(vp)]),
        "mmspx_server_destroy": (None, [vp]),
        "mmspx_server_submit": (C.c_int, [vp, P(Desc), dp, dp, C.c_int, C.c_int, i64p]),
        "mmspx_server_cancel": (C.c_int, [vp, C.c_int64]),
        "mmspx_server_wait": (C.c_int, [vp, C.c_int64, C.c_double, P(RequestInfo)]),
        "mmspx_server_fetch": (C.c_int, [vp, C.c_int64, dp, C.c_size_t]),
        "mmspx_server_info": (C.c_int, [vp, C.c_int64, P(RequestInfo)]),
        "mmspx_server_clock": (C.c_double, [vp]),
        "mmspx_server_stats": (C.c_int, [vp, i64p, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def _strerror(code):
    return lib.mmspx_strerror(code).decode()


def check(rc):
    if rc != 0:
        raise MmspxError(rc, lib.mmspx_last_error().decode())
    return rc


def declared_symbols(header: str = HEADER):
    """Function names declared in include/mmsp_amd.h."""
    with open(header) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(mmspx_[a-z_0-9]+)\s*\(", text)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_desc(shape, fields=1, vector=False, precision="f64", bc=None, dx=None, g0=None) -> Desc:
    d = Desc()
    dim = len(shape)
    d.dim = dim
    for k in range(3):
        d.n[k] = int(shape[k]) if k < dim else 1
        d.g0[k] = int(g0[k]) if g0 is not None and k < dim else 0
        lo, hi = bc[k] if bc is not None and k < dim else (PERIODIC, PERIODIC)
        d.b0[k], d.b1[k] = lo, hi
        d.dx[k] = float(dx[k]) if dx is not None and k < dim else 1.0
    d.fields = int(fields)
    d.vector = 1 if vector else 0
    d.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    return d


def _model(m):
    return MODELS[m] if isinstance(m, str) else int(m)


def _params(model, params):
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    return p


def fill_uniform(n, seed, lo, hi):
    out = np.empty(int(n), dtype=np.float64)
    lib.mmspx_fill_uniform(_dp(out), int(n), seed, lo, hi)
    return out


def host_digest(a) -> int:
    """The state digest of a host array (node-major, float64 or float32)."""
    a = np.ascontiguousarray(a)
    if a.dtype not in (np.float64, np.float32):
        raise ValueError("float64 or float32 arrays only")
    out = C.c_uint64()
    check(lib.mmspx_host_digest(a.ctypes.data_as(C.c_void_p), a.size, a.itemsize, C.byref(out)))
    return out.value


def footprint(shape, model, fields=1, vector=False, precision="f64"):
    d = make_desc(shape, fields, vector, precision)
    return lib.mmspx_footprint(C.byref(d), _model(model))


NOALLOC_SCRATCH_BYTES = 32 << 20  # MMSPX_NOALLOC_SCRATCH_BYTES


def step_noalloc(desc: Desc, buf0: int, buf1: int, model, params, steps: int, first_step: int = 0,
                 stream: int = 0) -> int:
    """mmspx_step_noalloc on caller-owned device buffers (addresses as ints,
    the device layout); enqueued on `stream` (a hipStream_t as an int, 0: the
    null stream). Returns 0 or 1: the buffer holding the state after `steps`."""
    m = _model(model)
    p = _params(m, params)
    res = C.c_int(-1)
    check(lib.mmspx_step_noalloc(C.byref(desc), C.c_void_p(buf0), C.c_void_p(buf1), m, _dp(p), len(p), int(steps),
                                 int(first_step), C.c_void_p(stream), C.byref(res)))
    return res.value


class Grid:
    """A device-resident grid<dim,T> (scalar) or grid<dim,vector<T> > (vector=True)."""

    def __init__(self, shape=None, fields=1, vector=False, precision="f64", bc=None, dx=None, g0=None, _handle=None):
        if _handle is not None:
            self._h = _handle
        else:
            d = make_desc(shape, fields, vector, precision, bc, dx, g0)
            h = C.c_void_p()
            check(lib.mmspx_create(C.byref(d), C.byref(h)))
            self._h = h
        d = Desc()
        check(lib.mmspx_get_desc(self._h, C.byref(d)))
        self.desc = d
        self.shape = tuple(d.n[k] for k in range(d.dim))
        self.vector = bool(d.vector)
        self.fields = d.fields if self.vector else 1
        self.dtype = np.float64 if d.precision == F64 else np.float32

    @classmethod
    def from_dat(cls, path, precision=None):
        h = C.c_void_p()
        p = -1 if precision is None else PRECISIONS[precision]
        check(lib.mmspx_read_dat(path.encode(), p, C.byref(h)))
        return cls(_handle=h)

    def close(self):
        if getattr(self, "_h", None):
            lib.mmspx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def host_shape(self):
        return self.shape + ((self.fields,) if self.vector else ())

    def upload(self, a, layout=NODE_MAJOR):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != int(np.prod(self.host_shape)):
            raise ValueError("array size %d does not match grid %s" % (a.size, self.host_shape))
        check(lib.mmspx_upload(self._h, a.ctypes.data_as(C.c_void_p), layout))

    def download(self, layout=NODE_MAJOR):
        if layout == NODE_MAJOR:
            out = np.empty(self.host_shape, dtype=self.dtype)
        else:
            out = np.empty(((self.fields,) if self.vector else ()) + self.shape, dtype=self.dtype)
        check(lib.mmspx_download(self._h, out.ctypes.data_as(C.c_void_p), layout))
        return out

    def step(self, model, params, steps):
        m = _model(model)
        p = _params(m, params)
        check(lib.mmspx_step(self._h, m, _dp(p), p.size, int(steps)))

    def step_timed(self, model, params, steps):
        m = _model(model)
        p = _params(m, params)
        sec = C.c_double()
        nl = C.c_int64()
        check(lib.mmspx_step_timed(self._h, m, _dp(p), p.size, int(steps), C.byref(sec), C.byref(nl)))
        return sec.value, nl.value

    def sync(self):
        check(lib.mmspx_sync(self._h))

    def set_variant(self, v):
        check(lib.mmspx_set_variant(self._h, int(v)))

    def set_tile(self, ty, tz, chunk):
        check(lib.mmspx_set_tile(self._h, ty, tz, chunk))

    def set_tile_options(self, register_queues=False, body=None):
        """Tile body (mmspx_set_tile_options): Cahn-Hilliard 0 LDS rings, 1
        register queues, 2 and 3 register queues on 768 and 1024 threads (3-D
        32 x 64 tiles), 6 and 7 two steps per pass (3-D periodic); 2-D Allen-Cahn
        and grain growth on 1 x 64 strips: 2 wave strips; 4 and 5 the split
        strips, 9, 10 and 11 the split strips two steps per pass (2-D
        periodic); 8 the 3-D plane march; 12 the 3-D plane march two steps per
        pass (periodic; Ostwald ripening: its two passes in one march)."""
        check(lib.mmspx_set_tile_options(self._h, int(body) if body is not None else (1 if register_queues else 0)))

    def get_tile(self):
        ty, tz, ch = C.c_int(), C.c_int(), C.c_int()
        v = lib.mmspx_get_tile(self._h, C.byref(ty), C.byref(tz), C.byref(ch))
        return dict(variant=v % 16, ty=ty.value, tz=tz.value, chunk=ch.value, opt=v // 16)

    def autotune(self, model, params, trial_steps=20):
        m = _model(model)
        p = _params(m, params)
        buf = C.create_string_buffer(8192)
        check(lib.mmspx_autotune(self._h, m, _dp(p), p.size, int(trial_steps), buf, 8192))
        return buf.value.decode()

    def kernel_time(self, model, params, reps):
        """(mean, min) device seconds of one step, each step bracketed by its own events."""
        m = _model(model)
        p = _params(m, params)
        mean, mn = C.c_double(), C.c_double()
        check(lib.mmspx_kernel_time(self._h, m, _dp(p), p.size, int(reps), C.byref(mean), C.byref(mn)))
        return mean.value, mn.value

    def digest(self):
        """64-bit digest of the state (equals host_digest of download())."""
        out = C.c_uint64()
        check(lib.mmspx_digest(self._h, C.byref(out)))
        return out.value

    def mass(self):
        out = np.zeros(self.fields, dtype=np.float64)
        check(lib.mmspx_reduce(self._h, MASS, 0, None, 0, _dp(out), out.size))
        return out

    def energy(self, model, params):
        m = _model(model)
        p = _params(m, params)
        out = np.zeros(1, dtype=np.float64)
        check(lib.mmspx_reduce(self._h, ENERGY, m, _dp(p), p.size, _dp(out), 1))
        return float(out[0])

    def write_dat(self, path, compress=True):
        check(lib.mmspx_write_dat(self._h, path.encode(), 1 if compress else 0))

    def checkpoint(self, path, compress=True, threads=0):
        """Starts a .dat checkpoint of the current state and returns at once;
        steps issued afterwards overlap the download and the writing. Call
        wait() on the result (threads: host threads deflating, 0 = one per
        core; 1 writes compress2's exact bytes)."""
        h = C.c_void_p()
        check(lib.mmspx_checkpoint_begin(self._h, path.encode(), 1 if compress else 0, int(threads), C.byref(h)))
        return Checkpoint(h)


class Checkpoint:
    """A checkpoint in flight (State.checkpoint)."""

    def __init__(self, h):
        self._h = h

    def wait(self):
        """Joins the writer; returns the seconds from begin to the file's close."""
        if self._h is None:
            raise RuntimeError("checkpoint already waited for")
        sec = C.c_double()
        h, self._h = self._h, None
        check(lib.mmspx_checkpoint_wait(h, C.byref(sec)))
        return sec.value

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib.mmspx_checkpoint_wait(self._h, None)
            self._h = None


def write_dat_host(path, array, shape, fields=1, vector=False, precision="f64", bc=None, dx=None, g0=None,
                   compress=True, threads=0):
    """Writes a host array (nodes in row-major order, each node's fields
    contiguous) as a serial-layout .dat with `threads` deflating threads."""
    d = make_desc(shape, fields, vector, precision, bc, dx, g0)
    dt = np.float64 if d.precision == F64 else np.float32
    a = np.ascontiguousarray(array, dtype=dt)
    n = int(np.prod(shape)) * (int(fields) if vector else 1)
    if a.size != n:
        raise ValueError("array size %d does not match grid %s x %d" % (a.size, tuple(shape), fields))
    check(lib.mmspx_write_dat_host(C.byref(d), a.ctypes.data_as(C.c_void_p), path.encode(), 1 if compress else 0,
                                   int(threads)))


class Batch:
    """Independent grids advanced together: one persistent launch per step over
    every member's tiles, multi-step runs replayed from a captured hipGraph."""

    def __init__(self, shapes, model, precision="f32", fields=1, vector=False, bcs=None, dxs=None):
        self.model = _model(model)
        descs = (Desc * len(shapes))()
        for m, sh in enumerate(shapes):
            descs[m] = make_desc(sh, fields, vector, precision, bcs[m] if bcs else None, dxs[m] if dxs else None)
        h = C.c_void_p()
        check(lib.mmspx_batch_create(descs, len(shapes), self.model, C.byref(h)))
        self._h = h
        self.shapes = [tuple(s) for s in shapes]
        self.fields = fields if vector else 1
        self.vector = vector
        self.dtype = np.float64 if make_desc((1,), precision=precision).precision == F64 else np.float32

    def close(self):
        if getattr(self, "_h", None):
            lib.mmspx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def host_shape(self, m):
        return self.shapes[m] + ((self.fields,) if self.vector else ())

    def upload(self, m, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != int(np.prod(self.host_shape(m))):
            raise ValueError("array does not match member %d" % m)
        check(lib.mmspx_batch_upload(self._h, m, a.ctypes.data_as(C.c_void_p), NODE_MAJOR))

    def download(self, m):
        out = np.empty(self.host_shape(m), dtype=self.dtype)
        check(lib.mmspx_batch_download(self._h, m, out.ctypes.data_as(C.c_void_p), NODE_MAJOR))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        if p.ndim != 2 or p.shape[0] != len(self.shapes):
            raise ValueError("params must have one row per member")
        check(lib.mmspx_batch_set_params(self._h, _dp(p), p.shape[1]))

    def step(self, steps, graph_steps=0):
        check(lib.mmspx_batch_step(self._h, int(steps), int(graph_steps)))

    def step_timed(self, steps, graph_steps=0):
        sec = C.c_double()
        check(lib.mmspx_batch_step_timed(self._h, int(steps), int(graph_steps), C.byref(sec)))
        return sec.value

    def mass(self, m):
        out = np.zeros(self.fields, dtype=np.float64)
        check(lib.mmspx_batch_reduce(self._h, m, MASS, _dp(out), out.size))
        return out

    def autotune(self, trial_steps=10):
        buf = C.create_string_buffer(16384)
        check(lib.mmspx_batch_autotune(self._h, int(trial_steps), buf, 16384))
        return buf.value.decode()

    def set_tile(self, strip_width=256, rows_per_item=32, blocks_per_cu=8, cells=0):
        """Strips (width 64/128/256), runs of `cells` nodes (Allen-Cahn), or with
        strip_width 48 temporal blocks taking rows_per_item steps per pass
        (periodic Allen-Cahn)."""
        check(lib.mmspx_batch_set_tile(self._h, strip_width, rows_per_item, blocks_per_cu, cells))

    def get_tile(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        tuned = lib.mmspx_batch_get_tile(self._h, C.byref(a), C.byref(b), C.byref(c))
        return dict(strip_width=a.value, rows_per_item=b.value, blocks_per_cu=c.value, tuned=bool(tuned))


QUEUED, RUNNING, DONE, REJECTED, CANCELLED = 0, 1, 2, 3, 4


class Server:
    """Continuous-batching solve server (a worker thread in the library owns the GPU)."""

    def __init__(self, model="ch_dw", precision="f32", hbm_budget=64 << 30, max_slots=256, chunk_steps=100,
                 fp64_above_steps=0):
        """fp64_above_steps > 0 (FP32 servers): requests of more steps run in FP64."""
        cfg = ServerConfig()
        cfg.hbm_budget = int(hbm_budget)
        cfg.max_slots = int(max_slots)
        cfg.chunk_steps = int(chunk_steps)
        cfg.model = _model(model)
        cfg.precision = PRECISIONS[precision]
        cfg.fp64_above_steps = int(fp64_above_steps)
        h = C.c_void_p()
        check(lib.mmspx_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.shapes = {}

    def close(self):
        if getattr(self, "_h", None):
            lib.mmspx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, init, params, steps, bc=None, dx=None):
        a = np.ascontiguousarray(init, dtype=np.float64)
        d = make_desc(a.shape, 1, False, "f64", bc, dx)
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
        rid = C.c_int64()
        check(lib.mmspx_server_submit(self._h, C.byref(d), _dp(a.reshape(-1)), _dp(p), p.size, int(steps),
                                      C.byref(rid)))
        self.shapes[rid.value] = a.shape
        return rid.value

    @staticmethod
    def _info(i):
        return dict(id=i.id, status=i.status, steps_done=i.steps_done, t_submit=i.t_submit, t_admit=i.t_admit,
                    t_done=i.t_done, mass0=i.mass0, mass1=i.mass1,
                    precision="f64" if i.precision == F64 else "f32")

    def wait(self, rid, timeout=-1.0):
        i = RequestInfo()
        check(lib.mmspx_server_wait(self._h, rid, float(timeout), C.byref(i)))
        return self._info(i)

    def info(self, rid):
        i = RequestInfo()
        check(lib.mmspx_server_info(self._h, rid, C.byref(i)))
        return self._info(i)

    def fetch(self, rid):
        shape = self.shapes.pop(rid)
        out = np.empty(shape, dtype=np.float64)
        check(lib.mmspx_server_fetch(self._h, rid, _dp(out.reshape(-1)), out.size))
        return out

    def cancel(self, rid):
        check(lib.mmspx_server_cancel(self._h, rid))

    def clock(self):
        return lib.mmspx_server_clock(self._h)

    def stats(self):
        out = (C.c_int64 * 8)()
        err = lib.mmspx_server_stats(self._h, out, 8)
        keys = ["admitted", "completed", "rejected", "rounds", "peak_bytes", "peak_concurrent", "launches", "fp64_routed"]
        d = dict(zip(keys, list(out)))
        d["worker_error"] = err
        return d


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """Attainable device copy rate (bytes read + written per second)."""
    out = C.c_double()
    check(lib.mmspx_copy_bandwidth(int(nbytes), int(reps), C.byref(out)))
    return out.value


def peek_dat(path) -> Desc:
    d = Desc()
    check(lib.mmspx_peek_dat(path.encode(), C.byref(d)))
    return d
