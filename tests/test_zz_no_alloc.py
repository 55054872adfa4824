"""mmspx_step_noalloc (explicit steps on caller-owned device buffers, the
two-pass kernels' scalar plane in the library's fixed device array) against
mmspx_step on a library state: bitwise, on two calls in a row, the second
with a smaller grid, and against the oracle.

The device buffers are the caller's: allocated here with hipMalloc from the
HIP runtime libmmsp_amd.so links (not through torch, which bundles a runtime
of its own); the entry point is called through the C ABI (ctypes)."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

mx = pytest.importorskip("mmsp_amd")

_hip = None


def hip():
    """The HIP runtime already loaded for libmmsp_amd.so (same soname)."""
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = h
    return _hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), self.nbytes) == 0
        self.ptr = p.value

    def put(self, host):
        assert host.nbytes == self.nbytes
        assert hip().hipMemcpy(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0

    def get(self, dtype):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

P, N, M, D = O.PERIODIC, O.NEUMANN, O.MIRROR, O.DIRICHLET


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def via_state(shape, prec, model, p, a, steps, bc=None, q=0):
    kw = dict(fields=q, vector=True) if q else {}
    with mx.Grid(shape, precision=prec, bc=bc, **kw) as g:
        g.upload(a)
        g.step(model, p, steps)
        return g.download()


def via_noalloc(shape, prec, model, p, a, steps, bc=None, q=0, calls=(None,)):
    """Runs `steps` through mmspx_step_noalloc, split into the step counts of
    `calls` (None: one call), continuing the step numbering."""
    desc = mx.make_desc(shape, q if q else 1, bool(q), prec, bc)
    dt = np.float64 if prec == "f64" else np.float32
    host = np.ascontiguousarray(np.moveaxis(a, -1, 0) if q else a, dtype=dt)  # device layout
    bufs = [DevBuf(host.nbytes), DevBuf(host.nbytes)]
    try:
        bufs[0].put(host)
        parts = [steps] if calls == (None,) else list(calls)
        assert sum(parts) == steps
        cur, done = 0, 0
        for n in parts:  # the null stream
            res = mx.step_noalloc(desc, bufs[cur].ptr, bufs[1 - cur].ptr, model, p, n, first_step=done)
            cur = cur if res == 0 else 1 - cur
            done += n
        assert hip().hipDeviceSynchronize() == 0
        out = bufs[cur].get(dt).reshape(host.shape)
    finally:
        for b in bufs:
            b.free()
    return np.ascontiguousarray(np.moveaxis(out, 0, -1)) if q else out


CASES = [
    # (name, model, params, prec, bc, q, (bigger shape, smaller shape), steps, (lo, hi))
    ("ch_neumann_3d", "ch_dw", [1.0, 1.0, 0.005], "f64", [(N, N), (M, D), (N, P)], 0, ((48, 40, 64), (20, 18, 30)), 13,
     (0.45, 0.55)),
    ("ch_poly_2d_f32", "ch", [1.0, 1.0, 1.0, 1.0, 0.01], "f32", [(N, M), (D, N)], 0, ((130, 170), (33, 40)), 9,
     (-1.0, 1.0)),
    # 1-D Cahn-Hilliard and an axis periodic on one side only run the two
    # passes, whose scalar plane lives in the library's fixed device array
    ("ch_two_pass_1d", "ch_dw", [1.0, 1.0, 0.005], "f64", [(N, M)], 0, ((5000,), (900,)), 10, (0.45, 0.55)),
    ("ch_two_pass_2d_f32", "ch", [1.0, 1.0, 1.0, 1.0, 0.01], "f32", [(P, N), (D, D)], 0, ((90, 120), (31, 17)), 9,
     (-1.0, 1.0)),
    ("ostwald_3d", "ostwald", [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0], "f64", [(M, N), (P, P), (D, P)], 3,
     ((16, 18, 20), (8, 9, 10)), 12, (0.0, 1.0)),
    # from 2^20 nodes 3-D Ostwald ripening runs two marches with w in the fixed array
    ("ostwald_3d_march", "ostwald", [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0], "f64", [(N, P), (P, P), (M, D)],
     2, ((104, 100, 101), (8, 9, 10)), 5, (0.0, 1.0)),
    # periodic with 2 fields: the one-pass march, w in LDS (no scalar plane)
    ("ostwald_3d_one_pass", "ostwald", [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0], "f32", None, 2,
     ((64, 128, 128), (8, 9, 10)), 5, (0.0, 1.0)),
    ("ch_periodic_3d", "ch_dw", [1.0, 1.0, 0.005], "f64", None, 0, ((40, 36, 70), (12, 10, 20)), 11, (0.45, 0.55)),
    ("ac_tblocks_2d", "ac", [1.0, 1.0, 1.0, 1.0, 0.01], "f64", None, 0, ((256, 256), (64, 48)), 13, (-0.01, 0.01)),
    ("gg_strips_2d_f32", "gg", [0.05], "f32", None, 8, ((70, 200), (20, 40)), 7, (0.0, 0.01)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_mmspx_step(case):
    name, model, p, prec, bc, q, shapes, steps, (lo, hi) = case
    for k, shape in enumerate(shapes):  # two calls in a row, the second smaller
        say("[gpu] noalloc %s %s call %d" % (name, shape, k))
        n = int(np.prod(shape)) * (q if q else 1)
        a = O.uniform(n, 1000 + 17 * k + len(name), lo, hi).reshape(tuple(shape) + ((q,) if q else ()))
        if prec != "f64":
            a = a.astype(np.float32)
        got = via_noalloc(shape, prec, model, p, a, steps, bc=bc, q=q)
        exp = via_state(shape, prec, model, p, a, steps, bc=bc, q=q)
        assert got.dtype == exp.dtype
        assert np.array_equal(got, exp), (name, shape)


def test_noalloc_noise_continues_the_step_count():
    """Model B split over two calls (first_step continues the noise counter)
    equals one mmspx_step run and the oracle."""
    shape, p = (20, 24, 30), [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 11]
    for bc in (None, [(N, N), (P, P), (M, D)]):
        say("[gpu] noalloc model B bc=%s" % (bc,))
        a = O.uniform(int(np.prod(shape)), 1200, -0.1, 0.1).reshape(shape)
        got = via_noalloc(shape, "f64", "ch_noise", p, a, 13, bc=bc, calls=(7, 6))
        exp = via_state(shape, "f64", "ch_noise", p, a, 13, bc=bc)
        assert np.array_equal(got, exp)
        assert np.array_equal(got, O.ch_noise(a, p, 13, bc=bc))


def test_noalloc_refuses_planes_beyond_the_array_and_bad_arguments():
    # 1-D FP64 two-pass Cahn-Hilliard of 5 000 000 nodes: a 40 MB plane > 32 MiB
    n1 = 5_000_000
    desc = mx.make_desc((n1,), precision="f64", bc=[(N, N)])
    a, b = DevBuf(n1 * 8), DevBuf(n1 * 8)
    try:
        with pytest.raises(mx.MmspxError) as e:
            mx.step_noalloc(desc, a.ptr, b.ptr, "ch_dw", [1.0, 1.0, 0.005], 1)
        assert e.value.code == -1 and "bytes" in str(e.value)
        with pytest.raises(mx.MmspxError):
            mx.step_noalloc(desc, a.ptr, a.ptr, "ch_dw", [1.0, 1.0, 0.005], 1)  # same buffer twice
        with pytest.raises(mx.MmspxError):
            mx.step_noalloc(desc, a.ptr, b.ptr, "gg", [0.05], 1)  # a vector model on a scalar grid
        # a 3-D grid of the same size takes the fused kernel, which needs no plane
        desc = mx.make_desc((200, 200, 125), precision="f64", bc=[(N, N), (P, P), (P, P)])
        assert mx.step_noalloc(desc, a.ptr, b.ptr, "ch_dw", [1.0, 1.0, 0.005], 0) == 0
        assert hip().hipDeviceSynchronize() == 0
    finally:
        a.free()
        b.free()
