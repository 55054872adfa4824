"""CPU tests of the C ABI library: it loads, exports every declared symbol,
and its host-only logic (validation, .dat header parsing, seeded inputs,
error reporting) behaves; no kernel is launched here."""
import ctypes as C
import os

import numpy as np
import pytest

import mmsp_amd as mx
import oracle as O
from conftest import golden_path


def test_library_exports_every_declared_symbol():
    names = mx.declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(mx.lib, n), n
    assert mx.lib.mmspx_abi_version() == 2


def test_strerror_names_every_code():
    for code in range(0, -9, -1):
        assert mx.lib.mmspx_strerror(code).decode() not in ("", "unknown error")
    assert mx.lib.mmspx_strerror(-99).decode() == "unknown error"


def test_fill_uniform_matches_oracle_generator():
    for seed, lo, hi in [(20260101, -0.01, 0.01), (7, 0.45, 0.55), (0, 0.0, 1.0)]:
        assert np.array_equal(mx.fill_uniform(4097, seed, lo, hi), O.uniform(4097, seed, lo, hi))


@pytest.mark.parametrize(
    "kw",
    [
        dict(shape=(0, 4)),
        dict(shape=(4, 4, 4, 4)),
        dict(shape=(4, 4), bc=[(mx.PARALLEL, mx.PERIODIC), (mx.PERIODIC, mx.PERIODIC)]),
        dict(shape=(1, 4), bc=[(mx.MIRROR, mx.MIRROR), (mx.PERIODIC, mx.PERIODIC)]),
        dict(shape=(4, 4), dx=[0.0, 1.0]),
        dict(shape=(4, 4), fields=0, vector=True),
        dict(shape=(65536, 65536)),
        dict(shape=(4, 4), precision=7),
    ],
)
def test_create_rejects_bad_descriptors_before_touching_the_device(kw):
    shape = kw.pop("shape")
    d = mx.make_desc(shape if len(shape) <= 3 else shape[:3], **kw)
    if len(shape) > 3:
        d.dim = 4
    h = C.c_void_p()
    rc = mx.lib.mmspx_create(C.byref(d), C.byref(h))
    assert rc == mx.lib.mmspx_create.restype(-1) or rc == -1
    assert not h.value
    assert mx.lib.mmspx_last_error()


def test_footprint_counts_buffers():
    f_ac = mx.footprint((256, 256, 256), "ac")
    f_ch = mx.footprint((256, 256, 256), "ch_dw")
    cells = 256 ** 3
    assert f_ac >= 2 * cells * 8 and f_ch >= 3 * cells * 8
    assert mx.footprint((2048, 2048), "gg", fields=32, vector=True, precision="f32") >= 2 * 2048 * 2048 * 32 * 4


def test_peek_dat_reads_golden_headers(cases):
    for c in cases:
        d = mx.peek_dat(golden_path(c["name"], "in"))
        assert d.dim == len(c["shape"])
        assert [d.n[k] for k in range(d.dim)] == c["shape"]
        assert d.vector == (1 if c["fields"] else 0)
        assert d.precision == (mx.F64 if c["dtype"] == "f64" else mx.F32)
        if c["bc"]:
            assert [[d.b0[k], d.b1[k]] for k in range(d.dim)] == c["bc"]
        if c["fields"]:
            assert d.fields == c["fields"]


def test_read_dat_reports_bad_files(tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"not a grid\n")
    h = C.c_void_p()
    assert mx.lib.mmspx_read_dat(str(bad).encode(), -1, C.byref(h)) == -5  # EFORMAT
    assert mx.lib.mmspx_read_dat(str(tmp_path / "missing.dat").encode(), -1, C.byref(h)) == -4  # EIO
    good = open(golden_path("ac2d_f64", "in"), "rb").read()
    trunc = tmp_path / "trunc.dat"
    trunc.write_bytes(good[: len(good) // 2])
    d = mx.Desc()
    # the header parses, the block table is cut short
    assert mx.lib.mmspx_peek_dat(str(trunc).encode(), C.byref(d)) == -5


def test_step_validates_before_launch():
    # a null state is rejected on the host
    p = np.array([1.0, 1, 1, 1, 0.01])
    assert mx.lib.mmspx_step(None, mx.AC, p.ctypes.data_as(C.POINTER(C.c_double)), 5, 1) == -1
