"""GPU: the plug-in boundary. The reference's own programs (main(), command
line, .dat reader/writer and checkpoint naming from MMSP.main.hpp and the MMSP
headers) compiled with their update() bodies replaced by include/mmsp_gpu.hpp
(oracle/_ref/*-gpu, built from mmsp_amd/dropin/) must write the same files as
the stock reference programs (oracle/_ref/allen-cahn ...), run side by side:
same names, same headers, data within 1e-12 (FP64) and bitwise equal to the
oracle with the device's cube."""
import math
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from oracle import dat
from conftest import golden_path, params_of

pytestmark = pytest.mark.gpu
mx = pytest.importorskip("mmsp_amd")

PROGRAMS = {"ac": ("allen-cahn", "allen-cahn-gpu"), "ch": ("cahn-hilliard", "cahn-hilliard-gpu"),
            "gg": ("graingrowth", "graingrowth-gpu"), "zn": ("zener", "zener-gpu"), "ow": ("ostwald", "ostwald-gpu"),
            "ga": ("graingrowth-aniso", "graingrowth-aniso-gpu")}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


DROPIN_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mmsp_amd", "dropin", "bin")


def where(name):
    """Stock reference programs in oracle/_ref, drop-in programs (-gpu) in mmsp_amd/dropin/bin."""
    return os.path.join(DROPIN_BIN if "-gpu" in name else O.REF_DIR, name)


def have(name):
    return os.path.exists(where(name))


def run(prog, cwd, args, env=None, stderr=False):
    r = subprocess.run([where(prog)] + args, cwd=cwd, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (prog, r.stdout[-2000:], r.stderr[-2000:])
    return (r.stdout, r.stderr) if stderr else r.stdout


def progress_shape(out):
    # the bar's dots and the structure of the lines, without clock values
    return [re.sub(r"[0-9]", "#", line) for line in out.splitlines()
            if "•" in line or line.startswith("No.")]


# stock-parameter golden cases: the drop-in programs use the examples' own parameters
# (the stock programs' typedefs are double grids, so the float fixtures are not used here)
CASES = [("ac2d_f64", "ac"), ("ac3d_bc_f64", "ac"), ("ch2d_stock_f64", "ch"),
         ("gg3d_f64", "gg"), ("gg2d_bc_f64", "gg"), ("gg1d_f64", "gg"),
         ("zn2d_f64", "zn"), ("zn3d_f64", "zn"), ("ow2d_f64", "ow"), ("ow3d_f64", "ow"), ("ga2d_f64", "ga"),
         ("ga3d_f64", "ga")]


@pytest.mark.parametrize("name,model", CASES)
def test_dropin_program_matches_stock_program(name, model, tmp_path):
    stock, gpu = PROGRAMS[model]
    if not (have(stock) and have(gpu)):
        pytest.skip("reference programs not built here")
    say("[gpu] drop-in %s via %s" % (name, gpu))
    import json

    case = [c for c in json.load(open(os.path.join(os.path.dirname(golden_path(name, "in")), "manifest.json")))
            if c["name"] == name][0]
    steps = case["steps"]
    inc = max(1, steps // 3)
    while steps % inc:
        inc += 1
    outs = {}
    for prog in (stock, gpu):
        d = tmp_path / prog
        d.mkdir()
        shutil.copy(golden_path(name, "in"), d / "start.dat")
        out = run(prog, str(d), ["start.dat", str(steps), str(inc)])
        outs[prog] = (sorted(p for p in os.listdir(d) if p != "start.dat"), out, d)
    files_s, out_s, ds = outs[stock]
    files_g, out_g, dg = outs[gpu]
    assert files_g == files_s and len(files_g) == steps // inc  # checkpoint names from MMSP.main.hpp
    assert progress_shape(out_g) == progress_shape(out_s)
    gin = dat.read_dat(golden_path(name, "in"))
    for f in files_s:
        a, b = dat.read_dat(str(ds / f)), dat.read_dat(str(dg / f))
        assert (a.dim, a.hdr_fields, a.b0, a.b1, a.dx, a.g0) == (b.dim, b.hdr_fields, b.b0, b.b1, b.dx, b.g0)
        ref = a.data.astype(np.float64)
        err = np.abs(b.data.astype(np.float64) - ref).max() / np.abs(ref).max()
        assert err <= (1e-12 if a.data.dtype == np.float64 else 1e-6), (f, err)
    # the last checkpoint equals the oracle run with the device's cube, bitwise
    form = "f64" if case["dtype"] == "f64" else "f32ref"
    p = params_of(case)
    bc = list(zip(gin.b0, gin.b1))
    if model == "ac":
        exp = O.ac(gin.data, p, steps, form=form, bc=bc, dx=gin.dx)
    elif model == "ch":
        exp = O.ch(gin.data, p, steps, mu=O.MU_POLY, form=form, bc=bc, dx=gin.dx)
    elif model == "zn":
        exp = O.zener(gin.data, p[0], steps, form=form, bc=bc, dx=gin.dx)
    elif model == "ow":
        exp = O.ostwald(gin.data, p, steps, form=form, bc=bc, dx=gin.dx)
    elif model == "ga":
        exp = O.gg_aniso(gin.data, p, steps, form=form, bc=bc, dx=gin.dx)
    else:
        exp = O.gg(gin.data, p[0], steps, form=form, bc=bc, dx=gin.dx)
    last = dat.read_dat(str(dg / files_g[-1])).data
    assert np.array_equal(last, exp)
    # and the stock program's last file equals the reference's golden output
    assert np.array_equal(dat.read_dat(str(ds / files_s[-1])).data, dat.read_dat(golden_path(name, "out")).data)



# The examples' own parameters: model_A.cpp:85-91, model_B.cpp:86-92, spinodal.cpp:86-88.
STOCHASTIC = {
    "model_A": (O.ac_noise, [1.0, 1.0, 1.0, 1.0, 0.01, math.sqrt(2.0 * 0.01 / 0.01)]),
    "model_B": (O.ch_noise, [1.0, 1.0, 1.0, 1.0, 0.01, math.sqrt(2.0 * 0.01 / 0.01)]),
    "spinodal": (O.ch_noise, [1.0, 1.0, 1.0, 0.5, 0.01, math.sqrt(0.05 * 0.01)]),
}


@pytest.mark.parametrize("prog", ["model_A", "model_B", "spinodal"])
def test_stochastic_dropin_programs_agree_in_distribution(prog, tmp_path):
    """model_A, model_B and spinodal with the GPU update beside the stock
    programs (which seed rand() from the clock): the same checkpoint names,
    headers and progress output, and from phi = 0 on 256^2 after 100 steps the
    mean of phi^2 of both within 5 % of the oracle's mean over four seeds
    (the spread of one run is about 1 %)."""
    gpu = prog + "-gpu"
    if not (have(prog) and have(gpu)):
        pytest.skip("reference programs not built here")
    say("[gpu] stochastic drop-in %s" % prog)
    n, steps, inc = 256, 100, 50
    outs = {}
    for p in (prog, gpu):
        d = tmp_path / p
        d.mkdir()
        dat.write_dat(str(d / "start.dat"), dat.make_grid(np.zeros((n, n))), compress=True)
        out = run(p, str(d), ["start.dat", str(steps), str(inc)])
        outs[p] = (sorted(f for f in os.listdir(d) if f != "start.dat"), out, d)
    files_s, out_s, ds = outs[prog]
    files_g, out_g, dg = outs[gpu]
    assert files_g == files_s and len(files_g) == steps // inc
    assert progress_shape(out_g) == progress_shape(out_s)
    f, p = STOCHASTIC[prog]
    expect = float(np.mean([(f(np.zeros((n, n)), p + [seed], steps) ** 2).mean() for seed in range(1, 5)]))
    for dd in (ds, dg):
        a = dat.read_dat(str(dd / files_s[-1]))
        assert (a.dim, a.dx, a.g0) == (2, [1.0, 1.0], [0, 0])
        v = float((a.data ** 2).mean())
        assert abs(v / expect - 1) < 0.05, (dd, v, expect)


def gpu_stats(err):
    """The counters mmsp_gpu.hpp prints at exit under MMSP_GPU_STATS=1."""
    m = re.search(r"update\(\) calls (\d+), steps (\d+), uploads (\d+), downloads (\d+)", err)
    assert m, err[-2000:]
    return dict(zip(("calls", "steps", "uploads", "downloads"), map(int, m.groups())))


def test_driver_loop_keeps_the_state_resident(tmp_path):
    """MMSP.main.hpp:304-325 calls update(grid, increment) in a loop with
    output() in between: the drop-in uploads the grid once and keeps it on the
    device (one download per call), and the checkpoints equal the oracle's
    states bitwise. A grid the caller changes between calls is uploaded again
    (in-process, through the library's digests)."""
    prog = "cahn-hilliard-dw-gpu"
    if not have(prog):
        pytest.skip("drop-in programs not built here")
    say("[gpu] drop-in driver loop, resident state")
    shape, steps, inc = (24, 20, 32), 300, 50
    c0 = O.uniform(int(np.prod(shape)), 41, 0.45, 0.55).reshape(shape)
    mx.write_dat_host(str(tmp_path / "start.dat"), c0, shape)
    env = dict(os.environ, MMSP_GPU_STATS="1")
    out, err = run(prog, str(tmp_path), ["start.dat", str(steps), str(inc)], env=env, stderr=True)
    st = gpu_stats(err)
    assert st == dict(calls=steps // inc, steps=steps, uploads=1, downloads=steps // inc), st
    files = sorted(p for p in os.listdir(tmp_path) if p != "start.dat")
    assert len(files) == steps // inc
    state = c0
    for f in files:
        state = O.ch(state, [1.0, 1.0, 0.005], inc, mu=O.MU_DW)
        assert np.array_equal(dat.read_dat(str(tmp_path / f)).data, state), f
    # the digests: equal for the state and its download, different after one changed bit
    with mx.Grid(shape) as g:
        g.upload(c0)
        g.step("ch_dw", [1.0, 1.0, 0.005], 7)
        h = g.digest()
        x = g.download()
        assert h == mx.host_digest(x)
        x.reshape(-1)[1234] = np.nextafter(x.reshape(-1)[1234], 2.0)
        assert mx.host_digest(x) != h
    q = O.uniform(30 * 40 * 8, 42, 0.0, 1.0).reshape(30, 40, 8).astype(np.float32)
    with mx.Grid((30, 40), fields=8, vector=True, precision="f32") as g:
        g.upload(q)
        assert g.digest() == mx.host_digest(q)  # node-major order for vector grids
