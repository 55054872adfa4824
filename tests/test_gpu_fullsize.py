"""GPU parity at the benchmark workloads' full sizes, against digests of the
oracle's states (tests/golden/fullsize.json, written by
tests/golden/make_fullsize.py from the oracle's C restatement, which the
golden cases pin bitwise to the reference's own update()).

Bars (written next to each comparison):
  * FP64 form: SHA-256 of the downloaded state equals the oracle's (bitwise).
    Where the reference's arithmetic calls libm pow for the cube (Allen-Cahn,
    grain growth) the oracle with the device's correctly rounded cube is the
    bitwise target, and its distance to the oracle with libm pow (bitwise the
    reference) is recorded in the fixture and asserted <= 1e-12 normwise.
  * FP32 form: bitwise equal to the oracle's FP32 restatement, and its
    normwise error per field against the FP64 state from the same
    float-rounded start asserted against the bound stated beside it.
  * Cahn-Hilliard mass: relative change <= 1e-12 (FP64).
Workloads: C1 512^2 Allen-Cahn, C2 256^3 Cahn-Hilliard (FP64 and FP32), C3
2048^2 x 32 grain growth, C4 the 64 batched Allen-Cahn members, and the 3-D
bench workloads on their default kernels: 256^3 Allen-Cahn, 256^3 x 8 grain
growth, 256^3 x 2 Ostwald ripening.
Normwise error: max_i |x_i - ref_i| / max_i |ref_i|, per field.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

mx = pytest.importorskip("mmsp_amd")

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize.json")
with open(FIXTURE) as f:
    FULL = json.load(f)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def digest(a):
    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()


def first_planes_differing(a, planes, k=5):
    got = [hashlib.sha256(memoryview(np.ascontiguousarray(a[i])).cast("B")).hexdigest()[:16] for i in range(a.shape[0])]
    return [i for i, (x, y) in enumerate(zip(got, planes)) if x != y][:k]


def normwise(x, ref, vector):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if vector:
        return [float(np.abs(x[..., j] - ref[..., j]).max() / np.abs(ref[..., j]).max()) for j in range(ref.shape[-1])]
    return [float(np.abs(x - ref).max() / np.abs(ref).max())]


def need(name):
    if name not in FULL:
        pytest.skip("tests/golden/fullsize.json has no %s entry (run tests/golden/make_fullsize.py)" % name)
    return FULL[name]


def test_c1_allen_cahn_512_fp64_digest():
    """512^2 FP64, 1000 steps: the oracle's digest; the correctly rounded cube
    stays within 1e-12 of libm pow's (the reference's arithmetic)."""
    w = need("ac2d")
    say("[gpu] full size ac2d fp64")
    a = O.uniform(512 * 512, w["seed"], w["lo"], w["hi"]).reshape(512, 512)
    with mx.Grid((512, 512)) as g:
        g.upload(a)
        g.step("ac", w["params"], w["steps"])
        got = g.download()
    assert digest(got) == w["f64"]["sha256"]
    assert w["f64"]["err_vs_reference_pow"] <= 1e-12  # GPU state vs the reference's arithmetic


def test_c2_cahn_hilliard_256_fp64_bitwise():
    """256^3 FP64, 2000 steps with the shipped default kernel: bitwise the
    oracle's state (the double-well mu has no cube: bitwise the reference's
    arithmetic too); mass conserved to 1e-12."""
    w = need("ch3d")
    say("[gpu] full size ch3d fp64 2000 steps")
    shape = tuple(w["shape"])
    a = O.uniform(int(np.prod(shape)), w["seed"], w["lo"], w["hi"]).reshape(shape)
    with mx.Grid(shape) as g:
        g.upload(a)
        m0 = g.mass()[0]
        g.step("ch_dw", w["params"], w["steps"])
        m1 = g.mass()[0]
        tile = g.get_tile()
        got = g.download()
    assert tile["variant"] == 2, tile  # the fused kernel the bench times
    d = digest(got)
    assert d == w["f64"]["sha256"], ("planes differing", first_planes_differing(got, w["f64"]["planes"]))
    assert abs(m1 - m0) / abs(m0) <= 1e-12, (m0, m1)


def test_c2_cahn_hilliard_256_fp32_error_vs_fp64():
    """256^3 FP32, 2000 steps: bitwise the oracle's FP32 restatement; the
    FP64 run from the same float-rounded start is bitwise the oracle's; the
    FP32 error against it is stated (spinodal decomposition amplifies the
    rounding of the stored state): <= 5e-5 normwise."""
    w = need("ch3d")
    say("[gpu] full size ch3d fp32 2000 steps")
    shape = tuple(w["shape"])
    a = O.uniform(int(np.prod(shape)), w["seed"], w["lo"], w["hi"]).reshape(shape).astype(np.float32)
    with mx.Grid(shape, precision="f32") as g:
        g.upload(a)
        m0 = g.mass()[0]
        g.step("ch_dw", w["params"], w["steps"])
        m1 = g.mass()[0]
        f32 = g.download()
    assert digest(f32) == w["f32"]["sha256"], ("planes differing", first_planes_differing(f32, w["f32"]["planes"]))
    with mx.Grid(shape) as g:
        g.upload(a.astype(np.float64))
        g.step("ch_dw", w["params"], w["steps"])
        f64 = g.download()
    assert digest(f64) == w["f64_from_f32_start"]["sha256"]
    err = normwise(f32, f64, False)[0]
    say("[gpu] ch3d fp32 err vs fp64 %.3e, mass rel change %.3e" % (err, abs(m1 - m0) / abs(m0)))
    assert err == pytest.approx(w["f32"]["err_vs_f64"], rel=1e-12)
    assert err <= 5e-5  # stated FP32 bound for 2000 steps of this spinodal decomposition
    assert abs(m1 - m0) / abs(m0) <= 1e-5  # FP32 storage, FP64 mass reduction


def test_c3_grain_growth_2048_q32_fp32_and_fp64():
    """2048^2 x 32 fields, 500 steps, the shipped default kernels (split wave
    strips): FP64 bitwise the oracle's (correctly rounded cube), itself within
    1e-12 of the reference's libm-pow arithmetic; FP32 bitwise the oracle's FP32
    restatement and within 1e-5 of the FP64 state per field."""
    w = need("gg2d")
    shape, q = tuple(w["shape"]), w["fields"]
    n = int(np.prod(shape))
    a = O.uniform(n * q, w["seed"], w["lo"], w["hi"]).astype(np.float32).reshape(shape + (q,))
    say("[gpu] full size gg2d fp32 500 steps")
    with mx.Grid(shape, fields=q, vector=True, precision="f32") as g:
        g.upload(a)
        g.step("gg", [w["dt"]], w["steps"])
        tile = g.get_tile()
        f32 = g.download()
    assert tile["variant"] == 2, tile
    assert digest(f32) == w["f32"]["sha256"], ("rows differing", first_planes_differing(f32, w["f32"]["planes"]))
    say("[gpu] full size gg2d fp64 500 steps")
    with mx.Grid(shape, fields=q, vector=True, precision="f64") as g:
        g.upload(a.astype(np.float64))
        g.step("gg", [w["dt"]], w["steps"])
        f64 = g.download()
    assert digest(f64) == w["f64"]["sha256"], ("rows differing", first_planes_differing(f64, w["f64"]["planes"]))
    assert max(w["f64"]["err_vs_reference_pow"]) <= 1e-12  # GPU FP64 vs the reference's arithmetic
    err = normwise(f32, f64, True)
    say("[gpu] gg2d fp32 err vs fp64 max over fields %.3e" % max(err))
    assert max(err) <= 1e-5  # the FP32 bar per field


def test_ac3d_allen_cahn_256_fp64_digest():
    """256^3 FP64 Allen-Cahn, 1000 steps on the shipped default (the plane
    march two steps per pass): the oracle's digest; the correctly rounded cube
    stays within 1e-12 of libm pow's (the reference's arithmetic)."""
    w = need("ac3d")
    say("[gpu] full size ac3d fp64 1000 steps")
    shape = tuple(w["shape"])
    a = O.uniform(int(np.prod(shape)), w["seed"], w["lo"], w["hi"]).reshape(shape)
    with mx.Grid(shape) as g:
        g.upload(a)
        g.step("ac", w["params"], w["steps"])
        tile = g.get_tile()
        got = g.download()
    assert (tile["variant"], tile["opt"]) == (2, 12), tile  # the kernel the bench times
    assert digest(got) == w["f64"]["sha256"], ("planes differing", first_planes_differing(got, w["f64"]["planes"]))
    assert w["f64"]["err_vs_reference_pow"] <= 1e-12  # GPU state vs the reference's arithmetic


def test_gg3d_grain_growth_256_q8_fp32_and_fp64():
    """256^3 x 8 fields, 200 steps, the shipped defaults (FP32: the plane
    march two steps per pass; FP64: the one-step march): FP32 bitwise the
    oracle's FP32 restatement, FP64 from the same float-rounded start bitwise
    the oracle's (correctly rounded cube); FP32 within 1e-5 of FP64 per field."""
    w = need("gg3d")
    shape, q = tuple(w["shape"]), w["fields"]
    a = O.uniform(int(np.prod(shape)) * q, w["seed"], w["lo"], w["hi"]).astype(np.float32).reshape(shape + (q,))
    say("[gpu] full size gg3d fp32 200 steps")
    with mx.Grid(shape, fields=q, vector=True, precision="f32") as g:
        g.upload(a)
        g.step("gg", [w["dt"]], w["steps"])
        tile = g.get_tile()
        f32 = g.download()
    assert (tile["variant"], tile["opt"]) == (2, 12), tile
    assert digest(f32) == w["f32"]["sha256"], ("planes differing", first_planes_differing(f32, w["f32"]["planes"]))
    say("[gpu] full size gg3d fp64 200 steps")
    with mx.Grid(shape, fields=q, vector=True, precision="f64") as g:
        g.upload(a.astype(np.float64))
        g.step("gg", [w["dt"]], w["steps"])
        f64 = g.download()
    assert digest(f64) == w["f64"]["sha256"], ("planes differing", first_planes_differing(f64, w["f64"]["planes"]))
    err = normwise(f32, f64, True)
    say("[gpu] gg3d fp32 err vs fp64 max over fields %.3e" % max(err))
    assert max(err) <= 1e-5  # the FP32 bar per field


def test_ostwald3d_256_q2_fp64_digest():
    """256^3 x 2 fields FP64 Ostwald ripening, 200 steps on the shipped
    default (the two passes in one plane march, w kept in LDS): bitwise the
    oracle's state."""
    w = need("ostwald3d")
    shape, q = tuple(w["shape"]), w["fields"]
    say("[gpu] full size ostwald3d fp64 200 steps")
    a = O.uniform(int(np.prod(shape)) * q, w["seed"], w["lo"], w["hi"]).reshape(shape + (q,))
    with mx.Grid(shape, fields=q, vector=True) as g:
        g.upload(a)
        g.step("ostwald", w["params"], w["steps"])
        tile = g.get_tile()
        got = g.download()
    assert (tile["variant"], tile["opt"]) == (2, 12), tile
    assert digest(got) == w["f64"]["sha256"], ("planes differing", first_planes_differing(got, w["f64"]["planes"]))


def test_c4_batched_allen_cahn_64_members():
    """The 64 uneven members (sides 128..2048) in one batch with the
    autotuned decomposition and a captured graph of 100 steps, 1000 steps:
    every member bitwise the oracle's FP32 run; the same batch in FP64
    bitwise the oracle's FP64 run; FP32 error per member <= 1e-5."""
    w = need("batch_ac")
    members = w["members"]
    shapes = [(m["side"], m["side"]) for m in members]
    params = np.array([[1.0, 1.0, m["kappa"], 1.0, 0.01] for m in members])
    inits = [O.uniform(m["side"] ** 2, m["seed"], w["lo"], w["hi"]).astype(np.float32).reshape(m["side"], m["side"])
             for m in members]
    out = {}
    for prec in ("f32", "f64"):
        say("[gpu] full size batch_ac %s %d steps" % (prec, w["steps"]))
        with mx.Batch(shapes, "ac", precision=prec) as b:
            b.set_params(params)
            for k, a in enumerate(inits):
                b.upload(k, a)
            if prec == "f32":
                b.autotune(trial_steps=3)  # the decomposition the benchmark runs
                for k, a in enumerate(inits):
                    b.upload(k, a)
            b.step(w["steps"], graph_steps=100)
            out[prec] = [b.download(k) for k in range(len(members))]
    bad32 = [k for k, m in enumerate(members) if digest(out["f32"][k]) != m["f32"]]
    bad64 = [k for k, m in enumerate(members) if digest(out["f64"][k]) != m["f64"]]
    assert not bad32 and not bad64, (bad32, bad64)
    errs = [normwise(out["f32"][k], out["f64"][k], False)[0] for k in range(len(members))]
    say("[gpu] batch_ac fp32 err vs fp64 max %.3e" % max(errs))
    assert max(errs) <= 1e-5  # the FP32 bar for batched solves
