"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's golden vectors.

Bars (the tolerance of each comparison is written next to it):
  * FP64 form vs the oracle with the correctly rounded cube: bitwise.
  * FP64 form vs the reference's own output (libm pow cube): normwise
    max|gpu-ref|/max|ref| <= 1e-12 per field.
  * F32REF form (the reference's grid<dim,float> semantics) vs oracle: bitwise;
    vs the reference: <= 1e-6 normwise.
  * F32 form (float arithmetic) vs the oracle's F32 restatement: bitwise.
  * .dat written from device state: byte-identical to the reference's file.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from oracle import dat
from conftest import golden_cases, golden_path, model_of, params_of

pytestmark = pytest.mark.gpu

mx = pytest.importorskip("mmsp_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def normwise(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def oracle_run(model, a, params, steps, form, bc=None, dx=None, cube=O.CUBE_CR):
    if model == "ac":
        return O.ac(a, params, steps, form=form, cube=cube, bc=bc, dx=dx)
    if model == "ch":
        return O.ch(a, params, steps, mu=O.MU_POLY, form=form, cube=cube, bc=bc, dx=dx)
    if model == "ch_dw":
        return O.ch(a, params, steps, mu=O.MU_DW, form=form, cube=cube, bc=bc, dx=dx)
    if model == "zener":
        return O.zener(a, params[0], steps, form=form, cube=cube, bc=bc, dx=dx)
    if model == "ostwald":
        return O.ostwald(a, params, steps, form=form, cube=cube, bc=bc, dx=dx)
    if model == "gg_aniso":
        return O.gg_aniso(a, params, steps, form=form, bc=bc, dx=dx)
    return O.gg(a, params[0], steps, form=form, cube=cube, bc=bc, dx=dx)


CASES = golden_cases()


@pytest.mark.parametrize("variant", [1, 0, 2])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_golden_case(case, variant):
    name = case["name"]
    prec = "f64" if case["dtype"] == "f64" else "f32ref"
    say("[gpu] golden %s variant=%d precision=%s" % (name, variant, prec))
    gin = dat.read_dat(golden_path(name, "in"))
    gout = dat.read_dat(golden_path(name, "out"))
    with mx.Grid.from_dat(golden_path(name, "in"), precision=prec) as g:
        g.set_variant(variant)
        try:
            g.step(model_of(case), params_of(case), case["steps"])
        except mx.MmspxError as e:
            assert variant == 2 and e.code == -8
            pytest.skip("no tiled kernel for this case")
        got = g.download()
    exp = oracle_run(model_of(case), gin.data, params_of(case), case["steps"], prec,
                     bc=list(zip(gin.b0, gin.b1)), dx=gin.dx)
    assert got.dtype == exp.dtype
    assert np.array_equal(got, exp), (name, normwise(got, exp))
    tol = 1e-12 if prec == "f64" else 1e-6
    if gout.vector:
        for j in range(gout.fields):
            assert normwise(got[..., j], gout.data[..., j]) <= tol
    else:
        assert normwise(got, gout.data) <= tol


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_dat_written_from_device_is_byte_identical(case, tmp_path):
    name = case["name"]
    say("[gpu] dat round trip %s" % name)
    src = golden_path(name, "out")
    out = str(tmp_path / "w.dat")
    with mx.Grid.from_dat(src) as g:
        g.write_dat(out, compress=True)
    assert open(out, "rb").read() == open(src, "rb").read()
    with mx.Grid.from_dat(src) as g:
        g.write_dat(out, compress=False)
    r = dat.read_dat(out)
    assert np.array_equal(r.data, dat.read_dat(src).data)


P, N, M, D = O.PERIODIC, O.NEUMANN, O.MIRROR, O.DIRICHLET
F32_CASES = [
    ("ac", (200,), None, [1.0, 1.0, 1.3, 1.0, 0.01], 40, (-1.0, 1.0)),
    ("ac", (37, 29), [(N, M), (P, D)], [1.0, 1.0, 1.0, 1.0, 0.01], 40, (-1.0, 1.0)),
    ("ac", (64, 64), None, [1.0, 1.0, 1.0, 1.0, 0.01], 100, (-0.01, 0.01)),
    ("ac", (11, 9, 13), [(M, N), (D, P), (P, P)], [1.0, 1.0, 1.0, 1.0, 0.01], 30, (-1.0, 1.0)),
    ("ch", (33, 31), None, [1.0, 1.0, 1.0, 1.0, 0.01], 30, (-1.0, 1.0)),
    ("ch_dw", (40, 36), [(N, N), (M, M)], [1.0, 1.0, 0.005], 60, (0.45, 0.55)),
    ("ch_dw", (16, 18, 20), None, [1.0, 1.0, 0.005], 60, (0.45, 0.55)),
    ("ch_dw", (32, 32, 64), None, [1.0, 1.0, 0.005], 40, (0.45, 0.55)),
    ("gg", (30, 26), None, [0.05], 30, (0.0, 0.01)),
    ("gg", (9, 10, 11), [(N, M), (P, P), (D, N)], [0.01], 20, (0.0, 1.0)),
    ("gg4", (70, 300), [(N, M), (D, P)], [0.05], 25, (0.0, 1.0)),
    ("gg32", (40, 150), None, [0.05], 20, (0.0, 0.01)),
    ("ac", (150, 600), [(M, N), (P, P)], [1.0, 1.0, 1.0, 1.0, 0.01], 50, (-1.0, 1.0)),
]


@pytest.mark.parametrize("prec", ["f32", "f64", "f32ref"])
@pytest.mark.parametrize("idx", range(len(F32_CASES)))
def test_seeded_forms_bitwise(idx, prec):
    model, shape, bc, params, steps, (lo, hi) = F32_CASES[idx]
    q = {"gg": 5, "gg4": 4, "gg32": 32}.get(model, 0)
    model = "gg" if q else model
    say("[gpu] seeded %s %s %s bc=%s steps=%d" % (model, shape, prec, bc, steps))
    n = int(np.prod(shape)) * max(q, 1)
    a = O.uniform(n, 1000 + idx, lo, hi).reshape(tuple(shape) + ((q,) if q else ()))
    if prec != "f64":
        a = a.astype(np.float32)
    exp = oracle_run(model, a, params, steps, prec, bc=bc)
    for variant in (1, 0, 2):
        with mx.Grid(shape, fields=max(q, 1), vector=q > 0, precision=prec, bc=bc) as g:
            g.set_variant(variant)
            g.upload(a)
            try:
                g.step(model, params, steps)
            except mx.MmspxError as e:
                assert variant == 2 and e.code == -8  # no tiled kernel for this model/shape/boundaries
                continue
            got = g.download()
        assert np.array_equal(got, exp), (model, shape, prec, variant, normwise(got, exp))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fused_ch_options_and_directions(prec):
    """Both fused tile bodies on every compiled tile (3-D and 2-D rows), and a
    long graph-replayed run."""
    shape = (40, 36, 70)
    say("[gpu] fused CH options %s" % prec)
    a = O.uniform(int(np.prod(shape)), 77, 0.45, 0.55).reshape(shape)
    if prec == "f32":
        a = a.astype(np.float32)
    p = [1.0, 1.0, 0.005]
    exp13 = O.ch(a, p, 13, mu=O.MU_DW, form=prec)
    for tile in [(16, 64, 16), (8, 32, 16), (8, 64, 32), (4, 128, 64), (32, 64, 12), (16, 64, 22)]:
        for rq in (False, True):
            with mx.Grid(shape, precision=prec) as g:
                g.set_variant(2)
                g.set_tile(*tile)
                g.set_tile_options(rq)
                g.upload(a)
                g.step("ch_dw", p, 13)
                assert np.array_equal(g.download(), exp13), (tile, rq)
    # register queues on 32 x 64 tiles with 768 and 1024 threads: partial
    # tiles in both axes, short last chunks, a grid smaller than one tile;
    # chunks of 2, 5, 7 and 16 planes end the four-plane unrolled march at
    # each of its four steps
    small = (12, 10, 20)
    a_s = O.uniform(int(np.prod(small)), 79, 0.45, 0.55).reshape(small)
    if prec == "f32":
        a_s = a_s.astype(np.float32)
    exp_s = O.ch(a_s, p, 13, mu=O.MU_DW, form=prec)
    for body in (2, 3):
        for chunk in (2, 5, 7, 16):
            for arr, ex in ((a, exp13), (a_s, exp_s)):
                with mx.Grid(arr.shape, precision=prec) as g:
                    g.set_variant(2)
                    g.set_tile(32, 64, chunk)
                    g.set_tile_options(body=body)
                    g.upload(arr)
                    g.step("ch_dw", p, 13)
                    assert np.array_equal(g.download(), ex), (arr.shape, body, chunk)
    # two steps per pass (body 6 on 32 x 64 tiles, 7 on 32 x 32): 13 steps are
    # six two-step passes and one single step; chunks of 2, 5 and 7 planes and
    # every chunk the default picks (6 .. 48: the last covers the whole
    # grid), partial tiles, a grid smaller than one tile
    for body, tile in ((6, (32, 64)), (7, (32, 32))):
        for chunk in (2, 5, 6, 7, 8, 12, 16, 24, 32, 48):
            for arr, ex in ((a, exp13), (a_s, exp_s)):
                with mx.Grid(arr.shape, precision=prec) as g:
                    g.set_variant(2)
                    g.set_tile(tile[0], tile[1], chunk)
                    g.set_tile_options(body=body)
                    g.upload(arr)
                    g.step("ch_dw", p, 13)
                    assert g.get_tile()["opt"] == body
                    assert np.array_equal(g.download(), ex), (arr.shape, body, chunk)
        # the polynomial mu and unequal spacing
        pp = [1.0, 1.0, 1.0, 1.0, 0.002]
        b0 = O.uniform(int(np.prod(shape)), 81, -1.0, 1.0).reshape(shape)
        if prec == "f32":
            b0 = b0.astype(np.float32)
        dx = [1.0, 0.5, 2.0]
        exp_b = O.ch(b0, pp, 10, mu=O.MU_POLY, form=prec, dx=dx)
        with mx.Grid(shape, precision=prec, dx=dx) as g:
            g.set_variant(2)
            g.set_tile(tile[0], tile[1], 16)
            g.set_tile_options(body=body)
            g.upload(b0)
            g.step("ch", pp, 10)
            assert np.array_equal(g.download(), exp_b), (body, "poly")
        # the double well with kappa != 1 (the weighted form of the body)
        pk = [0.8, 1.0, 0.005]
        exp_k = O.ch(a, pk, 9, mu=O.MU_DW, form=prec)
        with mx.Grid(shape, precision=prec) as g:
            g.set_variant(2)
            g.set_tile(tile[0], tile[1], 16)
            g.set_tile_options(body=body)
            g.upload(a)
            g.step("ch_dw", pk, 9)
            assert np.array_equal(g.download(), exp_k), (body, "kappa")
    with mx.Grid(shape, precision=prec) as g:  # the polynomial mu form too
        g.set_variant(2)
        g.set_tile(32, 64, 16)
        g.set_tile_options(body=2)
        pp = [1.0, 1.0, 1.0, 1.0, 0.005]
        b0 = O.uniform(int(np.prod(shape)), 80, -1.0, 1.0).reshape(shape)
        if prec == "f32":
            b0 = b0.astype(np.float32)
        g.upload(b0)
        g.step("ch", pp, 9)
        assert np.array_equal(g.download(), O.ch(b0, pp, 9, mu=O.MU_POLY, form=prec))
    for shape2 in [(40, 300), (30, 64), (17, 5)]:  # 2-D rows, wider and narrower than a tile
        a2 = O.uniform(int(np.prod(shape2)), 78, 0.45, 0.55).reshape(shape2)
        if prec == "f32":
            a2 = a2.astype(np.float32)
        exp2 = O.ch(a2, p, 11, mu=O.MU_DW, form=prec)
        for tz in (64, 128, 256):
            for rq in (False, True):
                with mx.Grid(shape2, precision=prec) as g:
                    g.set_variant(2)
                    g.set_tile(1, tz, 8)
                    g.set_tile_options(rq)
                    g.upload(a2)
                    g.step("ch_dw", p, 11)
                    assert np.array_equal(g.download(), exp2), (shape2, tz, rq)
        # wave strips (60 columns; odd chunks march downward): the eight-row
        # unrolled march ends at every ring position (chunks of 2, 3, 5, 7, 8,
        # 13 and 16 rows, 16 being the FP64 default above 2^21 nodes) and one
        # chunk takes the whole grid
        for chunk in (2, 3, 5, 7, 8, 13, 16, 64):
            with mx.Grid(shape2, precision=prec) as g:
                g.set_variant(2)
                g.set_tile(1, 60, chunk)
                g.upload(a2)
                g.step("ch_dw", p, 11)
                assert np.array_equal(g.download(), exp2), (shape2, chunk)
    exp = O.ch(a, p, 107, mu=O.MU_DW, form=prec)
    with mx.Grid(shape, precision=prec) as g:
        g.set_variant(2)
        g.set_tile_options(True)
        g.upload(a)
        g.step("ch_dw", p, 7)
        g.step("ch_dw", p, 100)  # graph replay starting at odd parity
        assert np.array_equal(g.download(), exp)
    with mx.Grid(shape, precision=prec) as g:  # two-step passes in a captured graph
        g.set_variant(2)
        g.set_tile(32, 32, 16)
        g.set_tile_options(body=7)
        g.upload(a)
        g.step("ch_dw", p, 7)
        g.step("ch_dw", p, 100)
        assert np.array_equal(g.download(), exp)
    with mx.Grid((8, 8, 8), precision=prec) as g:  # noise: the two-step bodies refuse it
        g.set_variant(2)
        g.set_tile(32, 32, 16)
        g.set_tile_options(body=7)
        with pytest.raises(mx.MmspxError):
            g.step("ch_noise", [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 5], 2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_two_step_defaults_on_mid_sizes(prec):
    """Periodic 3-D Cahn-Hilliard below 256^3 takes the two-step bodies with
    the chunk the cost model picks (odd sizes, non-cubic, 6- to 48-plane
    chunks, both tile shapes in FP32); 31 steps (15 passes and one single
    step), bitwise against the oracle."""
    p = [1.0, 1.0, 0.005]
    for shape in [(112, 112, 112), (100, 90, 130), (100, 224, 200), (30, 40, 50)]:
        say("[gpu] two-step default %s %s" % (shape, prec))
        a = O.uniform(int(np.prod(shape)), 83, 0.45, 0.55).reshape(shape)
        if prec == "f32":
            a = a.astype(np.float32)
        with mx.Grid(shape, precision=prec) as g:
            g.upload(a)
            g.step("ch_dw", p, 31)
            t = g.get_tile()
            got = g.download()
        assert t["variant"] == 2 and t["opt"] in (6, 7), t
        assert np.array_equal(got, O.ch(a, p, 31, mu=O.MU_DW, form=prec)), (shape, t)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_fused_ch_non_periodic(prec):
    """Fused Cahn-Hilliard with mirror, Neumann and Dirichlet sides (each axis
    periodic on both sides or on neither): the 3-D register-queue bodies on
    256, 768 and 1024 threads, the 2-D wave strips (chunks ending the
    eight-row march at each position, both directions) and register-queue row
    strips; the double-well and polynomial mu and model B's noise; bitwise
    against the oracle's two passes. An axis periodic on one side only stays
    on the two-pass kernels."""
    p = [1.0, 1.0, 0.005]
    pp = [1.0, 1.0, 1.0, 1.0, 0.005]
    pn = [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 9]
    cases3 = [((20, 36, 70), [(N, N), (M, M), (D, D)]), ((13, 9, 40), [(M, M), (P, P), (N, N)]),
              ((6, 5, 7), [(D, D), (N, N), (M, M)])]
    for shape, bc in cases3:
        say("[gpu] fused CH non-periodic 3-D %s %s %s" % (shape, bc, prec))
        a = O.uniform(int(np.prod(shape)), 1600 + shape[1], 0.45, 0.55).reshape(shape)
        b0 = O.uniform(int(np.prod(shape)), 1610 + shape[1], -1.0, 1.0).reshape(shape)
        if prec != "f64":
            a, b0 = a.astype(np.float32), b0.astype(np.float32)
        exp = O.ch(a, p, 9, mu=O.MU_DW, form=prec, bc=bc)
        expp = O.ch(b0, pp, 9, mu=O.MU_POLY, form=prec, bc=bc)
        expn = O.ch_noise(b0, pn, 9, form=prec, bc=bc)
        runs = [((8, 32, 5), 1), ((16, 64, 7), 1), ((32, 64, 16), 2), ((32, 64, 3), 3), (None, None)]
        for tile, body in runs:
            for model, arr, params, ex in (("ch_dw", a, p, exp), ("ch", b0, pp, expp), ("ch_noise", b0, pn, expn)):
                with mx.Grid(shape, precision=prec, bc=bc) as g:
                    if tile:
                        g.set_variant(2)
                        g.set_tile(*tile)
                        g.set_tile_options(body=body)
                    g.upload(arr)
                    g.step(model, params, 4)
                    g.step(model, params, 5)
                    assert g.get_tile()["variant"] == 2, (shape, tile, model)
                    assert np.array_equal(g.download(), ex), (shape, bc, tile, body, model)
    for shape, bc in [((40, 300), [(N, N), (M, M)]), ((17, 5), [(D, D), (N, N)]), ((33, 64), [(M, M), (P, P)]),
                      ((30, 61), [(P, P), (D, D)])]:
        say("[gpu] fused CH non-periodic 2-D %s %s %s" % (shape, bc, prec))
        a = O.uniform(int(np.prod(shape)), 1700 + shape[1], 0.45, 0.55).reshape(shape)
        if prec != "f64":
            a = a.astype(np.float32)
        exp = O.ch(a, p, 11, mu=O.MU_DW, form=prec, bc=bc)
        expn = O.ch_noise(a, pn, 11, form=prec, bc=bc)
        for tile, body in [((1, 60, c), 0) for c in (2, 3, 5, 7, 8, 13)] + [((1, 64, 8), 1), ((1, 128, 16), 1),
                                                                        (None, None)]:
            with mx.Grid(shape, precision=prec, bc=bc) as g:
                if tile:
                    g.set_variant(2)
                    g.set_tile(*tile)
                    if body:
                        g.set_tile_options(body=body)
                g.upload(a)
                g.step("ch_dw", p, 11)
                assert g.get_tile()["variant"] == 2
                assert np.array_equal(g.download(), exp), (shape, bc, tile)
            with mx.Grid(shape, precision=prec, bc=bc) as g:
                if tile:
                    g.set_variant(2)
                    g.set_tile(*tile)
                    if body:
                        g.set_tile_options(body=body)
                g.upload(a)
                g.step("ch_noise", pn, 11)
                assert np.array_equal(g.download(), expn), (shape, bc, tile, "noise")
    # periodic on one side only: the two passes
    shape, bc = (12, 14, 16), [(P, N), (P, P), (P, P)]
    a = O.uniform(int(np.prod(shape)), 1800, 0.45, 0.55).reshape(shape)
    if prec != "f64":
        a = a.astype(np.float32)
    with mx.Grid(shape, precision=prec, bc=bc) as g:
        g.upload(a)
        g.step("ch_dw", p, 3)
        assert g.get_tile()["variant"] == 1
        assert np.array_equal(g.download(), O.ch(a, p, 3, mu=O.MU_DW, form=prec, bc=bc))


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_grain_growth_2d_tiles(prec):
    """The 2-D grain-growth tiles (8x64 and 16x64 cells), partial tiles and
    every boundary condition, against the oracle bitwise."""
    cases = [((70, 300), 4, [(N, M), (D, P)]), ((40, 150), 32, None), ((9, 130), 8, [(M, M), (N, D)]),
             ((33, 64), 2, [(P, P), (M, N)])]
    for shape, q, bc in cases:
        say("[gpu] gg tile %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 300 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec != "f64":
            a = a.astype(np.float32)
        exp = O.gg(a, 0.01, 9, form=prec, bc=bc)
        for tx in (8, 16):
            with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                g.set_variant(2)
                g.set_tile(tx, 64, tx)
                g.upload(a)
                g.step("gg", [0.01], 9)
                assert np.array_equal(g.download(), exp), (shape, q, tx)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_wave_strips(prec):
    """The 64-column wave strips (2-D Allen-Cahn and grain growth): partial
    strips, grids narrower than a strip, chunks that leave a short last chunk,
    every boundary condition; bitwise against the oracle."""
    ac_cases = [((70, 300), [(N, M), (D, P)]), ((40, 64), None), ((9, 5), [(M, M), (N, D)]), ((33, 129), [(P, P), (M, N)]),
                ((20, 128), [(M, N), (N, M)])]
    for shape, bc in ac_cases:
        say("[gpu] wave strips ac %s %s" % (shape, prec))
        a = O.uniform(int(np.prod(shape)), 400, -1.0, 1.0).reshape(shape)
        if prec != "f64":
            a = a.astype(np.float32)
        p = [1.0, 1.0, 1.0, 1.0, 0.01]
        exp = O.ac(a, p, 11, form=prec, bc=bc)
        # chunks of 7, 8 and 24 rows (and the short last ones) end the
        # four-row unrolled march at each of its four rows
        for chunk in (7, 8, 24):
            with mx.Grid(shape, precision=prec, bc=bc) as g:
                g.set_variant(2)
                g.set_tile(1, 64, chunk)
                g.set_tile_options(body=2)  # wave strips
                g.upload(a)
                g.step("ac", p, 11)
                assert np.array_equal(g.download(), exp), (shape, chunk)
    gg_cases = [((70, 300), 4, [(N, M), (D, P)]), ((40, 150), 16, None), ((9, 130), 8, [(M, M), (N, D)]),
                ((33, 64), 2, [(P, P), (M, N)])]
    if prec != "f64":
        gg_cases.append(((37, 200), 32, [(P, P), (N, P)]))
    for shape, q, bc in gg_cases:
        say("[gpu] wave strips gg %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 500 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec != "f64":
            a = a.astype(np.float32)
        exp = O.gg(a, 0.01, 9, form=prec, bc=bc)
        # body 2: wave strips; bodies 4 and 5: split strips on 2 and 4 waves (8, 16, 32 fields)
        for body in ((2, 4, 5) if q in (8, 16, 32) else (2,)):
            for chunk in (7, 8, 16):
                with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                    g.set_variant(2)
                    g.set_tile(1, 64, chunk)
                    g.set_tile_options(body=body)
                    g.upload(a)
                    g.step("gg", [0.01], 9)
                    assert np.array_equal(g.download(), exp), (shape, q, chunk, body)
    if prec == "f64":  # 32 FP64 fields: the split strips hold them, the wave strips are refused
        q, shape = 32, (23, 70)
        a = O.uniform(int(np.prod(shape)) * q, 532, 0.0, 1.0).reshape(shape + (q,))
        with mx.Grid(shape, fields=q, vector=True, precision="f64") as g:
            g.set_variant(2)
            g.set_tile(1, 64, 5)
            g.set_tile_options(body=4)
            g.upload(a)
            g.step("gg", [0.01], 6)
            exp = O.gg(a, 0.01, 6, form="f64")
            assert np.array_equal(g.download(), exp)
            g.set_tile_options(body=5)
            g.set_tile(1, 64, 7)
            g.upload(a)
            g.step("gg", [0.01], 6)
            assert np.array_equal(g.download(), exp)
        with mx.Grid((16, 70), fields=32, vector=True, precision="f64") as g:
            g.set_variant(2)
            g.set_tile(1, 64, 8)
            g.set_tile_options(body=2)
            g.upload(np.zeros((16, 70, 32)))
            with pytest.raises(mx.MmspxError):
                g.step("gg", [0.01], 1)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_ac_temporal_blocks(prec):
    """Periodic Allen-Cahn temporal blocks (8, 16 or 32 rows x 48 columns, up to
    8 steps per pass): partial tiles, a grid smaller than one tile, step counts
    that leave a short last pass, and graph-replayed runs; bitwise against the
    oracle's single steps."""
    p = [1.0, 1.0, 1.0, 1.0, 0.01]
    for shape in [(64, 64), (70, 130), (9, 5), (512, 512)]:
        say("[gpu] temporal blocks %s %s" % (shape, prec))
        a = O.uniform(int(np.prod(shape)), 600, -1.0, 1.0).reshape(shape)
        if prec != "f64":
            a = a.astype(np.float32)
        steps = 13 if shape[0] < 512 else 107
        exp = O.ac(a, p, steps, form=prec)
        for tx in (8, 16, 32):
            for k in (4, 8):
                with mx.Grid(shape, precision=prec) as g:
                    g.set_variant(2)
                    g.set_tile(tx, 48, k)
                    g.upload(a)
                    g.step("ac", p, steps)
                    assert np.array_equal(g.download(), exp), (shape, tx, k)
                    assert g.get_tile()["opt"] == 3
    with mx.Grid((40, 50), precision=prec, bc=[(N, N), (P, P)]) as g:  # periodic boundaries only
        g.set_variant(2)
        g.set_tile(16, 48, 8)
        g.upload(np.zeros((40, 50)))
        with pytest.raises(mx.MmspxError):
            g.step("ac", p, 8)


def test_upload_download_layouts():
    say("[gpu] layouts")
    a = O.uniform(7 * 6 * 5 * 3, 4, 0, 1).reshape(7, 6, 5, 3)
    with mx.Grid((7, 6, 5), fields=3, vector=True) as g:
        g.upload(a)
        assert np.array_equal(g.download(), a)
        fm = g.download(layout=mx.FIELD_MAJOR)
        assert np.array_equal(fm, np.moveaxis(a, -1, 0))
        g.upload(fm, layout=mx.FIELD_MAJOR)
        assert np.array_equal(g.download(), a)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_reductions_match_oracle(prec):
    say("[gpu] reductions %s" % prec)
    rng_a = O.uniform(48 * 40, 3, 0.3, 0.7).reshape(48, 40)
    v = O.uniform(20 * 22 * 4, 5, 0.0, 1.0).reshape(20, 22, 4)
    bc = [(N, M), (P, D)]
    if prec == "f32":
        rng_a, v = rng_a.astype(np.float32), v.astype(np.float32)
    with mx.Grid((48, 40), precision=prec, bc=bc, dx=[0.5, 2.0]) as g:
        g.upload(rng_a)
        assert abs(g.mass()[0] - O.mass(rng_a)[0]) <= 1e-15 * abs(O.mass(rng_a)[0])
        for model, params, kind, ep in [("ac", [1.0, 1.0, 1.7, 1.0, 0.01], O.E_POLY, [1.0, 1.0, 1.7]),
                                        ("ch_dw", [0.8, 1.0, 0.005], O.E_DW, [0.8])]:
            e = g.energy(model, params)
            r = O.energy(rng_a, kind, ep, bc=bc, dx=[0.5, 2.0])
            assert abs(e - r) <= 1e-13 * abs(r), (model, e, r)
    with mx.Grid((20, 22), fields=4, vector=True, precision=prec) as g:
        g.upload(v)
        assert np.allclose(g.mass(), O.mass(v, vector=True), rtol=1e-15, atol=0)
        e = g.energy("gg", [0.05])
        r = O.energy(v, O.E_GG, [2.0], vector=True)
        assert abs(e - r) <= 1e-13 * abs(r)


def test_c1_allen_cahn_full_size_bitwise():
    """512^2 FP64, 1000 steps, phi0 ~ U(-0.01, 0.01) from seed 20260101."""
    say("[gpu] C1 512^2 AC 1000 steps")
    a = O.uniform(512 * 512, 20260101, -0.01, 0.01).reshape(512, 512)
    with mx.Grid((512, 512)) as g:
        g.upload(a)
        g.step("ac", [1.0, 1.0, 1.0, 1.0, 0.01], 1000)
        got = g.download()
    exp = O.ac(a, [1.0, 1.0, 1.0, 1.0, 0.01], 1000)
    assert np.array_equal(got, exp)


def test_c2_cahn_hilliard_full_size_mass():
    """256^3 FP64, 2000 steps: mass conserved to 1e-12, state finite and bounded,
    generic two-pass and fused paths agree bitwise after 50 steps."""
    say("[gpu] C2 256^3 CH mass")
    n = 256
    a = O.uniform(n ** 3, 20260102, 0.45, 0.55).reshape(n, n, n)
    p = [1.0, 1.0, 0.005]
    with mx.Grid((n, n, n)) as g:
        g.upload(a)
        m0 = g.mass()[0]
        g.step("ch_dw", p, 2000)
        m1 = g.mass()[0]
        got = g.download()
    assert abs(m1 - m0) / abs(m0) <= 1e-12, (m0, m1)
    assert np.all(np.isfinite(got)) and got.min() > -0.5 and got.max() < 1.5
    outs = []
    for variant in (1, 2):
        with mx.Grid((n, n, n)) as g:
            g.set_variant(variant)
            g.upload(a)
            g.step("ch_dw", p, 50)
            outs.append(g.download())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_split_strips_two_steps(prec):
    """Grain growth and Zener pinning (periodic) on the split strips two steps
    per pass: 8, 16 and 32 fields on 2 and 4 waves (bodies 9 and 10), 2, 4 and
    8 fields on one wave (body 11):
    partial strips, grids narrower than a strip and shorter than a chunk,
    chunks of 2, 3, 5, 8, 13 and 64 rows (the four-row march ends at each
    position), odd step counts (the last step on the one-step strips) and a
    graph-replayed run; bitwise against the oracle."""
    cases = [((41, 130), 8), ((33, 64), 16), ((9, 50), 8), ((5, 7), 16), ((70, 300), 32), ((41, 130), 2),
             ((9, 50), 4), ((5, 7), 2)]
    for shape, q in cases:
        if q == 32 and prec == "f32ref":
            continue
        a = O.uniform(int(np.prod(shape)) * q, 540 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec != "f64":
            a = a.astype(np.float32)
        for steps in (9, 10):
            say("[gpu] two-step split strips %s q=%d %s steps %d" % (shape, q, prec, steps))
            exp = O.gg(a, 0.01, steps, form=prec)
            # Zener pinning: FP64 and FP32 forms, 32 fields in FP32 only (as its other fast paths)
            expz = O.zener(a, 0.01, steps, form=prec) if prec != "f32ref" and (q < 32 or prec == "f32") else None
            bodies = ((9, 10) if q > 4 else ()) + ((11,) if q <= 8 else ())
            for body in bodies:
                for chunk in ((2, 3, 5, 8, 13, 64) if steps == 9 else (4, 7)):
                    with mx.Grid(shape, fields=q, vector=True, precision=prec) as g:
                        g.set_variant(2)
                        g.set_tile(1, 60, chunk)
                        g.set_tile_options(body=body)
                        g.upload(a)
                        g.step("gg", [0.01], steps)
                        assert g.get_tile()["opt"] == body
                        assert np.array_equal(g.download(), exp), (shape, q, body, chunk, steps)
                    if expz is not None and chunk in (3, 7, 64):
                        with mx.Grid(shape, fields=q, vector=True, precision=prec) as g:
                            g.set_variant(2)
                            g.set_tile(1, 60, chunk)
                            g.set_tile_options(body=body)
                            g.upload(a)
                            g.step("zener", [0.01], steps)
                            assert np.array_equal(g.download(), expz), (shape, q, body, chunk, steps, "zener")
    # a captured graph of two-step passes (100 steps on a small grid), from odd parity
    shape, q = (33, 64), 8
    a = O.uniform(int(np.prod(shape)) * q, 560, 0.0, 1.0).reshape(shape + (q,))
    if prec != "f64":
        a = a.astype(np.float32)
    with mx.Grid(shape, fields=q, vector=True, precision=prec) as g:
        g.set_variant(2)
        g.set_tile(1, 60, 8)
        g.set_tile_options(body=10)
        g.upload(a)
        g.step("gg", [0.01], 3)
        g.step("gg", [0.01], 100)
        assert np.array_equal(g.download(), O.gg(a, 0.01, 103, form=prec))
    # non-periodic grids are refused
    with mx.Grid((20, 70), fields=8, vector=True, precision=prec, bc=[(N, N), (P, P)]) as g:
        g.set_variant(2)
        g.set_tile(1, 60, 8)
        g.set_tile_options(body=9)
        g.upload(np.zeros((20, 70, 8)))
        with pytest.raises(mx.MmspxError):
            g.step("gg", [0.01], 2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_zener_and_ostwald(prec):
    """Zener pinning on the generic kernel, the wave strips and the split
    strips (every compiled field count, partial strips, several boundaries),
    Ostwald ripening on the
    generic two-pass kernels in 1-D, 2-D and 3-D: bitwise against the oracle."""
    zcases = [((70, 300), 4, [(N, M), (D, P)]), ((40, 130), 2, None), ((9, 70), 8, [(M, M), (N, D)]),
              ((33, 64), 16, [(P, P), (M, N)])]
    if prec == "f32":
        zcases.append(((20, 100), 32, [(P, P), (N, P)]))
    for shape, q, bc in zcases:
        say("[gpu] zener %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 600 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec == "f32":
            a = a.astype(np.float32)
        exp = O.zener(a, 0.01, 9, form=prec, bc=bc)
        runs = [(1, None, 0), (2, (1, 64, 7), 2), (2, (1, 64, 8), 2), (2, (1, 64, 24), 2), (0, None, 0)]
        if q in (8, 16, 32):
            runs += [(2, (1, 64, 6), 4), (2, (1, 64, 9), 4), (2, (1, 64, 5), 5)]  # the split strips
        for variant, tile, body in runs:
            with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                g.set_variant(variant)
                if tile:
                    g.set_tile(*tile)
                    g.set_tile_options(body=body)
                g.upload(a)
                g.step("zener", [0.01], 9)
                assert np.array_equal(g.download(), exp), (shape, q, variant, tile, body)
    p = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
    for shape, q, bc in [((90,), 2, [(N, D)]), ((40, 36), 3, None), ((10, 9, 12), 4, [(M, N), (P, P), (D, P)])]:
        say("[gpu] ostwald %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 700 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec == "f32":
            a = a.astype(np.float32)
        exp = O.ostwald(a, p, 12, form=prec, bc=bc)
        with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
            g.upload(a)
            g.step("ostwald", p, 5)
            g.step("ostwald", p, 7)
            assert np.array_equal(g.download(), exp), (shape, q)
    with mx.Grid((8, 8), fields=3, vector=True, precision="f32ref") as g:
        with pytest.raises(mx.MmspxError) as e:
            g.step("ostwald", p, 1)
        assert e.value.code == -8
    with mx.Grid((8, 8), fields=1, vector=True, precision=prec) as g:
        with pytest.raises(mx.MmspxError):
            g.step("ostwald", p, 1)  # no order parameter


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_multiphase_grain_growth(prec):
    """Multi-phase grain growth from one-hot grains: bitwise against the oracle
    in 1-D, 2-D and 3-D, several boundaries, up to 32 fields."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    p = [0.01, 8.0, 1.0, 1.0]
    for shape, q, bc in [((300,), 3, [(N, D)]), ((70, 90), 6, [(N, M), (D, P)]), ((64, 64), 32, None),
                         ((10, 12, 14), 4, [(M, N), (P, P), (D, P)])]:
        say("[gpu] multi-phase grain growth %s q=%d %s" % (shape, q, prec))
        a = mg.one_hot_grains(shape, q, 900 + q)
        if prec == "f32":
            a = a.astype(np.float32)
        exp = O.gg_aniso(a, p, 25, form=prec, bc=bc)
        with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
            g.upload(a)
            g.step("gg_aniso", p, 10)
            g.step("gg_aniso", p, 15)
            assert np.array_equal(g.download(), exp), (shape, q)
    with mx.Grid((8, 8), fields=33, vector=True, precision=prec) as g:
        with pytest.raises(mx.MmspxError):
            g.step("gg_aniso", p, 1)  # more than 32 fields


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ostwald_wave_strips(prec):
    """The 2-D Ostwald wave strips (periodic): partial strips, grids narrower
    than a strip, chunks of 2, 3, 8 and 64 rows, 2 to 8 fields; bitwise
    against the oracle (and so against the generic two-pass kernels)."""
    p = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
    for shape, q in [((40, 300), 4), ((17, 5), 2), ((70, 130), 3), ((33, 64), 8), ((9, 61), 5)]:
        say("[gpu] ostwald wave strips %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 800 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec == "f32":
            a = a.astype(np.float32)
        exp = O.ostwald(a, p, 9, form=prec)
        for chunk in (2, 3, 8, 64):
            with mx.Grid(shape, fields=q, vector=True, precision=prec) as g:
                g.set_variant(2)
                g.set_tile(1, 60, chunk)
                g.upload(a)
                g.step("ostwald", p, 9)
                assert np.array_equal(g.download(), exp), (shape, q, chunk)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_multiphase_wave_strips(prec):
    """The 2-D multi-phase wave strips (2, 4 and 8 fields, any boundary):
    one-hot grains and sparse random fields (every active-set size), partial
    strips, grids narrower than a strip, chunks of 2 to 32 rows; bitwise
    against the oracle, and the default choice against the generic kernel."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    p = [0.01, 8.0, 1.0, 1.0]
    cases = [((40, 300), 4, None, "grains"), ((17, 5), 2, [(N, D), (M, P)], "sparse"),
             ((70, 130), 8, [(P, P), (N, M)], "sparse"), ((33, 64), 8, None, "grains"),
             ((9, 61), 2, [(D, N), (P, D)], "grains"), ((50, 129), 4, [(M, M), (D, D)], "sparse"),
             ((40, 200), 8, [(P, P), (P, P)], "thin")]
    for shape, q, bc, kind in cases:
        n = int(np.prod(shape))
        if kind == "grains":
            a = mg.one_hot_grains(shape, q, 700 + q)
        else:
            v = O.uniform(n * q, 710 + q, 0.0, 1.0).reshape(shape + (q,))
            # "thin": 2 to 7 fields present at most nodes (both paths of the 8-field kernels)
            keep = O.uniform(n * q, 720 + q, 0.0, 1.0).reshape(shape + (q,)) < (0.1 if kind == "thin" else 0.35)
            a = np.where(keep, v, 0.0)
        if prec == "f32":
            a = a.astype(np.float32)
        exp = O.gg_aniso(a, p, 9, form=prec, bc=bc)
        for chunk in (2, 3, 8, 32):
            say("[gpu] multi-phase wave strips %s q=%d %s chunk %d %s" % (shape, q, kind, chunk, prec))
            with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                g.set_variant(2)
                g.set_tile(1, 64, chunk)
                g.set_tile_options(body=2)
                g.upload(a)
                g.step("gg_aniso", p, 9)
                assert np.array_equal(g.download(), exp), (shape, q, chunk)
        with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
            g.upload(a)
            g.step("gg_aniso", p, 4)
            g.step("gg_aniso", p, 5)
            assert np.array_equal(g.download(), exp), (shape, q, "default")
        with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
            g.set_variant(1)
            g.upload(a)
            g.step("gg_aniso", p, 9)
            assert np.array_equal(g.download(), exp), (shape, q, "generic")
    a = mg.one_hot_grains((20, 70), 6, 7)
    if prec == "f32":
        a = a.astype(np.float32)
    with mx.Grid((20, 70), fields=6, vector=True, precision=prec) as g:
        g.upload(a)
        g.set_variant(2)  # 6 fields: no strips compiled
        with pytest.raises(mx.MmspxError):
            g.step("gg_aniso", p, 1)
        g.set_variant(0)  # the default falls back to the generic kernel
        g.step("gg_aniso", p, 2)
        assert np.array_equal(g.download(), O.gg_aniso(a, p, 2, form=prec))


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_march_3d(prec):
    """The 3-D plane march (tile body 8): Allen-Cahn, model A, grain growth,
    Zener pinning and multi-phase grain growth with 2, 4 and 8 fields, Ostwald
    ripening with 2, 3 and 4 (two marches), every compiled tile, chunks of 2,
    3, 5, 7 and 16 planes (the four-plane unrolled march ends at each of its
    positions), partial tiles, grids smaller than one tile, every boundary
    condition on every axis; bitwise against the oracle."""
    bcs = [None, [(N, M), (D, P), (M, N)], [(P, D), (N, N), (D, M)], [(M, M), (P, P), (N, D)]]
    ac_tiles = [(8, 64), (16, 64), (32, 64), (16, 32)]
    vec_tiles = [(8, 64), (16, 32)]
    pa = [1.0, 1.0, 1.3, 1.0, 0.01]
    pn = [1.0, 1.0, 1.0, 1.0, 0.01, 0.4, 17]
    for shape, bc in [((13, 20, 70), bcs[1]), ((9, 40, 33), bcs[2]), ((6, 5, 7), bcs[3]), ((12, 18, 66), bcs[0])]:
        say("[gpu] march 3-D ac %s %s %s" % (shape, bc, prec))
        a = O.uniform(int(np.prod(shape)), 1300 + shape[2], -1.0, 1.0).reshape(shape)
        if prec != "f64":
            a = a.astype(np.float32)
        exp = O.ac(a, pa, 7, form=prec, bc=bc)
        expn = O.ac_noise(a, pn, 7, form=prec, bc=bc)
        for ti, tile in enumerate(ac_tiles):
            for chunk in ((2, 3, 5, 7, 16) if ti == 0 else (3, 16)):
                with mx.Grid(shape, precision=prec, bc=bc) as g:
                    g.set_variant(2)
                    g.set_tile(tile[0], tile[1], chunk)
                    g.set_tile_options(body=8)
                    g.upload(a)
                    g.step("ac", pa, 7)
                    assert np.array_equal(g.download(), exp), (shape, tile, chunk)
            with mx.Grid(shape, precision=prec, bc=bc) as g:
                g.set_variant(2)
                g.set_tile(tile[0], tile[1], 5)
                g.set_tile_options(body=8)
                g.upload(a)
                g.step("ac_noise", pn, 3)
                g.step("ac_noise", pn, 4)
                assert np.array_equal(g.download(), expn), (shape, tile, "noise")
    for shape, q, bc in [((11, 20, 70), 2, bcs[1]), ((9, 17, 33), 4, bcs[2]), ((10, 12, 40), 8, bcs[3]),
                         ((5, 6, 7), 8, bcs[0])]:
        say("[gpu] march 3-D gg/zener %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 1400 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec != "f64":
            a = a.astype(np.float32)
        exp = O.gg(a, 0.01, 7, form=prec, bc=bc)
        expz = O.zener(a, 0.01, 7, form=prec, bc=bc) if prec != "f32ref" else None
        for tile in vec_tiles:
            for chunk in (2, 3, 5, 16):
                with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                    g.set_variant(2)
                    g.set_tile(tile[0], tile[1], chunk)
                    g.set_tile_options(body=8)
                    g.upload(a)
                    g.step("gg", [0.01], 7)
                    assert np.array_equal(g.download(), exp), (shape, q, tile, chunk)
                if expz is not None:
                    with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                        g.set_variant(2)
                        g.set_tile(tile[0], tile[1], chunk)
                        g.set_tile_options(body=8)
                        g.upload(a)
                        g.step("zener", [0.01], 7)
                        assert np.array_equal(g.download(), expz), (shape, q, tile, chunk, "zener")
    # multi-phase grain growth: one-hot grains and sparse random fields (every
    # active-set size, both paths of the 8-field kernels)
    if prec != "f32ref":
        import importlib.util

        spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                                  "make_golden.py"))
        mg = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mg)
        pg = [0.01, 8.0, 1.0, 1.0]
        for shape, q, bc, kind in [((11, 20, 70), 2, bcs[1], "sparse"), ((9, 17, 33), 4, bcs[2], "grains"),
                                   ((10, 12, 40), 8, bcs[3], "sparse"), ((12, 18, 66), 8, bcs[0], "grains"),
                                   ((5, 6, 7), 4, bcs[1], "sparse")]:
            say("[gpu] march 3-D gg_aniso %s q=%d %s %s" % (shape, q, kind, prec))
            n = int(np.prod(shape))
            if kind == "grains":
                a = mg.one_hot_grains(shape, q, 1600 + q)
            else:
                v = O.uniform(n * q, 1610 + q, 0.0, 1.0).reshape(shape + (q,))
                keep = O.uniform(n * q, 1620 + q, 0.0, 1.0).reshape(shape + (q,)) < 0.3
                a = np.where(keep, v, 0.0)
            if prec != "f64":
                a = a.astype(np.float32)
            exp = O.gg_aniso(a, pg, 7, form=prec, bc=bc)
            for tile in vec_tiles:
                for chunk in (2, 3, 5, 16):
                    with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                        g.set_variant(2)
                        g.set_tile(tile[0], tile[1], chunk)
                        g.set_tile_options(body=8)
                        g.upload(a)
                        g.step("gg_aniso", pg, 7)
                        assert np.array_equal(g.download(), exp), (shape, q, tile, chunk, kind)
            with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                g.upload(a)
                g.step("gg_aniso", pg, 3)
                g.step("gg_aniso", pg, 4)
                assert np.array_equal(g.download(), exp), (shape, q, "default", kind)
        # the default takes the march from 2^21 nodes
        shape = (130, 128, 128)
        a = mg.one_hot_grains(shape, 4, 1630)
        if prec != "f64":
            a = a.astype(np.float32)
        with mx.Grid(shape, fields=4, vector=True, precision=prec) as g:
            g.upload(a)
            g.step("gg_aniso", pg, 3)
            assert g.get_tile()["opt"] == 8
            assert np.array_equal(g.download(), O.gg_aniso(a, pg, 3, form=prec))
    # Ostwald ripening (2, 3 and 4 fields): the two passes as marches
    if prec != "f32ref":
        po = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
        for shape, q, bc in [((11, 20, 70), 2, bcs[1]), ((9, 17, 33), 3, bcs[2]), ((10, 12, 40), 4, bcs[3]),
                             ((12, 18, 66), 4, bcs[0]), ((5, 6, 7), 2, bcs[0])]:
            say("[gpu] march 3-D ostwald %s q=%d %s" % (shape, q, prec))
            a = O.uniform(int(np.prod(shape)) * q, 1700 + q, 0.0, 1.0).reshape(shape + (q,))
            if prec != "f64":
                a = a.astype(np.float32)
            exp = O.ostwald(a, po, 7, form=prec, bc=bc)
            for tile in vec_tiles:
                for chunk in (2, 3, 5, 16):
                    with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                        g.set_variant(2)
                        g.set_tile(tile[0], tile[1], chunk)
                        g.set_tile_options(body=8)
                        g.upload(a)
                        g.step("ostwald", po, 7)
                        assert np.array_equal(g.download(), exp), (shape, q, tile, chunk, "ostwald")
            with mx.Grid(shape, fields=q, vector=True, precision=prec, bc=bc) as g:
                g.upload(a)
                g.step("ostwald", po, 3)
                g.step("ostwald", po, 4)
                assert np.array_equal(g.download(), exp), (shape, q, "default", "ostwald")
        # the default takes the marches from 2^20 nodes
        shape = (104, 100, 101)
        a = O.uniform(int(np.prod(shape)) * 2, 1750, 0.0, 1.0).reshape(shape + (2,))
        if prec != "f64":
            a = a.astype(np.float32)
        with mx.Grid(shape, fields=2, vector=True, precision=prec) as g:
            g.upload(a)
            g.step("ostwald", po, 3)
            assert g.get_tile()["opt"] == 8
            assert np.array_equal(g.download(), O.ostwald(a, po, 3, form=prec)), "ostwald default"
    # the defaults take the march; other field counts stay on the generic kernel
    a = O.uniform(10 * 12 * 14 * 3, 1500, 0.0, 1.0).reshape(10, 12, 14, 3)
    if prec != "f64":
        a = a.astype(np.float32)
    with mx.Grid((10, 12, 14), fields=3, vector=True, precision=prec) as g:
        g.upload(a)
        g.step("gg", [0.01], 4)
        assert g.get_tile()["variant"] == 1
        assert np.array_equal(g.download(), O.gg(a, 0.01, 4, form=prec))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_march_3d_two_steps(prec):
    """The 3-D plane march two steps per pass (tile body 12, periodic grids):
    Allen-Cahn on its three tiles, grain growth and Zener pinning with 2, 4
    and 8 fields on theirs, chunks of 2, 3, 5 and 16 planes (the four-plane
    unrolled march ends at each of its positions), partial tiles, grids
    smaller than one tile and thinner than the halo, odd step counts (the last
    step on the one-step march) and a graph-replayed run; Ostwald ripening
    with 2, 3 and 4 fields on the one-pass march (w kept in LDS); the
    defaults; bitwise against the oracle. Tiles whose planes do not fit the
    LDS, non-periodic grids and the F32REF form are refused."""
    pa = [1.0, 1.0, 1.3, 1.0, 0.01]
    for shape in [(13, 20, 70), (9, 40, 33), (6, 5, 7), (12, 18, 66), (3, 2, 130), (1, 9, 33), (2, 20, 70)]:
        a = O.uniform(int(np.prod(shape)), 1600 + shape[2], -1.0, 1.0).reshape(shape)
        if prec != "f64":
            a = a.astype(np.float32)
        for steps in (7, 8):
            say("[gpu] two-step march ac %s %s steps %d" % (shape, prec, steps))
            exp = O.ac(a, pa, steps, form=prec)
            for ti, tile in enumerate([(16, 32), (16, 64), (32, 64)]):
                for chunk in ((2, 3, 5, 16) if ti == 0 else (3, 16)):
                    with mx.Grid(shape, precision=prec) as g:
                        g.set_variant(2)
                        g.set_tile(tile[0], tile[1], chunk)
                        g.set_tile_options(body=12)
                        g.upload(a)
                        g.step("ac", pa, steps)
                        assert g.get_tile()["opt"] == 12
                        assert np.array_equal(g.download(), exp), (shape, tile, chunk, steps)
    for shape, q in [((11, 20, 70), 2), ((9, 17, 33), 4), ((10, 12, 40), 8), ((5, 6, 7), 8), ((4, 33, 65), 4),
                     ((1, 20, 70), 2), ((2, 9, 33), 8)]:
        say("[gpu] two-step march gg/zener %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 1700 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec != "f64":
            a = a.astype(np.float32)
        exp = O.gg(a, 0.01, 7, form=prec)
        expz = O.zener(a, 0.01, 7, form=prec)
        for tile in [(8, 64), (16, 32), (8, 32)]:
            fits = not (prec == "f64" and q == 8 and tile != (8, 32))
            for chunk in (2, 3, 5, 16):
                for model, want in (("gg", exp), ("zener", expz)):
                    if model == "zener" and chunk not in (3, 16):
                        continue
                    with mx.Grid(shape, fields=q, vector=True, precision=prec) as g:
                        g.set_variant(2)
                        g.set_tile(tile[0], tile[1], chunk)
                        g.set_tile_options(body=12)
                        g.upload(a)
                        if not fits:
                            with pytest.raises(mx.MmspxError):
                                g.step(model, [0.01], 2)
                            continue
                        g.step(model, [0.01], 7)
                        assert np.array_equal(g.download(), want), (shape, q, tile, chunk, model)
    # a captured graph of two-step passes (100 steps), from odd parity
    shape, q = (10, 20, 36), 4
    a = O.uniform(int(np.prod(shape)) * q, 1750, 0.0, 1.0).reshape(shape + (q,))
    if prec != "f64":
        a = a.astype(np.float32)
    with mx.Grid(shape, fields=q, vector=True, precision=prec) as g:
        g.set_variant(2)
        g.set_tile(16, 32, 8)
        g.set_tile_options(body=12)
        g.upload(a)
        g.step("gg", [0.01], 3)
        g.step("gg", [0.01], 100)
        assert np.array_equal(g.download(), O.gg(a, 0.01, 103, form=prec))
    # Ostwald ripening in one pass (body 12: w kept in LDS; 2, 3 and 4 fields)
    po = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
    for shape, q in [((11, 20, 70), 2), ((9, 17, 33), 3), ((10, 12, 40), 4), ((5, 6, 7), 2), ((4, 33, 65), 3),
                     ((1, 12, 40), 2), ((2, 9, 70), 4)]:
        say("[gpu] one-pass march ostwald %s q=%d %s" % (shape, q, prec))
        a = O.uniform(int(np.prod(shape)) * q, 1780 + q, 0.0, 1.0).reshape(shape + (q,))
        if prec != "f64":
            a = a.astype(np.float32)
        exp = O.ostwald(a, po, 7, form=prec)
        for tile in [(8, 64), (16, 32)]:
            for chunk in (2, 3, 5, 16):
                with mx.Grid(shape, fields=q, vector=True, precision=prec) as g:
                    g.set_variant(2)
                    g.set_tile(tile[0], tile[1], chunk)
                    g.set_tile_options(body=12)
                    g.upload(a)
                    g.step("ostwald", po, 7)
                    assert g.get_tile()["opt"] == 12
                    assert np.array_equal(g.download(), exp), (shape, q, tile, chunk, "ostwald")
    # the defaults: the two-step march where 64-node tile rows fit the last
    # axis (Allen-Cahn; 2 and 4 fields), the one-step march elsewhere
    # (Ostwald ripening with 2 fields from 2^20 nodes: the one-pass march)
    for shape, q, model, body in [((70, 40, 64), 1, "ac", 12), ((40, 24, 128), 4, "gg", 12),
                                  ((30, 20, 120), 2, "zener", 12), ((40, 24, 96), 1, "ac", 8),
                                  ((30, 20, 64), 8, "gg", 8), ((64, 128, 128), 2, "ostwald", 12)]:
        say("[gpu] two-step march default %s q=%d %s %s" % (shape, q, model, prec))
        vec = q > 1
        a = O.uniform(int(np.prod(shape)) * q, 1760 + q, 0.0 if vec else -1.0, 1.0)
        a = a.reshape(shape + ((q,) if vec else ()))
        if prec != "f64":
            a = a.astype(np.float32)
        params = po if model == "ostwald" else [0.01] if vec else pa
        exp = oracle_run(model, a, params, 5, prec)
        with mx.Grid(shape, fields=q, vector=vec, precision=prec) as g:
            g.upload(a)
            g.step(model, params, 5)
            t = g.get_tile()
            assert (t["variant"], t["opt"]) == (2, body), (shape, t)
            assert np.array_equal(g.download(), exp), (shape, q, model, "default")
    # non-periodic grids and the F32REF form are refused
    with mx.Grid((6, 20, 40), precision=prec, bc=[(N, N), (P, P), (P, P)]) as g:
        g.set_variant(2)
        g.set_tile(16, 32, 8)
        g.set_tile_options(body=12)
        g.upload(np.zeros((6, 20, 40)))
        with pytest.raises(mx.MmspxError):
            g.step("ac", pa, 2)
    for steps in (1, 2):
        with mx.Grid((6, 20, 40), precision="f32ref") as g:
            g.set_variant(2)
            g.set_tile(16, 32, 8)
            g.set_tile_options(body=12)
            g.upload(np.zeros((6, 20, 40), dtype=np.float32))
            with pytest.raises(mx.MmspxError):
                g.step("ac", pa, steps)
    with mx.Grid((6, 20, 40), fields=2, vector=True, precision=prec, bc=[(P, P), (N, N), (P, P)]) as g:
        g.set_variant(2)
        g.set_tile(8, 64, 8)
        g.set_tile_options(body=12)
        g.upload(np.zeros((6, 20, 40, 2)))
        with pytest.raises(mx.MmspxError):
            g.step("ostwald", po, 1)


def test_graph_replay_follows_new_parameters():
    """Runs of 100 steps replay a captured graph whose launches bake the
    model's constants in; a second run with other constants on the same grid
    must capture again. Ostwald ripening and multi-phase grain growth (their
    constants live outside the Allen-Cahn/Cahn-Hilliard fields), and Allen-Cahn
    with another K: set A for 100 steps, then set B for 100, bitwise against
    the oracle."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    ow_a = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
    ow_b = [0.002, 0.5, 2.0, 0.1, 0.9, 1.5, 0.5, 2.0, 1.0]
    ga_a = [0.01, 8.0, 1.0, 1.0]
    ga_b = [0.02, 6.0, 0.5, 2.0]
    ac_a = [1.0, 1.0, 1.0, 1.0, 0.01]
    ac_b = [1.0, 1.0, 1.6, 1.0, 0.01]
    for prec in ("f64", "f32"):
        cases = [("ostwald", (40, 36), 3, ow_a, ow_b, None), ("ostwald", (10, 9, 12), 3, ow_a, ow_b, None),
                 ("gg_aniso", (40, 70), 4, ga_a, ga_b, "grains"), ("gg_aniso", (8, 10, 12), 4, ga_a, ga_b, "grains"),
                 ("ac", (64, 64), 0, ac_a, ac_b, None)]
        for model, shape, q, pa, pb, kind in cases:
            say("[gpu] graph replay with new parameters %s %s %s" % (model, shape, prec))
            n = int(np.prod(shape))
            if kind == "grains":
                a = mg.one_hot_grains(shape, q, 950 + q)
            elif q:
                a = O.uniform(n * q, 960 + q, 0.0, 1.0).reshape(shape + (q,))
            else:
                a = O.uniform(n, 970, -0.01, 0.01).reshape(shape)
            if prec == "f32":
                a = a.astype(np.float32)
            kw = dict(fields=q, vector=True) if q else {}
            with mx.Grid(shape, precision=prec, **kw) as g:
                g.upload(a)
                g.step(model, pa, 100)
                mid = g.download()
                g.step(model, pb, 100)
                got = g.download()
            exp_mid = oracle_run(model, a, pa, 100, prec)
            exp = oracle_run(model, exp_mid, pb, 100, prec)
            assert np.array_equal(mid, exp_mid), (model, shape, prec, "set A")
            assert np.array_equal(got, exp), (model, shape, prec, "set B")


def test_zero_steps_leave_the_state():
    """step(model, p, 0) launches nothing and keeps the state, also for the
    stochastic forms, whose step count (the noise counter) stays where it was."""
    a = O.uniform(24 * 70, 31, 0.3, 0.7).reshape(24, 70)
    q = O.uniform(24 * 70 * 8, 32, 0.0, 1.0).reshape(24, 70, 8)
    for model, params, arr, kw in [("ac", [1.0, 1.0, 1.0, 1.0, 0.01], a, {}), ("ch_dw", [1.0, 1.0, 0.005], a, {}),
                                   ("gg", [0.01], q, dict(fields=8, vector=True)),
                                   ("ch_noise", [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 5], a, {})]:
        say("[gpu] zero steps %s" % model)
        with mx.Grid(arr.shape[:2], **kw) as g:
            g.upload(arr)
            g.step(model, params, 0)
            assert np.array_equal(g.download(), arr), model
            if model == "ch_noise":
                g.step(model, params, 3)
                assert np.array_equal(g.download(), O.ch_noise(arr, params, 3)), model


@pytest.mark.parametrize("case", ["gg32", "zener16", "gga8", "ostwald4", "ac_neumann", "ch3d", "gg4"])
def test_autotune_runs_every_candidate_and_keeps_the_state(case):
    """mmspx_autotune times every compiled candidate of the model on the state
    (the wave strips, the split strips on two and four waves, the tiles, the
    generic kernel), restores the state, and the tile it keeps gives the
    oracle's bits."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    say("[gpu] autotune %s" % case)
    if case == "gg32":
        shape, q, prec, model, p = (70, 200), 32, "f32", "gg", [0.01]
        a = O.uniform(int(np.prod(shape)) * q, 801, 0.0, 1.0).reshape(shape + (q,)).astype(np.float32)
        exp = O.gg(a, 0.01, 7, form=prec)
    elif case == "gg4":
        shape, q, prec, model, p = (66, 130), 4, "f64", "gg", [0.01]
        a = O.uniform(int(np.prod(shape)) * q, 806, 0.0, 1.0).reshape(shape + (q,))
        exp = O.gg(a, 0.01, 7, form=prec)
    elif case == "zener16":
        shape, q, prec, model, p = (50, 140), 16, "f32", "zener", [0.01]
        a = O.uniform(int(np.prod(shape)) * q, 802, 0.0, 1.0).reshape(shape + (q,)).astype(np.float32)
        exp = O.zener(a, 0.01, 7, form=prec)
    elif case == "gga8":
        shape, q, prec, model, p = (64, 200), 8, "f64", "gg_aniso", [0.01, 8.0, 1.0, 1.0]
        a = mg.one_hot_grains(shape, q, 803)
        exp = O.gg_aniso(a, p, 7, form=prec)
    elif case == "ostwald4":
        shape, q, prec, model, p = (64, 150), 4, "f64", "ostwald", [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
        a = O.uniform(int(np.prod(shape)) * q, 804, 0.0, 1.0).reshape(shape + (q,))
        exp = O.ostwald(a, p, 7, form=prec)
    elif case == "ac_neumann":
        shape, q, prec, model, p = (70, 200), 0, "f64", "ac", [1.0, 1.0, 1.0, 1.0, 0.01]
        a = O.uniform(int(np.prod(shape)), 805, -1.0, 1.0).reshape(shape)
        exp = O.ac(a, p, 7, form=prec, bc=[(N, N), (N, N)])
    else:
        shape, q, prec, model, p = (40, 36, 70), 0, "f64", "ch_dw", [1.0, 1.0, 0.005]
        a = O.uniform(int(np.prod(shape)), 807, 0.45, 0.55).reshape(shape)
        exp = O.ch(a, p, 7, mu=O.MU_DW, form=prec)
    kw = dict(fields=q, vector=True) if q else {}
    bc = [(N, N), (N, N)] if case == "ac_neumann" else None
    with mx.Grid(shape, precision=prec, bc=bc, **kw) as g:
        g.upload(a)
        rep = g.autotune(model, p, trial_steps=2)
        lines = [ln for ln in rep.splitlines() if ln.startswith("variant=")]
        assert len(lines) >= 2, rep  # the generic kernel and at least one tiled candidate
        assert np.array_equal(g.download(), a)  # restored
        g.step(model, p, 7)
        assert np.array_equal(g.download(), exp), (case, g.get_tile())
