"""GPU: batched solves (one persistent launch per step over every member's
tiles, hipGraph replay). Each member must equal, bitwise, the oracle's run
of that member alone and the single-grid path's run of it."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
mx = pytest.importorskip("mmsp_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


P, N, M, D = O.PERIODIC, O.NEUMANN, O.MIRROR, O.DIRICHLET


def seeded(shape, seed, lo, hi, q=0, f32=True):
    a = O.uniform(int(np.prod(shape)) * max(q, 1), seed, lo, hi).reshape(tuple(shape) + ((q,) if q else ()))
    return a.astype(np.float32) if f32 else a


@pytest.mark.parametrize("tile", [None, (128, 16, 4, 0), (256, 32, 8, 1024), (64, 8, 16, 300), (48, 8, 4, 0),
                                  (48, 3, 8, 0)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batch_allen_cahn_uneven_members(prec, tile):
    shapes = [(64, 96), (33, 47), (128, 128), (16, 200), (300, 20), (7, 5)]
    kappas = [0.5, 1.0, 1.3, 2.0, 0.8, 1.7]
    say("[gpu] batch AC %s members=%s" % (prec, shapes))
    params = np.array([[1.0, 1.0, k, 1.0, 0.01] for k in kappas])
    init = [seeded(s, 40 + m, -0.5, 0.5, f32=prec == "f32") for m, s in enumerate(shapes)]
    with mx.Batch(shapes, "ac", precision=prec) as b:
        b.set_params(params)
        if tile:
            b.set_tile(*tile)
        for m, a in enumerate(init):
            b.upload(m, a)
        b.step(37, graph_steps=10)  # three graph replays + 7 direct steps
        got = [b.download(m) for m in range(len(shapes))]
    for m, s in enumerate(shapes):
        exp = O.ac(init[m], params[m], 37, form=prec)
        assert np.array_equal(got[m], exp), (m, s)


def test_batch_allen_cahn_boundaries_and_alone_vs_in_mix():
    shapes = [(40, 70), (90, 33)]
    bcs = [[(N, M), (D, P)], [(M, M), (N, N)]]
    say("[gpu] batch AC boundaries")
    params = np.array([[1.0, 1.0, 1.0, 1.0, 0.01], [1.0, 1.0, 1.5, 1.0, 0.01]])
    init = [seeded(s, 60 + m, -1, 1) for m, s in enumerate(shapes)]
    with mx.Batch(shapes, "ac", precision="f32", bcs=bcs) as b:
        b.set_params(params)
        with pytest.raises(mx.MmspxError):
            b.set_tile(48, 8, 4, 0)  # temporal blocks need periodic members
        b.set_tile(256, 32, 8, 512)  # runs of consecutive nodes through the boundary mapping
        for m, a in enumerate(init):
            b.upload(m, a)
        b.step(20, graph_steps=4)
        got = [b.download(m) for m in range(2)]
    for m in range(2):
        with mx.Grid(shapes[m], precision="f32", bc=bcs[m]) as g:
            g.upload(init[m])
            g.step("ac", params[m], 20)
            alone = g.download()
        assert np.array_equal(got[m], alone)
        assert np.array_equal(got[m], O.ac(init[m], params[m], 20, form="f32", bc=bcs[m]))


# 2-D members by wave strips with chunks of 3, 5, 7, 13 and 16 rows (the
# eight-row unrolled march ends at every ring position), or by LDS row strips
@pytest.mark.parametrize("tile", [None, (60, 3, 4, 0), (60, 5, 4, 0), (60, 7, 4, 0), (60, 13, 4, 0), (60, 16, 4, 0),
                                  (64, 16, 8, 0)])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batch_cahn_hilliard_2d_and_3d_members(prec, tile):
    shapes = [(64, 64), (20, 36, 40), (100, 70), (16, 16, 16), (33, 65, 70), (40, 600), (9, 20, 300), (70, 257)]
    say("[gpu] batch CH %s tile=%s" % (prec, tile))
    params = np.array([[1.0, 1.0, 0.005]] * len(shapes))
    init = [seeded(s, 80 + m, 0.45, 0.55, f32=prec == "f32") for m, s in enumerate(shapes)]
    with mx.Batch(shapes, "ch_dw", precision=prec) as b:
        b.set_params(params)
        if tile:
            b.set_tile(*tile)  # 2-D members by 60-column wave strips (the default) or LDS row strips
        for m, a in enumerate(init):
            b.upload(m, a)
        m0 = [b.mass(m)[0] for m in range(len(shapes))]
        b.step(24, graph_steps=8)
        got = [b.download(m) for m in range(len(shapes))]
        m1 = [b.mass(m)[0] for m in range(len(shapes))]
    for m, s in enumerate(shapes):
        exp = O.ch(init[m], params[m], 24, mu=O.MU_DW, form=prec)
        assert np.array_equal(got[m], exp), (m, s)
        if prec == "f64":
            assert abs(m1[m] - m0[m]) <= 1e-12 * abs(m0[m])


def test_batch_grain_growth_q8():
    shapes = [(50, 70), (64, 130)]
    say("[gpu] batch GG q=8")
    params = np.array([[0.05], [0.03]])
    init = [seeded(s, 90 + m, 0.0, 1.0, q=8) for m, s in enumerate(shapes)]
    with mx.Batch(shapes, "gg", precision="f32", fields=8, vector=True) as b:
        b.set_params(params)
        for m, a in enumerate(init):
            b.upload(m, a)
        b.step(15, graph_steps=5)
        got = [b.download(m) for m in range(2)]
    for m in range(2):
        assert np.array_equal(got[m], O.gg(init[m], params[m][0], 15, form="f32"))


def test_batch_autotune_restores_state_and_keeps_results():
    shapes = [(256, 256), (128, 512), (64, 64)]
    say("[gpu] batch autotune")
    params = np.array([[1.0, 1.0, 1.0, 1.0, 0.01]] * 3)
    init = [seeded(s, 100 + m, -0.01, 0.01) for m, s in enumerate(shapes)]
    with mx.Batch(shapes, "ac", precision="f32") as b:
        b.set_params(params)
        for m, a in enumerate(init):
            b.upload(m, a)
        rep = b.autotune(trial_steps=3)
        assert rep.count("\n") >= 30 and b.get_tile()["tuned"]
        for m, a in enumerate(init):
            assert np.array_equal(b.download(m), a)
        b.step(12, graph_steps=6)
        got = [b.download(m) for m in range(3)]
    for m in range(3):
        assert np.array_equal(got[m], O.ac(init[m], params[m], 12, form="f32"))


def test_batch_rejects_unsupported_members():
    with pytest.raises(mx.MmspxError):
        mx.Batch([(8, 8, 8)], "ac")
    with pytest.raises(mx.MmspxError):
        mx.Batch([(8, 8)], "ch_dw", bcs=[[(N, N), (P, P)]])
    with pytest.raises(mx.MmspxError):
        mx.Batch([(8, 8), (8, 8)], "gg", fields=5, vector=True)
