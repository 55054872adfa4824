"""GPU: the stochastic forms (model A: MMSPX_AC_NOISE, model B / spinodal:
MMSPX_CH_NOISE) on every kernel that runs them, against the oracle.

The noise is a function of (seed, step, node), so every kernel, tile and chunk
must give the oracle's bits: bitwise in every precision form. The reference
itself is matched in distribution on the CPU (test_oracle_noise.py)."""
import math
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

mx = pytest.importorskip("mmsp_amd")

P, N, M, D = O.PERIODIC, O.NEUMANN, O.MIRROR, O.DIRICHLET
SIG = math.sqrt(2.0 * 0.01 / (0.01 * 1.0))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def seeded(shape, seed, lo, hi, prec):
    a = O.uniform(int(np.prod(shape)), seed, lo, hi).reshape(shape)
    return a if prec == "f64" else a.astype(np.float32)


def run(shape, prec, model, p, steps, a, bc=None, variant=0, tile=None, body=None, split=None):
    with mx.Grid(shape, precision=prec, bc=bc) as g:
        if variant:
            g.set_variant(variant)
        if tile:
            g.set_tile(*tile)
        if body is not None:
            g.set_tile_options(body=body)
        g.upload(a)
        for n in split or [steps]:
            g.step(model, p, n)
        return g.download()


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_model_a_generic_every_dimension_and_boundary(prec):
    p = [1.0, 1.0, 1.0, 1.0, 0.01, SIG, 17]
    for shape, bc in [((300,), [(N, D)]), ((33, 70), [(P, P), (M, N)]), ((9, 10, 11), None)]:
        say("[gpu] model A generic %s %s" % (shape, prec))
        a = seeded(shape, 31, -0.2, 0.2, prec)
        exp = O.ac_noise(a, p, 9, form=prec, bc=bc)
        assert np.array_equal(run(shape, prec, "ac_noise", p, 9, a, bc=bc, variant=1), exp)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_model_a_wave_strips(prec):
    p = [1.0, 1.0, 1.0, 1.0, 0.01, SIG, 23]
    for shape, bc in [((70, 300), [(N, M), (D, P)]), ((40, 64), None), ((9, 5), [(M, M), (N, D)])]:
        say("[gpu] model A wave strips %s %s" % (shape, prec))
        a = seeded(shape, 37, -0.2, 0.2, prec)
        exp = O.ac_noise(a, p, 11, form=prec, bc=bc)
        for chunk in (7, 8, 24):
            got = run(shape, prec, "ac_noise", p, 11, a, bc=bc, variant=2, tile=(1, 64, chunk), body=2)
            assert np.array_equal(got, exp), (shape, chunk)
        # the default for 2-D model A is the wave strips; steps split over calls continue the noise
        assert np.array_equal(run(shape, prec, "ac_noise", p, 11, a, bc=bc, split=[4, 7]), exp)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_model_b_generic(prec):
    p = [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 5]
    for shape, bc in [((40, 33), [(N, M), (P, D)]), ((10, 9, 12), None)]:
        say("[gpu] model B generic %s %s" % (shape, prec))
        a = seeded(shape, 41, 0.3, 0.7, prec)
        exp = O.ch_noise(a, p, 8, form=prec, bc=bc)
        assert np.array_equal(run(shape, prec, "ch_noise", p, 8, a, bc=bc, variant=1), exp)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ref"])
def test_model_b_fused_bodies(prec):
    """3-D: the LDS-ring body, register queues on 512, 768 and 1024 threads, partial
    tiles and short chunks; 2-D: row strips of both bodies and the wave strips
    (both marching directions)."""
    p = [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 6]
    shape = (40, 36, 70)
    a = seeded(shape, 43, 0.45, 0.55, prec)
    exp = O.ch_noise(a, p, 7, form=prec)
    for tile, body in [((16, 64, 16), 0), ((16, 64, 16), 1), ((32, 64, 12), 1), ((32, 64, 16), 2), ((8, 32, 7), 0),
                       ((32, 64, 5), 3)]:
        say("[gpu] model B fused %s body %d %s" % (tile, body, prec))
        got = run(shape, prec, "ch_noise", p, 7, a, variant=2, tile=tile, body=body)
        assert np.array_equal(got, exp), (tile, body)
    assert np.array_equal(run(shape, prec, "ch_noise", p, 7, a, split=[3, 4]), exp)  # default tile
    for shape2 in [(40, 300), (17, 5)]:
        a2 = seeded(shape2, 47, 0.45, 0.55, prec)
        exp2 = O.ch_noise(a2, p, 9, form=prec)
        for tile, body in [((1, 64, 8), 0), ((1, 128, 8), 1), ((1, 60, 3), 0), ((1, 60, 8), 0)]:
            say("[gpu] model B 2-D %s %s body %d %s" % (shape2, tile, body, prec))
            got = run(shape2, prec, "ch_noise", p, 9, a2, variant=2, tile=tile, body=body)
            assert np.array_equal(got, exp2), (shape2, tile, body)


def test_sigma_zero_is_the_deterministic_kernel():
    shape = (24, 20, 64)
    a = seeded(shape, 53, 0.4, 0.6, "f64")
    base = [1.0, 1.0, 1.0, 1.0, 0.005]
    det = run(shape, "f64", "ch", base, 12, a)
    assert np.array_equal(run(shape, "f64", "ch_noise", base + [0.0, 1], 12, a), det)
    a2 = seeded((50, 130), 54, -0.3, 0.3, "f64")
    det2 = run((50, 130), "f64", "ac", base, 12, a2, variant=2, tile=(1, 64, 8), body=2)
    assert np.array_equal(run((50, 130), "f64", "ac_noise", base + [0.0, 1], 12, a2), det2)


def test_model_b_conserves_mass_at_size():
    shape = (128, 128, 128)
    say("[gpu] model B 128^3 mass")
    with mx.Grid(shape, precision="f64") as g:
        g.upload(0.5 + 0.05 * O.uniform(int(np.prod(shape)), 59, -1, 1).reshape(shape))
        m0 = g.mass()[0]
        g.step("ch_noise", [1.0, 1.0, 1.0, 1.0, 0.005, SIG, 9], 200)
        m1 = g.mass()[0]
        out = g.download()
    assert abs(m1 - m0) / abs(m0) < 1e-12
    assert np.all(np.isfinite(out))


def test_autotune_keeps_the_noise_sequence():
    shape = (20, 24, 64)
    p = [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 8]
    a = seeded(shape, 61, 0.45, 0.55, "f64")
    exp = O.ch_noise(a, p, 6, form="f64")
    with mx.Grid(shape, precision="f64") as g:
        g.upload(a)
        g.autotune("ch_noise", p, trial_steps=2)
        g.upload(a)
        g.step("ch_noise", p, 6)
        assert np.array_equal(g.download(), exp)


def test_batches_and_servers_refuse_the_stochastic_forms():
    with pytest.raises(mx.MmspxError):
        mx.Batch([(64, 64), (32, 32)], "ac_noise", precision="f32")
    with pytest.raises(mx.MmspxError):
        mx.Grid((8, 8)).step("ac_noise", [1.0, 1.0, 1.0, 1.0, 0.01, 1.0, 0.5], 1)  # seed not an integer
