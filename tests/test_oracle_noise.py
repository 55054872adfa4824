"""CPU: the oracle's stochastic forms (model A, model B / spinodal) and their noise.

The reference draws its noise from rand() in node order (gaussian(),
examples/phase_transitions/model_A/model_A.cpp:13-39), which no parallel
implementation can follow, so parity with the reference is in distribution.
What pins the restatement:
  * Philox4x32-10 against the known-answer vectors published with the
    Random123 library (Salmon et al., SC'11);
  * the log and cosine of the Box-Muller transform against libm;
  * the deviates against N(0, 1) (Kolmogorov-Smirnov, moments, correlations);
  * sigma = 0 reduces model A / B to the deterministic Allen-Cahn /
    Cahn-Hilliard restatements bitwise (the operation order around the noise);
  * model B conserves mass with noise (the noise enters through lap(mu));
  * linearised models (u = 0): the mean of phi^2 at every step has a closed
    form per Fourier mode, E|phi_k|^2 = b_k^2 sigma^2 (1 - a_k^(2t)) / (1 - a_k^2);
  * the reference's own stock model A and model B (oracle/_ref, seeded with
    srand) against the restatement on the same grid and steps.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from oracle import dat

scipy_stats = pytest.importorskip("scipy.stats")


def test_philox_known_answers():
    """Random123's kat_vectors for philox4x32 with 10 rounds."""
    M = 0xFFFFFFFF
    assert O.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox4x32_10([M, M, M, M], [M, M]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_ln_and_cos_against_libm():
    rng = np.random.default_rng(1)
    us = np.concatenate([rng.random(20000), rng.random(2000) ** 30, [1.0, 0.5, 0.25, 2.0 ** -33, 0.70710677,
                                                                     0.70710683, 0.99999994]]).astype(np.float32)
    us = us[us >= 2.0 ** -33]  # the deviates' u1 never goes below 2^-33 (normal floats only)
    for u in us:
        got, want = O.ln_unit(float(u)), math.log(float(u))
        assert abs(got - want) <= 2e-7 * max(1.0, abs(want)), (u, got, want)
    assert O.ln_unit(1.0) == 0.0
    ts = np.concatenate([np.arange(0, 2 ** 24, 4099), [0, 2 ** 22, 2 ** 23, 3 * 2 ** 22, 2 ** 24 - 1]]) * 2.0 ** -24
    for t in ts:
        c, s = O.cos_sin_2pi(float(t))
        assert abs(c - math.cos(2 * math.pi * float(t))) <= 3e-7, (t, c)
        assert abs(s - math.sin(2 * math.pi * float(t))) <= 3e-7, (t, s)
    assert O.cos_sin_2pi(0.0) == (1.0, 0.0) and O.cos_sin_2pi(0.5) == (-1.0, 0.0)
    assert O.cos_sin_2pi(0.25) == (0.0, 1.0) and O.cos_sin_2pi(0.75) == (0.0, -1.0)


def test_deviates_are_standard_normal_and_independent():
    # node (x, r): x the axis-0 index, r the index within the plane; the two
    # planes x = 2k and 2k+1 share a Philox block (the cos and sin of one pair)
    n = 100000
    ge = np.array([O.normal(11, 3, 2 * (i % 50), i // 50) for i in range(n)])
    go = np.array([O.normal(11, 3, 2 * (i % 50) + 1, i // 50) for i in range(n)])
    g = np.concatenate([ge, go])
    m = g.size
    assert abs(g.mean()) < 5 / math.sqrt(m)
    assert abs(g.var() - 1.0) < 5 * math.sqrt(2.0 / m)
    assert abs(scipy_stats.skew(g)) < 5 * math.sqrt(6.0 / m)
    assert abs(scipy_stats.kurtosis(g)) < 5 * math.sqrt(24.0 / m)
    assert scipy_stats.kstest(g, "norm").pvalue > 1e-4
    lim = 5 / math.sqrt(n)
    # the two values of one pair, neighbouring r, the next step, another seed
    assert abs(np.corrcoef(ge, go)[0, 1]) < lim
    assert abs(np.corrcoef(ge ** 2, go ** 2)[0, 1]) < lim  # independent, not merely uncorrelated
    h = np.array([O.normal(11, 4, 2 * (i % 50), i // 50) for i in range(n)])
    k = np.array([O.normal(12, 3, 2 * (i % 50), i // 50) for i in range(n)])
    r1 = np.array([O.normal(11, 3, 2 * (i % 50), i // 50 + 7919) for i in range(n)])
    for other in (h, k, r1):
        assert abs(np.corrcoef(ge, other)[0, 1]) < lim
    assert O.normal(11, 3, 2 * 17, 0) == ge[17]  # a function of (seed, step, x, r) only
    assert np.abs(g).max() < 6.8  # u1 >= 2^-33 bounds |g| by sqrt(66 ln 2)


SHAPES = [((9, 13), [(O.PERIODIC, O.PERIODIC), (O.NEUMANN, O.MIRROR)]), ((5, 6, 7), None), ((31,), [(O.DIRICHLET, O.PERIODIC)])]


@pytest.mark.parametrize("form", ["f64", "f32", "f32ref"])
@pytest.mark.parametrize("shape,bc", SHAPES)
def test_sigma_zero_is_the_deterministic_model(form, shape, bc):
    a = O.uniform(int(np.prod(shape)), 21, -0.3, 0.3).reshape(shape)
    base = [0.7, 1.3, 0.9, 1.1, 0.01]
    ref_a = O.ac(a, base, 7, form=form, bc=bc)
    got_a = O.ac_noise(a, base + [0.0, 99], 7, form=form, bc=bc)
    assert np.array_equal(ref_a, got_a)
    ref_c = O.ch(a, base, 7, mu=O.MU_POLY, form=form, bc=bc)
    got_c = O.ch_noise(a, base + [0.0, 99], 7, form=form, bc=bc)
    assert np.array_equal(ref_c, got_c)


@pytest.mark.parametrize("which", ["ac", "ch"])
def test_steps_continue_the_noise_sequence(which):
    f = O.ac_noise if which == "ac" else O.ch_noise
    a = O.uniform(12 * 10, 4, -0.1, 0.1).reshape(12, 10)
    p = [1.0, 1.0, 1.0, 1.0, 0.01, 0.5, 3]
    whole = f(a, p, 10)
    split = f(f(a, p, 4), p, 6, step0=4)
    assert np.array_equal(whole, split)
    assert not np.array_equal(whole, f(a, p[:6] + [4], 10))  # another seed, another path


def test_model_b_conserves_mass_with_noise():
    a = 0.5 + 0.05 * O.uniform(24 * 20 * 16, 8, -1, 1).reshape(24, 20, 16)
    out = O.ch_noise(a, [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 5], 40)
    m0, m1 = O.mass(a)[0], O.mass(out)[0]
    assert abs(m1 - m0) / abs(m0) < 1e-12
    assert np.abs(out - O.ch(a, [1.0, 1.0, 1.0, 1.0, 0.005], 40, mu=O.MU_POLY)).max() > 1e-4  # the noise acted


def _lam(n, m):
    l1 = lambda k, N: 2 - 2 * np.cos(2 * np.pi * k / N)
    return l1(np.arange(n), n)[:, None] + l1(np.arange(m), m)[None, :]


def _mean_sq_path(step_fn, x, T):
    out = []
    for t in range(T):
        x = step_fn(x, t)
        out.append(float((x ** 2).mean()))
    return np.array(out)


@pytest.mark.parametrize("which", ["A", "B"])
def test_linearised_models_match_the_closed_form(which):
    """u = 0, r = -1 (stable): per Fourier mode phi_k' = a_k phi_k + b_k xi_k."""
    N, T = 64, 400
    lam = _lam(N, N)
    if which == "A":
        dtM = 0.1
        p = [-1.0, 0.0, 1.0, 1.0, dtM, 1.0, 11]
        a_k, b_k = 1 - dtM * (1 + lam), dtM + 0 * lam
        path = _mean_sq_path(lambda x, t: O.ac_noise(x, p, 1, step0=t), np.zeros((N, N)), T)
    else:
        dtM = 0.02
        p = [-1.0, 0.0, 1.0, 1.0, dtM, 1.0, 13]
        a_k, b_k = 1 - dtM * lam * (1 + lam), dtM * lam
        path = _mean_sq_path(lambda x, t: O.ch_noise(x, p, 1, step0=t), np.zeros((N, N)), T)
    t = np.arange(1, T + 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        per_mode = [np.where(b_k > 0, b_k ** 2 * (1 - a_k ** (2 * s)) / (1 - a_k ** 2), 0.0).mean() for s in t]
    exact = np.array(per_mode)
    ratio = path[100:].mean() / exact[100:].mean()
    # about 0.3 % statistical spread at this size; a misplaced dt*M or a noise
    # added to the update instead of mu misses by far more
    assert abs(ratio - 1) < 0.02, ratio


@pytest.mark.skipif(not os.path.exists(os.path.join(O.REF_DIR, "ref_ma")), reason="reference not built here")
@pytest.mark.parametrize("model", ["ma", "mb"])
def test_reference_stock_models_agree_in_distribution(tmp_path, model):
    """The stock model_A / model_B update (r = u = K = M = 1, dt = 0.01, kT = 0.01,
    dV = 1: sigma = sqrt(2 kT / (dt dV))) from phi = 0 on 128^2 for 100 steps,
    six rand() seeds, against six Philox seeds of the restatement: the mean of
    phi^2 agrees within four combined standard errors, and both agree with the
    linearised closed form (the cubic term is below 1 % at these amplitudes)."""
    N, T = 128, 100
    z = np.zeros((N, N))
    fin = str(tmp_path / "z.dat")
    dat.write_dat(fin, dat.make_grid(z), compress=False)
    refs = []
    for sd in range(1, 7):
        out = str(tmp_path / ("o%d.dat" % sd))
        env = dict(os.environ, HARNESS_SEED=str(sd))
        subprocess.run([os.path.join(O.REF_DIR, "ref_" + model), "run", fin, out, str(T), "stock"], check=True,
                       capture_output=True, env=env, timeout=300)
        refs.append(float((dat.read_dat(out).data ** 2).mean()))
    sigma = math.sqrt(2.0 * 0.01 / (0.01 * 1.0))
    f = O.ac_noise if model == "ma" else O.ch_noise
    ours = [float((f(z, [1.0, 1.0, 1.0, 1.0, 0.01, sigma, seed], T) ** 2).mean()) for seed in range(1, 7)]
    r, o = np.array(refs), np.array(ours)
    se = math.hypot(r.std(ddof=1) / math.sqrt(len(r)), o.std(ddof=1) / math.sqrt(len(o)))
    assert abs(r.mean() - o.mean()) < 4 * se, (r, o)
    lam = _lam(N, N)
    dtM = 0.01
    if model == "ma":
        a_k, b_k = 1 + dtM * (1 - lam), dtM + 0 * lam
    else:
        a_k, b_k = 1 + dtM * lam * (1 - lam), dtM * lam
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(np.abs(a_k ** 2 - 1) > 0, (a_k ** (2 * T) - 1) / (a_k ** 2 - 1), T)
    theory = float((b_k ** 2 * sigma ** 2 * g).mean())
    for v in (r, o):
        assert abs(v.mean() / theory - 1) < 0.03, (v.mean(), theory)
