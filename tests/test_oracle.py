"""CPU tests that pin the oracle.

The oracle (oracle/mmsp_oracle.c, oracle/dat.py) is checked
  * bitwise against golden vectors written by the reference itself
    (tests/golden/, produced by oracle/_ref from the reference's own update()
    loops and .dat writer), and
  * against what the mathematics fixes: constant/linear/Fourier-mode
    laplacians, scalar maps of uniform states, fixed points, symmetries,
    linear growth rates, mass invariance and closed-form energies,
so that a dropped term, a wrong sign, a swapped axis or a wrong neighbour
fails at least one test.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from oracle import dat
from conftest import bc_of, golden_path, params_of

AC_STOCK = [1.0, 1.0, 1.0, 1.0, 0.01]  # allen-cahn.cpp:58-62


def run_oracle(case, a, cube, bc=None, dx=None):
    form = "f64" if case["dtype"] == "f64" else "f32ref"
    p = params_of(case)
    if case["model"] == "ac":
        return O.ac(a, p, case["steps"], form=form, cube=cube, bc=bc, dx=dx)
    if case["model"] == "ch":
        mu = O.MU_DW if case["variant"] == "config" else O.MU_POLY
        return O.ch(a, p, case["steps"], mu=mu, form=form, cube=cube, bc=bc, dx=dx)
    if case["model"] == "zn":
        return O.zener(a, p[0], case["steps"], form=form, cube=cube, bc=bc, dx=dx)
    if case["model"] == "ow":
        return O.ostwald(a, p, case["steps"], form=form, cube=cube, bc=bc, dx=dx)
    if case["model"] == "ga":
        return O.gg_aniso(a, p, case["steps"], form=form, bc=bc, dx=dx)
    return O.gg(a, p[0], case["steps"], form=form, cube=cube, bc=bc, dx=dx)


# ---- golden vectors from the reference -----------------------------------------
def test_golden_manifest_complete(cases):
    names = {c["name"] for c in cases}
    for m in ("ac", "ch", "gg"):
        assert any(c["model"] == m for c in cases)
    # every boundary condition and every dimension appears
    bcs = {b for c in cases if c["bc"] for pair in c["bc"] for b in pair}
    assert bcs >= {O.MIRROR, O.NEUMANN, O.PERIODIC, O.DIRICHLET}
    assert {len(c["shape"]) for c in cases} == {1, 2, 3}
    for n in names:
        assert os.path.exists(golden_path(n, "in")) and os.path.exists(golden_path(n, "out"))


def test_oracle_bitwise_vs_reference(cases):
    """Oracle with libm pow reproduces the reference's output bits."""
    for c in cases:
        gin = dat.read_dat(golden_path(c["name"], "in"))
        gout = dat.read_dat(golden_path(c["name"], "out"))
        bc = list(zip(gin.b0, gin.b1))
        got = run_oracle(c, gin.data, O.CUBE_POW, bc=bc, dx=gin.dx)
        assert got.dtype == gout.data.dtype
        assert np.array_equal(got, gout.data), c["name"]
        assert bc == (bc_of(c) or [(O.PERIODIC, O.PERIODIC)] * len(c["shape"]))


def test_oracle_cr_cube_within_tolerance(cases):
    """The correctly rounded cube the device uses stays within 1e-12 of the reference."""
    for c in cases:
        gin = dat.read_dat(golden_path(c["name"], "in"))
        gout = dat.read_dat(golden_path(c["name"], "out"))
        got = run_oracle(c, gin.data, O.CUBE_CR, bc=list(zip(gin.b0, gin.b1)), dx=gin.dx)
        ref = gout.data.astype(np.float64)
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
        tol = 1e-12 if c["dtype"] == "f64" else 1e-6
        assert err <= tol, (c["name"], err)


def test_cube_cr_is_correctly_rounded():
    """cube_cr(x) equals x**3 rounded once (exact rational arithmetic)."""
    from fractions import Fraction

    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.uniform(-1.5, 1.5, 4000), rng.uniform(-0.02, 0.02, 2000), [0.0, 1.0, -1.0, 0.5]]):
        exact = Fraction(float(x)) ** 3
        got = O.cube_cr(float(x))
        # correctly rounded: no double is closer to the exact cube
        lo, hi = np.nextafter(got, -np.inf), np.nextafter(got, np.inf)
        e = abs(Fraction(got) - exact)
        assert e <= abs(Fraction(float(lo)) - exact) and e <= abs(Fraction(float(hi)) - exact), x


# ---- boundary mapping ------------------------------------------------------------
@pytest.mark.parametrize("n", [2, 3, 7])
def test_check_boundary_definitions(n):
    for x in range(-1, n + 1):
        assert O.check_boundary(x, 0, n, O.PERIODIC, O.PERIODIC) == x % n
        assert O.check_boundary(x, 0, n, O.NEUMANN, O.NEUMANN) == min(max(x, 0), n - 1)
        assert O.check_boundary(x, 0, n, O.DIRICHLET, O.DIRICHLET) == min(max(x, 0), n - 1)
        mir = -x if x < 0 else (2 * (n - 1) - x if x >= n else x)
        assert O.check_boundary(x, 0, n, O.MIRROR, O.MIRROR) == mir
    # each side uses its own condition
    assert O.check_boundary(-1, 0, n, O.NEUMANN, O.PERIODIC) == 0
    assert O.check_boundary(n, 0, n, O.NEUMANN, O.PERIODIC) == 0


# ---- laplacian closed forms ----------------------------------------------------------
@pytest.mark.parametrize("shape,dx", [((32,), [0.5]), ((16, 12), [1.0, 0.25]), ((8, 6, 10), [1.0, 2.0, 0.5])])
def test_laplacian_fourier_eigenvalue(shape, dx):
    ks = [1, 2, 3][: len(shape)]
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    theta = sum(2 * np.pi * k * x / n for k, x, n in zip(ks, grids, shape))
    a = np.cos(theta)
    lam = sum((2 * np.cos(2 * np.pi * k / n) - 2) / (h * h) for k, n, h in zip(ks, shape, dx))
    flat = a.reshape(-1)
    for i in range(0, flat.size, max(1, flat.size // 37)):
        got = O.laplacian(a, i, dx=dx)
        assert abs(got - lam * flat[i]) < 1e-12 * max(1.0, abs(lam))


@pytest.mark.parametrize("bc", [O.PERIODIC, O.NEUMANN, O.MIRROR, O.DIRICHLET])
def test_laplacian_constant_and_ramp(bc):
    a = np.full((5, 6), 0.3)
    for i in range(a.size):
        assert O.laplacian(a, i, bc=[(bc, bc)] * 2) == 0.0
    # a linear ramp along axis 0 has zero laplacian away from the faces
    r = np.add.outer(np.arange(7.0) * 3.0, np.zeros(4))
    for x in range(1, 6):
        for y in range(4):
            assert O.laplacian(r, x * 4 + y, bc=[(bc, bc)] * 2) == 0.0


def test_laplacian_axis_order_and_neighbours():
    """An impulse: the laplacian is -2*sum(w) at the spike and w_d at its 2*dim neighbours."""
    shape, dx = (5, 6, 7), [1.0, 0.5, 2.0]
    w = [1 / (h * h) for h in dx]
    a = np.zeros(shape)
    a[2, 3, 4] = 1.0
    lap = np.array([O.laplacian(a, i, dx=dx) for i in range(a.size)]).reshape(shape)
    assert lap[2, 3, 4] == -2 * sum(w)
    assert lap[1, 3, 4] == w[0] and lap[3, 3, 4] == w[0]
    assert lap[2, 2, 4] == w[1] and lap[2, 4, 4] == w[1]
    assert lap[2, 3, 3] == w[2] and lap[2, 3, 5] == w[2]
    assert np.count_nonzero(lap) == 7


# ---- Allen-Cahn --------------------------------------------------------------------
def test_ac_uniform_state_follows_scalar_map():
    c = 0.37
    a = np.full((6, 5), c)
    got = O.ac(a, AC_STOCK, 50, cube=O.CUBE_POW)
    for _ in range(50):
        c = c - 0.01 * 1.0 * (-1.0 * c + 1.0 * math.pow(c, 3) - 1.0 * 0.0)
    assert np.all(got == c)


def test_ac_fixed_points_and_odd_symmetry():
    for v in (0.0, 1.0, -1.0):
        a = np.full((8, 8), v)
        assert np.array_equal(O.ac(a, AC_STOCK, 20), a)
    a = O.uniform(64, 5, -1, 1).reshape(8, 8)
    assert np.array_equal(O.ac(-a, AC_STOCK, 20), -O.ac(a, AC_STOCK, 20))


def test_ac_linear_growth_rate():
    n, k, eps = 32, 3, 1e-9
    x = np.arange(n)
    a = eps * np.cos(2 * np.pi * k * x / n)[:, None] * np.ones((1, 4))
    K, dt = 1.7, 0.01
    lam = 2 - 2 * np.cos(2 * np.pi * k / n)
    got = O.ac(a, [1.0, 1.0, K, 1.0, dt], 1)
    g = 1 + dt * (1.0 - K * lam)
    assert np.allclose(got, g * a, rtol=1e-9, atol=1e-9 * eps)


# ---- Cahn-Hilliard -------------------------------------------------------------------
def test_ch_mass_invariance_and_uniform_fixed_point():
    a = O.uniform(12 * 10 * 8, 9, 0.45, 0.55).reshape(12, 10, 8)
    got = O.ch(a, [1.0, 1.0, 0.005], 100)
    m0, m1 = O.mass(a)[0], O.mass(got)[0]
    assert abs(m1 - m0) / m0 < 1e-13
    u = np.full((6, 7), 0.42)
    assert np.array_equal(O.ch(u, [1.0, 1.0, 0.005], 10), u)


def test_ch_single_mode_growth_factor():
    """Linear CH: a mode grows by 1 - dt*M*lam*(f''(cbar) + kappa*lam) per step."""
    n, k, eps, kap, M, dt = 32, 2, 1e-9, 1.0, 1.0, 0.005
    x = np.arange(n)
    a = 0.5 + eps * np.cos(2 * np.pi * k * x / n)
    lam = 2 - 2 * np.cos(2 * np.pi * k / n)
    fpp = 2 - 12 * 0.5 + 12 * 0.25  # f = c^2 (1-c)^2
    g = 1 - dt * M * lam * (fpp + kap * lam)
    got = O.ch(a, [kap, M, dt], 1)
    assert np.allclose(got - 0.5, g * (a - 0.5), rtol=1e-6, atol=1e-22)


def test_ch_poly_form_matches_reference_chemical_potential():
    """Stock CH: mu = -phi + phi^3 - lap; with M=0 nothing moves, with kappa... via one step by hand."""
    a = O.uniform(30, 3, -1, 1)
    got = O.ch(a, [1.0, 1.0, 1.0, 1.0, 0.01], 1, mu=O.MU_POLY, cube=O.CUBE_POW)
    lap = lambda v: np.roll(v, -1) - 2 * v + np.roll(v, 1)
    mu = -a + a ** 3 - lap(a)
    exp = a + 0.01 * lap(mu)
    assert np.allclose(got, exp, rtol=0, atol=1e-14)


# ---- grain growth -------------------------------------------------------------------
def test_gg_single_grain_fixed_point_and_zero_fields():
    a = np.zeros((6, 5, 4))
    a[..., 1] = 1.0
    assert np.array_equal(O.gg(a, 0.05, 10), a)
    z = np.zeros((6, 5, 3))
    assert np.array_equal(O.gg(z, 0.05, 10), z)


def test_gg_single_field_is_allen_cahn_with_kappa_2():
    """Q=1: -phi - phi^3 + 2*phi*phi^2 - 2 lap = -phi + phi^3 - 2 lap (AC, r=u=M=1, K=2)."""
    a = O.uniform(20 * 16, 4, -1, 1).reshape(20, 16)
    g = O.gg(a[..., None], 0.01, 30)[..., 0]
    c = O.ac(a, [1.0, 1.0, 2.0, 1.0, 0.01], 30)
    assert np.abs(g - c).max() < 1e-13


def test_gg_field_permutation_symmetry():
    a = O.uniform(10 * 9 * 5, 8, 0, 1).reshape(10, 9, 5)
    perm = [3, 0, 4, 1, 2]
    g1 = O.gg(a, 0.05, 20)[..., perm]
    g2 = O.gg(a[..., perm], 0.05, 20)
    assert np.abs(g1 - g2).max() < 1e-14


def test_gg_interface_shrinks_embedded_grain():
    """The reference's own invariant (graingrowth.cpp:150-170): an embedded grain's area does not grow."""
    # one 64x64 cell of generate(2)'s pattern (graingrowth.cpp:34-48): a grain of radius 16
    n = 64
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    inside = np.sqrt((32.0 - x % 64) ** 2 + (32.0 - y % 64) ** 2) < 16.0
    a = np.zeros((n, n, 2))
    a[..., 1] = inside
    a[..., 0] = ~inside
    s0 = a[..., 1].sum()
    g = O.gg(a, 0.01, 4000)
    assert np.all(np.isfinite(g)) and g[..., 1].sum() <= s0


# ---- FP32 forms --------------------------------------------------------------------
def test_f32_forms_track_f64():
    a = O.uniform(24 * 20, 2, -1, 1).reshape(24, 20)
    r64 = O.ac(a, AC_STOCK, 50)
    for form in ("f32", "f32ref"):
        r32 = O.ac(a.astype(np.float32), AC_STOCK, 50, form=form)
        assert r32.dtype == np.float32
        assert np.abs(r32 - r64).max() < 1e-5


# ---- reductions ---------------------------------------------------------------------
def test_mass_is_correctly_summed():
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, 100000) * 10.0 ** rng.integers(-8, 8, 100000)
    assert O.mass(a)[0] == math.fsum(a)
    v = rng.uniform(0, 1, (500, 4))
    assert np.all(O.mass(v, vector=True) == [math.fsum(v[:, j]) for j in range(4)])


def test_energy_closed_forms():
    # uniform AC field: bulk only
    c, shape = 0.3, (10, 12)
    e = O.energy(np.full(shape, c), O.E_POLY, [1.0, 1.0, 1.0], dx=[0.5, 2.0])
    assert math.isclose(e, 120 * (-0.5 * c * c + 0.25 * c ** 4) * 1.0, rel_tol=1e-14)
    # single mode along axis 1: sum (a(x+1)-a(x))^2 = N (1 - cos th) eps^2
    n0, n1, k, eps, K = 6, 16, 3, 0.01, 1.7
    th = 2 * np.pi * k / n1
    a = np.ones((n0, 1)) * (eps * np.cos(th * np.arange(n1)))[None, :]
    e = O.energy(a, O.E_POLY, [1.0, 1.0, K])
    bulk = n0 * n1 * (-0.5 * eps ** 2 / 2 + 0.25 * eps ** 4 * 3 / 8)
    grad = 0.5 * K * n0 * n1 * (1 - np.cos(th)) * eps ** 2
    assert math.isclose(e, bulk + grad, rel_tol=1e-12)
    # double well at its minima and maximum
    assert O.energy(np.zeros((4, 4)), O.E_DW, [1.0]) == 0.0
    assert math.isclose(O.energy(np.full((4, 4), 0.5), O.E_DW, [1.0]), 16 * 0.0625, rel_tol=1e-15)
    # grain growth, two uniform fields a, b: bulk terms plus the a^2 b^2 pair
    ga, gb = 0.6, 0.3
    v = np.zeros((3, 4, 2))
    v[..., 0], v[..., 1] = ga, gb
    exp = 12 * ((-ga * ga / 2 + ga ** 4 / 4) + (-gb * gb / 2 + gb ** 4 / 4) + ga * ga * gb * gb)
    assert math.isclose(O.energy(v, O.E_GG, [2.0], vector=True), exp, rel_tol=1e-14)


def test_ac_energy_decreases():
    a = O.uniform(32 * 32, 12, -0.5, 0.5).reshape(32, 32)
    e0 = O.energy(a, O.E_POLY, [1.0, 1.0, 1.0])
    e1 = O.energy(O.ac(a, AC_STOCK, 100), O.E_POLY, [1.0, 1.0, 1.0])
    assert e1 < e0


# ---- seeded generator -------------------------------------------------------------------
def test_splitmix64_known_answers():
    # SplitMix64 reference outputs (Vigna's splitmix64.c) for seed 1234567
    exp = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
           16408922859458223821]
    assert [O.lib().orc_splitmix64(1234567, i) for i in range(5)] == exp
    u = O.uniform(1000, 1234567, 0.0, 1.0)
    assert u[0] == (exp[0] >> 11) * 2.0 ** -53


# ---- .dat format ---------------------------------------------------------------------------
def test_dat_roundtrip_bytes(cases):
    """Re-writing what the reference wrote gives the same bytes (zlib level 9, same header)."""
    for c in cases:
        for which in ("in", "out"):
            p = golden_path(c["name"], which)
            raw = open(p, "rb").read()
            assert dat.write_dat(None, dat.read_dat(p)) == raw, p
        out = dat.read_dat(golden_path(c["name"], "out"))
        assert out.hdr_fields == (c["fields"] if c["fields"] else 0)


def test_dat_raw_block_reads_back():
    a = O.uniform(5 * 7 * 3, 2, 0, 1).reshape(5, 7, 3).astype(np.float32)
    g = dat.make_grid(a, vector=True, bc=[(O.NEUMANN, O.MIRROR), (O.PERIODIC, O.DIRICHLET)], dx=[0.5, 3.0])
    r = dat.read_dat(dat.write_dat(None, g, compress=False))
    assert np.array_equal(r.data, a) and r.b0 == g.b0 and r.b1 == g.b1 and r.dx == [0.5, 3.0]


@pytest.mark.skipif(not os.path.exists(os.path.join(O.REF_DIR, "mmsp2tsv")), reason="reference not built here")
def test_dat_readable_by_reference_reader(tmp_path):
    """The reference's own reader (via utility/mmsp2tsv) parses raw and compressed files from dat.py."""
    a = O.uniform(6 * 5, 3, -1, 1).reshape(6, 5)
    for comp in (True, False):
        f = tmp_path / ("g%d.dat" % comp)
        dat.write_dat(str(f), dat.make_grid(a), compress=comp)
        out = tmp_path / ("g%d.tsv" % comp)
        subprocess.run([os.path.join(O.REF_DIR, "mmsp2tsv"), str(f), str(out)], check=True, capture_output=True)
        vals = [float(line.split()[-1]) for line in open(out) if line.strip() and not line.startswith("#")]
        assert np.allclose(sorted(vals), sorted(a.reshape(-1)), atol=1e-12)


# ---- Zener pinning and Ostwald ripening --------------------------------------------
@pytest.mark.parametrize("form", ["f64", "f32"])
def test_zener_holds_the_particles_and_is_grain_growth_without_them(form):
    """Field 0 keeps its values bitwise; with field 0 zero everywhere the other
    fields follow grain growth exactly (field 0 then adds nothing to the sum)."""
    shape, q = (14, 17), 4
    a = O.uniform(int(np.prod(shape)) * q, 301, 0.0, 1.0).reshape(shape + (q,))
    if form == "f32":
        a = a.astype(np.float32)
    bc = [(O.NEUMANN, O.MIRROR), (O.PERIODIC, O.DIRICHLET)]
    z = O.zener(a, 0.01, 9, form=form, bc=bc)
    assert np.array_equal(z[..., 0], a[..., 0])
    a0 = a.copy()
    a0[..., 0] = 0
    assert np.array_equal(O.zener(a0, 0.01, 9, form=form, bc=bc), O.gg(a0, 0.01, 9, form=form, bc=bc))


@pytest.mark.parametrize("form", ["f64", "f32"])
def test_ostwald_conserves_the_concentration(form):
    """C' = C + dt*D*lap(w): on a periodic grid the concentration's sum is
    invariant (to rounding), while the order parameters are not."""
    shape, q = (20, 18, 16), 3
    a = O.uniform(int(np.prod(shape)) * q, 303, 0.0, 1.0).reshape(shape + (q,))
    if form == "f32":
        a = a.astype(np.float32)
    p = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
    out = O.ostwald(a, p, 40, form=form)
    m0, m1 = O.mass(a, vector=True), O.mass(out, vector=True)
    assert abs(m1[0] - m0[0]) / abs(m0[0]) < (1e-13 if form == "f64" else 1e-5)
    assert np.all(np.abs(m1[1:] - m0[1:]) / np.abs(m0[1:]) > 1e-4)


def test_ostwald_uniform_state_follows_the_local_map():
    """A spatially uniform state has zero laplacians: C stays and each order
    parameter follows phi' = phi - dt*L*(-gamma (C-Ca)^2 + delta phi^3 +
    epsilon phi (sum - phi^2)), evaluated here in Python floats from
    the model's definition (sum over the order parameters)."""
    q, steps = 3, 5
    vals = [0.3, 0.62, 0.17]
    a = np.broadcast_to(np.array(vals), (6, 5, q)).copy()
    p = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
    out = O.ostwald(a, p, steps)
    dt, L, Ca, Cb, delta, eps = 0.001, 1.0, 0.05, 0.95, 1.0, 3.0
    gamma = 2.0 / (Cb - Ca) ** 2
    C, phi = vals[0], list(vals[1:])
    for _ in range(steps):
        s = sum(v * v for v in phi)
        phi = [v - dt * L * (-gamma * (C - Ca) ** 2 + delta * v ** 3 + eps * v * (s - v * v)) for v in phi]
    assert np.all(out[..., 0] == C)
    for j in range(1, q):
        assert np.allclose(out[..., j], phi[j - 1], rtol=1e-13, atol=0)


def _grains(shape, q, seed):
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.one_hot_grains(shape, q, seed)


@pytest.mark.parametrize("form", ["f64", "f32"])
def test_multiphase_grain_growth_stays_on_the_simplex(form):
    """After every step each node's fields lie in [0, 1] and sum to 1 (the
    projection), nodes whose neighbourhood holds one grain keep their values,
    and relabelling the grains relabels the result (uniform energies)."""
    shape, q = (30, 26), 5
    a = _grains(shape, q, 41)
    if form == "f32":
        a = a.astype(np.float32)
    p = [0.01, 8.0, 1.0, 1.0]
    out = O.gg_aniso(a, p, 1, form=form)
    tol = 4e-16 if form == "f64" else 4e-7
    assert out.min() >= 0 and out.max() <= 1 and np.abs(out.sum(-1) - 1).max() <= tol
    # one grain in the whole 5-point neighbourhood: copied
    lone = np.ones(shape, dtype=bool)
    for ax in range(2):
        for sh in (-1, 1):
            lone &= (np.roll(a, sh, axis=ax) == a).all(-1)
    assert np.array_equal(out[lone], a[lone]) and lone.any() and (~lone).any()
    later = O.gg_aniso(a, p, 15, form=form)
    perm = [3, 0, 4, 1, 2]
    assert np.allclose(O.gg_aniso(a[..., perm], p, 15, form=form), later[..., perm], atol=50 * tol)


def test_thread_count_does_not_change_the_bits():
    """The scalar steppers' node loops run on OpenMP threads from 2^16 nodes;
    one thread and several give the same bits in every form."""
    a3 = O.uniform(48 * 40 * 40, 91, 0.45, 0.55).reshape(48, 40, 40)
    a2 = O.uniform(260 * 256, 92, -0.3, 0.3).reshape(260, 256)
    runs = [lambda f: O.ch(a3, [1.0, 1.0, 0.005], 3, mu=O.MU_DW, form=f),
            lambda f: O.ch(a2, [1.0, 1.0, 1.0, 1.0, 0.005], 3, mu=O.MU_POLY, form=f),
            lambda f: O.ac(a2, [1.0, 1.0, 1.0, 1.0, 0.01], 3, form=f),
            lambda f: O.ac_noise(a2, [1.0, 1.0, 1.0, 1.0, 0.01, 0.7, 5], 3, form=f),
            lambda f: O.ch_noise(a3, [1.0, 1.0, 1.0, 1.0, 0.005, 0.3, 6], 3, form=f)]
    before = O.set_threads(0)
    try:
        for form in ("f64", "f32", "f32ref"):
            O.set_threads(1)
            one = [r(form) for r in runs]
            O.set_threads(4)
            four = [r(form) for r in runs]
            for k, (x, y) in enumerate(zip(one, four)):
                assert np.array_equal(x, y), (form, k)
    finally:
        O.set_threads(before)
