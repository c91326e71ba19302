"""Pins of the oracle's trajectory analyses (g(r), MSD, VACF and the two
diffusion coefficients): the reference's own closed-form cases
(tests/test_analysis.cpp:19-157, 228-307), exact values the normalisation
fixes, brute-force numpy restatements on small random trajectories, and bit
equality with the reference compiled from its own sources where built."""
import numpy as np
import pytest

PI = 3.14159265358979323846


def doctest_approx(actual, expected, eps):
    """doctest::Approx(expected).epsilon(eps) as the reference's tests use it:
    |a - e| < eps * (1 + max(|a|, |e|)) (the default scale of 1)."""
    a, e = np.asarray(actual), np.asarray(expected)
    return np.all(np.abs(a - e) < eps * (1.0 + np.maximum(np.abs(a), np.abs(e))))


def _pairs_r(L, x):
    """All i<j minimum-image distances, numpy, a different evaluation order."""
    d = x[:, None, :] - x[None, :, :]
    d = d - L * np.floor(d / L + 0.5)
    r = np.sqrt((d * d).sum(axis=-1))
    iu = np.triu_indices(x.shape[0], 1)
    return r[iu]


def test_rdf_ideal_gas_is_one(orc):
    # test_analysis.cpp:19-38: 40 frames of 1500 uniform atoms in a 20^3 box
    rng = np.random.default_rng(4242)
    L = np.array([20.0, 20.0, 20.0])
    frames = (rng.uniform(size=(40, 1500, 3)) - 0.5) * 20.0
    g, r, _ = orc.rdf(L, frames, 8.0, 80)
    assert np.all(np.abs(g[5:] - 1.0) <= 0.06), np.max(np.abs(g[5:] - 1.0))


def test_rdf_two_atoms_one_bin_exact_value(orc):
    # test_analysis.cpp:40-61, plus the value the normalisation fixes:
    # g_k = 1 / (frames * N/2 * rho * shell_k) with N = 2, rho = 2 / V
    L = np.array([20.0, 20.0, 20.0])
    x = np.array([[[0.0, 0.0, 0.0], [3.05, 0.0, 0.0]]])
    g, r, h = orc.rdf(L, x, 8.0, 80)
    k = int(3.05 / 0.1)
    assert h[k] == 1.0 and h.sum() == 1.0
    w = 8.0 / 80
    shell = 4.0 / 3.0 * PI * (((k + 1) * w) ** 3 - (k * w) ** 3)
    assert g[k] == pytest.approx(1.0 / (0.5 * 2 * (2 / 8000.0) * shell), rel=1e-14)
    assert np.all(g[np.arange(80) != k] == 0.0)
    assert r[k] == pytest.approx((k + 0.5) * w, rel=1e-15)


def test_rdf_histogram_matches_numpy_brute(orc):
    rng = np.random.default_rng(7)
    L = np.array([7.0, 11.0, 13.0])
    frames = (rng.uniform(size=(3, 300, 3)) - 0.5) * L
    _, _, h = orc.rdf(L, frames, 3.4, 34)
    want = np.zeros(34)
    for f in frames:
        r = _pairs_r(L, f)
        r = r[r < 3.4]
        np.add.at(want, (r / 0.1).astype(int), 1.0)
    assert np.array_equal(h, want)


def test_rdf_refuses_rmax_beyond_half_box(orc):
    L = np.array([10.0, 6.0, 10.0])
    with pytest.raises(ValueError, match="rMax"):
        orc.rdf(L, np.zeros((1, 4, 3)), 3.1, 10)


def test_msd_ballistic_limit(orc):
    # test_analysis.cpp:63-96
    v = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.5, 0.5, 0.5]])
    dt, max_lag = 0.1, 50
    frames = np.stack([v * (f * dt) for f in range(max_lag + 1)])
    m, _ = orc.msd(np.full(3, 1e6), frames, max_lag)
    v2 = (v * v).sum(axis=1).mean()
    t = np.arange(max_lag + 1) * dt
    np.testing.assert_allclose(m, v2 * t * t, rtol=1e-10, atol=0)


def test_msd_unwraps_boundary_crossings(orc):
    # test_analysis.cpp:98-134: one atom at vx = 1.7 in a box of 4, 2000 frames
    Lb, dt, vx, max_lag = 4.0, 0.05, 1.7, 1500
    t = np.arange(2000) * dt
    x = vx * t
    xw = x - Lb * np.floor(x / Lb + 0.5)
    frames = np.zeros((2000, 1, 3))
    frames[:, 0, 0] = xw
    m, _ = orc.msd(np.full(3, Lb), frames, max_lag)
    lags = np.arange(0, max_lag + 1, 100)
    np.testing.assert_allclose(m[lags], (vx * lags * dt) ** 2, rtol=1e-8, atol=0)
    assert m[max_lag] > Lb * Lb


def test_msd_vacf_match_numpy_brute(orc):
    # every time origin, every lag, against a direct numpy average of the
    # unwrapped displacement (known here because the walk is generated
    # unwrapped and then folded into the box)
    rng = np.random.default_rng(11)
    n, nf, max_lag = 50, 40, 12
    L = np.array([3.0, 4.0, 5.0])
    steps = rng.normal(scale=0.4, size=(nf, n, 3))
    steps[0] = (rng.uniform(size=(n, 3)) - 0.5) * L
    true = np.cumsum(steps, axis=0)
    wrapped = true - L * np.floor(true / L + 0.5)
    m, _ = orc.msd(L, wrapped, max_lag)
    want = np.zeros(max_lag + 1)
    for lag in range(max_lag + 1):
        d = true[lag:] - true[:nf - lag]
        want[lag] = (d * d).sum(axis=-1).mean()
    np.testing.assert_allclose(m, want, rtol=1e-11, atol=1e-14)

    vel = rng.normal(size=(nf, n, 3))
    c, _ = orc.vacf(vel, max_lag)
    u = np.array([(vel[lag:] * vel[:nf - lag]).sum(axis=-1).mean() for lag in range(max_lag + 1)])
    np.testing.assert_allclose(c, u / u[0], rtol=1e-12, atol=1e-14)


def test_vacf_free_particles_is_one(orc):
    # test_analysis.cpp:136-157
    v = np.array([[1.0, 0.0, 0.0], [0.0, -2.0, 0.5], [0.3, 0.3, 0.3], [-1.5, 0.0, 2.0]])
    frames = np.stack([v] * 46)
    c, _ = orc.vacf(frames, 40)
    np.testing.assert_allclose(c, 1.0, rtol=1e-10, atol=0)


def _ou_trajectory(n, dt, gamma, kT, equil, steps, seed):
    """Free-particle BAOAB with zero force (exact OU velocity kernel between
    two half drifts), positions unwrapped in a box far larger than the walk."""
    rng = np.random.default_rng(seed)
    c1 = np.exp(-gamma * dt)
    c2 = np.sqrt((1 - c1 * c1) * kT)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    xs, vs = [], []
    for k in range(equil + steps):
        x += 0.5 * dt * v
        v = c1 * v + c2 * rng.normal(size=(n, 3))
        x += 0.5 * dt * v
        if k >= equil - 1:
            xs.append(x.copy())
            vs.append(v.copy())
    return np.stack(xs), np.stack(vs)


def test_ou_diffusion_coefficients(orc):
    # test_analysis.cpp:228-307: D = kT / (m gamma) = 1 from both the Einstein
    # fit and the Green-Kubo integral, within 3%; C(t) = exp(-gamma t)
    dt, max_lag = 0.2, 100
    xs, vs = _ou_trajectory(6000, dt, 1.0, 1.0, 100, 750, 20260921)
    _, d_msd = orc.msd(np.full(3, 1000.0), xs, max_lag, dt, int(10.0 / dt), max_lag + 1)
    c, d_vacf = orc.vacf(vs, max_lag, dt)
    assert d_msd == pytest.approx(1.0, rel=0.03)
    assert d_vacf == pytest.approx(1.0, rel=0.03)
    assert abs(d_msd - d_vacf) < 0.03
    lags = np.arange(0, max_lag + 1, 10)
    assert doctest_approx(c[lags], np.exp(-lags * dt), 0.05)


def test_diffusion_fit_edge_cases(orc):
    # fewer than two points in the fit window, and a single pushed frame
    xs = np.zeros((5, 2, 3))
    _, d = orc.msd(np.full(3, 10.0), xs, 3, 0.1, 3, 4)
    assert d == 0.0
    _, d = orc.vacf(np.ones((1, 2, 3)), 3, 0.1)
    assert d == 0.0
    c, _ = orc.vacf(np.zeros((3, 2, 3)), 3)
    assert np.all(c == 0.0)


def test_analysis_matches_reference_build(orc, ref):
    rng = np.random.default_rng(3)
    L = np.array([6.0, 7.0, 8.0])
    frames = (rng.uniform(size=(4, 200, 3)) - 0.5) * L * 1.3  # some outside the box
    g0, c0, _ = orc.rdf(L, frames, 2.9, 29)
    g1, c1, _ = ref.rdf(L, frames, 2.9, 29)
    assert np.array_equal(g0, g1) and np.array_equal(c0, c1)
    walk = np.cumsum(rng.normal(scale=0.5, size=(30, 64, 3)), axis=0)
    wrapped = walk - L * np.floor(walk / L + 0.5)
    m0, d0 = orc.msd(L, wrapped, 9, 0.05, 2, 10)
    m1, d1 = ref.msd(L, wrapped, 9, 0.05, 2, 10)
    assert np.array_equal(m0, m1) and d0 == d1
    v = rng.normal(size=(30, 64, 3))
    v0, e0 = orc.vacf(v, 9, 0.05)
    v1, e1 = ref.vacf(v, 9, 0.05)
    assert np.array_equal(v0, v1) and e0 == e1
