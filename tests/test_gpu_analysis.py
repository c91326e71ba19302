"""GPU g(r), MSD and VACF (md_rdf_*, md_msd_*, md_vacf_*) against the oracle
on frames of real device trajectories (atoms re-sorted by list rebuilds in
between), and the reference's closed-form cases (tests/test_analysis.cpp:
19-157, 228-307) run through the device path.

Bars: g(r) pair counts are integers and must equal the oracle's exactly (and
so must g, whose normalisation is the same host arithmetic). MSD and VACF
per-atom terms equal the oracle's bit for bit; their sums over atoms run in
a different order (a workgroup tree instead of a sequential loop), so the
per-lag values are compared at 1e-12 relative."""
import numpy as np
import pytest

from moleculardynamics_amd import native

pytestmark = pytest.mark.gpu

CORR_RTOL = 1e-12


def doctest_approx(actual, expected, eps):
    """doctest::Approx(expected).epsilon(eps): |a - e| < eps (1 + max(|a|, |e|))."""
    a, e = np.asarray(actual), np.asarray(expected)
    return bool(np.all(np.abs(a - e) < eps * (1.0 + np.maximum(np.abs(a), np.abs(e)))))


def _lj_context(fr, rc=2.5, skin=0.3):
    ctx = native.Context()
    ctx.upload(fr.L, np.stack([fr.x, fr.y, fr.z], axis=1), np.stack([fr.vx, fr.vy, fr.vz], axis=1))
    ctx.set_lj(1.0, 1.0, rc, True)
    ctx.config_list(rc, skin)
    ctx.build_list()
    ctx.forces(native.VERLET, native.FP64)
    return ctx


def test_rdf_counts_match_oracle_on_a_trajectory(gpu_ok, orc):
    fr = orc.fcc_lattice(6, 0.8442, 1.5, 77)  # 864 atoms, L = 10.08: 3 cells of 3.3
    ctx = _lj_context(fr)
    try:
        ctx.set_integrator(kind=native.VV, dt=0.005)
        r_all = native.Rdf(3.3, 66)
        r_cel = native.Rdf(3.3, 66)
        frames = []
        for k in range(4):
            ctx.run(25, thermo=False)
            print(f"rdf frame {k}", flush=True)
            r_all.accumulate(ctx, native.RDF_ALL_PAIRS)
            r_cel.accumulate(ctx, native.RDF_CELLS)
            frames.append(ctx.download()[0])
        assert ctx.rebuild_count() > 1
        g_o, c_o, h_o = orc.rdf(fr.L, np.stack(frames), 3.3, 66)
        for r in (r_all, r_cel):
            c, g, h, nf = r.result()
            assert nf == 4
            assert np.array_equal(h.astype(np.float64), h_o)
            assert np.array_equal(g, g_o) and np.array_equal(c, c_o)
    finally:
        ctx.close()


@pytest.mark.parametrize("source", [native.RDF_ALL_PAIRS, native.RDF_CELLS])
def test_rdf_orthorhombic_random_frame(gpu_ok, orc, source):
    # 12,000 uniform atoms in a (21, 33, 39) box, some outside it: all three
    # fold branches, every bin populated
    rng = np.random.default_rng(5)
    L = np.array([21.0, 33.0, 39.0])
    pos = (rng.uniform(size=(12000, 3)) - 0.5) * L * 1.2
    ctx = native.Context()
    try:
        ctx.upload(L, pos)
        r = native.Rdf(6.9, 345)
        r.accumulate(ctx, source)
        c, g, h, _ = r.result()
        _, _, h_o = orc.rdf(L, pos[None], 6.9, 345)
        assert np.array_equal(h.astype(np.float64), h_o)
    finally:
        ctx.close()


def test_rdf_two_atoms_and_refusals(gpu_ok):
    # test_analysis.cpp:40-61
    ctx = native.Context()
    try:
        ctx.upload(np.full(3, 20.0), np.array([[0.0, 0.0, 0.0], [3.05, 0.0, 0.0]]))
        r = native.Rdf(8.0, 80)
        r.accumulate(ctx)
        c, g, h, _ = r.result()
        k = int(3.05 / 0.1)
        assert h[k] == 1 and h.sum() == 1 and g[k] > 0.0
        assert np.all(g[np.arange(80) != k] == 0.0)
        with pytest.raises(native.MDError, match="rMax exceeds half"):
            native.Rdf(10.5, 10).accumulate(ctx)
        with pytest.raises(native.MDError, match="fewer than 3 cells"):
            native.Rdf(7.0, 10).accumulate(ctx, native.RDF_CELLS)
        with pytest.raises(native.MDError, match="atoms"):
            native.Msd(3, 4).push(ctx)
    finally:
        ctx.close()


def test_rdf_ideal_gas_is_one(gpu_ok):
    # test_analysis.cpp:19-38 (numpy's generator instead of mt19937_64)
    rng = np.random.default_rng(4242)
    ctx = native.Context()
    try:
        r = native.Rdf(8.0, 80)
        for _ in range(40):
            ctx.upload(np.full(3, 20.0), (rng.uniform(size=(1500, 3)) - 0.5) * 20.0)
            r.accumulate(ctx)
        _, g, _, _ = r.result()
        assert np.all(np.abs(g[5:] - 1.0) <= 0.06)
    finally:
        ctx.close()


def test_msd_vacf_match_oracle_on_a_trajectory(gpu_ok, orc):
    fr = orc.fcc_lattice(6, 0.7, 1.8, 91)  # a hot liquid: atoms diffuse and cross the box
    ctx = _lj_context(fr)
    try:
        ctx.set_integrator(kind=native.BAOAB, dt=0.005, friction=1.0, temperature=1.8, seed=3)
        n = fr.x.size
        msd, vacf = native.Msd(n, 12), native.Vacf(n, 12)
        pos, vel, L = [], [], []
        for k in range(40):
            if k:
                ctx.run(8, thermo=False)
            msd.push(ctx)
            vacf.push(ctx)
            p, v, _ = ctx.download()
            pos.append(p)
            vel.append(v)
        assert ctx.rebuild_count() > 3
        m_o, d_o = orc.msd(fr.L, np.stack(pos), 12, 0.04, 3, 13)
        c_o, e_o = orc.vacf(np.stack(vel), 12, 0.04)
        m = msd.result()
        np.testing.assert_allclose(m, m_o, rtol=CORR_RTOL, atol=1e-15)
        assert m[12] > 0.05  # it moved
        # C(t) changes sign: near its zeros the bound is absolute
        np.testing.assert_allclose(vacf.result(), c_o, rtol=CORR_RTOL, atol=1e-13)
        assert msd.diffusion(0.04, 3, 13) == pytest.approx(d_o, rel=1e-10)
        assert vacf.diffusion(0.04) == pytest.approx(e_o, rel=1e-10)
        u = vacf.result(normalised=False)
        assert u[0] == pytest.approx(np.mean([(v * v).sum(axis=1).mean() for v in vel]), rel=1e-12)
    finally:
        ctx.close()


def test_msd_ballistic_and_unwrapping(gpu_ok):
    # test_analysis.cpp:63-96: msd(t) = <v^2> t^2
    v = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.5, 0.5, 0.5]])
    ctx = native.Context()
    try:
        ctx.upload(np.full(3, 1e6), np.zeros((3, 3)))
        m = native.Msd(3, 50)
        for f in range(51):
            ctx.set_positions(v * (f * 0.1))
            m.push(ctx)
        t = np.arange(51) * 0.1
        assert doctest_approx(m.result(), (v * v).sum(axis=1).mean() * t * t, 1e-10)
        # test_analysis.cpp:98-134: one atom at vx = 1.7 crossing a box of 4
        ctx.upload(np.full(3, 4.0), np.zeros((1, 3)))
        m = native.Msd(1, 1500)
        for f in range(2000):
            x = 1.7 * f * 0.05
            ctx.set_positions(np.array([[x - 4.0 * np.floor(x / 4.0 + 0.5), 0.0, 0.0]]))
            m.push(ctx)
        r = m.result()
        lags = np.arange(0, 1501, 100)
        assert doctest_approx(r[lags], (1.7 * lags * 0.05) ** 2, 1e-8)
        assert r[1500] > 16.0
    finally:
        ctx.close()


def test_vacf_free_particles_is_one(gpu_ok):
    # test_analysis.cpp:136-157
    v = np.array([[1.0, 0.0, 0.0], [0.0, -2.0, 0.5], [0.3, 0.3, 0.3], [-1.5, 0.0, 2.0]])
    ctx = native.Context()
    try:
        ctx.upload(np.full(3, 100.0), np.zeros((4, 3)) + np.arange(4)[:, None], v)
        c = native.Vacf(4, 40)
        for _ in range(46):
            c.push(ctx)
        assert doctest_approx(c.result(), np.ones(41), 1e-10)
    finally:
        ctx.close()


def test_ou_diffusion_through_md_run(gpu_ok):
    """test_analysis.cpp:228-307 on the fused device loop: 6000 free particles
    under LangevinBAOAB (LJ with epsilon = 0, so every force is exactly 0),
    kT = m = gamma = 1, dt 0.2: D = kT / (m gamma) = 1 within 3% from both the
    Einstein fit and the Green-Kubo integral, C(t) = exp(-t) within the
    reference's tolerance. The noise is the counter-based stream, not the
    reference's mt19937_64 sequence, so this is the statistical pin."""
    rng = np.random.default_rng(20260921)
    n, dt, max_lag = 6000, 0.2, 100
    L = np.full(3, 1000.0)
    ctx = native.Context()
    try:
        ctx.upload(L, (rng.uniform(size=(n, 3)) - 0.5) * L)
        ctx.set_lj(1.0, 0.0, 25.0, False)
        ctx.config_list(25.0, 25.0)
        ctx.build_list()
        ctx.forces(native.VERLET, native.FP64)
        ctx.set_integrator(kind=native.BAOAB, dt=dt, friction=1.0, temperature=1.0, seed=20260921)
        ctx.run(100, thermo=False)
        msd, vacf = native.Msd(n, max_lag), native.Vacf(n, max_lag)
        msd.push(ctx)
        vacf.push(ctx)
        for _ in range(750):
            ctx.run(1, thermo=False)
            msd.push(ctx)
            vacf.push(ctx)
        d_msd = msd.diffusion(dt, int(10.0 / dt), max_lag + 1)
        d_vacf = vacf.diffusion(dt)
        print(f"D(msd) = {d_msd:.5f}  D(vacf) = {d_vacf:.5f}", flush=True)
        assert d_msd == pytest.approx(1.0, rel=0.03)
        assert d_vacf == pytest.approx(1.0, rel=0.03)
        assert abs(d_msd - d_vacf) < 0.03
        c = vacf.result()
        lags = np.arange(0, max_lag + 1, 10)
        assert doctest_approx(c[lags], np.exp(-lags * dt), 0.05)
    finally:
        ctx.close()
