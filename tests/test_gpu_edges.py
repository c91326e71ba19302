"""Edge cases of the plain path through the C ABI: the smallest systems, a
system with no pair inside the list radius, and the calls that are refused.

Closed forms (Truncated LJ, sigma = epsilon = 1): u = 4 (s^12 - s^6),
w = 24 (2 s^12 - s^6) with s = 1/r, and F_i = w / r^2 (r_i - r_j), the
reference's pair terms (forcefield/lennard_jones.cpp:18-40). Tolerances: FP64
to 1e-12 relative; mixed precision to 1e-5 relative, the FP32 rounding of
one pair term: at r = 1.1, w = 24 (2 s^12 - s^6) keeps one ninth of its
largest part, so a few FP32 roundings reach 2e-6 of w.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair_terms(d):
    r2 = float(d @ d)
    s6 = (1.0 / r2) ** 3
    u = 4.0 * (s6 * s6 - s6)
    w = 24.0 * (2.0 * s6 * s6 - s6)
    return u, w, (w / r2) * d


def _two_atoms():
    p0 = np.array([0.3, -0.2, 0.1])
    p1 = p0 + 1.1 * np.array([2.0, 3.0, 6.0]) / 7.0
    return np.array([10.0, 10.0, 10.0]), np.stack([p0, p1])


@pytest.mark.parametrize("source,precision", [(0, 0), (1, 0), (2, 0), (2, 1)])
def test_two_atoms_closed_form(ctx, source, precision):
    from moleculardynamics_amd import native

    L, x = _two_atoms()
    ctx.upload(L, x)
    ctx.set_lj(1.0, 1.0, 2.5, False)
    ctx.config_list(2.5, 0.3)
    ctx.build_list()
    e, w = ctx.forces(source, precision)
    f = ctx.download()[2]
    u0, w0, f0 = _pair_terms(x[0] - x[1])
    tol = 1e-12 if precision == native.FP64 else 1e-5
    assert abs(e - u0) <= tol * abs(u0)
    assert abs(w - w0) <= tol * abs(w0)
    assert np.max(np.abs(f[0] - f0)) <= tol * np.max(np.abs(f0))
    assert np.max(np.abs(f[1] + f0)) <= tol * np.max(np.abs(f0))
    assert ctx.pair_count() == 1
    if source == native.VERLET:
        assert ctx.pairs().tolist() == [[0, 1]]


def test_two_atoms_mixed_run_follows_fp64(ctx):
    """20 velocity-Verlet steps of a bound pair: mixed and FP64 trajectories
    agree to 1e-6, and both conserve the energy."""
    from moleculardynamics_amd import native

    L, x = _two_atoms()
    v = np.array([[0.1, 0.0, -0.2], [-0.1, 0.0, 0.2]])
    out = []
    for prec in (native.FP64, native.MIXED):
        ctx.upload(L, x, v)
        ctx.set_lj(1.0, 1.0, 2.5, False)
        ctx.config_list(2.5, 0.3)
        ctx.build_list()
        ctx.forces(native.VERLET, prec)
        ctx.set_integrator(kind=native.VV, dt=0.002)
        th = ctx.run(20, prec, thermo=True)
        out.append((ctx.download(), th))
    (s64, t64), (s32, t32) = out
    assert np.max(np.abs(s64[0] - s32[0])) < 1e-6
    assert np.max(np.abs(s64[1] - s32[1])) < 1e-6
    for th in (t64, t32):
        etot = th["pe"] + th["ke"]
        assert np.max(np.abs(etot - etot[0])) < 1e-3 * abs(etot[0])
        assert np.all(th["pairs"] == 1)


def test_dilute_gas_moves_ballistically(ctx):
    """27 atoms 4 apart (beyond rc + skin): no list entries, zero energy,
    virial and forces; 50 steps leave the velocities bit for bit unchanged and
    move the positions by 50 dt v (modulo the box)."""
    from moleculardynamics_amd import native

    g = np.arange(3) * 4.0 - 4.0
    x = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
    rng = np.random.default_rng(3)
    v = rng.normal(0.0, 0.5, size=x.shape)
    L = np.array([12.0, 12.0, 12.0])
    ctx.upload(L, x, v)
    ctx.set_lj(1.0, 1.0, 2.5, False)
    ctx.config_list(2.5, 0.3)
    ctx.build_list()
    e, w = ctx.forces(native.VERLET, native.FP64)
    assert e == 0.0 and w == 0.0 and ctx.pairs().shape == (0, 2)
    assert np.all(ctx.download()[2] == 0.0)
    ctx.set_integrator(kind=native.VV, dt=0.005)
    th = ctx.run(50, native.FP64, thermo=True)
    assert np.all(th["pe"] == 0.0) and np.all(th["pairs"] == 0)
    x1, v1, f1 = ctx.download()
    assert np.array_equal(v1, v) and np.all(f1 == 0.0)
    want = x + 50 * 0.005 * v
    d = (x1 - want) - L * np.round((x1 - want) / L)  # equal modulo the box
    assert np.max(np.abs(d)) < 1e-12


@pytest.mark.parametrize("precision", [0, 1])
def test_single_atom(ctx, precision):
    from moleculardynamics_amd import native

    x = np.array([[0.5, -1.0, 2.0]])
    v = np.array([[0.2, 0.1, -0.3]])
    ctx.upload(np.array([5.0, 5.0, 5.0]), x, v)
    ctx.set_lj(1.0, 1.0, 2.5, False)
    ctx.config_list(2.5, 0.3)
    ctx.build_list()
    e, w = ctx.forces(native.VERLET, precision)
    assert e == 0.0 and w == 0.0
    ctx.set_integrator(kind=native.VV, dt=0.01)
    ctx.run(10, precision, thermo=False)
    x1, v1, f1 = ctx.download()
    assert np.array_equal(v1, v) and np.all(f1 == 0.0)
    assert np.max(np.abs(x1 - (x + 0.1 * v))) < 1e-12


def test_refused_calls(ctx):
    """No atoms, and a run before the list is built, are errors that leave the
    context usable."""
    from moleculardynamics_amd import native

    with pytest.raises(native.MDError):
        ctx.upload(np.array([5.0, 5.0, 5.0]), np.zeros((0, 3)))
    L, x = _two_atoms()
    ctx.upload(L, x)
    ctx.set_lj(1.0, 1.0, 2.5, False)
    ctx.set_integrator(kind=native.VV, dt=0.005)
    with pytest.raises(native.MDError):
        ctx.run(5, native.FP64, thermo=False)
    ctx.config_list(2.5, 0.3)
    ctx.build_list()
    ctx.forces(native.VERLET, native.FP64)
    ctx.run(5, native.FP64, thermo=False)


def test_row_overflow_during_run_recovers(gpu_ok, orc):
    """A row capacity that fits the first build exactly overflows as the
    fluid's rows fluctuate: md_run restores its starting state, doubles the
    capacity, builds again and reruns the call. The result is the trajectory
    of a run with ample capacity (the forces are the same pair sums; the extra
    list build only changes summation order) and no MD_ERR_CAPACITY."""
    from moleculardynamics_amd import native

    fr = orc.fcc_lattice(6, 0.8442, 1.2, 77)
    x, v = np.stack([fr.x, fr.y, fr.z], 1), np.stack([fr.vx, fr.vy, fr.vz], 1)
    pairs = orc.verlet_pairs(fr, 2.5, 0.3)
    rows = np.bincount(pairs.ravel(), minlength=fr.n)
    tight = int(rows.max() + 7) // 8 * 8
    out = []
    for cap in (tight, 0):
        c = native.Context()
        c.upload(fr.L, x, v)
        c.set_lj(1.0, 1.0, 2.5, False)
        c.config_list(2.5, 0.3)
        c.set_capacity(cap)
        c.build_list()
        c.forces()
        c.set_integrator(kind=native.VV, dt=0.005)
        b0 = c.rebuild_count()
        th = c.run(300)
        out.append((th, c.download(), c.rebuild_count() - b0))
        c.close()
    (th_t, st_t, nb_t), (th_a, st_a, nb_a) = out
    assert nb_t >= nb_a + 1  # the recovery rebuilt at least once more
    assert np.max(np.abs(th_t["pe"] - th_a["pe"]) / np.abs(th_a["pe"])) < 1e-8
    d = st_t[0] - st_a[0]
    d -= fr.L * np.floor(d / fr.L + 0.5)
    assert np.max(np.abs(d)) < 1e-7
