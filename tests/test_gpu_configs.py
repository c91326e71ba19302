"""GPU parity at the workloads' own sizes, against the CPU oracle on the same
inputs:

* the 1,048,576-atom LJ fluid (fccLattice(64, 0.8442, 0.722, 20260001), rc
  2.5, skin 0.3): FP64 E, W and F (1e-12), the in-cutoff pair count and the
  sorted Verlet pair set (bit-exact) on the lattice and on the snapshot after
  200 Nose-Hoover-chain steps, five NHC steps against the oracle's, and the
  mixed path's force error (RMS <= 1e-6 relative) on the snapshot;
* orthorhombic boxes (sides in the ratio 7 : 11 : 13 as the reference's
  tests/test_box.cpp:18 uses) above and below the one-lane threshold;
* crystallites of 32,768 and 32,769 atoms, either side of the size where the
  pair kernels change their lanes per atom and the list build from eight
  lanes per slot to one;
* the 864-atom NVE run (1000 velocity-Verlet steps) against the oracle;
* all 256 replicas of the batched recipe, primed forces against the oracle;
* SPC/E water at its configuration (4096 molecules, 12,288 sites, rc 9 A,
  tuned Ewald alpha and kmax): forces and energy terms, and three
  constrained velocity-Verlet steps, against the oracle.
"""
import numpy as np
import pytest

from oracle.pyoracle import Frame

pytestmark = pytest.mark.gpu

REL = 1e-12
C2 = dict(cells=64, rho=0.8442, T=0.722, seed=20260001)


def _pos(fr):
    return np.stack([fr.x, fr.y, fr.z], axis=1)


def _vel(fr):
    return np.stack([fr.vx, fr.vy, fr.vz], axis=1)


def _frame(L, x, v=None):
    L = np.asarray(L, dtype=float)
    if v is None:
        return Frame(L, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
    return Frame(L, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy(), v[:, 0].copy(),
                 v[:, 1].copy(), v[:, 2].copy())


def _force_err(f, g):
    """max_i |f_i - g_i| / max(|g|_max, 1)"""
    return float(np.max(np.abs(f - g)) / max(np.max(np.abs(g)), 1.0))


def _sorted_pairs(p):
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def _rms(f, g):
    return float(np.sqrt(np.sum((f - g) ** 2) / np.sum(g ** 2)))


def _check_snapshot(ctx, orc, fr, rc, skin, mixed_rms=1e-6):
    """Pair set, FP64 E/W/F and pair count against the oracle on one
    snapshot; mixed-path errors against FP64. Returns the mixed errors.

    E and W are compared with the oracle's compensated sums of the same pair
    terms: at millions of pairs the reference-order sequential FP64 sum is
    itself off by up to ~1e-10 relative (on a lattice its rounding errors do
    not cancel), far above the 1e-12 gate."""
    from moleculardynamics_amd import native

    ctx.build_list()
    got = ctx.pairs()
    want = _sorted_pairs(orc.verlet_pairs(fr, rc, skin))
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    del got, want
    e, w = ctx.forces(native.VERLET, native.FP64)
    n64 = ctx.pair_count()
    _, _, f = ctx.download()
    f0, _, _ = orc.lj_forces(fr, rc=rc, mode="cells")
    e0, w0, np0 = orc.lj_ev_compensated(fr, rc=rc, mode="cells")
    assert abs(e - e0) <= REL * abs(e0)
    assert abs(w - w0) <= REL * abs(w0)
    assert _force_err(f, f0) <= REL
    assert n64 == np0
    e32, w32 = ctx.forces(native.VERLET, native.MIXED)
    assert ctx.pair_count() == n64
    _, _, f32 = ctx.download()
    frms = float(np.sqrt(np.mean(np.sum(f * f, axis=1))))
    if frms < 1e-6:
        # a perfect lattice: every FP64 net force vanishes by symmetry, so the
        # error is judged against a pair force instead (|f| ~ 1 at r ~ 1.2)
        assert np.max(np.abs(f32 - f)) <= 1e-5
        rms = mx = float("nan")
    else:
        rms = _rms(f32, f)
        mx = float(np.max(np.linalg.norm(f32 - f, axis=1)) / frms)
        assert rms <= mixed_rms, rms
    # FP32 pair terms carry ~1e-7 relative rounding each; on a lattice every
    # pair of a shell rounds alike, so the sums keep that error. W cancels
    # near the potential minimum (a crystal at spacing 1.12 sits on it), which
    # scales the same rounding up relative to W.
    assert abs(e32 - e) <= 1e-7 * abs(e)
    assert abs(w32 - w) <= 1e-6 * abs(w)
    return dict(rms=rms, max=mx, e=abs(e32 - e) / abs(e), w=abs(w32 - w) / abs(w))


@pytest.fixture(scope="module")
def c2_lattice(orc):
    return orc.fcc_lattice(C2["cells"], C2["rho"], C2["T"], C2["seed"])


def test_c2_lattice_forces_pairs(gpu_ok, orc, c2_lattice):
    from moleculardynamics_amd import native

    fr = c2_lattice
    assert fr.n == 1048576
    ctx = native.Context()
    try:
        ctx.upload(fr.L, _pos(fr), _vel(fr))
        ctx.set_lj(1.0, 1.0, 2.5, False)
        ctx.config_list(2.5, 0.3)
        _check_snapshot(ctx, orc, fr, 2.5, 0.3)
    finally:
        ctx.close()


def test_c2_nhc_trajectory_and_snapshot(gpu_ok, orc, c2_lattice):
    """Five NHC steps match the oracle's (PE and H' to 1e-9: the oracle's
    sequential sum over 28 M pair terms carries ~1e-10 of rounding near the
    lattice; KE to 1e-10); after 200 the melting lattice's snapshot passes the
    snapshot checks, the mixed path within RMS 1e-6 of FP64."""
    from moleculardynamics_amd import native

    fr = c2_lattice
    n = fr.n
    ctx = native.Context()
    try:
        ctx.upload(fr.L, _pos(fr), _vel(fr))
        ctx.set_lj(1.0, 1.0, 2.5, False)
        ctx.config_list(2.5, 0.3)
        ctx.build_list()
        ctx.forces(native.VERLET, native.FP64)
        ctx.set_integrator(kind=native.NHC, dt=0.005, temperature=0.722, tau=0.5, chain_length=3,
                           dof=3 * n - 3)
        th = ctx.run(5, native.FP64)
        sim = orc.sim(fr, kind="nhc", dt=0.005, rc=2.5, lfs=False, skin=0.3, T=0.722, tau=0.5,
                      M=3, dof=3 * n - 3)
        ot = sim.run(5)
        del sim
        assert np.max(np.abs(th["pe"] - ot["pe"]) / np.abs(ot["pe"])) < 1e-9
        assert np.max(np.abs(th["ke"] - ot["ke"]) / np.abs(ot["ke"])) < 1e-10
        h = th["pe"] + th["ke"] + th["conserved"]
        oh = ot["pe"] + ot["ke"] + ot["conserved"]
        assert np.max(np.abs(h - oh) / np.abs(oh)) < 1e-9
        ctx.run(195, native.FP64, thermo=False)
        x, v, _ = ctx.download()
        errs = _check_snapshot(ctx, orc, _frame(fr.L, x), 2.5, 0.3)
        print("C2 mixed vs FP64 after 200 NHC steps:", errs)
    finally:
        ctx.close()


def test_sc_lattice_closed_form(gpu_ok, orc):
    """64,000 atoms on a simple-cubic lattice (every pair a known shell): the
    GPU's FP64 E and W equal the closed form to 1e-12 and its pair count the
    shell count. The mixed path's E is within 1e-7 (one FP32 rounding per
    term, alike for every pair of a shell) and W within 1e-5: the nearest
    shell (r = 1.1) sits by the potential minimum, where 24 (2 s^12 - s^6)
    cancels to an eighth of s^6, so the same rounding is eight times larger
    relative to W."""
    from moleculardynamics_amd import native
    from test_oracle import _sc_lattice_closed_form

    fr, e_cf, w_cf, np_cf = _sc_lattice_closed_form(40, 1.1, 2.5)
    ctx = native.Context()
    try:
        ctx.upload(fr.L, _pos(fr))
        ctx.set_lj(1.0, 1.0, 2.5, False)
        ctx.config_list(2.5, 0.3)
        ctx.build_list()
        e, w = ctx.forces(native.VERLET, native.FP64)
        assert ctx.pair_count() == np_cf
        assert abs(e - e_cf) <= REL * abs(e_cf) and abs(w - w_cf) <= REL * abs(w_cf)
        e32, w32 = ctx.forces(native.VERLET, native.MIXED)
        assert ctx.pair_count() == np_cf
        print("sc lattice mixed: E rel %.3e, W rel %.3e" % (abs(e32 - e_cf) / abs(e_cf),
                                                            abs(w32 - w_cf) / abs(w_cf)))
        assert abs(e32 - e_cf) <= 1e-7 * abs(e_cf) and abs(w32 - w_cf) <= 1e-5 * abs(w_cf)
    finally:
        ctx.close()


def _ortho_frame(sites, spacing, jitter, seed):
    """A jittered simple-cubic crystal filling an orthorhombic box."""
    nx, ny, nz = sites
    L = np.array([nx, ny, nz], dtype=float) * spacing
    g = [(np.arange(k) + 0.5) * spacing - 0.5 * Ld for k, Ld in zip(sites, L)]
    x = np.stack(np.meshgrid(*g, indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    x = x + rng.uniform(-jitter, jitter, size=x.shape)
    x -= L * np.floor(x / L + 0.5)
    v = rng.normal(size=x.shape)
    v -= v.mean(axis=0)
    return _frame(L, x, v)


@pytest.mark.parametrize("sites", [(21, 33, 39), (35, 55, 65)], ids=["27k", "125k"])
def test_orthorhombic_box(gpu_ok, orc, sites):
    """Sides 7 : 11 : 13 (a different fixed-point unit on every axis): pair set,
    forces and mixed error on the crystal and after 20 velocity-Verlet steps,
    whose thermo matches the oracle's."""
    from moleculardynamics_amd import native

    fr = _ortho_frame(sites, 1.1, 0.08, 7 + sites[0])
    ctx = native.Context()
    try:
        ctx.upload(fr.L, _pos(fr), _vel(fr))
        ctx.set_lj(1.0, 1.0, 2.5, False)
        ctx.config_list(2.5, 0.3)
        _check_snapshot(ctx, orc, fr, 2.5, 0.3)
        ctx.forces(native.VERLET, native.FP64)
        ctx.set_integrator(kind=native.VV, dt=0.005)
        th = ctx.run(20, native.FP64)
        sim = orc.sim(fr, kind="vv", dt=0.005, rc=2.5, lfs=False, skin=0.3)
        ot = sim.run(20)
        assert np.max(np.abs(th["pe"] - ot["pe"]) / np.abs(ot["pe"])) < 1e-10
        assert np.max(np.abs(th["ke"] - ot["ke"]) / np.abs(ot["ke"])) < 1e-10
        x, _, _ = ctx.download()
        ox, _, _ = sim.state()
        assert np.max(np.abs(x - ox)) < 1e-10
        _check_snapshot(ctx, orc, _frame(fr.L, x), 2.5, 0.3)
    finally:
        ctx.close()


def test_lane_threshold(gpu_ok, orc):
    """A 32^3 simple-cubic crystallite (32,768 atoms) in a larger box runs both
    pair kernels with 8 lanes per atom and the list build with 8 lanes per
    slot; with one isolated extra atom (32,769) the mixed pair kernel runs 4
    lanes per atom, the FP64 kernel 8 and the build one thread per slot. Both
    match the oracle (FP64 to 1e-12, pair sets bit-exact), and the mixed
    forces of the two agree within eight units of the 2^-32 fixed-point grid
    (each lane's share of a row is rounded to the grid once: at most 8 + 4
    half units apart)."""
    from moleculardynamics_amd import native

    m = 32
    g = (np.arange(m) - 0.5 * (m - 1)) * 1.12
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(9)
    x = x + rng.uniform(-0.06, 0.06, size=x.shape)
    L = np.array([48.0, 48.0, 48.0])
    assert x.shape[0] == 32768
    out = []
    for pos in (x, np.vstack([x, [[23.0, 23.0, 23.0]]])):
        fr = _frame(L, pos)
        c = native.Context()
        try:
            c.upload(L, pos)
            c.set_lj(1.0, 1.0, 2.5, False)
            c.config_list(2.5, 0.3)
            errs = _check_snapshot(c, orc, fr, 2.5, 0.3)
            c.forces(native.VERLET, native.MIXED)
            f32 = c.download()[2][:32768]
        finally:
            c.close()
        out.append((f32, errs))
    assert np.max(np.abs(out[0][0] - out[1][0])) <= 8.0 * 2.0 ** -32


def test_lj864_nve_1000_steps(gpu_ok, orc):
    """fccLattice(6, 0.8442, 0.722, 12345), Truncated rc 2.5, skin 0.3, VV
    dt 0.005: the whole 1000-step NVE run against the oracle. Both sum the
    same pair terms in different orders, so the trajectories separate from
    the last bits at the rate set by the fluid's Lyapunov exponent; the first
    400 steps agree to 1e-9, the rest in energy conservation and rebuilds."""
    from moleculardynamics_amd import native

    fr = orc.fcc_lattice(6, 0.8442, 0.722, 12345)
    ctx = native.Context()
    try:
        ctx.upload(fr.L, _pos(fr), _vel(fr))
        ctx.set_lj(1.0, 1.0, 2.5, False)
        ctx.config_list(2.5, 0.3)
        ctx.build_list()
        e0 = ctx.forces(native.VERLET, native.FP64)[0] + ctx.kinetic_energy()
        ctx.set_integrator(kind=native.VV, dt=0.005)
        th = ctx.run(1000, native.FP64)
        sim = orc.sim(fr, kind="vv", dt=0.005, rc=2.5, lfs=False, skin=0.3)
        ot = sim.run(1000)
        k = 400
        assert np.max(np.abs(th["pe"][:k] - ot["pe"][:k]) / np.abs(ot["pe"][:k])) < 1e-9
        assert np.max(np.abs(th["ke"][:k] - ot["ke"][:k]) / np.abs(ot["ke"][:k])) < 1e-9
        # Truncated potential: the energy jumps when pairs cross rc; both runs
        # conserve the total to the same order
        h, oh = th["pe"] + th["ke"], ot["pe"] + ot["ke"]
        d_gpu = np.max(np.abs(h - e0)) / abs(e0)
        d_orc = np.max(np.abs(oh - e0)) / abs(e0)
        assert d_gpu < 3.0 * d_orc + 1e-6 and d_gpu < 5e-3, (d_gpu, d_orc)
        assert abs(ctx.rebuild_count() - sim.rebuilds) <= 2
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_full_recipe_primed_forces(gpu_ok, orc, precision):
    """All 256 replicas of the batched recipe (108..4000 atoms, rho* in
    [0.5, 0.9)), jittered off the lattice: the primed forces of every replica
    against the oracle (FP64 to 1e-12; mixed RMS <= 2e-6 per replica: near a
    lattice the net forces are small against the pair forces they sum, so
    the relative RMS of the same FP32 rounding is larger than in a fluid)."""
    from moleculardynamics_amd import native

    cells, rho, T, seeds = native.batch_draw(256, 4000)
    systems = [native.fcc_lattice(int(c), float(r), float(t), int(s))
               for c, r, t, s in zip(cells, rho, T, seeds)]
    assert max(s[1].shape[0] for s in systems) == 4000
    rng = np.random.default_rng(21)
    systems = [(L, x + rng.uniform(-0.04, 0.04, size=x.shape), v) for L, x, v in systems]
    prec = native.FP64 if precision == "fp64" else native.MIXED
    b = native.Batch(systems, T, seeds, precision=prec)
    try:
        b.prime()
        _, _, forces = b.download()
    finally:
        b.close()
    worst = 0.0
    for (L, x, _), f in zip(systems, forces):
        fr = _frame(L, x)
        mode = "cells" if orc.cells_usable(L, 2.5) else "brute"
        f0, _, _ = orc.lj_forces(fr, rc=2.5, mode=mode)
        if precision == "fp64":
            worst = max(worst, _force_err(f, f0))
        else:
            worst = max(worst, _rms(f, f0))
    assert worst <= (REL if precision == "fp64" else 2e-6), worst


@pytest.fixture(scope="module")
def c3():
    from moleculardynamics_amd import native

    L = 49.71853
    pos, vel = native.water_lattice(16, L, 2026, 300.0)
    q2 = 4096 * (0.8476 ** 2 + 2 * 0.4238 ** 2)
    alpha, kmax = native.ewald_parameters(12288, [L] * 3, q2, 9.0, 1e-5)
    return L, pos, vel, alpha, kmax


def test_water_config3_forces(gpu_ok, orc, c3):
    from moleculardynamics_amd import native

    L, pos, vel, alpha, kmax = c3
    fr = _frame([L] * 3, pos)
    w = native.Water([L] * 3, pos, vel,
                     params=native.water_params(cutoff=9.0, kmax=kmax, alpha=alpha))
    try:
        t, pe = w.forces()
        _, _, f = w.download()
    finally:
        w.close()
    q = np.array([-0.8476 if i % 3 == 0 else 0.4238 for i in range(fr.n)])
    o, _ = orc.ewald(fr, q, np.arange(fr.n) // 3, rc=9.0, kmax=kmax, alpha=alpha)
    assert t["real"] == pytest.approx(o["real"], rel=1e-12)
    assert t["self"] == pytest.approx(o["self"], rel=1e-13)
    assert t["intra"] == pytest.approx(o["intra"], rel=1e-12)
    assert t["recip"] == pytest.approx(o["recip"], rel=1e-10)
    f0, pe0 = orc.spce_forces(fr, rc=9.0, kmax=kmax, alpha=alpha)
    # the total carries the reciprocal term's 1e-10 (its structure factors sum
    # 12,288 sites per k-vector in another order)
    assert pe == pytest.approx(pe0, rel=1e-10)
    assert np.max(np.abs(f - f0)) <= 1e-10 * np.max(np.abs(f0))


def test_water_config3_dynamics(gpu_ok, orc, c3):
    from moleculardynamics_amd import native

    L, pos, vel, alpha, kmax = c3
    fr = _frame([L] * 3, pos, vel)
    dt = 0.0018236739  # 2 fs in units of 1.0966875353 ps
    wo, pe0, ke0 = orc.water_run(fr, rc=9.0, kmax=kmax, alpha=alpha, dt=dt, nsteps=3)
    w = native.Water([L] * 3, pos, vel,
                     params=native.water_params(cutoff=9.0, kmax=kmax, alpha=alpha))
    try:
        w.forces()
        pe, ke = w.run(3, dt)
        p, _, _ = w.download()
    finally:
        w.close()
    assert np.max(np.abs(pe - pe0) / np.abs(pe0)) < 1e-10
    assert np.max(np.abs(ke - ke0) / np.abs(ke0)) < 1e-10
    d = p - np.stack([wo.x, wo.y, wo.z], 1)
    d -= L * np.floor(d / L + 0.5)
    assert np.max(np.abs(d)) < 1e-10


def test_water_config3_settle_follows_shake(gpu_ok, c3):
    """Config 3 under SETTLE and under SHAKE/RATTLE on the device: the same
    trajectory to the SHAKE tolerance over 20 steps, and the bonds held to
    1e-12 after 200 SETTLE steps."""
    from moleculardynamics_amd import native

    L, pos, vel, alpha, kmax = c3
    dt = 0.0018236739
    out = {}
    for method in (native.RIGID_SHAKE, native.RIGID_SETTLE):
        w = native.Water([L] * 3, pos, vel,
                         params=native.water_params(cutoff=9.0, kmax=kmax, alpha=alpha))
        try:
            w.set_constraint(method)
            w.forces()
            print("case: config 3, 20 steps, constraint", method, flush=True)
            out[method] = w.run(20, dt) + (w.download()[0],)
            if method == native.RIGID_SETTLE:
                w.run(180, dt)
                out["long"] = w.download()[0]
        finally:
            w.close()
    (pa, ka, xa), (pb, kb, xb) = out[native.RIGID_SHAKE], out[native.RIGID_SETTLE]
    assert np.max(np.abs(pa - pb) / np.abs(pa)) < 1e-9
    assert np.max(np.abs(ka - kb) / np.abs(ka)) < 1e-9
    d = xa - xb
    d -= L * np.floor(d / L + 0.5)
    assert np.max(np.abs(d)) < 1e-9
    hh = 2.0 * np.sin(0.5 * np.deg2rad(109.47))
    p = out["long"]
    for a, b, t in ((0, 1, 1.0), (0, 2, 1.0), (1, 2, hh)):
        e = p[a::3] - p[b::3]
        e -= L * np.floor(e / L + 0.5)
        assert np.max(np.abs(np.linalg.norm(e, axis=1) - t)) < 1e-12

