"""Batched uneven replicas (one set of launches per step for all replicas,
each split into pieces, HIP-graph segments) against the oracle and against
the single-system path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems(count, master=4000, max_cells=None):
    from moleculardynamics_amd import native

    cells, rho, T, seeds = native.batch_draw(count, master)
    if max_cells is not None:
        cells = np.minimum(cells, max_cells)
    systems = [native.fcc_lattice(int(c), float(r), float(t), int(s))
               for c, r, t, s in zip(cells, rho, T, seeds)]
    return systems, T, seeds, cells, rho


def test_recipe_draw_is_deterministic_and_in_range():
    from moleculardynamics_amd import native

    c1, r1, t1, s1 = native.batch_draw(256, 4000)
    c2, r2, t2, s2 = native.batch_draw(256, 4000)
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2) and np.array_equal(t1, t2)
    assert set(np.unique(c1)) <= set(range(3, 11)) and len(np.unique(c1)) == 8
    assert r1.min() >= 0.5 and r1.max() < 0.9 and t1.min() >= 0.7 and t1.max() < 1.5
    assert np.array_equal(s1, 4000 + np.arange(256))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_primed_forces_match_oracle(gpu_ok, orc, precision):
    from moleculardynamics_amd import native
    from oracle.pyoracle import Frame

    systems, T, seeds, _, _ = _systems(12, master=4100, max_cells=7)
    # off the lattice: on a perfect fcc lattice every net force is zero by symmetry
    rng = np.random.default_rng(11)
    systems = [(L, x + rng.uniform(-0.05, 0.05, size=x.shape), v) for L, x, v in systems]
    prec = native.FP64 if precision == "fp64" else native.MIXED
    b = native.Batch(systems, T, seeds, precision=prec)
    b.prime()
    _, _, forces = b.download()
    for (L, x, v), f in zip(systems, forces):
        fr = Frame(L, x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
        f0, _, _ = orc.lj_forces(fr, rc=2.5, mode="brute")
        if precision == "fp64":
            assert np.max(np.abs(f - f0)) <= 1e-12 * max(1.0, np.max(np.abs(f0)))
        else:
            rms = np.sqrt(np.sum((f - f0) ** 2) / np.sum(f0 ** 2))
            assert rms < 2e-6
    b.close()


def test_small_box_single_image_semantics(gpu_ok, orc):
    """N=108 at rho 0.9: L = 4.93 < 2 (rc + skin); forces follow the
    reference's single nearest image."""
    from moleculardynamics_amd import native
    from oracle.pyoracle import Frame

    L, x, v = native.fcc_lattice(3, 0.9, 1.0, 77)
    b = native.Batch([(L, x, v)], [1.0], [77], precision=native.FP64)
    b.prime()
    b.run(40, graph=False)
    pos, _, f = b.download()
    fr = Frame(L, pos[0][:, 0].copy(), pos[0][:, 1].copy(), pos[0][:, 2].copy())
    f0, _, _ = orc.lj_forces(fr, rc=2.5, mode="brute")
    assert np.max(np.abs(f[0] - f0)) <= 1e-11 * max(1.0, np.max(np.abs(f0)))
    b.close()


def test_graph_replay_matches_plain_launches(gpu_ok):
    from moleculardynamics_amd import native

    systems, T, seeds, _, _ = _systems(16, master=4200, max_cells=6)
    outs = []
    for graph in (False, True):
        b = native.Batch(systems, T, seeds, precision=native.MIXED)
        b.prime()
        th = b.run(200, segment=100, graph=graph, thermo=True)
        outs.append((b.download(), th))
        b.close()
    (p0, v0, f0), th0 = outs[0]
    (p1, v1, f1), th1 = outs[1]
    for a, c in zip(p0 + v0 + f0, p1 + v1 + f1):
        assert np.array_equal(a, c)
    assert np.array_equal(th0, th1)


def test_piece_split_keeps_trajectory(gpu_ok):
    """Rows, slot order and fixed-point sums do not depend on how a replica is
    split into workgroups: one piece per replica and the largest split give
    the same trajectory, E and W bit for bit, through several rebuilds (the
    kinetic energy is an FP64 sum whose grouping follows the split)."""
    from moleculardynamics_amd import native

    systems, T, seeds, _, _ = _systems(10, master=4600)
    n = np.array([s[1].shape[0] for s in systems])
    outs = []
    for pieces in (np.ones_like(n), np.maximum(1, np.minimum(32, n // 32))):
        b = native.Batch(systems, T, seeds, precision=native.MIXED)
        b.set_pieces(pieces)
        b.prime()
        th = b.run(150, segment=50, thermo=True)
        outs.append((b.download(), th, b.rebuild_counts()))
        b.close()
    (s0, th0, c0), (s1, th1, c1) = outs
    assert np.all(c0 > 1) and np.array_equal(c0, c1)
    for a, c in zip(s0[0] + s0[1] + s0[2], s1[0] + s1[1] + s1[2]):
        assert np.array_equal(a, c)
    assert np.array_equal(th0[..., [0, 1, 3]], th1[..., [0, 1, 3]])
    assert np.allclose(th0[..., 2], th1[..., 2], rtol=1e-12, atol=0.0)


def test_invalid_split_and_lanes_are_errors(gpu_ok):
    """Piece counts outside [1, min(32, n / 32)] and lanes per atom outside
    {1, 2, 4, 8, 16} are refused with an error, and the batch keeps working."""
    from moleculardynamics_amd import native

    systems, T, seeds, _, _ = _systems(3, master=4700, max_cells=5)
    n = np.array([s[1].shape[0] for s in systems])
    b = native.Batch(systems, T, seeds)
    b.prime()
    maxp = np.maximum(1, np.minimum(32, n // 32))
    with pytest.raises(native.MDError):
        b.set_pieces(np.zeros_like(n))
    with pytest.raises(native.MDError):
        b.set_pieces(maxp + 1)
    with pytest.raises(native.MDError):
        b.set_tpa(np.full(len(n), 3), 512)
    with pytest.raises(native.MDError):
        b.set_tpa(np.full(len(n), 4), 384)
    b.set_pieces(maxp)
    b.run(10, graph=False)
    b.close()


def test_batch_matches_single_system_path(gpu_ok):
    """Same seed, same counter-based noise, fixed-point forces: a replica in
    the batch and the same system on the plain path follow one trajectory.
    The batch kernel forms pair differences from the FP64 positions, the
    plain mixed kernel from cell-relative quantised positions (both within
    1e-8 of the exact difference). An FP32 rounding of d can then differ by
    one ulp, so the two agree to 1e-5 after 30 steps rather than bit for bit
    (a different noise stream or force field would differ by 1e-2)."""
    from moleculardynamics_amd import native

    systems, T, seeds, _, _ = _systems(4, master=4300, max_cells=5)
    b = native.Batch(systems, T, seeds, precision=native.MIXED)
    b.prime()
    b.run(30, graph=False)
    pos, vel, _ = b.download()
    b.close()
    for k, (L, x, v) in enumerate(systems):
        ctx = native.Context()
        ctx.upload(L, x, v)
        ctx.set_lj(1.0, 1.0, 2.5, False)
        ctx.config_list(2.5, 0.3)
        ctx.build_list()
        ctx.forces(native.VERLET, native.MIXED)
        ctx.set_integrator(kind=native.BAOAB, dt=0.005, friction=1.0, temperature=float(T[k]),
                           seed=int(seeds[k]))
        ctx.run(30, native.MIXED, thermo=False)
        x1, v1, _ = ctx.download()
        ctx.close()
        assert np.max(np.abs(x1 - pos[k])) < 1e-5
        assert np.max(np.abs(v1 - vel[k])) < 1e-5


def test_batch_thermostat_equipartition(gpu_ok):
    """Every replica's mean kinetic temperature (dof = 3N) within 3% of its
    own target after relaxation (the reference's equipartition gate is 2% on
    a 10k-step window; this window is shorter)."""
    from moleculardynamics_amd import native

    systems, T, seeds, cells, _ = _systems(16, master=4400, max_cells=6)
    b = native.Batch(systems, T, seeds, precision=native.MIXED)
    b.prime()
    b.run(2000, segment=100)
    th = b.run(4000, segment=100, thermo=True)
    n = np.array([s[1].shape[0] for s in systems])
    Tm = np.mean(2.0 * th[:, :, 2] / (3 * n[None, :]), axis=0)
    assert np.all(np.abs(Tm / T - 1.0) < 0.03), (Tm / T)
    assert np.all(b.rebuild_counts() > 1)
    b.close()


def test_autotune_keeps_trajectory(gpu_ok):
    from moleculardynamics_amd import native

    systems, T, seeds, _, _ = _systems(24, master=4500)
    ref = native.Batch(systems, T, seeds)
    ref.prime()
    ref.run(20, graph=False)
    want = ref.download()
    ref.close()
    b = native.Batch(systems, T, seeds)
    b.prime()
    tpa, pieces, block, ms = b.autotune(trial_steps=3)
    assert block in (256, 512, 1024) and set(np.unique(tpa)) <= {1, 2, 4, 8, 16} and ms > 0
    n = np.array([s[1].shape[0] for s in systems])
    assert np.all(pieces >= 1) and np.all(pieces <= np.maximum(1, np.minimum(32, n // 32)))
    b.run(20, graph=False)
    got = b.download()
    b.close()
    for a, c in zip(want[0] + want[1], got[0] + got[1]):
        assert np.array_equal(a, c)
