"""Bricks of list cells (the LDS-staged pair kernels) against the slot-entry
list layout, on the same inputs.

With bricks on, a list entry is the partner's index in the halo its row's
brick workgroup stages in LDS; the rows, their order and the pair arithmetic
are the slot layout's. So the sorted pair set, the forces of both
arithmetic paths (bit for bit), the mixed path's energy and virial
(fixed-point sums, bit for bit), the FP64 energy and virial (up to the
grouping of the workgroup sums) and velocity-Verlet trajectories must agree
between the two layouts. Bricks are off unless requested
(md_neigh_set_bricks). Covered: the automatic brick width and widths 2 and 3, both
truncation schemes, cubic and orthorhombic boxes, a dense blob whose brick
halos overflow the first LDS rows (the build grows them and builds again),
and Nose-Hoover-chain runs through md_run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _liquidish(cells, rho, seed, aspect=(1.0, 1.0, 1.0), shake=0.15):
    """fcc sites scaled to an (optionally orthorhombic) box, displaced at
    random by up to `shake` of the lattice constant, with random velocities
    (dynamics tests take small displacements: close pairs would saturate the
    mixed path's fixed-point accumulators)."""
    rng = np.random.default_rng(seed)
    n = 4 * cells ** 3
    vol = n / rho
    a = np.asarray(aspect, dtype=float)
    L = a * (vol / np.prod(a)) ** (1.0 / 3.0)
    b = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    g = np.arange(cells)
    cell = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 1, 3)
    x = ((cell + b[None]).reshape(-1, 3) / cells) * L
    x += rng.uniform(-shake, shake, size=x.shape) * (L / cells)
    x -= L * np.floor(x / L + 0.5)
    v = rng.normal(size=x.shape)
    v -= v.mean(axis=0)
    return L, x, v


def _ctx(L, x, v, bricks, lfs=False):
    from moleculardynamics_amd import native

    c = native.Context()
    c.upload(L, x, v)
    c.set_lj(1.0, 1.0, 2.5, lfs)
    c.config_list(2.5, 0.3)
    c.set_bricks(bricks)
    c.build_list()
    return c


def _compare_snapshot(a, b):
    """a: slot entries, b: bricks. Pair sets, forces, E/W and counts."""
    from moleculardynamics_amd import native

    assert a.bricks() == 0
    assert b.bricks() > 0
    assert np.array_equal(a.pairs(), b.pairs())
    ea, wa = a.forces(native.VERLET, native.FP64)
    eb, wb = b.forces(native.VERLET, native.FP64)
    assert a.pair_count() == b.pair_count()
    assert np.array_equal(a.download()[2], b.download()[2])
    assert abs(ea - eb) <= 1e-13 * abs(ea)
    assert abs(wa - wb) <= 1e-13 * abs(wa)
    ea, wa = a.forces(native.VERLET, native.MIXED)
    eb, wb = b.forces(native.VERLET, native.MIXED)
    assert a.pair_count() == b.pair_count()
    assert ea == eb and wa == wb
    assert np.array_equal(a.download()[2], b.download()[2])


@pytest.mark.parametrize("width", [-1, 2, 3], ids=["auto", "w2", "w3"])
@pytest.mark.parametrize("lfs", [False, True], ids=["trunc", "lfs"])
def test_bricks_match_slot_entries(gpu_ok, width, lfs):
    L, x, v = _liquidish(24, 0.8442, 11)  # 55,296 atoms
    a = _ctx(L, x, v, 0, lfs)
    b = _ctx(L, x, v, width, lfs)
    try:
        _compare_snapshot(a, b)
    finally:
        a.close()
        b.close()


def test_bricks_orthorhombic(gpu_ok):
    """Sides 7 : 11 : 13: a different fixed-point unit per axis, bricks of
    unequal widths on every axis."""
    L, x, v = _liquidish(22, 0.8, 12, aspect=(7, 11, 13))  # 42,592 atoms
    a = _ctx(L, x, v, 0)
    b = _ctx(L, x, v, -1)
    try:
        _compare_snapshot(a, b)
    finally:
        a.close()
        b.close()


def test_bricks_velocity_verlet_trajectory(gpu_ok):
    """40 velocity-Verlet steps through md_run, list rebuilds included: the
    forces are the same bits in both layouts, so positions and velocities
    stay bit-identical (the mixed and the FP64 path)."""
    from moleculardynamics_amd import native

    L, x, v = _liquidish(24, 0.8442, 13, shake=0.03)
    for prec in (native.MIXED, native.FP64):
        a = _ctx(L, x, 0.9 * v, 0)
        b = _ctx(L, x, 0.9 * v, -1)
        try:
            ths = []
            for c in (a, b):
                c.forces(native.VERLET, prec)
                c.set_integrator(kind=native.VV, dt=0.005)
                ths.append(c.run(40, prec))
            assert a.rebuild_count() == b.rebuild_count() > 1
            xa, va, _ = a.download()
            xb, vb, _ = b.download()
            assert np.array_equal(xa, xb)
            assert np.array_equal(va, vb)
            if prec == native.MIXED:  # fixed-point sums: the same bits
                assert np.array_equal(ths[0]["pe"], ths[1]["pe"])
            else:  # FP64 sums grouped by workgroup
                assert np.max(np.abs(ths[0]["pe"] - ths[1]["pe"]) / np.abs(ths[0]["pe"])) <= 1e-13
            assert np.max(np.abs(ths[0]["ke"] - ths[1]["ke"]) / ths[0]["ke"]) <= 1e-13
            _compare_snapshot(a, b)
        finally:
            a.close()
            b.close()


def test_bricks_nose_hoover_run(gpu_ok):
    """60 Nose-Hoover-chain steps (mixed): the chain sees the kinetic energy,
    an FP64 sum whose workgroup grouping differs between the layouts, so the
    runs agree to rounding, not bit for bit."""
    from moleculardynamics_amd import native

    L, x, v = _liquidish(24, 0.8442, 14, shake=0.03)
    a = _ctx(L, x, v, 0)
    b = _ctx(L, x, v, -1)
    try:
        ths = []
        for c in (a, b):
            c.forces(native.VERLET, native.MIXED)
            c.set_integrator(kind=native.NHC, dt=0.005, temperature=0.722, tau=0.5,
                             chain_length=3, dof=3 * x.shape[0] - 3)
            ths.append(c.run(60, native.MIXED))
        for key in ("pe", "ke", "conserved"):
            d = np.abs(ths[0][key] - ths[1][key])
            assert np.max(d / np.maximum(np.abs(ths[0][key]), 1.0)) <= 1e-10, key
        xa = a.download()[0]
        xb = b.download()[0]
        assert np.max(np.abs(xa - xb)) <= 1e-10
    finally:
        a.close()
        b.close()


def test_bricks_dense_blob_grows_halo_rows(gpu_ok):
    """A dense blob (55,296 atoms at rho* 0.84) in a box three times its
    volume: the halos of the blob's bricks hold far more atoms than the mean
    density predicts, so the first build overflows the LDS rows, and the
    build grows them (or turns the bricks off) and builds again. The result
    equals the slot layout's."""
    from moleculardynamics_amd import native

    Lb, x, v = _liquidish(24, 0.84, 15)
    L = Lb * 3.0 ** (1.0 / 3.0)
    a = _ctx(L, x, v, 0)
    b = native.Context()
    try:
        b.upload(L, x, v)
        b.set_lj(1.0, 1.0, 2.5, False)
        b.config_list(2.5, 0.3)
        b.set_bricks(-1)
        b.build_list()
        assert np.array_equal(a.pairs(), b.pairs())
        for prec in (native.FP64, native.MIXED):
            ea, wa = a.forces(native.VERLET, prec)
            eb, wb = b.forces(native.VERLET, prec)
            assert np.array_equal(a.download()[2], b.download()[2])
            assert abs(ea - eb) <= 1e-13 * abs(ea)
            assert abs(wa - wb) <= 1e-13 * abs(wa)
    finally:
        a.close()
        b.close()
