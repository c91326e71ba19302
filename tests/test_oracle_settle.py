"""The oracle's closed-form rigid-water constraints (SETTLE), pinned against
the reference's own SHAKE/RATTLE, which the reference says SETTLE must
reproduce to machine precision (constrain/rigid_water.hpp:20-24), and against
the conditions that define the solution:

* positions equal SHAKE converged to 1e-14 (the oracle's and, when built, the
  reference's own RigidWater, oracle/_ref) to 1e-12 A over steps of 0.001 to
  0.1 A per site, including molecules stored across the periodic boundary;
* the constrained molecule has the rigid bond lengths to 1e-14, the
  unconstrained centre of mass, displacements in the plane of the reference
  molecule and no angular momentum about it (bond forces exert no torque);
* velocities equal RATTLE converged to 1e-14; every bond's r.v is zero, the
  momentum is kept and the kinetic energy drops by exactly the removed
  part's (the projection is orthogonal in the mass metric: nothing injected);
* the constrained velocity Verlet with SETTLE follows the one with SHAKE /
  RATTLE (and the reference build's) to 1e-11 over 20 SPC/E steps, with the
  bonds held to 1e-12 throughout;
* steps without a solution and bad arguments raise.
Runs on the CPU."""
import numpy as np
import pytest

from conftest import nist_spce
from oracle.pyoracle import load_nist

HH = 2.0 * np.sin(0.5 * np.deg2rad(109.47))


def _masses(n):
    return np.array([15.9994 if i % 3 == 0 else 1.008 for i in range(n)])


def _pos(fr):
    return np.stack([fr.x, fr.y, fr.z], axis=1)


def _set_pos(fr, p):
    fr.x, fr.y, fr.z = p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def _mi(d, L):
    return d - L * np.floor(d / L + 0.5)


def _bond_error(p, L):
    worst = 0.0
    for a, b, t in ((0, 1, 1.0), (0, 2, 1.0), (1, 2, HH)):
        d = _mi(p[a::3] - p[b::3], L)
        worst = max(worst, np.max(np.abs(np.linalg.norm(d, axis=1) - t) / t))
    return worst


@pytest.fixture(scope="module")
def rigid(orc):
    """NIST SPC/E configuration 1 made rigid to 1e-14 (its file holds 6 digits)."""
    fr = load_nist(nist_spce(1))
    m = _masses(fr.n)
    _, p = orc.rigid_shake(fr, m, fr, tol=1e-14, maxit=2000)
    _set_pos(fr, p)
    return fr, m


def _moved(fr, amp, seed):
    rng = np.random.default_rng(seed)
    mv = fr.copy()
    _set_pos(mv, _pos(fr) + rng.normal(size=(fr.n, 3)) * amp)
    return mv


@pytest.mark.parametrize("amp", [0.001, 0.01, 0.05, 0.1])
def test_positions_equal_converged_shake(orc, rigid, amp):
    fr, m = rigid
    mv = _moved(fr, amp, 11)
    it, ps = orc.rigid_shake(mv, m, fr, tol=1e-14, maxit=5000)
    assert it > 1
    one, pt = orc.settle_positions(mv, m, fr)
    assert one == 1
    assert np.max(np.abs(pt - ps)) < 1e-12
    assert _bond_error(pt, fr.L) < 1e-14


def test_positions_equal_reference_build_shake(orc, ref, rigid):
    fr, m = rigid
    mv = _moved(fr, 0.03, 12)
    _, ps = ref.rigid_shake(mv, m, fr, tol=1e-14, maxit=5000)
    _, pt = orc.settle_positions(mv, m, fr)
    assert np.max(np.abs(pt - ps)) < 1e-12


def test_position_solution_conditions(orc, rigid):
    fr, m = rigid
    mv = _moved(fr, 0.05, 13)
    _, pt = orc.settle_positions(mv, m, fr)
    p0, pu = _pos(fr), _pos(mv)
    D = pt - pu
    mm = m.reshape(-1, 3, 1)
    # centre of mass kept
    assert np.max(np.abs((D.reshape(-1, 3, 3) * mm).sum(1))) < 1e-13
    r0 = p0.reshape(-1, 3, 3)
    rel0 = np.stack([_mi(r0[:, k] - r0[:, 0], fr.L) for k in range(3)], axis=1)
    nrm = np.cross(rel0[:, 1], rel0[:, 2])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    Dm = D.reshape(-1, 3, 3)
    # displacements in the reference plane
    assert np.max(np.abs(np.einsum("mkd,md->mk", Dm, nrm))) < 1e-13
    # no angular momentum about the centre: sum m r0 x D = 0
    L = (np.cross(rel0, Dm) * mm).sum(1)
    assert np.max(np.abs(L)) < 1e-13


def test_positions_across_the_boundary(orc, rigid):
    fr, m = rigid
    mv = _moved(fr, 0.02, 14)
    _, pt = orc.settle_positions(mv, m, fr)
    # store every second hydrogen one box length away: the result moves with it
    sh = np.zeros((fr.n, 3))
    sh[1::6, 0] = fr.L[0]
    sh[2::6, 1] = -fr.L[1]
    mv2, fr2 = mv.copy(), fr.copy()
    _set_pos(mv2, _pos(mv) + sh)
    _set_pos(fr2, _pos(fr) - sh)
    _, pt2 = orc.settle_positions(mv2, m, fr2)
    assert np.max(np.abs(pt2 - sh - pt)) < 1e-12


def _random_velocities(fr, m, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(fr.n, 3)) * np.sqrt(300.0 / m)[:, None]
    w = fr.copy()
    w.vx, w.vy, w.vz = v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()
    return w, v


def test_velocities_equal_converged_rattle(orc, rigid):
    fr, m = rigid
    w, v = _random_velocities(fr, m, 15)
    it, vr = orc.rigid_rattle(w, m, tol=1e-14, maxit=5000)
    assert it > 1
    one, vs = orc.settle_velocities(w, m)
    assert one == 1
    scale = np.max(np.abs(v))
    assert np.max(np.abs(vs - vr)) < 1e-13 * scale
    p = _pos(fr)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        d = _mi(p[a::3] - p[b::3], fr.L)
        rv = np.sum(d * (vs[a::3] - vs[b::3]), axis=1)
        assert np.max(np.abs(rv)) < 1e-13 * scale
    mm = m.reshape(-1, 3, 1)
    dv = vs - v
    assert np.max(np.abs((dv.reshape(-1, 3, 3) * mm).sum(1))) < 1e-12 * scale
    ke = lambda u: 0.5 * np.sum(m[:, None] * u * u)  # noqa: E731
    assert ke(vs) < ke(v)
    assert ke(v) == pytest.approx(ke(vs) + ke(dv), rel=1e-12)


def test_velocities_equal_reference_build_rattle(orc, ref, rigid):
    fr, m = rigid
    w, v = _random_velocities(fr, m, 16)
    _, vref = ref.rigid_rattle(w, m, tol=1e-14, maxit=5000)
    _, vs = orc.settle_velocities(w, m)
    assert np.max(np.abs(vs - vref)) < 1e-13 * np.max(np.abs(v))


def test_dynamics_follow_shake(orc, rigid):
    fr, m = rigid
    rng = np.random.default_rng(3)
    v = rng.normal(size=(fr.n, 3)) * np.sqrt(298.0 / m)[:, None]
    v -= (v * m[:, None]).sum(0) / m.sum()
    w = fr.copy()
    w.vx, w.vy, w.vz = v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()
    _, vr = orc.settle_velocities(w, m)
    w.vx, w.vy, w.vz = vr[:, 0].copy(), vr[:, 1].copy(), vr[:, 2].copy()
    dt = 0.0018236739  # 2 fs in units of 1.0966875353 ps
    wa, pa, ka = orc.water_run(w, dt=dt, nsteps=20, settle=True)
    wb, pb, kb = orc.water_run(w, dt=dt, nsteps=20, settle=False)
    assert np.max(np.abs(pa - pb) / np.abs(pb)) < 1e-11
    assert np.max(np.abs(ka - kb) / np.abs(kb)) < 1e-11
    assert np.max(np.abs(_mi(_pos(wa) - _pos(wb), fr.L))) < 1e-10
    assert _bond_error(_pos(wa), fr.L) < 1e-12
    # held over a longer run
    wc, _, _ = orc.water_run(w, dt=dt, nsteps=200, settle=True)
    assert _bond_error(_pos(wc), fr.L) < 1e-12


def test_dynamics_follow_reference_build(orc, ref, rigid):
    fr, m = rigid
    rng = np.random.default_rng(4)
    v = rng.normal(size=(fr.n, 3)) * np.sqrt(298.0 / m)[:, None]
    w = fr.copy()
    w.vx, w.vy, w.vz = v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()
    _, vr = orc.settle_velocities(w, m)
    w.vx, w.vy, w.vz = vr[:, 0].copy(), vr[:, 1].copy(), vr[:, 2].copy()
    wa, pa, ka = orc.water_run(w, nsteps=10, settle=True)
    wb, pb, kb = ref.water_run(w, nsteps=10)
    assert np.max(np.abs(pa - pb) / np.abs(pb)) < 1e-11
    assert np.max(np.abs(ka - kb) / np.abs(kb)) < 1e-11


def test_unsolvable_steps_and_bad_arguments(orc, rigid):
    fr, m = rigid
    p0 = _pos(fr)
    # molecule 0's first hydrogen moved 3 A along the molecule's normal: no
    # rotation keeps the out-of-plane coordinates
    nrm = np.cross(_mi(p0[1] - p0[0], fr.L), _mi(p0[2] - p0[0], fr.L))
    nrm /= np.linalg.norm(nrm)
    p = p0.copy()
    p[1] += 3.0 * nrm
    mv = fr.copy()
    _set_pos(mv, p)
    with pytest.raises(RuntimeError, match="rotated"):
        orc.settle_positions(mv, m, fr)
    # degenerate reference (all three sites of molecule 0 on one line)
    q = p0.copy()
    q[2] = q[0] + 2.0 * _mi(q[1] - q[0], fr.L)
    bad = fr.copy()
    _set_pos(bad, q)
    with pytest.raises(RuntimeError, match="degenerate"):
        orc.settle_positions(fr, m, bad)
    m2 = m.copy()
    m2[2] = 2.0
    with pytest.raises(RuntimeError, match="hydrogen masses"):
        orc.settle_positions(fr, m2, fr)
    with pytest.raises(ValueError):
        orc.settle_positions(fr, m, fr, oh=-1.0)
    with pytest.raises(ValueError):
        orc.settle_positions(fr, m, fr, angle=180.0)
