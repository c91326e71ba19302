"""Smooth particle-mesh Ewald on the GPU (md_water_set_kspace with
MD_KSPACE_PME) against the oracle's restatement (oracle/pme.cpp), which the
CPU suite pins against the Ewald reciprocal sum over the same k-vectors and
against the gradient of its own energy (tests/test_oracle_pme.py).

* NIST SPC/E configurations 1-4 (rc 10 A, kmax 5) at two grids and orders:
  the reciprocal energy against the oracle's PME to 1e-10, the other terms
  unchanged, and the forces' PME part (PME minus Ewald forces on the GPU)
  against the oracle's (PME minus Ewald reciprocal forces) to 1e-9 of the
  largest force;
* the 12,288-site configuration (rc 9 A, tuned alpha, kmax 15, the automatic
  grid of 64 points and order 8): PME within 1e-6 of the Ewald reciprocal
  energy and forces, and the oracle's PME at the same grid to 1e-10;
* constrained dynamics (SETTLE, 2 fs) with PME against the same run with
  Ewald, and the argument checks.
"""
import numpy as np
import pytest

from conftest import nist_spce
from oracle.pyoracle import Frame, load_nist

pytestmark = pytest.mark.gpu

DT = 0.0018236739  # 2 fs in units of 1.0966875353 ps


def _pos(fr):
    return np.stack([fr.x, fr.y, fr.z], axis=1)


def _q_mol(n):
    return np.array([-0.8476 if i % 3 == 0 else 0.4238 for i in range(n)]), np.arange(n) // 3


def _frame(L, x):
    return Frame(np.asarray(L, dtype=float), x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())


def _ewald_recip_forces(orc, fr, q, mol, rc, kmax, alpha):
    _, f = orc.ewald(fr, q, mol, rc=rc, kmax=kmax, alpha=alpha, forces=True)
    _, f0 = orc.ewald(fr, q, mol, rc=rc, kmax=0, alpha=alpha, forces=True)
    return f - f0


@pytest.mark.parametrize("grid,order", [(24, 6), (32, 8)], ids=["g24p6", "g32p8"])
@pytest.mark.parametrize("cfg", [1, 2, 3, 4])
def test_pme_nist_vs_oracle(gpu_ok, orc, cfg, grid, order):
    from moleculardynamics_amd import native

    fr = load_nist(nist_spce(cfg))
    q, mol = _q_mol(fr.n)
    w = native.Water(fr.L, _pos(fr), params=native.water_params(cutoff=10.0, kmax=5))
    try:
        te, _ = w.forces()
        fe = w.download()[2]
        w.set_kspace(native.KSPACE_PME, grid, order)
        tp, _ = w.forces()
        fp = w.download()[2]
    finally:
        w.close()
    e_pme, f_pme = orc.pme(fr, q, kmax=5, grid=(grid,) * 3, order=order, forces=True)
    assert tp["recip"] == pytest.approx(e_pme, rel=1e-10)
    for k in ("disp", "lrc", "real", "self", "intra"):
        assert tp[k] == te[k]
    f_er = _ewald_recip_forces(orc, fr, q, mol, 10.0, 5, 0.0)
    scale = np.max(np.abs(fe))
    assert np.max(np.abs((fp - fe) - (f_pme - f_er))) <= 1e-9 * scale


@pytest.fixture(scope="module")
def c3():
    from moleculardynamics_amd import native

    L = 49.71853
    pos, vel = native.water_lattice(16, L, 2026, 300.0)
    q2 = 4096 * (0.8476 ** 2 + 2 * 0.4238 ** 2)
    alpha, kmax = native.ewald_parameters(12288, [L] * 3, q2, 9.0, 1e-5)
    return L, pos, vel, alpha, kmax


def test_pme_config3_accuracy_and_oracle(gpu_ok, orc, c3):
    from moleculardynamics_amd import native

    L, pos, vel, alpha, kmax = c3
    w = native.Water([L] * 3, pos, vel, params=native.water_params(cutoff=9.0, kmax=kmax, alpha=alpha))
    try:
        te, pe_e = w.forces()
        fe = w.download()[2]
        w.set_kspace(native.KSPACE_PME)  # automatic grid (64 points for kmax 15), order 8
        tp, pe_p = w.forces()
        fp = w.download()[2]
    finally:
        w.close()
    # PME against the Ewald sum over the same k-vectors
    assert abs(tp["recip"] - te["recip"]) <= 1e-6 * abs(te["recip"])
    assert np.sqrt(np.sum((fp - fe) ** 2) / np.sum(fe ** 2)) <= 1e-6
    # and against the oracle's PME at the same grid and order
    fr = _frame([L] * 3, pos)
    q, _ = _q_mol(fr.n)
    e_pme, f_pme = orc.pme(fr, q, kmax=kmax, alpha=alpha, grid=(64, 64, 64), order=8, forces=True)
    assert tp["recip"] == pytest.approx(e_pme, rel=1e-10)
    q_, mol = _q_mol(fr.n)
    f_er = _ewald_recip_forces(orc, fr, q_, mol, 9.0, kmax, alpha)
    assert np.max(np.abs((fp - fe) - (f_pme - f_er))) <= 1e-9 * np.max(np.abs(fe))


def test_pme_config3_dynamics_follows_ewald(gpu_ok, c3):
    """20 SETTLE steps of 2 fs with PME and with Ewald from the same state:
    the potential energies agree to the PME accuracy (1e-6) and the
    trajectories to 1e-5 A (the forces differ by about 1e-6 relative, and 20
    steps of a liquid grow that to about 1e-6 A)."""
    from moleculardynamics_amd import native

    L, pos, vel, alpha, kmax = c3
    out = {}
    for method in (native.KSPACE_EWALD, native.KSPACE_PME):
        w = native.Water([L] * 3, pos, vel,
                         params=native.water_params(cutoff=9.0, kmax=kmax, alpha=alpha))
        try:
            w.set_kspace(method)
            w.set_constraint(native.RIGID_SETTLE)
            w.forces()
            pe, ke = w.run(20, DT)
            x = w.download()[0]
        finally:
            w.close()
        out[method] = (pe, ke, x)
    pe_e, ke_e, x_e = out[native.KSPACE_EWALD]
    pe_p, ke_p, x_p = out[native.KSPACE_PME]
    assert np.max(np.abs(pe_p - pe_e) / np.abs(pe_e)) < 1e-6
    d = x_p - x_e
    d -= L * np.floor(d / L + 0.5)
    assert np.max(np.abs(d)) < 1e-5
    # the constrained run conserves energy as well with PME as with Ewald
    de = np.abs((pe_p + ke_p) - (pe_p[0] + ke_p[0]))
    de_e = np.abs((pe_e + ke_e) - (pe_e[0] + ke_e[0]))
    assert np.max(de) <= 2.0 * np.max(de_e) + 1e-6 * abs(pe_e[0])


def test_pme_arguments(gpu_ok):
    from moleculardynamics_amd import native

    fr = load_nist(nist_spce(1))
    w = native.Water(fr.L, _pos(fr), params=native.water_params(cutoff=10.0, kmax=5))
    try:
        for grid, order in [(10, 6), (24, 2), (24, 11), (5, 3)]:  # grid < 2 kmax + 1, bad orders
            with pytest.raises(native.MDError):
                w.set_kspace(native.KSPACE_PME, grid, order)
        with pytest.raises(native.MDError):
            w.set_kspace(7)
        # a refused call leaves the Ewald path in place
        t0, _ = w.forces()
        w.set_kspace(native.KSPACE_PME, 24, 6)
        w.set_kspace(native.KSPACE_EWALD)
        t1, _ = w.forces()
        assert t1["recip"] == t0["recip"]
    finally:
        w.close()
