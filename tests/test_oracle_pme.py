"""The oracle's smooth particle-mesh Ewald (oracle/pme.cpp) pinned on the CPU.

PME approximates the reference's reciprocal sum over its own k-vector set
(forcefield/ewald.cpp:112-172), so the oracle's Ewald reciprocal term (itself
pinned by the NIST SPC/E values) is its limit as the grid and the spline
order grow. Checked here on NIST SPC/E configuration 1 (300 sites):

* the energy error falls as the grid and the order refine, down to rounding
  (below 1e-12 relative at 64 points and order 10);
* the forces converge to the Ewald reciprocal forces (the difference of the
  oracle's Ewald forces with kmax 5 and with no k-vectors);
* the forces are minus the gradient of the PME energy (central differences);
* the spline correction |b(m)|^2 of order 4 equals its closed form
  9 / (2 + cos(2 pi m / K))^2;
* shifting every site by whole grid spacings leaves the energy unchanged
  (the mesh's own translation symmetry).
"""
import math

import numpy as np
import pytest

from conftest import nist_spce
from oracle.pyoracle import load_nist


def _water(cfg):
    fr = load_nist(nist_spce(cfg))
    q = np.array([-0.8476 if i % 3 == 0 else 0.4238 for i in range(fr.n)])
    return fr, q, np.arange(fr.n) // 3


def _ewald_recip(orc, fr, q, mol, kmax=5):
    t, f = orc.ewald(fr, q, mol, rc=10.0, kmax=kmax, alpha=0.0, forces=True)
    _, f0 = orc.ewald(fr, q, mol, rc=10.0, kmax=0, alpha=0.0, forces=True)
    return t["recip"], f - f0


def test_pme_converges_to_ewald_recip(orc):
    fr, q, mol = _water(1)
    e_ref, f_ref = _ewald_recip(orc, fr, q, mol)
    errs = []
    for K, p in [(16, 4), (16, 6), (24, 6), (32, 8), (48, 8), (64, 10)]:
        e, f = orc.pme(fr, q, kmax=5, grid=(K, K, K), order=p, forces=True)
        errs.append(abs(e - e_ref) / abs(e_ref))
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[0] > 1e-4 and errs[-1] < 1e-12, errs
    fscale = np.max(np.abs(f_ref))
    assert np.max(np.abs(f - f_ref)) / fscale < 1e-10


def test_pme_forces_are_energy_gradient(orc):
    fr, q, _ = _water(2)
    K, p, h = (24, 24, 24), 6, 1e-5
    _, f = orc.pme(fr, q, kmax=5, grid=K, order=p, forces=True)
    fscale = np.max(np.abs(f))
    for i in (0, 1, 2, 97, 298):
        for d, name in enumerate(("x", "y", "z")):
            a = getattr(fr, name)
            a0 = a[i]
            a[i] = a0 + h
            ep, _ = orc.pme(fr, q, kmax=5, grid=K, order=p)
            a[i] = a0 - h
            em, _ = orc.pme(fr, q, kmax=5, grid=K, order=p)
            a[i] = a0
            assert abs(-(ep - em) / (2 * h) - f[i, d]) / fscale < 1e-7


@pytest.mark.parametrize("K", [16, 33, 64])
def test_spline_correction_order4_closed_form(orc, K):
    # M_4(1), M_4(2), M_4(3) = 1/6, 2/3, 1/6: |sum_j M_4(j+1) e^{i j t}|^2 = ((2 + cos t)/3)^2
    for m in (-7, -1, 0, 1, 5, K // 3):
        t = 2.0 * math.pi * m / K
        assert orc.pme_bsq(4, K, m) == pytest.approx(9.0 / (2.0 + math.cos(t)) ** 2, rel=1e-14)


def test_pme_grid_translation_symmetry(orc):
    fr, q, _ = _water(3)
    K, p = 30, 6
    e0, _ = orc.pme(fr, q, kmax=5, grid=(K, K, K), order=p)
    sh = fr.L / K * np.array([1.0, 2.0, -3.0])
    fr.x += sh[0]
    fr.y += sh[1]
    fr.z += sh[2]
    e1, _ = orc.pme(fr, q, kmax=5, grid=(K, K, K), order=p)
    assert abs(e1 - e0) <= 1e-12 * abs(e0)
