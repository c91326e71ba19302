"""The C ABI library loads, exports every entry point include/moldyn_hip.h
declares, and its host-only entry points (no device needed) agree with the
oracle: the seeded fcc lattice bit for bit, the tail corrections, the cell
usability rule."""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def lib():
    from moleculardynamics_amd import build, native

    if not os.path.exists(native.LIB_PATH):
        build.build(verbose=False)
    return native.load()


def test_every_declared_symbol_is_exported(lib):
    from moleculardynamics_amd import native

    names = native.declared_symbols()
    assert len(names) >= 30
    missing = [nm for nm in names if not hasattr(lib, nm)]
    assert not missing, missing


def test_context_creation_without_gpu_fails_cleanly(lib):
    from moleculardynamics_amd import native

    hip = ctypes.CDLL("libamdhip64.so")
    cnt = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(native.MDError):
        native.Context()


@pytest.mark.parametrize("cells,rho,T,seed", [(3, 0.80, 1.0, 12345), (6, 0.8442, 0.722, 12345),
                                             (4, 0.5, 1.5, 4000), (10, 0.9, 0.7, 4123)])
def test_fcc_lattice_bit_identical_to_oracle(orc, cells, rho, T, seed):
    from moleculardynamics_amd import native

    L, x, v = native.fcc_lattice(cells, rho, T, seed)
    fr = orc.fcc_lattice(cells, rho, T, seed)
    assert np.array_equal(L, fr.L)
    assert np.array_equal(x, np.stack([fr.x, fr.y, fr.z], axis=1))
    assert np.array_equal(v, np.stack([fr.vx, fr.vy, fr.vz], axis=1))


def test_fcc_lattice_golden_kinetic_energy():
    """KE of fccLattice(3, 0.80, 1.0, 12345) is the reference's exp02 step-0
    value 152.6408823 (10 significant digits)."""
    from moleculardynamics_amd import native

    _, _, v = native.fcc_lattice(3, 0.80, 1.0, 12345)
    assert 0.5 * np.sum(v * v) == pytest.approx(152.6408823, rel=1e-9)
    assert np.all(np.abs(v.sum(axis=0)) < 1e-12)


def test_tail_corrections_match_oracle_and_hand_values(orc):
    from moleculardynamics_amd import native

    hand = json.load(open(os.path.join(GOLDEN, "nist_values.json")))["lrc_hand"]
    for rho, rc, u_n, p in hand:
        side = (500 / rho) ** (1 / 3)
        u, pr = native.lj_tail(500, [side] * 3, cutoff=rc)
        assert (u, pr) == orc.lj_lrc(500, [side] * 3, rc=rc)
        assert u / 500 == pytest.approx(u_n, rel=2e-3)
        assert pr == pytest.approx(p, rel=2e-3)
    assert native.lj_tail(500, [8.4] * 3, lfs=True) == (0.0, 0.0)


def test_cells_usable_rule(orc):
    from moleculardynamics_amd import native

    for L, rc in [([10, 10, 10], 3.0), ([5, 5, 5], 3.0), ([8.9, 9.1, 12], 3.0), ([9, 9, 9], 3.0)]:
        assert native.cells_usable(L, rc) == orc.cells_usable(L, rc)


def test_auto_capacity_arithmetic(lib):
    """ceil(2 (4/3) pi rl^3 rho + 32), at most n - 1, rounded up to 8: 192 at
    the 1M-atom fluid's rho* = 0.8442 with rl = 2.8 (mean row 77.6)."""
    from moleculardynamics_amd import native

    n = 1048576
    L = np.array([107.494156] * 3)
    mean = 4.0 / 3.0 * np.pi * 2.8 ** 3 * n / np.prod(L)
    assert abs(mean - 77.6) < 0.05
    want = int(np.ceil(2.0 * mean + 32.0) + 7) // 8 * 8
    assert native.auto_capacity(n, L, 2.5, 0.3) == want == 192
    # small systems are capped at n - 1 (then rounded up to the octet size)
    assert native.auto_capacity(10, np.array([2.0] * 3), 2.5, 0.3) == 16
    assert native.auto_capacity(0, L, 2.5, 0.3) == 0
    assert native.auto_capacity(n, L, -1.0, 0.3) == 0
