"""Pins of the oracle's per-type-pair LJ (no reference symbol; built from
forcefield/lennard_jones.cpp:50-88 with per-type-pair sigma, epsilon and
cutoff), its steepest descent and its velocity-Verlet run.

What fixes them independently of the restatement:
  * inclusion-exclusion over single-type energies: E = E(A; AA) + E(B; BB)
    + [E(A u B; AB) - E(A; AB) - E(B; AB)], each a single-type brute-force
    evaluation (the NIST-pinned oracle LJ, and the reference compiled from
    its own sources when oracle/_ref is present);
  * a mixture whose type pairs all share one parameter set is the
    single-type LJ, bit for bit;
  * closed forms at a fixed cutoff (dE/deps = E_ab/eps, dE/dsigma = W_ab/sigma,
    dW/deps = W_ab/eps, dW/dsigma = Y_ab/sigma) against central differences;
  * forces are minus the energy gradient (central differences);
  * steepest descent lowers the energy; velocity Verlet with the linear
    force shift conserves it.
"""
import numpy as np
import pytest

from conftest import KA_EPS, KA_RC, KA_SIGMA, ka_frame, ka_types


def _subset(fr, mask):
    from oracle.pyoracle import Frame

    return Frame(fr.L.copy(), fr.x[mask].copy(), fr.y[mask].copy(), fr.z[mask].copy())


def _incl_excl(o, fr, types):
    a, b = types == 0, types == 1
    fa, fb = _subset(fr, a), _subset(fr, b)
    s, e, r = np.array(KA_SIGMA), np.array(KA_EPS), np.array(KA_RC)
    ev = lambda f, i, j: np.array(o.lj_energy_virial(f, s[i, j], e[i, j], r[i, j]))  # noqa: E731
    return ev(fa, 0, 0) + ev(fb, 1, 1) + ev(fr, 0, 1) - ev(fa, 0, 1) - ev(fb, 0, 1)


@pytest.fixture(scope="module")
def mixture():
    fr = ka_frame(500, 1.2, 11, jitter=0.15)
    return fr, ka_types(fr.n)


def test_inclusion_exclusion_oracle(orc, mixture):
    fr, t = mixture
    m = orc.mix_lj(fr, t, KA_SIGMA, KA_EPS, KA_RC)
    want = _incl_excl(orc, fr, t)
    assert m["energy"] == pytest.approx(want[0], rel=1e-12)
    assert m["virial"] == pytest.approx(want[1], rel=1e-12)


def test_inclusion_exclusion_reference(orc, ref, mixture):
    """The same identity with the single-type energies from the reference
    compiled from its own sources."""
    fr, t = mixture
    m = orc.mix_lj(fr, t, KA_SIGMA, KA_EPS, KA_RC)
    want = _incl_excl(ref, fr, t)
    assert m["energy"] == pytest.approx(want[0], rel=1e-12)
    assert m["virial"] == pytest.approx(want[1], rel=1e-12)


@pytest.mark.parametrize("lfs", [False, True])
def test_uniform_parameters_reduce_to_single_type(orc, mixture, lfs):
    fr, t = mixture
    sg, ep, rc = [[1.1, 1.1], [1.1, 1.1]], [[0.7, 0.7], [0.7, 0.7]], [[2.4, 2.4], [2.4, 2.4]]
    m = orc.mix_lj(fr, t, sg, ep, rc, lfs=lfs, forces=True)
    e, w = orc.lj_energy_virial(fr, 1.1, 0.7, 2.4, lfs)
    f, _, _ = orc.lj_forces(fr, 1.1, 0.7, 2.4, lfs, mode="brute")
    assert m["energy"] == e and m["virial"] == w
    assert np.array_equal(m["forces"], f)


def _with(tab, a, b, v):
    t = np.array(tab, dtype=float)
    t[a, b] = t[b, a] = v
    return t


@pytest.mark.parametrize("pair", [(0, 0), (0, 1), (1, 1)])
def test_closed_form_parameter_derivatives(orc, mixture, pair):
    fr, t = mixture
    a, b = pair
    base = orc.mix_lj(fr, t, KA_SIGMA, KA_EPS, KA_RC)
    lo, hi = min(a, b), max(a, b)
    e_ab, w_ab, y_ab = base["e_ab"][lo, hi], base["w_ab"][lo, hi], base["y_ab"][lo, hi]
    if lo != hi:  # ordered tables: pairs whose lower index has type hi land in [hi, lo]
        e_ab += base["e_ab"][hi, lo]
        w_ab += base["w_ab"][hi, lo]
        y_ab += base["y_ab"][hi, lo]
    eps, sig = KA_EPS[a][b], KA_SIGMA[a][b]
    h = 1e-6
    for which, x0, closed_e, closed_w in [("eps", eps, e_ab / eps, w_ab / eps),
                                         ("sig", sig, w_ab / sig, y_ab / sig)]:
        vals = []
        for s in (+1, -1):
            v = x0 * (1 + s * h)
            tabs = dict(sig=KA_SIGMA, eps=KA_EPS)
            tabs[which] = _with(tabs[which], a, b, v)
            r = orc.mix_lj(fr, t, tabs["sig"], tabs["eps"], KA_RC)
            vals.append((r["energy"], r["virial"]))
        de = (vals[0][0] - vals[1][0]) / (2 * h * x0)
        dw = (vals[0][1] - vals[1][1]) / (2 * h * x0)
        assert de == pytest.approx(closed_e, rel=2e-6, abs=1e-6 * abs(base["energy"]) / x0)
        assert dw == pytest.approx(closed_w, rel=2e-6, abs=1e-6 * abs(base["virial"]) / x0)


@pytest.mark.parametrize("lfs", [False, True])
def test_forces_are_minus_gradient(orc, mixture, lfs):
    fr, t = mixture
    m = orc.mix_lj(fr, t, KA_SIGMA, KA_EPS, KA_RC, lfs=lfs, forces=True)
    h = 1e-6
    for i in (0, 7, 123, 450):
        for comp in ("x", "y", "z"):
            es = []
            for s in (+1, -1):
                g = fr.copy()
                getattr(g, comp)[i] += s * h
                es.append(orc.mix_lj(g, t, KA_SIGMA, KA_EPS, KA_RC, lfs=lfs)["energy"])
            fd = -(es[0] - es[1]) / (2 * h)
            f = m["forces"][i, "xyz".index(comp)]
            assert fd == pytest.approx(f, rel=1e-6, abs=1e-6)


def test_steepest_descent_relaxes(orc):
    fr = ka_frame(300, 1.2, 5, jitter=0.3)
    t = ka_types(fr.n)
    before = orc.mix_lj(fr, t, KA_SIGMA, KA_EPS, KA_RC, forces=True)
    out, e = orc.mix_minimize(fr, t, KA_SIGMA, KA_EPS, KA_RC, 60, 1e-3, 0.05)
    after = orc.mix_lj(out, t, KA_SIGMA, KA_EPS, KA_RC, forces=True)
    assert e[0] == before["energy"] and e[-1] == after["energy"]
    assert e[-1] < e[0]
    fmax = lambda r: np.max(np.linalg.norm(r["forces"], axis=1))  # noqa: E731
    assert fmax(after) < fmax(before)
    # the step cap: no coordinate moved by more than max_disp per step
    d = np.stack([out.x - fr.x, out.y - fr.y, out.z - fr.z], axis=1)
    d -= fr.L * np.floor(d / fr.L + 0.5)
    assert np.max(np.linalg.norm(d, axis=1)) <= 60 * 0.05 + 1e-12


def test_vv_conserves_energy_with_force_shift(orc):
    fr = ka_frame(256, 1.2, 9, jitter=0.05)
    t = ka_types(fr.n)
    relaxed, _ = orc.mix_minimize(fr, t, KA_SIGMA, KA_EPS, KA_RC, 100, 1e-3, 0.02, lfs=True)
    relaxed.vx, relaxed.vy, relaxed.vz = fr.vx.copy(), fr.vy.copy(), fr.vz.copy()
    out, e, w = orc.mix_vv(relaxed, t, KA_SIGMA, KA_EPS, KA_RC, 0.002, 200, lfs=True)
    assert e.shape == (201,) and w.shape == (201,)
    ke0 = 0.5 * np.sum(relaxed.vx ** 2 + relaxed.vy ** 2 + relaxed.vz ** 2)
    ke1 = 0.5 * np.sum(out.vx ** 2 + out.vy ** 2 + out.vz ** 2)
    h0, h1 = e[0] + ke0, e[-1] + ke1
    assert abs(h1 - h0) / abs(h0) < 1e-4
