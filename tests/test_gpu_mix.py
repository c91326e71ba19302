"""The differentiable per-type-pair LJ path on the GPU (md_mix_*) against the
oracle (forces, energy, virial, steepest descent, velocity Verlet), closed
forms and central finite differences (parameter derivatives, force VJP,
trajectory adjoint), and the full binary-mixture workload: 32,000 atoms,
80:20 A:B at rho* = 1.2, 500 steepest-descent steps, 200 velocity-Verlet
steps with checkpoints every 20, gradients of <E>, <W> and a final-frame
loss with respect to (eps_ab, sigma_ab) against central differences
(h = 1e-6).

Trajectory gradients use the linear force shift: with truncated forces a pair
crossing its cutoff between the two finite-difference runs makes a jump in
the force, so differences of trajectories are not defined there. Single-frame
derivatives are checked for both truncation schemes."""
import numpy as np
import pytest

from conftest import KA_EPS, KA_RC, KA_SIGMA, ka_frame, ka_types

pytestmark = pytest.mark.gpu

CANON = [(0, 0), (0, 1), (1, 1)]


def _pos(fr):
    return np.stack([fr.x, fr.y, fr.z], axis=1)


def _vel(fr):
    return np.stack([fr.vx, fr.vy, fr.vz], axis=1)


def _mix(fr, t, lfs, sig=KA_SIGMA, eps=KA_EPS, vel=True):
    from moleculardynamics_amd import native

    return native.Mix(fr.L, _pos(fr), t, sig, eps, KA_RC, lfs=lfs, vel=_vel(fr) if vel else None)


def _with(tab, a, b, v):
    t = np.array(tab, dtype=float)
    t[a, b] = t[b, a] = v
    return t


@pytest.fixture(scope="module")
def small():
    fr = ka_frame(500, 1.2, 21, jitter=0.15)
    return fr, ka_types(fr.n)


@pytest.mark.parametrize("lfs", [False, True])
def test_forces_energy_virial_match_oracle(gpu_ok, orc, small, lfs):
    fr, t = small
    m = _mix(fr, t, lfs)
    e, w = m.forces()
    _, _, f = m.download()
    o = orc.mix_lj(fr, t, KA_SIGMA, KA_EPS, KA_RC, lfs=lfs, forces=True)
    assert e == pytest.approx(o["energy"], rel=1e-12)
    assert w == pytest.approx(o["virial"], rel=1e-12)
    assert np.max(np.abs(f - o["forces"])) <= 1e-12 * np.max(np.abs(o["forces"]))
    m.close()


def test_parameter_tables_truncated_closed_forms(gpu_ok, orc, small):
    fr, t = small
    m = _mix(fr, t, False)
    e, w, tab = m.forces(derivatives=True)
    o = orc.mix_lj(fr, t, KA_SIGMA, KA_EPS, KA_RC)
    assert tab[:, 0].sum() == pytest.approx(e, rel=1e-12)
    assert tab[:, 1].sum() == pytest.approx(w, rel=1e-12)
    for k, (a, b) in enumerate(CANON):
        e_ab, w_ab, y_ab = o["e_ab"][a, b], o["w_ab"][a, b], o["y_ab"][a, b]
        if a != b:
            e_ab += o["e_ab"][b, a]
            w_ab += o["w_ab"][b, a]
            y_ab += o["y_ab"][b, a]
        eps, sig = KA_EPS[a][b], KA_SIGMA[a][b]
        want = [e_ab, w_ab, e_ab / eps, w_ab / sig, w_ab / eps, y_ab / sig]
        scale = [abs(e), abs(w), abs(e), abs(w), abs(w), abs(y_ab) + abs(w)]
        for got, exp, sc in zip(tab[k], want, scale):
            assert abs(got - exp) <= 1e-12 * max(sc, abs(exp))
    m.close()


@pytest.mark.parametrize("lfs", [False, True])
def test_parameter_tables_vs_oracle_finite_differences(gpu_ok, orc, small, lfs):
    """dE/dp and dW/dp for p = eps_ab, sigma_ab against central differences
    of the oracle energy at fixed cutoffs (h = 1e-6 relative)."""
    fr, t = small
    m = _mix(fr, t, lfs)
    e, w, tab = m.forces(derivatives=True)
    m.close()
    h = 1e-6
    for k, (a, b) in enumerate(CANON):
        for which, col_e, col_w in (("eps", 2, 4), ("sig", 3, 5)):
            x0 = (KA_EPS if which == "eps" else KA_SIGMA)[a][b]
            ev = []
            for s in (+1, -1):
                tabs = dict(sig=KA_SIGMA, eps=KA_EPS)
                tabs[which] = _with(tabs[which], a, b, x0 * (1 + s * h))
                r = orc.mix_lj(fr, t, tabs["sig"], tabs["eps"], KA_RC, lfs=lfs)
                ev.append((r["energy"], r["virial"]))
            de = (ev[0][0] - ev[1][0]) / (2 * h * x0)
            dw = (ev[0][1] - ev[1][1]) / (2 * h * x0)
            assert tab[k, col_e] == pytest.approx(de, rel=1e-6, abs=1e-7 * abs(e) / x0)
            assert tab[k, col_w] == pytest.approx(dw, rel=1e-6, abs=1e-7 * abs(w) / x0)


def test_force_vjp_is_hessian_vector_product(gpu_ok, orc, small):
    """(dF/dx)^T g against central differences of the GPU forces along g
    (the Jacobian is symmetric), and g . dF/dp against differences of the
    oracle forces in p; linear force shift, so forces are continuous."""
    fr, t = small
    rng = np.random.default_rng(3)
    g = rng.normal(size=(fr.n, 3))
    m = _mix(fr, t, True)
    o, th = m.force_vjp(g)
    h = 1e-6
    fs = []
    for s in (+1, -1):
        p = _pos(fr) + s * h * g
        m.set_state(p)
        m.forces()
        fs.append(m.download()[2])
    m.close()
    jg = (fs[0] - fs[1]) / (2 * h)
    assert np.max(np.abs(o - jg)) <= 1e-5 * np.max(np.abs(jg))
    for k, (a, b) in enumerate(CANON):
        for col, which in ((0, "eps"), (1, "sig")):
            x0 = (KA_EPS if which == "eps" else KA_SIGMA)[a][b]
            ff = []
            for s in (+1, -1):
                tabs = dict(sig=KA_SIGMA, eps=KA_EPS)
                tabs[which] = _with(tabs[which], a, b, x0 * (1 + s * h))
                ff.append(orc.mix_lj(fr, t, tabs["sig"], tabs["eps"], KA_RC, lfs=True,
                                     forces=True)["forces"])
            fd = np.sum(g * (ff[0] - ff[1])) / (2 * h * x0)
            assert th[k, col] == pytest.approx(fd, rel=1e-5, abs=1e-6 * np.sum(np.abs(g)))


def test_energy_and_virial_gradients(gpu_ok, small):
    """w_energy dE/dx + w_virial dW/dx against central differences."""
    fr, t = small
    m = _mix(fr, t, True)
    o, _ = m.force_vjp(None, 0.7, -0.3)
    rng = np.random.default_rng(8)
    u = rng.normal(size=(fr.n, 3))
    h = 1e-6
    vals = []
    for s in (+1, -1):
        m.set_state(_pos(fr) + s * h * u)
        e, w = m.forces()
        vals.append(0.7 * e - 0.3 * w)
    m.close()
    fd = (vals[0] - vals[1]) / (2 * h)
    assert np.sum(o * u) == pytest.approx(fd, rel=1e-6)


def test_minimize_matches_oracle(gpu_ok, orc, small):
    fr, t = small
    m = _mix(fr, t, False)
    e = m.minimize(40, 1e-3, 0.05)
    p, _, _ = m.download()
    m.close()
    out, e0 = orc.mix_minimize(fr, t, KA_SIGMA, KA_EPS, KA_RC, 40, 1e-3, 0.05)
    assert np.max(np.abs(e - e0) / np.abs(e0)) < 1e-10
    d = p - _pos(out)
    d -= fr.L * np.floor(d / fr.L + 0.5)
    assert np.max(np.abs(d)) < 1e-10


@pytest.mark.parametrize("lfs", [False, True])
def test_velocity_verlet_matches_oracle(gpu_ok, orc, small, lfs):
    fr, t = small
    relaxed, _ = orc.mix_minimize(fr, t, KA_SIGMA, KA_EPS, KA_RC, 50, 1e-3, 0.02, lfs=lfs)
    relaxed.vx, relaxed.vy, relaxed.vz = fr.vx.copy(), fr.vy.copy(), fr.vz.copy()
    m = _mix(relaxed, t, lfs)
    e, w = m.run(0.003, 40, checkpoint_every=7)
    p, v, _ = m.download()
    m.close()
    out, e0, w0 = orc.mix_vv(relaxed, t, KA_SIGMA, KA_EPS, KA_RC, 0.003, 40, lfs=lfs)
    assert np.max(np.abs(e - e0)) < 1e-9 * np.max(np.abs(e0))
    assert np.max(np.abs(w - w0)) < 1e-9 * np.max(np.abs(w0))
    d = p - _pos(out)
    d -= fr.L * np.floor(d / fr.L + 0.5)
    assert np.max(np.abs(d)) < 1e-9
    assert np.max(np.abs(v - _vel(out))) < 1e-9


def _loss(m, dt, nsteps, k):
    """<E> + 0.1 <W> + final kinetic energy."""
    e, w = m.run(dt, nsteps, k)
    _, v, _ = m.download()
    return e.mean() + 0.1 * w.mean() + 0.5 * np.sum(v * v), v


def _adjoint(m, nsteps, v):
    fr_w = 1.0 / (nsteps + 1)
    return m.adjoint(fr_w, 0.1 * fr_w, None, v)


def test_adjoint_matches_finite_differences(gpu_ok, orc, small):
    fr, t = small
    relaxed, _ = orc.mix_minimize(fr, t, KA_SIGMA, KA_EPS, KA_RC, 50, 1e-3, 0.02, lfs=True)
    relaxed.vx, relaxed.vy, relaxed.vz = fr.vx.copy(), fr.vy.copy(), fr.vz.copy()
    dt, nsteps, k = 0.004, 40, 10
    m = _mix(relaxed, t, True)
    x0, v0 = _pos(relaxed), _vel(relaxed)
    _, vT = _loss(m, dt, nsteps, k)
    th, dx0, dv0 = _adjoint(m, nsteps, vT)
    h = 1e-6
    for kk, (a, b) in enumerate(CANON):
        for col, which in ((0, "eps"), (1, "sig")):
            base = (KA_EPS if which == "eps" else KA_SIGMA)[a][b]
            ls = []
            for s in (+1, -1):
                tabs = dict(sig=KA_SIGMA, eps=KA_EPS)
                tabs[which] = _with(tabs[which], a, b, base * (1 + s * h))
                m.set_params(tabs["sig"], tabs["eps"], KA_RC, True)
                m.set_state(x0, v0)
                ls.append(_loss(m, dt, nsteps, k)[0])
            fd = (ls[0] - ls[1]) / (2 * h * base)
            assert th[kk, col] == pytest.approx(fd, rel=1e-5, abs=1e-7 * abs(ls[0]))
    m.set_params(KA_SIGMA, KA_EPS, KA_RC, True)
    rng = np.random.default_rng(17)
    for grad, which in ((dx0, "x"), (dv0, "v")):
        u = rng.normal(size=x0.shape)
        ls = []
        for s in (+1, -1):
            if which == "x":
                m.set_state(x0 + s * h * u, v0)
            else:
                m.set_state(x0, v0 + s * h * u)
            ls.append(_loss(m, dt, nsteps, k)[0])
        fd = (ls[0] - ls[1]) / (2 * h)
        assert np.sum(grad * u) == pytest.approx(fd, rel=1e-5, abs=1e-7 * abs(ls[0]))
    m.close()


def test_adjoint_independent_of_checkpoint_interval(gpu_ok, orc, small):
    fr, t = small
    relaxed, _ = orc.mix_minimize(fr, t, KA_SIGMA, KA_EPS, KA_RC, 50, 1e-3, 0.02, lfs=True)
    relaxed.vx, relaxed.vy, relaxed.vz = fr.vx.copy(), fr.vy.copy(), fr.vz.copy()
    res = []
    for k in (1, 7, 30):
        m = _mix(relaxed, t, True)
        _, vT = _loss(m, 0.004, 30, k)
        res.append(_adjoint(m, 30, vT))
        m.close()
    for th, dx, dv in res[1:]:
        assert np.allclose(th, res[0][0], rtol=1e-9, atol=0)
        assert np.max(np.abs(dx - res[0][1])) <= 1e-9 * np.max(np.abs(res[0][1]))
        assert np.max(np.abs(dv - res[0][2])) <= 1e-9 * np.max(np.abs(res[0][2]))


def test_autograd_function_wrappers(gpu_ok, small):
    import torch

    from moleculardynamics_amd.autograd import MixEnergy, MixForces, MixSystem, MixTrajectory

    fr, t = small
    sysm = MixSystem(fr.L, t, KA_RC, lfs=True)
    eps = torch.tensor([KA_EPS[a][b] for a, b in CANON], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor([KA_SIGMA[a][b] for a, b in CANON], dtype=torch.float64, requires_grad=True)
    pos = torch.tensor(_pos(fr), dtype=torch.float64, requires_grad=True)
    e, w = MixEnergy.apply(eps, sig, pos, sysm)
    (e + 0.5 * w).backward()
    h = 1e-6
    # position gradient against a central difference along a random direction
    u = torch.tensor(np.random.default_rng(4).normal(size=_pos(fr).shape), dtype=torch.float64)
    vals = []
    for sgn in (+1, -1):
        ee, ww = MixEnergy.apply(eps.detach(), sig.detach(), pos.detach() + sgn * h * u, sysm)
        vals.append(float(ee + 0.5 * ww))
    assert float((pos.grad * u).sum()) == pytest.approx((vals[0] - vals[1]) / (2 * h), rel=1e-6)
    # forces and their VJP: d(F . g)/dx = J^T g
    p2 = pos.detach().clone().requires_grad_(True)
    f = MixForces.apply(eps.detach(), sig.detach(), p2, sysm)
    # (a uniform g would be in the null space: the forces sum to zero)
    g = torch.tensor(np.random.default_rng(5).normal(size=_pos(fr).shape), dtype=torch.float64)
    (f * g).sum().backward()
    fs = []
    for sgn in (+1, -1):
        fs.append(MixForces.apply(eps.detach(), sig.detach(), pos.detach() + sgn * h * u, sysm))
    fd = float(((fs[0] - fs[1]) / (2 * h) * g).sum())
    assert float((p2.grad * u).sum()) == pytest.approx(fd, rel=1e-5)
    for i in range(3):
        for s_, tgt in ((eps, eps.grad), (sig, sig.grad)):
            d = torch.zeros(3, dtype=torch.float64)
            d[i] = h * float(s_[i])
            vals = []
            for sgn in (+1, -1):
                ee, ww = MixEnergy.apply(eps.detach() + (sgn * d if s_ is eps else 0),
                                         sig.detach() + (sgn * d if s_ is sig else 0),
                                         pos.detach(), sysm)
                vals.append(float(ee + 0.5 * ww))
            fd = (vals[0] - vals[1]) / (2 * h * float(s_[i]))
            assert float(tgt[i]) == pytest.approx(fd, rel=1e-6, abs=1e-8 * abs(vals[0]))
    # trajectory wrapper: gradient of <E> w.r.t. eps against the adjoint through the C ABI
    x0 = torch.tensor(_pos(fr), dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(_vel(fr), dtype=torch.float64, requires_grad=True)
    em, wm, xT, vT = MixTrajectory.apply(eps, sig, x0, v0, sysm, 0.002, 20, 5)
    loss = em + 0.5 * (vT * vT).sum()
    eps.grad = None
    loss.backward()
    assert eps.grad is not None and torch.all(torch.isfinite(eps.grad))
    assert x0.grad.shape == x0.shape and v0.grad.shape == v0.shape
    sysm.close()


def test_trajectory_adjoint_survives_interleaved_calls(gpu_ok, small):
    """A MixEnergy call at other parameters between a MixTrajectory's forward
    and backward reconfigures the shared handle; the backward then runs its
    forward again from its own inputs, so the gradients equal those of an
    uninterrupted forward/backward."""
    import torch

    from moleculardynamics_amd.autograd import MixEnergy, MixSystem, MixTrajectory

    fr, t = small
    sysm = MixSystem(fr.L, t, KA_RC, lfs=True)
    grads = []
    for interleave in (False, True):
        eps = torch.tensor([KA_EPS[a][b] for a, b in CANON], dtype=torch.float64,
                           requires_grad=True)
        sig = torch.tensor([KA_SIGMA[a][b] for a, b in CANON], dtype=torch.float64,
                           requires_grad=True)
        x0 = torch.tensor(_pos(fr), dtype=torch.float64, requires_grad=True)
        v0 = torch.tensor(_vel(fr), dtype=torch.float64, requires_grad=True)
        em, wm, xT, vT = MixTrajectory.apply(eps, sig, x0, v0, sysm, 0.002, 20, 5)
        if interleave:
            with torch.no_grad():
                MixEnergy.apply(1.3 * eps.detach(), 0.97 * sig.detach(), x0.detach() + 0.01, sysm)
        (em + 0.1 * wm + 0.5 * (vT * vT).sum()).backward()
        grads.append((eps.grad.clone(), sig.grad.clone(), x0.grad.clone(), v0.grad.clone()))
    for a, b in zip(*grads):
        assert torch.allclose(a, b, rtol=1e-12, atol=0), (a, b)
    sysm.close()


def test_binary_mixture_workload_gradients(gpu_ok):
    """The full differentiated workload: N = 32,000 (25,600 A + 6,400 B),
    L = (32000 / 1.2)^(1/3) = 29.876, seeded uniform positions relaxed by 500
    steepest-descent steps, Maxwell-Boltzmann velocities at T* = 1.0, 200
    velocity-Verlet steps (dt* = 0.005) with checkpoints every 20; gradients
    of L = <E> + 0.1 <W> + KE(T) in the six (eps_ab, sigma_ab) against
    central differences (h = 1e-6)."""
    from moleculardynamics_amd import native

    n, rho = 32000, 1.2
    box = (n / rho) ** (1.0 / 3.0)
    assert box == pytest.approx(29.876, abs=1e-3)
    t = ka_types(n)
    x, v = native.mix_random_state(n, box, 2027, 1.0)
    L = np.array([box] * 3)
    m = native.Mix(L, x, t, KA_SIGMA, KA_EPS, KA_RC, lfs=True, vel=v)
    e_sd = m.minimize(500, 1e-3, 0.05)
    assert e_sd[-1] < e_sd[0] and e_sd[-1] < 0
    xr, _, _ = m.download()
    dt, nsteps, k = 0.005, 200, 20
    m.set_state(xr, v)
    _, vT = _loss(m, dt, nsteps, k)
    th, _, _ = _adjoint(m, nsteps, vT)
    h = 1e-6
    for kk, (a, b) in enumerate(CANON):
        for col, which in ((0, "eps"), (1, "sig")):
            base = (KA_EPS if which == "eps" else KA_SIGMA)[a][b]
            ls = []
            for s in (+1, -1):
                tabs = dict(sig=KA_SIGMA, eps=KA_EPS)
                tabs[which] = _with(tabs[which], a, b, base * (1 + s * h))
                m.set_params(tabs["sig"], tabs["eps"], KA_RC, True)
                m.set_state(xr, v)
                ls.append(_loss(m, dt, nsteps, k)[0])
            fd = (ls[0] - ls[1]) / (2 * h * base)
            assert th[kk, col] == pytest.approx(fd, rel=1e-4, abs=1e-6 * abs(ls[0]))
    m.close()
