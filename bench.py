"""Benchmark: 1,048,576-atom Lennard-Jones fluid, Nose-Hoover chain, one MI355X.

Workload (the large single-GPU configuration): fccLattice(64, rho*=0.8442,
T*=0.722, seed 20260001) -> N = 1,048,576 atoms, LJ eps = sigma = 1, cutoff
2.5 sigma (Truncated), Verlet list with skin 0.3 and the displacement-triggered
rebuild, Nose-Hoover chain M = 3 at T* = 0.722 (tau 0.5, dof 3N-3),
dt* = 0.005. A step is one full integrator step of every atom (rebuild check,
rebuild when triggered, forces, both half-kicks, both half-chains).

Both arithmetic paths run from the identical initial state: the plain FP64
path and the mixed path (FP32 pair arithmetic, 64-bit fixed-point
accumulation, FP64 state). `value` is the mixed path's timesteps/s; the FP64
numbers, pair-interactions/s, the mixed-vs-FP64 force error and the roofline
of the dominant kernel (the LJ pair kernel, timed with HIP events on the
stream it runs on, against the device's peaks measured in the same run), a
2000-step NVE drift of both arithmetic paths, the mixed-Ewald error on the
SPC/E configuration, and the CPU baseline: the oracle restatement of the
reference CPU path (-O3, one thread) on 20 NHC steps of the same workload,
with the host CPU model and core count.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CELLS, RHO, TEMP, SEED = 64, 0.8442, 0.722, 20260001
RC, SKIN, DT, TAU, CHAIN = 2.5, 0.3, 0.005, 0.5, 3
F_PAIR = 31.0              # flops per pair-interaction (fixed convention)
# AMD-published MI355X figures, reported beside the peaks measured in the run
PUB_FP64 = 78.6e12         # vector FP64, FLOP/s
PUB_FP32 = 157.3e12        # vector FP32 (packed), FLOP/s
PUB_HBM = 8.0e12           # B/s


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(key: str):
    """HBM bytes per launch of the pair kernel from the committed rocprofv3
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "k_lj_lj1m_nhc_traffic.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def run_gpu(L, pos, vel, precision: int, steps: int, warmup: int):
    from moleculardynamics_amd import native

    n = pos.shape[0]
    ctx = native.Context()
    ctx.upload(L, pos, vel)
    ctx.set_lj(1.0, 1.0, RC, False)
    ctx.config_list(RC, SKIN)
    ctx.build_list()
    ctx.forces(native.VERLET, precision)
    ctx.set_integrator(kind=native.NHC, dt=DT, temperature=TEMP, tau=TAU, chain_length=CHAIN,
                       dof=3 * n - 3)
    if warmup:
        ctx.run(warmup, precision, thermo=False)
    builds0 = ctx.rebuild_count()
    ctx.synchronize()
    # HIP events bracket the pair kernel only: each bracket costs a few
    # microseconds of device time per step, so the step's other kernels run
    # unbracketed in the timed region
    ctx.timing_select(native.Context.TIME_PAIR)
    ctx.timing_reset()
    t0 = time.perf_counter()
    th = ctx.run(steps, precision, thermo=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_force, nforce = ctx.timing_read(0)
    ctx.timing(False)
    builds = ctx.rebuild_count() - builds0
    x, v, f = ctx.download()
    ctx.close()
    wall = t1 - t0
    return dict(wall_s=wall, steps=steps, ms_per_step=1e3 * wall / steps,
                pairs_per_step=float(np.mean(th["pairs"])),
                force_ms=ms_force / max(nforce, 1), force_launches=nforce,
                rebuilds=builds, thermo=th, final=(x, v, f))


def force_error(L, pos, vel):
    """RMS / max force error of the mixed path against FP64 on one snapshot."""
    from moleculardynamics_amd import native

    ctx = native.Context()
    ctx.upload(L, pos, vel)
    ctx.set_lj(1.0, 1.0, RC, False)
    ctx.config_list(RC, SKIN)
    ctx.build_list()
    e64, w64 = ctx.forces(native.VERLET, native.FP64)
    _, _, f64 = ctx.download()
    e32, w32 = ctx.forces(native.VERLET, native.MIXED)
    _, _, f32 = ctx.download()
    ctx.close()
    d = f32 - f64
    rms = float(np.sqrt(np.sum(d * d) / np.sum(f64 * f64)))
    mx = float(np.max(np.linalg.norm(d, axis=1)) / np.sqrt(np.mean(np.sum(f64 * f64, axis=1))))
    return dict(rms_rel=rms, max_rel=mx, energy_rel=abs(e32 - e64) / abs(e64),
                virial_rel=abs(w32 - w64) / abs(w64))


def host_description():
    """CPU model and the cores this process may use (and the machine's count)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count()}


def nve_drift(L, pos, vel, steps: int):
    """max_t |E(t) - E(0)| / |E(0)| of velocity-Verlet NVE on the LinearForceShift
    potential (continuous energy, so cutoff crossings do not mask precision),
    FP64 and mixed legs from the same state, plus the largest relative gap
    between the two legs' total energies."""
    from moleculardynamics_amd import native

    out = {}
    h = {}
    for name, prec in [("fp64", native.FP64), ("mixed", native.MIXED)]:
        ctx = native.Context()
        ctx.upload(L, pos, vel)
        ctx.set_lj(1.0, 1.0, RC, True)
        ctx.config_list(RC, SKIN)
        ctx.build_list()
        ctx.forces(native.VERLET, prec)
        ctx.set_integrator(kind=native.VV, dt=DT)
        th = ctx.run(steps, prec, thermo=True)
        ctx.close()
        h[name] = th["pe"] + th["ke"]
        out[name] = float(np.max(np.abs(h[name] - h[name][0])) / abs(h[name][0]))
    out["legs_max_rel_gap"] = float(np.max(np.abs(h["mixed"] - h["fp64"])) / abs(h["fp64"][0]))
    out.update(steps=steps, potential="LinearForceShift rc=2.5", integrator="VV dt*=0.005")
    return out


def water_mixed_error():
    """Mixed (FP32 real-space erfc) against FP64 SPC/E forces on the 4096-molecule
    configuration (12,288 sites, rc 9 A, alpha/kmax tuned for 1e-5)."""
    from moleculardynamics_amd import native

    L = 49.71853
    pos, vel = native.water_lattice(16, L, 2026, 300.0)
    q2 = 4096 * (0.8476 ** 2 + 2 * 0.4238 ** 2)
    alpha, kmax = native.ewald_parameters(12288, [L] * 3, q2, 9.0, 1e-5)
    w = native.Water([L] * 3, pos, vel, params=native.water_params(cutoff=9.0, kmax=kmax,
                                                                    alpha=alpha))
    t64, pe64 = w.forces(native.FP64)
    _, _, f64 = w.download()
    t32, pe32 = w.forces(native.MIXED)
    _, _, f32 = w.download()
    w.close()
    d = f32 - f64
    return {"sites": 12288, "alpha": alpha, "kmax": kmax,
            "rms_rel": float(np.sqrt(np.sum(d * d) / np.sum(f64 * f64))),
            "max_rel": float(np.max(np.linalg.norm(d, axis=1))
                             / np.sqrt(np.mean(np.sum(f64 * f64, axis=1)))),
            "energy_rel": abs(pe32 - pe64) / abs(pe64),
            "real_space_energy_rel": abs(t32["real"] - t64["real"]) / abs(t64["real"]),
            "pressure": "not computed: the water path has no virial (as the reference's)"}


def cpu_baseline(timed_steps: int = 20, warm_steps: int = 2):
    from oracle import pyoracle

    o = pyoracle.Oracle(pyoracle.ORACLE_SO)
    fr = o.fcc_lattice(CELLS, RHO, TEMP, SEED)
    n = fr.n
    t0 = time.perf_counter()
    sim = o.sim(fr, kind="nhc", dt=DT, rc=RC, lfs=False, skin=SKIN, T=TEMP, tau=TAU, M=CHAIN,
                dof=3 * n - 3)
    t_setup = time.perf_counter() - t0
    sim.run(warm_steps)
    t0 = time.perf_counter()
    sim.run(timed_steps)
    dt = time.perf_counter() - t0
    host = host_description()
    return dict(value=timed_steps / dt, unit="timesteps/s", cores=1, kind="port",
                sample=(f"N={n} NHC steps through the Verlet list, {warm_steps} warm-up + "
                        f"{timed_steps} timed steps after the initial build "
                        f"(setup {t_setup:.1f} s), one thread of {host['cpu_model']} "
                        f"({host['nproc']} usable cores); oracle restatement built -O3 -DNDEBUG"),
                s_per_step=dt / timed_steps, **host)


def _line(metric, value, unit, steps, warmup, ms, dtype, workload, extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded)",
           "config": {"workload": workload, "parallelism": "single GPU"}}
    out.update(extra)
    print(json.dumps(out))
    return 0


def bench_lj864(args) -> int:
    """864-atom LJ fluid, FP64, velocity Verlet NVE, 1000 steps (the reference's own
    CPU-runnable case): fccLattice(6, 0.8442, 0.722, 12345), rc 2.5 (Truncated), skin 0.3."""
    from moleculardynamics_amd import native

    L, pos, vel = native.fcc_lattice(6, RHO, TEMP, 12345)
    ctx = native.Context()
    ctx.upload(L, pos, vel)
    ctx.set_lj(1.0, 1.0, RC, False)
    ctx.config_list(RC, SKIN)
    ctx.build_list()
    ctx.forces(native.VERLET, native.FP64)
    ctx.set_integrator(kind=native.VV, dt=DT)
    ctx.run(args.warmup, native.FP64, thermo=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    th = ctx.run(1000, native.FP64, thermo=True)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.close()
    h = th["pe"] + th["ke"]
    extra = {"pairs_per_step": float(np.mean(th["pairs"])),
             "nve_drift": float(np.max(np.abs(h - h[0])) / abs(h[0]))}
    if not args.no_cpu:
        from oracle import pyoracle

        o = pyoracle.Oracle(pyoracle.ORACLE_SO)
        fr = o.fcc_lattice(6, RHO, TEMP, 12345)
        sim = o.sim(fr, kind="vv", dt=DT, rc=RC, skin=SKIN)
        c0 = time.perf_counter()
        sim.run(1000)
        cw = time.perf_counter() - c0
        extra["cpu_baseline"] = {"value": 1000 / cw, "unit": "timesteps/s", "cores": 1,
                                 "kind": "port", "sample": "all 1000 NVE steps, single thread",
                                 **host_description()}
    return _line("timesteps/s, 864-atom LJ NVE, FP64, 1 MI355X", 1000 / wall, "timesteps/s", 1000,
                 args.warmup, 1e3 * wall / 1000, "f64",
                 "LJ fluid N=864 (fcc 6^3x4), rho*=0.8442, T*=0.722, rc=2.5, skin=0.3, VV dt*=0.005",
                 extra)


def bench_batch(args) -> int:
    """256 uneven LJ replicas (108..4000 atoms, rho* in [0.5, 0.9], T* in [0.7, 1.5]),
    BAOAB gamma = 1, dt* = 0.005, 10,000 steps in HIP-graph segments of 100, mixed
    precision; against the plain single-system path launched once per replica."""
    from moleculardynamics_amd import native

    cells, rho, T, seeds = native.batch_draw(256, 4000)
    systems = [native.fcc_lattice(int(c), float(r), float(t), int(s))
               for c, r, t, s in zip(cells, rho, T, seeds)]
    atoms = int(sum(s[1].shape[0] for s in systems))
    b = native.Batch(systems, T, seeds, precision=native.MIXED)
    b.prime()
    # tune from a state whose list ages are spread, over trials that span rebuilds
    graph = not args.no_graph
    b.run(200, segment=100, graph=graph)
    tpa, pieces, block, _ = b.autotune(trial_steps=30)
    b.run(max(args.warmup, 100), segment=100, graph=graph)
    steps = 10000
    t0 = time.perf_counter()
    b.run(steps, segment=100, graph=graph)
    wall = time.perf_counter() - t0
    b.close()
    batch_rate = 256 * steps / wall  # replica-steps per second
    # the plain path, one replica at a time, 100 steps each
    plain_steps = 100
    plain_wall = plain_setup = 0.0
    for k, (L, x, v) in enumerate(systems):
        t0 = time.perf_counter()
        ctx = native.Context()
        ctx.upload(L, x, v)
        ctx.set_lj(1.0, 1.0, RC, False)
        ctx.config_list(RC, SKIN)
        ctx.build_list()
        ctx.forces(native.VERLET, native.MIXED)
        ctx.set_integrator(kind=native.BAOAB, dt=DT, friction=1.0, temperature=float(T[k]),
                           seed=int(seeds[k]))
        ctx.run(10, native.MIXED, thermo=False)  # graph capture and first launches, untimed
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.run(plain_steps, native.MIXED, thermo=False)
        ctx.synchronize()
        t2 = time.perf_counter()
        ctx.close()
        plain_setup += t1 - t0
        plain_wall += t2 - t1
    plain_rate = 256 * plain_steps / plain_wall
    extra = {"replicas": 256, "atoms_total": atoms, "block": int(block),
             "tpa_classes": sorted(set(int(t) for t in tpa)),
             "workgroups_per_step": int(sum(int(p) for p in pieces)),
             "plain_path_replica_steps_per_s": plain_rate,
             "batch_vs_plain": batch_rate / plain_rate,
             "plain_path_setup_s": plain_setup,
             "plain_path_sample": f"256 replicas x {plain_steps} timed steps, one md_ctx each, "
                                  "run one after another; context setup, upload, list build and "
                                  "10 warm-up steps per replica timed separately"}
    return _line("replica-steps/s, 256 uneven LJ replicas, BAOAB, mixed, 1 MI355X", batch_rate,
                 "replica-steps/s", steps, 100, 1e3 * wall / steps,
                 "f32 pair / i64 fixed-point accumulate / f64 state",
                 "256 LJ replicas of 108-4000 atoms, rho* U[0.5,0.9], T* U[0.7,1.5], rc=2.5, "
                 "BAOAB gamma=1 dt*=0.005, 10000 steps in "
                 + ("100-step graph segments" if graph else "plain launches"), extra)


def pme_scaling(sides=(16, 20, 24), calls=20):
    """Force evaluations (SPC/E, rc 9 A, alpha and kmax for 1e-5 at each size) with the
    k-vector sums and with PME, at m^3 molecules on the water density's lattice: the
    sums grow as N kmax^3 (kmax grows with L), PME as N plus its grid."""
    from moleculardynamics_amd import native

    out = []
    for m in sides:
        L = 49.71853 * m / 16.0
        n = 3 * m ** 3
        pos, vel = native.water_lattice(m, L, 2026, 300.0)
        q2 = m ** 3 * (0.8476 ** 2 + 2 * 0.4238 ** 2)
        alpha, kmax = native.ewald_parameters(n, [L] * 3, q2, 9.0, 1e-5)
        row = {"sites": n, "L": L, "kmax": kmax}
        w = native.Water([L] * 3, pos, vel, params=native.water_params(cutoff=9.0, kmax=kmax, alpha=alpha))
        try:
            for name, method in (("ewald", native.KSPACE_EWALD), ("pme", native.KSPACE_PME)):
                w.set_kspace(method)
                t, _ = w.forces()
                t0 = time.perf_counter()
                for _ in range(calls):
                    w.forces()
                row[name + "_ms_per_force_call"] = 1e3 * (time.perf_counter() - t0) / calls
                row[name + "_recip"] = t["recip"]
        finally:
            w.close()
        row["pme_speedup"] = row["ewald_ms_per_force_call"] / row["pme_ms_per_force_call"]
        row["recip_rel_diff"] = abs(row["pme_recip"] - row["ewald_recip"]) / abs(row["ewald_recip"])
        log(f"  PME scaling: {n} sites, kmax {kmax}: Ewald {row['ewald_ms_per_force_call']:.3f} ms, "
            f"PME {row['pme_ms_per_force_call']:.3f} ms per force call ({row['pme_speedup']:.2f}x)")
        out.append(row)
    return out


def bench_water(args) -> int:
    """4096 SPC/E molecules (12,288 sites), L = 49.71853 A, rc 9 A, alpha and kmax for a
    1e-5 relative force error, 500 constrained velocity-Verlet steps of 2 fs at 300 K."""
    from moleculardynamics_amd import native

    L = 49.71853
    pos, vel = native.water_lattice(16, L, 2026, 300.0)
    q2 = 4096 * (0.8476 ** 2 + 2 * 0.4238 ** 2)
    alpha, kmax = native.ewald_parameters(12288, [L] * 3, q2, 9.0, 1e-5)
    prm = native.water_params(cutoff=9.0, kmax=kmax, alpha=alpha)
    w = native.Water([L] * 3, pos, vel, params=prm)
    dt = 0.0018236739  # 2 fs in units of 1.0966875353 ps
    w.forces()
    if args.warmup:
        w.run(args.warmup, dt)
    t0 = time.perf_counter()
    pe, ke = w.run(500, dt)
    wall = time.perf_counter() - t0
    w.close()
    drift = lambda p, k: float(np.max(np.abs((p + k) - (p[0] + k[0]))) / abs(p[0] + k[0]))  # noqa: E731
    extra = {"alpha": alpha, "kmax": kmax, "sites": 12288, "constraint": "SHAKE/RATTLE (RigidWater)",
             "energy_drift": drift(pe, ke)}
    # the same run with SETTLE, the closed-form constraint (same fixed points)
    ws = native.Water([L] * 3, pos, vel, params=prm)
    ws.set_constraint(native.RIGID_SETTLE)
    ws.forces()
    if args.warmup:
        ws.run(args.warmup, dt)
    t0 = time.perf_counter()
    pes, kes = ws.run(500, dt)
    wall_s = time.perf_counter() - t0
    ws.close()
    extra["settle"] = {"timesteps_per_s": 500 / wall_s, "ms_per_step": 1e3 * wall_s / 500,
                       "energy_drift": drift(pes, kes), "speedup_vs_shake": wall / wall_s}
    # SETTLE with the reciprocal term by smooth PME over the same k-vectors
    # (automatic grid, order 8), and its error against the Ewald sum on the
    # starting configuration
    wp = native.Water([L] * 3, pos, vel, params=prm)
    te, _ = wp.forces()
    fe = wp.download()[2]
    wp.set_kspace(native.KSPACE_PME)
    tp, _ = wp.forces()
    fp = wp.download()[2]
    wp.set_constraint(native.RIGID_SETTLE)
    if args.warmup:
        wp.run(args.warmup, dt)
    t0 = time.perf_counter()
    pep, kep = wp.run(500, dt)
    wall_p = time.perf_counter() - t0
    wp.close()
    extra["settle_pme"] = {
        "timesteps_per_s": 500 / wall_p, "ms_per_step": 1e3 * wall_p / 500,
        "energy_drift": drift(pep, kep), "speedup_vs_settle_ewald": wall_s / wall_p,
        "grid": 2 * ((21 * kmax + 9) // 10), "order": 8,
        "recip_energy_rel_error": abs(tp["recip"] - te["recip"]) / abs(te["recip"]),
        "force_rms_rel_error": float(np.sqrt(np.sum((fp - fe) ** 2) / np.sum(fe ** 2)))}
    log(f"  SETTLE + PME: {500 / wall_p:.1f} steps/s ({wall_s / wall_p:.2f}x SETTLE + Ewald), "
        f"recip error {extra['settle_pme']['recip_energy_rel_error']:.2e}, "
        f"force RMS error {extra['settle_pme']['force_rms_rel_error']:.2e}")
    extra["pme_scaling"] = pme_scaling()
    if not args.no_cpu:
        from oracle import pyoracle
        from oracle.pyoracle import Frame

        # the reference's Ewald fixes alpha = 5.6/L, so the tuned alpha runs on the port
        o = pyoracle.Oracle(pyoracle.ORACLE_SO)
        fr = Frame(np.array([L] * 3), pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(),
                   vel[:, 0].copy(), vel[:, 1].copy(), vel[:, 2].copy())
        c0 = time.perf_counter()
        o.water_run(fr, rc=9.0, kmax=kmax, alpha=alpha, dt=dt, nsteps=2)
        cw = time.perf_counter() - c0
        extra["cpu_baseline"] = {"value": 2 / cw, "unit": "timesteps/s", "cores": 1, "kind": "port",
                                 "sample": "2 constrained VV steps (O(N^2) real space), single thread"}
    return _line("timesteps/s, SPC/E Ewald 12,288 sites, 1 MI355X", 500 / wall, "timesteps/s", 500,
                 args.warmup, 1e3 * wall / 500, "f64",
                 "SPC/E 4096 molecules, L=49.71853 A, rc=9 A, Ewald 1e-5, CVV dt=2 fs, 300 K", extra)


def bench_mixture(args) -> int:
    """The differentiated binary mixture: N = 32,000 (80:20), rho* = 1.2, 500
    steepest-descent steps, then 200 velocity-Verlet steps with checkpoints every 20 and
    the adjoint of <E> + 0.1 <W> + KE(T) (linear force shift)."""
    from moleculardynamics_amd import native

    n = 32000
    box = (n / 1.2) ** (1.0 / 3.0)
    t = np.zeros(n, dtype=np.int32)
    t[int(0.8 * n):] = 1
    sig, eps, rc = [[1.0, 0.8], [0.8, 0.88]], [[1.0, 1.5], [1.5, 0.5]], [[2.5, 2.0], [2.0, 2.2]]
    x, v = native.mix_random_state(n, box, 2027, 1.0)
    m = native.Mix([box] * 3, x, t, sig, eps, rc, lfs=True, vel=v)
    t0 = time.perf_counter()
    m.minimize(500, 1e-3, 0.05)
    t_sd = time.perf_counter() - t0
    t0 = time.perf_counter()
    e, w = m.run(0.005, 200, 20)
    t_fwd = time.perf_counter() - t0
    _, vT, _ = m.download()
    t0 = time.perf_counter()
    m.adjoint(1.0 / 201, 0.1 / 201, None, vT)
    t_adj = time.perf_counter() - t0
    m.close()
    extra = {"minimize_s": t_sd, "forward_s": t_fwd, "adjoint_s": t_adj,
             "adjoint_over_forward": t_adj / t_fwd}
    if not args.no_cpu:
        from oracle import pyoracle
        from oracle.pyoracle import Frame

        o = pyoracle.Oracle(pyoracle.ORACLE_SO)
        fr = Frame(np.array([box] * 3), x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
        c0 = time.perf_counter()
        o.mix_lj(fr, t, sig, eps, rc, lfs=True)
        cw = time.perf_counter() - c0
        extra["cpu_baseline"] = {"value": 1.0 / cw, "unit": "energy evaluations/s", "cores": 1,
                                 "kind": "port",
                                 "sample": "one brute-force mixture energy at N=32000, single thread"}
    return _line("timesteps/s, 32k binary-mixture VV forward + adjoint, FP64, 1 MI355X",
                 200 / (t_fwd + t_adj), "timesteps/s", 200, 0, 1e3 * (t_fwd + t_adj) / 200, "f64",
                 "binary LJ mixture N=32000 80:20 rho*=1.2, 500 SD + 200 VV (K=20) + adjoint",
                 extra)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch workload: plain launches instead of HIP-graph segments "
                         "(for kernel traces)")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--nve-steps", type=int, default=2000,
                    help="VV steps of the NVE drift check per leg (0 skips it)")
    ap.add_argument("--no-water", action="store_true",
                    help="skip the mixed-Ewald error check (for kernel traces of the LJ legs)")
    ap.add_argument("--workload", default="lj1m",
                    choices=["lj1m", "lj864", "water", "batch", "mixture"],
                    help="lj1m (default): the 1M-atom LJ line; the others are the remaining "
                         "configurations, each printing its own JSON line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run three device-to-device copies of MIB MiB (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path. A single large "
              "system would shard by spatial domain decomposition (one brick of the box per "
              "GPU, an rc + skin halo exchanged each step, atoms migrating at list rebuilds).",
              file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        import torch

        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}))
        return 0
    other = {"lj864": bench_lj864, "water": bench_water, "batch": bench_batch,
             "mixture": bench_mixture}
    if args.workload in other:
        return other[args.workload](args)

    from moleculardynamics_amd import native

    log("measuring device peaks ...")
    peak_hbm, peak_fp64, peak_fp32 = native.device_peaks()
    log(f"  HBM {peak_hbm / 1e12:.3f} TB/s, FP64 {peak_fp64 / 1e12:.2f} TF, "
        f"FP32 {peak_fp32 / 1e12:.2f} TF")
    L, pos, vel = native.fcc_lattice(CELLS, RHO, TEMP, SEED)
    n = pos.shape[0]
    log(f"workload: N={n} L={L[0]:.6f}")
    res = {}
    for name, prec in [("fp64", native.FP64), ("mixed", native.MIXED)]:
        log(f"gpu leg {name}: warmup {args.warmup}, timed {args.steps}")
        res[name] = run_gpu(L, pos, vel, prec, args.steps, args.warmup)
        r = res[name]
        log(f"  {name}: {r['ms_per_step']:.4f} ms/step, force kernel {r['force_ms']:.4f} ms, "
            f"pairs/step {r['pairs_per_step']:.0f}, rebuilds {r['rebuilds']}")
    # physics checks of the mixed run against FP64 (same initial state)
    xf, vf, _ = res["fp64"]["final"]
    err = force_error(L, xf, vf)
    log(f"  mixed vs fp64 force error: {err}")
    th64, th32 = res["fp64"]["thermo"], res["mixed"]["thermo"]
    h64 = th64["pe"] + th64["ke"] + th64["conserved"]
    h32 = th32["pe"] + th32["ke"] + th32["conserved"]
    drift64 = float(np.max(np.abs(h64 - h64[0])) / abs(h64[0]))
    drift32 = float(np.max(np.abs(h32 - h32[0])) / abs(h32[0]))

    drift = werr = None
    if args.nve_steps > 0:
        log(f"NVE drift, {args.nve_steps} VV steps per leg ...")
        drift = nve_drift(L, pos, vel, args.nve_steps)
        log(f"  {drift}")
    if not args.no_water:
        log("mixed-Ewald error (SPC/E, 12,288 sites) ...")
        werr = water_mixed_error()
        log(f"  {werr}")

    m = res["mixed"]
    d = res["fp64"]
    tps = 1e3 / m["ms_per_step"]
    tps64 = 1e3 / d["ms_per_step"]
    pairs = m["pairs_per_step"]
    achieved = F_PAIR * pairs / (m["force_ms"] * 1e-3)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_steps)
        log(f"  cpu: {cpu['value']:.4f} steps/s ({cpu['kind']})")
    out = {
        "metric": "pair-interactions/s + timesteps/s, 1M-atom LJ, mixed vs FP64, 1 MI355X",
        "value": tps,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair / i64 fixed-point accumulate / f64 state",
        "data": "synthetic (seeded fcc lattice + Maxwell-Boltzmann velocities)",
        "config": {"workload": "LJ fluid N=1048576 (fcc 64^3x4), rho*=0.8442, T*=0.722, "
                               "rc=2.5, skin=0.3, NHC M=3 tau=0.5, dt*=0.005",
                   "atoms": n, "parallelism": "single GPU"},
        "pair_interactions_per_s": pairs * tps,
        "pairs_per_step": pairs,
        "fp64": {"timesteps_per_s": tps64, "ms_per_step": d["ms_per_step"],
                 "pair_interactions_per_s": d["pairs_per_step"] * tps64,
                 "force_kernel_ms": d["force_ms"],
                 "roofline_frac_valu_fp64": F_PAIR * d["pairs_per_step"] / (d["force_ms"] * 1e-3) / peak_fp64,
                 "nhc_drift": drift64},
        "mixed_vs_fp64_speedup": tps / tps64,
        "mixed_error": {**err, "nhc_drift": drift32, "nhc_drift_fp64": drift64},
        "nve_drift": drift,
        "water_mixed_error": werr,
        "rebuilds": {"fp64": d["rebuilds"], "mixed": m["rebuilds"]},
        "peaks_measured": {"hbm_TB_s": peak_hbm / 1e12, "fp64_TFLOP_s": peak_fp64 / 1e12,
                           "fp32_TFLOP_s": peak_fp32 / 1e12},
        "peaks_published": {"hbm_TB_s": PUB_HBM / 1e12, "fp64_TFLOP_s": PUB_FP64 / 1e12,
                            "fp32_TFLOP_s": PUB_FP32 / 1e12},
        "roofline": {"bound": "valu", "kernel": "k_lj_lanes (LJ pair kernel, mixed, 4 lanes per atom)",
                     "achieved": achieved / 1e12, "peak": peak_fp32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak_fp32,
                     "frac_of_published": achieved / PUB_FP32,
                     "traffic": profiled_traffic("traffic_bytes_per_launch_mixed"),
                     "traffic_source": "profiles/k_lj_lj1m_nhc_traffic.md (rocprofv3 --pmc "
                                       "FETCH_SIZE, WRITE_SIZE, copy-calibrated)",
                     "launch_ms": m["force_ms"],
                     "step_frac": (max(168.0 * n / peak_hbm, F_PAIR * pairs / peak_fp32)
                                   / (m["ms_per_step"] * 1e-3))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
