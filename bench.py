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
stream it runs on) are in the same JSON line, next to the CPU baseline: the
reference compiled from its own sources (oracle/_ref) when present, else the
oracle restatement, single-threaded, on a bounded sample of the same workload.

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
PEAK_FP64 = 78.6e12        # MI355X vector FP64, FLOP/s
PEAK_FP32 = 157.3e12       # MI355X vector FP32, FLOP/s
PEAK_HBM = 8.0e12          # B/s


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


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
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    th = ctx.run(steps, precision, thermo=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_force, nforce = ctx.timing_read(0)
    ms_build, nbuild = ctx.timing_read(1)
    ms_pre, npre = ctx.timing_read(2)
    ctx.timing(False)
    builds = ctx.rebuild_count() - builds0
    x, v, f = ctx.download()
    ctx.close()
    wall = t1 - t0
    return dict(wall_s=wall, steps=steps, ms_per_step=1e3 * wall / steps,
                pairs_per_step=float(np.mean(th["pairs"])),
                force_ms=ms_force / max(nforce, 1), force_launches=nforce,
                build_ms_per_step=ms_build / max(npre, 1), pre_ms=ms_pre / max(npre, 1),
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


def cpu_baseline(timed_steps: int = 10, warm_steps: int = 2):
    from oracle import pyoracle

    path = pyoracle.REF_SO if os.path.exists(pyoracle.REF_SO) else pyoracle.ORACLE_SO
    kind = "reference" if path == pyoracle.REF_SO else "port"
    o = pyoracle.Oracle(path)
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
    return dict(value=timed_steps / dt, unit="timesteps/s", cores=1, kind=kind,
                sample=(f"N={n} NHC steps through the Verlet list, {warm_steps} warm-up + "
                        f"{timed_steps} timed steps after the initial build "
                        f"(setup {t_setup:.1f} s), single thread"),
                s_per_step=dt / timed_steps)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run three device-to-device copies of MIB MiB (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(see DESIGN.md for how the workload would shard)", file=sys.stderr)
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

    from moleculardynamics_amd import native

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
                 "roofline_frac_valu_fp64": F_PAIR * d["pairs_per_step"] / (d["force_ms"] * 1e-3) / PEAK_FP64,
                 "nhc_drift": drift64},
        "mixed_vs_fp64_speedup": tps / tps64,
        "mixed_error": {**err, "nhc_drift": drift32, "nhc_drift_fp64": drift64},
        "rebuilds": {"fp64": d["rebuilds"], "mixed": m["rebuilds"]},
        "roofline": {"bound": "valu", "kernel": "k_lj (LJ pair kernel, mixed)",
                     "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": None,
                     "launch_ms": m["force_ms"],
                     "step_frac": (max(168.0 * n / PEAK_HBM, F_PAIR * pairs / PEAK_FP32)
                                   / (m["ms_per_step"] * 1e-3))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
