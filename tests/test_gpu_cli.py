"""moldyn_run_gpu, the reference driver's command line on the device path,
run end to end: the reference's experiment gates and its committed results
(tests/golden/results, written by the reference's own moldyn_run).

* exp01 (experiments/exp01_nist_table.sh:30-70): --nist-config exits 0 with
  12 PASS rows, and each row's computed U, W and U_LRC equal the reference's
  table to the 4 significant figures both print.
* exp02 (experiments/exp02_energy_conservation.sh:26-94): four NVE runs of
  fcc(3), rho 0.80, T 1, seed 12345, LinearForceShift, skin 0.4, t = 1.
  The fitted slope of log(max |E - E0| / |E0|) against log(dt) lies in
  [1.85, 2.15]; and since velocity Verlet from the reference's own lattice
  is deterministic, every thermo row equals the reference's CSV row to the
  10 significant figures it prints (2e-9 relative; a wrong LRC, dof,
  pressure term or step order would miss by far more).
* A Nose-Hoover run with every output: the files' layout and row counts.
Argument errors are checked without a GPU."""
import csv
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "moleculardynamics_amd", "bin", "moldyn_run_gpu")
GOLDEN = os.path.join(ROOT, "tests", "golden", "results")


def _need_cli():
    if not os.path.exists(CLI):
        pytest.skip("moldyn_run_gpu is not built (python -c 'import __graft_entry__ as g; g.build()')")


def _run(args, timeout=300):
    print("running:", " ".join(args), flush=True)
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout)


def _rows(path):
    with open(path) as fh:
        lines = [l for l in fh if not l.startswith("#")]
    return list(csv.DictReader(lines))


@pytest.mark.parametrize("args,msg", [
    (["--bogus", "1"], "unknown argument"),
    (["--cells"], "missing value"),
    (["--cells", "x"], "invalid integer"),
    (["--cells", "3", "--density", "0.8"], "missing required flag"),
    (["--cells", "3", "--density", "0.8", "--temperature", "1", "--cutoff", "2.5",
      "--skin", "0.3", "--dt", "0.005", "--equilibrate", "0", "--production", "1",
      "--thermostat", "hot", "--seed", "1"], "--thermostat must be one of"),
    (["--cells", "3", "--density", "0.8", "--temperature", "1", "--cutoff", "2.5",
      "--skin", "0.3", "--dt", "0.005", "--equilibrate", "0", "--production", "1",
      "--thermostat", "none", "--seed", "-4"], "invalid --seed"),
])
def test_argument_errors(args, msg):
    _need_cli()
    r = _run(args)
    assert r.returncode == 1
    assert msg in r.stderr and "Usage:" in r.stderr


@pytest.mark.gpu
def test_exp01_nist_table(gpu_ok):
    _need_cli()
    r = _run(["--nist-config", os.path.join(ROOT, "tests", "golden", "nist", "lj")])
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines() if l.split() and l.split()[0].isdigit()]
    rows = [f for f in rows if len(f) == 16]
    assert len(rows) == 12 and all(f[15] == "PASS" for f in rows)
    ref = _rows(os.path.join(GOLDEN, "exp01_nist_table.csv"))
    for f, g in zip(rows, ref):
        assert (f[0], f[1]) == (g["cfg"], g["trunc"])
        for col, k in (("u_calc", 3), ("w_calc", 7), ("ulrc_calc", 11)):
            assert f[k] == g[col], (col, f[k], g[col])


def _exp02_slope(errs, dts):
    x, y = np.log(dts), np.log(errs)
    return float(np.sum((x - x.mean()) * (y - y.mean())) / np.sum((x - x.mean()) ** 2))


@pytest.mark.gpu
def test_exp02_energy_conservation(gpu_ok, tmp_path):
    _need_cli()
    dts = [0.001, 0.002, 0.004, 0.008]
    errs = []
    for dt in dts:
        steps = int(1.0 / dt + 0.5)
        out = str(tmp_path / f"thermo_dt{dt}.csv")
        r = _run(["--cells", "3", "--density", "0.80", "--temperature", "1.0", "--cutoff", "2.5",
                  "--skin", "0.4", "--dt", str(dt), "--equilibrate", "0", "--production",
                  str(steps), "--thermostat", "none", "--seed", "12345", "--truncation",
                  "linear-force-shift", "--thermo-out", out])
        assert r.returncode == 0, r.stderr
        mine = _rows(out)
        ref = _rows(os.path.join(GOLDEN, f"exp02_thermo_dt{dt}.csv"))
        assert len(mine) == len(ref) == steps + 1
        for a, b in zip(mine, ref):
            assert (a["step"], a["phase"]) == (b["step"], b["phase"])
            for col in ("time", "potential_energy", "kinetic_energy", "total_energy",
                        "temperature", "pressure"):
                x, y = float(a[col]), float(b[col])
                assert math.isclose(x, y, rel_tol=2e-9, abs_tol=1e-12), (dt, a["step"], col, x, y)
        e = np.array([float(a["total_energy"]) for a in mine])
        errs.append(np.max(np.abs(e - e[0])) / abs(e[0]))
    slope = _exp02_slope(np.array(errs), np.array(dts))
    print(f"exp02 slope {slope:.6f} (reference 2.003996)")
    assert 1.85 < slope < 2.15
    assert slope == pytest.approx(2.003996, abs=2e-3)


@pytest.mark.gpu
def test_nose_hoover_run_with_every_output(gpu_ok, tmp_path):
    _need_cli()
    paths = {k: str(tmp_path / f"{k}.out") for k in ("thermo", "rdf", "msd", "traj")}
    r = _run(["--cells", "4", "--density", "0.8442", "--temperature", "0.722", "--cutoff", "2.5",
              "--skin", "0.3", "--dt", "0.005", "--equilibrate", "200", "--production", "200",
              "--thermostat", "nose-hoover", "--seed", "7", "--thermo-out", paths["thermo"],
              "--rdf-out", paths["rdf"], "--msd-out", paths["msd"], "--traj-out", paths["traj"]])
    assert r.returncode == 0, r.stderr
    head = open(paths["thermo"]).read().splitlines()
    assert head[0].startswith("# command: ") and head[1].startswith("# git_sha: ")
    assert head[4] == "# truncation: truncated"
    th = _rows(paths["thermo"])
    assert len(th) == 401
    assert [t["phase"] for t in th[:2]] == ["start", "equilibrate"] and th[-1]["phase"] == "production"
    t_prod = np.array([float(t["temperature"]) for t in th[201:]])
    assert abs(t_prod.mean() / 0.722 - 1.0) < 0.1
    rdf = open(paths["rdf"]).read().splitlines()
    assert "# frames: 200" in rdf
    g = _rows(paths["rdf"])
    assert len(g) == 200
    r_peak = float(max(g, key=lambda q: float(q["g"]))["r"])
    assert 1.0 < r_peak < 1.2  # first shell of the LJ liquid
    msd = _rows(paths["msd"])
    assert len(msd) == 101 and float(msd[0]["msd"]) == 0.0 and float(msd[0]["vacf"]) == 1.0
    lines = open(paths["traj"]).read().splitlines()
    assert len(lines) == 200 * (256 + 2) and lines[0] == "256"
    assert lines[1].startswith("step=201 time=") and 'command="' in lines[1]
