import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long-running statistical or full-size checks")


def nist_lj(cfg: int) -> str:
    return os.path.join(GOLDEN, "nist", "lj", f"lj_sample_config_periodic{cfg}.txt")


def nist_spce(cfg: int) -> str:
    return os.path.join(GOLDEN, "nist", "spce", f"spce_sample_config_periodic{cfg}.txt")


# Binary mixture of the differentiated workload: 80:20 A:B, eps_AA = 1,
# sigma_AA = 1, eps_AB = 1.5, sigma_AB = 0.8, eps_BB = 0.5, sigma_BB = 0.88,
# cutoff 2.5 sigma_ab.
KA_SIGMA = [[1.0, 0.8], [0.8, 0.88]]
KA_EPS = [[1.0, 1.5], [1.5, 0.5]]
KA_RC = [[2.5, 2.0], [2.0, 2.2]]


def ka_types(n: int):
    import numpy as np

    t = np.zeros(n, dtype=np.int32)
    t[int(round(0.8 * n)):] = 1
    return t


def ka_frame(n: int, rho: float, seed: int, jitter: float = 0.0):
    """A mixture configuration on a jittered simple-cubic lattice (no
    overlaps, so the pair terms stay moderate), types by ka_types."""
    import numpy as np
    from oracle.pyoracle import Frame

    L = (n / rho) ** (1.0 / 3.0)
    m = int(np.ceil(n ** (1.0 / 3.0)))
    g = (np.arange(m) + 0.5) * (L / m) - 0.5 * L
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)[:n]
    rng = np.random.default_rng(seed)
    pts = pts + rng.uniform(-jitter, jitter, size=pts.shape)
    pts = pts - L * np.floor(pts / L + 0.5)
    v = rng.normal(size=(n, 3))
    v -= v.mean(axis=0)
    return Frame(np.array([L, L, L]), pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(),
                 v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy())


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle

    if not os.path.exists(pyoracle.ORACLE_SO):
        pyoracle.build()
    return pyoracle.Oracle(pyoracle.ORACLE_SO)


@pytest.fixture(scope="session")
def ref():
    from oracle import pyoracle

    if not os.path.exists(pyoracle.REF_SO):
        pytest.skip("reference build (oracle/_ref) not present")
    return pyoracle.Oracle(pyoracle.REF_SO)


@pytest.fixture(scope="session")
def gpu_ok():
    import ctypes

    try:
        hip = ctypes.CDLL("libamdhip64.so")
        cnt = ctypes.c_int(0)
        if hip.hipGetDeviceCount(ctypes.byref(cnt)) != 0 or cnt.value < 1:
            pytest.skip("no GPU visible")
    except OSError:
        pytest.skip("HIP runtime not loadable")
    return True


@pytest.fixture()
def ctx(gpu_ok):
    from moleculardynamics_amd import native

    c = native.Context()
    yield c
    c.close()
