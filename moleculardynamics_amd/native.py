"""ctypes binding of the C ABI in include/moldyn_hip.h (libmoldyn_hip.so).

This is plumbing for tests and the benchmark: it calls the HIP library and
nothing else. There is no CPU fallback: a missing library raises at load.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmoldyn_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "moldyn_hip.h")

MD_OK, MD_ERR_INVALID_ARGUMENT, MD_ERR_RUNTIME, MD_ERR_HIP, MD_ERR_CAPACITY, MD_ERR_STATE = range(6)
BRUTE, CELLS, VERLET = 0, 1, 2
FP64, MIXED = 0, 1
VV, BAOAB, NHC = 0, 1, 2

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


class LJParams(C.Structure):
    _fields_ = [("sigma", C.c_double), ("epsilon", C.c_double), ("cutoff", C.c_double),
                ("linear_force_shift", C.c_int)]


class Integrator(C.Structure):
    _fields_ = [("kind", C.c_int), ("dt", C.c_double), ("friction", C.c_double),
                ("temperature", C.c_double), ("tau", C.c_double), ("chain_length", C.c_int),
                ("dof", C.c_long), ("seed", C.c_uint64)]


class Thermo(C.Structure):
    _fields_ = [("potential", C.c_double), ("virial", C.c_double), ("kinetic", C.c_double),
                ("conserved", C.c_double), ("pairs", C.c_double)]


class WaterParams(C.Structure):
    _fields_ = [("qH", C.c_double), ("sigma_o", C.c_double), ("epsilon_o", C.c_double),
                ("r_oh", C.c_double), ("angle_hoh", C.c_double), ("mass_o", C.c_double),
                ("mass_h", C.c_double), ("cutoff", C.c_double), ("kmax", C.c_int),
                ("n2max", C.c_int), ("alpha", C.c_double), ("skin", C.c_double),
                ("tolerance", C.c_double), ("max_iterations", C.c_int)]


def water_params(cutoff=10.0, kmax=5, alpha=0.0, skin=1.0, n2max=0, tolerance=1e-12,
                 max_iterations=500, r_oh=1.0, angle_hoh=109.47):
    """SPC/E parameters of the reference (forcefield/spce.hpp:16-24)."""
    return WaterParams(0.42380, 3.16555789, 78.19743111, r_oh, angle_hoh, 15.9994, 1.008, cutoff,
                       kmax, n2max, alpha, skin, tolerance, max_iterations)


class MDError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads libmoldyn_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m moleculardynamics_amd.build`")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.md_ctx_create.argtypes = [vp, C.POINTER(vp)]
    lib.md_ctx_destroy.argtypes = [vp]
    lib.md_last_error.argtypes = [vp]
    lib.md_last_error.restype = C.c_char_p
    lib.md_ctx_synchronize.argtypes = [vp]
    lib.md_system_upload.argtypes = [vp, C.c_long, _dp] + [_dp] * 7
    lib.md_system_download.argtypes = [vp] + [_dp] * 9
    lib.md_system_set_positions.argtypes = [vp, _dp, _dp, _dp]
    lib.md_system_set_velocities.argtypes = [vp, _dp, _dp, _dp]
    lib.md_system_kinetic_energy.argtypes = [vp, _dp]
    lib.md_system_remove_com_motion.argtypes = [vp]
    lib.md_system_size.argtypes = [vp, C.POINTER(C.c_long)]
    lib.md_lj_set.argtypes = [vp, C.POINTER(LJParams)]
    lib.md_lj_forces.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
    lib.md_lj_energy_virial.argtypes = [vp, _dp, _dp]
    lib.md_lj_tail.argtypes = [C.POINTER(LJParams), C.c_long, _dp, _dp, _dp]
    lib.md_cells_usable.argtypes = [_dp, C.c_double]
    lib.md_cells_usable.restype = C.c_int
    lib.md_neigh_config.argtypes = [vp, C.c_double, C.c_double]
    lib.md_neigh_build.argtypes = [vp]
    lib.md_neigh_needs_rebuild.argtypes = [vp, C.POINTER(C.c_int)]
    lib.md_neigh_rebuild_if_needed.argtypes = [vp, C.POINTER(C.c_int)]
    lib.md_neigh_rebuild_count.argtypes = [vp, C.POINTER(C.c_long)]
    lib.md_neigh_export_pairs.argtypes = [vp, _u32p, C.c_long, C.POINTER(C.c_long)]
    lib.md_neigh_set_capacity.argtypes = [vp, C.c_int]
    lib.md_neigh_set_bricks.argtypes = [vp, C.c_int]
    lib.md_neigh_bricks.argtypes = [vp, C.POINTER(C.c_int)]
    lib.md_neigh_auto_capacity.argtypes = [C.c_long, _dp, C.c_double, C.c_double]
    lib.md_neigh_auto_capacity.restype = C.c_int
    lib.md_integrator_set.argtypes = [vp, C.POINTER(Integrator)]
    lib.md_run.argtypes = [vp, C.c_long, C.c_int, C.POINTER(Thermo)]
    lib.md_nhc_state.argtypes = [vp, _dp, _dp, _dp]
    lib.md_timing_enable.argtypes = [vp, C.c_int]
    lib.md_timing_select.argtypes = [vp, C.c_int]
    lib.md_timing_read.argtypes = [vp, C.c_int, _dp, C.POINTER(C.c_long)]
    lib.md_timing_reset.argtypes = [vp]
    lib.md_last_pair_count.argtypes = [vp, C.POINTER(C.c_long)]
    lib.md_fcc_lattice.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint64] + [_dp] * 7
    lib.md_system_set_forces.argtypes = [vp, _dp, _dp, _dp]
    lib.md_half_kick.argtypes = [vp, C.c_double]
    lib.md_drift.argtypes = [vp, C.c_double]
    lib.md_langevin_o.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64]
    lib.md_scale_velocities.argtypes = [vp, C.c_double]
    lib.md_lj_forces_cells.argtypes = [vp, C.c_double, _dp, _dp]
    for kind in ("rdf", "msd", "vacf"):
        getattr(lib, f"md_{kind}_destroy").argtypes = [vp]
        getattr(lib, f"md_{kind}_last_error").argtypes = [vp]
        getattr(lib, f"md_{kind}_last_error").restype = C.c_char_p
        if kind != "rdf":
            getattr(lib, f"md_{kind}_create").argtypes = [C.c_long, C.c_long, C.POINTER(vp)]
            getattr(lib, f"md_{kind}_push").argtypes = [vp, vp]
    lib.md_rdf_create.argtypes = [C.c_double, C.c_long, C.POINTER(vp)]
    lib.md_rdf_accumulate.argtypes = [vp, vp, C.c_int]
    lib.md_rdf_result.argtypes = [vp, _dp, _dp, _u64p, C.POINTER(C.c_long)]
    lib.md_msd_result.argtypes = [vp, _dp]
    lib.md_msd_diffusion.argtypes = [vp, C.c_double, C.c_long, C.c_long, _dp]
    lib.md_vacf_result.argtypes = [vp, C.c_int, _dp]
    lib.md_vacf_diffusion.argtypes = [vp, C.c_double, _dp]
    lib.md_device_peaks.argtypes = [_dp, _dp, _dp]
    lib.md_lj_forces_noalloc.argtypes = ([C.c_int, C.POINTER(LJParams), C.c_long, _dp] +
                                         [vp] * 6 + [_dp, _dp, vp])
    lib.md_batch_draw.argtypes = [C.c_int, C.c_uint64, _ip, _dp, _dp, _u64p]
    lib.md_batch_create.argtypes = ([vp, C.c_int, _lp, _dp] + [_dp] * 6 +
                                    [_dp, _u64p, C.POINTER(LJParams), C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.POINTER(vp)])
    lib.md_batch_destroy.argtypes = [vp]
    lib.md_batch_last_error.argtypes = [vp]
    lib.md_batch_last_error.restype = C.c_char_p
    lib.md_batch_prime.argtypes = [vp]
    lib.md_batch_run.argtypes = [vp, C.c_long, C.c_long, C.c_int, _dp]
    lib.md_batch_set_tpa.argtypes = [vp, _ip, C.c_int]
    lib.md_batch_set_pieces.argtypes = [vp, _ip]
    lib.md_batch_autotune.argtypes = [vp, C.c_int, _ip, _ip, _ip, _dp]
    lib.md_batch_download.argtypes = [vp] + [_dp] * 9
    lib.md_batch_rebuild_counts.argtypes = [vp, _lp]
    wp = C.POINTER(WaterParams)
    lib.md_ewald_parameters.argtypes = [C.c_long, _dp, C.c_double, C.c_double, C.c_double, _dp, _ip]
    lib.md_water_lattice.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_double, wp] + [_dp] * 6
    lib.md_water_create.argtypes = [vp, C.c_long, _dp] + [_dp] * 6 + [wp, C.POINTER(vp)]
    lib.md_water_destroy.argtypes = [vp]
    lib.md_water_last_error.argtypes = [vp]
    lib.md_water_last_error.restype = C.c_char_p
    lib.md_water_forces.argtypes = [vp, C.c_int, _dp, _dp]
    lib.md_water_alpha.argtypes = [vp, _dp, _ip]
    lib.md_water_build_list.argtypes = [vp]
    lib.md_water_shake.argtypes = [vp, _dp, _dp, _dp, _ip]
    lib.md_water_rattle.argtypes = [vp, _ip]
    lib.md_water_settle.argtypes = [vp, _dp, _dp, _dp]
    lib.md_water_settle_velocities.argtypes = [vp]
    lib.md_water_set_constraint.argtypes = [vp, C.c_int]
    lib.md_water_set_kspace.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.md_water_run.argtypes = [vp, C.c_long, C.c_double, C.c_int, _dp, _dp]
    lib.md_water_download.argtypes = [vp] + [_dp] * 9
    lib.md_water_set_state.argtypes = [vp] + [_dp] * 6
    lib.md_mix_create.argtypes = [vp, C.POINTER(vp)]
    lib.md_mix_destroy.argtypes = [vp]
    lib.md_mix_last_error.argtypes = [vp]
    lib.md_mix_last_error.restype = C.c_char_p
    lib.md_mix_upload.argtypes = [vp, C.c_long, _dp] + [_dp] * 7 + [_ip]
    lib.md_mix_set_params.argtypes = [vp, C.c_int, _dp, _dp, _dp, C.c_int]
    lib.md_mix_set_skin.argtypes = [vp, C.c_double]
    lib.md_mix_forces.argtypes = [vp, _dp, _dp, _dp]
    lib.md_mix_force_vjp.argtypes = [vp, C.c_double, C.c_double] + [_dp] * 7
    lib.md_mix_minimize.argtypes = [vp, C.c_int, C.c_double, C.c_double, _dp]
    lib.md_mix_run.argtypes = [vp, C.c_double, C.c_int, C.c_int, _dp, _dp]
    lib.md_mix_adjoint.argtypes = [vp, C.c_double, C.c_double] + [_dp] * 13
    lib.md_mix_download.argtypes = [vp] + [_dp] * 9
    lib.md_mix_set_state.argtypes = [vp] + [_dp] * 6
    lib.md_mix_list_builds.argtypes = [vp, _lp]
    lib.md_mix_random_state.argtypes = [C.c_long, C.c_double, C.c_uint64, C.c_double] + [_dp] * 7
    _lib = lib
    return lib


RDF_AUTO, RDF_ALL_PAIRS, RDF_CELLS = 0, 1, 2
RIGID_SHAKE, RIGID_SETTLE = 0, 1
KSPACE_EWALD, KSPACE_PME = 0, 1


def device_peaks():
    """Measured (HBM bytes/s, FP64 flops/s, FP32 flops/s) of the current device."""
    lib = load()
    h, d, f = C.c_double(), C.c_double(), C.c_double()
    st = lib.md_device_peaks(C.byref(h), C.byref(d), C.byref(f))
    if st != MD_OK:
        raise MDError(st, "md_device_peaks failed")
    return h.value, d.value, f.value


class _Handle:
    """An md_rdf / md_msd / md_vacf handle on the current device."""

    kind = ""

    def _make(self, *args):
        self.lib = load()
        h = C.c_void_p()
        st = getattr(self.lib, f"md_{self.kind}_create")(*args, C.byref(h))
        if st != MD_OK:
            raise MDError(st, f"md_{self.kind}_create failed")
        self.h = h

    def _chk(self, st):
        if st != MD_OK:
            msg = getattr(self.lib, f"md_{self.kind}_last_error")(self.h).decode()
            raise MDError(st, msg)

    def close(self):
        if getattr(self, "h", None):
            getattr(self.lib, f"md_{self.kind}_destroy")(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Rdf(_Handle):
    """Rdf(r_max, bins) accumulated from a Context's device system."""

    kind = "rdf"

    def __init__(self, r_max, bins):
        self.bins = int(bins)
        self._make(C.c_double(r_max), C.c_long(self.bins))

    def accumulate(self, ctx, source=RDF_AUTO):
        self._chk(self.lib.md_rdf_accumulate(self.h, ctx.h, source))

    def result(self):
        """(centres, g, hist, frames)."""
        c, g = np.zeros(self.bins), np.zeros(self.bins)
        h = np.zeros(self.bins, dtype=np.uint64)
        f = C.c_long()
        self._chk(self.lib.md_rdf_result(self.h, _p(c), _p(g), h.ctypes.data_as(_u64p),
                                         C.byref(f)))
        return c, g, h, f.value


class Msd(_Handle):
    kind = "msd"

    def __init__(self, natoms, max_lag):
        self.max_lag = int(max_lag)
        self._make(C.c_long(natoms), C.c_long(self.max_lag))

    def push(self, ctx):
        self._chk(self.lib.md_msd_push(self.h, ctx.h))

    def result(self):
        m = np.zeros(self.max_lag + 1)
        self._chk(self.lib.md_msd_result(self.h, _p(m)))
        return m

    def diffusion(self, dt, fit_from, fit_to):
        d = C.c_double()
        self._chk(self.lib.md_msd_diffusion(self.h, dt, fit_from, fit_to, C.byref(d)))
        return d.value


class Vacf(_Handle):
    kind = "vacf"

    def __init__(self, natoms, max_lag):
        self.max_lag = int(max_lag)
        self._make(C.c_long(natoms), C.c_long(self.max_lag))

    def push(self, ctx):
        self._chk(self.lib.md_vacf_push(self.h, ctx.h))

    def result(self, normalised=True):
        c = np.zeros(self.max_lag + 1)
        self._chk(self.lib.md_vacf_result(self.h, int(normalised), _p(c)))
        return c

    def diffusion(self, dt):
        d = C.c_double()
        self._chk(self.lib.md_vacf_diffusion(self.h, dt, C.byref(d)))
        return d.value


NOALLOC_MAX_ATOMS = 1 << 20


def lj_forces_noalloc(source, sigma, epsilon, cutoff, lfs, n, L, x, y, z, fx=0, fy=0, fz=0,
                      stream=None):
    """md_lj_forces_noalloc on device arrays given as integer addresses (for
    example torch tensors' data_ptr()); fx = fy = fz = 0 asks for E and W
    only. Returns (E, W); raises MDError with md_last_error(NULL)'s text."""
    lib = load()
    p = LJParams(sigma, epsilon, cutoff, int(lfs))
    Lb = _col(L)
    e, w = C.c_double(), C.c_double()
    ptr = [C.c_void_p(int(v)) if v else None for v in (x, y, z, fx, fy, fz)]
    st = lib.md_lj_forces_noalloc(source, C.byref(p), n, _p(Lb), *ptr, C.byref(e), C.byref(w),
                                  stream)
    if st != MD_OK:
        raise MDError(st, lib.md_last_error(None).decode())
    return e.value, w.value


def declared_symbols(header: str = HEADER) -> list[str]:
    """Function names declared in include/moldyn_hip.h."""
    import re

    txt = open(header).read()
    return sorted(set(re.findall(r"\b(md_[a-z0-9_]+)\s*\(", txt)))


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _col(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One md_ctx: a device-resident system plus its list and integrator."""

    def __init__(self, stream=None):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.md_ctx_create(stream, C.byref(h))
        if st != MD_OK:
            raise MDError(st, "md_ctx_create failed (is a GPU visible?)")
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.md_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        if st != MD_OK:
            raise MDError(st, self.lib.md_last_error(self.h).decode())

    # ---- state ------------------------------------------------------------
    def upload(self, L, pos, vel=None, mass=None):
        pos = np.asarray(pos, dtype=np.float64)
        n = pos.shape[0]
        cols = [_col(pos[:, d]) for d in range(3)]
        vcols = [None] * 3 if vel is None else [_col(np.asarray(vel)[:, d]) for d in range(3)]
        m = None if mass is None else _col(mass)
        Lb = _col(L)
        self._keep = (cols, vcols, m, Lb)
        self._chk(self.lib.md_system_upload(self.h, n, _p(Lb), *[_p(c) for c in cols],
                                            *[_p(c) for c in vcols], _p(m)))
        self.n = n

    def download(self):
        a = [np.zeros(self.n) for _ in range(9)]
        self._chk(self.lib.md_system_download(self.h, *[_p(v) for v in a]))
        return (np.stack(a[0:3], axis=1), np.stack(a[3:6], axis=1), np.stack(a[6:9], axis=1))

    def set_positions(self, pos):
        cols = [_col(np.asarray(pos)[:, d]) for d in range(3)]
        self._chk(self.lib.md_system_set_positions(self.h, *[_p(c) for c in cols]))

    def set_velocities(self, vel):
        cols = [_col(np.asarray(vel)[:, d]) for d in range(3)]
        self._chk(self.lib.md_system_set_velocities(self.h, *[_p(c) for c in cols]))

    def kinetic_energy(self) -> float:
        ke = C.c_double()
        self._chk(self.lib.md_system_kinetic_energy(self.h, C.byref(ke)))
        return ke.value

    def remove_com_motion(self):
        self._chk(self.lib.md_system_remove_com_motion(self.h))

    # ---- LJ ---------------------------------------------------------------
    def set_lj(self, sigma=1.0, epsilon=1.0, cutoff=2.5, lfs=False):
        self.ljp = LJParams(sigma, epsilon, cutoff, int(lfs))
        self._chk(self.lib.md_lj_set(self.h, C.byref(self.ljp)))

    def forces(self, source=VERLET, precision=FP64):
        e, w = C.c_double(), C.c_double()
        self._chk(self.lib.md_lj_forces(self.h, source, precision, C.byref(e), C.byref(w)))
        return e.value, w.value

    def energy_virial(self):
        e, w = C.c_double(), C.c_double()
        self._chk(self.lib.md_lj_energy_virial(self.h, C.byref(e), C.byref(w)))
        return e.value, w.value

    # ---- neighbour list ----------------------------------------------------
    def config_list(self, cutoff, skin):
        self._chk(self.lib.md_neigh_config(self.h, cutoff, skin))

    def build_list(self):
        self._chk(self.lib.md_neigh_build(self.h))

    def needs_rebuild(self) -> bool:
        f = C.c_int()
        self._chk(self.lib.md_neigh_needs_rebuild(self.h, C.byref(f)))
        return bool(f.value)

    def rebuild_if_needed(self) -> bool:
        f = C.c_int()
        self._chk(self.lib.md_neigh_rebuild_if_needed(self.h, C.byref(f)))
        return bool(f.value)

    def rebuild_count(self) -> int:
        c = C.c_long()
        self._chk(self.lib.md_neigh_rebuild_count(self.h, C.byref(c)))
        return c.value

    def set_capacity(self, per_atom: int):
        self._chk(self.lib.md_neigh_set_capacity(self.h, per_atom))

    def set_bricks(self, cells: int):
        """Cells per brick side of the LDS-staged pair kernels (-1 automatic,
        0 off, the default); the list must be built again afterwards."""
        self._chk(self.lib.md_neigh_set_bricks(self.h, cells))

    def bricks(self) -> int:
        """Bricks of the current list layout (0: bricks off)."""
        v = C.c_int(0)
        self._chk(self.lib.md_neigh_bricks(self.h, C.byref(v)))
        return int(v.value)

    def pairs(self) -> np.ndarray:
        npairs = C.c_long()
        st = self.lib.md_neigh_export_pairs(self.h, None, 0, C.byref(npairs))
        if st not in (MD_OK, MD_ERR_CAPACITY):
            self._chk(st)
        out = np.zeros(2 * max(1, npairs.value), dtype=np.uint32)
        self._chk(self.lib.md_neigh_export_pairs(self.h, out.ctypes.data_as(_u32p), npairs.value,
                                                 C.byref(npairs)))
        return out[: 2 * npairs.value].reshape(-1, 2)

    # ---- integration -------------------------------------------------------
    def set_integrator(self, kind=VV, dt=0.005, friction=1.0, temperature=1.0, tau=0.5,
                       chain_length=3, dof=0, seed=1):
        self.itg = Integrator(kind, dt, friction, temperature, tau, chain_length, dof, seed)
        self._chk(self.lib.md_integrator_set(self.h, C.byref(self.itg)))

    def run(self, nsteps, precision=FP64, thermo=True):
        th = (Thermo * nsteps)() if thermo else None
        self._chk(self.lib.md_run(self.h, nsteps, precision, th))
        if not thermo:
            return None
        a = np.frombuffer(th, dtype=np.float64).reshape(nsteps, 5)
        return dict(pe=a[:, 0].copy(), virial=a[:, 1].copy(), ke=a[:, 2].copy(),
                    conserved=a[:, 3].copy(), pairs=a[:, 4].copy())

    def nhc_state(self, M=3):
        cq = C.c_double()
        xi, p = np.zeros(M), np.zeros(M)
        self._chk(self.lib.md_nhc_state(self.h, C.byref(cq), _p(xi), _p(p)))
        return cq.value, xi, p

    # ---- measurement -------------------------------------------------------
    def timing(self, on=True):
        self._chk(self.lib.md_timing_enable(self.h, int(on)))

    # kernel groups of timing_read / timing_select
    TIME_PAIR, TIME_BUILD, TIME_STEP = 1, 2, 4

    def timing_select(self, mask):
        """Times only the kernel groups in mask (TIME_PAIR | TIME_BUILD | TIME_STEP)."""
        self._chk(self.lib.md_timing_select(self.h, int(mask)))

    def timing_read(self, kernel_id=0):
        ms, k = C.c_double(), C.c_long()
        self._chk(self.lib.md_timing_read(self.h, kernel_id, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    def timing_reset(self):
        self._chk(self.lib.md_timing_reset(self.h))

    def pair_count(self) -> int:
        c = C.c_long()
        self._chk(self.lib.md_last_pair_count(self.h, C.byref(c)))
        return c.value

    def synchronize(self):
        self._chk(self.lib.md_ctx_synchronize(self.h))


def auto_capacity(n, L, cutoff, skin) -> int:
    """The automatic neighbour-row capacity (md_neigh_auto_capacity)."""
    return int(load().md_neigh_auto_capacity(n, _p(_col(L)), cutoff, skin))


def cells_usable(L, cutoff) -> bool:
    return bool(load().md_cells_usable(_p(_col(L)), cutoff))


def batch_draw(count: int, master_seed: int):
    """The replica recipe through the C ABI: (cells, rho, T, seeds) per replica."""
    cells = np.zeros(count, dtype=np.int32)
    rho, T = np.zeros(count), np.zeros(count)
    seeds = np.zeros(count, dtype=np.uint64)
    st = load().md_batch_draw(count, master_seed, cells.ctypes.data_as(C.POINTER(C.c_int)),
                              _p(rho), _p(T), seeds.ctypes.data_as(C.POINTER(C.c_uint64)))
    if st != MD_OK:
        raise MDError(st, "md_batch_draw failed")
    return cells, rho, T, seeds


class Batch:
    """md_batch: many independent replicas advanced by one set of launches per step."""

    def __init__(self, systems, temperatures, seeds, *, sigma=1.0, epsilon=1.0, cutoff=2.5,
                 skin=0.3, dt=0.005, friction=1.0, precision=MIXED):
        """systems: list of (L[3], pos (n,3), vel (n,3))."""
        self.lib = load()
        self.sizes = np.array([s[1].shape[0] for s in systems], dtype=np.int64)
        self.nrep = len(systems)
        L = np.concatenate([np.asarray(s[0], dtype=float) for s in systems])
        pos = np.concatenate([s[1] for s in systems])
        vel = np.concatenate([s[2] for s in systems])
        cols = [_col(pos[:, d]) for d in range(3)] + [_col(vel[:, d]) for d in range(3)]
        T = _col(temperatures)
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        self.ljp = LJParams(sigma, epsilon, cutoff, 0)
        h = C.c_void_p()
        st = self.lib.md_batch_create(None, self.nrep, self.sizes.ctypes.data_as(C.POINTER(C.c_long)),
                                      _p(L), *[_p(c) for c in cols], _p(T),
                                      sd.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(self.ljp),
                                      skin, dt, friction, precision, C.byref(h))
        self.h = h
        if st != MD_OK:
            msg = self.lib.md_batch_last_error(h).decode() if h else "md_batch_create failed"
            self.close()
            raise MDError(st, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.md_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        if st != MD_OK:
            raise MDError(st, self.lib.md_batch_last_error(self.h).decode())

    def prime(self):
        self._chk(self.lib.md_batch_prime(self.h))

    def run(self, nsteps, segment=100, graph=True, thermo=False):
        out = np.zeros(nsteps * self.nrep * 4) if thermo else None
        self._chk(self.lib.md_batch_run(self.h, nsteps, segment, int(graph), _p(out)))
        return None if out is None else out.reshape(nsteps, self.nrep, 4)

    def set_tpa(self, tpa, block=512):
        t = np.ascontiguousarray(tpa, dtype=np.int32)
        self._chk(self.lib.md_batch_set_tpa(self.h, t.ctypes.data_as(C.POINTER(C.c_int)), block))

    def set_pieces(self, pieces):
        p = np.ascontiguousarray(pieces, dtype=np.int32)
        self._chk(self.lib.md_batch_set_pieces(self.h, p.ctypes.data_as(C.POINTER(C.c_int))))

    def autotune(self, trial_steps=5):
        """Returns (lanes per atom, pieces per replica, block, ms per step)."""
        t = np.zeros(self.nrep, dtype=np.int32)
        p = np.zeros(self.nrep, dtype=np.int32)
        blk = C.c_int()
        ms = C.c_double()
        self._chk(self.lib.md_batch_autotune(self.h, trial_steps, t.ctypes.data_as(C.POINTER(C.c_int)),
                                             p.ctypes.data_as(C.POINTER(C.c_int)), C.byref(blk),
                                             C.byref(ms)))
        return t, p, blk.value, ms.value

    def download(self):
        n = int(self.sizes.sum())
        a = [np.zeros(n) for _ in range(9)]
        self._chk(self.lib.md_batch_download(self.h, *[_p(v) for v in a]))
        split = np.cumsum(self.sizes)[:-1]
        out = []
        for k in range(3):
            arr = np.stack(a[3 * k:3 * k + 3], axis=1)
            out.append(np.split(arr, split))
        return out  # [positions per replica], [velocities], [forces]

    def rebuild_counts(self):
        c = np.zeros(self.nrep, dtype=np.int64)
        self._chk(self.lib.md_batch_rebuild_counts(self.h, c.ctypes.data_as(C.POINTER(C.c_long))))
        return c


def ewald_parameters(n, L, sum_q2, cutoff, accuracy):
    a, k = C.c_double(), C.c_int()
    st = load().md_ewald_parameters(n, _p(_col(L)), sum_q2, cutoff, accuracy, C.byref(a), C.byref(k))
    if st != MD_OK:
        raise MDError(st, "md_ewald_parameters failed")
    return a.value, k.value


def water_lattice(m, L, seed, temperature, params=None):
    params = params or water_params()
    n = 3 * m ** 3
    a = [np.zeros(n) for _ in range(6)]
    st = load().md_water_lattice(m, L, seed, temperature, C.byref(params), *[_p(v) for v in a])
    if st != MD_OK:
        raise MDError(st, "md_water_lattice failed")
    return np.stack(a[0:3], axis=1), np.stack(a[3:6], axis=1)


class Water:
    """md_water: rigid SPC/E sites with Ewald on the device."""

    def __init__(self, L, pos, vel=None, params=None):
        self.lib = load()
        self.params = params or water_params()
        pos = np.asarray(pos, dtype=float)
        self.n = pos.shape[0]
        cols = [_col(pos[:, d]) for d in range(3)]
        vcols = [None] * 3 if vel is None else [_col(np.asarray(vel)[:, d]) for d in range(3)]
        h = C.c_void_p()
        st = self.lib.md_water_create(None, self.n, _p(_col(L)), *[_p(c) for c in cols],
                                      *[_p(c) for c in vcols], C.byref(self.params), C.byref(h))
        self.h = h
        if st != MD_OK:
            msg = self.lib.md_water_last_error(h).decode() if h else "md_water_create failed"
            self.close()
            raise MDError(st, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.md_water_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        if st != MD_OK:
            raise MDError(st, self.lib.md_water_last_error(self.h).decode())

    def forces(self, precision=FP64):
        t = np.zeros(6)
        pe = C.c_double()
        self._chk(self.lib.md_water_forces(self.h, precision, _p(t), C.byref(pe)))
        return dict(disp=t[0], lrc=t[1], real=t[2], recip=t[3], self=t[4], intra=t[5]), pe.value

    def alpha(self):
        a, k = C.c_double(), C.c_int()
        self._chk(self.lib.md_water_alpha(self.h, C.byref(a), C.byref(k)))
        return a.value, k.value

    def shake(self, ref_pos):
        cols = [_col(np.asarray(ref_pos)[:, d]) for d in range(3)]
        it = C.c_int()
        self._chk(self.lib.md_water_shake(self.h, *[_p(c) for c in cols], C.byref(it)))
        return it.value

    def rattle(self):
        it = C.c_int()
        self._chk(self.lib.md_water_rattle(self.h, C.byref(it)))
        return it.value

    def settle(self, ref_pos):
        """Closed-form position constraint against ref_pos (SETTLE)."""
        cols = [_col(np.asarray(ref_pos)[:, d]) for d in range(3)]
        self._chk(self.lib.md_water_settle(self.h, *[_p(c) for c in cols]))
        return 1

    def settle_velocities(self):
        self._chk(self.lib.md_water_settle_velocities(self.h))
        return 1

    def set_constraint(self, method):
        """RIGID_SHAKE (default) or RIGID_SETTLE for run()."""
        self._chk(self.lib.md_water_set_constraint(self.h, method))

    def set_kspace(self, method, grid=0, order=0):
        """KSPACE_EWALD (default) or KSPACE_PME with `grid` points per axis
        (0: automatic) and B-spline `order` (0: 8) for forces() and run()."""
        self._chk(self.lib.md_water_set_kspace(self.h, method, grid, order))

    def run(self, nsteps, dt, precision=FP64):
        pe, ke = np.zeros(nsteps), np.zeros(nsteps)
        self._chk(self.lib.md_water_run(self.h, nsteps, dt, precision, _p(pe), _p(ke)))
        return pe, ke

    def download(self):
        a = [np.zeros(self.n) for _ in range(9)]
        self._chk(self.lib.md_water_download(self.h, *[_p(v) for v in a]))
        return (np.stack(a[0:3], axis=1), np.stack(a[3:6], axis=1), np.stack(a[6:9], axis=1))

    def set_state(self, pos=None, vel=None):
        pc = [None] * 3 if pos is None else [_col(np.asarray(pos)[:, d]) for d in range(3)]
        vc = [None] * 3 if vel is None else [_col(np.asarray(vel)[:, d]) for d in range(3)]
        self._chk(self.lib.md_water_set_state(self.h, *[_p(c) for c in pc], *[_p(c) for c in vc]))


def fcc_lattice(cells: int, density: float, temperature: float, seed: int):
    """fccLattice through the C ABI: returns (L[3], positions (N,3), velocities (N,3))."""
    n = 4 * cells ** 3
    L = np.zeros(3)
    a = [np.zeros(n) for _ in range(6)]
    st = load().md_fcc_lattice(cells, density, temperature, seed, _p(L), *[_p(v) for v in a])
    if st != MD_OK:
        raise MDError(st, "md_fcc_lattice failed")
    return L, np.stack(a[0:3], axis=1), np.stack(a[3:6], axis=1)


def lj_tail(n, L, sigma=1.0, epsilon=1.0, cutoff=2.5, lfs=False):
    p = LJParams(sigma, epsilon, cutoff, int(lfs))
    u, pr = C.c_double(), C.c_double()
    st = load().md_lj_tail(C.byref(p), n, _p(_col(L)), C.byref(u), C.byref(pr))
    if st != MD_OK:
        raise MDError(st, "md_lj_tail failed")
    return u.value, pr.value


def _vec3(a, n):
    """(n, 3) array -> three contiguous columns (or three Nones)."""
    if a is None:
        return [None] * 3
    a = np.asarray(a, dtype=np.float64).reshape(n, 3)
    return [_col(a[:, d]) for d in range(3)]


def type_pairs(ntypes: int):
    """Canonical order of the unordered type pairs (a <= b), row-major."""
    return [(a, b) for a in range(ntypes) for b in range(a, ntypes)]


def mix_random_state(n: int, box_length: float, seed: int, temperature: float, mass=None):
    """md_mix_random_state: (positions (n,3), velocities (n,3))."""
    a = [np.zeros(n) for _ in range(6)]
    m = None if mass is None else _col(mass)
    st = load().md_mix_random_state(n, box_length, seed, temperature, _p(m), *[_p(v) for v in a])
    if st != MD_OK:
        raise MDError(st, "md_mix_random_state failed")
    return np.stack(a[0:3], axis=1), np.stack(a[3:6], axis=1)


class Mix:
    """md_mix: per-type-pair LJ with parameter derivatives, force VJPs,
    steepest descent, velocity-Verlet runs with checkpoints and their adjoint."""

    def __init__(self, L, pos, types, sigma, epsilon, cutoff, lfs=False, vel=None, mass=None,
                 skin=0.3, stream=None):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.md_mix_create(stream, C.byref(h))
        self.h = h
        if st != MD_OK:
            msg = self.lib.md_mix_last_error(h).decode() if h else "md_mix_create failed"
            self.close()
            raise MDError(st, msg)
        self.upload(L, pos, types, vel, mass)
        self.set_params(sigma, epsilon, cutoff, lfs)
        self._chk(self.lib.md_mix_set_skin(self.h, skin))

    def close(self):
        if getattr(self, "h", None):
            self.lib.md_mix_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        if st != MD_OK:
            raise MDError(st, self.lib.md_mix_last_error(self.h).decode())

    def upload(self, L, pos, types, vel=None, mass=None):
        pos = np.asarray(pos, dtype=np.float64)
        self.n = pos.shape[0]
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        cols, vcols = _vec3(pos, self.n), _vec3(vel, self.n)
        m = None if mass is None else _col(mass)
        Lb = _col(L)
        self._chk(self.lib.md_mix_upload(self.h, self.n, _p(Lb), *[_p(c) for c in cols],
                                         *[_p(c) for c in vcols], _p(m),
                                         self.types.ctypes.data_as(C.POINTER(C.c_int))))

    def set_params(self, sigma, epsilon, cutoff, lfs=False):
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        self.ntypes = int(sg.shape[0])
        ep = np.ascontiguousarray(epsilon, dtype=np.float64)
        rc = np.ascontiguousarray(cutoff, dtype=np.float64)
        self._chk(self.lib.md_mix_set_params(self.h, self.ntypes, _p(sg.reshape(-1)),
                                             _p(ep.reshape(-1)), _p(rc.reshape(-1)), int(lfs)))

    @property
    def npairs(self):
        return self.ntypes * (self.ntypes + 1) // 2

    def forces(self, derivatives=False):
        """(E, W) or (E, W, tables) with tables[k] = (E_ab, W_ab, dE/deps,
        dE/dsigma, dW/deps, dW/dsigma) for the canonical type pairs."""
        e, w = C.c_double(), C.c_double()
        t = np.zeros(6 * self.npairs) if derivatives else None
        self._chk(self.lib.md_mix_forces(self.h, C.byref(e), C.byref(w), _p(t)))
        if derivatives:
            return e.value, w.value, t.reshape(self.npairs, 6)
        return e.value, w.value

    def force_vjp(self, g=None, w_energy=0.0, w_virial=0.0):
        """(dF/dx)^T g + wE dE/dx + wW dW/dx, and the same cotangent on
        (eps_ab, sigma_ab) as an (npairs, 2) array."""
        gc = _vec3(g, self.n)
        o = [np.zeros(self.n) for _ in range(3)]
        th = np.zeros(2 * self.npairs)
        self._chk(self.lib.md_mix_force_vjp(self.h, w_energy, w_virial, *[_p(c) for c in gc],
                                            *[_p(v) for v in o], _p(th)))
        return np.stack(o, axis=1), th.reshape(self.npairs, 2)

    def minimize(self, nsteps, step=1e-3, max_disp=0.05):
        e = np.zeros(nsteps + 1)
        self._chk(self.lib.md_mix_minimize(self.h, int(nsteps), step, max_disp, _p(e)))
        return e

    def run(self, dt, nsteps, checkpoint_every=20):
        e, w = np.zeros(nsteps + 1), np.zeros(nsteps + 1)
        self._chk(self.lib.md_mix_run(self.h, dt, int(nsteps), int(checkpoint_every), _p(e), _p(w)))
        return e, w

    def adjoint(self, w_energy=0.0, w_virial=0.0, gx=None, gv=None):
        """(dL/dtheta (npairs, 2: eps, sigma), dL/dx0, dL/dv0) of the last run."""
        gxc, gvc = _vec3(gx, self.n), _vec3(gv, self.n)
        th = np.zeros(2 * self.npairs)
        dx = [np.zeros(self.n) for _ in range(3)]
        dv = [np.zeros(self.n) for _ in range(3)]
        self._chk(self.lib.md_mix_adjoint(self.h, w_energy, w_virial, *[_p(c) for c in gxc],
                                          *[_p(c) for c in gvc], _p(th), *[_p(v) for v in dx],
                                          *[_p(v) for v in dv]))
        return th.reshape(self.npairs, 2), np.stack(dx, axis=1), np.stack(dv, axis=1)

    def download(self):
        a = [np.zeros(self.n) for _ in range(9)]
        self._chk(self.lib.md_mix_download(self.h, *[_p(v) for v in a]))
        return (np.stack(a[0:3], axis=1), np.stack(a[3:6], axis=1), np.stack(a[6:9], axis=1))

    def set_state(self, pos=None, vel=None):
        pc, vc = _vec3(pos, self.n), _vec3(vel, self.n)
        self._chk(self.lib.md_mix_set_state(self.h, *[_p(c) for c in pc], *[_p(c) for c in vc]))

    def list_builds(self):
        b = C.c_long()
        self._chk(self.lib.md_mix_list_builds(self.h, C.byref(b)))
        return b.value
