"""ctypes driver for the CPU oracle (oracle/_build/liboracle.so) and, where it
was built, the reference compiled from its own sources
(oracle/_ref/libmoldyn_ref.so). Both export the same ``orc_*`` C symbols.

TEST INFRASTRUCTURE ONLY: used by tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of bench.py as the checker / CPU baseline. The product
(``moleculardynamics_amd``) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmoldyn_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_up = C.POINTER(C.c_uint)


def _ptr(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


@dataclass
class Frame:
    """Positions (and optionally velocities) of one periodic system."""

    L: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray | None = None
    vy: np.ndarray | None = None
    vz: np.ndarray | None = None
    types: list = field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "Frame":
        c = lambda a: None if a is None else a.copy()  # noqa: E731
        return Frame(self.L.copy(), self.x.copy(), self.y.copy(), self.z.copy(),
                     c(self.vx), c(self.vy), c(self.vz), list(self.types))


def build(ref: bool = False) -> None:
    """Compile the oracle (and the reference when ``ref`` and its sources are
    present) with the Makefile beside this file."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir(os.environ.get("REF", "/root/reference") + "/src/moldyn"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class Oracle:
    """Wrapper over one ``orc_*`` library (the oracle or the reference build)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        self.path = path
        lib = C.CDLL(path)
        self.lib = lib
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_is_reference.restype = C.c_int
        lib.orc_fcc_lattice.argtypes = [C.c_int, C.c_double, C.c_double, C.c_ulonglong] + [_dp] * 7
        lib.orc_lj_ev_brute.argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                        C.c_double, C.c_int, _dp]
        lib.orc_lj_forces.argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        lib.orc_lj_lrc.argtypes = [C.c_long, _dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        if hasattr(lib, "orc_lj_ev_compensated"):
            lib.orc_lj_ev_compensated.argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_double,
                                                  C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        lib.orc_cells_usable.argtypes = [_dp, C.c_double]
        lib.orc_rdf.argtypes = [C.c_long, C.c_long, _dp, _dp, C.c_double, C.c_long, _dp, _dp, _dp]
        lib.orc_msd.argtypes = [C.c_long, C.c_long, C.c_long, _dp, _dp, C.c_double, C.c_long,
                                C.c_long, _dp, _dp]
        lib.orc_vacf.argtypes = [C.c_long, C.c_long, C.c_long, _dp, C.c_double, _dp, _dp]
        lib.orc_verlet_pairs.argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                         _up, C.c_long]
        lib.orc_verlet_pairs.restype = C.c_long
        lib.orc_needs_rebuild.argtypes = [C.c_long, _dp] + [_dp] * 6 + [C.c_double, C.c_double]
        lib.orc_sim_create.argtypes = ([C.c_long, _dp] + [_dp] * 7 +
                                       [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                        C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_long, C.c_ulonglong])
        lib.orc_sim_create.restype = C.c_void_p
        lib.orc_sim_destroy.argtypes = [C.c_void_p]
        lib.orc_sim_run.argtypes = [C.c_void_p, C.c_long, _dp, _dp, _dp, _dp]
        lib.orc_sim_get.argtypes = [C.c_void_p] + [_dp] * 9
        lib.orc_sim_set_velocities.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.orc_sim_rebuilds.argtypes = [C.c_void_p]
        lib.orc_sim_rebuilds.restype = C.c_long
        lib.orc_sim_conserved.argtypes = [C.c_void_p]
        lib.orc_sim_conserved.restype = C.c_double
        lib.orc_coulomb_prefactor.restype = C.c_double
        lib.orc_ewald.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, _lp, C.c_double, C.c_int,
                                  C.c_double, C.c_int, _dp, _dp, _dp, _dp]
        lib.orc_spce_forces.argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                        C.c_double, _dp, _dp, _dp, _dp]
        if hasattr(lib, "orc_pme"):  # the oracle's own (the reference has no PME)
            lib.orc_pme.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                    C.POINTER(C.c_int), C.c_int, C.c_int, _dp, _dp, _dp, _dp]
            lib.orc_pme_bsq.argtypes = [C.c_int, C.c_int, C.c_int]
            lib.orc_pme_bsq.restype = C.c_double
        lib.orc_rigid_shake.argtypes = [C.c_long, _dp] + [_dp] * 7 + [C.c_double, C.c_double,
                                                                      C.c_double, C.c_int]
        lib.orc_rigid_rattle.argtypes = [C.c_long, _dp] + [_dp] * 7 + [C.c_double, C.c_double,
                                                                       C.c_double, C.c_int]
        lib.orc_water_run.argtypes = ([C.c_long, _dp] + [_dp] * 6 +
                                      [C.c_double, C.c_int, C.c_double, C.c_double, C.c_long,
                                       _dp, _dp])
        if hasattr(lib, "orc_settle_positions"):
            lib.orc_settle_positions.argtypes = [C.c_long, _dp] + [_dp] * 7 + [C.c_double,
                                                                               C.c_double]
            lib.orc_settle_velocities.argtypes = [C.c_long, _dp] + [_dp] * 7 + [C.c_double,
                                                                                C.c_double]
            lib.orc_water_run_settle.argtypes = lib.orc_water_run.argtypes
        if hasattr(lib, "orc_mix_lj"):
            lib.orc_mix_lj.argtypes = ([C.c_long, _dp, _dp, _dp, _dp, _ip, C.c_int, _dp, _dp, _dp,
                                        C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp])
            lib.orc_mix_minimize.argtypes = ([C.c_long, _dp, _dp, _dp, _dp, _ip, C.c_int, _dp,
                                              _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                              _dp])
            lib.orc_mix_vv.argtypes = ([C.c_long, _dp] + [_dp] * 7 + [_ip, C.c_int, _dp, _dp, _dp,
                                                                      C.c_int, C.c_double, C.c_int,
                                                                      _dp, _dp])

    @property
    def is_reference(self) -> bool:
        return bool(self.lib.orc_is_reference())

    def error(self) -> str:
        return self.lib.orc_last_error().decode()

    # ---- inputs -------------------------------------------------------------
    def fcc_lattice(self, cells: int, rho: float, T: float, seed: int) -> Frame:
        n = 4 * cells ** 3
        L = np.zeros(3)
        a = [np.zeros(n) for _ in range(6)]
        self.lib.orc_fcc_lattice(cells, rho, T, seed, _ptr(L), *[_ptr(v) for v in a])
        return Frame(L, *a)

    # ---- Lennard-Jones ------------------------------------------------------
    def lj_energy_virial(self, fr: Frame, sig=1.0, eps=1.0, rc=2.5, lfs=False):
        ev = np.zeros(2)
        self.lib.orc_lj_ev_brute(fr.n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z),
                                 sig, eps, rc, int(lfs), _ptr(ev))
        return float(ev[0]), float(ev[1])

    def lj_forces(self, fr: Frame, sig=1.0, eps=1.0, rc=2.5, lfs=False, mode="brute", skin=0.3):
        n = fr.n
        f = [np.zeros(n) for _ in range(3)]
        ev = np.zeros(2)
        m = {"brute": 0, "cells": 1, "verlet": 2}[mode]
        st = self.lib.orc_lj_forces(n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z), sig, eps,
                                    rc, int(lfs), m, skin, *[_ptr(v) for v in f], _ptr(ev))
        if st == -1:
            raise ValueError(self.error())
        if st != 0:
            raise RuntimeError(self.error())
        return np.stack(f, axis=1), float(ev[0]), float(ev[1])

    def lj_ev_compensated(self, fr: Frame, sig=1.0, eps=1.0, rc=2.5, lfs=False, mode="cells"):
        """(E, W, pairs inside rc) of the same pair terms as lj_forces, E and W
        summed with compensated summation (oracle library only)."""
        out = np.zeros(3)
        st = self.lib.orc_lj_ev_compensated(fr.n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z),
                                            sig, eps, rc, int(lfs), {"brute": 0, "cells": 1}[mode],
                                            _ptr(out))
        if st != 0:
            raise RuntimeError(self.error())
        return float(out[0]), float(out[1]), int(out[2])

    def lj_lrc(self, n: int, L, sig=1.0, eps=1.0, rc=2.5, lfs=False):
        out = np.zeros(2)
        self.lib.orc_lj_lrc(n, _ptr(np.asarray(L, dtype=float)), sig, eps, rc, int(lfs), _ptr(out))
        return float(out[0]), float(out[1])

    # ---- trajectory analysis -------------------------------------------------
    @staticmethod
    def _frames(L, xyz):
        """L: (frames, 3) or (3,); xyz: (frames, n, 3) -> packed (frames, 3, n)."""
        xyz = np.asarray(xyz, dtype=np.float64)
        nf, n = xyz.shape[0], xyz.shape[1]
        Lf = np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.float64), (nf, 3)))
        packed = np.ascontiguousarray(np.transpose(xyz, (0, 2, 1)))
        return nf, n, Lf, packed

    def rdf(self, L, xyz, r_max: float, bins: int):
        """(g, centres, hist) of Rdf(r_max, bins) accumulated over the frames
        (hist is all zero from the reference build, whose Rdf keeps it private)."""
        nf, n, Lf, packed = self._frames(L, xyz)
        g, c, h = np.zeros(bins), np.zeros(bins), np.zeros(bins)
        st = self.lib.orc_rdf(nf, n, _ptr(Lf), _ptr(packed), r_max, bins, _ptr(g), _ptr(c), _ptr(h))
        if st != 0:
            raise (ValueError if st == -1 else RuntimeError)(self.error())
        return g, c, h

    def msd(self, L, xyz, max_lag: int, dt=1.0, fit_from=0, fit_to=0):
        """(msd per lag, D) of Msd(n, max_lag) pushed once per frame."""
        nf, n, Lf, packed = self._frames(L, xyz)
        m, D = np.zeros(max_lag + 1), np.zeros(1)
        st = self.lib.orc_msd(nf, n, max_lag, _ptr(Lf), _ptr(packed), dt, fit_from, fit_to,
                              _ptr(m), _ptr(D))
        if st != 0:
            raise RuntimeError(self.error())
        return m, float(D[0])

    def vacf(self, vxyz, max_lag: int, dt=1.0):
        """(normalised vacf per lag, D) of Vacf(n, max_lag) pushed once per frame."""
        nf, n, _, packed = self._frames(np.ones(3), vxyz)
        c, D = np.zeros(max_lag + 1), np.zeros(1)
        st = self.lib.orc_vacf(nf, n, max_lag, _ptr(packed), dt, _ptr(c), _ptr(D))
        if st != 0:
            raise RuntimeError(self.error())
        return c, float(D[0])

    def cells_usable(self, L, cutoff: float) -> bool:
        return bool(self.lib.orc_cells_usable(_ptr(np.asarray(L, dtype=float)), cutoff))

    def verlet_pairs(self, fr: Frame, rc: float, skin: float) -> np.ndarray:
        """(P, 2) uint32 array of (i<j) pairs in the build order."""
        cap = max(1024, fr.n * 64)
        while True:
            out = np.zeros(2 * cap, dtype=np.uint32)
            np_ = self.lib.orc_verlet_pairs(fr.n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z),
                                            rc, skin, _ptr(out, _up), cap)
            if np_ >= 0:
                return out[: 2 * np_].reshape(-1, 2)
            cap = -np_ - 1

    def needs_rebuild(self, ref: Frame, cur: Frame, rc: float, skin: float) -> bool:
        return bool(self.lib.orc_needs_rebuild(ref.n, _ptr(ref.L), _ptr(ref.x), _ptr(ref.y),
                                               _ptr(ref.z), _ptr(cur.x), _ptr(cur.y), _ptr(cur.z),
                                               rc, skin))

    # ---- dynamics -----------------------------------------------------------
    def sim(self, fr: Frame, *, kind="vv", dt=0.005, sig=1.0, eps=1.0, rc=2.5, lfs=False,
            skin=-1.0, gamma=1.0, T=1.0, tau=0.5, M=3, dof=None, seed=1, mass=None):
        return Sim(self, fr, kind=kind, dt=dt, sig=sig, eps=eps, rc=rc, lfs=lfs, skin=skin,
                   gamma=gamma, T=T, tau=tau, M=M, dof=dof, seed=seed, mass=mass)

    # ---- electrostatics -----------------------------------------------------
    def coulomb_prefactor(self) -> float:
        return float(self.lib.orc_coulomb_prefactor())

    def ewald(self, fr: Frame, q, mol, rc=10.0, kmax=5, alpha=0.0, forces=False):
        n = fr.n
        q = np.ascontiguousarray(q, dtype=float)
        mol = np.ascontiguousarray(mol, dtype=np.int64)
        f = [np.zeros(n) for _ in range(3)]
        t = np.zeros(4)
        st = self.lib.orc_ewald(n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z), _ptr(q),
                                _ptr(mol, _lp), rc, kmax, alpha, int(forces),
                                *[_ptr(v) for v in f], _ptr(t))
        if st != 0:
            raise RuntimeError(self.error())
        return dict(real=t[0], recip=t[1], self=t[2], intra=t[3]), np.stack(f, axis=1)

    def pme(self, fr: Frame, q, kmax=5, alpha=0.0, grid=(32, 32, 32), order=6, forces=False):
        """Smooth PME reciprocal energy (and forces) over the Ewald k-set."""
        n = fr.n
        q = np.ascontiguousarray(q, dtype=float)
        K = (C.c_int * 3)(*[int(g) for g in grid])
        f = [np.zeros(n) for _ in range(3)]
        e = np.zeros(1)
        st = self.lib.orc_pme(n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z), _ptr(q), kmax,
                              alpha, K, order, int(forces), *[_ptr(v) for v in f], _ptr(e))
        if st != 0:
            raise RuntimeError(self.error())
        return float(e[0]), np.stack(f, axis=1)

    def pme_bsq(self, order, K, m):
        return float(self.lib.orc_pme_bsq(order, K, m))

    def spce_forces(self, fr: Frame, rc=10.0, kmax=5, alpha=0.0):
        n = fr.n
        f = [np.zeros(n) for _ in range(3)]
        pe = np.zeros(1)
        st = self.lib.orc_spce_forces(n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z), rc, kmax,
                                      alpha, *[_ptr(v) for v in f], _ptr(pe))
        if st != 0:
            raise RuntimeError(self.error())
        return np.stack(f, axis=1), float(pe[0])

    def rigid_shake(self, fr: Frame, m, ref: Frame, oh=1.0, angle=109.47, tol=1e-12, maxit=500):
        x, y, z = fr.x.copy(), fr.y.copy(), fr.z.copy()
        m = np.ascontiguousarray(m, dtype=float)
        it = self.lib.orc_rigid_shake(fr.n, _ptr(fr.L), _ptr(x), _ptr(y), _ptr(z), _ptr(m),
                                      _ptr(ref.x), _ptr(ref.y), _ptr(ref.z), oh, angle, tol, maxit)
        if it == -1:
            raise ValueError(self.error())
        if it < 0:
            raise RuntimeError(self.error())
        return it, np.stack([x, y, z], axis=1)

    def rigid_rattle(self, fr: Frame, m, oh=1.0, angle=109.47, tol=1e-12, maxit=500):
        vx, vy, vz = fr.vx.copy(), fr.vy.copy(), fr.vz.copy()
        m = np.ascontiguousarray(m, dtype=float)
        it = self.lib.orc_rigid_rattle(fr.n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z),
                                       _ptr(m), _ptr(vx), _ptr(vy), _ptr(vz), oh, angle, tol, maxit)
        if it == -1:
            raise ValueError(self.error())
        if it < 0:
            raise RuntimeError(self.error())
        return it, np.stack([vx, vy, vz], axis=1)

    def settle_positions(self, fr: Frame, m, ref: Frame, oh=1.0, angle=109.47):
        """Closed-form constraint of the positions against ``ref`` (oracle only:
        the reference has no SETTLE)."""
        x, y, z = fr.x.copy(), fr.y.copy(), fr.z.copy()
        m = np.ascontiguousarray(m, dtype=float)
        it = self.lib.orc_settle_positions(fr.n, _ptr(fr.L), _ptr(x), _ptr(y), _ptr(z), _ptr(m),
                                           _ptr(ref.x), _ptr(ref.y), _ptr(ref.z), oh, angle)
        if it == -1:
            raise ValueError(self.error())
        if it < 0:
            raise RuntimeError(self.error())
        return it, np.stack([x, y, z], axis=1)

    def settle_velocities(self, fr: Frame, m, oh=1.0, angle=109.47):
        vx, vy, vz = fr.vx.copy(), fr.vy.copy(), fr.vz.copy()
        m = np.ascontiguousarray(m, dtype=float)
        it = self.lib.orc_settle_velocities(fr.n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z),
                                            _ptr(m), _ptr(vx), _ptr(vy), _ptr(vz), oh, angle)
        if it == -1:
            raise ValueError(self.error())
        if it < 0:
            raise RuntimeError(self.error())
        return it, np.stack([vx, vy, vz], axis=1)

    def water_run(self, fr: Frame, rc=10.0, kmax=5, alpha=0.0, dt=0.00182367, nsteps=1,
                  settle=False):
        w = fr.copy()
        pe = np.zeros(nsteps)
        ke = np.zeros(nsteps)
        fn = self.lib.orc_water_run_settle if settle else self.lib.orc_water_run
        st = fn(w.n, _ptr(w.L), _ptr(w.x), _ptr(w.y), _ptr(w.z), _ptr(w.vx), _ptr(w.vy),
                _ptr(w.vz), rc, kmax, alpha, dt, nsteps, _ptr(pe), _ptr(ke))
        if st != 0:
            raise RuntimeError(self.error())
        return w, pe, ke

    def mix_lj(self, fr: Frame, types, sig, eps, rc, lfs=False, forces=False):
        n = fr.n
        nt = int(np.asarray(sig).shape[0])
        t = np.ascontiguousarray(types, dtype=np.int32)
        sg = np.ascontiguousarray(sig, dtype=float).reshape(-1)
        ep = np.ascontiguousarray(eps, dtype=float).reshape(-1)
        rcs = np.ascontiguousarray(rc, dtype=float).reshape(-1)
        f = [np.zeros(n) for _ in range(3)]
        out = np.zeros(2)
        tabs = [np.zeros(nt * nt) for _ in range(3)]
        self.lib.orc_mix_lj(n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z), _ptr(t, _ip), nt,
                            _ptr(sg), _ptr(ep), _ptr(rcs), int(lfs), int(forces),
                            *[_ptr(v) for v in f], _ptr(out), *[_ptr(v) for v in tabs])
        return dict(energy=float(out[0]), virial=float(out[1]),
                    e_ab=tabs[0].reshape(nt, nt), w_ab=tabs[1].reshape(nt, nt),
                    y_ab=tabs[2].reshape(nt, nt), forces=np.stack(f, axis=1))

    @staticmethod
    def _mix_tables(types, sig, eps, rc):
        t = np.ascontiguousarray(types, dtype=np.int32)
        tabs = [np.ascontiguousarray(a, dtype=float).reshape(-1) for a in (sig, eps, rc)]
        return t, int(np.asarray(sig).shape[0]), tabs

    def mix_minimize(self, fr: Frame, types, sig, eps, rc, nsteps, step, max_disp, lfs=False):
        """Steepest descent; returns (relaxed Frame, energies[nsteps+1])."""
        out = fr.copy()
        t, nt, (sg, ep, rcs) = self._mix_tables(types, sig, eps, rc)
        e = np.zeros(nsteps + 1)
        self.lib.orc_mix_minimize(out.n, _ptr(out.L), _ptr(out.x), _ptr(out.y), _ptr(out.z),
                                  _ptr(t, _ip), nt, _ptr(sg), _ptr(ep), _ptr(rcs), int(lfs),
                                  int(nsteps), step, max_disp, _ptr(e))
        return out, e

    def mix_vv(self, fr: Frame, types, sig, eps, rc, dt, nsteps, lfs=False, mass=None):
        """Velocity Verlet on the mixture; returns (final Frame, E[nsteps+1], W[nsteps+1])."""
        out = fr.copy()
        t, nt, (sg, ep, rcs) = self._mix_tables(types, sig, eps, rc)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=float)
        e, w = np.zeros(nsteps + 1), np.zeros(nsteps + 1)
        self.lib.orc_mix_vv(out.n, _ptr(out.L), _ptr(out.x), _ptr(out.y), _ptr(out.z),
                            _ptr(out.vx), _ptr(out.vy), _ptr(out.vz),
                            None if m is None else _ptr(m), _ptr(t, _ip), nt, _ptr(sg), _ptr(ep),
                            _ptr(rcs), int(lfs), dt, int(nsteps), _ptr(e), _ptr(w))
        return out, e, w


class Sim:
    """One integrator run on the oracle (or reference) library."""

    KINDS = {"vv": 0, "baoab": 1, "nhc": 2}

    def __init__(self, orc: Oracle, fr: Frame, *, kind, dt, sig, eps, rc, lfs, skin, gamma, T,
                 tau, M, dof, seed, mass):
        self.o = orc
        self.n = fr.n
        k = self.KINDS[kind]
        if dof is None:
            dof = 3 * fr.n - 3
        p1, p2 = (gamma, T) if k == 1 else (T, tau)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=float)
        vx = fr.vx if fr.vx is not None else np.zeros(fr.n)
        vy = fr.vy if fr.vy is not None else np.zeros(fr.n)
        vz = fr.vz if fr.vz is not None else np.zeros(fr.n)
        self.h = orc.lib.orc_sim_create(fr.n, _ptr(fr.L), _ptr(fr.x), _ptr(fr.y), _ptr(fr.z),
                                        _ptr(vx), _ptr(vy), _ptr(vz), _ptr(m), sig, eps, rc,
                                        int(lfs), skin, k, dt, p1, p2, M, dof, seed)
        if not self.h:
            raise ValueError(orc.error())

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.orc_sim_destroy(self.h)
            self.h = None

    def run(self, nsteps: int):
        pe, ke, w, hc = (np.zeros(nsteps) for _ in range(4))
        st = self.o.lib.orc_sim_run(self.h, nsteps, _ptr(pe), _ptr(ke), _ptr(w), _ptr(hc))
        if st != 0:
            raise RuntimeError(self.o.error())
        return dict(pe=pe, ke=ke, virial=w, conserved=hc)

    def state(self):
        a = [np.zeros(self.n) for _ in range(9)]
        self.o.lib.orc_sim_get(self.h, *[_ptr(v) for v in a])
        return (np.stack(a[0:3], axis=1), np.stack(a[3:6], axis=1), np.stack(a[6:9], axis=1))

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=float)
        cols = [np.ascontiguousarray(v[:, d]) for d in range(3)]
        self.o.lib.orc_sim_set_velocities(self.h, *[_ptr(c) for c in cols])

    @property
    def rebuilds(self) -> int:
        return int(self.o.lib.orc_sim_rebuilds(self.h))

    @property
    def conserved(self) -> float:
        return float(self.o.lib.orc_sim_conserved(self.h))


def load_nist(path: str) -> Frame:
    """NIST SRS configuration reader: box line, count line (not trusted as an
    atom count), then ``index x y z [type]`` lines to end of file, blank lines
    skipped; velocities zero (io/nist_config.cpp:30-86)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    if not lines:
        raise RuntimeError(f"empty file {path}")
    try:
        L = np.array([float(t) for t in lines[0].split()[:3]])
        int(lines[1].split()[0])
    except (ValueError, IndexError) as e:
        raise RuntimeError(f"malformed header in {path}") from e
    xs, ys, zs, types = [], [], [], []
    for ln in lines[2:]:
        parts = ln.split()
        if not parts:
            continue
        try:
            int(parts[0])
            xs.append(float(parts[1])); ys.append(float(parts[2])); zs.append(float(parts[3]))
        except (ValueError, IndexError) as e:
            raise RuntimeError(f"malformed atom line in {path}") from e
        types.append(parts[4] if len(parts) > 4 else "")
    n = len(xs)
    fr = Frame(L, np.array(xs), np.array(ys), np.array(zs), np.zeros(n), np.zeros(n), np.zeros(n))
    fr.types = types if any(types) else []
    return fr
