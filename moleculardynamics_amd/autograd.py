"""torch.autograd.Function wrappers over the differentiable per-type-pair LJ
path (md_mix_* in include/moldyn_hip.h).

The arithmetic runs in the HIP library: forward passes call md_mix_forces /
md_mix_run, backward passes call md_mix_force_vjp / md_mix_adjoint. torch
only carries the tensors. Parameters are the per-type-pair epsilon and sigma
in the canonical order of unordered type pairs (a <= b, row-major); cutoffs
are held fixed (the derivatives are taken at fixed cutoffs).

    system = MixSystem(L, types, cutoff_table, lfs=True)
    E, W = MixEnergy.apply(eps, sigma, pos, system)
    F = MixForces.apply(eps, sigma, pos, system)
    Emean, Wmean, xT, vT = MixTrajectory.apply(eps, sigma, x0, v0, system, dt, nsteps, K)
"""
from __future__ import annotations

import numpy as np
import torch

from . import native


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).numpy()


def _like(a, ref: torch.Tensor) -> torch.Tensor:
    return torch.as_tensor(np.asarray(a, dtype=np.float64), dtype=torch.float64, device=ref.device)


def symmetric_table(values, ntypes: int) -> np.ndarray:
    """Canonical per-pair values (a <= b, row-major) -> symmetric (nt, nt) table."""
    t = np.zeros((ntypes, ntypes))
    k = 0
    for a in range(ntypes):
        for b in range(a, ntypes):
            t[a, b] = t[b, a] = values[k]
            k += 1
    return t


class MixSystem:
    """One md_mix handle with the fixed part of the model: box, types,
    masses, cutoff table, force shift and list skin."""

    def __init__(self, L, types, cutoff, lfs: bool = False, mass=None, skin: float = 0.3):
        self.L = np.asarray(L, dtype=np.float64).reshape(3)
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.cutoff = np.asarray(cutoff, dtype=np.float64)
        self.ntypes = int(self.cutoff.shape[0])
        self.lfs = bool(lfs)
        self.mass = None if mass is None else np.asarray(mass, dtype=np.float64)
        self.skin = skin
        self.mix: native.Mix | None = None
        self.run_id = 0

    def configure(self, eps: np.ndarray, sigma: np.ndarray, pos: np.ndarray, vel=None) -> None:
        # any reconfiguration replaces the state the last run's checkpoints
        # describe, so a pending trajectory adjoint must run its forward again
        self.run_id += 1
        e = symmetric_table(eps, self.ntypes)
        s = symmetric_table(sigma, self.ntypes)
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        if self.mix is None or self.mix.n != pos.shape[0]:
            if self.mix is not None:
                self.mix.close()
            self.mix = native.Mix(self.L, pos, self.types, s, e, self.cutoff, lfs=self.lfs,
                                  vel=vel, mass=self.mass, skin=self.skin)
        else:
            self.mix.set_params(s, e, self.cutoff, self.lfs)
            self.mix.set_state(pos, vel if vel is not None else np.zeros_like(pos))

    def close(self) -> None:
        if self.mix is not None:
            self.mix.close()
            self.mix = None


class MixEnergy(torch.autograd.Function):
    """(E, W) of a configuration; gradients on eps, sigma and positions."""

    @staticmethod
    def forward(ctx, eps, sigma, pos, system: MixSystem):
        system.configure(_np(eps), _np(sigma), _np(pos))
        e, w = system.mix.forces()
        ctx.system = system
        ctx.save_for_backward(eps, sigma, pos)
        return _like(e, pos), _like(w, pos)

    @staticmethod
    def backward(ctx, ge, gw):
        eps, sigma, pos = ctx.saved_tensors
        system = ctx.system
        system.configure(_np(eps), _np(sigma), _np(pos))
        o, th = system.mix.force_vjp(None, float(ge), float(gw))
        return _like(th[:, 0], eps), _like(th[:, 1], sigma), _like(o, pos), None


class MixForces(torch.autograd.Function):
    """Forces of a configuration; the backward pass is the force VJP."""

    @staticmethod
    def forward(ctx, eps, sigma, pos, system: MixSystem):
        system.configure(_np(eps), _np(sigma), _np(pos))
        system.mix.forces()
        _, _, f = system.mix.download()
        ctx.system = system
        ctx.save_for_backward(eps, sigma, pos)
        return _like(f, pos)

    @staticmethod
    def backward(ctx, gf):
        eps, sigma, pos = ctx.saved_tensors
        system = ctx.system
        system.configure(_np(eps), _np(sigma), _np(pos))
        o, th = system.mix.force_vjp(_np(gf))
        return _like(th[:, 0], eps), _like(th[:, 1], sigma), _like(o, pos), None


class MixTrajectory(torch.autograd.Function):
    """nsteps of velocity Verlet from (x0, v0): returns the frame means
    <E>, <W> over t = 0..nsteps and the final (x_T, v_T). The backward pass
    is the adjoint with checkpoints every checkpoint_every steps."""

    @staticmethod
    def forward(ctx, eps, sigma, x0, v0, system: MixSystem, dt: float, nsteps: int,
                checkpoint_every: int = 20):
        system.configure(_np(eps), _np(sigma), _np(x0), _np(v0))
        e, w = system.mix.run(dt, nsteps, checkpoint_every)
        xt, vt, _ = system.mix.download()
        system.run_id += 1
        ctx.system = system
        ctx.run_id = system.run_id
        ctx.args = (dt, nsteps, checkpoint_every)
        ctx.save_for_backward(eps, sigma, x0, v0)
        return _like(e.mean(), x0), _like(w.mean(), x0), _like(xt, x0), _like(vt, v0)

    @staticmethod
    def backward(ctx, ge, gw, gx, gv):
        eps, sigma, x0, v0 = ctx.saved_tensors
        system = ctx.system
        dt, nsteps, k = ctx.args
        if system.run_id != ctx.run_id:  # another run replaced the checkpoints: redo this one
            system.configure(_np(eps), _np(sigma), _np(x0), _np(v0))
            system.mix.run(dt, nsteps, k)
            system.run_id += 1
            ctx.run_id = system.run_id
        frames = nsteps + 1
        gxn = None if gx is None else _np(gx)
        gvn = None if gv is None else _np(gv)
        th, dx0, dv0 = system.mix.adjoint(float(ge) / frames, float(gw) / frames, gxn, gvn)
        return (_like(th[:, 0], eps), _like(th[:, 1], sigma), _like(dx0, x0), _like(dv0, v0),
                None, None, None, None)
