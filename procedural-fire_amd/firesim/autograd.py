"""A torch.autograd.Function over the step's adjoint (fs_adjoint_* in libfiresim.so).

    vel, S, S2 = rollout(adj, dt, nsteps, spot_flux, face_flux=None, state=None)

runs `nsteps` steps of the adjoint's FireSim with the given pyrolysis-spot fluxes (kg/s,
one per spot, in the order set by FireSim.set_spots), optionally a floor fuel flux
(kg/m^2/s, shape (nz, nx)) and optionally a starting state (vel, S, S2 of shape
(nz, ny, nx, 4)); it returns the final fields as float32 tensors on spot_flux's device.
Any scalar built from them (a temperature/soot mismatch, a heat-release target) can be
back-propagated: the backward seeds the adjoint with the incoming gradients and runs the
reverse sweep on the GPU, returning dL/d(spot_flux), dL/d(face_flux) and dL/d(state).

The library owns the simulation's memory; tensors cross the boundary through host
buffers (the C ABI takes host pointers).  One Adjoint holds one recorded run: call
backward before the next rollout on the same Adjoint.
"""
import numpy as np
import torch

from . import Adjoint


def _np(t):
    return t.detach().to("cpu", torch.float32).contiguous().numpy()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, adj, dt, nsteps, solve_tol, spot_flux, face_flux, vel0, S0, S20):
        sim = adj.sim
        rows = sim.spots()[:, :6].astype(np.float32)
        flux = _np(spot_flux).reshape(-1)
        if flux.size != len(rows):
            raise ValueError(f"spot_flux has {flux.size} entries, the sim has {len(rows)} spots")
        rows[:, 4] = flux
        sim.set_spots(rows)
        if face_flux is not None:
            sim.set_face_flux(_np(face_flux))
        if vel0 is not None:
            sim.write("vel", _np(vel0))
            sim.write("S", _np(S0))
            sim.write("S2", _np(S20))
        adj.forward(dt, nsteps)
        ctx.adj = adj
        ctx.solve_tol = solve_tol
        ctx.n_spots = len(rows)
        ctx.face = face_flux is not None
        ctx.state = vel0 is not None
        ctx.dtype, ctx.device = spot_flux.dtype, spot_flux.device
        dev = spot_flux.device
        return tuple(torch.from_numpy(sim.read(f)).to(dev) for f in ("vel", "S", "S2"))

    @staticmethod
    def backward(ctx, g_vel, g_S, g_S2):
        adj = ctx.adj
        for name, g in (("vel", g_vel), ("S", g_S), ("S2", g_S2)):
            if g is not None:
                adj.seed(name, _np(g))
        adj.backward(solve_tol=ctx.solve_tol)
        gs, gf = adj.grads(ctx.n_spots, face=ctx.face)
        d_flux = torch.from_numpy(gs).to(ctx.device, ctx.dtype)
        d_face = torch.from_numpy(gf).to(ctx.device, ctx.dtype) if ctx.face else None
        d_state = [None, None, None]
        if ctx.state:
            d_state = [torch.from_numpy(adj.read(f)).to(ctx.device) for f in ("vel", "S", "S2")]
        return (None, None, None, None, d_flux, d_face, *d_state)


def rollout(adj: Adjoint, dt, nsteps, spot_flux, face_flux=None, state=None, solve_tol=0.0):
    """Final (vel, S, S2) after `nsteps` steps, differentiable in spot_flux, face_flux and state."""
    vel0, S0, S20 = state if state is not None else (None, None, None)
    return _Rollout.apply(adj, float(dt), int(nsteps), float(solve_tol), spot_flux, face_flux, vel0, S0, S20)


def mismatch(S, target_theta, target_soot):
    """The mismatch loss of the adjoint workload: sum (theta - theta*)^2 + (soot - soot*)^2."""
    return ((S[..., 0] - target_theta) ** 2).sum() + ((S[..., 2] - target_soot) ** 2).sum()
