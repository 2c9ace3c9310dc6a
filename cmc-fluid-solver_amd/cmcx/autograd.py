"""torch.autograd binding of the reverse mode over the C ABI.

    final = Rollout.apply(node_values, params, state0, solver, steps, dt, num_global, every)

runs `steps` ADI steps (UpdateBoundaries + TimeStep, num_local = 1) of a built
:class:`cmcx.solver.AdiSolver3D` from `state0` with the given node values
(BOUND/VALVE data, shape (4, N): u, v, w, T) and FluidParams (v_T, v_vis,
t_vis, t_phi), and returns the final state (4, N).  Its backward is the HIP
adjoint (cmcx_adjoint_backward): checkpoints every `every` steps, each interval
recomputed once, transposed line solves in reverse sweep order.

Tensors cross the boundary through host memory, so torch never touches the GPU
here (the solver owns the device state); gradients come back in float64.  A
solver holds one recording: the backward of a rollout raises if another
rollout was recorded on the same solver since, or its parameters or node
values were changed (use one solver per live rollout).
:func:`channel_rollout` builds the inlet-control problem on top of it: node
values from a coarse inlet profile (each control value fills a block of inlet
nodes) and the kinematic viscosity, both differentiable.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .params import FluidParams

_lib.lib()  # libcmcx binds its HIP runtime before torch is imported

import torch  # noqa: E402


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, node_values, params, state0, solver, steps, dt, num_global, every):
        g = solver.grid
        nv = node_values.detach().to("cpu", torch.float64).numpy()
        solver.set_boundary(g.bc_vel, g.bc_temp, nv[0], nv[1], nv[2], nv[3])
        p = params.detach().to("cpu", torch.float64).numpy()
        solver.set_params(FluidParams(*[float(x) for x in p]))
        s0 = state0.detach().to("cpu").numpy().astype(solver.np_dtype)
        solver.set_fields(s0[0], s0[1], s0[2], s0[3])
        solver.adjoint_forward(int(steps), float(dt), int(num_global), int(every))
        u, v, w, _, T = solver.fields()
        ctx.solver = solver
        ctx.generation = solver.adjoint_generation()
        ctx.devices = (node_values.device, params.device, state0.device)
        return torch.from_numpy(np.stack([u, v, w, T]).astype(np.float64)).to(node_values.device)

    @staticmethod
    def backward(ctx, gout):
        # the solver keeps one recording: another rollout recorded on it since,
        # or new parameters / node values, would give another rollout's gradient
        if ctx.solver.adjoint_generation() != ctx.generation:
            raise RuntimeError("Rollout.backward: the solver's recording is no longer this "
                               "rollout's (another Rollout.apply, set_params or set_boundary "
                               "on the same solver since); use one solver per live rollout")
        seed = gout.detach().to("cpu", torch.float64).numpy()
        gb, sb, pb = ctx.solver.adjoint_backward(seed)
        dn, dp, ds = ctx.devices
        return (torch.from_numpy(gb.astype(np.float64)).to(dn),
                torch.from_numpy(pb.astype(np.float64)).to(dp),
                torch.from_numpy(sb.astype(np.float64)).to(ds),
                None, None, None, None, None)


def inlet_nodes(dom) -> np.ndarray:
    """Node ids of the x = 0 face, (ny, nz) order."""
    ny, nz = dom.ny, dom.nz
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    return (j * nz + k).ravel()


def channel_rollout(solver, dom, control, nu, steps, every, block: int = 3, state0=None):
    """Final state of the inlet-control problem as a differentiable function of
    `control` (torch, (ny/block, nz/block)) and `nu` (torch scalar, v_vis).
    The inlet u of node (0, j, k) is control[j // block, k // block]; t_vis,
    t_phi and v_T stay at the domain's values."""
    base = torch.tensor(np.stack([dom.vx, dom.vy, dom.vz, dom.T]), dtype=torch.float64)
    prof = torch.kron(control.to(torch.float64), torch.ones(block, block, dtype=torch.float64))
    prof = prof[:dom.ny, :dom.nz].reshape(-1)
    ids = torch.from_numpy(inlet_nodes(dom))
    u = base[0].index_put((ids,), prof)
    node_values = torch.stack([u, base[1], base[2], base[3]])
    p = dom.params()
    params = torch.stack([torch.tensor(p.v_T, dtype=torch.float64), nu.to(torch.float64).reshape(()),
                          torch.tensor(p.t_vis, dtype=torch.float64),
                          torch.tensor(p.t_phi, dtype=torch.float64)])
    if state0 is None:
        state0 = torch.tensor(np.stack([dom.vx, dom.vy, dom.vz, dom.T]), dtype=torch.float64)
    return Rollout.apply(node_values, params, state0, solver, steps, dom.dt, dom.num_global, every)
