"""Differentiable Coulomb field of the tree solve for torch.autograd.

``coulomb_field(x, y, z, q, solver)`` returns the field ``E`` (n, 3) and the
potential ``pot`` (n) at every particle, in the input order, from one grow
and traverse on ``solver``.  Its backward runs the HIP adjoint
(``pepc_gpu_adjoint_device``): gradients with respect to the positions and
charges with the tree and the interaction lists of that solve held fixed.
The Barnes-Hut acceptance decisions are piecewise constant in the positions,
so this is the derivative of the solve wherever the decisions do not change
(everywhere at theta = 0).

The permutation between input order and the solver's key order is applied
with torch index operations; every arithmetic step is in libpepc_gpu.so.
"""
from __future__ import annotations

import torch


class _CoulombField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, q, solver):
        solver.grow(x.detach(), y.detach(), z.detach(), q.detach())
        res = solver.traverse()
        perm = torch.from_numpy(solver.sorted_host()[1]).to(x.device)
        n = perm.numel()
        e = torch.empty((n, 3), dtype=torch.float64, device=x.device)
        pot = torch.empty(n, dtype=torch.float64, device=x.device)
        e[perm] = res.e.T
        pot[perm] = res.pot
        solver._generation = getattr(solver, "_generation", 0) + 1
        ctx.solver = solver
        ctx.generation = solver._generation
        ctx.save_for_backward(perm, x.detach(), y.detach(), z.detach(), q.detach())
        return e, pot

    @staticmethod
    def backward(ctx, ge, gpot):
        perm, x, y, z, q = ctx.saved_tensors
        solver = ctx.solver
        if getattr(solver, "_generation", None) != ctx.generation:
            # the solver grew another tree since this forward: rebuild this
            # one (the build is deterministic, so tree and lists are the same)
            solver.grow(x, y, z, q)
            solver._generation += 1
            ctx.generation = solver._generation
        n = perm.numel()
        dev = perm.device
        if ge is None:
            ge = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        if gpot is None:
            gpot = torch.zeros(n, dtype=torch.float64, device=dev)
        ges = ge.to(torch.float64)[perm].T.contiguous()
        gps = gpot.to(torch.float64)[perm].contiguous()
        gx, gy, gz, gq = solver.adjoint(ges, gps)
        outs = []
        for g in (gx, gy, gz, gq):
            o = torch.empty_like(g)
            o[perm] = g
            outs.append(o)
        return outs[0], outs[1], outs[2], outs[3], None


def coulomb_field(x, y, z, q, solver):
    """Field (n, 3) and potential (n) at the particles, input order,
    differentiable in x, y, z and q (float64 device tensors)."""
    return _CoulombField.apply(x, y, z, q, solver)


def energy(q, pot):
    """Electrostatic energy 1/2 sum q phi."""
    return 0.5 * torch.sum(q * pot)


__all__ = ["coulomb_field", "energy"]
