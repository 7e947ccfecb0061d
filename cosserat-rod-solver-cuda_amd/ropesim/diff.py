"""Differentiation form: rollouts of a RopeBatch behind torch.autograd.

    x = ropesim.diff.rollout(batch, ea, ei, mass, drag, p0, v0, p1, dt, substeps)   # [F, num_vertices, 3]

The rollout starts from the batch's construction state (every environment is reset). It drives the roots
with the root program (p0, v0, p1 [F, num_ropes, 3], one row per frame) and returns the positions after
each frame. Gradients of any loss on those positions flow back to:
- per-vertex EA, EI (per segment, the entry at a rope's last vertex unused) and node masses (pinned
  entries unused), each [num_vertices] float32 on the GPU;
- the linear drag (a scalar);
- the root program p0, v0, p1.

The forward records one checkpoint per frame. The backward is the hand-written adjoint of the substep
(RopeBatch.adjoint). It recomputes each frame's substeps from its checkpoint, so memory grows with
the number of frames, not substeps. The active set of the forward is frozen: contact on/off, friction
branch, orientation signs, broad-phase lists, rope-rope candidate lists (the barrier, its friction and the
step limit are differentiated with the lists fixed). Not covered: quadratic drag; recording refuses it, and
rope contact groups above 512 vertices. The reference has no adjoint; this module has no counterpart in it.
"""
from __future__ import annotations

import numpy as np
import torch


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, dt, substeps, ea, ei, mass, drag, p0, v0, p1):
        F = p0.shape[0]
        nv = batch.num_vertices
        dev = p0.device
        batch.reset_envs(np.arange(batch.num_envs, dtype=np.int32))
        f32 = lambda t: t.detach().to(device=dev, dtype=torch.float32).contiguous()
        batch.set_material(f32(ea), f32(ei), f32(mass))
        batch.drag = float(drag)
        P0, V0, P1 = f32(p0), f32(v0), f32(p1)
        out = torch.empty(F, nv, 3, device=dev, dtype=torch.float32)
        batch.tape_begin(F)
        try:
            for f in range(F):
                batch.step(P0[f], V0[f], dt, substeps, root_pos_end=P1[f])
                out[f].copy_(torch.as_tensor(batch.positions(), device=dev).reshape(nv, 3))
        finally:
            batch.tape_end()
        ctx.batch = batch
        ctx.shapes = (F, batch.num_envs, nv, dev, drag.dtype if torch.is_tensor(drag) else torch.float32)
        return out

    @staticmethod
    def backward(ctx, g):
        F, N, nv, dev, ddt = ctx.shapes
        b = ctx.batch
        g = g.detach().to(dtype=torch.float32).contiguous()
        gea = torch.empty(nv, device=dev)
        gei = torch.empty(nv, device=dev)
        gm = torch.empty(nv, device=dev)
        gp0 = torch.empty(F, N, 3, device=dev)
        gv0 = torch.empty(F, N, 3, device=dev)
        gp1 = torch.empty(F, N, 3, device=dev)
        gdrag, mismatch = b.adjoint(g, gea, gei, gm, gp0, gv0, gp1)
        if mismatch:
            raise RuntimeError(f"adjoint: {mismatch} vertices recomputed differently from the recorded rollout "
                               "(material or meshes changed after the forward?)")
        return (None, None, None, gea, gei, gm, torch.tensor(gdrag, dtype=ddt, device=dev), gp0, gv0, gp1)


def rollout(batch, ea, ei, mass, drag, p0, v0, p1, dt, substeps):
    """Positions after each frame [F, num_vertices, 3] of a root-program rollout from the construction state;
    differentiable with respect to ea, ei, mass, drag, p0, v0 and p1."""
    if not torch.is_tensor(drag):
        drag = torch.tensor(float(drag), device=p0.device)
    return _Rollout.apply(batch, float(dt), int(substeps), ea, ei, mass, drag, p0, v0, p1)
