"""PyTorch binding of the adjoint: gap-field enhancement as a differentiable function of the
design-region permittivity.

    J = GapFieldObjective.apply(eps, scn, steps, checkpoint_every, gap)
    J.backward()          # eps.grad = dJ/d eps from the discrete adjoint (sjf_adjoint_run)

    J = BatchedGapFieldObjective.apply(eps_stack, scns, steps, gaps)   # one J per variant
    J.sum().backward()    # eps_stack.grad from one batched adjoint run (sjf_batch_adjoint_run)

`eps` is a tensor shaped like the scenario's design box (any device; values move through host
memory into the engine).  The forward runs the whole simulation once on the GPU with state
checkpoints and the backward pass of the adjoint in the same call; autograd's backward only
scales the stored gradient.
"""
from __future__ import annotations

import numpy as np

from .engine import Batch, Engine
from .scenario import Scenario


def run_adjoint(scn: Scenario, eps: np.ndarray, steps: int, checkpoint_every: int, gap,
                precision: str | None = None):
    """(J, dJ/d eps) for the scenario with design permittivities `eps`."""
    s = Scenario(**{**scn.__dict__})
    lo, hi = scn.design[0], scn.design[1]
    s.design = (lo, hi, np.asarray(eps, dtype=np.float64))
    eng = s.make(Engine, precision=precision or scn.precision)
    try:
        J, g = eng.adjoint_run(steps, checkpoint_every, gap, int(np.prod(np.asarray(hi) - np.asarray(lo))))
    finally:
        eng.close()
    return J, g.reshape(np.asarray(eps).shape)


def run_adjoint_batch(scns, eps, steps: int, gaps, precision: str | None = None, checkpoint_every: int = 0):
    """(J per scenario, list of dJ/d eps) for scenario i with design permittivities eps[i], all
    in one batched adjoint run (the scenarios share nz and their materials' pole count)."""
    engs = []
    try:
        for s0, e in zip(scns, eps):
            s = Scenario(**{**s0.__dict__})
            s.design = (s0.design[0], s0.design[1], np.asarray(e, dtype=np.float64))
            engs.append(s.make(Engine, precision=precision or s0.precision))
        b = Batch(engs)
        try:
            J, g = b.adjoint_run(steps, gaps, [np.asarray(e).size for e in eps], checkpoint_every)
        finally:
            b.close()
    finally:
        for e in engs:
            e.close()
    return J, [gi.reshape(np.asarray(e).shape) for gi, e in zip(g, eps)]


try:
    import torch

    class GapFieldObjective(torch.autograd.Function):
        """J = dt * sum_n |E^n(gap)|^2 as a function of the design permittivity tensor."""

        @staticmethod
        def forward(ctx, eps, scn, steps, checkpoint_every, gap):
            J, g = run_adjoint(scn, eps.detach().cpu().double().numpy(), steps, checkpoint_every, gap)
            ctx.save_for_backward(torch.as_tensor(g, dtype=eps.dtype, device=eps.device))
            return torch.tensor(J, dtype=torch.float64, device=eps.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out.to(g.dtype) * g, None, None, None, None

    class BatchedGapFieldObjective(torch.autograd.Function):
        """J_i = dt * sum_n |E^n(gap_i)|^2 of design variant i, for a stack of design
        permittivity tensors eps[i] (all of one shape), from one batched adjoint run."""

        @staticmethod
        def forward(ctx, eps, scns, steps, gaps):
            J, g = run_adjoint_batch(scns, list(eps.detach().cpu().double().numpy()), steps, gaps)
            ctx.save_for_backward(torch.as_tensor(np.stack(g), dtype=eps.dtype, device=eps.device))
            return torch.as_tensor(J, dtype=torch.float64, device=eps.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            w = grad_out.to(g.dtype).reshape((-1,) + (1,) * (g.dim() - 1))
            return w * g, None, None, None

except ImportError:  # torch is optional for the engine itself
    GapFieldObjective = None
    BatchedGapFieldObjective = None
