"""lnZ(beta) of the FTLM estimator as a differentiable torch function of the Hamiltonian parameters.

theta = (t, t', U, V, phi_x, phi_y, mu) (a float64 tensor of 7 entries) -> lnZ(beta_b) of the sector
(n_up, n_dn), or of its momentum-K block, for H - mu N with the converged Gauss rule
(ftlm_log_partition_complex, src/ftlm_thermo.cpp:213-256).  The forward pass runs the device Lanczos
recurrence once and keeps a tape (checkpointed Krylov vectors, scalars, quadrature derivatives);
backward(grad_out) runs the adjoint sweep for the weights grad_out, with the derivative operators of
t, t', U, V, phi_x, phi_y as one term-split pass over a sector's hop lists.  d/dmu is beta N.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from . import api

THETA = ("t", "tp", "U", "V", "phi_x", "phi_y", "mu")


@dataclass(frozen=True)
class FtlmModel:
    Lx: int
    Ly: int
    n_up: int
    n_dn: int
    K: api.MomentumSector | None = None


def _operators(model: FtlmModel, th: np.ndarray):
    p = api.HubbardParams(model.Lx, model.Ly, float(th[0]), float(th[1]), float(th[2]), float(th[3]), float(th[4]),
                          float(th[5]))
    if model.K is None:
        op = api.HubbardOperator.from_params(p, model.n_up, model.n_dn)
    else:
        op = api.KBlockOperator.from_params(p, model.n_up, model.n_dn, model.K)
    dops = [api.derivative_operator(p, model.n_up, model.n_dn, q, model.K) for q in THETA[:6]]
    return op, dops


class FtlmLogPartition(torch.autograd.Function):
    """lnZ(betas) for theta = (t, t', U, V, phi_x, phi_y, mu); see the module docstring."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, model: FtlmModel, betas, par: api.FtlmParams, checkpoint: int = 0,
                mode: int = L.FTLM_TS_AUTO):
        if theta.shape != (len(THETA),):
            raise ValueError(f"theta must hold {len(THETA)} entries {THETA}")
        th = theta.detach().to("cpu", torch.float64).numpy()
        b = np.ascontiguousarray(np.atleast_1d(np.asarray(betas, np.float64)))
        op, dops = _operators(model, th)
        tape = api.FtlmTape(op, dops, b, par, checkpoint, mode)
        n = model.n_up + model.n_dn
        ctx.tape, ctx.betas, ctx.n = tape, b, n
        ctx.theta_meta = (theta.dtype, theta.device)
        return torch.as_tensor(tape.lnZ + b * th[6] * n, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        w = grad_out.detach().to("cpu", torch.float64).numpy()
        g = ctx.tape.backward(w)
        gmu = float(np.dot(w, ctx.betas) * ctx.n)
        dtype, device = ctx.theta_meta
        grad = torch.tensor(list(g) + [gmu], dtype=torch.float64).to(device=device, dtype=dtype)
        return grad, None, None, None, None, None


def ftlm_log_partition(theta: torch.Tensor, model: FtlmModel, betas, par: api.FtlmParams, checkpoint: int = 0,
                       mode: int = L.FTLM_TS_AUTO) -> torch.Tensor:
    """Differentiable lnZ(betas) of H(theta) - mu N on the model's sector or momentum block."""
    return FtlmLogPartition.apply(theta, model, betas, par, checkpoint, mode)
