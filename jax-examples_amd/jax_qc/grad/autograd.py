"""SCF energy as a differentiable function of the nuclear coordinates.

forward: integrals + SCF on the GPU at the given coordinates;
backward: the analytic gradient kernels (derivative ERIs and one-electron
derivatives contracted with D and W on the device), valid because the SCF
energy is stationary in the orbitals.
"""

from __future__ import annotations

import torch

from jax_qc.basis import build_basis_set
from jax_qc.core.types import CalcConfig
from jax_qc.grad.analytic import scf_gradient
from jax_qc.grad.numerical_grad import displaced
from jax_qc.integrals import compute_integrals
from jax_qc.scf import run_scf


class SCFEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, mol, basis_name, config, opts):
        m = displaced(mol, coords.detach().cpu().double().numpy())
        bs = build_basis_set(m, basis_name)
        ints = compute_integrals(m, bs, eri=opts.get("eri", "store"), tau=opts.get("tau", 1e-12),
                                 tau32=opts.get("tau32"))
        res = run_scf(m, ints, config)
        ctx.state = (m, bs, res, opts.get("grad_tau", 1e-13))
        ctx.dev = coords.device
        ctx.dtype = coords.dtype
        return torch.tensor(res.energy, dtype=coords.dtype, device=coords.device)

    @staticmethod
    def backward(ctx, g_out):
        m, bs, res, gtau = ctx.state
        g = scf_gradient(m, bs, res, gtau).to(device=ctx.dev, dtype=ctx.dtype)
        return g_out * g, None, None, None, None


def scf_energy(coords: torch.Tensor, mol, basis_name="sto-3g", config: CalcConfig = CalcConfig(),
               **opts) -> torch.Tensor:
    """E(coords) with a hand-written backward (see module docstring)."""
    return SCFEnergy.apply(coords, mol, basis_name, config, opts)
