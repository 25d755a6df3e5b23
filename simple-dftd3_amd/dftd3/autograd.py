"""Differentiable D3 energies of a batch on PyTorch tensors.

``dispersion_energy(spec, positions, lattice)`` returns the energy of every
system of a batch as a tensor that autograd differentiates with respect to
the positions (N, 3) and the lattices (nsys, 3, 3), both in bohr:

* the forward pass runs the energy-only kernels of the library (no gradient
  work, nothing is kept but the inputs);
* the backward pass runs the energy / gradient / virial kernels on the saved
  inputs (the coordination-number derivatives are recomputed pair by pair,
  no 3 x N x N dCN/dR array exists) and turns gradient G and virial sigma
  into the adjoints of the inputs.

Lattice adjoint: with the Cartesian positions X (rows x_i) and the lattice L
(rows = lattice vectors) as independent inputs, the strain derivative of the
energy (the reference's virial, disp.f90:56-185 with sigma = dE/d(epsilon)
for x -> (1+epsilon) x, L_k -> (1+epsilon) L_k) is

    sigma = G^T X + (dE/dL)^T L    =>    dE/dL = L^{-T} (sigma - G^T X)^T.

Inputs in float32 run the mixed-precision kernels (FP32 arithmetic, FP64
accumulation) by default, float64 inputs the FP64 kernels; adjoints are
returned in the dtype of the inputs.  Positions are used as given (not
wrapped into the cell).
"""
from typing import Optional

import numpy as np
import torch

from . import torch_api


def _spec_with_lattice(spec: torch_api.BatchSpec, lattice: Optional[torch.Tensor]) -> torch_api.BatchSpec:
    if lattice is None:
        return spec
    lat = lattice.detach().to(torch.float64).cpu().numpy().reshape(len(spec), 3, 3)
    periodic = spec.periodic if spec.periodic is not None else np.ones((len(spec), 3), dtype=bool)
    return torch_api.BatchSpec(spec.natoms, spec.numbers, lat, periodic, spec.params, spec.cutoffs)


def lattice_adjoint(gradient: torch.Tensor, sigma: torch.Tensor, positions: torch.Tensor, lattice: torch.Tensor,
                    owner: torch.Tensor) -> torch.Tensor:
    """dE/dL (nsys, 3, 3) from the gradient G (N, 3), the virial sigma
    (nsys, 3, 3) as the library returns it (sigma(a, b) at [b][a]), the
    positions X (N, 3), the lattices L (nsys, 3, 3, rows = vectors) and the
    system index of every atom: dE/dL = L^{-T} (sigma - G^T X)^T."""
    sig = sigma.transpose(1, 2)
    sig = 0.5 * (sig + sig.transpose(1, 2))
    gx = torch.zeros_like(sig)
    gx.index_add_(0, owner, gradient.unsqueeze(2) * positions.unsqueeze(1))  # sum_i G_ia x_ib
    return torch.linalg.solve(lattice.transpose(1, 2), (sig - gx).transpose(1, 2))


class _Dispersion(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, lattice, spec, precision):
        pos64 = positions.detach().to(torch.float64).contiguous()
        run_spec = _spec_with_lattice(spec, lattice)
        out = torch_api.run_batch(run_spec, pos64, precision=precision, grad=False)
        ctx.spec = run_spec
        ctx.precision = precision
        ctx.has_lattice = lattice is not None
        ctx.save_for_backward(positions, lattice if lattice is not None else positions.new_empty(0))
        return out["energy"].to(positions.dtype)

    @staticmethod
    def backward(ctx, grad_energy):
        positions, lattice = ctx.saved_tensors
        pos64 = positions.detach().to(torch.float64).contiguous()
        out = torch_api.run_batch(ctx.spec, pos64, precision=ctx.precision, grad=True)
        natoms = torch.as_tensor(ctx.spec.natoms, dtype=torch.long, device=positions.device)
        owner = torch.repeat_interleave(torch.arange(len(ctx.spec), device=positions.device), natoms)
        ge = grad_energy.to(torch.float64)
        g = out["gradient"]
        grad_pos = (g * ge[owner].unsqueeze(1)).to(positions.dtype) if ctx.needs_input_grad[0] else None
        grad_lat = None
        if ctx.has_lattice and ctx.needs_input_grad[1]:
            d = lattice_adjoint(g, out["sigma"], pos64, lattice.detach().to(torch.float64), owner)
            grad_lat = (d * ge.view(-1, 1, 1)).to(lattice.dtype)
        return grad_pos, grad_lat, None, None


def dispersion_energy(spec: torch_api.BatchSpec, positions: torch.Tensor, lattice: Optional[torch.Tensor] = None,
                      precision: Optional[str] = None) -> torch.Tensor:
    """Energies (nsys,) of a batch, differentiable in positions and lattice.

    ``spec`` is the host description of the batch (torch_api.batch_spec);
    ``positions`` (N, 3) and ``lattice`` (nsys, 3, 3, optional, replacing the
    lattices of ``spec``) are tensors on the GPU in bohr.  ``precision`` is
    'fp64' or 'mixed' (default: 'mixed' for float32 inputs, else 'fp64').
    """
    if positions.dim() != 2 or positions.shape[1] != 3 or positions.shape[0] != int(np.sum(spec.natoms)):
        raise ValueError("positions must have shape (sum of natoms, 3)")
    if lattice is not None and tuple(lattice.shape) != (len(spec), 3, 3):
        raise ValueError("lattice must have shape (nsys, 3, 3)")
    if precision is None:
        precision = "mixed" if positions.dtype == torch.float32 else "fp64"
    return _Dispersion.apply(positions, lattice, spec, precision)


def forces_and_stress(spec: torch_api.BatchSpec, positions: torch.Tensor, lattice: torch.Tensor,
                      precision: Optional[str] = None):
    """Energies, forces -dE/dx (N, 3) and the stress sigma / V (nsys, 3, 3) of
    periodic systems through the autograd path (sigma = dE/d epsilon from the
    input adjoints: G^T X + (dE/dL)^T L)."""
    pos = positions.detach().requires_grad_(True)
    lat = lattice.detach().requires_grad_(True)
    energy = dispersion_energy(spec, pos, lat, precision)
    gpos, glat = torch.autograd.grad(energy.sum(), (pos, lat))
    natoms = torch.as_tensor(spec.natoms, dtype=torch.long, device=positions.device)
    owner = torch.repeat_interleave(torch.arange(len(spec), device=positions.device), natoms)
    p64, l64 = pos.detach().double(), lat.detach().double()
    sig = torch.zeros(len(spec), 3, 3, dtype=torch.float64, device=positions.device)
    sig.index_add_(0, owner, gpos.double().unsqueeze(2) * p64.unsqueeze(1))
    sig += glat.double().transpose(1, 2) @ l64
    vol = torch.linalg.det(l64).abs().view(-1, 1, 1)
    return energy.detach(), -gpos, (sig / vol).to(positions.dtype)
