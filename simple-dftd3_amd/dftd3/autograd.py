"""Differentiable D3 energies of a batch on PyTorch tensors.

``dispersion_energy(spec, positions, lattice)`` returns the energy of every
system of a batch as a tensor that autograd differentiates with respect to
the positions (N, 3) and the lattices (nsys, 3, 3), both in bohr:

* when an input requires a gradient, the forward pass runs the energy /
  gradient / virial kernels once (the coordination-number derivatives are
  recomputed pair by pair, no 3 x N x N dCN/dR array exists) and keeps the
  gradient G and the virial sigma; otherwise it runs the energy-only
  kernels;
* the backward pass turns G and sigma into the adjoints of the inputs
  (no second evaluation).

The lattices are read on the host (the lattice images are enumerated
there); a lattice tensor passed again unchanged (the same tensor object with
the same version counter) reuses the previous host copy, so a step with a
fixed cell does not synchronise on it.  (Writes through ``.data`` do not
advance the version counter: pass a new tensor after such a write.)

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
    # the same tensor object (held here, so its memory is not reused) with
    # the same version counter holds the same values
    cached = getattr(spec, "_lattice_spec", None)
    if cached is not None and cached[0] is lattice and cached[1] == lattice._version:
        return cached[2]
    lat = lattice.detach().to(torch.float64).cpu().numpy().reshape(len(spec), 3, 3)
    periodic = spec.periodic if spec.periodic is not None else np.ones((len(spec), 3), dtype=bool)
    run_spec = torch_api.BatchSpec(spec.natoms, spec.numbers, lat, periodic, spec.params, spec.cutoffs)
    spec._lattice_spec = (lattice, lattice._version, run_spec)
    return run_spec


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


def _owner(spec: torch_api.BatchSpec, device) -> torch.Tensor:
    """System index of every atom (cached on the spec per device)."""
    cache = spec.__dict__.setdefault("_owner", {})
    if device not in cache:
        natoms = torch.as_tensor(spec.natoms, dtype=torch.long, device=device)
        cache[device] = torch.repeat_interleave(torch.arange(len(spec), device=device), natoms)
    return cache[device]


class _Dispersion(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, lattice, spec, precision):
        pos64 = positions.detach().to(torch.float64).contiguous()
        run_spec = _spec_with_lattice(spec, lattice)
        need = ctx.needs_input_grad[0] or (lattice is not None and ctx.needs_input_grad[1])
        out = torch_api.run_batch(run_spec, pos64, precision=precision, grad=need)
        ctx.has_lattice = lattice is not None
        ctx.spec = run_spec
        if need:
            lat64 = lattice.detach().to(torch.float64) if lattice is not None else pos64.new_empty(0)
            ctx.save_for_backward(out["gradient"], out["sigma"], pos64, lat64)
        ctx.dtypes = (positions.dtype, lattice.dtype if lattice is not None else None)
        return out["energy"].to(positions.dtype)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_energy):
        g, sigma, pos64, lat64 = ctx.saved_tensors
        owner = _owner(ctx.spec, g.device)
        ge = grad_energy.to(torch.float64)
        grad_pos = (g * ge[owner].unsqueeze(1)).to(ctx.dtypes[0]) if ctx.needs_input_grad[0] else None
        grad_lat = None
        if ctx.has_lattice and ctx.needs_input_grad[1]:
            d = lattice_adjoint(g, sigma, pos64, lat64, owner)
            grad_lat = (d * ge.view(-1, 1, 1)).to(ctx.dtypes[1])
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
    owner = _owner(spec, positions.device)
    p64, l64 = pos.detach().double(), lat.detach().double()
    sig = torch.zeros(len(spec), 3, 3, dtype=torch.float64, device=positions.device)
    sig.index_add_(0, owner, gpos.double().unsqueeze(2) * p64.unsqueeze(1))
    sig += glat.double().transpose(1, 2) @ l64
    vol = torch.linalg.det(l64).abs().view(-1, 1, 1)
    return energy.detach(), -gpos, (sig / vol).to(positions.dtype)
