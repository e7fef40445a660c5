"""FP64 torch restatement of elemental generation, differentiated by torch autograd.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the package
``easydistillation_amd``) may import this module; tests use it as the FP64
autograd reference for the hand-written HIP adjoint.

It follows the same reference lines as ``oracle/elemental_oracle.py``:
``hop`` / ``nD`` follow lattice/generator/elemental.py:279-288, the split
loop :309-321, the phased contraction :322-329, blending :330-337, the phase
lattice/insertion/phase.py:41-46.  V is used as given (no complex64
rounding), so that derivatives with respect to V are those of the exact
map.  Pinned against the numpy oracle in tests/test_autograd_oracle.py;
the gradients are torch's own autograd of these operations.

Layouts: U_t [3, Lz, Ly, Lx, 3, 3], V_t [Ne, Lz, Ly, Lx, 3],
E_t [D, P, Ne, Ne]; all complex128 tensors on any device.
"""

from __future__ import annotations

import math
from typing import Sequence

import torch

from .elemental_oracle import derivative_list, splits


def momentum_phase(latt_size: Sequence[int], n: Sequence[int], device) -> torch.Tensor:
    """exp(+2 pi i n.x / L) on [Lz, Ly, Lx] (phase.py:41-46), products reduced mod L."""
    Lx, Ly, Lz = latt_size[0], latt_size[1], latt_size[2]
    x = torch.arange(Lx, device=device).view(1, 1, Lx)
    y = torch.arange(Ly, device=device).view(1, Ly, 1)
    z = torch.arange(Lz, device=device).view(Lz, 1, 1)
    f64 = torch.float64
    frac = (n[0] * x % Lx).to(f64) / Lx + (n[1] * y % Ly).to(f64) / Ly + (n[2] * z % Lz).to(f64) / Lz
    ang = 2 * math.pi * frac
    return torch.polar(torch.ones_like(ang), ang)


def hop(V: torch.Tensor, U_t: torch.Tensor, mu: int) -> torch.Tensor:
    """(nabla_mu V)(x) = U_mu(x) V(x + mu) - U_mu(x - mu)^dag V(x - mu) (elemental.py:282-287)."""
    ax = 3 - mu  # axis of direction mu in [e, z, y, x, c]
    Vf = torch.roll(V, -1, ax)
    fwd = torch.matmul(U_t[mu], Vf.unsqueeze(-1)).squeeze(-1)
    Ud = U_t[mu].conj().transpose(-1, -2).contiguous()
    bwd = torch.roll(torch.matmul(Ud, V.unsqueeze(-1)).squeeze(-1), 1, ax)
    return fwd - bwd


def nD(V, U_t, deriv):
    out = V
    for mu in deriv:
        out = hop(out, U_t, mu)
    return out


def elemental_timeslice(V_t, U_t, latt_size, num_nabla, momentum_list, stocastic_coeff=None):
    Ne = V_t.shape[0]
    dev = V_t.device
    phases = [momentum_phase(latt_size, m, dev) for m in momentum_list]
    cache = {(): V_t}

    def W(dirs):
        if dirs not in cache:
            cache[dirs] = nD(V_t, U_t, dirs)
        return cache[dirs]

    rows = []
    for deriv in derivative_list(num_nabla):
        per_p = [None] * len(phases)
        for left_dirs, right_dirs, coeff in splits(deriv):
            L = W(left_dirs).conj()
            Rt = W(right_dirs).reshape(Ne, -1).transpose(0, 1).contiguous()
            for pi, ph in enumerate(phases):
                term = coeff * torch.matmul((L * ph[None, :, :, :, None]).reshape(Ne, -1), Rt)
                per_p[pi] = term if per_p[pi] is None else per_p[pi] + term
        rows.append(torch.stack(per_p))
    E = torch.stack(rows)
    if stocastic_coeff is not None:
        E = E * torch.as_tensor(stocastic_coeff, dtype=torch.float64, device=dev)
    return E


def elemental(V, U, latt_size, num_nabla, momentum_list, stocastic_coeff=None):
    """[Lt, D, P, Ne, Ne] from U [3, Lt, Lz, Ly, Lx, 3, 3] and V [Lt, Ne, Lz, Ly, Lx, 3]."""
    return torch.stack(
        [
            elemental_timeslice(V[t], U[:, t], latt_size, num_nabla, momentum_list, stocastic_coeff)
            for t in range(V.shape[0])
        ]
    )


def profile_trace(E, w):
    """T[t, d, p] = sum_k w_k E[t, d, p, k, k]."""
    return (torch.diagonal(E, dim1=-2, dim2=-1) * w).sum(-1)


def profiled_loss(E, w):
    """L = sum_{t,d,p} |sum_k w_k E[t,d,p,k,k]|^2."""
    return (profile_trace(E, w).abs() ** 2).sum()
