"""Pins the torch FP64 restatement used as the autograd reference, and the
profile-trace adjoint, on the CPU.

The torch restatement must reproduce the numpy oracle's forward (which is
itself pinned by tests/test_oracle.py); its gradients are then torch's own
autograd.  ``ProfileTraceFunction``'s written-out backward is checked against
torch autograd of the plain expression.
"""

import numpy as np
import pytest
import torch

from easydistillation_amd import synthetic as S
from easydistillation_amd.autograd import ProfileTraceFunction
from oracle import elemental_oracle as O
from oracle import elemental_torch as T


@pytest.mark.parametrize("num_nabla,blend", [(0, False), (1, True), (2, False)])
def test_torch_restatement_matches_numpy_oracle(num_nabla, blend):
    latt = [3, 4, 5, 2]
    Ne = 5
    rng = np.random.default_rng(13)
    U = S.gauge_np(rng, latt, "haar")
    V = S.eigvecs_np(rng, latt, Ne).astype(np.complex128)
    V = V + 1e-3 * (rng.normal(size=V.shape) + 1j * rng.normal(size=V.shape))  # not c64-representable
    moms = S.momenta_by_norm(2)[:9]
    w = rng.uniform(0.5, 1.5, size=(Ne, Ne)) if blend else None
    Et = T.elemental(torch.from_numpy(V), torch.from_numpy(U), latt, num_nabla, moms, w).numpy()
    En = np.stack(
        [O.elemental_timeslice(V[t], U[:, t], latt, num_nabla, moms, w, round_v_to_c64=False) for t in range(latt[3])]
    )
    assert np.abs(Et - En).max() / np.abs(En).max() < 1e-13


def test_torch_phase_matches_closed_form():
    latt = [5, 4, 3, 1]
    for n in [(1, 0, 0), (-2, 3, 1), (0, -1, -4)]:
        got = T.momentum_phase(latt, n, "cpu").numpy()
        want = O.momentum_phase(latt, n)
        assert np.abs(got - want).max() < 1e-14


def test_profile_trace_backward_matches_autograd():
    g = torch.Generator().manual_seed(3)
    E = torch.randn((2, 4, 3, 6, 6), dtype=torch.complex128, generator=g).requires_grad_()
    w = (torch.rand(6, dtype=torch.float64, generator=g) + 0.5).requires_grad_()
    L1 = (ProfileTraceFunction.apply(E, w).abs() ** 2).sum()
    gE1, gw1 = torch.autograd.grad(L1, [E, w])
    L2 = T.profiled_loss(E, w)
    gE2, gw2 = torch.autograd.grad(L2, [E, w])
    assert torch.allclose(L1, L2, rtol=1e-14, atol=0)
    assert (gE1 - gE2).abs().max() < 1e-12
    assert (gw1 - gw2).abs().max() < 1e-12


def test_torch_autograd_matches_finite_differences():
    # sanity of the reference's differentiability (complex roll / matmul / conj)
    latt = [3, 3, 4, 1]
    Ne = 3
    rng = np.random.default_rng(2)
    U = torch.from_numpy(S.gauge_np(rng, latt, "haar")).requires_grad_()
    V = torch.from_numpy(S.eigvecs_np(rng, latt, Ne).astype(np.complex128)).requires_grad_()
    w = torch.from_numpy(rng.uniform(0.5, 1.5, Ne))
    moms = S.unit_momenta()

    def loss(U, V):
        return T.profiled_loss(T.elemental(V, U, latt, 1, moms), w)

    gU, gV = torch.autograd.grad(loss(U, V), [U, V])
    h = 1e-6
    for k in range(4):
        dU = torch.from_numpy(rng.normal(size=U.shape) + 1j * rng.normal(size=U.shape))
        dV = torch.from_numpy(rng.normal(size=V.shape) + 1j * rng.normal(size=V.shape))
        with torch.no_grad():
            fd = (loss(U + h * dU, V + h * dV) - loss(U - h * dU, V - h * dV)) / (2 * h)
        an = (gU.conj() * dU).real.sum() + (gV.conj() * dV).real.sum()
        assert abs(float(fd - an)) < 1e-6 * abs(float(an))
