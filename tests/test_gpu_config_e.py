"""Differentiation form at full size: 24^3 x 64, Nev = 128, Haar links, orthonormal
V, profile weights w ~ U(0.5, 1.5), 7 momenta, 4 derivative combinations,
loss L = sum_{t,d,p} |sum_k w_k E[t,d,p,k,k]|^2.

Checks dL/dV, dL/dU (HIP adjoint) and dL/dw (profile-trace adjoint):
  * against FP64 central finite differences, h = 1e-6, on 64 random
    directions; each direction is supported on one random timeslice t, and
    since L is a sum of per-timeslice terms and E[t] depends only on U[:, t]
    and V[t], the central difference needs only the term of t;
  * against FP64 torch autograd of the restatement in oracle/elemental_torch.py
    on two timeslices (the gradient restricted to t is that of term t).
"""

import sys
import time

import numpy as np
import pytest
import torch

from easydistillation_amd import synthetic as S
from easydistillation_amd.autograd import ElementalEngine, profiled_loss
from oracle import elemental_torch as T

pytestmark = pytest.mark.gpu

LATT = [24, 24, 24, 64]
NE = 128
NABLA = 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def nrm(a, b):
    return float((a - b).abs().max() / b.abs().max())


def test_config_e_gradients():
    dev = torch.device("cuda", 0)
    moms = S.unit_momenta()
    t_start = time.time()
    U = S.gauge_torch(LATT, "haar", seed=0, device=dev)
    V = S.eigvecs_torch(LATT, NE, seed=1, device=dev).to(torch.complex128)
    g = torch.Generator(device=dev).manual_seed(2)
    w = (torch.rand(NE, dtype=torch.float64, device=dev, generator=g) + 0.5)
    say(f"config E inputs built in {time.time() - t_start:.1f} s")

    eng = ElementalEngine(LATT, NE, NABLA, moms)
    Ug, Vg, wg = U.clone().requires_grad_(), V.clone().requires_grad_(), w.clone().requires_grad_()
    torch.cuda.synchronize()
    t0 = time.time()
    L = profiled_loss(Ug, Vg, wg, eng)
    torch.cuda.synchronize()
    t1 = time.time()
    gU, gV, gw = torch.autograd.grad(L, [Ug, Vg, wg])
    torch.cuda.synchronize()
    t2 = time.time()
    say(f"loss {float(L.detach()):.10e}; forward {t1 - t0:.3f} s, backward {t2 - t1:.3f} s")

    # finite differences, 64 directions
    h = 1e-6
    rng = np.random.default_rng(3)
    worst = 0.0
    with torch.no_grad():
        for k in range(64):
            t = int(rng.integers(LATT[3]))
            dU = torch.randn((3,) + tuple(U.shape[2:]), dtype=torch.complex128, device=dev, generator=g)
            dV = torch.randn(tuple(V.shape[1:]), dtype=torch.complex128, device=dev, generator=g)
            vals = []
            for sgn in (1, -1):
                Up = U.clone()
                Up[:, t] += sgn * h * dU
                Vp = V.clone()
                Vp[t] += sgn * h * dV
                E = eng.forward(Up, Vp, t0=t, nt=1)
                vals.append(float(T.profiled_loss(E, w)))
            fd = (vals[0] - vals[1]) / (2 * h)
            an = float((gU[:, t].conj() * dU).real.sum() + (gV[t].conj() * dV).real.sum())
            rel = abs(fd - an) / abs(an)
            worst = max(worst, rel)
            if k % 16 == 15:
                say(f"  FD direction {k + 1}/64 (t={t}): fd {fd:.10e} analytic {an:.10e} rel {rel:.2e}")
    say(f"FD worst relative difference over 64 (U, V) directions: {worst:.2e}")
    assert worst < 1e-6

    # dL/dw: w enters after E, so the central difference reuses one forward
    with torch.no_grad():
        E = eng.forward(U, V)
        worst_w = 0.0
        for k in range(64):
            dw = torch.randn(NE, dtype=torch.float64, device=dev, generator=g)
            fd = (float(T.profiled_loss(E, w + h * dw)) - float(T.profiled_loss(E, w - h * dw))) / (2 * h)
            an = float((gw * dw).sum())
            worst_w = max(worst_w, abs(fd - an) / abs(an))
        del E
    say(f"FD worst relative difference over 64 w directions: {worst_w:.2e}")
    assert worst_w < 1e-6

    # FP64 torch autograd on two timeslices
    for t in (0, 37):
        Ut = U[:, t].clone().requires_grad_()
        Vt = V[t].clone().requires_grad_()
        Et = T.elemental_timeslice(Vt, Ut, LATT, NABLA, moms)
        Lt = T.profiled_loss(Et.unsqueeze(0), w)
        rU, rV = torch.autograd.grad(Lt, [Ut, Vt])
        eu, ev = nrm(gU[:, t], rU), nrm(gV[t], rV)
        say(f"autograd t={t}: dL/dU {eu:.2e} dL/dV {ev:.2e}")
        assert eu < 1e-11 and ev < 1e-11
        del Ut, Vt, Et, Lt, rU, rV
        torch.cuda.empty_cache()
