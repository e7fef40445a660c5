"""Headline shape: 32^3, Nev = 200, 33 momenta, 13 derivatives.

``test_config_c_full_launch_against_oracle`` runs eight timeslices, which the
library contracts in one automatic launch (about 17 GB of workspace, each
timeslice's block just under 2^31 bytes, so the last one sits at the highest
buffer offsets), in both forms, and checks blocks of the first and the last
timeslice against the oracle (a subset of derivatives and momenta: the
whole timeslice would take minutes in numpy), V^dag V = 1 on every
timeslice and the mixed form against the FP64 form everywhere.

``test_config_c_shape_properties`` (one timeslice) checks properties that
hold at any size:

* V^dag V = 1 (d = 0, p = 0) to complex64 rounding;
* E_{rev d}[p] = (-1)^{|d|} E_d[-p]^dagger with every derivative computed
  directly (Hermitian shortcut off), and the shortcut equal to it;
* translation: rolling U and V by a lattice vector a multiplies E[p] by
  exp(+2 pi i n.a / L) (elemental.py:290-338 with phase.py:41-46);
* the mixed form within 1e-6 of the FP64 form, blockwise per (d, p).
"""

import sys

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import _native
from easydistillation_amd import synthetic as S
from oracle import elemental_oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def blockwise(E, R):
    """max over (d, p) of max|E - R| / max|R| (device tensors [t, D, P, Ne, Ne])."""
    import torch

    D, P = R.shape[1], R.shape[2]
    d = (E - R).abs().permute(1, 2, 0, 3, 4).reshape(D, P, -1).amax(-1)
    r = R.abs().permute(1, 2, 0, 3, 4).reshape(D, P, -1).amax(-1)
    row = r.amax(1, keepdim=True)
    den = torch.where(r >= 1e-8 * row, r, row.expand_as(r))
    return float((d / den.clamp_min(1e-300)).max())


def test_config_c_shape_properties():
    import torch

    latt, Ne, nn = [32, 32, 32, 1], 200, 2
    Lx, Ly, Lz, _ = latt
    moms = S.momenta_by_norm(4)
    D, P = ed.num_derivative(nn), len(moms)
    dev = torch.device("cuda:0")
    say("config C shape: inputs")
    U = S.gauge_torch(latt, "haar", seed=13, device=dev)
    V = S.eigvecs_torch(latt, Ne, seed=14, device=dev)

    def run(U, V, mode="fp64", hermitian=1):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.set_option(_native.ED_OPT_HERMITIAN, hermitian)
        g.load_arrays(U, V)
        out = g.calc_all(out=torch.empty((1, D, P, Ne, Ne), dtype=torch.complex128, device=dev))
        torch.cuda.synchronize()
        return out

    say("config C shape: fp64, every derivative computed")
    E = run(U, V, hermitian=0)
    idx = {m: i for i, m in enumerate(moms)}
    eye = torch.eye(Ne, dtype=torch.complex128, device=dev)
    assert float((E[0, 0, idx[(0, 0, 0)]] - eye).abs().max()) < 1e-5

    dl = O.derivative_list(nn)
    worst = 0.0
    for di, d in enumerate(dl):
        rd = dl.index(tuple(reversed(d)))
        for m in [(0, 0, 0), (1, 0, 0), (1, 1, -1), (0, 2, 0), (-1, 0, 1)]:
            mm = tuple(-c for c in m)
            lhs = E[0, rd, idx[m]]
            rhs = (-1) ** len(d) * E[0, di, idx[mm]].conj().transpose(0, 1)
            worst = max(worst, float((lhs - rhs).abs().max() / lhs.abs().max()))
    say(f"  Hermiticity relations, direct: {worst:.2e}")
    assert worst <= 1e-12

    say("config C shape: fp64 with the Hermitian shortcut")
    E1 = run(U, V, hermitian=1)
    err = blockwise(E1, E)
    say(f"  shortcut vs direct {err:.2e}")
    assert err <= 1e-12
    del E1

    say("config C shape: translated inputs")
    a = (5, 3, 2)  # (x, y, z) lattice vector
    Ur = torch.roll(U, shifts=(a[2], a[1], a[0]), dims=(2, 3, 4))
    Vr = torch.roll(V, shifts=(a[2], a[1], a[0]), dims=(2, 3, 4))
    Et = run(Ur, Vr)
    del Ur, Vr
    n = torch.tensor(moms, dtype=torch.float64, device=dev)
    arg = 2 * np.pi * (n[:, 0] * a[0] / Lx + n[:, 1] * a[1] / Ly + n[:, 2] * a[2] / Lz)
    phase = torch.polar(torch.ones_like(arg), arg).view(1, 1, P, 1, 1)
    err = blockwise(Et, phase * E)
    say(f"  translation covariance {err:.2e}")
    # two FP64 evaluations (different summation order over the sites), each
    # within the 1e-12 bar of the exact value
    assert err <= 2e-12
    del Et

    say("config C shape: mixed")
    Em = run(U, V, mode="mixed")
    err = blockwise(Em, E)
    say(f"  mixed vs fp64 {err:.2e}")
    assert err <= 1e-6


def test_config_c_full_launch_against_oracle():
    import torch

    latt, Ne, nn = [32, 32, 32, 8], 200, 2
    Lt = latt[3]
    moms = S.momenta_by_norm(4)
    D, P = ed.num_derivative(nn), len(moms)
    idx = {m: i for i, m in enumerate(moms)}
    dev = torch.device("cuda:0")
    say("config C, 8 timeslices: inputs")
    U = S.gauge_torch(latt, "haar", seed=41, device=dev)
    V = S.eigvecs_torch(latt, Ne, seed=42, device=dev)

    outs = {}
    for mode in ("fp64", "mixed"):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.set_option(_native.ED_OPT_PROFILE, 1)
        g.load_arrays(U, V)
        say(f"config C, 8 timeslices: {mode}")
        outs[mode] = g.calc_all(out=torch.empty((Lt, D, P, Ne, Ne), dtype=torch.complex128, device=dev))
        torch.cuda.synchronize()
        ms, launches = g.context.profile_read(_native.ED_KERNEL_CONTRACT)
        say(f"  {launches} contraction launch(es), {ms:.1f} ms")
        # all eight timeslices in one automatic launch, as the benchmark times it
        assert launches == 1
        del g
    E64, Em = outs["fp64"], outs["mixed"]

    eye = torch.eye(Ne, dtype=torch.complex128, device=dev)
    vv = float((E64[:, 0, idx[(0, 0, 0)]] - eye).abs().max())
    say(f"  V^dag V - 1 over 8 timeslices: {vv:.2e}")
    assert vv < 1e-5
    err = blockwise(Em, E64)
    say(f"  mixed vs fp64 over 8 timeslices: {err:.2e}")
    assert err <= 1e-6

    # derivatives 5 (x then y) and 7 (y then x) are a Hermitian pair: one is
    # contracted, the other written as its partner; 12 (z z) is palindromic
    checks = [(0, [0, 5, 7, 12], [(0, 0, 0), (1, 0, 0), (1, 1, -1), (0, -2, 0)]),
              (Lt - 1, [2, 9], [(-1, 0, 1), (0, 0, 2)])]
    for t, dsel, msel in checks:
        say(f"  oracle at t={t}: derivatives {dsel}, momenta {msel}")
        Vt = V[t].cpu().numpy()
        Ut = U[:, t].cpu().numpy()
        ref = O.elemental_timeslice(Vt, Ut, latt, nn, msel, derivative_indices=dsel)
        pidx = [idx[m] for m in msel]
        for mode, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
            got = outs[mode][t][dsel][:, pidx].cpu().numpy()
            d = np.abs(got - ref).reshape(len(dsel), len(msel), -1).max(-1)
            r = np.abs(ref).reshape(len(dsel), len(msel), -1).max(-1)
            e = float((d / r).max())
            say(f"    {mode}: max blockwise {e:.2e}")
            assert e <= tol
