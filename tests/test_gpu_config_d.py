"""Batched uneven form at its benchmark workload: 16^3 x 32 / Nev 64,
24^3 x 48 / Nev 128 and 32^3 x 64 / Nev 256, smooth links (eps = 0.3),
19 momenta with |n|^2 <= 2, num_nabla = 1 (4 derivatives), mixed form, every
timeslice of the three lattices in one graph replay of the autotuned batch.

Checks, per lattice:

* the batch output equals the lattice's own single-context mixed output bit
  for bit (each work item runs the same arithmetic);
* the mixed output is within 1e-6 blockwise of the FP64 form;
* V^dag V = 1 (d = 0, p = 0) on every timeslice;
* E_d[p] = (-1)^{|d|} E_d[-p]^dagger for every d and p (with num_nabla = 1
  every derivative is its own reverse), on the batch output;
* oracle values at t = 0 for a subset of momenta (16^3 lattice: every
  derivative; 24^3 lattice: two derivatives).
"""

import sys

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import _native
from easydistillation_amd import synthetic as S
from easydistillation_amd.batch import ElementalBatch
from oracle import elemental_oracle as O

pytestmark = pytest.mark.gpu

LATTICES = [([16, 16, 16, 32], 64), ([24, 24, 24, 48], 128), ([32, 32, 32, 64], 256)]


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def blockwise(E, R):
    """max over (d, p) of max|E - R| / max|R| over device tensors [t, D, P, Ne, Ne]."""
    import torch

    D, P = R.shape[1], R.shape[2]
    d = torch.zeros((D, P), dtype=torch.float64, device=R.device)
    r = torch.zeros_like(d)
    for t in range(R.shape[0]):  # one timeslice at a time keeps the temporaries small
        d = torch.maximum(d, (E[t] - R[t]).abs().reshape(D, P, -1).amax(-1))
        r = torch.maximum(r, R[t].abs().reshape(D, P, -1).amax(-1))
    row = r.amax(1, keepdim=True)
    den = torch.where(r >= 1e-8 * row, r, row.expand_as(r))
    return float((d / den.clamp_min(1e-300)).max())


def test_config_d_workload():
    import torch

    dev = torch.device("cuda:0")
    moms = S.momenta_by_norm(2)
    nn = 1
    D, P = ed.num_derivative(nn), len(moms)
    idx = {m: i for i, m in enumerate(moms)}
    neg = [idx[tuple(-c for c in m)] for m in moms]
    gens, inputs = [], []
    for i, (latt, Ne) in enumerate(LATTICES):
        say(f"config D: inputs {latt} Nev {Ne}")
        U = S.gauge_torch(latt, "smooth", 0.3, seed=51 + i, device=dev)
        V = S.eigvecs_torch(latt, Ne, seed=61 + i, device=dev)
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode="mixed")
        g.load_arrays(U, V)
        gens.append(g)
        inputs.append((U, V))
    torch.cuda.synchronize()

    batch = ElementalBatch(gens)
    ms = batch.tune(1)
    h16 = batch.option(_native.ED_BOPT_H16)
    say(f"config D: tuned batch {ms:.1f} ms per replay, FP16 split sum {h16}")
    outs = batch.run()
    # the single-context runs below take the split-sum kernel the tuner chose
    for g in gens:
        g.set_option(_native.ED_OPT_MIXED_WS, 2 if h16 == 1 else 1)

    for (latt, Ne), g, out, (U, V) in zip(LATTICES, gens, outs, inputs):
        Lt = latt[3]
        tag = f"{latt[0]}^3x{Lt}/Nev{Ne}"
        single = g.calc_all(out=torch.empty_like(out))
        same = bool(torch.equal(single, out))
        say(f"{tag}: batch == single bit for bit: {same}")
        assert same
        del single

        E64 = g.calc_all(mode="fp64", out=torch.empty_like(out))
        err = blockwise(out, E64)
        say(f"{tag}: mixed vs fp64 {err:.2e}")
        assert err <= 1e-6
        del E64

        eye = torch.eye(Ne, dtype=torch.complex128, device=dev)
        vv = float((out[:, 0, idx[(0, 0, 0)]] - eye).abs().max())
        say(f"{tag}: V^dag V - 1 {vv:.2e}")
        assert vv < 1e-5

        herm = 0.0
        for d, dirs in enumerate(O.derivative_list(nn)):
            lhs = out[:, d]
            rhs = (-1) ** len(dirs) * out[:, d, neg].conj().transpose(-1, -2)
            herm = max(herm, float((lhs - rhs).abs().max() / lhs.abs().max()))
        say(f"{tag}: E_d[p] vs (-1)^|d| E_d[-p]^dagger {herm:.2e}")
        assert herm <= 1e-6

        if Ne <= 128:
            dsel = list(range(D)) if Ne <= 64 else [0, 2]
            msel = [(0, 0, 0), (1, 0, 0), (0, -1, 1), (-1, 1, 0)]
            Vt = V[0].cpu().numpy()
            Ut = U[:, 0].cpu().numpy()
            ref = O.elemental_timeslice(Vt, Ut, latt, nn, msel, derivative_indices=dsel)
            got = out[0][dsel][:, [idx[m] for m in msel]].cpu().numpy()
            dd = np.abs(got - ref).reshape(len(dsel), len(msel), -1).max(-1)
            rr = np.abs(ref).reshape(len(dsel), len(msel), -1).max(-1)
            e = float((dd / rr).max())
            say(f"{tag}: oracle at t=0, derivatives {dsel}, momenta {msel}: {e:.2e}")
            assert e <= 1e-6
