"""ed_compute_noalloc (fixed workspace, no allocation) against ed_compute.

Both run the same kernels on the same inputs, so the outputs must be equal
bit for bit, in both forms.  Calls in a row: a 16^3 x 4 lattice (Nev 40, two
derivatives, batches of two timeslices in the fixed workspace), then a
smaller one (16^3 x 2, Nev 20, one derivative, different padding) on another
context, then the first again -- each call must clear what the previous one
left.  Error paths: the plan not built yet, a timeslice larger than the
fixed workspace.
"""

import sys

import pytest

import easydistillation_amd as ed
from easydistillation_amd import _native
from easydistillation_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def make(latt, Ne, nn, mode, seed):
    import torch

    moms = S.momenta_by_norm(2)
    g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
    U = S.gauge_torch(latt, "haar", seed=seed, device="cuda:0")
    V = S.eigvecs_torch(latt, Ne, seed=seed + 1, device="cuda:0")
    g.load_arrays(U, V)
    shape = (latt[3], g.num_derivative, g.num_momentum, Ne, Ne)
    return g, shape, torch


def noalloc(g, shape, torch, mode):
    out = torch.full(shape, float("nan"), dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    g.context.compute_noalloc(0, shape[0], _native.ED_MODE_MIXED if mode == "mixed" else _native.ED_MODE_FP64,
                              out.data_ptr())
    g.context.synchronize()
    return out


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_noalloc_equals_compute(mode):
    big, big_shape, torch = make([16, 16, 16, 4], 40, 2, mode, seed=71)
    small, small_shape, _ = make([16, 16, 16, 2], 20, 1, mode, seed=81)
    ref_big = big.calc_all(out=torch.empty(big_shape, dtype=torch.complex128, device="cuda:0"))
    ref_small = small.calc_all(out=torch.empty(small_shape, dtype=torch.complex128, device="cuda:0"))
    torch.cuda.synchronize()
    for name, g, shape, ref in [("16^3x4/Nev40", big, big_shape, ref_big), ("16^3x2/Nev20", small, small_shape, ref_small),
                                ("16^3x4/Nev40 again", big, big_shape, ref_big)]:
        got = noalloc(g, shape, torch, mode)
        same = bool(torch.equal(got, ref))
        say(f"{mode} {name}: noalloc == compute bit for bit: {same}")
        assert same


def test_noalloc_error_paths():
    import torch

    latt, Ne = [16, 16, 16, 1], 8
    g = ed.ElementalGenerator(latt, None, None, 1, [(0, 0, 0)], Ne=Ne, mode="mixed")
    U = S.gauge_torch(latt, "haar", seed=3, device="cuda:0")
    V = S.eigvecs_torch(latt, Ne, seed=4, device="cuda:0")
    g.load_arrays(U, V)
    out = torch.empty((1, g.num_derivative, 1, Ne, Ne), dtype=torch.complex128, device="cuda:0")
    # the plan is built lazily by ed_compute / ed_query; before that the call refuses
    with pytest.raises(_native.NativeError) as e:
        g.context.compute_noalloc(0, 1, _native.ED_MODE_MIXED, out.data_ptr())
    assert e.value.code == _native.ED_ERR_STATE
    g.context.query(_native.ED_Q_SPLITS)
    g.context.compute_noalloc(0, 1, _native.ED_MODE_MIXED, out.data_ptr())
    g.context.synchronize()
    with pytest.raises(_native.NativeError) as e:
        g.context.compute_noalloc(0, 1, _native.ED_MODE_MIXED, 0)
    assert e.value.code == _native.ED_ERR_ARG

    # one timeslice of 32^3 with Nev 200 and two derivatives needs ~2.1 GB: refused
    latt, Ne = [32, 32, 32, 1], 200
    h = ed.ElementalGenerator(latt, None, None, 2, [(0, 0, 0)], Ne=Ne, mode="mixed")
    h.load_arrays(S.gauge_torch(latt, "haar", seed=5, device="cuda:0"),
                  S.eigvecs_torch(latt, Ne, seed=6, device="cuda:0"))
    h.context.query(_native.ED_Q_SPLITS)
    big = torch.empty((1, h.num_derivative, 1, Ne, Ne), dtype=torch.complex128, device="cuda:0")
    with pytest.raises(_native.NativeError) as e:
        h.context.compute_noalloc(0, 1, _native.ED_MODE_MIXED, big.data_ptr())
    say(f"oversized timeslice: {e.value}")
    assert e.value.code == _native.ED_ERR_UNSUPPORTED
