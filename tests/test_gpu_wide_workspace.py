"""The warp-specialized mixed contraction past 2^31 bytes of workspace per
timeslice.

Its 64-bit-address form (taken from 2^31 bytes, or forced with
ED_OPT_WIDE_WS) must give the 32-bit-offset form's outputs bit for bit: a
padded eigenvector row reads a row of the same tile instead of zeros and an
unused (split, colour) slot reads live operands under a zero coefficient, so
only dropped products and exact zeros differ.  Past the line (32^3, Nev 256,
two derivatives: 2.6 GB per timeslice) the mixed form stays within 1e-6 of
the FP64 form, blockwise as in the other parity tests.
"""

import sys

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import _native
from easydistillation_amd import synthetic as S
from oracle import elemental_oracle as O

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def blockwise(E, Eref):
    E, Eref = np.asarray(E), np.asarray(Eref)
    d = np.abs(E - Eref).transpose(1, 2, 0, 3, 4).reshape(E.shape[1], E.shape[2], -1).max(-1)
    r = np.abs(Eref).transpose(1, 2, 0, 3, 4).reshape(E.shape[1], E.shape[2], -1).max(-1)
    row = r.max(1, keepdims=True)
    den = np.where(r >= 1e-8 * row, r, row)
    return float((d / np.maximum(den, 1e-300)).max())


CASES = [
    # padded rows (37 = 2 x 16 + 5), 1, 2, 3 and 4 split terms, 33 momenta
    ([16, 16, 16, 2], 37, 2, S.momenta_by_norm(4), "haar"),
    # 57 momenta (two momentum blocks), odd extent, smooth links
    ([16, 16, 17, 1], 16, 1, S.momenta_by_norm(5)[:57], "smooth"),
    # one tile with 11 padded rows, no derivative
    ([16, 16, 16, 1], 5, 0, [(0, 0, 0), (1, 0, 0), (0, -1, 1)], "haar"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_wide_addresses_match_32bit_offsets(case):
    latt, Ne, nn, moms, links = CASES[case]
    say(f"wide vs 32-bit: latt={latt} Ne={Ne} nabla={nn} P={len(moms)} {links}")
    U, V = S.make_inputs_np(latt, Ne, links, seed=300 + case)
    outs = []
    for wide in (0, 1):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode="mixed")
        g.set_option(_native.ED_OPT_WIDE_WS, wide)
        g.load_arrays(U, V)
        assert g.context.query(_native.ED_Q_MIXED_KERNEL) == (4 if wide else 3)
        outs.append(g.calc_all())
    err = blockwise(outs[1], O.elemental(V, U, latt, nn, moms))
    say(f"  identical: {np.array_equal(outs[0], outs[1])}, vs oracle {err:.2e}")
    assert np.array_equal(outs[0], outs[1])
    assert err <= MIXED_TOL


def test_mixed_past_2gib_per_timeslice():
    import torch

    latt, Ne, nn = [32, 32, 32, 1], 256, 2
    moms = S.momenta_by_norm(4)
    D, P = ed.num_derivative(nn), len(moms)
    # slots x padded Nev x colours x sites x complex64
    assert 13 * Ne * 3 * 32**3 * 8 >= 1 << 31
    # the 32-bit-only kinds give way to the 64-bit warp-specialized kernel here
    g0 = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode="mixed")
    for kind in (0, 1, 2):
        g0.set_option(_native.ED_OPT_MIXED_WS, kind)
        assert g0.context.query(_native.ED_Q_MIXED_KERNEL) == 4
    del g0
    dev = torch.device("cuda:0")
    say(f"past 2 GiB: latt={latt} Ne={Ne} nabla={nn} P={P}")
    U = S.gauge_torch(latt, "smooth", 0.3, seed=5, device=dev)
    V = S.eigvecs_torch(latt, Ne, seed=6, device=dev)
    outs = {}
    for mode in ("mixed", "fp64"):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.load_arrays(U, V)
        if mode == "mixed":
            assert g.context.query(_native.ED_Q_MIXED_KERNEL) == 4
        outs[mode] = g.calc_all(out=torch.empty((1, D, P, Ne, Ne), dtype=torch.complex128, device=dev))
        torch.cuda.synchronize()
        del g
    E, R = outs["mixed"], outs["fp64"]
    d = (E - R).abs().permute(1, 2, 0, 3, 4).reshape(D, P, -1).amax(-1)
    r = R.abs().permute(1, 2, 0, 3, 4).reshape(D, P, -1).amax(-1)
    row = r.amax(1, keepdim=True)
    den = torch.where(r >= 1e-8 * row, r, row.expand_as(r))
    err = float((d / den.clamp_min(1e-300)).max())
    say(f"  mixed vs fp64 blockwise {err:.2e}")
    assert err <= MIXED_TOL
    assert float(r.max()) > 0


def test_batch_wide_form_equals_single_lattices():
    """A batch holding a lattice past 2^31 bytes per timeslice (and one forced
    wide) runs the grouped kernels' 64-bit-address form: every output equals
    the lattice's single-lattice output bit for bit, for both launch shapes."""
    import torch

    from easydistillation_amd.batch import ElementalBatch

    dev = torch.device("cuda:0")
    moms = S.momenta_by_norm(4)
    gens = []
    big = [32, 32, 32, 1]
    U = S.gauge_torch(big, "smooth", 0.3, seed=7, device=dev)
    V = S.eigvecs_torch(big, 256, seed=8, device=dev)
    g = ed.ElementalGenerator(big, None, None, 2, moms, Ne=256, mode="mixed")
    g.load_arrays(U, V)
    gens.append(g)
    del U, V
    for i, (latt, Ne, wide) in enumerate((([16, 16, 16, 2], 21, 1), ([16, 16, 18, 1], 40, 0))):
        U, V = S.make_inputs_np(latt, Ne, "haar", seed=40 + i)
        g = ed.ElementalGenerator(latt, None, None, 2, moms, Ne=Ne, mode="mixed")
        g.set_option(_native.ED_OPT_WIDE_WS, wide)
        g.load_arrays(U, V)
        gens.append(g)
    singles = [g.calc_all(out=torch.empty((g.latt_size[3], 13, len(moms), g.Ne, g.Ne), dtype=torch.complex128,
                                          device=dev)) for g in gens]
    b = ElementalBatch(gens)
    for grid in (0, 2):
        b.set_option(_native.ED_BOPT_GRID_PER_CU, grid)
        outs = b.run()
        torch.cuda.synchronize()
        same = [bool(torch.equal(o, s)) for o, s in zip(outs, singles)]
        say(f"batch with a 2.6 GB-per-timeslice lattice, grid/CU={grid}: identical {same}")
        assert all(same)


def test_batch_wide_form_with_long_split_lists_is_refused():
    """The grouped kernels take derivatives of more than four split terms
    (three derivatives) only in the 32-bit form: a batch that needs both
    raises instead of running."""
    from easydistillation_amd.batch import ElementalBatch

    moms = [(0, 0, 0), (1, 0, 0)]
    gens = []
    for latt, Ne, nn, wide in (([16, 16, 16, 1], 8, 1, 1), ([16, 16, 16, 1], 8, 3, 0)):
        U, V = S.make_inputs_np(latt, Ne, "haar", seed=11)
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode="mixed")
        g.set_option(_native.ED_OPT_WIDE_WS, wide)
        g.load_arrays(U, V)
        gens.append(g)
    b = ElementalBatch(gens)
    with pytest.raises(_native.NativeError, match="split terms") as err:
        b.run()
    assert err.value.code == _native.ED_ERR_UNSUPPORTED
