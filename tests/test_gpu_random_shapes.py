"""Seeded random shapes through the C ABI: both forms against the oracle, and
batches of random lattices against the single-lattice outputs.

Lattice extents 2..9 (and 16^3 with odd extents in the last axis), Ne 1..40,
num_nabla 0..2, random momentum lists with repeats, rough and smooth links.
Bars: FP64 <= 1e-12; mixed <= 1e-6 (below V3 = 4096 the mixed form runs the
FP64 kernels; on the 16^3 cases its 64-bit-address form equals it bit for
bit); batch outputs equal the single-lattice outputs bit for bit for every
launch shape.
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


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def blockwise(E, Eref):
    E, Eref = np.asarray(E), np.asarray(Eref)
    d = np.abs(E - Eref).transpose(1, 2, 0, 3, 4).reshape(E.shape[1], E.shape[2], -1).max(-1)
    r = np.abs(Eref).transpose(1, 2, 0, 3, 4).reshape(E.shape[1], E.shape[2], -1).max(-1)
    row = r.max(1, keepdims=True)
    den = np.where(r >= 1e-8 * row, r, row)
    return float((d / np.maximum(den, 1e-300)).max())


def random_spec(rng, allm, big):
    if big:
        latt = [16, 16, int(rng.integers(16, 19)), 1]
    else:
        latt = [int(rng.integers(2, 10)) for _ in range(3)] + [int(rng.integers(1, 4))]
    Ne = int(rng.integers(1, 41))
    nn = int(rng.integers(0, 3))
    moms = [tuple(allm[i]) for i in rng.choice(len(allm), size=int(rng.integers(1, 40)))]
    links = "haar" if rng.random() < 0.6 else "smooth"
    return latt, Ne, nn, moms, links


@pytest.mark.parametrize("case", range(12))
def test_random_shape_against_oracle(case):
    rng = np.random.default_rng(1000 + case)
    big = case % 4 == 3
    latt, Ne, nn, moms, links = random_spec(rng, S.momenta_by_norm(5), big=big)
    U, V = S.make_inputs_np(latt, Ne, links, seed=int(rng.integers(0, 10**6)))
    ref = O.elemental(V, U, latt, nn, moms)
    errs, outs = {}, {}
    for mode in ("fp64", "mixed"):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.load_arrays(U, V)
        outs[mode] = g.calc_all()
        errs[mode] = blockwise(outs[mode], ref)
    say(f"random {case}: latt={latt} Ne={Ne} nabla={nn} P={len(moms)} {links}: "
        f"fp64 {errs['fp64']:.1e} mixed {errs['mixed']:.1e}")
    assert errs["fp64"] <= 1e-12
    assert errs["mixed"] <= 1e-6
    if big:
        # the 64-bit-address form of the mixed kernel gives the same outputs
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode="mixed")
        g.set_option(_native.ED_OPT_WIDE_WS, 1)
        g.load_arrays(U, V)
        assert g.context.query(_native.ED_Q_MIXED_KERNEL) == 4
        assert np.array_equal(g.calc_all(), outs["mixed"])


@pytest.mark.parametrize("case", range(4))
def test_random_batch_equals_single(case):
    rng = np.random.default_rng(2000 + case)
    allm = S.momenta_by_norm(4)
    gens = []
    for _ in range(int(rng.integers(2, 5))):
        latt, Ne, nn, moms, links = random_spec(rng, allm, big=rng.random() < 0.35)
        latt[3] = min(latt[3], 2)
        mode = "mixed" if rng.random() < 0.6 else "fp64"
        U, V = S.make_inputs_np(latt, Ne, links, seed=int(rng.integers(0, 10**6)))
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.load_arrays(U, V)
        gens.append(g)
    singles = [g.calc_all() for g in gens]
    b = ElementalBatch(gens)
    for grid, order, graph in ((0, 1, 1), (2, 0, 1), (1, 1, 0)):
        b.set_option(_native.ED_BOPT_GRID_PER_CU, grid)
        b.set_option(_native.ED_BOPT_ORDER, order)
        b.set_option(_native.ED_BOPT_GRAPH, graph)
        outs = [o.cpu().numpy() for o in b.run()]
        diff = max(float(np.abs(o - s).max()) for o, s in zip(outs, singles))
        say(f"random batch {case} ({len(gens)} lattices) grid/CU={grid} order={order} graph={graph}: {diff:.1e}")
        assert diff == 0
