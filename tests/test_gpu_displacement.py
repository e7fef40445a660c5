"""Displacement elementals on the device (ed_set_displacement) against the oracle.

Reference: lattice/generator/displacement_elemental.py:53-96; the reference
test's shape (tests/test_displacement_elemental.py: 4^3 x 8, Ne = 20,
distance 8, six momenta) plus uneven shapes.  Bars: FP64 <= 1e-12, mixed
<= 1e-6 (normwise per (n, p) block).
"""

import sys

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import synthetic as S
from easydistillation_amd.batch import ElementalBatch
from oracle import elemental_oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def blockwise(E, Eref):
    E, Eref = np.asarray(E), np.asarray(Eref)
    d = np.abs(E - Eref).transpose(1, 2, 0, 3, 4).reshape(E.shape[1], E.shape[2], -1).max(-1)
    r = np.abs(Eref).transpose(1, 2, 0, 3, 4).reshape(E.shape[1], E.shape[2], -1).max(-1)
    row = r.max(1, keepdims=True)
    den = np.where(r >= 1e-8 * row, r, row)
    return float((d / np.maximum(den, 1e-300)).max())


REF_MOMS = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1), (0, 1, 2), (1, 1, 2)]
CASES = [
    ([4, 4, 4, 8], 20, 8, REF_MOMS, "haar"),
    ([5, 6, 3, 2], 17, 3, S.momenta_by_norm(2), "haar"),
    ([6, 4, 5, 3], 9, 1, S.unit_momenta(), "smooth"),
    ([4, 5, 6, 2], 12, 0, [(1, 0, 0), (0, 0, 0)], "haar"),
    ([16, 16, 16, 2], 24, 4, S.momenta_by_norm(2), "haar"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_displacement_against_oracle(case, mode):
    latt, Ne, dist, moms, kind = CASES[case]
    say(f"displacement {latt} Ne={Ne} distance={dist} P={len(moms)} {mode}")
    U, V = S.make_inputs_np(latt, Ne, kind, seed=60 + case)
    g = ed.DisplacementElementalGenerator(latt, None, None, dist, moms, Ne=Ne, mode=mode)
    g.load_arrays(U, V)
    E = g.calc_all()
    assert E.shape == (latt[3], dist + 1, len(moms), Ne, Ne)
    err = blockwise(E, O.displacement_elemental(V, U, latt, dist, moms))
    say(f"  err={err:.3e}")
    assert err <= (1e-12 if mode == "fp64" else 1e-6)


def test_displacement_single_timeslice_and_switch_back():
    latt, Ne, dist = [4, 4, 4, 3], 6, 2
    moms = S.unit_momenta()
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=3)
    g = ed.DisplacementElementalGenerator(latt, None, None, dist, moms, Ne=Ne)
    g.load_arrays(U, V)
    E1 = g.calc(1)
    ref = O.displacement_timeslice(V[1], U[:, 1], latt, dist, moms)
    assert np.abs(E1 - ref).max() / np.abs(ref).max() < 1e-12
    # the same context back in derivative mode
    g.context.set_derivatives(1)
    out = np.empty((latt[3], 4, len(moms), Ne, Ne), np.complex128)
    g.context.compute(0, latt[3], 0, out.ctypes.data, False)
    assert np.abs(out - O.elemental(V, U, latt, 1, moms)).max() < 1e-12


def test_displacement_in_a_batch():
    a_latt, b_latt = [4, 4, 4, 4], [6, 5, 4, 2]
    Ua, Va = S.make_inputs_np(a_latt, 10, "haar", seed=11)
    Ub, Vb = S.make_inputs_np(b_latt, 7, "haar", seed=12)
    ga = ed.DisplacementElementalGenerator(a_latt, None, None, 5, REF_MOMS, Ne=10)
    ga.load_arrays(Ua, Va)
    gb = ed.ElementalGenerator(b_latt, None, None, 2, S.unit_momenta(), Ne=7)
    gb.load_arrays(Ub, Vb)
    want = [ga.calc_all(), gb.calc_all()]
    outs = [o.cpu().numpy() for o in ElementalBatch([ga, gb]).run()]
    for w, o in zip(want, outs):
        assert np.abs(w - o).max() == 0
