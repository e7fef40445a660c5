"""Contraction autotuner (ed_tune, easydistillation_amd.tune).

* every strip height and item order gives the same elementals bit for bit
  (they only reorder which workgroup runs which tile);
* the tuner leaves a valid choice on the context and results stay within
  the bars against the oracle;
* the JSON table: a second generator of the same shape reads the stored
  choice instead of timing, a different shape is timed again;
* the batch tuner's strip heights keep batch == single bit for bit.
"""

import sys

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import _native
from easydistillation_amd import synthetic as S
from easydistillation_amd.batch import ElementalBatch
from easydistillation_amd.tune import autotune, load_table
from oracle import elemental_oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def blockwise(E, R):
    d = np.abs(E - R).transpose(1, 2, 0, 3, 4).reshape(R.shape[1], R.shape[2], -1).max(-1)
    r = np.abs(R).transpose(1, 2, 0, 3, 4).reshape(R.shape[1], R.shape[2], -1).max(-1)
    row = r.max(1, keepdims=True)
    return float((d / np.maximum(np.where(r >= 1e-8 * row, r, row), 1e-300)).max())


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_tile_orders_agree_and_tuner_choice_is_valid(mode, tmp_path):
    latt, Ne, nn = [16, 16, 16, 2], 40, 2
    moms = S.momenta_by_norm(2)
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=23)
    g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
    g.load_arrays(U, V)
    base = g.calc_all()
    for strip in (1, 3, 8):
        for order in (0, 1):
            g.set_option(_native.ED_OPT_STRIP, strip)
            g.set_option(_native.ED_OPT_ITEM_ORDER, order)
            assert np.array_equal(g.calc_all(), base), (strip, order)
    g.set_option(_native.ED_OPT_STRIP, 0)
    g.set_option(_native.ED_OPT_ITEM_ORDER, 0)
    cache = str(tmp_path / "tune.json")
    got = autotune(g, cache=cache, nt=2)
    say(f"{mode}: tuned {got}")
    assert got["source"] == "timed" and got["strip"] in (1, 2, 4, 8) and got["item_order"] in (0, 1)
    ref = O.elemental(V, U, latt, nn, moms)
    err = blockwise(g.calc_all(), ref)
    say(f"{mode}: tuned result vs oracle {err:.2e}")
    assert err <= (1e-12 if mode == "fp64" else 1e-6)

    # same shape: read from the table, no timing
    h = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
    h.load_arrays(U, V)
    again = autotune(h, cache=cache)
    assert again["source"] == "table"
    for k in ("strip", "item_order", "mixed_ws"):
        assert again[k] == got[k] and h.context.get_option(
            {"strip": _native.ED_OPT_STRIP, "item_order": _native.ED_OPT_ITEM_ORDER,
             "mixed_ws": _native.ED_OPT_MIXED_WS}[k]) == got[k]
    assert np.array_equal(h.calc_all(), g.calc_all())
    # another shape is timed and stored beside the first
    k = ed.ElementalGenerator(latt, None, None, 1, moms, Ne=Ne, mode=mode)
    k.load_arrays(U, V)
    assert autotune(k, cache=cache, nt=1)["source"] == "timed"
    assert len(load_table(cache)) == 2


def test_batch_strip_heights_keep_single_results():
    gens = []
    singles = []
    for i, (latt, Ne, nn) in enumerate([([16, 16, 16, 2], 20, 1), ([8, 8, 8, 2], 24, 2)]):
        U, V = S.make_inputs_np(latt, Ne, "smooth", seed=30 + i)
        g = ed.ElementalGenerator(latt, None, None, nn, S.momenta_by_norm(2), Ne=Ne, mode="mixed")
        g.load_arrays(U, V)
        singles.append(g.calc_all())
        gens.append(g)
    b = ElementalBatch(gens)
    for strip in (0, 1, 8):
        b.set_option(_native.ED_BOPT_STRIP, strip)
        outs = [o.cpu().numpy() for o in b.run()]
        for o, s in zip(outs, singles):
            assert np.array_equal(o, s), strip
    best = b.tune(1)
    h16 = b.option(_native.ED_BOPT_H16)
    say(f"batch tuned: {best:.2f} ms, strip {b.option(_native.ED_BOPT_STRIP)}, FP16 split sum {h16}")
    assert b.option(_native.ED_BOPT_STRIP) in (1, 2, 4, 8)
    outs = [o.cpu().numpy() for o in b.run()]
    # the single-context runs take the split-sum kernel the tuner chose
    for g in gens:
        g.set_option(_native.ED_OPT_MIXED_WS, 2 if h16 == 1 else 1)
    for o, g in zip(outs, gens):
        assert np.array_equal(o, g.calc_all())
