"""Batched form (ed_batch_*): several uneven lattices in one graph launch.

Each work item of the grouped persistent contraction runs the same arithmetic
as the single-lattice kernel, so batch outputs must equal each generator's
own calc_all bit for bit, for every launch shape, item order and with or
without graph capture; graph replays (with other device work between them)
must all give the same result; and the outputs meet the FP64 / mixed bars
against the oracle.
"""

import sys

import numpy as np
import pytest
import torch

import easydistillation_amd as ed
from easydistillation_amd import _native
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


SPECS = [
    # latt, Ne, num_nabla, momenta, mode, links
    ([4, 4, 4, 3], 10, 1, S.momenta_by_norm(2), "fp64", "smooth"),
    ([6, 5, 4, 2], 24, 2, S.unit_momenta(), "mixed", "haar"),
    ([8, 8, 8, 2], 40, 1, S.momenta_by_norm(2), "fp64", "haar"),
    ([16, 16, 16, 2], 20, 1, S.momenta_by_norm(1), "mixed", "smooth"),
]


def make_generators(specs, seed=0):
    gens, inputs = [], []
    for i, (latt, Ne, nn, moms, mode, kind) in enumerate(specs):
        rng = np.random.default_rng(seed + i)
        U = S.gauge_np(rng, latt, kind)
        V = S.eigvecs_np(rng, latt, Ne)
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.load_arrays(U, V)
        gens.append(g)
        inputs.append((U, V))
    return gens, inputs


def test_batch_matches_single_and_oracle():
    gens, inputs = make_generators(SPECS)
    singles = [g.calc_all() for g in gens]
    b = ElementalBatch(gens)
    outs = [o.cpu().numpy() for o in b.run()]
    for (latt, Ne, nn, moms, mode, _), g, s, o, (U, V) in zip(SPECS, gens, singles, outs, inputs):
        diff = np.abs(o - s).max()
        ref = O.elemental(V, U, latt, nn, moms)
        err = blockwise(o, ref)
        say(f"batch {latt} Ne={Ne} nabla={nn} {mode}: |batch - single| = {diff:.1e}, vs oracle {err:.2e}")
        assert diff == 0
        assert err <= (1e-12 if mode == "fp64" else 1e-6)


@pytest.mark.parametrize("grid,order,graph", [(0, 0, 1), (1, 1, 1), (4, 0, 0), (2, 1, 0)])
def test_launch_shapes_agree(grid, order, graph):
    gens, _ = make_generators(SPECS, seed=10)
    singles = [g.calc_all() for g in gens]
    b = ElementalBatch(gens)
    b.set_option(_native.ED_BOPT_GRID_PER_CU, grid)
    b.set_option(_native.ED_BOPT_ORDER, order)
    b.set_option(_native.ED_BOPT_GRAPH, graph)
    outs = [o.cpu().numpy() for o in b.run()]
    say(f"grid/CU={grid} order={order} graph={graph}")
    for s, o in zip(singles, outs):
        assert np.abs(o - s).max() == 0


def test_graph_replays_with_work_between():
    gens, _ = make_generators(SPECS[:3], seed=20)
    b = ElementalBatch(gens)
    first = [o.clone() for o in b.run()]
    junk = torch.zeros(1 << 20, device="cuda")
    for r in range(4):
        for o in b.outputs:
            o.zero_()  # the replay must rewrite every element
        junk.add_(1.0)
        torch.cuda.synchronize()
        outs = b.run()
        for f, o in zip(first, outs):
            assert torch.equal(f, o), f"replay {r} differs"
    say("4 replays identical")


def test_recapture_after_input_change():
    gens, inputs = make_generators(SPECS[:2], seed=30)
    b = ElementalBatch(gens)
    b.run()
    gens[0].stout_smear(1, 0.1)  # replaces the resident links
    outs = [o.cpu().numpy() for o in b.run()]
    want = gens[0].calc_all()
    assert np.abs(outs[0] - want).max() == 0


def test_tune_picks_a_valid_shape():
    gens, _ = make_generators(SPECS, seed=40)
    b = ElementalBatch(gens)
    ms = b.tune(1)
    grid, order = b.option(_native.ED_BOPT_GRID_PER_CU), b.option(_native.ED_BOPT_ORDER)
    h16 = b.option(_native.ED_BOPT_H16)
    say(f"tuned: {ms:.3f} ms per run, grid/CU={grid} order={order} h16={h16}")
    assert ms > 0 and grid in (0, 1, 2, 4) and order in (0, 1) and h16 in (0, 1)
    outs = [o.cpu().numpy() for o in b.run()]
    # the single-context runs take the split-sum kernel the tuner chose
    for g in gens:
        g.set_option(_native.ED_OPT_MIXED_WS, 2 if h16 == 1 else 1)
    for g, o in zip(gens, outs):
        assert np.abs(o - g.calc_all()).max() == 0


H16_SPECS = [
    # the FP16 split sum fits the 16^3 mixed members (V3 >= 4096); the others keep their kinds
    ([16, 16, 16, 1], 40, 2, S.momenta_by_norm(2), "mixed", "haar"),
    ([6, 5, 4, 2], 24, 2, S.unit_momenta(), "mixed", "haar"),
    ([16, 16, 16, 3], 24, 1, S.momenta_by_norm(1), "mixed", "smooth"),
    ([8, 8, 8, 2], 20, 1, S.momenta_by_norm(2), "fp64", "haar"),
]


@pytest.mark.parametrize("grid,order", [(0, 1), (2, 0)])
def test_batch_fp16_split_sum(grid, order):
    """ED_BOPT_H16 1 runs the FP16 split sum (hi/lo FP16 workspace,
    contract_mixed16 arithmetic) for the mixed members it fits: their outputs
    equal the single-context ED_OPT_MIXED_WS 2 runs bit for bit (kernel 5 by
    ED_Q_MIXED_KERNEL), differ from the FP32 split sum of ED_BOPT_H16 0, and
    meet the mixed bar against the oracle; ED_BOPT_H16 -1 follows each
    context's ED_OPT_MIXED_WS."""
    gens, inputs = make_generators(H16_SPECS, seed=70)
    b = ElementalBatch(gens)
    b.set_option(_native.ED_BOPT_GRID_PER_CU, grid)
    b.set_option(_native.ED_BOPT_ORDER, order)
    b.set_option(_native.ED_BOPT_H16, 0)
    off = [o.cpu().numpy() for o in b.run()]
    b.set_option(_native.ED_BOPT_H16, 1)
    on = [o.cpu().numpy() for o in b.run()]
    for (latt, Ne, nn, moms, mode, _), g, o0, o1, (U, V) in zip(H16_SPECS, gens, off, on, inputs):
        fits = mode == "mixed" and np.prod(latt[:3]) >= 4096
        g.set_option(_native.ED_OPT_MIXED_WS, 2)
        single = g.calc_all()
        kernel = g.context.query(_native.ED_Q_MIXED_KERNEL) if mode == "mixed" else None
        err = blockwise(o1, O.elemental(V, U, latt, nn, moms))
        say(f"h16 batch {latt} Ne={Ne} nabla={nn} {mode}: kernel {kernel}, |batch - single| "
            f"{np.abs(o1 - single).max():.1e}, |h16 - fp32 split| {np.abs(o1 - o0).max():.1e}, vs oracle {err:.2e}")
        assert np.array_equal(o1, single)
        assert err <= (1e-12 if mode == "fp64" else 1e-6)
        if fits:
            assert kernel == 5
            assert not np.array_equal(o1, o0)
        else:
            assert np.array_equal(o1, o0)
    # -1: as each context's ED_OPT_MIXED_WS (now 2 on every member)
    b.set_option(_native.ED_BOPT_H16, -1)
    follow = [o.cpu().numpy() for o in b.run()]
    for o1, f in zip(on, follow):
        assert np.array_equal(o1, f)


def test_batch_rejects_duplicates_and_empty():
    gens, _ = make_generators(SPECS[:1], seed=50)
    nb = _native.Batch(0)
    with pytest.raises(RuntimeError):
        nb.run()
    out = torch.empty(1, dtype=torch.complex128, device="cuda")
    nb.add(gens[0].context, _native.ED_MODE_FP64, out.data_ptr())
    with pytest.raises(RuntimeError):
        nb.add(gens[0].context, _native.ED_MODE_FP64, out.data_ptr())


def test_batch_is_ordered_with_member_streams():
    """Calls through the C ABI that return before their work is done (device
    output, device eigenvectors) are ordered with a batch run: the batch
    waits for each member context's stream and the streams wait for the
    batch, with no host synchronization in between."""
    gens, inputs = make_generators(SPECS[2:], seed=40)
    singles = [g.calc_all() for g in gens]
    b = ElementalBatch(gens)
    b.run()
    torch.cuda.synchronize()
    # enqueue new eigenvectors from device memory and an ed_compute into a device buffer on
    # each context's own stream, then run the batch at once
    outs_dev = []
    for g, (U, V) in zip(gens, inputs):
        Vd = torch.from_numpy(V).to("cuda:0")
        torch.cuda.synchronize()
        g.context.set_eigvecs(Vd.data_ptr(), _native.ED_C64, True)
        o = torch.empty_like(b.outputs[gens.index(g)])
        g.context.compute(0, g.latt_size[3], _native.ED_MODE_FP64 if g.mode == "fp64" else _native.ED_MODE_MIXED,
                          o.data_ptr(), True)
        outs_dev.append((o, Vd))
    b.batch.run()
    b.batch.synchronize()
    for g in gens:
        g.context.synchronize()
    for g, s, o, (od, _) in zip(gens, singles, b.outputs, outs_dev):
        assert np.array_equal(o.cpu().numpy(), s)
        assert np.array_equal(od.cpu().numpy(), s)
