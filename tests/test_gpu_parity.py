"""Parity of the HIP path (through the C ABI) against the CPU oracle and fixtures.

Bars: FP64 form <= 1e-12 and mixed form <= 1e-6, both as the normwise error
max|dE| / max|E| of every (d, p) block (elementwise relative error is
ill-conditioned on near-zero elementals).
"""

import sys

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import _native
from easydistillation_amd import synthetic as S
from oracle import elemental_oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def blockwise(E, Eref):
    """max over (d, p) blocks of max|dE| / max|Eref| for arrays [t, D, P, Ne, Ne].

    A block whose reference maximum is below 1e-8 of the largest block of the
    same derivative vanishes by symmetry (e.g. a Nyquist momentum component
    along the derivative's direction); its relative error is 0/0, so it is
    measured against that derivative's largest block instead.
    """
    E = np.asarray(E)
    Eref = np.asarray(Eref)
    d = np.abs(E - Eref)
    # move t into the block: [t, D, P, Ne, Ne] -> [D, P, t*Ne*Ne]
    d = d.transpose(1, 2, 0, 3, 4).reshape(d.shape[1], d.shape[2], -1).max(-1)
    r = np.abs(Eref).transpose(1, 2, 0, 3, 4).reshape(d.shape[0], d.shape[1], -1).max(-1)
    row = r.max(1, keepdims=True)
    den = np.where(r >= 1e-8 * row, r, row)
    return float((d / np.maximum(den, 1e-300)).max())


def run(latt, Ne, num_nabla, moms, U, V, mode="fp64", hermitian=1, tbatch=0, dilution=None, stage2=None):
    g = ed.ElementalGenerator(
        latt, None, None, num_nabla, moms, dilution=dilution, is_blending=dilution is not None, Ne=Ne, mode=mode
    )
    g.set_option(_native.ED_OPT_HERMITIAN, hermitian)
    if stage2 is not None:
        g.set_option(_native.ED_OPT_MIXED_STAGE2, stage2)
    if tbatch:
        g.set_option(_native.ED_OPT_TBATCH, tbatch)
    g.load_arrays(U, V)
    return g.calc_all()


def load_fixture(golden_dir, name):
    z = np.load(f"{golden_dir}/{name}.npz")
    latt = [int(v) for v in z["latt"]]
    moms = [tuple(int(c) for c in m) for m in z["momenta"]]
    return latt, int(z["num_nabla"]), moms, z["U"], z["V"], z["E"].transpose(2, 0, 1, 3, 4)


@pytest.mark.parametrize("name", ["config_a", "ref_shape", "brute_force"])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_fixtures(golden_dir, name, mode):
    latt, nn, moms, U, V, Eref = load_fixture(golden_dir, name)
    say(f"fixture {name} mode={mode} latt={latt} Ne={V.shape[1]} nabla={nn} P={len(moms)}")
    E = run(latt, V.shape[1], nn, moms, U, V, mode)
    err = blockwise(E, Eref)
    say(f"  err={err:.3e}")
    assert err <= (FP64_TOL if mode == "fp64" else MIXED_TOL)


def test_hermitian_shortcut_is_exact(golden_dir):
    latt, nn, moms, U, V, Eref = load_fixture(golden_dir, "ref_shape")
    say("hermitian on/off")
    E1 = run(latt, V.shape[1], nn, moms, U, V, hermitian=1)
    E0 = run(latt, V.shape[1], nn, moms, U, V, hermitian=0)
    assert blockwise(E0, Eref) <= FP64_TOL
    assert blockwise(E1, E0) <= 1e-13


CASES = [
    # (latt, Ne, num_nabla, momenta, links)
    ([3, 5, 7, 2], 37, 2, S.momenta_by_norm(2), "haar"),
    ([6, 4, 2, 1], 1, 1, S.unit_momenta(), "haar"),
    ([5, 5, 5, 3], 16, 0, S.momenta_by_norm(3), "smooth"),
    ([4, 3, 3, 1], 7, 3, [(0, 0, 0), (1, 0, 0), (0, 0, -1)], "haar"),
    ([4, 4, 6, 2], 20, 2, S.momenta_by_norm(5), "haar"),  # 57 momenta: two momentum blocks
    ([4, 4, 4, 2], 17, 1, [(1, 0, 0), (1, 0, 0), (0, 1, 1), (-1, 0, 0)], "haar"),  # repeats, no p = 0
    ([4, 4, 4, 1], 9, 2, [(0, 0, 0), (0, 0, 0), (0, 2, 0)], "smooth"),  # p = 0 twice
    ([8, 8, 8, 2], 40, 2, S.momenta_by_norm(4), "haar"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_uneven_shapes_against_oracle(case):
    latt, Ne, nn, moms, kind = CASES[case]
    say(f"case {case}: latt={latt} Ne={Ne} nabla={nn} P={len(moms)} links={kind}")
    U, V = S.make_inputs_np(latt, Ne, kind, seed=100 + case)
    Eref = O.elemental(V, U, latt, nn, moms)
    E = run(latt, Ne, nn, moms, U, V, "fp64")
    err = blockwise(E, Eref)
    say(f"  fp64 err={err:.3e}")
    assert err <= FP64_TOL
    Em = run(latt, Ne, nn, moms, U, V, "mixed")
    errm = blockwise(Em, Eref)
    say(f"  mixed err={errm:.3e}")
    assert errm <= MIXED_TOL
    # below V3 = 4096 the mixed form runs the FP64 kernels by default; the
    # FP32 mixed kernel stays reachable and is checked on the same shapes
    Em32 = run(latt, Ne, nn, moms, U, V, "mixed", stage2=0)
    errm32 = blockwise(Em32, Eref)
    say(f"  mixed FP32 kernel err={errm32:.3e}")
    assert errm32 <= MIXED_TOL


def test_blending_against_oracle():
    latt, Ne = [4, 4, 4, 2], 10
    dil = ([8, 12], [4, 6])
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=21)
    moms = S.unit_momenta()
    W = O.blending_coeff(Ne, dil)
    Eref = O.elemental(V, U, latt, 2, moms, stocastic_coeff=W)
    say("blending")
    assert blockwise(run(latt, Ne, 2, moms, U, V, dilution=dil), Eref) <= FP64_TOL
    assert blockwise(run(latt, Ne, 2, moms, U, V, dilution=dil, hermitian=0), Eref) <= FP64_TOL


def test_calc_paths_agree():
    latt, Ne = [4, 6, 4, 5], 24
    U, V = S.make_inputs_np(latt, Ne, "smooth", seed=5)
    moms = S.momenta_by_norm(2)
    say("calc paths")
    g = ed.ElementalGenerator(latt, None, None, 2, moms, Ne=Ne)
    g.load_arrays(U, V)
    single = np.stack([g.calc(t) for t in range(latt[3])])
    g.set_option(_native.ED_OPT_TBATCH, 2)
    batched = g.calc_all()
    again = g.calc(3)
    assert np.array_equal(single, batched)
    assert np.array_equal(again, batched[3])
    Eref = O.elemental(V, U, latt, 2, moms)
    assert blockwise(batched, Eref) <= FP64_TOL


def test_calc_windows_any_order_and_reload():
    """calc(t) computes aligned windows of timeslices in one launch; any call
    order, a partial last window, a mode switch and new inputs all return the
    right timeslice."""
    latt, Ne, nn = [4, 5, 3, 11], 8, 1
    moms = S.momenta_by_norm(1)
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=15)
    U2, V2 = S.make_inputs_np(latt, Ne, "smooth", seed=16)
    ref = O.elemental(V, U, latt, nn, moms)
    ref2 = O.elemental(V2, U2, latt, nn, moms)
    g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne)
    g.WINDOW = 4  # windows [0, 4), [4, 8), [8, 11)
    g.load_arrays(U, V)
    for t in [7, 0, 10, 3, 8, 4, 9, 10]:
        E = g.calc(t)
        assert E.shape == ref.shape[1:] and E.flags.c_contiguous
        assert blockwise(E[None], ref[t][None]) <= FP64_TOL, t
    a, b = g.calc(5), g.calc(5)
    a[:] = 0  # fresh arrays: one does not alias the other
    assert np.array_equal(b, g.calc(5)) and not np.array_equal(a, b)
    assert blockwise(g.calc(2, mode="mixed")[None], ref[2][None]) <= MIXED_TOL
    g.load_arrays(U2, V2)
    for t in [6, 1]:
        assert blockwise(g.calc(t)[None], ref2[t][None]) <= FP64_TOL, t


def test_calc_loop_overlaps_and_matches_calc_range():
    """The reference driver's loop, for t: data[t] = calc(t), with the next
    window computing and the next timeslice copying to page-locked memory
    while the host takes the current one: every timeslice equals calc_range's
    bit for bit, also across an option change in the middle of a window (the
    work in flight is waited for) and with the "torch" backend."""
    import torch

    from easydistillation_amd.backend import get_backend, set_backend

    latt, Ne, nn = [8, 8, 8, 11], 16, 2
    moms = S.momenta_by_norm(2)
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=17)
    Lt = latt[3]
    for mode in ("fp64", "mixed"):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.WINDOW = 4  # windows [0, 4), [4, 8), [8, 11)
        g.load_arrays(U, V)
        ref = g.calc_range(0, Lt)
        data = np.zeros_like(ref)
        for t in range(Lt):
            data[t] = g.calc(t)
            if t == 5:
                g.set_option(_native.ED_OPT_HERMITIAN, 1)  # the default again: same results
        say(f"calc loop {mode}: equal to calc_range {np.array_equal(data, ref)}")
        assert np.array_equal(data, ref)
        for t in [9, 10, 2, 3, 4]:
            assert np.array_equal(g.calc(t), ref[t]), t
        before = get_backend()
        set_backend("torch")
        try:
            dd = torch.zeros(ref.shape, dtype=torch.complex128, device="cuda")
            for t in range(Lt):
                dd[t] = g.calc(t)
            assert np.array_equal(dd.cpu().numpy(), ref)
        finally:
            set_backend(before)


def test_load_from_containers(tmp_path):
    latt = [4, 4, 4, 3]
    Lx, Ly, Lz, Lt = latt
    U, V = S.make_inputs_np(latt, 12, "haar", seed=31)
    full = np.zeros((Lt, Lz, Ly, Lx, 4, 3, 3), complex)
    full[..., :3, :, :] = U.transpose(1, 2, 3, 4, 0, 5, 6)
    full[..., 3, :, :] = np.eye(3)
    np.save(tmp_path / "c0.gauge.npy", full)
    Vtot = np.concatenate([V, V[:, :2]], axis=1).astype(np.complex128)  # file holds more vectors than used
    np.save(tmp_path / "c0.evec.npy", Vtot)
    gauge = ed.GaugeFieldNpy(f"{tmp_path}/", ".gauge.npy", list(full.shape))
    evec = ed.EigenvectorNpy(f"{tmp_path}/", ".evec.npy", list(Vtot.shape), 12)
    moms = S.unit_momenta()
    say("containers")
    g = ed.ElementalGenerator(latt, gauge, evec, 1, moms)
    g.load("c0")
    E = np.stack([g.calc(t) for t in range(Lt)])
    assert blockwise(E, O.elemental(V, U, latt, 1, moms)) <= FP64_TOL


def test_device_inputs_and_outputs():
    import torch

    latt, Ne = [6, 4, 4, 2], 20
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=41)
    moms = S.momenta_by_norm(2)
    say("device tensors")
    ed.set_backend("torch")
    try:
        g = ed.ElementalGenerator(latt, None, None, 2, moms, Ne=Ne)
        g.load_arrays(torch.from_numpy(U).cuda(), torch.from_numpy(V).cuda())
        E = g.calc_all()
        assert E.is_cuda and E.dtype == torch.complex128
        E = E.cpu().numpy()
    finally:
        ed.set_backend("hip")
    assert blockwise(E, O.elemental(V, U, latt, 2, moms)) <= FP64_TOL


def test_config_b_shape_properties():
    """16^3, Ne = 100, 33 momenta, 13 derivatives: invariants at the real per-t shape."""
    import torch

    latt, Ne = [16, 16, 16, 2], 100
    moms = S.momenta_by_norm(4)
    say("config B shape: inputs")
    U = S.gauge_torch(latt, "smooth", 0.3, seed=3, device="cuda")
    V = S.eigvecs_torch(latt, Ne, seed=4, device="cuda")
    g = ed.ElementalGenerator(latt, None, None, 2, moms, Ne=Ne)
    g.load_arrays(U, V)
    say("config B shape: fp64")
    E = g.calc_all()
    say("config B shape: mixed")
    Em = g.calc_all(mode="mixed")
    idx = {m: i for i, m in enumerate(moms)}
    # V^dag V = 1 to complex64 rounding
    assert np.abs(E[:, 0, idx[(0, 0, 0)]] - np.eye(Ne)).max() < 1e-5
    # E_{rev d}[p] = (-1)^{|d|} E_d[-p]^dagger
    dl = O.derivative_list(2)
    for di, d in enumerate(dl):
        rd = dl.index(tuple(reversed(d)))
        for m in [(0, 0, 0), (1, 0, 0), (1, 1, -1), (0, 2, 0)]:
            mm = tuple(-c for c in m)
            lhs = E[:, rd, idx[m]]
            rhs = (-1) ** len(d) * np.conj(np.swapaxes(E[:, di, idx[mm]], -1, -2))
            assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(lhs).max()
    err = blockwise(Em, E)
    say(f"  mixed vs fp64 err={err:.3e}")
    assert err <= MIXED_TOL
    # the tuner's choice at this shape: the FP16 split sum (unit plan)
    g.set_option(_native.ED_OPT_MIXED_WS, 2)
    Eh = g.calc_all(mode="mixed")
    assert g.context.query(_native.ED_Q_MIXED_KERNEL) == 5
    err16 = blockwise(Eh, E)
    say(f"  FP16 split sum vs fp64 err={err16:.3e}")
    assert err16 <= MIXED_TOL
    # one timeslice against the oracle
    say("config B shape: oracle t=0")
    Uh, Vh = U.cpu().numpy(), V.cpu().numpy()
    sub = [(0, 0, 0), (1, 0, 0), (0, -1, 1), (-1, 1, -1), (0, 0, -2)]
    Eref0 = O.elemental_timeslice(Vh[0], Uh[:, 0], latt, 2, sub)
    got = E[:1][:, :, [idx[m] for m in sub]]
    assert blockwise(got, Eref0[None]) <= FP64_TOL


@pytest.mark.parametrize("stage2", [0, 1])
def test_mixed_stage2_variants(stage2):
    """Both phase-contraction variants of the mixed form (FP32 MFMA, FP16 split) at 16^3."""
    latt, Ne = [16, 16, 16, 1], 24
    moms = S.momenta_by_norm(4)
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=77)
    say(f"mixed stage-2 variant {stage2}")
    g = ed.ElementalGenerator(latt, None, None, 2, moms, Ne=Ne, mode="mixed")
    g.set_option(_native.ED_OPT_MIXED_STAGE2, stage2)
    g.load_arrays(U, V)
    E = g.calc_all()
    err = blockwise(E, O.elemental(V, U, latt, 2, moms))
    say(f"  err={err:.3e}")
    assert err <= MIXED_TOL


WS_CASES = [
    # latt, Ne, num_nabla, momenta, links: padded sites (V3 % 32 != 0),
    # Ne % 16 != 0, two momentum blocks, every split count up to 12 (split,
    # colour) terms, rough and smooth links at V3 >= 4096 (where the FP16
    # stage-1 kernel runs)
    ([16, 16, 16, 2], 24, 2, S.momenta_by_norm(4), "haar"),
    ([5, 7, 9, 2], 37, 2, S.momenta_by_norm(5)[:57], "haar"),
    ([6, 6, 6, 3], 16, 1, S.unit_momenta(), "haar"),
    ([9, 4, 3, 1], 9, 0, [(1, 0, 0), (0, 0, 0), (-1, 0, 0)], "haar"),
    ([16, 16, 24, 1], 40, 2, S.momenta_by_norm(3), "smooth"),
]


@pytest.mark.parametrize("case", range(len(WS_CASES)))
def test_mixed_warp_specialized_matches_uniform(case):
    """The warp-specialized FP16 stage-2 kernels (FP16 stage 1 with hi/lo
    workspace pairs; FP32 stage 1) and the uniform kernel (per-tile scale),
    each against the oracle, and against each other within twice the bound."""
    latt, Ne, nn, moms, links = WS_CASES[case]
    U, V = S.make_inputs_np(latt, Ne, links, seed=90 + case)
    say(f"ws vs uniform: latt={latt} Ne={Ne} nabla={nn} P={len(moms)} links={links}")
    outs = {}
    for ws in (2, 1, 0):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode="mixed")
        g.set_option(_native.ED_OPT_MIXED_STAGE2, 1)
        g.set_option(_native.ED_OPT_MIXED_WS, ws)
        g.load_arrays(U, V)
        outs[ws] = g.calc_all()
    ref = O.elemental(V, U, latt, nn, moms)
    errs = {ws: blockwise(outs[ws], ref) for ws in outs}
    say("  vs oracle: " + ", ".join(f"kind {ws} {e:.3e}" for ws, e in errs.items()))
    for ws in (2, 1):
        d = blockwise(outs[ws], outs[0])
        say(f"  kind {ws} vs uniform blockwise {d:.2e}")
        assert d <= 2 * MIXED_TOL
    assert max(errs.values()) <= MIXED_TOL


@pytest.mark.parametrize("case", [0, 4])
def test_fp16_split_sum_unit_plan(case):
    """The FP16 split-sum kernel (ED_OPT_MIXED_WS 2) on its unit plan (the
    workspace holds i^k times each level-k derivative slot, every split
    coefficient becomes +1 with its multiplicity as repeated terms, results
    are multiplied by i^|d|) and on the scaled plan (coefficients multiply the
    operands): each against the oracle, and against each other."""
    latt, Ne, nn, moms, links = WS_CASES[case]
    U, V = S.make_inputs_np(latt, Ne, links, seed=90 + case)
    say(f"unit plan: latt={latt} Ne={Ne} nabla={nn} P={len(moms)} links={links}")
    outs = {}
    for unit in (1, 0):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode="mixed")
        g.set_option(_native.ED_OPT_MIXED_STAGE2, 1)
        g.set_option(_native.ED_OPT_MIXED_WS, 2)
        g.set_option(_native.ED_OPT_H16_UNIT, unit)
        assert g.context.get_option(_native.ED_OPT_H16_UNIT) == unit
        g.load_arrays(U, V)
        assert g.context.query(_native.ED_Q_MIXED_KERNEL) == 5
        outs[unit] = g.calc_all()
    ref = O.elemental(V, U, latt, nn, moms)
    errs = {unit: blockwise(outs[unit], ref) for unit in outs}
    d = blockwise(outs[1], outs[0])
    say(f"  vs oracle: unit {errs[1]:.3e}, scaled {errs[0]:.3e}; unit vs scaled {d:.2e}")
    assert max(errs.values()) <= MIXED_TOL
    assert d <= 2 * MIXED_TOL


LDS_HOP_CASES = [
    # latt, Ne, num_nabla, links, (mode, ED_OPT_MIXED_WS): 16-byte (FP64), 8-byte complex64
    # (FP32 split sum) and FP16 hi/lo pair workspaces; Lz = 3, the smallest z ring; Ne not a
    # multiple of the eigenvector chunk; a lattice that does not tile (falls back)
    ([16, 16, 16, 2], 40, 2, "haar", [("fp64", None), ("mixed", 1), ("mixed", 2)]),
    ([8, 8, 3, 2], 24, 2, "haar", [("fp64", None)]),
    ([16, 8, 32, 1], 37, 1, "smooth", [("fp64", None), ("mixed", 2)]),
    ([6, 8, 8, 1], 20, 2, "haar", [("fp64", None)]),
    # three derivatives (level 3: nine parents) and a single eigenvector
    ([8, 8, 4, 1], 9, 3, "haar", [("fp64", None), ("mixed", 2)]),
    ([8, 8, 3, 1], 1, 2, "smooth", [("fp64", None)]),
]


@pytest.mark.parametrize("case", range(len(LDS_HOP_CASES)))
def test_lds_staged_hop_is_bit_equal(case):
    """The LDS-staged hop (ED_OPT_HOP_LDS 1: (x, y) tile with halo and a ring of
    three z planes in LDS) writes the workspace the one-wave-per-site hop
    (ED_OPT_HOP_LDS 0) writes, bit for bit, so every elemental is equal; so
    does the default (-1), which picks one of them by element size."""
    latt, Ne, nn, links, forms = LDS_HOP_CASES[case]
    moms = S.momenta_by_norm(2)
    U, V = S.make_inputs_np(latt, Ne, links, seed=120 + case)
    for mode, ws in forms:
        outs = {}
        for lds in (0, 1, -1):
            g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
            if ws is not None:
                g.set_option(_native.ED_OPT_MIXED_STAGE2, 1)
                g.set_option(_native.ED_OPT_MIXED_WS, ws)
            g.set_option(_native.ED_OPT_HOP_LDS, lds)
            g.load_arrays(U, V)
            outs[lds] = g.calc_all()
        same = all(np.array_equal(outs[0], outs[v]) for v in (1, -1))
        say(f"LDS hop: latt={latt} Ne={Ne} nabla={nn} {mode} ws={ws}: bit-equal {same}")
        assert same
        if mode == "fp64" and np.prod(latt) <= 1024:
            err = blockwise(outs[1], O.elemental(V, U, latt, nn, moms))
            say(f"  fp64 vs oracle {err:.2e}")
            assert err <= FP64_TOL


def test_load_from_ildg_and_qdp_files(tmp_path):
    """load(key) through the reference's on-disk formats: ILDG/LIME links and
    QDP LazyDiskMapObj eigenvectors (elemental.py:102-105, preset.py:55-64,151-160)."""
    from easydistillation_amd.filedata import write_ildg, write_qdp_lazy

    latt, Ne, nn = [4, 5, 3, 2], 7, 1
    Lx, Ly, Lz, Lt = latt
    moms = S.unit_momenta()
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=5)
    full = np.empty((Lt, Lz, Ly, Lx, 4, 3, 3), np.complex128)
    full[..., :3, :, :] = U.transpose(1, 2, 3, 4, 0, 5, 6)
    full[..., 3, :, :] = np.eye(3)
    write_ildg(str(tmp_path / "cfg.lime"), full)
    write_qdp_lazy(str(tmp_path / "cfg.mod"), {(t, e): V[t, e] for t in range(Lt) for e in range(Ne)}, latt)
    gauge = ed.GaugeFieldIldg(str(tmp_path) + "/", ".lime", [Lt, Lz, Ly, Lx, 4, 3, 3])
    eig = ed.EigenvectorTimeSlice(str(tmp_path) + "/", ".mod", [Lt, Ne, Lz, Ly, Lx, 3], Ne)
    g = ed.ElementalGenerator(latt, gauge, eig, nn, moms)
    g.load("cfg")
    E = g.calc_all()
    err = blockwise(E, O.elemental(V, U, latt, nn, moms))
    say(f"ILDG + QDP containers: err={err:.3e}")
    assert err <= FP64_TOL


def test_load_from_gauge_timeslice_and_binary_files(tmp_path):
    """load(key) through the other two gauge containers of the reference:
    QDP LazyDiskMapObj records per (t, mu) (GaugeFieldTimeSlice, preset.py:44-53;
    read back in complex64, so the oracle gets the same rounded links) and a
    headerless binary array of interleaved float64 (GaugeFieldBinary,
    preset.py:162-170), with npy eigenvectors."""
    from easydistillation_amd.filedata import write_binary, write_qdp_lazy

    latt, Ne, nn = [4, 3, 5, 2], 6, 2
    Lx, Ly, Lz, Lt = latt
    V3 = Lx * Ly * Lz
    moms = S.momenta_by_norm(2)
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=9)
    full = np.empty((Lt, Lz, Ly, Lx, 4, 3, 3), np.complex128)
    full[..., :3, :, :] = U.transpose(1, 2, 3, 4, 0, 5, 6)
    full[..., 3, :, :] = np.eye(3)
    per_t_mu = full.reshape(Lt, V3, 4, 3, 3).transpose(0, 2, 1, 3, 4)
    write_qdp_lazy(str(tmp_path / "cfg.mod"), {(t, mu): per_t_mu[t, mu] for t in range(Lt) for mu in range(4)},
                   latt, dtype=">c16")
    inter = np.stack([full.real, full.imag], -1).reshape(Lt, V3, 4, 3, 3, 2)
    write_binary(str(tmp_path / "cfg.dat"), inter, "<f8")
    np.save(tmp_path / "cfg.npy", V)
    eig = ed.EigenvectorNpy(str(tmp_path) + "/", ".npy", None, Ne)
    cases = [
        ("timeslice", ed.GaugeFieldTimeSlice(str(tmp_path) + "/", ".mod", [Lt, 4, V3, 3, 3]),
         U.astype(np.complex64).astype(np.complex128)),
        ("binary", ed.GaugeFieldBinary(str(tmp_path) + "/", ".dat", [Lt, V3, 4, 3, 3, 2]), U),
    ]
    for name, gauge, Uref in cases:
        g = ed.ElementalGenerator(latt, gauge, eig, nn, moms)
        g.load("cfg")
        err = blockwise(g.calc_all(), O.elemental(V, Uref, latt, nn, moms))
        say(f"{name} gauge container: err={err:.3e}")
        assert err <= FP64_TOL


def test_mixed_below_4096_sites_runs_fp64_kernels():
    """Below V3 = 4096 the mixed form's default is the FP64 kernels (bit-identical to mode
    "fp64"); an explicit ED_OPT_MIXED_STAGE2 selects the mixed kernels again."""
    latt, Ne, nn = [6, 5, 7, 2], 12, 2
    moms = S.momenta_by_norm(2)
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=5)
    say("mixed routing below V3 = 4096")
    E64 = run(latt, Ne, nn, moms, U, V, "fp64")
    Em = run(latt, Ne, nn, moms, U, V, "mixed")
    assert np.array_equal(E64, Em)
    Em32 = run(latt, Ne, nn, moms, U, V, "mixed", stage2=0)
    assert not np.array_equal(E64, Em32)
    # the forced FP32 kernel on a lattice this small sits near the mixed bound
    # (the reason for the routing); here it only has to be the mixed kernel
    assert blockwise(Em32, E64) <= 1e-5


def test_mixed_kernel_query_follows_the_routing():
    """ED_Q_MIXED_KERNEL names the contraction the mixed form runs for the settings."""
    U, V = S.make_inputs_np([6, 5, 7, 1], 4, "haar", seed=2)
    cases = [
        ([6, 5, 7, 1], {}, 0),
        ([6, 5, 7, 1], {_native.ED_OPT_MIXED_STAGE2: 0}, 1),
        ([6, 5, 7, 1], {_native.ED_OPT_MIXED_STAGE2: 1, _native.ED_OPT_MIXED_WS: 0}, 2),
        ([16, 16, 16, 1], {}, 3),
        ([16, 16, 16, 1], {_native.ED_OPT_WIDE_WS: 1}, 4),
        ([16, 16, 16, 1], {_native.ED_OPT_MIXED_WS: 2}, 5),
    ]
    for latt, opts, want in cases:
        g = ed.ElementalGenerator(latt, None, None, 2, S.momenta_by_norm(2), Ne=4, mode="mixed")
        for k, v in opts.items():
            g.set_option(k, v)
        Ui, Vi = (U, V) if latt[0] == 6 else S.make_inputs_np(latt, 4, "haar", seed=3)
        g.load_arrays(Ui, Vi)
        got = g.context.query(_native.ED_Q_MIXED_KERNEL)
        say(f"mixed kernel for {latt} {opts}: {got}")
        assert got == want


def test_orthonormality_gate():
    """ed_set_eigvecs measures max|V^dag V - 1| in FP64 on what it stored (the
    complex64-rounded vectors); ED_OPT_ORTHO_GATE turns it into a refusal."""
    latt, Ne = [6, 5, 4, 3], 9
    U, V = S.make_inputs_np(latt, Ne, "haar", seed=12)
    g = ed.ElementalGenerator(latt, None, None, 1, [(0, 0, 0)], Ne=Ne)
    g.load_arrays(U, V)
    Vc = V.astype(np.complex64).astype(np.complex128).reshape(latt[3], Ne, -1)
    want = max(np.abs(Vc[t].conj() @ Vc[t].T - np.eye(Ne)).max() for t in range(latt[3]))
    got = g.context.orthonormality()
    say(f"orthonormal V: device {got:.3e}, host {want:.3e}")
    # FP64 sums in another order: a few ulps of the unit diagonal
    assert got < 1e-6 and abs(got - want) <= 1e-14
    # one eigenvector of the last timeslice scaled by 1.01: |V^dag V - 1| = 0.0201 there
    Vb = V.copy()
    Vb[2, 4] *= 1.01
    Vbc = Vb.astype(np.complex64).astype(np.complex128).reshape(latt[3], Ne, -1)
    want = max(np.abs(Vbc[t].conj() @ Vbc[t].T - np.eye(Ne)).max() for t in range(latt[3]))
    g.load_arrays(U, Vb)
    got = g.context.orthonormality()
    say(f"scaled eigenvector: device {got:.6e}, host {want:.6e}")
    assert abs(got - want) <= 1e-12 and abs(got - 0.0201) < 1e-5
    g.set_option(_native.ED_OPT_ORTHO_GATE, 6)
    with pytest.raises(_native.NativeError) as e:
        g.load_arrays(U, Vb)
    assert e.value.code == _native.ED_ERR_ARG
    with pytest.raises(_native.NativeError) as e:
        g.calc_all()
    assert e.value.code == _native.ED_ERR_STATE
    g.load_arrays(U, V)  # orthonormal again: accepted
    E = g.calc_all()
    assert np.abs(E[:, 0, 0] - np.eye(Ne)).max() < 1e-6
