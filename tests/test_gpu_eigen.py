"""Laplacian and its low modes on the device (ed_laplacian, EigenvectorGenerator)
against the oracle (lattice/generator/eigenvector.py:11-26,234-257), and the
device eigenvectors fed straight into elemental generation."""

import sys

import numpy as np
import pytest
import torch

import easydistillation_amd as ed
from easydistillation_amd import synthetic as S
from easydistillation_amd.eigen import EigenvectorGenerator
from oracle import eigen_oracle as EO
from oracle import elemental_oracle as O
from oracle import gauge_oracle as G

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_laplacian_matches_oracle_host_and_device():
    latt = [4, 5, 6, 2]
    Lx, Ly, Lz, Lt = latt
    rng = np.random.default_rng(1)
    U = S.gauge_np(rng, latt, "haar")
    F = rng.normal(size=(5, Lz, Ly, Lx, 3)) + 1j * rng.normal(size=(5, Lz, Ly, Lx, 3))
    gen = EigenvectorGenerator(latt, None, 4)
    gen.load_links(U)
    for t in range(Lt):
        want = np.moveaxis(EO.laplacian(np.moveaxis(F, 0, -1), U[:, t]), -1, 0)
        out = np.empty_like(F)
        gen._ctx.laplacian(t, np.ascontiguousarray(F).ctypes.data, out.ctypes.data, 5, False)
        dev = gen.laplacian(t, torch.from_numpy(F).cuda()).cpu().numpy()
        assert np.abs(out - want).max() < 1e-13 and np.abs(dev - want).max() < 1e-13


@pytest.mark.parametrize("latt,Ne,kind", [([4, 4, 4, 2], 8, "haar"), ([6, 5, 4, 1], 12, "smooth")])
def test_low_modes_match_oracle(latt, Ne, kind):
    rng = np.random.default_rng(7)
    U = S.gauge_np(rng, latt, kind)
    gen = EigenvectorGenerator(latt, None, Ne, tol=1e-11)
    gen.load_links(U)
    for t in range(latt[3]):
        vecs, vals = gen.calc(t)
        rv, rw = EO.lowest_modes(U[:, t], Ne, tol=1e-12)
        dv = np.abs(vecs.reshape(Ne, -1) - rv.reshape(Ne, -1)).max()
        say(f"{latt} {kind} t={t}: {gen.iterations} iterations, |dlambda| {np.abs(vals - rw).max():.1e}, |dv| {dv:.1e}")
        assert np.abs(vals - rw).max() < 1e-9
        assert dv < 1e-6


def test_device_eigenvectors_feed_elementals():
    latt, Ne, nn = [4, 4, 5, 2], 6, 1
    rng = np.random.default_rng(9)
    U = S.gauge_np(rng, latt, "haar")
    gen = EigenvectorGenerator(latt, None, Ne, tol=1e-11)
    gen.load_links(U)
    gen.stout_smear(2, 0.1)
    V, vals = gen.calc_all()
    assert V.is_cuda and V.shape == (latt[3], Ne, 5, 4, 4, 3)
    Us = G.stout_smear(U, 2, 0.1)
    for t in range(latt[3]):
        _, rw = EO.lowest_modes(Us[:, t], Ne, tol=1e-12)
        assert np.abs(vals[t] - rw).max() < 1e-9
    moms = S.unit_momenta()
    g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne)
    g.load_arrays(torch.from_numpy(Us).cuda(), V)
    E = g.calc_all()
    Vn = V.cpu().numpy()
    ref = O.elemental(Vn, Us, latt, nn, moms)
    assert np.abs(E - ref).max() / np.abs(ref).max() < 1e-12
    # V^dag V = 1 through the p = 0, d = 0 elemental
    assert np.abs(E[:, 0, 0] - np.eye(Ne)).max() < 1e-6


@pytest.mark.parametrize("kind", ["smooth", "haar"])
def test_mixed_form_on_laplacian_modes_16c(kind):
    """The mixed kernels (they run from V3 = 4096 on) fed with Laplacian low modes
    from the device eigensolver at 16^3, Nev 48: smooth low modes give a large
    coherent site-independent part of Q, the case the chunk-offset subtraction
    of the FP16 phase contraction is for.  Eigenvalues against scipy eigsh,
    residuals against the oracle Laplacian, then mixed elementals within 1e-6
    of the oracle for every (d, p) at t = 0 and of the FP64 form at every t."""
    from easydistillation_amd import _native

    latt, Ne, nn = [16, 16, 16, 2], 48, 2
    rng = np.random.default_rng(21)
    U = S.gauge_np(rng, latt, kind)
    gen = EigenvectorGenerator(latt, None, Ne, tol=1e-10)
    gen.load_links(U)
    V, vals = gen.calc_all()
    _, rw = EO.lowest_modes(U[:, 0], Ne, tol=1e-11)
    dl = np.abs(vals[0] - rw).max()
    Vn = V.cpu().numpy()
    res = max(
        np.abs(np.moveaxis(EO.laplacian(np.moveaxis(Vn[t], 0, -1), U[:, t]), -1, 0)
               - vals[t][:, None, None, None, None] * Vn[t]).max()
        for t in range(latt[3])
    )
    say(f"16^3 {kind}: eigenvalues {vals[0][0]:.4f}..{vals[0][-1]:.4f}, |dlambda| {dl:.1e}, residual {res:.1e}")
    assert dl < 1e-9 and res < 1e-8

    moms = S.momenta_by_norm(2)
    out = {}
    for mode in ("fp64", "mixed"):
        g = ed.ElementalGenerator(latt, None, None, nn, moms, Ne=Ne, mode=mode)
        g.load_arrays(torch.from_numpy(U).cuda(), V)
        if mode == "mixed":
            assert g.context.query(_native.ED_Q_MIXED_KERNEL) == 3
            assert g.context.orthonormality() < 1e-6
        out[mode] = g.calc_all()
    ref = O.elemental_timeslice(Vn[0], U[:, 0], latt, nn, moms)

    def blockwise(E, R):
        d = np.abs(E - R).reshape(R.shape[0], R.shape[1], -1).max(-1)
        r = np.abs(R).reshape(R.shape[0], R.shape[1], -1).max(-1)
        row = r.max(1, keepdims=True)
        return float((d / np.where(r >= 1e-8 * row, r, row)).max())

    e64 = blockwise(out["fp64"][0], ref)
    emx = blockwise(out["mixed"][0], ref)
    emf = max(blockwise(out["mixed"][t], out["fp64"][t]) for t in range(latt[3]))
    say(f"16^3 {kind}: vs oracle fp64 {e64:.2e} mixed {emx:.2e}; mixed vs fp64 {emf:.2e}")
    assert e64 <= 1e-12 and emx <= 1e-6 and emf <= 1e-6
