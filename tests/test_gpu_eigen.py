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
