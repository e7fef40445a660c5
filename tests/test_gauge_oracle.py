"""Pins the link-preprocessing oracle (unitary projection, 3-D stout smearing).

References: lattice/generator/elemental.py:107-117 (project_SU3),
:175-242 (numpy stout smearing).  Each check is independent of the
oracle's own formulas: scipy's matrix exponential, explicit per-site staple
loops, gauge covariance, the abelian closed form and the SVD polar factor.
"""

import numpy as np
import pytest
import scipy.linalg

from easydistillation_amd import synthetic as S
from oracle import gauge_oracle as G


def rand_traceless_hermitian(rng, n, scale):
    A = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    H = 0.5 * (A + np.conj(np.swapaxes(A, 1, 2)))
    H -= np.trace(H, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    return H * scale


@pytest.mark.parametrize("scale", [1e-4, 1e-2, 0.3, 1.0, 2.5])
def test_expi_matches_expm(scale):
    rng = np.random.default_rng(5)
    Q = rand_traceless_hermitian(rng, 40, scale)
    for sgn in (1, -1):  # -Q flips the sign of det Q (the parity branch)
        got = G.expi_traceless_hermitian(sgn * Q)
        want = np.stack([scipy.linalg.expm(1j * sgn * q) for q in Q])
        assert np.abs(got - want).max() < 1e-13


def test_expi_zero_is_identity():
    assert np.abs(G.expi_traceless_hermitian(np.zeros((2, 3, 3))) - np.eye(3)).max() == 0


def test_staples_match_brute_force():
    rng = np.random.default_rng(7)
    latt = [3, 4, 5, 2]
    U = S.gauge_np(rng, latt, "haar")
    U = U + 0.1 * (rng.normal(size=U.shape) + 1j * rng.normal(size=U.shape))  # generic matrices
    assert np.abs(G.staple_sum(U) - G.brute_force_staples(U)).max() < 1e-13


def test_stout_rho_zero_is_identity():
    rng = np.random.default_rng(1)
    U = S.gauge_np(rng, [3, 3, 4, 2], "haar")
    assert np.abs(G.stout_smear(U, 2, 0.0) - U).max() == 0


def test_stout_unitary_and_special():
    rng = np.random.default_rng(2)
    U = S.gauge_np(rng, [4, 3, 3, 2], "haar")
    Us = G.stout_smear(U, 3, 0.12)
    I = np.eye(3)
    assert np.abs(Us @ np.conj(np.swapaxes(Us, -1, -2)) - I).max() < 1e-13
    assert np.abs(np.linalg.det(Us) - 1).max() < 1e-13
    assert np.abs(Us - U).max() > 1e-2


def test_stout_gauge_covariance():
    rng = np.random.default_rng(4)
    latt = [3, 4, 5, 2]
    U = S.gauge_np(rng, latt, "haar")
    g = S.gauge_np(rng, latt, "haar")[0]  # g[t, z, y, x] in SU(3)
    Ug = np.empty_like(U)
    for mu in range(3):
        g_fwd = np.roll(g, -1, 3 - mu)  # g(x + mu) on [t, z, y, x]
        Ug[mu] = g @ U[mu] @ np.conj(np.swapaxes(g_fwd, -1, -2))
    Us, Ugs = G.stout_smear(U, 2, 0.1), G.stout_smear(Ug, 2, 0.1)
    for mu in range(3):
        g_fwd = np.roll(g, -1, 3 - mu)
        want = g @ Us[mu] @ np.conj(np.swapaxes(g_fwd, -1, -2))
        assert np.abs(Ugs[mu] - want).max() < 1e-12


def test_stout_abelian_closed_form():
    # diagonal links: Omega is diagonal, Q = Im(Omega_cc) - mean_c, exp(iQ) elementwise
    rng = np.random.default_rng(9)
    latt = [3, 4, 5, 2]
    Lx, Ly, Lz, Lt = latt
    phi = rng.uniform(-np.pi, np.pi, size=(3, Lt, Lz, Ly, Lx, 3))
    U = np.zeros((3, Lt, Lz, Ly, Lx, 3, 3), complex)
    for c in range(3):
        U[..., c, c] = np.exp(1j * phi[..., c])
    rho = 0.17
    C = G.brute_force_staples(U)
    om = rho * np.einsum("...cc->...c", C) * np.exp(-1j * phi)
    q = om.imag - om.imag.mean(axis=-1, keepdims=True)
    want = np.zeros_like(U)
    for c in range(3):
        want[..., c, c] = np.exp(1j * (q[..., c] + phi[..., c]))
    assert np.abs(G.stout_smear(U, 1, rho) - want).max() < 1e-13


def test_stout_raises_plaquette():
    rng = np.random.default_rng(12)
    U = S.gauge_np(rng, [4, 4, 4, 2], "haar")
    p0 = G.plaquette(U)
    p1 = G.plaquette(G.stout_smear(U, 1, 0.1))
    p2 = G.plaquette(G.stout_smear(U, 3, 0.1))
    assert p0 < p1 < p2


def test_project_matches_svd_polar():
    rng = np.random.default_rng(6)
    U0 = S.gauge_np(rng, [3, 3, 3, 2], "haar")
    M = U0 @ (np.eye(3) + 0.05 * (rng.normal(size=U0.shape) + 1j * rng.normal(size=U0.shape)))
    P, it = G.project_su3(M)
    W, _, Vh = np.linalg.svd(M)
    assert np.abs(P - W @ Vh).max() < 1e-13
    assert 2 <= it < 20
    # a unitary input is a fixed point
    P0, it0 = G.project_su3(P)
    assert it0 == 0 and np.abs(P0 - P).max() == 0


def test_project_raises_when_capped():
    rng = np.random.default_rng(8)
    M = S.gauge_np(rng, [3, 3, 3, 1], "haar") * 1.3
    with pytest.raises(RuntimeError):
        G.project_su3(M, max_iter=1)
