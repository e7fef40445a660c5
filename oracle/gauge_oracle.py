"""CPU oracle for the link preprocessing of elemental generation.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the package
``easydistillation_amd``) may import this module; it is imported only by
``tests/`` and by ``__graft_entry__.smoke()`` as the checker.

numpy (complex128) restatement of:

* ``project_su3``   follows ``lattice/generator/elemental.py:107-117``: the
                    polar iteration U <- (U + (U^-1)^dagger)/2 over the whole
                    array until max|U - (U^-1)^dagger| <= 1e-15 and
                    max|U U^dagger - 1| <= 1e-15 (``numpy.linalg.inv``, as
                    the reference).
* ``stout_smear``   follows the behaviour of
                    ``lattice/generator/elemental.py:175-242`` (3-D stout
                    smearing of the spatial links, staples inside the link's
                    timeslice), written from Morningstar and Peardon,
                    Phys. Rev. D 69 (2004) 054501, eqs. 1-3 (staples, Omega,
                    Q) and 29-35 (exp(iQ) by Cayley-Hamilton).
* ``expi_traceless_hermitian`` the Cayley-Hamilton exponential alone.

Pinning (tests/test_gauge_oracle.py): ``expi_traceless_hermitian`` against
``scipy.linalg.expm`` on random traceless Hermitian matrices of both signs of
det Q and small |Q|; ``staple_sum`` against ``brute_force_staples``
(explicit per-site loops with index arithmetic, no ``roll``);
``stout_smear`` by rho = 0 identity, unitarity / det = 1, gauge covariance
under random g(x) in SU(3), and the abelian closed form (diagonal links:
exp(iQ) is the elementwise exponential of the traceless imaginary part of
Omega); ``project_su3`` against the SVD polar factor W V^dagger.

Layout: U[3, Lt, Lz, Ly, Lx, 3, 3] (mu = x, y, z), U[..][a][b] acting on
colour b -- the array of elemental.py:103.
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "project_su3",
    "staple_sum",
    "brute_force_staples",
    "expi_traceless_hermitian",
    "stout_smear",
    "plaquette",
]


def _dag(M):
    return np.conj(np.swapaxes(M, -1, -2))


def project_su3(U: np.ndarray, max_iter: int = 100):
    """(projected U, number of updates); elemental.py:107-117."""
    U = np.array(U, dtype=np.complex128)
    eye = np.eye(3)
    for it in range(max_iter + 1):
        Uinv = np.linalg.inv(U)
        bad = np.max(np.abs(U - _dag(Uinv))) > 1e-15 or np.max(np.abs(U @ _dag(U) - eye)) > 1e-15
        if not bad:
            return U, it
        if it == max_iter:
            raise RuntimeError(f"unitary projection did not converge in {max_iter} iterations")
        U = 0.5 * (U + _dag(Uinv))
    raise AssertionError("unreachable")


def _shift(A, mu, sgn):
    """A(x + sgn*mu) for A[..., Lz, Ly, Lx, 3, 3] with mu = 0 (x), 1 (y), 2 (z)."""
    axis = A.ndim - 3 - mu
    return np.roll(A, -sgn, axis=axis)


def staple_sum(U: np.ndarray) -> np.ndarray:
    """C_mu(x) = sum_{nu != mu} [U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag
    + U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu)], mu, nu spatial (elemental.py:181-191)."""
    C = np.zeros_like(U)
    for mu in range(3):
        for nu in range(3):
            if nu == mu:
                continue
            C[mu] += U[nu] @ _shift(U[mu], nu, +1) @ _dag(_shift(U[nu], mu, +1))
            Um = _shift(U[nu], nu, -1)
            C[mu] += _dag(Um) @ _shift(U[mu], nu, -1) @ _shift(Um, mu, +1)
    return C


def brute_force_staples(U: np.ndarray) -> np.ndarray:
    """The same staple sum by explicit site loops and index arithmetic (tiny lattices)."""
    _, Lt, Lz, Ly, Lx = U.shape[:5]
    L = (Lx, Ly, Lz)
    C = np.zeros_like(U)

    def link(mu, t, x):
        return U[mu, t, x[2] % Lz, x[1] % Ly, x[0] % Lx]

    for t in range(Lt):
        for z in range(Lz):
            for y in range(Ly):
                for x in range(Lx):
                    site = np.array([x, y, z])
                    for mu in range(3):
                        emu = np.eye(3, dtype=int)[mu]
                        acc = np.zeros((3, 3), complex)
                        for nu in range(3):
                            if nu == mu:
                                continue
                            enu = np.eye(3, dtype=int)[nu]
                            acc += link(nu, t, site) @ link(mu, t, site + enu) @ link(nu, t, site + emu).conj().T
                            acc += (
                                link(nu, t, site - enu).conj().T
                                @ link(mu, t, site - enu)
                                @ link(nu, t, site - enu + emu)
                            )
                        C[mu, t, z, y, x] = acc
    del L
    return C


def expi_traceless_hermitian(Q: np.ndarray) -> np.ndarray:
    """exp(iQ) for traceless Hermitian Q[..., 3, 3] (Morningstar-Peardon eqs. 29-35)."""
    Q = np.asarray(Q, dtype=np.complex128)
    Q2 = Q @ Q
    c0 = np.real(np.trace(Q @ Q2, axis1=-2, axis2=-1)) / 3.0
    c1 = np.real(np.trace(Q2, axis1=-2, axis2=-1)) / 2.0
    zero = ~(c1 > 0)
    c1s = np.where(zero, 1.0, c1)
    neg = c0 < 0
    c0max = 2.0 * (c1s / 3.0) ** 1.5
    theta = np.arccos(np.minimum(np.abs(c0) / c0max, 1.0))
    u = np.sqrt(c1s / 3.0) * np.cos(theta / 3.0)
    w = np.sqrt(c1s) * np.sin(theta / 3.0)
    u2, w2 = u * u, w * w
    small = np.abs(w) <= 0.05
    xi0 = np.where(small, 1 - w2 / 6 * (1 - w2 / 20 * (1 - w2 / 42 * (1 - w2 / 72))), np.sin(w) / np.where(small, 1, w))
    em, e2 = np.exp(-1j * u), np.exp(2j * u)
    cw = np.cos(w)
    den = 1.0 / (9 * u2 - w2)
    f0 = ((u2 - w2) * e2 + em * (8 * u2 * cw + 2j * u * (3 * u2 + w2) * xi0)) * den
    f1 = (2 * u * e2 - em * (2 * u * cw - 1j * (3 * u2 - w2) * xi0)) * den
    f2 = (e2 - em * (cw + 3j * u * xi0)) * den
    f0 = np.where(neg, np.conj(f0), f0)
    f1 = np.where(neg, -np.conj(f1), f1)
    f2 = np.where(neg, np.conj(f2), f2)
    f0 = np.where(zero, 1.0, f0)
    f1 = np.where(zero, 0.0, f1)
    f2 = np.where(zero, 0.0, f2)
    eye = np.eye(3)
    return f0[..., None, None] * eye + f1[..., None, None] * Q + f2[..., None, None] * Q2


def stout_smear(U: np.ndarray, nstep: int, rho: float) -> np.ndarray:
    """nstep 3-D stout steps U <- exp(iQ) U (elemental.py:175-242)."""
    U = np.array(U, dtype=np.complex128)
    for _ in range(nstep):
        Om = rho * staple_sum(U) @ _dag(U)
        Q = 0.5j * (_dag(Om) - Om)
        tr = np.trace(Q, axis1=-2, axis2=-1) / 3.0
        Q = Q - tr[..., None, None] * np.eye(3)
        U = expi_traceless_hermitian(Q) @ U
    return U


def plaquette(U: np.ndarray) -> float:
    """Average spatial plaquette Re tr(U_mu U_nu(x+mu) U_mu(x+nu)^dag U_nu^dag) / 3."""
    acc, n = 0.0, 0
    for mu in range(3):
        for nu in range(mu + 1, 3):
            P = U[mu] @ _shift(U[nu], mu, +1) @ _dag(_shift(U[mu], nu, +1)) @ _dag(U[nu])
            acc += np.real(np.trace(P, axis1=-2, axis2=-1)).mean() / 3.0
            n += 1
    return acc / n
