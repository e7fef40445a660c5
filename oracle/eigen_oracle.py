"""CPU oracle for the Laplacian eigenvectors that feed elemental generation.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the package
``easydistillation_amd``) may import this module; it is imported only by
``tests/``.

* ``laplacian``     follows lattice/generator/eigenvector.py:11-26:
                    (L F)(x) = 6 F(x) - sum_mu [U_mu(x) F(x+mu) + U_mu(x-mu)^dag F(x-mu)]
                    on F[Lz, Ly, Lx, Nc, nvec].
* ``lowest_modes``  follows eigenvector.py:234-257: scipy ``eigsh`` with
                    which="SA" on that operator, eigenvalues sorted, and the
                    phase of each eigenvector fixed so that its first component
                    is real and non-negative.

Pinning (tests/test_eigen_oracle.py): ``laplacian`` against per-site loops,
Hermiticity and positivity, gauge covariance, and the free-field spectrum
4 sum_mu sin^2(pi k_mu / L_mu); ``lowest_modes`` against a dense ``eigh`` of
the assembled matrix on a tiny lattice.
"""

from __future__ import annotations

import numpy as np


def laplacian(F: np.ndarray, U_t: np.ndarray) -> np.ndarray:
    """F[Lz, Ly, Lx, 3, ...] -> L F (eigenvector.py:11-26); U_t[3, Lz, Ly, Lx, 3, 3]."""
    out = 6.0 * F
    for mu in range(3):
        ax = 2 - mu  # direction mu acts on axis 2 - mu of [z, y, x]
        out = out - np.einsum("zyxab,zyxb...->zyxa...", U_t[mu], np.roll(F, -1, ax))
        out = out - np.roll(np.einsum("zyxba,zyxb...->zyxa...", U_t[mu].conj(), F), 1, ax)
    return out


def lowest_modes(U_t: np.ndarray, Ne: int, tol: float = 1e-10, apply_renorm_phase: bool = True):
    """(evecs [Ne, Lz, Ly, Lx, 3], evals [Ne]) of the Laplacian (eigenvector.py:234-257)."""
    from scipy.sparse import linalg

    Lz, Ly, Lx = U_t.shape[1:4]
    n = Lz * Ly * Lx * 3

    def mv(v):
        v = np.asarray(v).reshape(Lz, Ly, Lx, 3, -1)
        return laplacian(v, U_t).reshape(n, -1)

    A = linalg.LinearOperator((n, n), matvec=mv, matmat=mv, dtype=np.complex128)
    rng = np.random.default_rng(0)
    v0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    evals, evecs = linalg.eigsh(A, Ne, which="SA", tol=tol, v0=v0)
    evecs = evecs.T.reshape(Ne, -1)
    order = np.argsort(evals)
    evals, evecs = evals[order], evecs[order]
    if apply_renorm_phase:
        evecs = evecs * np.exp(-1j * np.angle(evecs[:, 0]))[:, None]
    return evecs.reshape(Ne, Lz, Ly, Lx, 3), evals


def dense_laplacian(U_t: np.ndarray) -> np.ndarray:
    """The Laplacian as a dense [n, n] matrix by per-site loops (tiny lattices only)."""
    Lz, Ly, Lx = U_t.shape[1:4]
    L = (Lx, Ly, Lz)
    n = Lz * Ly * Lx * 3
    M = np.zeros((n, n), complex)

    def idx(x, y, z, c):
        return ((z % Lz) * Ly + (y % Ly)) * Lx * 3 + (x % Lx) * 3 + c

    for z in range(Lz):
        for y in range(Ly):
            for x in range(Lx):
                for a in range(3):
                    M[idx(x, y, z, a), idx(x, y, z, a)] += 6.0
                for mu in range(3):
                    e = [0, 0, 0]
                    e[mu] = 1
                    u = U_t[mu, z, y, x]
                    bx, by, bz = x - e[0], y - e[1], z - e[2]
                    ub = U_t[mu, bz % Lz, by % Ly, bx % Lx]
                    for a in range(3):
                        for b in range(3):
                            M[idx(x, y, z, a), idx(x + e[0], y + e[1], z + e[2], b)] -= u[a, b]
                            M[idx(x, y, z, a), idx(bx, by, bz, b)] -= np.conj(ub[b, a])
    del L
    return M
