"""Synthetic gauge links, eigenvectors and momentum sets.

The reference ships its test configuration only as Git-LFS pointers
(``tests/weak_field.*``), so parity and benchmarks run on seeded synthetic
inputs of the same layouts:

  U  [3, Lt, Lz, Ly, Lx, 3, 3]   complex128 spatial links (elemental.py:103)
  V  [Lt, Ne, Lz, Ly, Lx, 3]     eigenvectors, orthonormal per timeslice over
                                 (site, colour) and complex64-representable,
                                 because the reference stages V in complex64
                                 (elemental.py:55,297-298)

Two generator families:

* ``*_np``: numpy ``default_rng(seed)``, draw order U (all 4 directions,
  spatial 3 kept) then V then profile weights; used for committed fixtures
  and CPU tests.
* ``*_torch``: the same recipes with ``torch`` on any device, used to build
  benchmark-size inputs directly in HBM.
"""

from __future__ import annotations

import itertools
from typing import List, Sequence, Tuple

import numpy as np

Momentum = Tuple[int, int, int]


# ----------------------------------------------------------------------------
# momentum sets
# ----------------------------------------------------------------------------
def momenta_by_norm(max_n2: int) -> List[Momentum]:
    """All integer n with |n|^2 <= max_n2, sorted by |n|^2 then lexicographically.

    |n|^2 <= 2 gives 19 momenta, <= 3 gives 27, <= 4 gives 33.
    """
    r = int(np.floor(np.sqrt(max_n2)))
    out = [
        (x, y, z)
        for x, y, z in itertools.product(range(-r, r + 1), repeat=3)
        if x * x + y * y + z * z <= max_n2
    ]
    out.sort(key=lambda n: (n[0] ** 2 + n[1] ** 2 + n[2] ** 2, n))
    return out


def unit_momenta() -> List[Momentum]:
    """(0,0,0) and the six unit vectors."""
    return [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


# ----------------------------------------------------------------------------
# numpy generators
# ----------------------------------------------------------------------------
def _haar_su3_from_gaussian(Z: np.ndarray) -> np.ndarray:
    """QR of complex Gaussian 3x3 matrices, phase-fixed, then scaled to det = 1."""
    Q, R = np.linalg.qr(Z)
    d = np.diagonal(R, axis1=-2, axis2=-1)
    Q = Q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(Q)
    Q = Q / (det ** (1.0 / 3.0))[..., None, None]
    return Q


def _smooth_su3_from_gaussian(A: np.ndarray, eps: float) -> np.ndarray:
    """exp(i eps H), H = traceless Hermitian part of A, via Hermitian eigendecomposition."""
    H = 0.5 * (A + np.conj(np.swapaxes(A, -1, -2)))
    tr = np.trace(H, axis1=-2, axis2=-1).real / 3.0
    H = H - tr[..., None, None] * np.eye(3)
    w, Q = np.linalg.eigh(H)
    return (Q * np.exp(1j * eps * w)[..., None, :]) @ np.conj(np.swapaxes(Q, -1, -2))


def gauge_np(rng: np.random.Generator, latt_size: Sequence[int], kind: str = "haar", eps: float = 0.3) -> np.ndarray:
    """Spatial links U[3, Lt, Lz, Ly, Lx, 3, 3] complex128.

    Draws all 4 directions (the time links are discarded) so the draw order
    matches a full [Lt, Lz, Ly, Lx, 4, 3, 3] configuration.
    """
    Lx, Ly, Lz, Lt = latt_size
    shape = (Lt, Lz, Ly, Lx, 4, 3, 3)
    Z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)
    if kind == "haar":
        U = _haar_su3_from_gaussian(Z)
    elif kind == "smooth":
        U = _smooth_su3_from_gaussian(Z, eps)
    elif kind == "unit":
        U = np.broadcast_to(np.eye(3, dtype=np.complex128), shape).copy()
    else:
        raise ValueError(f"unknown link kind {kind!r}")
    return np.ascontiguousarray(U.transpose(4, 0, 1, 2, 3, 5, 6)[:3]).astype(np.complex128)


def eigvecs_np(rng: np.random.Generator, latt_size: Sequence[int], Ne: int) -> np.ndarray:
    """V[Lt, Ne, Lz, Ly, Lx, 3]: complex Gaussian, orthonormalised per t in FP64, rounded to complex64."""
    Lx, Ly, Lz, Lt = latt_size
    K = Lx * Ly * Lz * 3
    if Ne > K:
        raise ValueError("Ne exceeds 3*V3")
    shape = (Lt, Ne, K)
    G = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)
    out = np.empty((Lt, Ne, K), dtype=np.complex64)
    for t in range(Lt):
        Q, _ = np.linalg.qr(G[t].T)  # K x Ne, orthonormal columns
        out[t] = Q.T.astype(np.complex64)
    return out.reshape(Lt, Ne, Lz, Ly, Lx, 3)


def profile_weights_np(rng: np.random.Generator, Ne: int) -> np.ndarray:
    return rng.uniform(0.5, 1.5, size=Ne)


def make_inputs_np(latt_size, Ne, kind="haar", seed=0, eps=0.3):
    """(U, V) drawn from one ``default_rng(seed)`` in the documented order."""
    rng = np.random.default_rng(seed)
    U = gauge_np(rng, latt_size, kind, eps)
    V = eigvecs_np(rng, latt_size, Ne)
    return U, V


# ----------------------------------------------------------------------------
# torch generators (device-resident benchmark inputs)
# ----------------------------------------------------------------------------
def gauge_torch(latt_size, kind="haar", eps=0.3, seed=0, device="cuda", chunk=1 << 22):
    """Spatial links [3, Lt, Lz, Ly, Lx, 3, 3] complex128 built on ``device``.

    haar:   Gram-Schmidt on the first two columns of a complex Gaussian 3x3
            matrix, third column conj(u1 x u2), so det = 1 (Haar on SU(3));
    smooth: exp(i eps H) with H the traceless Hermitian part of a complex
            Gaussian (batched matrix exponential);
    unit:   identity links.
    """
    import torch

    Lx, Ly, Lz, Lt = latt_size
    n = 3 * Lt * Lz * Ly * Lx
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = torch.empty((n, 3, 3), dtype=torch.complex128, device=device)
    eye = torch.eye(3, dtype=torch.complex128, device=device)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        re = torch.randn((m, 3, 3), generator=g, dtype=torch.float64, device=device)
        im = torch.randn((m, 3, 3), generator=g, dtype=torch.float64, device=device)
        Z = torch.complex(re, im) / np.sqrt(2.0)
        if kind == "haar":
            z1, z2 = Z[:, :, 0], Z[:, :, 1]
            u1 = z1 / torch.linalg.vector_norm(z1, dim=1, keepdim=True)
            u2 = z2 - (u1.conj() * z2).sum(1, keepdim=True) * u1
            u2 = u2 / torch.linalg.vector_norm(u2, dim=1, keepdim=True)
            u3 = torch.stack(
                [
                    u1[:, 1] * u2[:, 2] - u1[:, 2] * u2[:, 1],
                    u1[:, 2] * u2[:, 0] - u1[:, 0] * u2[:, 2],
                    u1[:, 0] * u2[:, 1] - u1[:, 1] * u2[:, 0],
                ],
                dim=1,
            ).conj()
            Q = torch.stack([u1, u2, u3], dim=2)
        elif kind == "smooth":
            H = 0.5 * (Z + Z.conj().transpose(-1, -2))
            tr = torch.diagonal(H, dim1=-2, dim2=-1).sum(-1).real / 3.0
            H = H - tr[:, None, None] * eye
            Q = torch.linalg.matrix_exp((1j * eps) * H)
        elif kind == "unit":
            Q = eye.expand(m, 3, 3)
        else:
            raise ValueError(kind)
        out[s : s + m] = Q
    return out.view(3, Lt, Lz, Ly, Lx, 3, 3)


def eigvecs_torch(latt_size, Ne, seed=1, device="cuda", log=None):
    """V[Lt, Ne, Lz, Ly, Lx, 3] orthonormal per t (CholeskyQR2 in FP64), stored complex64."""
    import torch

    Lx, Ly, Lz, Lt = latt_size
    K = 3 * Lx * Ly * Lz
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = torch.empty((Lt, Ne, K), dtype=torch.complex64, device=device)
    for t in range(Lt):
        re = torch.randn((Ne, K), generator=g, dtype=torch.float64, device=device)
        im = torch.randn((Ne, K), generator=g, dtype=torch.float64, device=device)
        A = torch.complex(re, im)
        for _ in range(2):  # rows of A become orthonormal
            G = (A.conj() @ A.transpose(0, 1)).cpu()  # Gram matrix [Ne, Ne], G_ij = <a_i, a_j>
            L = torch.linalg.cholesky(G.transpose(0, 1))  # G^T = conj(G) = L L^dagger
            A = torch.linalg.inv(L).to(A.device) @ A  # small factor on the host, GEMM on the device
        out[t] = A.to(torch.complex64)
        if log is not None and (t + 1) % 16 == 0:
            log(f"eigenvectors {t + 1}/{Lt}")
    return out.view(Lt, Ne, Lz, Ly, Lx, 3)
