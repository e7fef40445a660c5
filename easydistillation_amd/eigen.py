"""Laplacian eigenvectors on MI355X: the feed of elemental generation.

Mirrors lattice/generator/eigenvector.py ``EigenvectorGenerator``:
constructor (``:29-47``: lattice size, gauge field container, Ne, tol),
``load`` (``:49-53``), ``project_SU3`` / ``stout_smear`` (``:55-232``) and
``calc(t)`` (``:356-``) returning the Ne lowest eigenpairs of the 3-D
gauge-covariant Laplacian of timeslice t (``_Laplacian``, ``:11-26``),
eigenvalues ascending and each eigenvector's phase fixed so that its first
component is real and non-negative (``:234-257``).

The reference calls ARPACK (``eigsh`` on a LinearOperator) on the host or
QUDA.  Here the operator is the library's HIP Laplacian (``ed_laplacian``)
applied to a block of vectors in HBM, and the solver is a Chebyshev-filtered
subspace iteration: Rayleigh-Ritz on a block of Ne + guard vectors, a
Chebyshev polynomial filter that damps the spectrum above the block's
largest Ritz value up to the Laplacian's bound 12 (unitary links), and
CholeskyQR2 re-orthonormalisation, until every wanted Ritz pair has residual
norm <= tol * max(|lambda|, 1).  The eigenvectors stay on the device
([Ne, Lz, Ly, Lx, 3] complex128), ready for ``ElementalGenerator.load_arrays``.
"""

from __future__ import annotations

from typing import List

import numpy as np
import torch

from . import _native
from .backend import get_backend
from .constant import Nc, Nd

__all__ = ["EigenvectorGenerator", "chebyshev_subspace_eigs"]

LAPLACIAN_BOUND = 12.0  # ||L|| <= 6 + 6 for unitary links


def _orthonormalize_rows(X: torch.Tensor) -> torch.Tensor:
    """CholeskyQR2 of the rows of X [m, n]; the m x m factor is formed on the host."""
    for _ in range(2):
        K = (X @ X.conj().transpose(0, 1).contiguous()).cpu()
        L = torch.linalg.cholesky(K)
        X = torch.linalg.inv(L).to(X.device) @ X
    return X


def chebyshev_subspace_eigs(apply, n: int, k: int, tol: float, device, guard: int | None = None, degree: int = 12,
                            max_iter: int = 1000, seed: int = 0, upper: float = LAPLACIAN_BOUND):
    """Lowest k eigenpairs of a Hermitian operator on C^n with spectrum in [0, upper].

    ``apply(X)`` maps a block X [m, n] (rows are vectors) to [m, n].  Returns
    (evals [k] float64 host, evecs [k, n] on ``device``, iterations).
    """
    m = min(n, k + (guard if guard is not None else max(8, k // 4)))
    g = torch.Generator(device=device).manual_seed(seed)
    X = torch.complex(torch.randn((m, n), generator=g, dtype=torch.float64, device=device),
                      torch.randn((m, n), generator=g, dtype=torch.float64, device=device))
    X = _orthonormalize_rows(X)
    for it in range(1, max_iter + 1):
        AX = apply(X)
        H = (X.conj() @ AX.transpose(0, 1).contiguous()).cpu()
        H = 0.5 * (H + H.conj().transpose(0, 1))
        theta, C = torch.linalg.eigh(H)
        Ct = C.transpose(0, 1).contiguous().to(device)
        X = Ct @ X
        AX = Ct @ AX
        res = torch.linalg.vector_norm(AX - theta.to(device)[:, None] * X, dim=1).cpu()
        ok = res[:k] <= tol * torch.clamp(theta[:k].abs(), min=1.0)
        if bool(ok.all()):
            return theta[:k].numpy().astype(np.float64), X[:k], it
        # filter: damp [a, upper], a = largest Ritz value of the block
        a = float(theta[-1])
        if a >= upper:
            a = 0.5 * (float(theta[k - 1]) + upper)
        e, c = 0.5 * (upper - a), 0.5 * (upper + a)
        Y0, Y1 = X, (apply(X) - c * X) / e
        for _ in range(2, degree + 1):
            Y0, Y1 = Y1, 2.0 * (apply(Y1) - c * Y1) / e - Y0
        X = _orthonormalize_rows(Y1)
    raise RuntimeError(f"Chebyshev subspace iteration did not converge in {max_iter} iterations")


class EigenvectorGenerator:
    def __init__(self, latt_size: List[int], gauge_field=None, Ne: int = 70, tol: float = 1e-10, *, device: int = 0,
                 degree: int = 12, guard: int | None = None, max_iter: int = 1000, seed: int = 0) -> None:
        Lx, Ly, Lz, Lt = (int(v) for v in latt_size)
        self.latt_size = [Lx, Ly, Lz, Lt]
        self.gauge_field = gauge_field
        self.Ne = int(Ne)
        self.tol = float(tol)
        self.degree, self.guard, self.max_iter, self.seed = int(degree), guard, int(max_iter), int(seed)
        self.device = device
        self._ctx = _native.Context(device)
        self._ctx.set_lattice(Lx, Ly, Lz, Lt, 1)
        self._stream = torch.cuda.Stream(device=torch.device("cuda", device))  # default priority
        self._ctx.set_stream(self._stream.cuda_stream)
        self._loaded = False
        self.iterations = None

    # -------------------------------------------------------------- links
    def load(self, key: str):
        """Links of configuration ``key`` (eigenvector.py:49-53): gauge_field.load(key)[:] is [Lt,Lz,Ly,Lx,Nd,Nc,Nc]."""
        U = np.asarray(self.gauge_field.load(key)[:])
        Lx, Ly, Lz, Lt = self.latt_size
        U = U.reshape(Lt, Lz, Ly, Lx, Nd, Nc, Nc).transpose(4, 0, 1, 2, 3, 5, 6)[: Nd - 1]
        self.load_links(U)

    def load_links(self, U) -> None:
        """Spatial links U[3, Lt, Lz, Ly, Lx, 3, 3] (numpy or device tensor)."""
        from .elemental import _device_ptr

        ptr, dev, code, keep = _device_ptr(U)
        self._ctx.set_gauge(ptr, code, dev)
        del keep
        self._loaded = True

    def project_SU3(self, max_iter: int = 100) -> int:
        """Unitary projection of the links (eigenvector.py:55-65)."""
        return self._ctx.project_su3(max_iter)

    def stout_smear(self, nstep: int, rho: float) -> None:
        """3-D stout smearing of the links (eigenvector.py:215-232)."""
        self._ctx.stout_smear(nstep, rho)

    # ------------------------------------------------------------ operator
    def laplacian(self, t: int, F: torch.Tensor) -> torch.Tensor:
        """L F for F [nvec, Lz, Ly, Lx, 3] (or [nvec, 3 V3]) complex128 on the device."""
        if not self._loaded:
            raise RuntimeError("call load(key) or load_links(U) first")
        F = F.contiguous()
        out = torch.empty_like(F)
        nvec = F.shape[0]
        # the library's stream starts after the caller's stream has produced F
        # and the caller's stream continues after the library has written out
        caller = torch.cuda.current_stream(F.device)
        self._stream.wait_stream(caller)
        with torch.cuda.stream(self._stream):
            self._ctx.laplacian(t, F.data_ptr(), out.data_ptr(), nvec, True)
        F.record_stream(self._stream)
        out.record_stream(self._stream)
        caller.wait_stream(self._stream)
        return out

    # ------------------------------------------------------------- solver
    def calc(self, t: int, apply_renorm_phase: bool = True):
        """(evecs [Ne, Lz, Ly, Lx, Nc], evals [Ne]) of timeslice t (eigenvector.py:234-257).

        The eigenvectors are a complex128 device tensor (a numpy array with the
        "hip" backend); eigenvalues ascending.
        """
        if not self._loaded:
            raise RuntimeError("call load(key) or load_links(U) first")
        Lx, Ly, Lz, Lt = self.latt_size
        if not 0 <= t < Lt:
            raise IndexError(f"timeslice {t} outside [0, {Lt})")
        n = Lz * Ly * Lx * Nc
        dev = torch.device("cuda", self.device)
        torch.cuda.current_stream(dev).synchronize()
        with torch.cuda.stream(self._stream):

            def apply(X):
                out = torch.empty_like(X)
                self._ctx.laplacian(t, X.data_ptr(), out.data_ptr(), X.shape[0], True)
                return out

            evals, evecs, it = chebyshev_subspace_eigs(apply, n, self.Ne, self.tol, dev, self.guard, self.degree,
                                                       self.max_iter, self.seed + t)
            if apply_renorm_phase:
                ph = torch.angle(evecs[:, 0])
                evecs = evecs * torch.exp(-1j * ph)[:, None]
            evecs = evecs.reshape(self.Ne, Lz, Ly, Lx, Nc).contiguous()
        self._stream.synchronize()
        self.iterations = it
        if get_backend() == "torch":
            return evecs, evals
        return evecs.cpu().numpy(), evals

    def calc_all(self, apply_renorm_phase: bool = True):
        """V [Lt, Ne, Lz, Ly, Lx, 3] on the device and eigenvalues [Lt, Ne]: every timeslice,
        in the layout ``ElementalGenerator.load_arrays`` takes."""
        Lx, Ly, Lz, Lt = self.latt_size
        V = torch.empty((Lt, self.Ne, Lz, Ly, Lx, Nc), dtype=torch.complex128, device=torch.device("cuda", self.device))
        vals = np.empty((Lt, self.Ne))
        from .backend import set_backend

        prev = get_backend()
        set_backend("torch")
        try:
            for t in range(Lt):
                V[t], vals[t] = self.calc(t, apply_renorm_phase)
        finally:
            set_backend(prev)
        return V, vals
