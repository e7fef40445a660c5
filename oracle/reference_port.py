"""Faithful CPU port of the reference loop nest, used as the CPU baseline.

TEST INFRASTRUCTURE ONLY (the ``cpu_baseline`` leg of ``bench.py`` and the
tests).  It keeps the reference's work pattern of
``lattice/generator/elemental.py:290-338`` so that its timing stands for the
reference's numpy backend:

* V is staged through a complex64 buffer, one eigenvector at a time
  (``elemental.py:55,297-298``);
* for every derivative and every split, both displaced vectors are recomputed
  from V (``elemental.py:309-321``, 78 hop applications per timeslice when
  num_nabla = 2), each hop doing two ``roll`` copies and two colour
  contractions (``elemental.py:279-288``);
* for every momentum the three-operand contraction "zyx,ezyxc,fzyxc->ef" is
  evaluated separately (``elemental.py:322-329``).  ``opt_einsum.contract``
  is not installed here, so ``numpy.einsum(..., optimize=True)`` stands in;
  both route the pairwise product to BLAS.

The result equals ``elemental_oracle.elemental_timeslice`` to rounding.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

from .elemental_oracle import derivative_list, momentum_phase


class ReferencePortGenerator:
    """Per-timeslice generator with the reference's constructor-time state."""

    def __init__(self, latt_size, U, V, num_nabla, momentum_list, stocastic_coeff=None):
        Lx, Ly, Lz, Lt = latt_size
        self.latt_size = latt_size
        self.U = U  # [3, Lt, Lz, Ly, Lx, 3, 3] complex128
        self.eigenvector = V  # [Lt, Ne, Lz, Ly, Lx, 3]
        self.derivative_list = derivative_list(num_nabla)
        self.momentum_list = list(momentum_list)
        self.Ne = V.shape[1]
        self._V = np.zeros((self.Ne, Lz, Ly, Lx, 3), "<c8")
        self._VPV = np.zeros((len(self.derivative_list), len(self.momentum_list), self.Ne, self.Ne), "<c16")
        self._phase = {tuple(m): momentum_phase(latt_size, m) for m in self.momentum_list}
        self.stocastic_coeff = stocastic_coeff

    @staticmethod
    def _nD(V, U, deriv):
        for d in deriv:
            Vf = np.roll(V, -1, 3 - d)
            UVf = np.einsum("zyxab,ezyxb->ezyxa", U[d], Vf, optimize=True)
            UdV = np.einsum("zyxba,ezyxb->ezyxa", U[d].conj(), V, optimize=True)
            V = UVf - np.roll(UdV, 1, 3 - d)
        return V

    def calc(self, t: int, momentum_subset=None, timers=None) -> np.ndarray:
        """calc(t) of the reference; ``momentum_subset`` (indices) and
        ``timers`` (dict receiving "hop" and "contract" seconds) serve the
        bounded CPU-baseline sample."""
        from time import perf_counter

        V = self._V
        VPV = self._VPV
        U_t = self.U[:, t]
        for e in range(self.Ne):
            V[e] = self.eigenvector[t, e]
        moms = list(enumerate(self.momentum_list))
        if momentum_subset is not None:
            moms = [moms[i] for i in momentum_subset]
        th = tc = 0.0
        for di, deriv in enumerate(self.derivative_list):
            VPV[di] = 0
            for pick in range(2 ** len(deriv)):
                right, left = [], []
                p = pick
                for direction in deriv:
                    (right if p & 1 else left).append(direction)
                    p >>= 1
                coeff = (-1) ** len(right)
                s0 = perf_counter()
                r = self._nD(V, U_t, right)
                l = self._nD(V, U_t, left[::-1])
                s1 = perf_counter()
                th += s1 - s0
                for pi, mom in moms:
                    ph = coeff * self._phase[tuple(mom)]
                    if self.stocastic_coeff is None:
                        VPV[di, pi] += np.einsum("zyx,ezyxc,fzyxc->ef", ph, l.conj(), r, optimize=True)
                    else:
                        VPV[di, pi] += np.einsum(
                            "zyx,ezyxc,fzyxc,ef->ef", ph, l.conj(), r, self.stocastic_coeff, optimize=True
                        )
                tc += perf_counter() - s1
        if timers is not None:
            timers["hop"] = timers.get("hop", 0.0) + th
            timers["contract"] = timers.get("contract", 0.0) + tc
        return VPV


def elemental_port(V, U, latt_size: Sequence[int], num_nabla, momentum_list, ts=None):
    """[len(ts), D, P, Ne, Ne] through the faithful port (copying the aliased buffer per t)."""
    gen = ReferencePortGenerator(latt_size, U, V, num_nabla, momentum_list)
    ts = range(V.shape[0]) if ts is None else ts
    return np.stack([gen.calc(t).copy() for t in ts])
