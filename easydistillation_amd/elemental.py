"""Drop-in ``ElementalGenerator`` computing on MI355X through the HIP library.

Mirrors lattice/generator/elemental.py:
  * constructor (``:17-59``): lattice size, gauge field and eigenvector
    containers, ``num_nabla``, momentum list, optional ``dilution`` /
    ``is_blending`` weights (``:61-100``);
  * ``load(key)`` (``:102-105``): reads the gauge field [Lt,Lz,Ly,Lx,Nd,Nc,Nc],
    keeps the three spatial directions and hands links and eigenvectors to the
    device;
  * ``calc(t)`` (``:290-338``): returns E[d, p, e, f] for timeslice t as
    complex128 [D, P, Ne, Ne].  Unlike the reference the returned array is a
    fresh array, not an internal buffer reused by the next call.

Additions: ``calc_all`` computes every timeslice in batched launches (the
fast path; ``calc(t)`` serves from its result when present), ``load_arrays``
takes links and eigenvectors directly (numpy or device tensors), and
``mode`` selects the FP64 form or the mixed-precision form.
"""

from __future__ import annotations

from typing import List, Tuple

import numpy as np

from . import _native
from .backend import get_backend
from .constant import Nc, Nd
from .insertion import derivative, num_derivative

_MODES = {"fp64": _native.ED_MODE_FP64, "mixed": _native.ED_MODE_MIXED}


def blending_weights(Ne: int, dilution) -> np.ndarray:
    """Stochastic blending weights [Ne, Ne] for ``dilution = (totNe_list, usedNe)``.

    Same assignment as lattice/generator/elemental.py:61-100: between
    different dilution blocks i != j the weight is tot_i tot_j / (used_i
    used_j); inside block i the diagonal gets tot_i/used_i and the
    off-diagonal tot_i (tot_i - 1) / (used_i (used_i - 1)).
    """
    tot = [int(v) for v in dilution[0]]
    used = [int(dilution[1])] * len(tot) if isinstance(dilution[1], int) else [int(v) for v in dilution[1]]
    assert len(used) == len(tot)
    assert all(u <= t for u, t in zip(used, tot))
    assert sum(used) == Ne
    edges = np.cumsum([0] + used)
    W = np.empty((Ne, Ne))
    for i, (ti, ui) in enumerate(zip(tot, used)):
        for j, (tj, uj) in enumerate(zip(tot, used)):
            blk = W[edges[i] : edges[i + 1], edges[j] : edges[j + 1]]
            if i != j:
                blk[:] = ti * tj / ui / uj
            else:
                blk[:] = (ti / ui) * (ti - 1) / (ui - 1)
                np.fill_diagonal(blk, ti / ui)
    return W


def _device_ptr(x):
    """(pointer, on_device, dtype code, keepalive) for a numpy array or torch tensor."""
    try:
        import torch
    except Exception:  # pragma: no cover - torch is part of the image
        torch = None
    if torch is not None and isinstance(x, torch.Tensor):
        if x.dtype not in (torch.complex128, torch.complex64):
            raise TypeError("complex128 or complex64 tensor required")
        x = x.contiguous()
        if x.is_cuda:
            # the library reads on its own stream: finish torch's producers first
            torch.cuda.current_stream(x.device).synchronize()
        code = _native.ED_C128 if x.dtype == torch.complex128 else _native.ED_C64
        return x.data_ptr(), x.is_cuda, code, x
    a = np.asarray(x)
    if a.dtype == np.complex128 or a.dtype == np.complex64:
        a = np.ascontiguousarray(a)
    else:
        a = np.ascontiguousarray(a, dtype=np.complex128)
    code = _native.ED_C128 if a.dtype == np.complex128 else _native.ED_C64
    return a.ctypes.data, False, code, a


class ElementalGenerator:
    def __init__(
        self,
        latt_size: List[int],
        gauge_field=None,
        eigenvector=None,
        num_nabla: int = 0,
        momentum_list: List[Tuple[int]] = [(0, 0, 0)],
        dilution: Tuple = None,
        is_blending: bool = False,
        *,
        Ne: int | None = None,
        mode: str = "fp64",
        device: int = 0,
    ) -> None:
        if mode not in _MODES:
            raise ValueError(f"mode must be one of {tuple(_MODES)}")
        if not 0 <= int(num_nabla) <= 3:
            raise ValueError("num_nabla must be in [0, 3]")
        Lx, Ly, Lz, Lt = (int(v) for v in latt_size)
        self.latt_size = [Lx, Ly, Lz, Lt]
        self.gauge_field = gauge_field
        self.eigenvector = eigenvector
        self.num_nabla = int(num_nabla)
        self.num_derivative = num_derivative(self.num_nabla)
        self.derivative_list = [derivative(n) for n in range(self.num_derivative)]
        self.momentum_list = [tuple(int(c) for c in m) for m in momentum_list]
        if any(len(m) != 3 for m in self.momentum_list):
            raise ValueError("momenta are (nx, ny, nz) triples")
        self.num_momentum = len(self.momentum_list)
        if Ne is None:
            if eigenvector is None:
                raise ValueError("Ne is required when no eigenvector container is given")
            Ne = eigenvector.Ne
        self.Ne = int(Ne)
        self.mode = mode

        if is_blending and dilution is not None:
            self.stocastic_coeff = blending_weights(self.Ne, dilution)
        elif is_blending:
            raise ValueError("Dilution tuple is not defined.")
        else:
            self.stocastic_coeff = None

        self._ctx = _native.Context(device)
        self._ctx.set_lattice(Lx, Ly, Lz, Lt, self.Ne)
        self._ctx.set_derivatives(self.num_nabla)
        self._ctx.set_momenta(self.momentum_list)
        if self.stocastic_coeff is not None:
            w = np.ascontiguousarray(self.stocastic_coeff, dtype=np.float64)
            self._ctx.set_weights(w.ctypes.data)
        self._loaded = False
        self._all = None  # (mode, [Lt, D, P, Ne, Ne]) of the last calc_all
        self._win = None  # (mode, t0, device tensor [n, D, P, Ne, Ne]) of the last calc(t) window
        self._next = None  # the following window, its compute enqueued and not waited for
        self._pins = None  # two page-locked host buffers of one timeslice
        self._pre = None  # ((mode, t), pin index, event) of the copy to page-locked memory in flight
        self._copy_stream = None

    # ------------------------------------------------------------------ input
    def load(self, key: str):
        """Read links and eigenvectors of configuration ``key`` (elemental.py:102-105)."""
        Lx, Ly, Lz, Lt = self.latt_size
        gauge = self.gauge_field.load(key)
        by_record = getattr(self.gauge_field, "spatial_links", None)
        if by_record is not None:
            # containers laid out per (t, mu) assemble the spatial links themselves
            U = by_record(key, self.latt_size)
        else:
            U = np.asarray(gauge[:])
            n = Lt * Lz * Ly * Lx * Nd * Nc * Nc
            if U.dtype.kind == "f" and U.size == 2 * n:
                # a real array of interleaved (re, im) pairs
                U = np.ascontiguousarray(U, dtype=np.float64).view(np.complex128)
            if U.size != n:
                raise ValueError(f"gauge field shape {U.shape} is not [Lt,Lz,Ly,Lx,Nd,Nc,Nc]")
            U = U.reshape(Lt, Lz, Ly, Lx, Nd, Nc, Nc).transpose(4, 0, 1, 2, 3, 5, 6)[: Nd - 1]
        self._gauge_field_path = getattr(gauge, "file", None)
        eig = self.eigenvector.load(key)
        V = np.empty((Lt, self.Ne, Lz, Ly, Lx, Nc), dtype=np.complex64)
        for t in range(Lt):
            for e in range(self.Ne):
                # one eigenvector per read, staged in complex64 (elemental.py:55,297-298)
                V[t, e] = np.asarray(eig[t, e]).reshape(Lz, Ly, Lx, Nc)
        self.load_arrays(U, V)

    def load_arrays(self, U, V) -> None:
        """Links U[3, Lt, Lz, Ly, Lx, 3, 3] and eigenvectors V[Lt, Ne, Lz, Ly, Lx, 3].

        numpy arrays or torch tensors (device tensors are read in place).
        V is rounded to complex64 as the reference does.
        """
        Lx, Ly, Lz, Lt = self.latt_size
        if tuple(U.shape) != (Nd - 1, Lt, Lz, Ly, Lx, Nc, Nc):
            raise ValueError(f"U shape {tuple(U.shape)} is not [3, Lt, Lz, Ly, Lx, 3, 3]")
        if tuple(V.shape) != (Lt, self.Ne, Lz, Ly, Lx, Nc):
            raise ValueError(f"V shape {tuple(V.shape)} is not [Lt, Ne, Lz, Ly, Lx, 3]")
        self._drain()
        ptr, dev, code, keep = _device_ptr(U)
        self._ctx.set_gauge(ptr, code, dev)
        del keep
        ptr, dev, code, keep = _device_ptr(V)
        self._ctx.set_eigvecs(ptr, code, dev)
        del keep
        self._loaded = True
        self._invalidate()

    # ---------------------------------------------------------- link smearing
    def project_SU3(self, max_iter: int = 100) -> int:
        """Unitary projection of the loaded links in device memory (elemental.py:107-117).

        Polar iteration U <- (U + (U^-1)^dagger)/2 until every link passes the
        reference's 1e-15 tests; returns the number of updates.  The reference
        loops without bound; here more than ``max_iter`` updates raise.
        """
        if not self._loaded:
            raise RuntimeError("call load(key) or load_arrays(U, V) first")
        self._invalidate()
        return self._ctx.project_su3(max_iter)

    def stout_smear(self, nstep: int, rho: float) -> None:
        """``nstep`` 3-D stout-smearing steps of weight ``rho`` on the loaded links
        (elemental.py:264-277, numpy form :175-242), computed on the device."""
        if not self._loaded:
            raise RuntimeError("call load(key) or load_arrays(U, V) first")
        self._invalidate()
        self._ctx.stout_smear(nstep, rho)

    def gauge(self):
        """Current links U[3, Lt, Lz, Ly, Lx, 3, 3] complex128 (numpy, or a device
        tensor with the "torch" backend)."""
        if not self._loaded:
            raise RuntimeError("call load(key) or load_arrays(U, V) first")
        Lx, Ly, Lz, Lt = self.latt_size
        shape = (Nd - 1, Lt, Lz, Ly, Lx, Nc, Nc)
        if get_backend() == "torch":
            import torch

            out = torch.empty(shape, dtype=torch.complex128, device=f"cuda:{self._ctx.device}")
            self._ctx.get_gauge(out.data_ptr(), True)
            return out
        out = np.empty(shape, dtype=np.complex128)
        self._ctx.get_gauge(out.ctypes.data, False)
        return out

    # ---------------------------------------------------------------- compute
    def _mode_code(self, mode):
        return _MODES[mode or self.mode]

    WINDOW = 8  # timeslices per device window of calc(t)

    def _drain(self) -> None:
        """Waits for what calc(t) left in flight (the next window's compute on the
        library stream, the next timeslice's copy to page-locked memory) and
        drops it, so that no buffer is released or overwritten while it is used."""
        if self._pre is not None:
            self._pre[2].synchronize()
            self._pre = None
        if self._next is not None:
            self._ctx.synchronize()
            self._next = None

    def _invalidate(self) -> None:
        self._drain()
        self._all = None
        self._win = None

    def calc(self, t: int, mode: str | None = None):
        """E[d, p, e, f] of timeslice t, complex128 [D, P, Ne, Ne] (elemental.py:290-338).

        Callers loop ``for t: data[t] = calc(t)`` (tests/test_elemental.py:42-45).
        The first call in an aligned group of ``WINDOW`` timeslices computes the
        whole group in one batched launch into device memory; the calls of
        that group copy their timeslice out (through a page-locked staging
        buffer) and return a fresh array.
        """
        if not self._loaded:
            raise RuntimeError("call load(key) or load_arrays(U, V) first")
        Lt = self.latt_size[3]
        if not 0 <= t < Lt:
            raise IndexError(f"timeslice {t} outside [0, {Lt})")
        m = mode or self.mode
        if self._all is not None and m == self._all[0]:
            res = self._all[1][t]
            return res.clone() if get_backend() == "torch" else np.array(res, copy=True)
        import torch

        host = get_backend() != "torch"
        dev = f"cuda:{self._ctx.device}"

        def window_shape(t0):
            return (min(self.WINDOW, Lt - t0), self.num_derivative, self.num_momentum, self.Ne, self.Ne)

        w = self._win
        if w is None or w[0] != m or not w[1] <= t < w[1] + w[2].shape[0]:
            nx = self._next
            t0 = t // self.WINDOW * self.WINDOW
            if nx is not None and nx[0] == m and nx[1] == t0:
                # the window enqueued by the previous calls
                if self._pre is not None:
                    self._pre[2].synchronize()  # the last copy out of the old window
                    self._pre = None
                self._ctx.synchronize()
                self._next = None
                spare = w[2] if w is not None else None
                self._win = w = nx
            else:
                self._drain()
                spare = nx[2] if nx is not None else None
                buf = w[2] if w is not None and tuple(w[2].shape) == window_shape(t0) else None
                self._win = None
                if buf is None:
                    buf = torch.empty(window_shape(t0), dtype=torch.complex128, device=dev)
                torch.cuda.current_stream(dev).synchronize()  # buf's memory is idle on torch's stream
                self._ctx.compute(t0, buf.shape[0], self._mode_code(m), buf.data_ptr(), True)
                self._ctx.synchronize()
                spare = spare if spare is not None else (w[2] if w is not None and w[2] is not buf else None)
                self._win = w = (m, t0, buf)
            t1 = w[1] + self.WINDOW
            if host and t1 < Lt:
                # the following window computes while this one is copied to the host
                # (skipped when device memory cannot hold a second window)
                shape = window_shape(t1)
                nb = spare if spare is not None and tuple(spare.shape) == shape else None
                if nb is None:
                    try:
                        nb = torch.empty(shape, dtype=torch.complex128, device=dev)
                    except torch.cuda.OutOfMemoryError:
                        nb = None
                    torch.cuda.current_stream(dev).synchronize()
                if nb is not None:
                    self._ctx.compute(t1, shape[0], self._mode_code(m), nb.data_ptr(), True)
                    self._next = (m, t1, nb)
        res = w[2][t - w[1]]
        if not host:
            return res.clone()
        # timeslice t in page-locked memory: copied ahead by the previous call, or now
        if self._pins is None or tuple(self._pins[0].shape) != tuple(res.shape):
            self._pins = [torch.empty(tuple(res.shape), dtype=torch.complex128, pin_memory=True) for _ in range(2)]
            self._copy_stream = torch.cuda.Stream(device=dev)
        pre = self._pre
        self._pre = None
        if pre is not None:
            pre[2].synchronize()
        if pre is not None and pre[0] == (m, t):
            k = pre[1]
        else:
            k = 0
            self._pins[k].copy_(res)
        # start the copy of t + 1 into the other buffer, then copy t to fresh host memory
        if t + 1 < w[1] + w[2].shape[0]:
            src = w[2][t + 1 - w[1]]
            with torch.cuda.stream(self._copy_stream):
                self._pins[1 - k].copy_(src, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(self._copy_stream)
            self._pre = ((m, t + 1), 1 - k, ev)
        # multi-threaded host copy into fresh memory, returned as a numpy array
        return torch.empty(tuple(res.shape), dtype=torch.complex128).copy_(self._pins[k]).numpy()

    def calc_range(self, t0: int, nt: int, mode: str | None = None, out=None):
        """[nt, D, P, Ne, Ne] for timeslices t0..t0+nt-1 in batched launches.

        ``out`` may be a preallocated numpy array or device tensor of that
        shape (complex128).  With the "torch" backend and no ``out`` the
        result stays in device memory.
        """
        shape = (nt, self.num_derivative, self.num_momentum, self.Ne, self.Ne)
        code = self._mode_code(mode)
        if out is None and get_backend() == "torch":
            import torch

            out = torch.empty(shape, dtype=torch.complex128, device=f"cuda:{self._ctx.device}")
        if out is None:
            out = np.empty(shape, dtype=np.complex128)
        if tuple(out.shape) != shape:
            raise ValueError(f"out shape {tuple(out.shape)} != {shape}")
        if isinstance(out, np.ndarray):
            if out.dtype != np.complex128 or not out.flags.c_contiguous:
                raise ValueError("out must be C-contiguous complex128")
            self._ctx.compute(t0, nt, code, out.ctypes.data, False)
        else:
            import torch

            if out.dtype != torch.complex128 or not out.is_contiguous() or not out.is_cuda:
                raise ValueError("out must be a contiguous complex128 device tensor")
            torch.cuda.current_stream(out.device).synchronize()
            self._ctx.compute(t0, nt, code, out.data_ptr(), True)
            self._ctx.synchronize()
        return out

    def calc_all(self, mode: str | None = None, out=None):
        """All timeslices, [Lt, D, P, Ne, Ne]; later ``calc(t)`` calls read from it."""
        res = self.calc_range(0, self.latt_size[3], mode, out)
        self._all = (mode or self.mode, res)
        return res

    def set_option(self, option: int, value: int) -> None:
        self._drain()
        self._ctx.set_option(option, value)
        self._invalidate()

    @property
    def context(self) -> _native.Context:
        return self._ctx
