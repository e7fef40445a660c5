"""ctypes binding of ``libelemental_hip.so`` (declared in ``include/elemental_hip.h``).

The library is built in-tree (``easydistillation_amd/lib``) by
``easydistillation_amd.build.build()``.  There is no fallback: if the
library is missing or a call fails, an exception is raised.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libelemental_hip.so")

ED_OK = 0
ED_ERR_ARG, ED_ERR_STATE, ED_ERR_HIP, ED_ERR_ALLOC, ED_ERR_UNSUPPORTED = -1, -2, -3, -4, -5
ED_C128, ED_C64 = 0, 1
ED_MODE_FP64, ED_MODE_MIXED = 0, 1
ED_OPT_HERMITIAN, ED_OPT_TBATCH, ED_OPT_WS_MB, ED_OPT_PROFILE, ED_OPT_MIXED_STAGE2 = 1, 2, 3, 4, 5
ED_OPT_EXACT_V = 6
ED_OPT_MIXED_WS = 7
ED_OPT_WIDE_WS = 8
ED_OPT_ORTHO_GATE = 9
ED_OPT_STRIP = 10
ED_OPT_ITEM_ORDER = 11
ED_OPT_H16_UNIT = 12
ED_OPT_HOP_LDS = 13
ED_KERNEL_STENCIL, ED_KERNEL_CONTRACT, ED_KERNEL_STAGE, ED_KERNEL_ADJOINT = 0, 1, 2, 3

# Every exported symbol and its ctypes signature (restype, argtypes).
_VP = ctypes.c_void_p
_I = ctypes.c_int
SIGNATURES = {
    "ed_version": (ctypes.c_char_p, []),
    "ed_create": (_I, [_I, ctypes.POINTER(_VP)]),
    "ed_destroy": (_I, [_VP]),
    "ed_last_error": (ctypes.c_char_p, [_VP]),
    "ed_set_stream": (_I, [_VP, _VP]),
    "ed_set_lattice": (_I, [_VP, _I, _I, _I, _I, _I]),
    "ed_set_gauge": (_I, [_VP, _VP, _I, _I]),
    "ed_set_eigvecs": (_I, [_VP, _VP, _I, _I]),
    "ed_set_momenta": (_I, [_VP, ctypes.POINTER(ctypes.c_int32), _I]),
    "ed_set_derivatives": (_I, [_VP, _I]),
    "ed_set_displacement": (_I, [_VP, _I]),
    "ed_set_weights": (_I, [_VP, ctypes.POINTER(ctypes.c_double)]),
    "ed_compute": (_I, [_VP, _I, _I, _I, _VP, _I]),
    "ed_compute_noalloc": (_I, [_VP, _I, _I, _I, _VP]),
    "ed_synchronize": (_I, [_VP]),
    "ed_orthonormality": (_I, [_VP, ctypes.POINTER(ctypes.c_double)]),
    "ed_set_option": (_I, [_VP, _I, ctypes.c_int64]),
    "ed_get_option": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_int64)]),
    "ed_tune": (_I, [_VP, _I, _I, _I, _I, ctypes.POINTER(ctypes.c_double)]),
    "ed_profile_read": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    "ed_profile_reset": (_I, [_VP]),
    "ed_query": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_int64)]),
    "ed_project_su3": (_I, [_VP, _I, ctypes.POINTER(_I)]),
    "ed_stout_smear": (_I, [_VP, _I, ctypes.c_double]),
    "ed_get_gauge": (_I, [_VP, _VP, _I]),
    "ed_laplacian": (_I, [_VP, _I, _VP, _VP, _I, _I]),
    "ed_backward": (_I, [_VP, _I, _I, _VP, _I, _VP, _VP, _I]),
    "ed_batch_create": (_I, [_I, ctypes.POINTER(_VP)]),
    "ed_batch_destroy": (_I, [_VP]),
    "ed_batch_last_error": (ctypes.c_char_p, [_VP]),
    "ed_batch_add": (_I, [_VP, _VP, _I, _VP]),
    "ed_batch_set_option": (_I, [_VP, _I, ctypes.c_int64]),
    "ed_batch_get_option": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_int64)]),
    "ed_batch_run": (_I, [_VP]),
    "ed_batch_synchronize": (_I, [_VP]),
    "ed_batch_time": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_double)]),
    "ed_batch_tune": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_double)]),
}
ED_BOPT_GRID_PER_CU, ED_BOPT_ORDER, ED_BOPT_GRAPH, ED_BOPT_STRIP, ED_BOPT_H16 = 1, 2, 3, 4, 5
ED_Q_CONTRACT_FLOPS_PER_T, ED_Q_ALGO_FLOPS_PER_T, ED_Q_PRIMARY_DERIVATIVES = 1, 2, 3
ED_Q_MOMENTUM_BLOCKS, ED_Q_HOPS_PER_T, ED_Q_SPLITS = 4, 5, 6
ED_Q_MIXED_F32_FLOPS_PER_T, ED_Q_MIXED_F16_FLOPS_PER_T = 7, 8
ED_Q_MIXED_KERNEL = 9

_lib: Optional[ctypes.CDLL] = None


class NativeError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with easydistillation_amd.build.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback"
        )
    try:
        # torch bundles its own HIP runtime with the same soname; loading it
        # first makes this library bind to that runtime, so one process never
        # holds two HIP runtimes (the second would find no device).
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Context:
    """One device context (``ed_create``); methods mirror the C entry points."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _VP()
        rc = self.lib.ed_create(device, ctypes.byref(h))
        if rc != ED_OK:
            raise NativeError("ed_create", rc, "could not create a context (is a GPU visible?)")
        self.h = h
        self.device = device

    def _chk(self, name: str, rc: int) -> None:
        if rc != ED_OK:
            msg = self.lib.ed_last_error(self.h)
            raise NativeError(name, rc, msg.decode() if msg else "")

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.ed_destroy(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr: int | None) -> None:
        self._chk("ed_set_stream", self.lib.ed_set_stream(self.h, _VP(stream_ptr or 0)))

    def set_lattice(self, Lx, Ly, Lz, Lt, Ne) -> None:
        self._chk("ed_set_lattice", self.lib.ed_set_lattice(self.h, Lx, Ly, Lz, Lt, Ne))

    def set_gauge(self, ptr: int, dtype: int, on_device: bool) -> None:
        self._chk("ed_set_gauge", self.lib.ed_set_gauge(self.h, _VP(ptr), dtype, int(on_device)))

    def set_eigvecs(self, ptr: int, dtype: int, on_device: bool) -> None:
        self._chk("ed_set_eigvecs", self.lib.ed_set_eigvecs(self.h, _VP(ptr), dtype, int(on_device)))

    def set_momenta(self, moms) -> None:
        flat = [int(c) for m in moms for c in m]
        arr = (ctypes.c_int32 * len(flat))(*flat)
        self._chk("ed_set_momenta", self.lib.ed_set_momenta(self.h, arr, len(flat) // 3))

    def set_derivatives(self, num_nabla: int) -> None:
        self._chk("ed_set_derivatives", self.lib.ed_set_derivatives(self.h, num_nabla))

    def set_displacement(self, distance: int) -> None:
        self._chk("ed_set_displacement", self.lib.ed_set_displacement(self.h, distance))

    def set_weights(self, coeff_ptr: int | None) -> None:
        p = ctypes.cast(_VP(coeff_ptr), ctypes.POINTER(ctypes.c_double)) if coeff_ptr else None
        self._chk("ed_set_weights", self.lib.ed_set_weights(self.h, p))

    def set_option(self, option: int, value: int) -> None:
        self._chk("ed_set_option", self.lib.ed_set_option(self.h, option, int(value)))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64()
        self._chk("ed_get_option", self.lib.ed_get_option(self.h, option, ctypes.byref(v)))
        return v.value

    def tune(self, mode: int, t0: int, nt: int, reps: int = 1) -> float:
        """ed_tune: time the contraction's tile orders and kernels, keep the fastest; ms per run."""
        ms = ctypes.c_double()
        self._chk("ed_tune", self.lib.ed_tune(self.h, mode, t0, nt, reps, ctypes.byref(ms)))
        return ms.value

    def compute(self, t0: int, nt: int, mode: int, out_ptr: int, out_on_device: bool) -> None:
        self._chk("ed_compute", self.lib.ed_compute(self.h, t0, nt, mode, _VP(out_ptr), int(out_on_device)))

    def compute_noalloc(self, t0: int, nt: int, mode: int, out_ptr: int) -> None:
        """ed_compute_noalloc: device output, fixed workspace, no allocation (enqueues only)."""
        self._chk("ed_compute_noalloc", self.lib.ed_compute_noalloc(self.h, t0, nt, mode, _VP(out_ptr)))

    def synchronize(self) -> None:
        self._chk("ed_synchronize", self.lib.ed_synchronize(self.h))

    def orthonormality(self) -> float:
        """FP64 max|V^dag V - 1| of the eigenvectors last set."""
        v = ctypes.c_double()
        self._chk("ed_orthonormality", self.lib.ed_orthonormality(self.h, ctypes.byref(v)))
        return v.value

    def profile_read(self, kernel: int):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk("ed_profile_read", self.lib.ed_profile_read(self.h, kernel, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def query(self, what: int) -> int:
        v = ctypes.c_int64()
        self._chk("ed_query", self.lib.ed_query(self.h, what, ctypes.byref(v)))
        return v.value

    def profile_reset(self) -> None:
        self._chk("ed_profile_reset", self.lib.ed_profile_reset(self.h))

    def project_su3(self, max_iter: int = 100) -> int:
        it = _I()
        self._chk("ed_project_su3", self.lib.ed_project_su3(self.h, int(max_iter), ctypes.byref(it)))
        return it.value

    def stout_smear(self, nstep: int, rho: float) -> None:
        self._chk("ed_stout_smear", self.lib.ed_stout_smear(self.h, int(nstep), float(rho)))

    def laplacian(self, t: int, f_ptr: int, out_ptr: int, nvec: int, on_device: bool) -> None:
        self._chk("ed_laplacian", self.lib.ed_laplacian(self.h, int(t), _VP(f_ptr), _VP(out_ptr), int(nvec), int(on_device)))

    def get_gauge(self, ptr: int, on_device: bool) -> None:
        self._chk("ed_get_gauge", self.lib.ed_get_gauge(self.h, _VP(ptr), int(on_device)))

    def backward(self, t0: int, nt: int, grad_ptr: int, grad_on_device: bool, gv_ptr, gu_ptr, out_on_device: bool):
        self._chk(
            "ed_backward",
            self.lib.ed_backward(
                self.h, t0, nt, _VP(grad_ptr), int(grad_on_device), _VP(gv_ptr or 0), _VP(gu_ptr or 0), int(out_on_device)
            ),
        )


class Batch:
    """A batch of contexts computed by one graph launch (``ed_batch_*``)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _VP()
        rc = self.lib.ed_batch_create(device, ctypes.byref(h))
        if rc != ED_OK:
            raise NativeError("ed_batch_create", rc, "could not create a batch (is a GPU visible?)")
        self.h = h
        self.device = device
        self._members = []  # keeps the contexts alive while the batch refers to them

    def _chk(self, name: str, rc: int) -> None:
        if rc != ED_OK:
            msg = self.lib.ed_batch_last_error(self.h)
            raise NativeError(name, rc, msg.decode() if msg else "")

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.ed_batch_destroy(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, ctx: Context, mode: int, out_ptr: int) -> None:
        self._chk("ed_batch_add", self.lib.ed_batch_add(self.h, ctx.h, mode, _VP(out_ptr)))
        self._members.append(ctx)

    def set_option(self, option: int, value: int) -> None:
        self._chk("ed_batch_set_option", self.lib.ed_batch_set_option(self.h, option, int(value)))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64()
        self._chk("ed_batch_get_option", self.lib.ed_batch_get_option(self.h, option, ctypes.byref(v)))
        return v.value

    def run(self) -> None:
        self._chk("ed_batch_run", self.lib.ed_batch_run(self.h))

    def synchronize(self) -> None:
        self._chk("ed_batch_synchronize", self.lib.ed_batch_synchronize(self.h))

    def time(self, reps: int = 1) -> float:
        ms = ctypes.c_double()
        self._chk("ed_batch_time", self.lib.ed_batch_time(self.h, int(reps), ctypes.byref(ms)))
        return ms.value

    def tune(self, reps: int = 1) -> float:
        ms = ctypes.c_double()
        self._chk("ed_batch_tune", self.lib.ed_batch_tune(self.h, int(reps), ctypes.byref(ms)))
        return ms.value
