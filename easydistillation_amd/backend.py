"""Backend selection (mirrors lattice/backend.py:7-33).

The reference switches the array module between numpy and cupy.  This
package computes on MI355X through its HIP library only; the backend name
selects what ``ElementalGenerator.calc`` returns:

* ``"hip"``   (default) -- numpy arrays in host memory;
* ``"torch"`` -- torch tensors left in device memory (HBM).

``"numpy"`` and ``"cupy"`` are the reference's own backends and are
rejected here with an error rather than served by a CPU fallback.
"""

from __future__ import annotations

_BACKEND = "hip"
_ALLOWED = ("hip", "torch")


def set_backend(backend) -> None:
    global _BACKEND
    if not isinstance(backend, str):
        backend = backend.__name__
    name = backend.lower()
    if name not in _ALLOWED:
        raise ValueError(
            f'backend must be one of {_ALLOWED}; "{name}" is served by the reference package, '
            "this package has no CPU path"
        )
    _BACKEND = name


def get_backend() -> str:
    return _BACKEND
