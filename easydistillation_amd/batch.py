"""Batched elemental generation over several lattices in one graph launch.

The reference computes one configuration and one timeslice per ``calc(t)``
call (lattice/generator/elemental.py:290-338) and its callers loop
(tests/test_elemental.py:43-45).  ``ElementalBatch`` takes several loaded
``ElementalGenerator`` instances -- lattices of different size, Ne, momenta
and derivative depth -- and computes every timeslice of all of them with one
replay of a captured hipGraph: per lattice the staging and hop kernels, then
one persistent grouped contraction over the work of every lattice.  The
launch shape and work order are picked by the library's autotuner
(``tune``).
"""

from __future__ import annotations

from typing import List, Sequence

import torch

from . import _native
from .elemental import _MODES, ElementalGenerator

__all__ = ["ElementalBatch"]


class ElementalBatch:
    def __init__(self, generators: Sequence[ElementalGenerator], modes: Sequence[str] | None = None, device: int = 0):
        if not generators:
            raise ValueError("at least one generator is required")
        modes = list(modes) if modes is not None else [g.mode for g in generators]
        if len(modes) != len(generators):
            raise ValueError("one mode per generator")
        self.generators = list(generators)
        self.modes = modes
        self.batch = _native.Batch(device)
        self.outputs: List[torch.Tensor] = []
        for g, m in zip(self.generators, modes):
            if not g._loaded:
                raise RuntimeError("every generator must be loaded (load or load_arrays) before batching")
            Lx, Ly, Lz, Lt = g.latt_size
            out = torch.empty(
                (Lt, g.num_derivative, g.num_momentum, g.Ne, g.Ne), dtype=torch.complex128, device=f"cuda:{device}"
            )
            self.batch.add(g.context, _MODES[m], out.data_ptr())
            self.outputs.append(out)

    def run(self) -> List[torch.Tensor]:
        """Compute every timeslice of every generator; returns the [Lt, D, P, Ne, Ne] device tensors."""
        torch.cuda.current_stream().synchronize()
        self.batch.run()
        self.batch.synchronize()
        return self.outputs

    def tune(self, reps: int = 1) -> float:
        """Autotune launch shape and work order; returns the best time per run in ms."""
        return self.batch.tune(reps)

    def time(self, reps: int = 1) -> float:
        """Mean time per run in ms (one warm-up run first)."""
        return self.batch.time(reps)

    def set_option(self, option: int, value: int) -> None:
        self.batch.set_option(option, value)

    def option(self, option: int) -> int:
        return self.batch.get_option(option)

    def elementals(self) -> List[int]:
        """Elementals per run of each generator (Lt * P * D * Ne^2)."""
        return [g.latt_size[3] * g.num_momentum * g.num_derivative * g.Ne * g.Ne for g in self.generators]
