"""Drop-in ``DisplacementElementalGenerator`` computing on MI355X.

Mirrors lattice/generator/displacement_elemental.py:
  * constructor (``:11-51``): lattice size, gauge field and eigenvector
    containers, ``distance``, momentum list;
  * ``load(key)`` (``:73-76``), ``project_SU3`` / ``stout_smear`` (``:98-214``);
  * ``calc(t)`` (``:78-96``): E[n, p, e, f] for n = 0 .. distance as complex128
    [distance + 1, P, Ne, Ne], with right vector M_n the mean of the six
    straight Wilson-line transports of V over n sites (``_D``, ``:53-71``) and
    left vector V.

The transports run in the library's path kernel (one launch per distance,
all six directions fused), the contraction in the same MFMA kernels as the
derivative elementals.
"""

from __future__ import annotations

from typing import List, Tuple

from .elemental import ElementalGenerator


class DisplacementElementalGenerator(ElementalGenerator):
    def __init__(
        self,
        latt_size: List[int],
        gauge_field=None,
        eigenvector=None,
        distance: int = 0,
        momentum_list: List[Tuple[int]] = [(0, 0, 0)],
        *,
        Ne: int | None = None,
        mode: str = "fp64",
        device: int = 0,
    ) -> None:
        if not 0 <= int(distance) <= 64:
            raise ValueError("distance must be in [0, 64]")
        super().__init__(latt_size, gauge_field, eigenvector, 0, momentum_list, Ne=Ne, mode=mode, device=device)
        self.distance = int(distance)
        self._ctx.set_displacement(self.distance)
        # outputs are indexed by distance n = 0 .. distance
        self.num_derivative = self.distance + 1
        self.derivative_list = None

    def calc(self, t: int, mode: str | None = None):
        """E[n, p, e, f] of timeslice t, complex128 [distance + 1, P, Ne, Ne] (displacement_elemental.py:78-96)."""
        return super().calc(t, mode)
