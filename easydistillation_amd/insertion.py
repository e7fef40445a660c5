"""Derivative index map and momentum phases of the elemental path.

``derivative`` and ``output`` follow lattice/insertion/derivative.py:1-33;
``MomentumPhase.get`` follows lattice/insertion/phase.py:11-13,41-46;
``momentum_dict`` / ``mom_dict_to_list`` give the momentum orders of
lattice/insertion/mom_dict.py:1-175 (the order in which downstream readers
of an elemental file, insertion/__init__.py:195-198, index momenta).
These are host-side helpers: the HIP library rebuilds the same index map
and an FP64 phase table itself.
"""

from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

_NAMES = ("dx", "dy", "dz")


def _level(n: int) -> Tuple[int, int]:
    """(k, r): n is the r-th string of length k (strings of length k start at (3^k-1)/2)."""
    if n < 0:
        raise AssertionError("derivative index must be a non-negative int")
    k = 0
    while n >= 3**k:
        n -= 3**k
        k += 1
    return k, n


def derivative(n: int) -> Tuple[int, ...]:
    """Directions (0=x, 1=y, 2=z) of derivative index n, in the order they are applied."""
    k, r = _level(int(n))
    digits = []
    for _ in range(k):
        digits.append(r % 3)
        r //= 3
    return tuple(digits[::-1])


def derivative_index(dirs: Sequence[int]) -> int:
    """Inverse of ``derivative``."""
    k = len(dirs)
    v = 0
    for d in dirs:
        if d not in (0, 1, 2):
            raise ValueError("directions are 0, 1, 2")
        v = v * 3 + d
    return (3**k - 1) // 2 + v


def num_derivative(num_nabla: int) -> int:
    return (3 ** (num_nabla + 1) - 1) // 2


def output(derivative_coeff_index) -> str:
    """Operator name such as "dxdy" or "-dz" for (coeff, index) (derivative.py:1-20).

    The name lists the nablas as an operator product, i.e. the last applied
    direction first.
    """
    c, n = derivative_coeff_index
    if not (isinstance(n, int) and n >= 0):
        raise AssertionError("derivative index must be a non-negative int")
    name = "".join(_NAMES[d] for d in reversed(derivative(n)))
    if c == 1:
        return name
    if c == -1:
        return "-" + name
    return str(c) + name


class MomentumPhase:
    """exp(+2 pi i (nx x/Lx + ny y/Ly + nz z/Lz)) on [Lz, Ly, Lx], cached per momentum."""

    def __init__(self, latt_size: List[int]) -> None:
        self.latt_size = list(latt_size)
        self.cache: Dict[Tuple[int, int, int], np.ndarray] = {}

    def get(self, n: Sequence[int]) -> np.ndarray:
        key = tuple(int(v) for v in n)
        if key not in self.cache:
            Lx, Ly, Lz = self.latt_size[:3]
            x = (key[0] * np.arange(Lx) % Lx) / Lx
            y = (key[1] * np.arange(Ly) % Ly) / Ly
            z = (key[2] * np.arange(Lz) % Lz) / Lz
            frac = z[:, None, None] + y[None, :, None] + x[None, None, :]
            self.cache[key] = np.exp(2j * np.pi * frac)
        return self.cache[key]


def momentum_dict(mom: int = 9) -> Dict[int, str]:
    """Momentum table {index: "nx ny nz"} in the reference's order (mom_dict.py:1-175).

    mom = 1: the zero momentum.  mom = 3: the 27 n with components in
    {-1, 0, 1}, by |n|^2, then by which components are nonzero (mask
    4 [nx != 0] + 2 [ny != 0] + [nz != 0], ascending), then + before - in
    x, y, z order.  mom = 9: the 123 n with |n|^2 <= 9, by nz, then ny, then
    nx, ascending.
    """
    import itertools

    if mom == 1:
        moms = [(0, 0, 0)]
    elif mom == 3:
        mask = lambda n: 4 * (n[0] != 0) + 2 * (n[1] != 0) + (n[2] != 0)  # noqa: E731
        moms = sorted(itertools.product((-1, 0, 1), repeat=3),
                      key=lambda n: (sum(c * c for c in n), mask(n), tuple(-c for c in n)))
    elif mom == 9:
        moms = sorted((n for n in itertools.product(range(-3, 4), repeat=3) if sum(c * c for c in n) <= 9),
                      key=lambda n: (n[2], n[1], n[0]))
    else:
        raise ValueError(f"Unknown mom max = {mom}")
    return {i: f"{n[0]} {n[1]} {n[2]}" for i, n in enumerate(moms)}


def mom_dict_to_list(mom: int = 9) -> List[Tuple[int, int, int]]:
    """The momenta of ``momentum_dict(mom)`` as integer triples, in order
    (mom_dict.py:162-175): the momentum_list to pass to ElementalGenerator for
    files read with that table."""
    return [tuple(int(p) for p in v.split(" ")) for v in momentum_dict(mom).values()]
