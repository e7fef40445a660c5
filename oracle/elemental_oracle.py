"""CPU oracle for distillation elemental generation.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the package
``easydistillation_amd``) may import this module; it is imported only by
``tests/``, by ``__graft_entry__.smoke()`` as the checker, and by the
``cpu_baseline`` leg of ``bench.py``.

This is a clean-room numpy restatement (complex128 arithmetic) of the
reference's elemental generation:

* ``derivative``        follows ``lattice/insertion/derivative.py:23-33``
* ``num_derivative``    follows ``lattice/generator/elemental.py:48``
* ``momentum_phase``    follows ``lattice/insertion/phase.py:11-13,41-46``
                        (phase exp(+2 pi i n.x / L), x fastest in [z, y, x])
* ``hop`` / ``nD``      follow ``lattice/generator/elemental.py:279-288``
                        (nabla_mu psi = U_mu(x) psi(x+mu) - U_mu^dag(x-mu) psi(x-mu))
* ``splits``            follows ``lattice/generator/elemental.py:309-321``
                        (bit j of the pick mask sends the j-th direction to the
                        right vector; the left list is applied reversed;
                        coefficient (-1)^{#right})
* ``elemental_timeslice`` follows ``lattice/generator/elemental.py:290-338``
                        including the complex64 rounding of V (``:55,297-298``)
                        and the blending Hadamard factor (``:330-337``)
* ``blending_coeff``    follows ``lattice/generator/elemental.py:61-100``
* ``displacement_right`` / ``displacement_timeslice`` follow
                        ``lattice/generator/displacement_elemental.py:53-71,78-96``
                        (six straight paths of length n from V, their mean as
                        the right vector, V as the left vector)

Pinning (every function is checked by tests that do not retype its formula):
``derivative`` by the exhaustive index table and the operator names of
``lattice/insertion/derivative.py:1-20``; ``hop`` by the diagonal constant
background closed form and gauge covariance; ``splits`` by the free-field
closed form and the Hermiticity relations; ``elemental_timeslice`` by
``brute_force_timeslice`` (explicit per-site loops), the V^dag V known answer,
translation covariance and gauge invariance; ``blending_coeff`` by a
hand-computed table; ``displacement_timeslice`` by ``brute_force_displacement``
(explicit Wilson lines), the free-field closed form, the constant diagonal
background and gauge invariance.  See ``tests/test_oracle.py``.

Array layouts (reference conventions):
  U_t  [3, Lz, Ly, Lx, 3, 3]   spatial links of one timeslice, U[mu][z,y,x][a,b]
  V_t  [Ne, Lz, Ly, Lx, 3]     eigenvectors of one timeslice
  E_t  [D, P, Ne, Ne]          output of one timeslice (complex128)
"""

from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

__all__ = [
    "derivative",
    "num_derivative",
    "derivative_list",
    "momentum_phase",
    "hop",
    "nD",
    "splits",
    "elemental_timeslice",
    "elemental",
    "brute_force_timeslice",
    "blending_coeff",
    "displacement_right",
    "displacement_timeslice",
    "displacement_elemental",
    "brute_force_displacement",
]


def derivative(n: int) -> Tuple[int, ...]:
    """Direction tuple of derivative index ``n`` (lattice/insertion/derivative.py:23-33).

    Index n enumerates base-3 strings by increasing length: length-k strings
    occupy indices [(3^k-1)/2, (3^(k+1)-1)/2).  The digits of n - (3^k-1)/2
    are read least-significant first and the resulting list is reversed, so
    the most significant digit is the first direction applied.
    """
    if n < 0:
        raise ValueError("derivative index must be >= 0")
    k = 0
    while n >= 3**k:
        n -= 3**k
        k += 1
    digits = []
    for _ in range(k):
        digits.append(n % 3)
        n //= 3
    return tuple(reversed(digits))


def num_derivative(num_nabla: int) -> int:
    """Number of derivative combinations with up to ``num_nabla`` nablas (elemental.py:48)."""
    return (3 ** (num_nabla + 1) - 1) // 2


def derivative_list(num_nabla: int) -> List[Tuple[int, ...]]:
    return [derivative(n) for n in range(num_derivative(num_nabla))]


def momentum_phase(latt_size: Sequence[int], n: Sequence[int]) -> np.ndarray:
    """exp(+2 pi i (nx x/Lx + ny y/Ly + nz z/Lz)) on [Lz, Ly, Lx] (phase.py:11-13,41-46)."""
    Lx, Ly, Lz = latt_size[0], latt_size[1], latt_size[2]
    nx, ny, nz = n
    x = np.arange(Lx).reshape(1, 1, Lx)
    y = np.arange(Ly).reshape(1, Ly, 1)
    z = np.arange(Lz).reshape(Lz, 1, 1)
    arg = 2j * np.pi * (nx * x / Lx + ny * y / Ly + nz * z / Lz)
    return np.exp(arg).astype(np.complex128)


def _axis(mu: int) -> int:
    # direction 0=x, 1=y, 2=z acts on axis 3-mu of [e, z, y, x, c] (elemental.py:283,286)
    return 3 - mu


def hop(V: np.ndarray, U_t: np.ndarray, mu: int) -> np.ndarray:
    """One covariant derivative along ``mu`` (elemental.py:282-287), periodic boundaries.

    (nabla_mu V)(x) = U_mu(x) V(x + mu) - U_mu(x - mu)^dag V(x - mu)
    """
    ax = _axis(mu)
    Vf = np.roll(V, -1, ax)                               # V(x + mu)
    forward = np.einsum("zyxab,ezyxb->ezyxa", U_t[mu], Vf)
    UdV = np.einsum("zyxba,ezyxb->ezyxa", U_t[mu].conj(), V)  # U^dag(x) V(x)
    backward = np.roll(UdV, 1, ax)                         # evaluated at x - mu
    return forward - backward


def nD(V: np.ndarray, U_t: np.ndarray, deriv: Sequence[int]) -> np.ndarray:
    """Apply the directions of ``deriv`` left to right (elemental.py:279-288)."""
    out = V
    for mu in deriv:
        out = hop(out, U_t, mu)
    return out


def splits(deriv: Sequence[int]) -> List[Tuple[Tuple[int, ...], Tuple[int, ...], int]]:
    """(left_dirs, right_dirs, coeff) for every split of ``deriv`` (elemental.py:309-321).

    Bit j of the pick mask sends deriv[j] to the right list; the others go to
    the left list, which is then reversed.  coeff = (-1)^{len(right)}.
    """
    out = []
    for pick in range(2 ** len(deriv)):
        right, left = [], []
        for j, mu in enumerate(deriv):
            if (pick >> j) & 1:
                right.append(mu)
            else:
                left.append(mu)
        out.append((tuple(left[::-1]), tuple(right), (-1) ** len(right)))
    return out


def blending_coeff(Ne: int, dilution) -> np.ndarray:
    """Stochastic blending weights [Ne, Ne] (elemental.py:61-100).

    dilution = (tot_list, used) with used an int or a list; the block (i, j)
    gets tot_i*tot_j/(used_i*used_j) off the diagonal blocks and, on a
    diagonal block, tot/used on its diagonal and tot(tot-1)/(used(used-1))
    elsewhere.
    """
    tot = list(dilution[0])
    used = [dilution[1]] * len(tot) if isinstance(dilution[1], int) else list(dilution[1])
    if len(used) != len(tot) or any(u > t for u, t in zip(used, tot)) or sum(used) != Ne:
        raise AssertionError("invalid dilution tuple")
    W = np.zeros((Ne, Ne))
    offs = np.concatenate([[0], np.cumsum(used)])
    for i in range(len(tot)):
        for j in range(len(tot)):
            si, sj = slice(offs[i], offs[i + 1]), slice(offs[j], offs[j + 1])
            if i != j:
                W[si, sj] = tot[i] * tot[j] / used[i] / used[j]
            else:
                c1 = tot[i] / used[i]
                c2 = c1 * (tot[i] - 1) / (used[i] - 1)  # used == 1 raises, as the reference does
                blk = np.full((used[i], used[i]), c2)
                np.fill_diagonal(blk, c1)
                W[si, sj] = blk
    return W


def elemental_timeslice(
    V_t: np.ndarray,
    U_t: np.ndarray,
    latt_size: Sequence[int],
    num_nabla: int,
    momentum_list: Sequence[Sequence[int]],
    stocastic_coeff: np.ndarray | None = None,
    round_v_to_c64: bool = True,
    derivative_indices: Sequence[int] | None = None,
) -> np.ndarray:
    """E_t[d, p, e, f] for one timeslice (elemental.py:290-338), complex128.

    V is rounded to complex64 first, as the reference stages it in a c8
    buffer (elemental.py:55,297-298); everything after is complex128.
    The per-momentum contraction "zyx,ezyxc,fzyxc->ef" is evaluated as a
    complex GEMM (BLAS) over the flattened (site, colour) index.
    ``derivative_indices`` restricts the output to those derivatives (in the
    given order; the default is all of them), for spot checks on lattices
    where the whole timeslice would take minutes.
    """
    V = np.asarray(V_t)
    if round_v_to_c64:
        V = V.astype(np.complex64)
    V = V.astype(np.complex128)
    U_t = np.asarray(U_t, dtype=np.complex128)
    Ne = V.shape[0]
    dl = derivative_list(num_nabla)
    if derivative_indices is not None:
        dl = [dl[i] for i in derivative_indices]
    P = len(momentum_list)
    phases = [momentum_phase(latt_size, m) for m in momentum_list]
    E = np.zeros((len(dl), P, Ne, Ne), dtype=np.complex128)
    cache = {(): V}

    def W(dirs):
        if dirs not in cache:
            cache[dirs] = nD(V, U_t, dirs)
        return cache[dirs]

    for di, deriv in enumerate(dl):
        for left_dirs, right_dirs, coeff in splits(deriv):
            L = W(left_dirs).conj()
            R = W(right_dirs).reshape(Ne, -1)
            for pi, ph in enumerate(phases):
                LP = (L * ph[None, :, :, :, None]).reshape(Ne, -1)
                E[di, pi] += coeff * (LP @ R.T)
    if stocastic_coeff is not None:
        E *= np.asarray(stocastic_coeff)[None, None]
    return E


def elemental(
    V: np.ndarray,
    U: np.ndarray,
    latt_size: Sequence[int],
    num_nabla: int,
    momentum_list: Sequence[Sequence[int]],
    stocastic_coeff: np.ndarray | None = None,
) -> np.ndarray:
    """All timeslices, stacked as [Lt, D, P, Ne, Ne] (the caller's stacking, tests/test_elemental.py:37-44).

    U is [3, Lt, Lz, Ly, Lx, 3, 3] (elemental.py:103), V is [Lt, Ne, Lz, Ly, Lx, 3].
    The on-disk layout [D, P, Lt, Ne, Ne] is ``out.transpose(1, 2, 0, 3, 4)``.
    """
    Lt = V.shape[0]
    return np.stack(
        [
            elemental_timeslice(V[t], U[:, t], latt_size, num_nabla, momentum_list, stocastic_coeff)
            for t in range(Lt)
        ]
    )


def brute_force_timeslice(
    V_t: np.ndarray,
    U_t: np.ndarray,
    latt_size: Sequence[int],
    num_nabla: int,
    momentum_list: Sequence[Sequence[int]],
) -> np.ndarray:
    """Explicit per-site, per-colour loops (pure Python, tiny lattices only).

    Independent of ``hop``/``nD``/``elemental_timeslice``: neighbours are
    found by modular index arithmetic, matrix-vector products are written as
    colour loops and the phase is computed per site from its integer
    coordinates.  Directions, split rule and phase sign follow Appendix-level
    statements of the reference semantics (elemental.py:279-338, phase.py:41-46).
    """
    import cmath

    Lx, Ly, Lz = latt_size[0], latt_size[1], latt_size[2]
    dims = (Lx, Ly, Lz)
    V = np.asarray(V_t).astype(np.complex64).astype(np.complex128)
    Ne = V.shape[0]

    def site_shift(z, y, x, mu, s):
        c = [x, y, z]
        c[mu] = (c[mu] + s) % dims[mu]
        return c[2], c[1], c[0]

    def apply_hop(F, mu):
        G = np.zeros_like(F)
        for e in range(Ne):
            for z in range(Lz):
                for y in range(Ly):
                    for x in range(Lx):
                        zf, yf, xf = site_shift(z, y, x, mu, +1)
                        zb, yb, xb = site_shift(z, y, x, mu, -1)
                        for a in range(3):
                            acc = 0j
                            for b in range(3):
                                acc += U_t[mu, z, y, x, a, b] * F[e, zf, yf, xf, b]
                                acc -= np.conj(U_t[mu, zb, yb, xb, b, a]) * F[e, zb, yb, xb, b]
                            G[e, z, y, x, a] = acc
        return G

    def applied(dirs):
        F = V
        for mu in dirs:
            F = apply_hop(F, mu)
        return F

    D = num_derivative(num_nabla)
    E = np.zeros((D, len(momentum_list), Ne, Ne), dtype=np.complex128)
    for di in range(D):
        deriv = derivative(di)
        for pick in range(2 ** len(deriv)):
            right = [deriv[j] for j in range(len(deriv)) if (pick >> j) & 1]
            left = [deriv[j] for j in range(len(deriv)) if not (pick >> j) & 1][::-1]
            coeff = -1 if len(right) % 2 else 1
            Lv, Rv = applied(left), applied(right)
            for pi, (nx, ny, nz) in enumerate(momentum_list):
                for z in range(Lz):
                    for y in range(Ly):
                        for x in range(Lx):
                            ph = cmath.exp(2j * cmath.pi * (nx * x / Lx + ny * y / Ly + nz * z / Lz))
                            for e in range(Ne):
                                for f in range(Ne):
                                    s = 0j
                                    for c in range(3):
                                        s += np.conj(Lv[e, z, y, x, c]) * Rv[f, z, y, x, c]
                                    E[di, pi, e, f] += coeff * ph * s
    return E


# --------------------------------------------------------------------------
# Displacement elementals (lattice/generator/displacement_elemental.py)
# --------------------------------------------------------------------------
def displacement_right(V: np.ndarray, U_t: np.ndarray, distance: int) -> List[np.ndarray]:
    """Right operands M_0 .. M_distance of displacement_elemental.py:53-71.

    M_0 = V; for n >= 1, M_n is the mean over the six straight paths of
    length n: forward P_mu(x) = U_mu(x) P_mu(x + mu) and backward
    P_-mu(x) = U_mu(x - mu)^dag P_-mu(x - mu), both started from V.
    """
    out = [V]
    fwd = [V, V, V]
    bwd = [V, V, V]
    for n in range(1, distance + 1):
        fwd = [np.einsum("zyxab,ezyxb->ezyxa", U_t[d], np.roll(fwd[d], -1, _axis(d))) for d in range(3)]
        bwd = [np.roll(np.einsum("zyxba,ezyxb->ezyxa", U_t[d].conj(), bwd[d]), 1, _axis(d)) for d in range(3)]
        out.append(sum(fwd + bwd) / 6.0)
    return out


def displacement_timeslice(
    V_t: np.ndarray,
    U_t: np.ndarray,
    latt_size: Sequence[int],
    distance: int,
    momentum_list: Sequence[Sequence[int]],
    round_v_to_c64: bool = True,
) -> np.ndarray:
    """E_t[n, p, e, f] = sum_x Phi_p(x) conj(V[e](x)) . M_n[f](x) (displacement_elemental.py:78-96)."""
    V = np.asarray(V_t)
    if round_v_to_c64:
        V = V.astype(np.complex64)
    V = V.astype(np.complex128)
    Ne = V.shape[0]
    rights = displacement_right(V, np.asarray(U_t, dtype=np.complex128), distance)
    E = np.zeros((distance + 1, len(momentum_list), Ne, Ne), dtype=np.complex128)
    L = V.conj()
    for n, R in enumerate(rights):
        Rm = R.reshape(Ne, -1)
        for pi, m in enumerate(momentum_list):
            ph = momentum_phase(latt_size, m)
            E[n, pi] = (L * ph[None, :, :, :, None]).reshape(Ne, -1) @ Rm.T
    return E


def displacement_elemental(V, U, latt_size, distance, momentum_list):
    """[Lt, distance + 1, P, Ne, Ne]; U [3, Lt, ...], V [Lt, Ne, ...]."""
    return np.stack(
        [displacement_timeslice(V[t], U[:, t], latt_size, distance, momentum_list) for t in range(V.shape[0])]
    )


def brute_force_displacement(V_t, U_t, latt_size, distance, momentum_list):
    """Explicit Wilson lines per site and direction (tiny lattices only)."""
    Lx, Ly, Lz = latt_size[0], latt_size[1], latt_size[2]
    L = (Lx, Ly, Lz)
    V = np.asarray(V_t).astype(np.complex64).astype(np.complex128)
    Ne = V.shape[0]
    E = np.zeros((distance + 1, len(momentum_list), Ne, Ne), complex)

    def link(d, s):
        return U_t[d, s[2] % Lz, s[1] % Ly, s[0] % Lx]

    for z in range(Lz):
        for y in range(Ly):
            for x in range(Lx):
                site = np.array([x, y, z])
                for n in range(distance + 1):
                    if n == 0:
                        M = V[:, z, y, x, :]
                    else:
                        M = np.zeros((Ne, 3), complex)
                        for d in range(3):
                            step = np.eye(3, dtype=int)[d]
                            # forward: U(x) U(x+d) ... U(x+(n-1)d) V(x+nd)
                            W = np.eye(3, dtype=complex)
                            for k in range(n):
                                W = W @ link(d, site + k * step)
                            t = (site + n * step) % L
                            M += (W @ V[:, t[2], t[1], t[0], :].T).T
                            # backward: U(x-d)^dag U(x-2d)^dag ... V(x-nd)
                            W = np.eye(3, dtype=complex)
                            for k in range(1, n + 1):
                                W = W @ link(d, site - k * step).conj().T
                            t = (site - n * step) % L
                            M += (W @ V[:, t[2], t[1], t[0], :].T).T
                        M /= 6.0
                    for pi, m in enumerate(momentum_list):
                        ph = np.exp(2j * np.pi * (m[0] * x / Lx + m[1] * y / Ly + m[2] * z / Lz))
                        E[n, pi] += ph * (V[:, z, y, x, :].conj() @ M.T)
    return E
