"""Generate the committed golden fixtures of the elemental path.

The reference's own golden files (tests/weak_field.elemental.npy and the
inputs it was made from) are Git-LFS pointers without content, so the
fixtures here are seeded synthetic inputs plus outputs of the CPU oracle.
Each oracle function they rely on is pinned independently in
tests/test_oracle.py (closed forms, invariants, brute force).

  config_a.npz     4^3 x 8, Ne = 20, Haar links (seed 0), num_nabla = 1,
                   momenta (0,0,0) and the six unit vectors;
                   E in the file layout [D=4, P=7, Lt=8, Ne, Ne]
  ref_shape.npz    the reference test's setup (tests/test_elemental.py:15-26:
                   num_nabla = 2 and its 7 momenta) on 4^3 x 2 with smooth
                   links (eps = 0.3, seed 1); E [13, 7, 2, 20, 20]
  brute_force.npz  anisotropic 3 x 4 x 5 x 2, Ne = 5, num_nabla = 2, five
                   momenta including negative and repeated entries; E from
                   the per-site loop oracle brute_force_timeslice

Run from the repository root: python tests/golden/make_fixtures.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from easydistillation_amd import synthetic as S  # noqa: E402
from oracle import elemental_oracle as O  # noqa: E402

REF_TEST_MOMENTA = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1), (0, 0, 2), (0, 1, 2), (1, 1, 2)]
BRUTE_MOMENTA = [(0, 0, 0), (1, 0, 0), (-1, 2, 0), (0, -1, 1), (1, 0, 0)]


def config_a():
    latt = [4, 4, 4, 8]
    U, V = S.make_inputs_np(latt, 20, "haar", seed=0)
    moms = S.unit_momenta()
    E = O.elemental(V, U, latt, 1, moms)  # [Lt, D, P, Ne, Ne]
    np.savez(
        os.path.join(HERE, "config_a.npz"),
        latt=np.array(latt),
        num_nabla=np.array(1),
        momenta=np.array(moms, dtype=np.int32),
        U=U,
        V=V,
        E=E.transpose(1, 2, 0, 3, 4),
    )


def ref_shape():
    latt = [4, 4, 4, 2]
    U, V = S.make_inputs_np(latt, 20, "smooth", seed=1)
    E = O.elemental(V, U, latt, 2, REF_TEST_MOMENTA)
    np.savez(
        os.path.join(HERE, "ref_shape.npz"),
        latt=np.array(latt),
        num_nabla=np.array(2),
        momenta=np.array(REF_TEST_MOMENTA, dtype=np.int32),
        U=U,
        V=V,
        E=E.transpose(1, 2, 0, 3, 4),
    )


def brute_force():
    latt = [3, 4, 5, 2]
    U, V = S.make_inputs_np(latt, 5, "haar", seed=2)
    E = np.stack([O.brute_force_timeslice(V[t], U[:, t], latt, 2, BRUTE_MOMENTA) for t in range(latt[3])])
    np.savez(
        os.path.join(HERE, "brute_force.npz"),
        latt=np.array(latt),
        num_nabla=np.array(2),
        momenta=np.array(BRUTE_MOMENTA, dtype=np.int32),
        U=U,
        V=V,
        E=E.transpose(1, 2, 0, 3, 4),
    )


if __name__ == "__main__":
    config_a()
    ref_shape()
    brute_force()
    print("fixtures written to", HERE)
