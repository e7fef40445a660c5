"""Link preprocessing on the device (ed_project_su3, ed_stout_smear) against the oracle.

References: lattice/generator/elemental.py:107-117 (project_SU3), :175-242 and
:264-277 (stout_smear).  Bars: links within 1e-12 (max abs) of the FP64
oracle; elementals of smeared links within the FP64 bar of 1e-12.
"""

import sys

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import synthetic as S
from oracle import elemental_oracle as O
from oracle import gauge_oracle as G

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def generator(latt, Ne, U, V, num_nabla=1, moms=None):
    moms = moms or S.unit_momenta()
    g = ed.ElementalGenerator(latt, None, None, num_nabla, moms, Ne=Ne)
    g.load_arrays(U, V)
    return g


@pytest.mark.parametrize("latt,nstep,rho", [([4, 6, 5, 3], 1, 0.1), ([8, 8, 8, 2], 3, 0.125), ([3, 3, 3, 1], 2, 0.3)])
def test_stout_matches_oracle(latt, nstep, rho):
    say(f"stout latt={latt} nstep={nstep} rho={rho}")
    rng = np.random.default_rng(21)
    U = S.gauge_np(rng, latt, "haar")
    V = S.eigvecs_np(rng, latt, 4)
    g = generator(latt, 4, U, V)
    g.stout_smear(nstep, rho)
    got = g.gauge()
    want = G.stout_smear(U, nstep, rho)
    err = np.abs(got - want).max()
    say(f"  max|dU|={err:.3e}")
    assert err < 1e-12


def test_stout_zero_steps_and_zero_rho():
    latt = [4, 4, 4, 2]
    rng = np.random.default_rng(2)
    U = S.gauge_np(rng, latt, "haar")
    g = generator(latt, 3, U, S.eigvecs_np(rng, latt, 3))
    g.stout_smear(0, 0.1)
    assert np.abs(g.gauge() - U).max() == 0
    g.stout_smear(2, 0.0)
    assert np.abs(g.gauge() - U).max() == 0


def test_project_matches_oracle():
    latt = [4, 5, 3, 2]
    rng = np.random.default_rng(6)
    U0 = S.gauge_np(rng, latt, "haar")
    M = U0 @ (np.eye(3) + 0.05 * (rng.normal(size=U0.shape) + 1j * rng.normal(size=U0.shape)))
    g = generator(latt, 3, M, S.eigvecs_np(rng, latt, 3))
    it = g.project_SU3()
    want, it_ref = G.project_su3(M)
    err = np.abs(g.gauge() - want).max()
    say(f"project: iterations {it} (oracle {it_ref}) max|dU|={err:.3e}")
    assert abs(it - it_ref) <= 1
    assert err < 1e-13


def test_project_cap_raises():
    latt = [3, 3, 3, 1]
    rng = np.random.default_rng(8)
    M = S.gauge_np(rng, latt, "haar") * 1.3
    g = generator(latt, 3, M, S.eigvecs_np(rng, latt, 3))
    with pytest.raises(RuntimeError):
        g.project_SU3(max_iter=1)


def test_elemental_after_smearing():
    latt = [4, 5, 6, 2]
    Ne, nn = 6, 2
    rng = np.random.default_rng(31)
    U = S.gauge_np(rng, latt, "haar")
    V = S.eigvecs_np(rng, latt, Ne)
    moms = S.momenta_by_norm(2)
    g = generator(latt, Ne, U, V, nn, moms)
    g.stout_smear(2, 0.1)
    E = g.calc_all()
    Us = G.stout_smear(U, 2, 0.1)
    Eref = O.elemental(V, Us, latt, nn, moms)
    err = np.abs(E - Eref).max() / np.abs(Eref).max()
    say(f"elemental after stout: rel err {err:.3e}")
    assert err < 1e-12
