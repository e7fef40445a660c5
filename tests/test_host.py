"""Host logic and C-ABI surface, on the CPU (no GPU compute calls)."""

import os
import re

import numpy as np
import pytest

import easydistillation_amd as ed
from easydistillation_amd import _native
from oracle import elemental_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "elemental_hip.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ed_\w+)\s*\(", text, re.M)))


def test_library_exports_every_header_symbol():
    lib = _native.load_library()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_native.SIGNATURES)
    assert lib.ed_version().startswith(b"elemental_hip")


def test_binding_constants_match_header_enums():
    """Every ED_* enum value the header declares has the same value in the Python binding."""
    text = open(HEADER).read()
    values = {}
    for body in re.findall(r"enum\s+\w+\s*\{(.*?)\}", text, re.S):
        for name, val in re.findall(r"(ED_\w+)\s*=\s*(-?\d+)", body):
            values[name] = int(val)
    assert {"ED_OPT_WIDE_WS", "ED_Q_MIXED_KERNEL", "ED_KERNEL_CONTRACT"} <= set(values)
    for name, val in values.items():
        assert getattr(_native, name) == val, name


def test_no_silent_cpu_fallback():
    try:
        import torch

        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is visible")
    with pytest.raises(_native.NativeError):
        _native.Context(0)
    with pytest.raises(_native.NativeError):
        ed.ElementalGenerator([4, 4, 4, 2], None, None, 1, [(0, 0, 0)], Ne=4)


def test_c_abi_argument_errors_without_context():
    lib = _native.load_library()
    assert lib.ed_create(0, None) == -1
    assert lib.ed_destroy(None) == -1
    assert lib.ed_compute(None, 0, 1, 0, None, 0) == -1
    assert lib.ed_last_error(None) == b"null context"


def test_derivative_map_matches_oracle():
    for n in range(40):
        assert ed.derivative(n) == O.derivative(n)
        assert ed.derivative_index(ed.derivative(n)) == n
    assert [ed.num_derivative(k) for k in range(4)] == [O.num_derivative(k) for k in range(4)]


def test_blending_weights_match_oracle():
    for Ne, dil in [(5, ([4, 6], [2, 3])), (6, ([10, 10], 3)), (4, ([8], [4])), (7, ([3, 9, 5], [2, 3, 2]))]:
        assert np.array_equal(ed.blending_weights(Ne, dil), O.blending_coeff(Ne, dil))
    with pytest.raises(AssertionError):
        ed.blending_weights(5, ([4, 6], [2, 2]))


def test_momentum_phase_matches_oracle():
    mp = ed.MomentumPhase([4, 5, 3, 8])
    for n in [(0, 0, 0), (1, 0, 0), (0, -2, 1), (3, 7, -5)]:
        assert np.abs(mp.get(n) - O.momentum_phase([4, 5, 3, 8], n)).max() < 1e-13


def test_backend_rejects_cpu_backends():
    with pytest.raises(ValueError):
        ed.set_backend("numpy")
    ed.set_backend("torch")
    assert ed.get_backend() == "torch"
    ed.set_backend("hip")
    assert ed.get_backend() == "hip"


def test_containers_follow_load_contract(tmp_path):
    V = (np.arange(2 * 3 * 2 * 2 * 2 * 3) * (1 + 1j)).reshape(2, 3, 2, 2, 2, 3).astype(np.complex128)
    np.save(tmp_path / "cfg.evec.npy", V)
    ev = ed.EigenvectorNpy(f"{tmp_path}/", ".evec.npy", list(V.shape), 2)
    data = ev.load("cfg")
    assert ev.Ne == 2
    assert np.array_equal(data[1, 0], V[1, 0])
    assert np.array_equal(data[0, :2], V[0, :2])
    Ug = np.zeros((2, 2, 2, 2, 4, 3, 3), complex)
    np.save(tmp_path / "cfg.gauge.npy", Ug)
    g = ed.GaugeFieldNpy(f"{tmp_path}/", ".gauge.npy")
    assert g.load("cfg")[:].shape == Ug.shape
    with pytest.raises(ValueError):
        ed.EigenvectorNpy(f"{tmp_path}/", ".evec.npy", [9, 9], 2).load("cfg")
    E = np.ones((2, 4, 3, 5, 5), complex)
    ed.save_elemental(tmp_path / "cfg.elem.npy", E)
    back = ed.ElementalNpy(f"{tmp_path}/", ".elem.npy", [4, 3, 2, 5, 5], 5).load("cfg")[:]
    assert back.shape == (4, 3, 2, 5, 5)


def test_momentum_sets():
    from easydistillation_amd import synthetic as S

    assert len(S.momenta_by_norm(2)) == 19
    assert len(S.momenta_by_norm(3)) == 27
    assert len(S.momenta_by_norm(4)) == 33
    m = S.momenta_by_norm(4)
    assert m[0] == (0, 0, 0) and set(m) == {tuple(-c for c in n) for n in m}


def test_synthetic_inputs_are_valid():
    from easydistillation_amd import synthetic as S

    for kind in ("haar", "smooth"):
        U, V = S.make_inputs_np([3, 4, 2, 2], 6, kind, seed=3)
        u = U.reshape(-1, 3, 3)
        assert np.abs(u @ u.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-13
        assert np.abs(np.linalg.det(u) - 1).max() < 1e-13
        assert V.dtype == np.complex64
        G = np.einsum("tek,tfk->tef", V.reshape(2, 6, -1).conj().astype(complex), V.reshape(2, 6, -1).astype(complex))
        assert np.abs(G - np.eye(6)).max() < 1e-6


TILE_ORDER_SRC = r"""
#include "ed_kernels.hpp"
#include <cstdio>
#include <set>
#include <utility>
// every tile of an nt x nt grid (or of its upper triangle) exactly once, in
// strips of TS e-tile rows walked f-tile by f-tile
int check(int TS) {
  for (int nt = 1; nt <= 40; ++nt)
    for (int m = 0; m < 2; ++m) {
      const int per = m ? nt * (nt + 1) / 2 : nt * nt;
      std::set<std::pair<int, int>> seen;
      int prev_strip = 0;
      for (int k = 0; k < per; ++k) {
        int te = -1, tf = -1;
        ed::decode_tile(k, nt, nt, m, TS, te, tf);
        if (te < 0 || te >= nt || tf < 0 || tf >= nt || (m && te > tf)) return 1;
        if (te / TS < prev_strip) return 2;  // strips in order
        prev_strip = te / TS;
        seen.insert({te, tf});
      }
      if ((int)seen.size() != per) return 3;
    }
  return 0;
}
int main() {
  const int strips[] = {1, 2, 3, 4, 8, ed::TSTRIP, ed::TSTRIP_FP};
  for (int i = 0; i < 7; ++i)
    if (int r = check(strips[i])) return 10 * i + r;
  int te, tf;
  ed::decode_tile(5, 13, 13, false, 4, te, tf);  // strip 0: (0,0) (1,0) (2,0) (3,0) (0,1) (1,1)
  if (te != 1 || tf != 1) return 100;
  ed::decode_tile(3, 13, 13, true, 4, te, tf);   // triangle strip 0: (0,0) (0,1) (1,1) (0,2)
  if (te != 0 || tf != 2) return 101;
  std::printf("ok\n");
  return 0;
}
"""


def test_contraction_tile_order_is_a_bijection(tmp_path):
    """decode_tile (the contraction's work-item -> tile map) visits every tile of
    the full grid and of the self-partnered upper triangle exactly once."""
    import shutil
    import subprocess

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "tile_order.cpp"
    src.write_text(TILE_ORDER_SRC)
    exe = tmp_path / "tile_order"
    csrc = os.path.join(ROOT, "easydistillation_amd", "csrc")
    subprocess.run([hipcc, "-w", "-O1", "-std=c++17", "--cuda-host-only", "-I", csrc, "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0 and res.stdout.strip() == "ok", res.returncode


def test_momentum_dict_orders_match_the_reference_tables():
    """momentum_dict reproduces the reference's momentum tables
    (lattice/insertion/mom_dict.py:1-175, extracted as data by
    tests/golden/make_mom_tables.py) and each table holds every momentum of
    its set exactly once."""
    import itertools
    import json

    from easydistillation_amd.insertion import mom_dict_to_list, momentum_dict

    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "mom_dict_tables.json")))["tables"]
    for mom in (1, 3, 9):
        ours = mom_dict_to_list(mom)
        assert [list(n) for n in ours] == ref[f"mom{mom}"], mom
        assert momentum_dict(mom) == {i: " ".join(str(c) for c in n) for i, n in enumerate(ours)}
    assert set(mom_dict_to_list(3)) == set(itertools.product((-1, 0, 1), repeat=3))
    ball = {n for n in itertools.product(range(-3, 4), repeat=3) if sum(c * c for c in n) <= 9}
    assert len(mom_dict_to_list(9)) == len(ball) == 123 and set(mom_dict_to_list(9)) == ball
    with pytest.raises(ValueError):
        momentum_dict(5)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_tune_table_round_trip_and_keys(tmp_path):
    """The tuner's JSON table (no GPU): atomic save/load, and shape keys that
    separate every plan-relevant difference."""
    from types import SimpleNamespace

    from easydistillation_amd.tune import load_table, save_table, shape_key

    path = str(tmp_path / "sub" / "tune.json")
    assert load_table(path) == {}
    save_table(path, {"a": {"strip": 2}})
    assert load_table(path) == {"a": {"strip": 2}}
    (tmp_path / "bad.json").write_text("not json")
    assert load_table(str(tmp_path / "bad.json")) == {}

    def gen(**kw):
        base = dict(latt_size=[16, 16, 16, 8], Ne=40, num_nabla=2, num_momentum=2,
                    momentum_list=[(0, 0, 0), (1, 0, 0)], stocastic_coeff=None)
        base.update(kw)
        return SimpleNamespace(**base)

    keys = {
        shape_key(gen(), "mixed"),
        shape_key(gen(), "fp64"),
        shape_key(gen(Ne=41), "mixed"),
        shape_key(gen(num_nabla=1), "mixed"),
        shape_key(gen(latt_size=[16, 16, 8, 8]), "mixed"),
        shape_key(gen(momentum_list=[(0, 0, 0), (0, 1, 0)]), "mixed"),
        shape_key(gen(stocastic_coeff=np.ones((40, 40))), "mixed"),
    }
    assert len(keys) == 7
    # Lt does not change the per-timeslice contraction
    assert shape_key(gen(latt_size=[16, 16, 16, 64]), "mixed") == shape_key(gen(), "mixed")
