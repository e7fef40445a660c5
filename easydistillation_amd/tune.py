"""Autotuned contraction shapes, kept per problem shape in a JSON file.

The library's tuner (``ed_tune``) times the contraction of a few timeslices
for every tile-order strip height, item order over the XCDs and, in the
mixed form, every mixed kernel the shape admits, and leaves the fastest set
on the context.  ``autotune`` runs it once per shape and stores the choice in
a JSON table keyed by the shape (lattice, Ne, derivatives, momenta, form);
later generators of the same shape read the choice from the table instead of
timing again.  The reference has no tuner: its ``calc(t)``
(lattice/generator/elemental.py:290-338) runs one fixed einsum path.

The table lives at ``$ED_TUNE_CACHE`` or ``~/.cache/easydistillation_amd/tune.json``.
Shapes missing there are looked up in ``tune_gfx950.json`` beside this module,
the choices this tuner measured for the benchmark shapes on MI355X; a shape in
neither is timed.
"""

from __future__ import annotations

import json
import os
import tempfile
from typing import Dict, Optional

from . import _native

TUNED_OPTIONS = {
    "strip": _native.ED_OPT_STRIP,
    "item_order": _native.ED_OPT_ITEM_ORDER,
    "mixed_ws": _native.ED_OPT_MIXED_WS,
}


SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_gfx950.json")


def default_cache_path() -> str:
    return os.environ.get("ED_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "easydistillation_amd", "tune.json")


def shape_key(gen, mode: str) -> str:
    """Key of a generator's contraction shape: everything the plan depends on."""
    Lx, Ly, Lz, _ = gen.latt_size
    disp = getattr(gen, "distance", None)
    form = f"disp{disp}" if disp is not None else f"nabla{gen.num_nabla}"
    moms = ";".join(",".join(str(c) for c in m) for m in gen.momentum_list)
    blend = "blend" if getattr(gen, "stocastic_coeff", None) is not None else "plain"
    return f"{Lx}x{Ly}x{Lz}/Ne{gen.Ne}/{form}/P{gen.num_momentum}[{moms}]/{blend}/{mode}"


def load_table(path: str) -> Dict[str, dict]:
    try:
        with open(path) as f:
            table = json.load(f)
        return table if isinstance(table, dict) else {}
    except (OSError, ValueError):
        return {}


def save_table(path: str, table: Dict[str, dict]) -> None:
    """Write atomically (a temporary file renamed over the table)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
    with os.fdopen(fd, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(gen, mode: Optional[str] = None, cache: Optional[str] = None, nt: Optional[int] = None,
             reps: int = 1, retune: bool = False) -> dict:
    """Pick the contraction shape for ``gen`` (a loaded ElementalGenerator) and apply it.

    Returns the chosen options with ``"source"``: ``"table"`` (read from the
    JSON table), ``"shipped"`` (from the package's table) or ``"timed"``
    (measured now and stored).  ``nt`` timeslices (default: up to 8) are
    timed from t = 0.
    """
    if not gen._loaded:
        raise RuntimeError("load the generator (load or load_arrays) before tuning")
    mode = mode or gen.mode
    path = cache or default_cache_path()
    key = shape_key(gen, mode)
    if not retune:
        for source, table in (("table", load_table(path)), ("shipped", load_table(SHIPPED))):
            entry = table.get(key)
            if entry is not None:
                for name, opt in TUNED_OPTIONS.items():
                    gen.set_option(opt, int(entry[name]))
                return dict(entry, source=source)
    Lt = gen.latt_size[3]
    n = min(Lt, nt or 8)
    ms = gen.context.tune(gen._mode_code(mode), 0, n, reps)
    gen._invalidate()
    entry = {name: gen.context.get_option(opt) for name, opt in TUNED_OPTIONS.items()}
    entry["ms_per_run"] = ms
    entry["timeslices_timed"] = n
    table = load_table(path)  # another process may have added shapes meanwhile
    table[key] = entry
    save_table(path, table)
    return dict(entry, source="timed")
