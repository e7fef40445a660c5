"""Extract the reference's momentum order tables as data.

lattice/insertion/mom_dict.py holds three dictionaries {index: "nx ny nz"}
(momDict_mom1, momDict_mom3, momDict_mom9).  This script reads that file as
text (it does not import or run it), collects each table's (index, momentum)
entries with a regular expression, and writes them as integer triples in
index order to mom_dict_tables.json, the fixture that pins
easydistillation_amd.insertion.momentum_dict.

Run from the repository root, where the reference is available:
    python tests/golden/make_mom_tables.py
"""

import json
import os
import re

SRC = "/root/reference/lattice/insertion/mom_dict.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mom_dict_tables.json")


def main():
    text = open(SRC).read()
    tables = {}
    for mom in (1, 3, 9):
        name = f"momDict_mom{mom}"
        start = text.index(name + " = {")
        body = text[start:text.index("}", start)]
        entries = [(int(k), [int(v) for v in s.split()]) for k, s in re.findall(r'(\d+):\s*"([^"]+)"', body)]
        assert [k for k, _ in entries] == list(range(len(entries))), name
        tables[f"mom{mom}"] = [n for _, n in entries]
    with open(OUT, "w") as f:
        json.dump({"source": "lattice/insertion/mom_dict.py", "tables": tables}, f)
    print({k: len(v) for k, v in tables.items()})


if __name__ == "__main__":
    main()
