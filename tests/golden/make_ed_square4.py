"""Exact-diagonalisation fixture for the 4x4 periodic square lattice
(tests/golden/ed_square4.json): thermal averages of oracle/ed.py (the
restatement of the reference's diag.C) at two temperatures / anisotropies.
Each case diagonalises the 17 S^z sectors densely (the largest holds 12870
states), which takes minutes, so the values are computed once and committed.
Run from the repository root: python tests/golden/make_ed_square4.py"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import ed  # noqa: E402

CASES = [(1.0, 1.0, 1.0), (0.5, 1.0, 0.7)]   # (T, Jxy, Jz)
KEYS = ["Energy", "Energy Density", "Specific Heat", "Magnetization^2", "Susceptibility",
        "Staggered Magnetization^2", "Staggered Susceptibility", "Staggered Structure Factor"]


def main():
    out = []
    for T, jxy, jz in CASES:
        r = ed.diagonalize("square", 4, T, jxy, jz)
        out.append({"lattice": "square", "L": 4, "T": T, "Jxy": jxy, "Jz": jz,
                    "results": {k: float(r[k]) for k in KEYS}})
        print(out[-1], flush=True)
    with open(os.path.join(HERE, "ed_square4.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
