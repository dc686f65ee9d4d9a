"""Writes tests/golden/lattice_unit_cube.json: the lattice coefficients the
reference loads for a 3-D periodic unit cube (load_lattice_coefficients,
src/interaction_specific/coulomb/module_fmm_framework.f90:260-508; computed
there with Lmax_multipole = 50 and 32 iterations).

Each entry is [l, m, value] for the real coefficient M(l, m) of the Taylor
table; every entry not listed is zero.  The values are parsed from the
decimal literals of the reference source (correctly rounded to float64, as a
Fortran compiler rounds a real(8) literal).

    python tests/golden/make_lattice_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REL = "src/interaction_specific/coulomb/module_fmm_framework.f90"


def main(ref="/root/reference"):
    path = os.path.join(ref, REL)
    lines = open(path).read().splitlines()
    beg = next(i for i, s in enumerate(lines) if re.match(r"\s*subroutine load_lattice_coefficients", s))
    end = next(i for i, s in enumerate(lines) if i > beg and "end subroutine load_lattice_coefficients" in s)
    pat = re.compile(r"M\(tblinv\(\s*(\d+),\s*(\d+),\s*Lmax_taylor\)\)\s*=\s*([-+0-9.E]+)_kind_physics")
    entries = []
    for s in lines[beg:end]:
        m = pat.search(s)
        if m:
            entries.append([int(m.group(1)), int(m.group(2)), float(m.group(3))])
    out = {"source": f"{REL}:{beg + 1}-{end + 1}", "lmax_multipole_used": 50, "maxiter": 32,
           "entries": entries}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lattice_unit_cube.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print(f"{len(entries)} entries -> {dst}")


if __name__ == "__main__":
    main(*sys.argv[1:])
