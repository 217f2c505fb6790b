"""Write ``ref_element.h``: the reference-element mass, d/dxi and d/deta
matrices and the first edge matrix of degrees 1-3, as
``integration.reference_element`` computes them, as compile-time tables for
the fused binary32 sweep (the constants become instruction literals and the
zero entries drop out).  A context checks its own matrices against these
tables before that sweep is switched on.

Run from the repository root: ``python <package>/csrc/gen_ref_element.py``.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]


def main() -> None:
    sys.path.insert(0, str(ROOT))
    from __graft_entry__ import load_package
    load_package()
    from pbte_amd.basis import build_basis
    from pbte_amd.integration import reference_element

    def arr(name, a):
        a = np.where(np.abs(a) < 1e-14 * np.max(np.abs(a)), 0.0, a)
        vals = ", ".join(repr(float(v)) for v in a.ravel())
        return f"    static constexpr double {name}[{a.size}] = {{{vals}}};\n"

    out = ["// generated by gen_ref_element.py from integration.reference_element;\n",
           "// do not edit.  Row-major [P][P]: mass, d/dxi, d/deta, edge 1 (local face 0).\n",
           "#pragma once\n\n",
           "template <int P> struct RefElem;\n"]
    for deg in (1, 2, 3):
        basis = build_basis(deg)
        _, m, dx, dy, e1, _, _ = reference_element(basis)
        P = basis.ndof_tri
        out.append(f"template <> struct RefElem<{P}> {{\n")
        out += [arr("M", m), arr("Dx", dx), arr("Dy", dy), arr("E0", e1)]
        out.append("};\n")
    (HERE / "ref_element.h").write_text("".join(out))
    print("wrote", HERE / "ref_element.h")


if __name__ == "__main__":
    main()
