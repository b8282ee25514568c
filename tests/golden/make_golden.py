"""Writes the small golden fixtures under tests/golden/ (run from the
repository root: python tests/golden/make_golden.py).

Each fixture holds one problem's inputs-by-recipe (mesh kind, cells, rule
degree, material) and the oracle's outputs: the CSC pattern (colptr, rowval)
of build_csc_pattern, the FP64 stiffness values, the consistent-mass values
on the mass pattern, the traction load fy = -1 on the right face, and the
displacement of the left-clamped problem from a sparse direct solve
(scipy SuperLU) of the Dirichlet-eliminated system (src/linear.jl:3711-3743
with iterative=false).  The oracle values themselves are pinned by
tests/test_oracle.py against the reference's recorded numbers and closed
forms; the fixtures freeze them so later changes to the oracle, the mesh
generator or the reference-element tables cannot drift silently.
"""
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "lowlevelfem.jl_amd"))
sys.path.insert(0, ROOT)

from lowlevelfem import mesh as M  # noqa: E402
from lowlevelfem import refelem as R  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {
    "hex8_3x2x2": dict(mesh=dict(lx=1.5, ly=1.0, lz=1.0, cells=(3, 2, 2)), deg=3, fdeg=3),
    "tet10_2x2x2": dict(mesh=dict(cells=(2, 2, 2), elem="tet", order=2), deg=2, fdeg=5, mdeg=4),
    "hex27_2x1x1_perturbed": dict(mesh=dict(lx=2.0, cells=(2, 1, 1), order=2, perturb=0.05, seed=3),
                                  deg=5, fdeg=5),
}
E, NU, RHO = 2e5, 0.3, 7.85e-9


def hooke():
    mu = E / (2 * (1 + NU))
    lam = 2 * mu * NU / (1 - 2 * NU)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return D


def compute(spec):
    m = M.structured_box_mesh(**spec["mesh"])
    tab = R.rule_tables(m.etype, spec["deg"])
    cp, rv = O.csc_pattern(m)
    K = O.stiffness_csc(m, tab, hooke(), cp, rv)
    mtab = R.rule_tables(m.etype, spec.get("mdeg", spec["deg"]))
    Ms = O.mass_legacy(m, mtab, rho=RHO).tocsc()
    Ms.sort_indices()
    face = m.faces["right"]
    f = O.traction(m, face, R.rule_tables(face.etype, spec["fdeg"]), [0.0, -1.0, 0.0])
    left = m.nodes_on("left")
    fixed = np.sort((3 * left[:, None] + np.arange(3)[None, :]).ravel())
    free = np.setdiff1d(np.arange(3 * m.N), fixed)
    A = sp.csc_matrix((K, rv, cp), shape=(3 * m.N, 3 * m.N))
    u = np.zeros(3 * m.N)
    u[free] = spla.spsolve(A[free][:, free].tocsc(), f[free])
    return dict(colptr=cp, rowval=rv, K=K, M_colptr=Ms.indptr.astype(np.int64),
                M_rowval=Ms.indices.astype(np.int64), M=Ms.data, f=f, fixed=fixed, u=u)


def main():
    for name, spec in CASES.items():
        out = compute(spec)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{name}: N={len(out['u']) // 3} nnz(K)={len(out['K'])} -> {os.path.relpath(path, ROOT)}")


if __name__ == "__main__":
    main()
