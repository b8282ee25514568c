"""A mesh of two volume element types (hex8 + prism6) through the mirror on
the GPU, against the oracle: the reference's stiffnessMatrixSolid appends
the triplets of every (material group, element type) before one sparse()
(src/linear.jl:190-330), so the oracle is the sum of its legacy IJV
assembly over the (group, type) element sets.  Checked: stiffness (two
materials spanning both types, normwise 1e-12), lumped and consistent mass,
body force on one group, traction and pressure on a face group of quads and
triangles, the solve, and the same stiffness from the mesh written to and
read back from an MSH 4.1 file."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import lowlevelfem as L
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from lowlevelfem.msh import read_msh
from mixed_mesh import hex_prism_box, write_msh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MATS = {"steel": dict(E=2.1e5, nu=0.3, rho=7.85e-9), "alu": dict(E=7.0e4, nu=0.33, rho=2.7e-9)}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(m):
    return L.Problem([L.Material(g, **MATS[g]) for g in MATS], mesh=m)


def _sub(m, blk, g):
    return M.Mesh(m.coords, blk.etype, blk.conn[blk.volumes[g]])


def _oracle_K(m):
    K = None
    for g, p in MATS.items():
        for blk in m.element_blocks():
            tab = R.rule_tables(blk.etype, 3)
            Kg = O.stiffness_legacy(_sub(m, blk, g), tab, p["E"], p["nu"], dropzeros=False, nthreads=4)
            K = Kg if K is None else K + Kg
    return K.tocsc()


def _face_tables(fs):
    order = R.element_properties(fs.etype)[2]
    return R.rule_tables(fs.etype, 2 * order + 1), R.basis(fs.etype, R.node_coords(fs.etype))[1]


def test_mixed_hex_prism_matches_oracle(gpu_ctx):
    say("[mixed] hex8 + prism6, two materials over both types")
    m = hex_prism_box()
    P = _problem(m)
    K = L.stiffnessMatrix(P)
    Ko = _oracle_K(m)
    err = sp.linalg.norm(K.A - Ko) / sp.linalg.norm(Ko)
    say(f"  stiffness rel err {err:.2e}, nnz {K.A.nnz}")
    assert err < 1e-12
    # mass
    Ml = L.massMatrix(P, lumped=True).A.diagonal()
    Mc = L.massMatrix(P, lumped=False).A
    Mlo, Mco = 0.0, None
    for g, p in MATS.items():
        for blk in m.element_blocks():
            tab = R.rule_tables(blk.etype, 3)
            Mlo = Mlo + O.mass_legacy(_sub(m, blk, g), tab, rho=p["rho"], lumped=True).diagonal()
            Mg = O.mass_legacy(_sub(m, blk, g), tab, rho=p["rho"])
            Mco = Mg if Mco is None else Mco + Mg
    assert np.abs(Ml - Mlo).max() <= 1e-12 * np.abs(Mlo).max()
    assert sp.linalg.norm(Mc - Mco) <= 1e-12 * sp.linalg.norm(Mco)
    # loads: body force on "alu", traction and pressure on the quad + triangle group "front"
    f = L.loadVector(P, [L.load("alu", fy=-2.0), L.load("front", fx=0.5, fz=1.0), L.load("right", p=3.0)])
    fo = np.zeros(3 * m.N)
    for blk in m.element_blocks():
        fo += O.body_force(_sub(m, blk, "alu"), R.rule_tables(blk.etype, 3), [0.0, -2.0, 0.0])
    for fs in m.face_sets("front"):
        fo += O.traction(m, fs, _face_tables(fs)[0], [0.5, 0.0, 1.0])
    fs = m.faces["right"]
    fo += O.pressure(m, fs, *_face_tables(fs), 3.0)[0]
    assert np.abs(f.a[:, 0] - fo).max() <= 1e-13 * np.abs(fo).max()
    # solve
    u = L.solveDisplacement(K, f, support=[L.displacementConstraint("left", ux=0, uy=0, uz=0)])
    fixed = L.constrainedDoFs(P, [L.displacementConstraint("left", ux=0, uy=0, uz=0)])
    uo, _, _ = O.pcg(Ko, fo, fixed, reltol=1e-13)
    say(f"  solve: {u.stats['iterations']} it, |u - u_oracle| / |u_oracle| "
        f"{np.linalg.norm(u.a[:, 0] - uo) / np.linalg.norm(uo):.2e}")
    assert np.linalg.norm(u.a[:, 0] - uo) <= 1e-8 * np.linalg.norm(uo)
    with pytest.raises(NotImplementedError, match="one element type"):
        L.solveDisplacement(K, f, support=[L.displacementConstraint("left", ux=0, uy=0, uz=0)],
                            preconditioner="amg")


def test_mixed_mesh_from_msh_file(gpu_ctx, tmp_path):
    m = hex_prism_box(cells=(4, 2, 3))
    path = tmp_path / "hexprism.msh"
    write_msh(m, path)
    r = read_msh(path)
    say(f"[mixed] read back: blocks {[b.etype for b in r.element_blocks()]}")
    K1 = L.stiffnessMatrix(_problem(m)).A
    K2 = L.stiffnessMatrix(_problem(r)).A
    assert sp.linalg.norm(K1 - K2) <= 1e-13 * sp.linalg.norm(K1)
    Ko = _oracle_K(r)
    assert sp.linalg.norm(K2 - Ko) <= 1e-12 * sp.linalg.norm(Ko)
