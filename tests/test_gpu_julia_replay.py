"""The Julia binding's call sequence, replayed through ctypes on the GPU.

The functions below follow lowlevelfem.jl_amd/julia/LowLevelFEMHIPExt.jl
statement for statement (type_mesh, pattern, group_scale, assemble_type,
assemble, gpu_solveDisplacement), with a stand-in for the gmsh queries: a
hexahedral mesh whose elements carry gmsh-style tags (unique, not
contiguous, not in element order) and sit in two volume entities, one per
material group, plus a third group with no elements.  The entry points the
replay calls, in order, are those the binding's ccalls name, and the
results match the oracle: the two-group stiffness (normwise 1e-12) and
lumped mass, and the direct-solve substitute (iterative=false: PCG to 1e-12)."""
import ctypes as C
import re
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from lowlevelfem import _lib
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from oracle import oracle as O
from test_julia_binding import _src, julia_ccalls

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Recorder:
    """The library with every entry point call recorded by name."""

    def __init__(self, lib):
        self.lib, self.names = lib, []

    def __getattr__(self, name):
        fn = getattr(self.lib, name)

        def call(*a):
            self.names.append(name)
            return fn(*a)
        return call


class FakeGmsh:
    """What the binding asks gmsh for, on a generated mesh: element tags in
    getElementsByType order, entity membership and physical groups."""

    def __init__(self, mesh, seed=7):
        rng = np.random.default_rng(seed)
        self.mesh = mesh
        self.elem_tags = (rng.permutation(mesh.E) * 3 + 101).astype(np.int64)   # gaps, shuffled
        self.entity = np.where(rng.random(mesh.E) < 0.45, 1, 2)                  # volume entity per element
        self.groups = {"steel": [(3, 1)], "alu": [(3, 2)], "void": [(3, 9)]}

    def getElementsByType(self, et, tag=-1):
        sel = np.ones(self.mesh.E, bool) if tag == -1 else self.entity == tag
        return self.elem_tags[sel], (self.mesh.conn[sel] + 1).ravel()           # element tags, node tags

    def getEntitiesForPhysicalName(self, name):
        return self.groups[name]


# ---- the binding, statement for statement --------------------------------

def type_mesh(lib, ctx, g, et, nn):
    elemTags, nodeTags = g.getElementsByType(et)
    xyz = np.ascontiguousarray(g.mesh.coords, dtype=np.float64)          # [N][3] = Julia's 3 x N
    conn = np.ascontiguousarray(nodeTags, dtype=np.int64)                 # 1-based, element-major
    h = C.c_void_p()
    _lib.check(lib.llfem_mesh_create(ctx.h, _lib.ptr(xyz), g.mesh.N, et, nn, _lib.ptr(conn), elemTags.size,
                                     3, 1, C.byref(h)))
    etags = np.asarray(elemTags, dtype=np.int64)
    return dict(et=et, nn=nn, handle=h, elemTags=etags, pos={int(t): i for i, t in enumerate(etags)})


def pattern(lib, m, op, ndof):
    nnz = C.c_int64(0)
    _lib.check(lib.llfem_mesh_pattern(m["handle"], op, 1, C.byref(nnz), None, None))
    colptr, rowval = np.empty(ndof + 1, np.int64), np.empty(nnz.value, np.int64)
    _lib.check(lib.llfem_mesh_pattern(m["handle"], op, 1, C.byref(nnz), _lib.ptr(colptr), _lib.ptr(rowval)))
    return colptr, rowval


def group_scale(g, m, phName):
    scale = np.zeros(m["elemTags"].size)
    for edim, etag in g.getEntitiesForPhysicalName(phName):
        elemTags, _ = g.getElementsByType(m["et"], etag)
        for t in elemTags:
            scale[m["pos"][int(t)]] = 1.0
    return scale


def assemble_type(lib, g, materials, m, op, tab, ndof, lumped=False):
    lump = op == _lib.OP_MASS and lumped
    colptr, rowval = (None, None) if lump else pattern(lib, m, op, ndof)
    nzval = np.zeros(ndof if lump else rowval.size)
    groups = [(mat, group_scale(g, m, mat["phName"])) for mat in materials]
    groups = [gr for gr in groups if np.any(gr[1] != 0.0)]
    for i, (mat, scale) in enumerate(groups, start=1):
        whole = len(groups) == 1 and np.all(scale == 1.0)
        acc = (not lump) and i > 1
        out = np.zeros(ndof) if lump else nzval
        copy_out = lump or i == len(groups)
        desc = _lib.OpDesc(op, tab.q, tab.weights.ctypes.data, tab.N.ctypes.data, tab.dN.ctypes.data, None,
                           mat["E"], mat["nu"], mat["rho"], int(lumped), 0,
                           None if whole else scale.ctypes.data, int(acc))
        _lib.check(lib.llfem_assemble(m["handle"], C.byref(desc), _lib.ptr(out) if copy_out else None,
                                      C.byref(_lib.Stats())))
        if lump:
            nzval += out
    if lump:
        return nzval
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(ndof, ndof))


def gpu_solve(lib, ctx, A, f, fixed1, ufix, iterative=False, reltol=np.sqrt(np.finfo(float).eps),
              maxiter=None):
    tol = reltol if iterative else 1e-12
    its = maxiter if iterative else max(maxiter, 10 * A.shape[0])
    colptr = np.ascontiguousarray(A.indptr + 1, dtype=np.int64)
    rowval = np.ascontiguousarray(A.indices + 1, dtype=np.int64)
    sol = np.zeros(A.shape[0])
    _lib.check(lib.llfem_solve_csc(ctx.h, A.shape[0], _lib.ptr(colptr), _lib.ptr(rowval),
                                   _lib.ptr(np.ascontiguousarray(A.data)), 1, _lib.ptr(f), _lib.ptr(fixed1),
                                   fixed1.size, _lib.ptr(ufix), _lib.PCG_JACOBI_FP64, tol, its, _lib.ptr(sol),
                                   C.byref(_lib.Stats())))
    return sol


def _julia_order(*funcs):
    src = _src()
    names = []
    for fn in funcs:
        body = re.search(r"^function " + fn + r"\(.*?^end", src, flags=re.S | re.M).group(0)
        for name, _, _, _ in julia_ccalls(body):
            if name != "llfem_last_error" and (not names or names[-1] != name):
                names.append(name)
    return names


def _dedup(seq):
    out = []
    for n in seq:
        if not out or out[-1] != n:
            out.append(n)
    return out


def test_replay_two_material_groups_by_element_tag(gpu_ctx):
    say("[julia replay] hex8 5x4x6, two groups by element tag")
    mesh = M.structured_box_mesh(lx=1.0, ly=0.8, lz=1.2, cells=(5, 4, 6))
    g = FakeGmsh(mesh)
    tab = R.rule_tables(mesh.etype, 3)
    assert all(a.dtype == np.float64 and a.flags.c_contiguous for a in (tab.weights, tab.N, tab.dN))
    mats = [dict(phName="steel", E=2.1e5, nu=0.3, rho=7.85e-9), dict(phName="alu", E=7e4, nu=0.33, rho=2.7e-9),
            dict(phName="void", E=1.0, nu=0.1, rho=1.0)]
    rec = Recorder(_lib.load_library())
    ndof = 3 * mesh.N
    m = type_mesh(rec, gpu_ctx, g, mesh.etype, mesh.npe)
    try:
        K = assemble_type(rec, g, mats, m, _lib.OP_STIFFNESS, tab, ndof)
        # the entry points, in the binding's order (type_mesh, pattern, assemble_type)
        assert _dedup(rec.names) == _julia_order("type_mesh", "pattern", "assemble_type")
        assert rec.names.count("llfem_assemble") == 2 and rec.names.count("llfem_mesh_pattern") == 2
        # oracle: the reference's triplets of each group's elements, summed
        Ko = None
        for mat, ent in (("steel", 1), ("alu", 2)):
            sub = M.Mesh(mesh.coords, mesh.etype, mesh.conn[g.entity == ent])
            spec = next(x for x in mats if x["phName"] == mat)
            Kg = O.stiffness_legacy(sub, tab, spec["E"], spec["nu"], dropzeros=False, nthreads=4)
            Ko = Kg if Ko is None else Ko + Kg
        err = sp.linalg.norm(K - Ko) / sp.linalg.norm(Ko)
        say(f"  stiffness rel err {err:.2e}; calls {_dedup(rec.names)}")
        assert err < 1e-12
        # lumped mass: the groups' diagonals added on the host
        Md = assemble_type(rec, g, mats, m, _lib.OP_MASS, tab, ndof, lumped=True)
        Mo = sum(O.mass_legacy(M.Mesh(mesh.coords, mesh.etype, mesh.conn[g.entity == ent]), tab,
                               rho=next(x for x in mats if x["phName"] == nm)["rho"], lumped=True).diagonal()
                 for nm, ent in (("steel", 1), ("alu", 2)))
        assert np.abs(Md - Mo).max() <= 1e-12 * np.abs(Mo).max()
        # gpu_solveDisplacement with iterative=false: PCG to 1e-12
        left = mesh.nodes_on("left")
        fixed = np.sort((3 * left[:, None] + np.arange(3)[None, :]).ravel())
        f = O.traction(mesh, mesh.faces["right"], R.rule_tables(mesh.faces["right"].etype, 3), [0.0, 0.0, 1.0])
        u = gpu_solve(rec, gpu_ctx, K, f, fixed + 1, np.zeros(fixed.size), iterative=False, maxiter=3 * mesh.N)
        uo, _, _ = O.pcg(Ko.tocsc(), f, fixed, reltol=1e-13)
        free = np.setdiff1d(np.arange(ndof), fixed)
        r = (Ko @ u - f)[free]
        say(f"  solve: relres {np.linalg.norm(r) / np.linalg.norm(f[free]):.2e}")
        assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(f[free])   # not sqrt(eps): the direct-solve substitute
        assert np.linalg.norm(u - uo) <= 1e-8 * np.linalg.norm(uo)
    finally:
        _lib.load_library().llfem_mesh_destroy(m["handle"])
