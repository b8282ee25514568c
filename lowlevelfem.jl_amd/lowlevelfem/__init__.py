"""Host-side mirror of LowLevelFEM.jl's operator-assembly API, backed by the
MI355X library libllfem.so.

Same names, argument meaning and error behaviour as the reference for the
hot path (LowLevelFEM.jl 1.13.31):

    structured_box_mesh  src/general.jl:8910-8996
    Material             src/general.jl:99-140, elastic_constants :142-203
    Problem              src/general.jl:478-632 (type=:Solid)
    SystemMatrix         src/general.jl:751-780
    VectorField          src/general.jl:1554-1563 (container only)
    stiffnessMatrix      src/linear.jl:70-87 -> stiffnessMatrixSolid :190-330
    massMatrix           src/linear.jl:1843-2004
    loadVector           src/linear.jl:3088-3434 (traction, line and point loads, pressure, body force)
    solveDisplacement    src/linear.jl:3711-3743
    displacementConstraint / load / BoundaryCondition / LoadCondition
                         src/general.jl:2480-2605, 4179-4233
    constrainedDoFs / freeDoFs / allDoFs / nodesOnPhysicalGroup
                         src/general.jl:5249-5492
    build_csc_pattern    src/multifield.jl:2083-2250
    SymGrad / Id / integral
                         the ∫(SymGrad(Pu) ⋅ D ⋅ SymGrad(Pu); ...) and ∫(Id(Pu) ⋅ c ⋅ Id(Pu); ...)
                         DSL paths, src/multifield.jl:5295-5391, assemble_operator :2630-3280;
                         ∫(Pu ⋅ g; Ω / Γ), assemble_linear src/multifield.jl:3496-3748

Differences that follow from the host language: DoF and node indices are
0-based (dof = 3*node + c); sparse matrices are scipy.sparse.csc_matrix with
int64 indices; the mesh is an explicit object (the reference keeps it in
gmsh's global state -- structured_box_mesh also sets the current mesh here).
There is no CPU fallback: every operator runs through libllfem.so.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import _lib
from . import refelem
from .mesh import Mesh, structured_box_mesh as _sbm, structural_nnz

__all__ = [
    "Material", "Problem", "SystemMatrix", "VectorField", "BoundaryCondition", "LoadCondition",
    "structured_box_mesh", "stiffnessMatrix", "massMatrix", "loadVector", "solveDisplacement",
    "displacementConstraint", "load", "constrainedDoFs", "freeDoFs", "allDoFs",
    "nodesOnPhysicalGroup", "build_csc_pattern", "SymGrad", "Id", "integral", "dropzeros",
    "elastic_constants", "Mesh", "read_msh", "solveField", "Symmetric",
]

_current_mesh: Mesh | None = None


def structured_box_mesh(**kw) -> Mesh:
    """Build a structured box mesh and make it the current model (the
    reference's gmsh.model.add("structured_box"), src/general.jl:8910)."""
    global _current_mesh
    _current_mesh = _sbm(**kw)
    return _current_mesh


def read_msh(path) -> Mesh:
    """Read a gmsh MSH 4.1 ASCII file and make it the current model (the
    reference's gmsh.open(path); node tags 1..N are the DoF base,
    src/multifield.jl:1602-1611)."""
    from .msh import read_msh as _read

    global _current_mesh
    _current_mesh = _read(path)
    return _current_mesh


def elastic_constants(E=None, nu=None, lam=None, mu=None, kappa=None):
    """Restates elastic_constants (src/general.jl:142-203)."""
    given = {k: v for k, v in dict(E=E, nu=nu, lam=lam, mu=mu, kappa=kappa).items() if v is not None}
    n = len(given)
    if n == 0 or (n == 1 and "E" in given):
        E = given.get("E", 2.0e5)
        nu = 0.3
        mu = E / (2 * (1 + nu)); lam = 2 * mu * nu / (1 - 2 * nu); kappa = E / (3 * (1 - 2 * nu))
    elif n == 1 and "kappa" in given:
        kappa = given["kappa"]; E = 2.0e5; nu = (1 - E / 3 / kappa) / 2
        mu = E / (2 * (1 + nu)); lam = 2 * mu * nu / (1 - 2 * nu)
    elif n == 2 and "E" in given and "nu" in given:
        E, nu = given["E"], given["nu"]
        mu = E / (2 * (1 + nu)); lam = 2 * mu * nu / (1 - 2 * nu); kappa = E / (3 * (1 - 2 * nu))
    elif n == 2 and "mu" in given and "nu" in given:
        mu, nu = given["mu"], given["nu"]
        E = 2 * mu * (1 + nu); lam = 2 * mu * nu / (1 - 2 * nu); kappa = E / (3 * (1 - 2 * nu))
    elif n == 2 and "lam" in given and "mu" in given:
        lam, mu = given["lam"], given["mu"]
        nu = lam / (2 * (lam + mu)); E = mu * (3 * lam + 2 * mu) / (lam + mu); kappa = lam + 2 * mu / 3
    elif n == 2 and "kappa" in given and "mu" in given:
        kappa, mu = given["kappa"], given["mu"]
        lam = kappa - 2 * mu / 3; nu = lam / (2 * (lam + mu)); E = 2 * mu * (1 + nu)
    elif n < 2:
        raise ValueError("Insufficient elastic parameters. Specify at least E+ν, μ+ν, λ+μ, κ+μ, or E alone.")
    else:
        raise ValueError(f"Elastic constants are overdetermined: {tuple(given)}")
    return dict(E=E, nu=nu, lam=lam, mu=mu, kappa=kappa)


@dataclass
class Material:
    phName: str = "body"
    type: str = "Hooke"
    E: float | None = None
    nu: float | None = None
    lam: float | None = None
    mu: float | None = None
    kappa: float | None = None
    rho: float = 7.85e-9

    def __post_init__(self):
        if self.type not in ("Hooke", "StVenantKirchhoff", "NeoHookeCompressible", "JFO"):
            raise ValueError(f"Invalid material type: {self.type}")
        ec = elastic_constants(self.E, self.nu, self.lam, self.mu, self.kappa)
        self.E, self.nu, self.lam, self.mu, self.kappa = (ec["E"], ec["nu"], ec["lam"], ec["mu"],
                                                          ec["kappa"])


class Problem:
    """A :Solid problem on a mesh (src/general.jl:478-632).  ``non`` is the
    node count; dim = pdim = 3."""

    def __init__(self, materials, type="Solid", mesh: Mesh | None = None, context=None, dim=3,
                 field="u", bandwidth=None):
        # :VectorField with dim = 3 and field :u is the DSL's spelling of the
        # same 3D vector problem (src/general.jl:478-632)
        if type == "VectorField" and dim == 3 and field == "u":
            type = "Solid"
        if type not in ("Solid", "dummy"):
            raise ValueError("only :Solid (3D linear elasticity) problems are on the GPU path")
        if bandwidth not in (None, "none"):
            raise NotImplementedError("Problem: node renumbering (bandwidth=...) is not on the GPU path")
        self.material = list(materials)
        self.type = type
        self.mesh = mesh if mesh is not None else _current_mesh
        if self.mesh is None and type != "dummy":
            raise ValueError("Problem: no mesh (call structured_box_mesh first or pass mesh=)")
        self.dim = 3
        self.pdim = 3
        self.field = "u"
        self.name = self.mesh.name if self.mesh is not None else "dummy"
        self.non = self.mesh.N if self.mesh is not None else 0
        self._ctx = context
        self._devs = None

    def ndofs(self):
        return self.pdim * self.non

    def device(self) -> _lib.DeviceMesh:
        """The device-resident mesh/pattern of the first element block."""
        return self.devices()[0][1]

    def devices(self) -> list:
        """(Block, DeviceMesh) per volume element type (built once, reused);
        every device mesh holds all node coordinates."""
        if self._devs is None:
            ctx = self._ctx or _lib.default_context()
            self._devs = [(b, _lib.DeviceMesh(ctx, self.mesh.coords, b.etype, b.conn))
                          for b in self.mesh.element_blocks()]
        return self._devs


@dataclass
class SystemMatrix:
    A: sp.csc_matrix
    model: Problem
    stats: dict = field(default_factory=dict)


@dataclass
class VectorField:
    a: np.ndarray                  # (ndof, nsteps)
    model: Problem
    type: str = "v3D"
    t: np.ndarray = field(default_factory=lambda: np.array([0.0]))
    nsteps: int = 1
    stats: dict = field(default_factory=dict)

    def copy(self):
        return VectorField(self.a.copy(), self.model, self.type, self.t.copy(), self.nsteps)


@dataclass
class BoundaryCondition:
    phName: str
    values: dict
    problem: Problem | None = None


@dataclass
class LoadCondition:
    phName: str
    values: dict
    problem: Problem | None = None


def displacementConstraint(name, **kw):
    """src/general.jl:4179-4183; keys ux, uy, uz."""
    return BoundaryCondition(name, {k: v for k, v in kw.items() if v is not None})


def load(name, **kw):
    """src/general.jl:4207-4233; traction keys fx, fy, fz."""
    return LoadCondition(name, {k: v for k, v in kw.items() if v is not None})


def nodesOnPhysicalGroup(problem: Problem, phName: str) -> np.ndarray:
    return problem.mesh.nodes_on(phName)


_COMP = {"x": 0, "y": 1, "z": 2}


def constrainedDoFs(problem: Problem, supports) -> np.ndarray:
    """unique(3*node + c) over the supports (src/general.jl:5289-5322)."""
    dofs = []
    for bc in supports:
        if bc.problem is not None and bc.problem is not problem:
            continue
        nodes = nodesOnPhysicalGroup(problem, bc.phName)
        for key in bc.values:
            if not key.startswith(problem.field):
                continue
            c = _COMP[key[len(problem.field):]]
            dofs.append(3 * nodes + c)
    if not dofs:
        return np.zeros(0, dtype=np.int64)
    d = np.concatenate(dofs)
    _, first = np.unique(d, return_index=True)
    return d[np.sort(first)].astype(np.int64)


def allDoFs(problem: Problem) -> np.ndarray:
    """All x DoFs, then y, then z (src/general.jl:5457-5472)."""
    nodes = np.arange(problem.non, dtype=np.int64)
    return np.concatenate([3 * nodes + c for c in range(3)])


def freeDoFs(problem: Problem, supports) -> np.ndarray:
    """setdiff(allDoFs, constrainedDoFs) keeping allDoFs order (src/general.jl:5487-5492)."""
    a = allDoFs(problem)
    return a[~np.isin(a, constrainedDoFs(problem, supports))]


def dropzeros(A: sp.csc_matrix) -> sp.csc_matrix:
    """SparseArrays.dropzeros!: remove stored entries that are exactly 0.0."""
    B = A.copy()
    B.eliminate_zeros()
    B.indptr = B.indptr.astype(np.int64)
    B.indices = B.indices.astype(np.int64)
    return B


def _csc(n, colptr, rowval, val):
    A = sp.csc_matrix((val, rowval, colptr), shape=(n, n))
    A.indptr = np.asarray(colptr, dtype=np.int64)
    A.indices = np.asarray(rowval, dtype=np.int64)
    return A


_PREC = {"fp64": _lib.FP64, "fp32": _lib.FP32_INT_FP64_SUM, "bf16x3": _lib.BF16X3}


def _rule(problem: Problem, gauss=None, etype=None):
    et = problem.mesh.etype if etype is None else etype
    name, dim, order, npe, family = refelem.element_properties(et)
    return refelem.rule_tables(et, refelem.gauss_degree(order, gauss))


def _sum_parts(parts):
    """Operators of the element blocks added in block order (the reference
    appends every element type's triplets before one sparse(),
    src/linear.jl:225-321)."""
    A = parts[0]
    for P in parts[1:]:
        A = A + P
    return A.tocsc()


def build_csc_pattern(Pu: Problem, Ps: Problem | None = None) -> sp.csc_matrix:
    """Structural CSC pattern with zero values (src/multifield.jl:2083-2250);
    with several element types, the union of their patterns."""
    if Ps is not None and Ps.mesh is not Pu.mesh:
        raise ValueError("Both problems must refer to the same gmsh model/mesh.")
    parts = []
    for _, dev in Pu.devices():
        colptr, rowval = dev.pattern(_lib.OP_STIFFNESS)
        parts.append(_csc(Pu.ndofs(), colptr, rowval, np.ones(rowval.size)))
    A = _sum_parts(parts)
    A.data[:] = 0.0
    return A


def _material_groups(problem: Problem, blk):
    """(material, element scale or None) per material group that has
    elements in element block `blk`.  Each material names a volume physical
    group (a generated box has the one group "body"); a single material
    whose group holds every element of the block runs the plain assembly,
    otherwise each group's elements get scale 1 and the others 0 (the
    reference loops over problem.material, then the group's element types
    and elements, src/linear.jl:190-330)."""
    vols = blk.volumes
    E = blk.conn.shape[0]
    known = set().union(*[b.volumes.keys() for b in problem.mesh.element_blocks()])
    if len(problem.material) == 1:
        name = problem.material[0].phName
        if name in vols and len(np.unique(vols[name])) == E:
            return [(problem.material[0], None)]
        if name not in known:
            if known or name != "body":     # a generated box has the one volume group "body"
                raise KeyError(f"Physical group '{name}' not found.")
            return [(problem.material[0], None)]
    out = []
    for mat in problem.material:
        if mat.phName not in known:
            raise KeyError(f"Physical group '{mat.phName}' not found.")
        if mat.phName not in vols:
            continue                        # no elements of this type in the group
        sc = np.zeros(E)
        sc[vols[mat.phName]] = 1.0
        out.append((mat, sc))
    return out


def stiffnessMatrix(problem: Problem, elements=(), forceOneThread=True, gauss=None,
                    precision="fp64", dropzeros_=True) -> SystemMatrix | None:
    """stiffnessMatrix(problem) (src/linear.jl:70-87): Hooke material of each
    material group, rule "Gauss"*(2*order+1), CSC with dropzeros! applied
    (src/linear.jl:323-324).  Several groups: one device assembly per group
    (element_scale 1 on its elements, 0 elsewhere), each added on the device
    to the groups before it (accumulate), so the device copy is the full K."""
    if problem.type == "dummy":
        return None
    parts, st = [], None
    for blk, dev in problem.devices():
        tab = _rule(problem, gauss, blk.etype)
        groups = _material_groups(problem, blk)
        last = len(groups) - 1
        for i, (m, sc) in enumerate(groups):
            val, st = dev.assemble(_lib.OP_STIFFNESS, tab, E=m.E, nu=m.nu, precision=_PREC[precision],
                                   element_scale=sc, accumulate=i > 0, copy_out=i == last)
        if groups:
            colptr, rowval = dev.pattern(_lib.OP_STIFFNESS)
            parts.append(_csc(problem.ndofs(), colptr, rowval, val))
    if not parts:
        raise ValueError("stiffnessMatrix: no material group has elements")
    A = _sum_parts(parts)
    if dropzeros_:
        A = dropzeros(A)
    return SystemMatrix(A, problem, _device_k(problem, st, resident=len(problem.devices()) == 1))


def _device_k(problem: Problem, st, resident=True) -> dict:
    """Stats of a stiffness assembly, tagged with the device copy it left on
    the problem's mesh (solveDisplacement's preconditioner="amg" solves that
    copy; a mesh of several element types has no single device copy)."""
    problem._k_version = getattr(problem, "_k_version", 0) + 1
    d = st.as_dict()
    d["device_k_version"] = problem._k_version if resident else None
    return d


class SymGrad:
    """SymGrad(P): symmetric gradient operator applied to a vector field
    (src/multifield.jl:4460)."""

    def __init__(self, P: Problem):
        self.P = P


class Id:
    """Id(P): identity operator applied to a vector field (src/multifield.jl:4463-4475,
    IdOp :144; build_B! :500)."""

    def __init__(self, P: Problem):
        self.P = P


def _integral_id(a, c, b, gauss, csc_matrix):
    """∫(Id(Pu) ⋅ c ⋅ Id(Pu)) with a scalar c: the consistent-mass shape
    ρ Σ_q N_a N_b |J| w δ_ij (src/multifield.jl:5295-5391 with IdOp).  Without
    csc_matrix: dropzeros! of the block-diagonal entries; with it the
    structural pattern is kept and the off-diagonal entries of every 3x3
    block are stored zeros, as the reference's CSC scatter leaves them."""
    if not (isinstance(a, Id) and isinstance(b, Id) and a.P is b.P):
        raise NotImplementedError("only ∫(Id(Pu) ⋅ c ⋅ Id(Pu)) with one field is on the GPU path")
    if np.size(c) != 1:
        raise NotImplementedError("∫(Id ⋅ C ⋅ Id): only a scalar coefficient is on the GPU path")
    rho = float(np.asarray(c, dtype=np.float64).reshape(-1)[0])
    P = a.P
    if len(P.devices()) > 1:
        raise NotImplementedError("∫(Id ⋅ c ⋅ Id) on a mesh of several element types is not on the GPU path")
    n = P.ndofs()
    dev = P.device()
    val, st = dev.assemble(_lib.OP_MASS, _rule(P, gauss), rho=rho, lumped=False)
    if csc_matrix is None:
        mcp, mrv = dev.pattern(_lib.OP_MASS)
        return SystemMatrix(dropzeros(_csc(n, mcp, mrv, val)), P, st.as_dict())
    K = csc_matrix.A if isinstance(csc_matrix, SystemMatrix) else csc_matrix
    cp, rv = dev.pattern(_lib.OP_STIFFNESS)
    if K.shape != (n, n) or not (np.array_equal(K.indptr, cp) and np.array_equal(K.indices, rv)):
        raise ValueError("assemble_operator: csc_matrix has an incompatible size/pattern")
    # the mass values sit where the row component equals the column component;
    # per column those entries are in the same order as the mass pattern's rows
    col = np.repeat(np.arange(n, dtype=np.int64), np.diff(cp))
    diag = (rv % 3) == (col % 3)
    K.data[:] = 0.0
    K.data[diag] = val
    return SystemMatrix(K, P, st.as_dict())


def integral(a, D, b=None, gauss=None, csc_matrix=None, precision="fp64", threads=None, Γ=None, Ω=None):
    """∫(SymGrad(Pu) ⋅ D ⋅ SymGrad(Pu); gauss, csc_matrix) (src/multifield.jl:5295-5339).
    D is the 6x6 coefficient in SymGrad row order [xx,yy,zz,xy,xz,yz].
    Without csc_matrix the result gets dropzeros! (:5336); with it the
    structural pattern is kept and the caller's matrix is filled in place.
    integral(Id(Pu), c, Id(Pu), ...) is the mass-shaped form (scalar c).
    integral(Pu, g; Γ=name or Ω=name) -- no third operand -- is the linear
    form ∫(Pu ⋅ g) with g a 3-vector or a function of (x, y, z).
    ``threads`` (the reference's worker count) has no meaning on the GPU path."""
    if b is None:
        P = a.P if isinstance(a, Id) else a
        if not isinstance(P, Problem):
            raise NotImplementedError("only ∫(Pu ⋅ g) is on the GPU path as a linear form")
        return _linear_form(P, D, gauss=gauss, Γ=Γ, Ω=Ω)
    if isinstance(a, Id) or isinstance(b, Id):
        return _integral_id(a, D, b, gauss, csc_matrix)
    if not isinstance(a, SymGrad) or not isinstance(b, SymGrad) or a.P is not b.P:
        raise NotImplementedError("only ∫(SymGrad(Pu) ⋅ D ⋅ SymGrad(Pu)) is on the GPU path")
    P = a.P
    D = np.asarray(D, dtype=np.float64).reshape(6, 6)
    if len(P.devices()) > 1:
        if csc_matrix is not None:
            raise NotImplementedError("∫(SymGrad ⋅ D ⋅ SymGrad; csc_matrix) on a mesh of several element "
                                      "types is not on the GPU path")
        parts = []
        for blk, dv in P.devices():
            v, st = dv.assemble(_lib.OP_STIFFNESS, _rule(P, gauss, blk.etype), D=D, precision=_PREC[precision])
            cp, rv = dv.pattern(_lib.OP_STIFFNESS)
            parts.append(_csc(P.ndofs(), cp, rv, v))
        return SystemMatrix(dropzeros(_sum_parts(parts)), P, _device_k(P, st, resident=False))
    dev = P.device()
    tab = _rule(P, gauss)
    val, st = dev.assemble(_lib.OP_STIFFNESS, tab, D=D, precision=_PREC[precision])
    if csc_matrix is not None:
        K = csc_matrix.A if isinstance(csc_matrix, SystemMatrix) else csc_matrix
        if K.shape != (P.ndofs(), P.ndofs()) or K.nnz != val.size:
            raise ValueError("assemble_operator: csc_matrix has an incompatible size/pattern")
        K.data[:] = val
        return SystemMatrix(K, P, _device_k(P, st))
    colptr, rowval = dev.pattern(_lib.OP_STIFFNESS)
    return SystemMatrix(dropzeros(_csc(P.ndofs(), colptr, rowval, val)), P, _device_k(P, st))


def massMatrix(problem: Problem, lumped=True, gauss=None) -> SystemMatrix | None:
    """massMatrix(problem; lumped=true) (src/linear.jl:1843-2004)."""
    if problem.type == "dummy":
        return None
    n = problem.ndofs()
    parts, st = [], None
    for blk, dev in problem.devices():
        tab = _rule(problem, gauss, blk.etype)
        val = None
        for m, sc in _material_groups(problem, blk):
            v, st = dev.assemble(_lib.OP_MASS, tab, rho=m.rho, lumped=lumped, element_scale=sc)
            val = v if val is None else val + v
        if val is None:
            continue
        if lumped:
            parts.append(sp.diags(val, 0, shape=(n, n), format="csc"))
        else:
            colptr, rowval = dev.pattern(_lib.OP_MASS)
            parts.append(_csc(n, colptr, rowval, val))
    if not parts:
        raise ValueError("massMatrix: no material group has elements")
    return SystemMatrix(dropzeros(_sum_parts(parts)), problem, st.as_dict())


def _eval_at(fn, P):
    """A load component at the Gauss points P (F, q, 3): a number, or a
    function of (x, y, z) as _loadvec_helper calls it (src/linear.jl:3288-3299);
    called once on whole arrays when it accepts them, else point by point."""
    shape = P.shape[:2]
    if not callable(fn):
        return np.full(shape, float(fn))
    try:
        v = np.asarray(fn(P[..., 0], P[..., 1], P[..., 2]), dtype=np.float64)
        return np.broadcast_to(v, shape).copy()
    except Exception:
        flat = P.reshape(-1, 3)
        return np.array([float(fn(*x)) for x in flat]).reshape(shape)


def _volume_elems(problem: Problem, blk, name: str):
    """Element indices of volume group `name` in element block `blk` (None:
    every element, the generated box's one group "body"; an empty array: the
    group has no elements of this type)."""
    blocks = problem.mesh.element_blocks()
    if blk.volumes.get(name) is not None:
        return np.asarray(blk.volumes[name], dtype=np.int64)
    if any(name in b.volumes for b in blocks):
        return np.zeros(0, dtype=np.int64)
    if name == "body" and not any(b.volumes for b in blocks):
        return None
    raise KeyError(f"Physical group '{name}' not found.")


def _eval_vec_at(g, P):
    """A vector coefficient at the Gauss points P (n, q, 3): a 3-vector of
    numbers, or a function of (x, y, z) returning one (called on whole
    arrays when it accepts them, else point by point)."""
    shape = P.shape[:2]
    if not callable(g):
        v = np.asarray(g, dtype=np.float64).reshape(-1)
        if v.size != 3:
            raise ValueError(f"assemble_linear: vector coefficient length ({v.size}) does not match "
                             "operator output dimension (3).")
        return np.broadcast_to(v, shape + (3,)).copy()
    try:
        out = np.stack([np.broadcast_to(np.asarray(c, dtype=np.float64), shape)
                        for c in g(P[..., 0], P[..., 1], P[..., 2])], axis=-1)
        if out.shape == shape + (3,):
            return out
    except Exception:
        pass
    flat = P.reshape(-1, 3)
    return np.array([np.asarray(g(*x), dtype=np.float64).reshape(3) for x in flat]).reshape(shape + (3,))


def _body_load(problem: Problem, name: str, comps, gauss=None):
    """Body force on volume group `name` (components numbers or functions),
    summed over the element blocks."""
    f = np.zeros(problem.ndofs())
    for blk, dev in problem.devices():
        elems = _volume_elems(problem, blk, name)
        if elems is not None and elems.size == 0:
            continue
        tab = _rule(problem, gauss, blk.etype)
        if any(callable(c) for c in comps):
            P = dev.volume_points(tab, elems)
            G = np.stack([_eval_at(c, P) for c in comps], axis=-1)
            f += dev.body(tab, g_gp=G, elems=elems)
        else:
            f += dev.body(tab, g=[float(c) for c in comps], elems=elems)
    return f


def _face_rule(fs):
    name, dim, order, npe, family = refelem.element_properties(fs.etype)
    return refelem.rule_tables(fs.etype, 2 * order + 1)


def loadVector(problem: Problem, loads) -> VectorField | None:
    """loadVector(problem, loads) (src/linear.jl:3088-3434) for a 3D solid:
    on a surface group, traction load(name; fx, fy, fz) or pressure
    load(name; p) (src/linear.jl:3196-3224, p replaces fx, fy, fz); on a
    curve group, a line load (force per length, Ja = |dx/dxi|), and on a
    point group, a concentrated force (Ja = 1) (src/linear.jl:3385-3388);
    on a volume group, a body force (src/linear.jl:3054-3068).  Components
    and p are numbers or functions of (x, y, z): traction and body-force
    functions are evaluated at the Gauss points the library returns, p at
    the nodes (scalarField, src/general.jl:4477-4515)."""
    if problem.type == "dummy":
        return None
    if not isinstance(loads, (list, tuple)):
        raise TypeError("loadVector: loads must be in a vector.")
    dev = problem.device()
    f = np.zeros(problem.ndofs())
    for ld in loads:
        if ld.problem is not None and ld.problem is not problem:
            continue
        comps = [0.0, 0.0, 0.0]
        p = None
        for k, v in ld.values.items():
            if k == "p":
                p = v
            elif k in ("fx", "fy", "fz"):
                comps["xyz".index(k[1])] = v
            else:
                raise NotImplementedError(f"loadVector: load key {k} is not on the GPU path")
        name = ld.phName
        for fs in (problem.mesh.face_sets(name) if name in problem.mesh.faces else []):
            tab = _face_rule(fs)
            if p is not None and tab.dim != 2:
                raise ValueError("loadVector: pressure can be given on a surface of a 3D solid.")
            if p is not None:
                dNn = refelem.basis(fs.etype, refelem.node_coords(fs.etype))[1]
                if callable(p):
                    x = problem.mesh.coords
                    pn = np.zeros(problem.non)
                    nodes = np.unique(fs.conn)
                    pn[nodes] = _eval_at(p, x[nodes][None, :, :])[0]
                    f += dev.pressure(fs.etype, fs.conn, tab, dNn, p_nodal=pn)
                else:
                    f += dev.pressure(fs.etype, fs.conn, tab, dNn, p=float(p))
            elif any(callable(c) for c in comps):
                P = dev.face_points(fs.etype, fs.conn, tab)
                T = np.stack([_eval_at(c, P) for c in comps], axis=-1)
                f += dev.traction_gp(fs.etype, fs.conn, tab, T)
            else:
                f += dev.traction(fs.etype, fs.conn, tab, np.array([float(c) for c in comps]))
        if name in problem.mesh.faces:
            continue
        if p is not None:
            _volume_elems(problem, problem.mesh.element_blocks()[0], name)
            raise ValueError("loadVector: pressure can be given on a surface of a 3D solid.")
        f += _body_load(problem, name, comps)
    return VectorField(f.reshape(-1, 1), problem)


def _linear_form(P: Problem, g, gauss=None, Γ=None, Ω=None) -> VectorField:
    """∫(Pu ⋅ g; Ω / Γ) (assemble_linear, src/multifield.jl:3496-3748):
    N_a g det J w over the volume groups of the problem's materials (or Ω),
    or N_a g |t1 x t2| w over the surface group Γ; rule Gauss(2 order + 1)."""
    dev = P.device()
    f = np.zeros(P.ndofs())
    if Γ is not None:
        for fs in P.mesh.face_sets(Γ):
            tab = _face_rule(fs)
            Pq = dev.face_points(fs.etype, fs.conn, tab)
            f += dev.traction_gp(fs.etype, fs.conn, tab, _eval_vec_at(g, Pq))
        return VectorField(f.reshape(-1, 1), P)
    names = [Ω] if Ω is not None else [m.phName for m in P.material]
    for name in names:
        for blk, dv in P.devices():
            elems = _volume_elems(P, blk, name)
            if elems is not None and elems.size == 0:
                continue
            tab = _rule(P, gauss, blk.etype)
            if callable(g):
                Pq = dv.volume_points(tab, elems)
                f += dv.body(tab, g_gp=_eval_vec_at(g, Pq), elems=elems)
            else:
                f += dv.body(tab, g=_eval_vec_at(g, np.zeros((1, 1, 3)))[0, 0], elems=elems)
    return VectorField(f.reshape(-1, 1), P)


def solveDisplacement(K: SystemMatrix, f: VectorField, support=(), iterative=False,
                      reltol=math.sqrt(np.finfo(float).eps), maxiter=None, preconditioner="jacobi",
                      ordering=True, method="fp64") -> VectorField:
    """solveDisplacement(K, f; support, ...) (src/linear.jl:3711-3743).

    The GPU path is Jacobi-PCG (method "fp64", or "mixed" = FP32 inner CG
    with FP64 refinement).  iterative=false (the reference's direct solve)
    runs the same PCG to reltol = 1e-12; iterative=true honours reltol.
    preconditioner (the reference passes it to cg as Pl, :3736): "jacobi", or
    "amg" — mixed-precision CG with the aggregation-AMG V-cycle on the
    stiffness kept on the problem's device mesh; K must be the result of the
    last stiffnessMatrix / integral(SymGrad...) call on that problem."""
    problem = K.model
    if f.type != "v3D":
        raise ValueError(f"solveDisplacement: wrong type of 'f': ({f.type})")
    support = list(support)
    fixed = constrainedDoFs(problem, support)
    uval = np.zeros(problem.ndofs())
    for bc in support:     # applyBoundaryConditions! (src/linear.jl:3505-3563)
        if bc.problem is not None and bc.problem is not problem:
            continue
        nodes = nodesOnPhysicalGroup(problem, bc.phName)
        for key, v in bc.values.items():
            if callable(v):
                raise NotImplementedError("function-valued supports are not on the GPU path")
            uval[3 * nodes + _COMP[key[1:]]] = float(v)
    ufix = uval[fixed]
    tol = reltol if iterative else 1e-12
    maxiter = maxiter or max(problem.non * problem.dim, 1000)
    meth = _lib.PCG_MIXED_REFINE if method == "mixed" else _lib.PCG_JACOBI_FP64
    dev = problem.device()
    if preconditioner == "amg":
        if len(problem.devices()) > 1:
            raise NotImplementedError("solveDisplacement: preconditioner=\"amg\" needs a mesh of one element type")
        if K.stats.get("device_k_version") is None or \
                K.stats["device_k_version"] != getattr(problem, "_k_version", None):
            raise ValueError("solveDisplacement: preconditioner=\"amg\" solves the stiffness kept on the "
                             "device; K is not the last one assembled for this problem")
        fa = np.ascontiguousarray(f.a[:, 0], dtype=np.float64)
        u, st = dev.solve(fa, fixed, ufix, method=_lib.PCG_AMG_RELIABLE, reltol=tol, maxiter=maxiter)
        out = f.copy()
        out.a[:, 0] = u
        out.stats = st.as_dict()
        return out
    if preconditioner not in ("jacobi", None):
        raise NotImplementedError(f"solveDisplacement: preconditioner {preconditioner!r} is not on the GPU path")
    A = K.A
    fa = np.ascontiguousarray(f.a[:, 0], dtype=np.float64)
    u, st = _lib.solve_csc(dev.ctx, A.indptr, A.indices, A.data, fa, fixed, ufix, method=meth,
                           reltol=tol, maxiter=maxiter)
    out = f.copy()
    out.a[:, 0] = u
    out.stats = st.as_dict()
    return out


class SymmetricSystemMatrix:
    """Symmetric(K): K flagged symmetric (the reference then solves with
    CHOLMOD, src/multifield.jl:4166-4268); the stored K is symmetric here,
    so the wrapper only carries it."""

    def __init__(self, K: SystemMatrix, uplo="U"):
        self.parent, self.uplo = K, uplo


def Symmetric(K: SystemMatrix, uplo="U") -> SymmetricSystemMatrix:
    return SymmetricSystemMatrix(K, uplo)


def solveField(K, f: VectorField, support=(), iterative=False, reltol=math.sqrt(np.finfo(float).eps),
               maxiter=None, preconditioner="jacobi", ordering=True) -> VectorField:
    """solveField(K, f; support, ...) and solveField(Symmetric(K), f; ...)
    (src/poisson.jl:1432-1546, src/multifield.jl:4166-4268): the same
    Dirichlet elimination as solveDisplacement, on the same device solve."""
    if isinstance(K, SymmetricSystemMatrix):
        K = K.parent
    return solveDisplacement(K, f, support=support, iterative=iterative, reltol=reltol, maxiter=maxiter,
                             preconditioner=preconditioner, ordering=ordering)
