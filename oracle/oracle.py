"""Python front of the CPU oracle (oracle.c).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker; never by the product path.

Each wrapper names the reference function the C code restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        P, I64, I, D = C.c_void_p, C.c_int64, C.c_int, C.c_double
        L.oracle_csc_pattern.argtypes = [I64, I, I64, P, I, P, P, P, P]
        L.oracle_csc_pattern.restype = I64
        L.oracle_hooke_D.argtypes = [D, D, P]
        L.oracle_stiffness_ijv.argtypes = [I64, P, I, I64, P, I, P, P, D, D, P, P, P, I]
        L.oracle_stiffness_ijv.restype = I
        L.oracle_sparse.argtypes = [I64, I64, P, P, P, I, P, P, P]
        L.oracle_sparse.restype = I64
        L.oracle_stiffness_csc.argtypes = [I64, P, I, I64, P, I, P, P, P, P, P, P, I, I64]
        L.oracle_stiffness_csc.restype = I
        L.oracle_mass_ijv.argtypes = [I64, P, I, I64, P, I, P, P, P, D, P, P, P]
        L.oracle_mass_ijv.restype = I
        L.oracle_traction.argtypes = [I64, P, I, I, I64, P, I, P, P, P, P, P]
        L.oracle_traction.restype = I
        L.oracle_pcg.argtypes = [I64, P, P, P, P, P, P, D, I64, P, P]
        L.oracle_pcg.restype = I64
        L.oracle_num_threads.restype = I
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def hooke_D(E, nu):
    """constitutive_matrix(3, :Solid, E, ν) (src/linear.jl:154-165)."""
    D = np.zeros(36)
    lib().oracle_hooke_D(E, nu, _p(D))
    return D.reshape(6, 6)


def csc_pattern(mesh, with_adj=False):
    """build_csc_pattern (src/multifield.jl:2083-2250) -> (colptr, rowval)."""
    conn = _c(mesh.conn, np.int64)
    N, E, n = mesh.N, conn.shape[0], conn.shape[1]
    nnz = lib().oracle_csc_pattern(N, n, E, _p(conn), 3, None, None, None, None)
    colptr = np.empty(3 * N + 1, np.int64)
    rowval = np.empty(nnz, np.int64)
    adj_ptr = np.empty(N + 1, np.int64)
    lib().oracle_csc_pattern(N, n, E, _p(conn), 3, _p(colptr), _p(rowval), _p(adj_ptr), None)
    if with_adj:
        adj = np.empty(adj_ptr[-1], np.int64)
        lib().oracle_csc_pattern(N, n, E, _p(conn), 3, None, None, _p(adj_ptr), _p(adj))
        return colptr, rowval, adj_ptr, adj
    return colptr, rowval


def sparse(n, I, J, V, dropzeros=True):
    """SparseArrays.sparse(I,J,V,n,n) (+ dropzeros!) as the reference calls
    it (src/linear.jl:323-324)."""
    I, J, V = _c(I, np.int64), _c(J, np.int64), _c(V, np.float64)
    colptr = np.empty(n + 1, np.int64)
    rowval = np.empty(I.size, np.int64)
    nzval = np.empty(I.size, np.float64)
    nnz = lib().oracle_sparse(n, I.size, _p(I), _p(J), _p(V), int(dropzeros), _p(colptr), _p(rowval),
                              _p(nzval))
    A = sp.csc_matrix((nzval[:nnz], rowval[:nnz], colptr), shape=(n, n))
    return A


def stiffness_legacy(mesh, tab, E=2e5, nu=0.3, dropzeros=True, nthreads=0):
    """stiffnessMatrix(problem) = stiffnessMatrixSolid (src/linear.jl:190-330)."""
    xyz, conn = _c(mesh.coords, np.float64), _c(mesh.conn, np.int64)
    n3 = 3 * conn.shape[1]
    nt = conn.shape[0] * n3 * n3
    I, J, V = np.empty(nt, np.int64), np.empty(nt, np.int64), np.empty(nt, np.float64)
    w, dN = _c(tab.weights, np.float64), _c(tab.dN, np.float64)
    lib().oracle_stiffness_ijv(mesh.N, _p(xyz), conn.shape[1], conn.shape[0], _p(conn), tab.q, _p(w),
                               _p(dN), E, nu, _p(I), _p(J), _p(V), nthreads)
    return sparse(3 * mesh.N, I, J, V, dropzeros)


def stiffness_csc(mesh, tab, D, colptr=None, rowval=None, nthreads=1, chunk=4096):
    """∫(SymGrad⋅D⋅SymGrad; csc_matrix) = assemble_operator(...; assembly=:csc)
    (src/multifield.jl:2394-2499, 2345-2380).  Returns nzval on the
    build_csc_pattern pattern."""
    if colptr is None:
        colptr, rowval = csc_pattern(mesh)
    xyz, conn = _c(mesh.coords, np.float64), _c(mesh.conn, np.int64)
    w, dN = _c(tab.weights, np.float64), _c(tab.dN, np.float64)
    Dm = _c(D, np.float64).reshape(36)
    nz = np.empty(rowval.size, np.float64)
    st = lib().oracle_stiffness_csc(mesh.N, _p(xyz), conn.shape[1], conn.shape[0], _p(conn), tab.q,
                                    _p(w), _p(dN), _p(Dm), _p(colptr), _p(rowval), _p(nz), nthreads,
                                    chunk)
    if st == -1:
        raise RuntimeError("compute_element_geometry!: singular geometry at integration point.")
    if st == -2:
        raise RuntimeError("find_csc_position: row is missing from the CSC pattern.")
    return nz


def mass_legacy(mesh, tab, rho=7.85e-9, lumped=False):
    """massMatrix(problem; lumped) = massMatrixSolid (src/linear.jl:1854-2004)."""
    xyz, conn = _c(mesh.coords, np.float64), _c(mesh.conn, np.int64)
    n3 = 3 * conn.shape[1]
    nt = conn.shape[0] * n3 * n3
    I, J, V = np.empty(nt, np.int64), np.empty(nt, np.int64), np.empty(nt, np.float64)
    w, Nr, dN = _c(tab.weights, np.float64), _c(tab.N, np.float64), _c(tab.dN, np.float64)
    lib().oracle_mass_ijv(mesh.N, _p(xyz), conn.shape[1], conn.shape[0], _p(conn), tab.q, _p(w),
                          _p(Nr), _p(dN), rho, _p(I), _p(J), _p(V))
    n = 3 * mesh.N
    M = sparse(n, I, J, V, dropzeros=False)
    if lumped:
        M = sp.diags(np.asarray(M.sum(axis=1)).ravel(), 0, shape=(n, n), format="csc")
    M.eliminate_zeros()
    return M


def traction(mesh, faceset, tab, t):
    """loadVector(problem, [load(name; fx,fy,fz)]) traction branch
    (src/linear.jl:3246-3413) on a surface, curve or point element set
    (tab.dim = 2, 1, 0)."""
    xyz, fc = _c(mesh.coords, np.float64), _c(faceset.conn, np.int64)
    w, Nr, dN = _c(tab.weights, np.float64), _c(tab.N, np.float64), _c(tab.dN, np.float64)
    f = np.empty(3 * mesh.N)
    tt = _c(t, np.float64)
    lib().oracle_traction(mesh.N, _p(xyz), fc.shape[1], tab.dim, fc.shape[0], _p(fc), tab.q, _p(w), _p(Nr),
                          _p(dN), _p(tt), _p(f))
    return f


def face_points(mesh, faceset, tab):
    """Physical Gauss-point coordinates of every face element, (F, q, 3):
    x = h[:, j]' * ncoord (src/linear.jl:3278-3280)."""
    X = mesh.coords[faceset.conn]                       # (F, nf, 3)
    return np.einsum("qa,fai->fqi", tab.N, X)


def traction_fn(mesh, faceset, tab, fns):
    """loadVector with function-valued components (fx, fy, fz callables of
    x, y, z, or numbers; None = 0), per face element and Gauss point as
    src/linear.jl:3277-3408 does it: t from _loadvec_helper (:3288-3299),
    Ja = |t1 x t2| on a surface (:3380-3384), |t1| on a curve (:3385-3386),
    1 at a point (:3387-3388), f1 += H1' t Ja w, fp[nn2] += f1 in element
    order.  Plain numpy loops over the face elements (small meshes)."""
    P = face_points(mesh, faceset, tab)
    F, q = P.shape[:2]
    T = np.zeros((F, q, 3))
    for c, fn in enumerate(fns):
        if fn is None:
            continue
        for e in range(F):
            for j in range(q):
                T[e, j, c] = fn(*P[e, j]) if callable(fn) else float(fn)
    f = np.zeros(3 * mesh.N)
    for e in range(F):
        nodes = faceset.conn[e]
        X = mesh.coords[nodes]
        f1 = np.zeros((len(nodes), 3))
        for j in range(q):
            t1 = tab.dN[j, :, 0] @ X
            t2 = tab.dN[j, :, 1] @ X
            Ja = (np.linalg.norm(np.cross(t1, t2)) if tab.dim == 2
                  else np.linalg.norm(t1) if tab.dim == 1 else 1.0)
            f1 += np.outer(tab.N[j], T[e, j]) * Ja * tab.weights[j]
        for a, v in enumerate(nodes):
            f[3 * v:3 * v + 3] += f1[a]
    return f, T


def volume_points(mesh, tab, elems=None):
    """Physical Gauss-point coordinates of volume elements, (ne, q, 3):
    x = h[:, j]' * ncoord (src/linear.jl:3278-3280)."""
    conn = mesh.conn if elems is None else mesh.conn[np.asarray(elems, dtype=np.int64)]
    return np.einsum("qa,eai->eqi", tab.N, mesh.coords[conn])


def body_force(mesh, tab, g=None, g_gp=None, elems=None):
    """loadVector on a volume group, the "Volume -> body force" case
    (src/linear.jl:3054-3068): per element and Gauss point f from
    _loadvec_helper (a constant g, or g_gp [ne][q][3] at volume_points),
    Ja = jacDet[j] (det J, DIM == dim == 3, src/linear.jl:3374-3375),
    f1 += H1' f Ja w, fp[nn2] += f1 in element order (src/linear.jl:3404-3413).
    The same integral is ∫(Pu ⋅ g; Ω) (assemble_element_vector!,
    src/multifield.jl:3386-3462).  Plain numpy loops (small meshes)."""
    elist = np.arange(mesh.E) if elems is None else np.asarray(elems, dtype=np.int64)
    f = np.zeros(3 * mesh.N)
    for k, e in enumerate(elist):
        nodes = mesh.conn[e]
        X = mesh.coords[nodes]
        f1 = np.zeros((len(nodes), 3))
        for j in range(tab.q):
            J = X.T @ tab.dN[j]                        # J[i][d] = sum_b x_b,i dN_b,d
            Ja = np.linalg.det(J)
            gq = np.asarray(g, dtype=np.float64) if g_gp is None else g_gp[k, j]
            f1 += np.outer(tab.N[j], gq) * Ja * tab.weights[j]
        for a, v in enumerate(nodes):
            f[3 * v:3 * v + 3] += f1[a]
    return f


def face_node_normals(mesh, faceset, dN_nodes):
    """normalVector (src/general.jl:4848-4940): per face element and local
    node k, the unit normal (∂x/∂ξ × ∂x/∂η)/|…| with the derivatives of the
    face map at the face's own reference node k.  (F, nf, 3)."""
    X = mesh.coords[faceset.conn]                      # (F, nf, 3)
    t1 = np.einsum("ka,fai->fki", dN_nodes[:, :, 0], X)
    t2 = np.einsum("ka,fai->fki", dN_nodes[:, :, 1], X)
    n = np.cross(t1, t2)
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


def pressure(mesh, faceset, tab, dN_nodes, p=0.0, p_nodal=None):
    """loadVector(problem, [load(name, p=p)]) (src/linear.jl:3196-3224):
    nv = -normalVector, the nodal traction t = elementsToNodes((nv ⋅ e_c) p)
    -- per node the sum over the face elements at it of -n_c p, times
    1/count (src/general.jl:5519-5640) -- then the traction branch with a
    nodal field, t(x_q) = h[:, j]' t[nnet] (_loadvec_helper,
    src/linear.jl:3033-3044), f1 += H1' t Ja w (:3404-3413).  p is a number
    or p_nodal [N] (a function's values at the nodes, scalarField,
    src/general.jl:4477-4515)."""
    nrm = face_node_normals(mesh, faceset, dN_nodes)
    F, nf = faceset.conn.shape
    acc = np.zeros((mesh.N, 3))
    cnt = np.zeros(mesh.N, dtype=np.int64)
    for e in range(F):
        for k, v in enumerate(faceset.conn[e]):
            pv = p if p_nodal is None else p_nodal[v]
            acc[v] += (-nrm[e, k]) * pv
            cnt[v] += 1
    tn = np.zeros((mesh.N, 3))
    hit = cnt > 0
    tn[hit] = acc[hit] * (1.0 / cnt[hit])[:, None]
    f = np.zeros(3 * mesh.N)
    for e in range(F):
        nodes = faceset.conn[e]
        X = mesh.coords[nodes]
        f1 = np.zeros((nf, 3))
        for j in range(tab.q):
            t1 = tab.dN[j, :, 0] @ X
            t2 = tab.dN[j, :, 1] @ X
            Ja = np.linalg.norm(np.cross(t1, t2))
            tq = tab.N[j] @ tn[nodes]
            f1 += np.outer(tab.N[j], tq) * Ja * tab.weights[j]
        for a, v in enumerate(nodes):
            f[3 * v:3 * v + 3] += f1[a]
    return f, tn


def pcg(K: sp.csc_matrix, f, fixed, ufix=None, reltol=1e-10, maxiter=100000):
    """solveDisplacement(K, f; support, iterative=true, Pl=Jacobi)
    (src/linear.jl:3711-3743) with IterativeSolvers.cg semantics."""
    n = K.shape[0]
    mask = np.zeros(n, np.uint8)
    mask[np.asarray(fixed, dtype=np.int64)] = 1
    uf = np.zeros(n)
    if ufix is not None:
        uf[np.asarray(fixed, dtype=np.int64)] = ufix
    cp, rv, v = _c(K.indptr, np.int64), _c(K.indices, np.int64), _c(K.data, np.float64)
    u = np.empty(n)
    rel = C.c_double()
    f = _c(f, np.float64)
    it = lib().oracle_pcg(n, _p(cp), _p(rv), _p(v), _p(f), _p(mask), _p(uf), reltol, maxiter, _p(u),
                          C.byref(rel))
    return u, int(it), rel.value


def num_threads():
    return lib().oracle_num_threads()
