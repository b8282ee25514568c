"""The MSH 4.1 reader on the reference's own benchmark mesh
(examples/dev/compare_LLFEM_Ferrite_dealII_Gridap/elasticity_benchmark.msh,
written by gmsh for structured_box_mesh(n = 30), gzip-compressed under
tests/golden/).  Pins: the counts gmsh reported reading it (results.txt:10-12:
29,791 nodes, 32,400 elements = 27,000 hex8 + 6 x 900 quads); the node set
equals the 31^3 lattice of the unit cube; every hex has a positive Jacobian
with the refelem hex8 order; the structural nnz is 9 (3n+1)^3 for n = 30;
and the oracle's stiffness on gmsh's own numbering reproduces the
reference's recorded sum(abs, K) = 4.0073076923076982e9 (results.txt:7)."""
import os

import numpy as np
import pytest

from conftest import hooke_symgrad_D
from lowlevelfem import mesh as M
from lowlevelfem import refelem as R
from lowlevelfem.msh import read_msh
from oracle import oracle as O

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "elasticity_benchmark_n30.msh.gz")


@pytest.fixture(scope="module")
def mesh():
    return read_msh(FIXTURE)


def test_counts_and_groups(mesh):
    assert (mesh.N, mesh.E, mesh.etype, mesh.npe) == (29791, 27000, 5, 8)
    assert sorted(mesh.faces) == ["bottom", "front", "left", "rear", "right", "top"]
    for name, fs in mesh.faces.items():
        assert fs.etype == 3 and fs.conn.shape == (900, 4), name
    assert mesh.nodes_on("left").size == 31 * 31
    assert np.allclose(mesh.coords[mesh.nodes_on("left"), 0], 0.0)
    assert np.allclose(mesh.coords[mesh.nodes_on("right"), 0], 1.0)


def test_node_set_is_the_lattice(mesh):
    ref = M.structured_box_mesh(n=30).coords
    a = np.round(mesh.coords * 30).astype(np.int64)
    assert np.allclose(mesh.coords * 30, a, atol=1e-9)
    b = np.round(ref * 30).astype(np.int64)
    key = lambda p: np.sort(p[:, 0] * 10000 + p[:, 1] * 100 + p[:, 2])
    assert np.array_equal(key(a), key(b))


def test_hex_orientation(mesh):
    dN = R.rule_tables(5, 1).dN[0]                       # (8, 3) at the cell centre
    X = mesh.coords[mesh.conn]                           # (E, 8, 3)
    J = np.einsum("eai,ak->eik", X, dN)
    assert np.all(np.linalg.det(J) > 0)


def test_stiffness_checksum_on_gmsh_numbering(mesh):
    cp, rv = O.csc_pattern(mesh)
    assert cp[-1] == 9 * 91 ** 3
    K = O.stiffness_csc(mesh, R.rule_tables(5, 3), hooke_symgrad_D(), cp, rv, nthreads=4)
    assert abs(np.abs(K).sum() / 4.0073076923076982e9 - 1) < 1e-12


def test_reader_errors(tmp_path):
    bad = tmp_path / "old.msh"
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    with pytest.raises(NotImplementedError):
        read_msh(bad)
    gap = tmp_path / "gap.msh"
    gap.write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 2 1 3\n3 1 0 2\n1\n3\n0 0 0\n1 0 0\n"
                   "$EndNodes\n$Elements\n0 0 0 0\n$EndElements\n")
    with pytest.raises(ValueError, match="node tags must be 1..N"):
        read_msh(gap)


def test_mixed_volume_types_round_trip(tmp_path):
    """A file with hexahedra and wedges: one block per volume element type
    (types in order of first appearance, elements in file order), material
    groups per block, and a face group of two face types as a list."""
    from mixed_mesh import hex_prism_box, write_msh
    m = hex_prism_box()
    path = tmp_path / "hexprism.msh"
    write_msh(m, path)
    r = read_msh(path)
    blocks = r.element_blocks()
    assert [b.etype for b in blocks] == [5, 6]
    for got, want in zip(blocks, m.element_blocks()):
        for g in ("steel", "alu"):
            a = got.conn[got.volumes[g]]
            b = want.conn[want.volumes[g]]
            assert np.array_equal(a, b)
        assert got.conn.shape == want.conn.shape
    assert np.allclose(r.coords, m.coords, rtol=0, atol=1e-15)
    front = r.face_sets("front")
    assert [fs.etype for fs in front] == [3, 2]
    assert np.array_equal(np.unique(r.nodes_on("front")), np.unique(m.nodes_on("front")))
    assert np.array_equal(r.nodes_on("steel"), m.nodes_on("steel"))


def _same_mesh(a, b):
    assert np.array_equal(a.coords, b.coords)
    for x, y in zip(a.element_blocks(), b.element_blocks()):
        assert x.etype == y.etype and np.array_equal(x.conn, y.conn)
        assert sorted(x.volumes) == sorted(y.volumes)
        for g in x.volumes:
            assert np.array_equal(x.volumes[g], y.volumes[g])
    assert sorted(a.faces) == sorted(b.faces)
    for f in a.faces:
        for p, q in zip(a.face_sets(f), b.face_sets(f)):
            assert p.etype == q.etype and np.array_equal(p.conn, q.conn)


def test_binary_matches_ascii(tmp_path):
    """The same mesh written as MSH 4.1 ASCII and binary (little-endian, data
    size 8) reads back identically, both equal to the written mesh."""
    from mixed_mesh import hex_prism_box, write_msh
    m = hex_prism_box()
    write_msh(m, tmp_path / "a.msh")
    write_msh(m, tmp_path / "b.msh", binary=True)
    a, b = read_msh(tmp_path / "a.msh"), read_msh(tmp_path / "b.msh")
    _same_mesh(a, b)
    assert [x.etype for x in b.element_blocks()] == [5, 6]
    assert np.array_equal(b.coords, m.coords)


def test_binary_reader_errors(tmp_path):
    from mixed_mesh import hex_prism_box, write_msh
    m = hex_prism_box(cells=(2, 2, 2))
    write_msh(m, tmp_path / "b.msh", binary=True)
    data = (tmp_path / "b.msh").read_bytes()
    (tmp_path / "d4.msh").write_bytes(data.replace(b"4.1 1 8", b"4.1 1 4", 1))
    with pytest.raises(NotImplementedError, match="data size 4"):
        read_msh(tmp_path / "d4.msh")
    (tmp_path / "cut.msh").write_bytes(data[: len(data) // 2])
    with pytest.raises(ValueError):
        read_msh(tmp_path / "cut.msh")
    (tmp_path / "noend.msh").write_bytes(data.replace(b"$MeshFormat\n4.1 1 8\n\x01\x00\x00\x00",
                                                      b"$MeshFormat\n4.1 1 8\n\x07\x00\x00\x00", 1))
    with pytest.raises(ValueError, match="endianness"):
        read_msh(tmp_path / "noend.msh")


def test_gmsh_recursive_order_reproduces_documented_hex27_quad9():
    """At order 2 the recursive rule (vertices, edges, face and volume
    interiors as lower-order elements) gives gmsh's documented hex27 and quad9
    orders, which the element-local order follows."""
    from lowlevelfem import refelem as RE
    from lowlevelfem.msh import gmsh_hex_lattice, gmsh_quad_lattice, gmsh_to_local
    assert np.array_equal(np.array(gmsh_hex_lattice(2)), RE.hex_node_lattice(2))
    assert np.array_equal(np.array(gmsh_quad_lattice(2)), RE.quad_node_lattice(2))
    for et, npe in ((92, 64), (93, 125), (36, 16), (37, 25)):
        perm = gmsh_to_local(et)
        assert np.array_equal(np.sort(perm), np.arange(npe))
    # vertices and edge nodes keep their places; only face and volume interiors move
    assert np.array_equal(gmsh_to_local(93)[:8 + 12 * 3], np.arange(8 + 12 * 3))
    assert not np.array_equal(gmsh_to_local(93), np.arange(125))


@pytest.mark.parametrize("order", [3, 4])
@pytest.mark.parametrize("binary", [False, True], ids=["ascii", "binary"])
def test_high_order_hex_round_trip(tmp_path, order, binary):
    """hex64 / hex125 meshes (quad16 / quad25 faces) written in gmsh's node
    order read back in the element-local order: the same connectivity, the
    same faces, and the same stiffness from the oracle.  In the written file,
    every element's gmsh-ordered nodes sit at the gmsh lattice positions of
    its own corner box (structured, unperturbed mesh)."""
    from mixed_mesh import write_msh
    from lowlevelfem.msh import gmsh_hex_lattice
    from lowlevelfem import refelem as RE
    from conftest import hooke_symgrad_D
    m = M.structured_box_mesh(cells=(2, 1, 2), lx=1.0, ly=0.5, lz=1.0, order=order)
    m.volumes["body"] = np.arange(m.E)
    path = tmp_path / f"hex{order}.msh"
    write_msh(m, path, binary=binary)
    r = read_msh(path)
    _same_mesh(r, m)
    # gmsh positions: node j of an element in gmsh order at lo + lattice_j / p * (hi - lo)
    from lowlevelfem.msh import gmsh_to_local
    gconn = m.conn[:, np.argsort(gmsh_to_local(m.etype))]
    lat = np.array(gmsh_hex_lattice(order), dtype=np.float64) / order
    for e in range(m.E):
        X = m.coords[gconn[e]]
        lo, hi = X.min(axis=0), X.max(axis=0)
        assert np.allclose(X, lo + lat * (hi - lo), atol=1e-14)
    tab = RE.rule_tables(m.etype, 2 * order + 1)
    cp, rv = O.csc_pattern(m)
    D = hooke_symgrad_D()
    assert np.array_equal(O.stiffness_csc(r, tab, D, *O.csc_pattern(r)), O.stiffness_csc(m, tab, D, cp, rv))


@pytest.mark.parametrize("order", [1, 2, 4])
@pytest.mark.parametrize("binary", [False, True], ids=["ascii", "binary"])
def test_curve_and_point_groups_round_trip(tmp_path, order, binary):
    """Physical curves (line2 / line3 / line5) and points (one point element,
    and two with a node listed twice) are kept as boundary groups with their
    element type and node order; their nodes are the group's nodes."""
    from mixed_mesh import with_edge_and_corner, write_msh
    m = with_edge_and_corner(M.structured_box_mesh(n=2, order=order))
    m.volumes = {"body": np.arange(m.E)}
    path = tmp_path / f"edges{order}.msh"
    write_msh(m, path, binary=binary)
    r = read_msh(path)
    for name in ("edge", "corner", "corners2", "left"):
        assert r.faces[name].etype == m.faces[name].etype
        assert np.array_equal(r.faces[name].conn, m.faces[name].conn), name
    assert np.array_equal(r.nodes_on("corner"), m.faces["corner"].conn[0])
    assert r.nodes_on("edge").size == 2 * order + 1
