"""Copies the reference's benchmark mesh into tests/golden/ (run once, from
the repository root, where the reference tree is mounted):

    python tests/golden/fetch_reference_mesh.py

examples/dev/compare_LLFEM_Ferrite_dealII_Gridap/elasticity_benchmark.msh is
the file gmsh wrote for ``structured_box_mesh(n = 30)``
(lowlevelfem_elasticity_benchmark.jl:26-27): 29,791 nodes, 27,000 hex8 and
the six face groups, MSH 4.1 ASCII.  The reference's recorded checksums on it
(results.txt:7) are sum(abs, K) = 4.0073076923076982e9 and sum(u) =
0.07057241535937334 (left clamped, traction (1, 0, 0) on right).  The copy is
gzip-compressed mesh data; tests read it with lowlevelfem.msh.read_msh.
"""
import gzip
import os
import shutil

SRC = "/root/reference/examples/dev/compare_LLFEM_Ferrite_dealII_Gridap/elasticity_benchmark.msh"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elasticity_benchmark_n30.msh.gz")

if __name__ == "__main__":
    with open(SRC, "rb") as fi, gzip.GzipFile(DST, "wb", compresslevel=9, mtime=0) as fo:
        shutil.copyfileobj(fi, fo)
    print(DST, os.path.getsize(DST))
