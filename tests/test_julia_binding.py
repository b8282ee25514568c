"""The Julia binding (lowlevelfem.jl_amd/julia/LowLevelFEMHIPExt.jl) against
include/llfem.h, without Julia: every struct it mirrors has the C struct's
fields in the C order with matching types, and the C layout that gcc gives
(offsetof / sizeof) equals the layout Julia gives those field types; every
constructor call passes one value per field; every ccall names a declared
entry point with one Julia argument type per C parameter, each compatible
with the C type, and passes as many arguments as it declares types.  The
ctypes structs of the Python binding are held to the same C layout, and the
copy of the binding in INTEGRATION.md is the file itself."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT
from lowlevelfem import _lib

JL = os.path.join(ROOT, "lowlevelfem.jl_amd", "julia", "LowLevelFEMHIPExt.jl")
HDR = os.path.join(ROOT, "include", "llfem.h")
PAIRS = {"OpDesc": "llfem_op_desc", "Stats": "llfem_stats", "Request": "llfem_request",
         "RequestStats": "llfem_request_stats"}


def _src():
    return open(JL).read()


def _hdr():
    return re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)


def split_top(s, sep=","):
    """Split s at separators outside (), {}, [] and string literals."""
    out, depth, cur, q = [], 0, "", None
    for ch in s:
        if q:
            cur += ch
            if ch == q:
                q = None
            continue
        if ch == '"':
            q = ch
        elif ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == sep and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def julia_structs(src):
    out = {}
    for m in re.finditer(r"^(?:mutable )?struct (\w+)[^\n]*\n(.*?)^end", src, flags=re.S | re.M):
        fields = []
        for line in m.group(2).splitlines():
            line = line.split("#")[0].strip()
            if not line or "=" in line:
                continue
            for item in line.split(";"):
                item = item.strip()
                if item:
                    name, typ = item.split("::")
                    fields.append((name.strip(), typ.strip()))
        out[m.group(1)] = fields
    return out


def c_structs(hdr):
    out = {}
    for m in re.finditer(r"typedef struct \{(.*?)\}\s*(\w+);", hdr, flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            mm = re.match(r"((?:const )?[\w ]+?\**)\s*(\w+(?:\[\d+\])?(?:\s*,\s*\w+(?:\[\d+\])?)*)$", decl)
            assert mm, decl
            base = mm.group(1).replace(" ", "") if "*" in mm.group(1) else mm.group(1).strip()
            for nm in mm.group(2).split(","):
                nm = nm.strip()
                arr = re.match(r"(\w+)\[(\d+)\]", nm)
                fields.append((arr.group(1), base, int(arr.group(2))) if arr else (nm, base, 0))
        out[m.group(2)] = fields
    return out


C_TO_JL = {"int32_t": "Int32", "int64_t": "Int64", "uint64_t": "UInt64", "double": "Float64",
           "constdouble*": "Ptr{Float64}", "double*": "Ptr{Float64}", "constint64_t*": "Ptr{Int64}",
           "llfem_op_desc": "OpDesc"}


def julia_layout(typ, structs):
    """(size, alignment) of a Julia field type as Julia lays out isbits structs."""
    prim = {"Int32": 4, "UInt32": 4, "Int64": 8, "UInt64": 8, "Float64": 8}
    if typ in prim:
        return prim[typ], prim[typ]
    if typ.startswith("Ptr{"):
        return 8, 8
    m = re.match(r"NTuple\{(\d+),(\w+)\}", typ.replace(" ", ""))
    if m:
        s, a = julia_layout(m.group(2), structs)
        return int(m.group(1)) * s, a
    offs, size, align = julia_offsets(structs[typ], structs)
    return size, align


def julia_offsets(fields, structs):
    off, align, offs = 0, 1, []
    for _, t in fields:
        s, a = julia_layout(t, structs)
        off = (off + a - 1) // a * a
        offs.append(off)
        off += s
        align = max(align, a)
    return offs, (off + align - 1) // align * align, align


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    """offsetof / sizeof of every field of the mirrored structs, from gcc."""
    cs = c_structs(_hdr())
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "llfem.h"', "int main(void) {"]
    for cname in list(PAIRS.values()) + ["llfem_tune_result"]:
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _, _ in cs[cname]:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    d = tmp_path_factory.mktemp("layout")
    (d / "layout.c").write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", str(d / "layout"), str(d / "layout.c")], check=True)
    out = subprocess.run([str(d / "layout")], check=True, capture_output=True, text=True).stdout
    lay = {}
    for line in out.splitlines():
        s, f, v = line.split()
        lay.setdefault(s, {})[f] = int(v)
    return lay


def test_struct_fields_match_header():
    js, cs = julia_structs(_src()), c_structs(_hdr())
    for jname, cname in PAIRS.items():
        jf, cf = js[jname], cs[cname]
        assert [n for n, _ in jf] == [n for n, _, _ in cf], jname
        for (n, jt), (_, ct, arr) in zip(jf, cf):
            want = f"NTuple{{{arr},{C_TO_JL[ct]}}}" if arr else C_TO_JL[ct]
            assert jt.replace(" ", "") == want, (jname, n, jt, ct)


def test_struct_layout_matches_gcc(c_layout):
    js = julia_structs(_src())
    for jname, cname in PAIRS.items():
        offs, size, _ = julia_offsets(js[jname], js)
        assert size == c_layout[cname]["sizeof"], jname
        for (n, _), o in zip(js[jname], offs):
            assert o == c_layout[cname][n], (jname, n)


def test_ctypes_layout_matches_gcc(c_layout):
    for cls, cname in [(_lib.OpDesc, "llfem_op_desc"), (_lib.Stats, "llfem_stats"),
                       (_lib.RequestSpec, "llfem_request"), (_lib.RequestStats, "llfem_request_stats"),
                       (_lib.TuneResult, "llfem_tune_result")]:
        lay = dict(c_layout[cname])
        assert ctypes.sizeof(cls) == lay.pop("sizeof"), cname
        assert [n for n, _ in cls._fields_] == list(lay), cname
        for n, _ in cls._fields_:
            assert getattr(cls, n).offset == lay[n], (cname, n)


def _calls(src, fname):
    """Argument lists of every call `fname(...)` (not definitions)."""
    out = []
    for m in re.finditer(r"(?<![\w.])" + fname + r"\(", src):
        line_start = src.rfind("\n", 0, m.start()) + 1
        if re.match(r"\s*(?:mutable )?struct\b", src[line_start:m.start()]):
            continue
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        out.append(split_top(src[m.end():i - 1]))
    return out


def test_constructors_pass_every_field():
    src = _src()
    js = julia_structs(src)
    for name in ("Stats", "RequestStats"):
        body = re.search(r"^mutable struct " + name + r"\b.*?^end", src, flags=re.S | re.M).group(0)
        args = _calls(body, "new")
        assert len(args) == 1 and len(args[0]) == len(js[name]), name
    descs = _calls(src, "OpDesc")
    assert len(descs) >= 2
    for args in descs:
        assert len(args) == len(js["OpDesc"]), args


def header_prototypes():
    out = {}
    for m in re.finditer(r"LLFEM_API\s+([\w\s\*]+?)\s*\b(llfem_\w+)\s*\((.*?)\)\s*;", _hdr(), flags=re.S):
        ret = m.group(1).replace(" ", "")
        params = []
        for p in split_top(" ".join(m.group(3).split())):
            if p == "void":
                continue
            arr = re.search(r"\[(\d+)\]$", p)
            p = re.sub(r"\[\d+\]$", "", p)
            typ = re.sub(r"\b\w+$", "", p).replace(" ", "")
            params.append(typ + ("[]" if arr else ""))
        out[m.group(2)] = (ret, params)
    return out


HANDLES = ("llfem_ctx", "llfem_mesh", "llfem_server")
STRUCT_J = {"llfem_op_desc": "OpDesc", "llfem_stats": "Stats", "llfem_request": "Request",
            "llfem_request_stats": "RequestStats", "llfem_tune_result": "TuneResult"}


def allowed_julia(ctype):
    t = ctype.replace("const", "")
    scalar = {"int": {"Cint", "Int32"}, "int32_t": {"Int32", "Cint"}, "int64_t": {"Int64"},
              "double": {"Float64", "Cdouble"}, "size_t": {"Csize_t", "UInt64"}}
    if t in scalar:
        return scalar[t]
    if t == "double[]":
        return {"Ptr{Float64}", "Ref{NTuple{3,Float64}}"}
    if any(t == h + "*" for h in HANDLES):
        return {"Ptr{Cvoid}"}
    if any(t == h + "**" for h in HANDLES):
        return {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"}
    if t == "llfem_mesh**" or t == "llfem_mesh*const*" or t == "llfem_mesh**":
        return {"Ptr{Ptr{Cvoid}}"}
    base = t.rstrip("*")
    if t.count("*") == 1:
        if base in STRUCT_J:
            return {f"Ref{{{STRUCT_J[base]}}}", f"Ptr{{{STRUCT_J[base]}}}"}
        prim = {"double": "Float64", "int64_t": "Int64", "int32_t": "Int32", "uint8_t": "UInt8"}
        if base in prim:
            return {f"Ptr{{{prim[base]}}}", f"Ref{{{prim[base]}}}"}
    raise AssertionError(f"no Julia mapping for C type {ctype}")


def julia_ccalls(src):
    out = []
    for args in _calls(src, "ccall"):
        name = re.match(r"\(:(\w+),\s*lib\)", args[0]).group(1)
        ret = args[1]
        tup = args[2].strip()
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = [t.replace(" ", "") for t in split_top(tup[1:-1])]
        out.append((name, ret, types, args[3:]))
    return out


def test_ccalls_match_prototypes():
    protos = header_prototypes()
    calls = julia_ccalls(_src())
    assert len(calls) >= 15
    ret_map = {"int": "Cint", "constchar*": "Cstring", "size_t": "Csize_t"}
    seen = set()
    for name, ret, types, args in calls:
        assert name in protos, f"{name} is not declared in llfem.h"
        cret, params = protos[name]
        assert ret == ret_map[cret], (name, ret, cret)
        assert len(types) == len(params), (name, types, params)
        assert len(args) == len(types), (name, len(args), len(types))
        for jt, ct in zip(types, params):
            assert jt in allowed_julia(ct), (name, jt, ct)
        seen.add(name)
    # the entry points the binding must reach
    for need in ("llfem_ctx_create", "llfem_mesh_create", "llfem_mesh_pattern", "llfem_assemble",
                 "llfem_load_traction", "llfem_solve", "llfem_solve_csc", "llfem_assemble_batch",
                 "llfem_server_submit", "llfem_server_wait"):
        assert need in seen, need


def test_integration_doc_holds_the_binding():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```julia\n(.*?)```", doc, flags=re.S)
    assert _src() in blocks


def test_element_scale_is_indexed_by_element_tags():
    """group_scale looks element tags up in the map built from the element
    tags of getElementsByType (its first return value), never node tags."""
    src = _src()
    assert re.search(r"elemTags, nodeTags = gmsh\.model\.mesh\.getElementsByType\(et\)", src)
    assert re.search(r"pos = Dict|Dict\(t => i for \(i, t\) in enumerate\(etags\)\)", src)
    body = re.search(r"^function group_scale.*?^end", src, flags=re.S | re.M).group(0)
    assert "elemTags, _ = gmsh.model.mesh.getElementsByType(m.et, etag)" in body
    assert "m.pos[Int64(t)]" in body
