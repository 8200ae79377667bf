"""The Julia binding (sphtogrid.jl_amd/julia/SPHtoGridAMD.jl) cannot run here
(no Julia), so its foreign calls are checked mechanically against the C ABI:
every `ccall` names a symbol include/sph2grid.h declares, with the return
type, the number and the order of argument types of that prototype, and
passes as many arguments as its type tuple lists; the Julia mirrors of the
ABI structs have the C structs' fields in the same order.  CPU only.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "sphtogrid.jl_amd", "julia", "SPHtoGridAMD.jl")
HDR = os.path.join(ROOT, "include", "sph2grid.h")

# C parameter type -> Julia ccall types accepted for it
C2JL = {
    "s2g_ctx*": {"Ptr{Cvoid}"},
    "const s2g_ctx*": {"Ptr{Cvoid}"},
    "s2g_ctx**": {"Ptr{Ptr{Cvoid}}"},
    "int": {"Cint"},
    "int64_t": {"Int64"},
    "uint32_t": {"UInt32"},
    "const void*": {"Ptr{T}", "Ptr{Cvoid}", "Ptr{Float32}", "Ptr{Float64}"},
    "void*": {"Ptr{Cvoid}"},
    "double*": {"Ptr{Float64}"},
    "const void*const*": {"Ptr{Ptr{T}}", "Ptr{Ptr{Cvoid}}"},
    "double*const*": {"Ptr{Ptr{Float64}}"},
    "const double*": {"Ptr{Float64}"},
    "const s2g_params*": {"Ref{S2GParams}", "Ptr{S2GParams}"},
    "s2g_stats*": {"Ptr{Cvoid}"},
    "const s2g_map_desc*": {"Ptr{S2GMapDesc}"},
    "const s2g_tuning*": {"Ref{S2GTuning}", "Ptr{S2GTuning}"},
    "double": {"Cdouble", "Float64"},
    "s2g_plan**": {"Ptr{Ptr{Cvoid}}"},
    "s2g_plan*": {"Ptr{Cvoid}"},
}
CRET = {"int": "Cint", "void": "Cvoid", "const char*": "Cstring"}
# C struct field type -> Julia field type
CFIELD = {"int64_t": "Int64", "int32_t": "Int32", "uint32_t": "UInt32", "double": "Float64",
          "s2g_params": "S2GParams", "const void*": "Ptr{Cvoid}", "double*": "Ptr{Float64}"}


def _strip_comments(src):
    return re.sub(r"/\*.*?\*/", "", src, flags=re.S)


def _norm(t):
    t = re.sub(r"\s+", " ", t.strip())
    return re.sub(r"\s*\*", "*", t)


def prototypes():
    src = _strip_comments(open(HDR).read())
    out = {}
    for ret, name, args in re.findall(r"(int|void|const char \*)\s*(s2g_\w+)\s*\(([^;]*?)\)\s*;",
                                      src, flags=re.S):
        types = []
        a = args.strip()
        if a and a != "void":
            for p in a.split(","):
                p = p.strip()
                m = re.match(r"^(.*?)([A-Za-z_]\w*)\s*(\[\d*\])?$", p)
                assert m, p
                t = _norm(m.group(1))
                if m.group(3):
                    t += "*"
                types.append(t)
        out[name] = (_norm(ret), types)
    return out


def _split_top(s):
    """Split s at commas outside (), {} and []."""
    parts, depth, cur = [], 0, []
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        parts.append("".join(cur).strip())
    return parts


def ccalls():
    src = open(JL).read()
    src = re.sub(r"#[^\n]*", "", src)
    out = []
    for m in re.finditer(r"ccall\(", src):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        parts = _split_top(src[m.end():i - 1])
        name = re.match(r"\(\s*:(\w+)\s*,\s*libsph2grid\s*\)", parts[0])
        assert name, parts[0]
        tup = parts[2].strip()
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = [re.sub(r"\s+", "", t) for t in _split_top(tup[1:-1])]
        out.append((name.group(1), parts[1].strip(), types, parts[3:]))
    return out


def c_struct_fields(name):
    src = _strip_comments(open(HDR).read())
    m = re.search(r"typedef struct %s \{(.*?)\}\s*%s;" % (name, name), src, flags=re.S)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        base = re.match(r"^((?:const\s+)?\w+)\s*(.*)$", decl)
        for d in base.group(2).split(","):
            d = d.strip()
            fields.append((d.lstrip("*").strip(), _norm(base.group(1) + "*" * d.count("*"))))
    return fields


def jl_struct_fields(name):
    src = open(JL).read()
    m = re.search(r"struct %s\n(.*?)\nend" % name, src, flags=re.S)
    assert m, name
    return [tuple(x.strip() for x in line.split("::")) for line in m.group(1).splitlines()
            if "::" in line]


def test_header_parses():
    p = prototypes()
    assert p["s2g_deposit"][1][:3] == ["s2g_ctx*", "int", "int"]
    assert len(p["s2g_deposit"][1]) == 17 and p["s2g_create"] == ("int", ["s2g_ctx**", "int"])


def test_every_ccall_matches_its_prototype():
    protos = prototypes()
    calls = ccalls()
    assert len(calls) >= 10
    for name, ret, types, args in calls:
        assert name in protos, f"ccall of {name}, which include/sph2grid.h does not declare"
        cret, ctypes_ = protos[name]
        assert ret == CRET[cret], (name, ret, cret)
        assert len(types) == len(ctypes_), (name, types, ctypes_)
        for k, (jt, ct) in enumerate(zip(types, ctypes_)):
            assert jt in C2JL[ct], f"{name} argument {k}: Julia {jt} for C {ct}"
        assert len(args) == len(types), (name, len(args), len(types))


@pytest.mark.parametrize("cname, jname", [("s2g_params", "S2GParams"),
                                          ("s2g_map_desc", "S2GMapDesc"),
                                          ("s2g_tuning", "S2GTuning")])
def test_struct_mirrors_match(cname, jname):
    c = c_struct_fields(cname)
    j = jl_struct_fields(jname)
    assert [f for f, _ in c] == [f for f, _ in j], (c, j)
    for (f, ct), (_, jt) in zip(c, j):
        assert CFIELD[ct] == jt, (f, ct, jt)


def test_binding_covers_the_drop_in_surface():
    src = open(JL).read()
    for fn in ("function sphMapping(", "function healpix_map(", "function healpix_deposit!(",
               "function cic_mapping_2D(", "function cic_mapping_3D(",
               "function healpix_deposit_maps!(",
               "function reduce_image_2D(", "function reduce_image_3D(",
               "function mass_density(", "function TSCInterpolation(",
               "function deposit_batch!("):
        assert fn in src, fn
    # parallel = true is one GPU call on the calling process: no Distributed
    # fan-out in the code (docstrings and comments removed)
    code = re.sub(r'"""[\s\S]*?"""', "", src)
    code = re.sub(r"#[^\n]*", "", code)
    assert "@spawnat" not in code and "fetch(" not in code and "Distributed" not in code
