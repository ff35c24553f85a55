"""Structural checks of the Julia ccall binding (compactles.jl_amd/julia/
CompactLESHIP.jl) without a Julia toolchain:

  * every record that mirrors a C struct (the comment after `struct X` names
    it) has the C struct's fields, in the header's order, and Julia's C-layout
    rule for isbits structs (natural alignment, NTuple = C array) gives the
    same offsets and size as gcc's offsetof / sizeof;
  * every ccall names a function the header declares and passes as many
    argument types as the declaration has parameters.
"""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "compactles_hip.h")
JULIA = os.path.join(ROOT, "compactles.jl_amd", "julia", "CompactLESHIP.jl")

PRIM = {"Cint": (4, 4), "Cdouble": (8, 8), "Int64": (8, 8), "Csize_t": (8, 8), "Ptr{Cvoid}": (8, 8)}


def julia_structs():
    """name -> (c_name, [(field, type)]) for `struct X  # c_name` blocks."""
    src = open(JULIA).read()
    out = {}
    for m in re.finditer(r"^struct (\w+)\s*(?:#\s*(\w+))?\s*\n(.*?)^end", src, re.M | re.S):
        fields = []
        for line in m.group(3).splitlines():
            line = line.split("#")[0].strip()
            if not line:
                continue
            fm = re.fullmatch(r"(\w+)::(.+)", line)
            assert fm, (m.group(1), line)
            fields.append((fm.group(1), fm.group(2).strip()))
        out[m.group(1)] = (m.group(2), fields)
    return out


def layout(tname, structs, cache):
    """(size, align) of a Julia field type under the C layout rule."""
    if tname in PRIM:
        return PRIM[tname]
    m = re.fullmatch(r"NTuple\{(\d+),\s*(.+)\}", tname)
    if m:
        sz, al = layout(m.group(2), structs, cache)
        return int(m.group(1)) * sz, al
    if tname in cache:
        return cache[tname][0], cache[tname][1]
    _, fields = structs[tname]
    off, al, offs = 0, 1, []
    for _, ft in fields:
        sz, fa = layout(ft, structs, cache)
        off = (off + fa - 1) // fa * fa
        offs.append(off)
        off += sz
        al = max(al, fa)
    size = (off + al - 1) // al * al
    cache[tname] = (size, al, offs)
    return size, al


def header_struct_fields(c_name):
    """Member names of `typedef struct {...} c_name;` in declaration order."""
    src = open(HEADER).read()
    m = re.search(r"typedef struct\s*\{([^{}]*)\}\s*" + c_name + r"\s*;", src, re.S)
    assert m, c_name
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        fp = re.match(r".*\(\s*\*\s*(\w+)\s*\)", decl, re.S)       # function pointer
        if fp:
            names.append(fp.group(1))
            continue
        decl = re.sub(r"^(const\s+)?\w+(\s*\*)*\s*", "", decl, count=1)
        for d in decl.split(","):
            names.append(re.match(r"\s*\**\s*(\w+)", d).group(1))
    return names


def test_mirrored_records_match_the_header():
    structs = julia_structs()
    mirrored = {j: c for j, (c, _) in structs.items() if c and c.startswith("cles_")}
    assert {"ClesSolverDesc", "ClesSchemeDesc", "ClesNscbcDesc", "ClesNasa9Interval", "ClesNasa9Species",
            "ClesBatchInstance", "ClesRunHooks"} <= set(mirrored)
    cache = {}
    lines, want = [], []
    for jname, cname in sorted(mirrored.items()):
        fields = structs[jname][1]
        assert [f for f, _ in fields] == header_struct_fields(cname), jname
        size, _ = layout(jname, structs, cache)
        offs = cache[jname][2]
        lines.append('  printf("%%zu\\n", sizeof(%s));' % cname)
        want.append(size)
        for (f, _), o in zip(fields, offs):
            lines.append('  printf("%%zu\\n", offsetof(%s, %s));' % (cname, f))
            want.append(o)
    prog = "#include <stdio.h>\n#include <stddef.h>\n#include \"compactles_hip.h\"\nint main(void) {\n" + \
           "\n".join(lines) + "\n  return 0;\n}\n"
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        exe = os.path.join(td, "p")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got == want


def _split_top(s):
    """Split s at commas outside (), {} and []."""
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return [p.strip() for p in parts if p.strip()]


def _balanced(s, i):
    """s[i] == '(': the text inside the matching parenthesis."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return s[i + 1:j]
    raise ValueError("unbalanced")


def header_param_counts():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:int|const char \*|void \*)\s*(cles_\w+)\s*\(", src, re.M):
        params = _balanced(src, m.end() - 1)
        ps = _split_top(params)
        out[m.group(1)] = 0 if ps == ["void"] else len(ps)
    return out


def test_ccalls_match_the_header():
    decl = header_param_counts()
    src = open(JULIA).read()
    calls = list(re.finditer(r"ccall\(\(:(cles_\w+),\s*LIB\),", src))
    assert len(calls) >= 30
    for m in calls:
        name = m.group(1)
        assert name in decl, name
        rest = src[m.end():]
        args = _split_top(_balanced("(" + rest, 0))
        ret, types = args[0], args[1]
        assert ret in ("Cint", "Cstring", "Cvoid", "Ptr{Cvoid}"), (name, ret)
        assert types.startswith("(") and types.endswith(")"), (name, types)
        n_types = len(_split_top(types[1:-1]))
        assert n_types == decl[name], (name, n_types, decl[name])
        assert len(args) - 2 == n_types, (name, "argument values vs types")


@pytest.mark.parametrize("jname,cname", [("ClesSolverDesc", "cles_solver_desc")])
def test_flattened_arrays_cover_the_c_arrays(jname, cname):
    """Each NTuple field holds as many elements as the C member's array."""
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    consts = {"CLES_MAXSP": 8, "CLES_MAXCL": 8, "CLES_MAXRHS": 16, "CLES_MAXM": 8, "CLES_NASA9_MAXINT": 3}
    body = re.search(r"typedef struct\s*\{([^{}]*)\}\s*" + cname + r"\s*;", src, re.S).group(1)
    counts = {}
    for decl in body.split(";"):
        decl = re.sub(r"^\s*\w+\s+", "", decl.strip(), count=1)
        for d in decl.split(","):
            m = re.match(r"\s*(\w+)((?:\[[^\]]+\])*)", d)
            if not m:
                continue
            n = 1
            for dim in re.findall(r"\[([^\]]+)\]", m.group(2)):
                n *= eval(dim, {}, consts)
            counts[m.group(1)] = n
    for f, t in julia_structs()[jname][1]:
        m = re.fullmatch(r"NTuple\{(\d+),.*\}", t)
        assert (int(m.group(1)) if m else 1) == counts[f], f
