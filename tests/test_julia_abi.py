"""The reference's Julia package (bindings/julia/src/*.jl) binds the C ABI
through 42 `ccall`s; tests/golden/julia_ccalls.json records each one's
name, return type and argument types and the layouts of the result structs
it passes by reference (written from the Julia sources as text by
tests/golden/make_julia_ccalls.py). Julia is not installed here, so this
checks, on the CPU, what `QKRYLOV_LIB_PATH=libqkrylov_amd.so` needs:
every called symbol is exported, and its prototype in include/qkrylov/c_api.h
has the same arity, return type and argument types (and the same struct
field types) under the Julia-to-C type map."""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "julia_ccalls.json")))
HEADER = open(os.path.join(ROOT, "include", "qkrylov", "c_api.h")).read()
LIB = os.path.join(ROOT, "qkrylov_amd", "libqkrylov_amd.so")

JULIA = {"Cvoid": "void", "Cint": "int", "Cfloat": "float", "Cdouble": "double", "UInt64": "uint64",
         "Int64": "int64", "Csize_t": "size", "Cstring": "char*", "Ptr{Cvoid}": "handle", "Ptr{Cfloat}": "float*",
         "Ptr{Cdouble}": "double*", "Ptr{Cint}": "int*", "Ptr{Ptr{Cchar}}": "char**", "Ref{Cint}": "int*",
         "Ref{Cfloat}": "float*"}
STRUCTS = {"LanczosResultC": "qkrylov_lanczos_result_c_t", "DavidsonResultC": "qkrylov_davidson_result_c_t",
           "FTLMResultC": "qkrylov_ftlm_result_c_t"}


def julia_type(t):
    m = re.fullmatch(r"Ref\{(\w+C)\}", t)
    if m:
        return STRUCTS[m.group(1)] + "*"
    return JULIA[t]


def c_type(t):
    t = " ".join(t.replace("const", " ").split())
    t = t.replace(" *", "*")
    if re.fullmatch(r"qkrylov_\w+_h", t) or t == "void*":
        return "handle"
    return {"uint64_t": "uint64", "int64_t": "int64", "size_t": "size"}.get(t, t)


def prototypes():
    out = {}
    flat = re.sub(r"/\*.*?\*/", " ", HEADER, flags=re.S)
    for m in re.finditer(r"([\w\s\*]+?)\b(qkrylov_\w+)\s*\(([^;{]*?)\)\s*;", flat):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3).strip()
        params = [] if args in ("", "void") else [a.strip() for a in args.split(",")]
        # drop the parameter name: the last identifier
        types = [re.sub(r"\b\w+\s*$", "", p).strip() or p for p in params]
        out[name] = (c_type(ret), [c_type(t) for t in types])
    return out


def struct_fields(cname):
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + cname + r";", HEADER)
    assert m, cname
    return [tuple(f.split()) for f in re.findall(r"^\s*(\w+\s+\w+);", m.group(1), re.M)]


def test_fixture_is_complete():
    assert len(FIX["calls"]) == 42 and set(FIX["structs"]) == set(STRUCTS)


@pytest.mark.parametrize("name", sorted(FIX["calls"]))
def test_julia_call_matches_header(name):
    sig = FIX["calls"][name]
    protos = prototypes()
    assert name in protos, f"{name} is not declared in c_api.h"
    ret, args = protos[name]
    assert ret == julia_type(sig["ret"]), (name, ret, sig["ret"])
    assert args == [julia_type(a) for a in sig["args"]], (name, args, sig["args"])


def test_julia_structs_match_header():
    jmap = {"Cfloat": "float", "Cint": "int", "Cdouble": "double"}
    for jname, fields in FIX["structs"].items():
        want = [(jmap[t], f) for f, t in fields]
        assert struct_fields(STRUCTS[jname]) == want, jname


def test_every_julia_symbol_is_exported():
    if not os.path.exists(LIB):
        pytest.skip("engine library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted(set(FIX["calls"]) - exported)
    assert not missing, missing
