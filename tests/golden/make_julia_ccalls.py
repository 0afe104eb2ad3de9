"""Writes tests/golden/julia_ccalls.json: every foreign call the reference's
Julia package makes (bindings/julia/src/*.jl, `ccall((:name, libqkrylov),
ret, (args...), ...)`), as (name, return type, argument types), read from
the Julia sources as text.

Run in the build container (where the reference checkout exists):
    python tests/golden/make_julia_ccalls.py [/root/reference]
"""
import glob
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def split_top(s):
    """Split at commas that are not inside (), {} or []."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def balanced(s, i):
    """Index just past the parenthesised group that starts at s[i] == '('."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j + 1
    raise ValueError("unbalanced")


def calls(src):
    for m in re.finditer(r"ccall\(\s*\(:(\w+),\s*libqkrylov\),\s*", src):
        name, pos = m.group(1), m.end()
        # return type up to the next top-level comma
        depth, j = 0, pos
        while not (src[j] == "," and depth == 0):
            depth += src[j] in "({[" and 1 or (src[j] in ")}]" and -1 or 0)
            j += 1
        ret = src[pos:j].strip()
        k = j + 1
        while src[k].isspace():
            k += 1
        end = balanced(src, k)
        args = split_top(src[k + 1:end - 1])
        yield name, ret, args


def structs(src):
    """Julia structs passed by reference to the library (names ending in C):
    field names and types in declaration order."""
    for m in re.finditer(r"^struct (\w+C)\n(.*?)^end", src, re.M | re.S):
        fields = re.findall(r"^\s*(\w+)::(\w+)\s*$", m.group(2), re.M)
        yield m.group(1), [list(f) for f in fields]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    seen, layouts = {}, {}
    for f in sorted(glob.glob(os.path.join(ref, "bindings", "julia", "src", "*.jl"))):
        src = open(f).read()
        for name, ret, args in calls(src):
            sig = {"ret": ret, "args": args, "file": os.path.relpath(f, ref)}
            if name in seen and (seen[name]["ret"], seen[name]["args"]) != (ret, args):
                raise SystemExit(f"{name}: two different signatures")
            seen.setdefault(name, sig)
        for name, fields in structs(src):
            layouts[name] = fields
    out = os.path.join(HERE, "julia_ccalls.json")
    json.dump({"calls": dict(sorted(seen.items())), "structs": layouts}, open(out, "w"), indent=1)
    print(f"wrote {out}: {len(seen)} foreign calls, {len(layouts)} structs")


if __name__ == "__main__":
    main()
