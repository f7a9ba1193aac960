"""The Julia binding (bernsteinbasis.jl_amd/julia/BernsteinBasisAMD.jl) against the C ABI.

No Julia runtime exists in this environment, so the binding cannot run here.  What can be
checked is what a wrong binding would get wrong first: every ``ccall((:bb_..., LIB), R,
(T1, ..., Tn), ...)`` names a function that include/bbdg.h declares, with the same number
of arguments, a Julia type that passes each C parameter type, and the right return type;
every function the header declares is bound; and the methods the reference exports
(src/BernsteinBasis.jl:19-23, src/BernsteinLift.jl:166-179,233-274,
src/ElevationMatrix.jl:11-24) have device-array methods.
"""
from __future__ import annotations

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bbdg.h")
SHIM = os.path.join(ROOT, "bernsteinbasis.jl_amd", "julia", "BernsteinBasisAMD.jl")


def _strip_c_comments(s: str) -> str:
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def header_prototypes() -> dict:
    """name -> (return type, [parameter types]) with parameter names removed"""
    text = _strip_c_comments(open(HEADER).read())
    protos = {}
    for m in re.finditer(r"((?:const\s+)?[A-Za-z_]\w*\s*\**)\s*\b(bb_\w+)\s*\(([^;{}]*?)\)\s*;", text, flags=re.S):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        ptypes = []
        for p in [x.strip() for x in params.split(",")]:
            if p in ("", "void"):
                continue
            p = re.sub(r"\s+", " ", p)
            arr = re.search(r"\[(\d*)\]$", p)
            if arr:  # "double l_j[20]" -> pointer
                p = p[:arr.start()]
            # drop the parameter name (last identifier), keep stars
            t = re.sub(r"\b\w+$", "", p).strip()
            if arr:
                t += " *"
            ptypes.append(_norm_c(t))
        protos[name] = (_norm_c(ret), ptypes)
    return protos


def _norm_c(t: str) -> str:
    t = t.replace("const", " ").replace("*", " * ")
    return " ".join(t.split())


def _split_top(s: str) -> list:
    """split a Julia type tuple body at top-level commas"""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "{(":
            depth += 1
        elif ch in "})":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def shim_ccalls() -> list:
    """[(name, return type, [argument types], number of arguments passed, line)]"""
    text = open(SHIM).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(bb_\w+),\s*LIB\)\s*,", text):
        i = m.end()
        # the call's full argument list: up to the matching ')' of "ccall("
        depth, j = 1, m.start() + len("ccall(")
        while depth:
            ch = text[j]
            depth += ch in "({"
            depth -= ch in ")}"
            j += 1
        body = text[i:j - 1]
        parts = _split_top(body)
        ret, types = parts[0], parts[1]
        assert types.startswith("(") and types.endswith(")"), (m.group(1), types)
        targs = _split_top(types[1:-1])
        calls.append((m.group(1), ret, targs, len(parts) - 2, text.count("\n", 0, m.start()) + 1))
    return calls


_OPAQUE = {"void", "bb_ctx", "bb_solver", "bb_batch"}
_SCALAR = {"int": {"Cint"}, "int64_t": {"Int64"}, "size_t": {"Csize_t"}, "double": {"Cdouble", "Float64"},
           "bb_layout": {"Layout"}}
_PTR = {"double": {"Ptr{Cdouble}", "Ref{Cdouble}"}, "int64_t": {"Ptr{Int64}", "Ref{Int64}"},
        "int32_t": {"Ptr{Int32}"}, "int8_t": {"Ptr{Int8}"}, "int": {"Ptr{Cint}", "Ref{Cint}"},
        "float": {"Ref{Cfloat}", "Ptr{Cfloat}"}, "char": {"Ptr{UInt8}", "Cstring"}}


def julia_types_for(ctype: str) -> set:
    stars = ctype.count("*")
    base = ctype.replace("*", "").strip()
    if stars == 0:
        return _SCALAR[base]
    if stars == 1:
        return {"Ptr{Cvoid}"} if base in _OPAQUE else _PTR[base]
    if stars == 2:
        if base in _OPAQUE:
            return {"Ref{Ptr{Cvoid}}"}
        return {f"Ref{{Ptr{{{t}}}}}" for t in ("Cdouble",) if base == "double"} | (
            {"Ptr{Ptr{Cdouble}}"} if base == "double" else set())
    raise AssertionError(ctype)


def test_header_parses():
    p = header_prototypes()
    assert len(p) > 50
    assert p["bb_lift_apply"][1] == ["bb_ctx *", "int", "int", "int64_t", "int", "int", "void *", "bb_layout",
                                     "void *", "bb_layout"]
    assert p["bb_rhs_combine"][1][5] == "double * *"  # const double *const d[9]
    assert p["bb_last_error"][0] == "char *"


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = shim_ccalls()
    assert len(calls) >= len(protos)
    for name, ret, targs, nargs, line in calls:
        assert name in protos, f"line {line}: {name} is not declared in bbdg.h"
        cret, cparams = protos[name]
        assert len(targs) == len(cparams), f"line {line}: {name} has {len(targs)} types, C has {len(cparams)}"
        assert nargs == len(cparams), f"line {line}: {name} passes {nargs} arguments, C takes {len(cparams)}"
        want_ret = {"int": {"Cint"}, "char *": {"Cstring"}}[cret]
        assert ret in want_ret, f"line {line}: {name} returns {ret}, C returns {cret}"
        for k, (jt, ct) in enumerate(zip(targs, cparams)):
            assert jt in julia_types_for(ct), f"line {line}: {name} argument {k + 1} is {jt}, C has {ct}"


def test_every_header_function_is_bound():
    bound = {c[0] for c in shim_ccalls()}
    missing = sorted(set(header_prototypes()) - bound)
    assert not missing, missing


def test_checker_rejects_a_wrong_binding():
    """a planted mistake (a dropped argument, a wrong integer width) fails the comparison"""
    protos = header_prototypes()
    cret, cparams = protos["bb_deriv_apply"]
    wrong_arity = ["Ptr{Cvoid}", "Cint", "Cint", "Cint", "Int64", "Cint", "Ptr{Cvoid}", "Layout", "Ptr{Cvoid}"]
    assert len(wrong_arity) != len(cparams)
    wrong_width = ["Ptr{Cvoid}", "Cint", "Cint", "Cint", "Cint", "Cint", "Ptr{Cvoid}", "Layout", "Ptr{Cvoid}", "Layout"]
    assert any(jt not in julia_types_for(ct) for jt, ct in zip(wrong_width, cparams))


def test_reference_methods_have_device_methods():
    text = open(SHIM).read()
    for pat in (r"function LinearAlgebra\.mul!\(out::ROCArray, D::DerivOp, x::ROCArray\)",
                r"LinearAlgebra\.mul!\(out::ROCArray, L::BernsteinLift, x::ROCArray\)",
                r"threaded_mul!\(out::ROCArray, L::MultithreadedBernsteinLift, x::ROCArray, thread\)",
                r"function face_lift_multiply!\(out::ROCArray, L::BernsteinLift, x::ROCArray\)",
                r"function Base\.:\*\(L::BernsteinLift, x::ROCMatrix",
                r"function LinearAlgebra\.mul!\(out::ROCArray, ::ElevationMatrix\{N\}, x::ROCArray\)",
                r"function Base\.:\*\(E::ElevationMatrix\{N\}, x::ROCArray"):
        assert re.search(pat, text), pat


@pytest.mark.parametrize("const", ["BB_F64", "BB_F32", "BB_AXIS_R", "BB_AXIS_S", "BB_AXIS_T", "BB_PDE_WAVE",
                                   "BB_PDE_MAXWELL", "BB_FLUX_REFERENCE", "BB_FLUX_UPWIND", "BB_LIFT_PLAIN",
                                   "BB_LIFT_ACC64", "BB_LIFT_TWOSUM", "BB_LIFT_ACC64_MFMA", "BB_PREC_F64", "BB_PREC_MIXED",
                                   "BB_PATH_PLAIN", "BB_PATH_FUSED"])
def test_enum_values_match(const):
    text = _strip_c_comments(open(HEADER).read())
    enums = {}
    for body in re.findall(r"enum\s+\w+\s*\{([^}]*)\}", text):
        for k, v in re.findall(r"(BB_\w+)\s*=\s*(\d+)", body):
            enums[k] = int(v)
    m = re.search(rf"const {const} = Cint\((\d+)\)", open(SHIM).read())
    assert m and int(m.group(1)) == enums[const]
