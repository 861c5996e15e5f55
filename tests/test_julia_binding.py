"""The Julia front end (julia/LowRankIntegratorsHIP.jl) against the C ABI, without Julia.

Julia is not installed here, so the module is checked statically: every
``ccall((:name, libhip), Ret, (T1, ...), args...)`` must name a function that
include/lri.h declares, with as many argument types as the C declaration has
parameters, each Julia type compatible with the C type (Cint for int, Int64 for
int64_t, Cdouble/Float64 for double, Ptr/Ref for pointers, Cstring for a
returned const char*), and as many arguments as types.  The module must also
define the methods through which the reference's unmodified ``solve``
(primitives.jl:68-80) reaches the engine: ``init``, ``alg_cache``,
``step!(integrator, alg, dt)`` and ``update_sol!`` for ``HIPAlgorithm``.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lri.h")
JULIA = os.path.join(ROOT, "julia", "LowRankIntegratorsHIP.jl")


def split_top(s):
    """Split s at commas that are not inside (), {} or []."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    tail = "".join(cur).strip()
    if tail:
        out.append(tail)
    return out


def c_declarations():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"#.*", "", text)
    decls = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(lri_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        ret, name, params = " ".join(m.group(1).split()), m.group(2), " ".join(m.group(3).split())
        ps = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
        # drop the parameter names, keep the types
        types = [re.sub(r"\s*\b\w+$", "", p).strip() if not p.endswith("*") else p for p in ps]
        decls[name] = (ret, types)
    return decls


def julia_ccalls():
    src = open(JULIA).read()
    src = re.sub(r"#[^\n]*", "", src)  # line comments
    calls = []
    for m in re.finditer(r"ccall\(", src):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        parts = split_top(src[m.end():i - 1])
        calls.append(parts)
    return calls


def compatible(ctype: str, jtype: str, ret=False) -> bool:
    c = ctype.replace("const ", "").replace(" ", "")
    j = jtype.replace(" ", "")
    if c == "int":
        return j == "Cint"
    if c == "int64_t":
        return j == "Int64"
    if c == "double":
        return j in ("Cdouble", "Float64")
    if ret and c == "char*":
        return j == "Cstring"
    if c in ("lri_ctx**", "lri_batch**"):
        return j == "Ref{Ptr{Cvoid}}"
    if c in ("lri_ctx*", "lri_batch*", "void*"):
        return j == "Ptr{Cvoid}"
    if c == "void*const*":
        return j in ("Ptr{Ptr{Cvoid}}", "Ref{Ptr{Cvoid}}")
    if c == "double*":
        return j in ("Ptr{Float64}", "Ref{Float64}", "Ptr{Cdouble}")
    if c == "int64_t*":
        return j in ("Ptr{Int64}", "Ref{Int64}")
    if c == "int*":
        return j in ("Ptr{Cint}", "Ref{Cint}")
    if c == "float*":
        return j in ("Ptr{Float32}", "Ref{Float32}")
    return False


def test_header_parser_sees_every_declaration():
    decls = c_declarations()
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    assert set(decls) == set(re.findall(r"\b(lri_[a-z_0-9]+)\s*\(", text))
    assert decls["lri_step_ksl"] == ("int", ["lri_ctx*", "int", "int64_t", "int64_t"])
    assert decls["lri_last_error"] == ("const char*", ["const lri_ctx*"])


def test_every_ccall_matches_the_header():
    decls = c_declarations()
    calls = julia_ccalls()
    assert len(calls) >= 30
    seen = set()
    for parts in calls:
        target, ret, types, args = parts[0], parts[1], parts[2], parts[3:]
        m = re.fullmatch(r"\(:(lri_\w+),\s*libhip\)", target)
        assert m, f"ccall target {target!r} is not (:lri_name, libhip)"
        name = m.group(1)
        assert name in decls, f"{name} is not declared in include/lri.h"
        cret, cparams = decls[name]
        assert types.startswith("(") and types.endswith(")"), f"{name}: argument types {types!r} are not a tuple"
        jtypes = split_top(types[1:-1])
        assert len(jtypes) == len(cparams), f"{name}: {len(jtypes)} Julia argument types, {len(cparams)} in C"
        assert compatible(cret, ret, ret=True), f"{name}: return {ret} for C {cret}"
        for k, (ct, jt) in enumerate(zip(cparams, jtypes)):
            assert compatible(ct, jt), f"{name}: argument {k + 1} is {jt} for C {ct}"
        assert len(args) == len(jtypes), f"{name}: {len(args)} arguments for {len(jtypes)} types"
        seen.add(name)
    # the entry points that replace the reference's step bodies are all bound
    for required in ("lri_create", "lri_destroy", "lri_set_state", "lri_get_state", "lri_snapshots_reserve",
                     "lri_snapshot_upload", "lri_step_ksl", "lri_step_ksl_strang", "lri_step_bug", "lri_step_bug_ra",
                     "lri_step_greedy", "lri_step_greedy_twofactor", "lri_set_operator_lyapunov",
                     "lri_set_operator_quadratic", "lri_step_de_ksl", "lri_step_de_bug", "lri_step_de_bug_ra",
                     "lri_step_bug_adjoint", "lri_step_ksl_adjoint", "lri_step_ksl_dual_adjoint", "lri_step_bug_ra_adjoint",
                     "lri_step_de_bug_adjoint", "lri_step_de_bug_ra_adjoint", "lri_step_de_ksl_adjoint", "lri_trajectory_reserve",
                     "lri_trajectory_save", "lri_trajectory_get", "lri_trajectory_get_range", "lri_last_error"):
        assert required in seen, f"{required} is not called from the Julia module"


def test_solve_reaches_the_engine():
    """The methods the reference's solve loop dispatches to, for the wrapper type."""
    src = open(JULIA).read()
    assert re.search(r"struct HIPAlgorithm\{A<:AbstractDLRAlgorithm\} <: AbstractDLRAlgorithm", src)
    assert re.search(r"import LowRankIntegrators: .*\binit\b.*\bstep!.*\balg_cache\b.*\bupdate_sol!", src)
    assert re.search(r"^function init\(prob::Union\{MatrixDataProblem,MatrixDEProblem\}, alg::HIPAlgorithm, dt\)",
                     src, flags=re.M)
    assert re.search(r"^function alg_cache\(prob::MatrixDataProblem, alg::HIPAlgorithm, u, dt; t0", src, flags=re.M)
    assert re.search(r"^function alg_cache\(prob::MatrixDEProblem, alg::HIPAlgorithm, u, dt; t0", src, flags=re.M)
    # three positional arguments, as solve calls it: step!(integrator, alg, dt)
    assert re.search(r"^function step!\(integrator::DLRIntegrator, alg::HIPAlgorithm, dt\)", src, flags=re.M)
    assert re.search(r"^function update_sol!\(integrator::DLRIntegrator\{<:Any,<:Any,<:HIPAlgorithm\}\)", src,
                     flags=re.M)
    assert re.search(r"function ChainRulesCore\.rrule\(::typeof\(final_state\), prob::MatrixDataProblem, "
                     r"alg::HIPAlgorithm", src)
    # balanced block structure (a cheap syntax check: every block opener has its `end`)
    body = re.sub(r'"""(.*?)"""', "", src, flags=re.S)
    body = re.sub(r'"(\\.|[^"\\])*"', '""', body)
    body = re.sub(r"#[^\n]*", "", body)
    body = re.sub(r"\[[^\[\]]*\bend\b[^\[\]]*\]", "[]", body)  # `end` used as an index
    openers = len(re.findall(r"^\s*(?:mutable struct|struct|function|module|if|for|while|let|begin|try)\b",
                             body, flags=re.M))
    openers += len(re.findall(r"=\s*(?:if|begin|let)\b", body))
    ends = len(re.findall(r"\bend\b", body))
    assert openers == ends, f"{openers} block openers, {ends} `end`s"


@pytest.mark.parametrize("name", ["lri_step_ksl", "lri_step_bug_ra", "lri_step_de_bug_adjoint"])
def test_compatibility_rules_reject_mismatches(name):
    """A mistyped binding must fail the checks above (planted on the parsed declaration)."""
    ret, params = c_declarations()[name]
    assert not compatible(params[-1], "Cint" if params[-1] != "int" else "Int64")
