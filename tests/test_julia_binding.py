"""The Julia binding (sparsedirectsolver.jl_amd/julia/SparseDirectSolverMI355XExt.jl) cannot run
here (no Julia toolchain), so its `ccall`s are checked against include/sds_hip.h as text: every
called symbol is declared, the argument-type tuple and the arguments passed have the header's
parameter count, each type has the header's kind (pointer, Int32, Int64, Float64), the return
type is Cint (Cstring for sds_last_error), and the entry points a drop-in needs are all bound."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "sparsedirectsolver.jl_amd", "julia", "SparseDirectSolverMI355XExt.jl")
HDR = os.path.join(ROOT, "include", "sds_hip.h")


def _split_top(s):
    """Split on the commas of nesting depth 0 (parentheses and braces)."""
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


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    protos = {}
    for ret, name, params in re.findall(r"^(int|const char\*)\s+(sds_\w+)\(([^;]*)\);", text, flags=re.M | re.S):
        ps = [p.strip() for p in params.split(",") if p.strip() and p.strip() != "void"]
        protos[name] = (ret, ps)
    return protos


def _kind_c(param):
    if "*" in param or param.startswith("sds_handle_t"):
        return "ptr"
    t = param.rsplit(" ", 1)[0]
    return {"int32_t": "i32", "int64_t": "i64", "double": "f64"}[t]


def _kind_jl(t):
    t = t.strip()
    if t.startswith(("Ptr{", "Ref{")) or t == "Cstring":
        return "ptr"
    return {"Cint": "i32", "Int32": "i32", "Int64": "i64", "Float64": "f64"}[t]


def julia_ccalls():
    src = open(JL).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(sds_\w+),", src):
        i = m.start() + len("ccall")
        depth = 0
        for j in range(i, len(src)):
            depth += src[j] == "("
            depth -= src[j] == ")"
            if depth == 0:
                break
        parts = _split_top(src[i + 1:j])
        ret, types = parts[1], parts[2]
        assert types.startswith("(") and types.endswith(")"), f"{m.group(1)}: argument types are not a tuple"
        calls.append((m.group(1), ret, _split_top(types[1:-1]), parts[3:]))
    return calls


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 15
    for name, ret, types, args in calls:
        assert name in protos, f"{name} is not declared in include/sds_hip.h"
        cret, params = protos[name]
        assert ret == ("Cstring" if cret == "const char*" else "Cint"), f"{name}: return type {ret}"
        assert len(types) == len(params), f"{name}: {len(types)} argument types, the header has {len(params)}"
        assert len(args) == len(params), f"{name}: {len(args)} arguments, the header has {len(params)}"
        for k, (t, p) in enumerate(zip(types, params)):
            assert _kind_jl(t) == _kind_c(p), f"{name} argument {k + 1}: Julia {t} for C {p}"


REQUIRED = ("sds_create", "sds_destroy", "sds_last_error", "sds_set_stream", "sds_upload_symbolic",
            "sds_alloc_numeric", "sds_factorize", "sds_solve", "sds_set_ir_mode", "sds_autotune", "sds_graph_capture",
            "sds_graph_launch", "sds_graph_info", "sds_graph_destroy", "sds_adjoint_scatter", "sds_get_stats",
            "sds_get_pivots", "sds_get_diag", "sds_memory_estimates")


@pytest.mark.parametrize("name", REQUIRED)
def test_binding_covers_entry_point(name):
    assert name in {c[0] for c in julia_ccalls()}


def test_binding_uses_the_reference_names_it_needs():
    """IR_FGMRES and InterruptedError are imported from SparseDirectSolver (neither is defined in
    the extension), host pointers are taken inside GC.@preserve, and the interrupt flag is passed."""
    src = open(JL).read()
    using = re.search(r"using SparseDirectSolver:([^\n]*(?:\n\s+[^\n]*)*)", src).group(1)
    for name in ("IR_FGMRES", "InterruptedError", "PHASE_SOLVE"):
        assert re.search(rf"\b{name}\b", using), name
    assert "pointer_from_objref(flag)" in src and "user_host_interrupt" in src
    for fn in ("factorize!", "autotune!", "graph_capture!"):
        body = src[src.index(f"function {fn}("):]
        body = body[:body.index("\nend\n")]
        assert "GC.@preserve opts" in body, fn
    assert "ChainRulesCore.rrule" in src and "solve_mode = 1" in src and "adjoint_scatter!" in src
