"""The Julia binding (julia/DNSROCm.jl) against the C ABI it calls.

Julia is not installed here, so the binding cannot run; this test checks
what can be checked from the text: every `ccall` names a function that
include/dns_rocm.h declares, with the same number of arguments and
compatible argument and return types, the parameter structs match the C
structs field by field, and the binding defines the methods through which
the reference's unchanged `solve!` (src/timestep.jl:92-111) and
`spectral_solve!` (src/spectral/step.jl:118-170) reach the device."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "julia", "DNSROCm.jl")
HDR = os.path.join(ROOT, "include", "dns_rocm.h")

# C parameter type -> Julia types that pass it
C_TO_JL = {
    "dns_fv *": {"Ptr{Cvoid}"},
    "dns_spec *": {"Ptr{Cvoid}"},
    "dns_cnn *": {"Ptr{Cvoid}"},
    "dns_fv **": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "dns_spec **": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "const dns_fv_params *": {"Ref{FVParams}"},
    "const dns_spec_params *": {"Ref{SpecParams}"},
    "double": {"Cdouble", "Float64"},
    "float": {"Cfloat", "Float32"},
    "int32_t": {"Int32", "Cint"},
    "int64_t": {"Int64"},
    "uint64_t": {"UInt64"},
    "double *": {"Ptr{Cdouble}", "Ref{Cdouble}", "Ptr{ComplexF64}", "Ptr{Float64}"},
    "const double *": {"Ptr{Cdouble}", "Ref{Cdouble}", "Ptr{ComplexF64}", "Ptr{Float64}"},
    "int32_t *": {"Ref{Int32}", "Ptr{Int32}"},
    "int64_t *": {"Ref{Int64}", "Ptr{Int64}"},
    "char *": {"Ptr{UInt8}"},
    "const char *": {"Cstring", "Ptr{UInt8}"},
}
RET = {"int": "Cint", "const char *": "Cstring"}


def _norm(t):
    t = re.sub(r"\s+", " ", t.strip())
    return re.sub(r"\s*\*", " *", t).replace("* *", "**")


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"^\s*((?:const\s+)?\w+\s*\*?)\s*(dns_\w+)\s*\(([^)]*)\)\s*;", text, re.M):
        ret, name, args = _norm(m.group(1)), m.group(2), m.group(3).strip()
        types = []
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                a = re.sub(r"\w+$", "", a)  # drop the parameter name
                types.append(_norm(a))
        protos[name] = (ret, types)
    return protos


def split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def julia_ccalls():
    text = open(JL).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*libdns\),\s*([\w{}]+),\s*\(([^()]*)\)", text, re.S):
        calls.append((m.group(1), m.group(2), split_top(m.group(3))))
    return text, calls


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    _, calls = julia_ccalls()
    assert len(calls) >= 30
    for name, ret, args in calls:
        assert name in protos, f"{name} is not declared in include/dns_rocm.h"
        cret, cargs = protos[name]
        assert RET[cret] == ret, (name, cret, ret)
        assert len(args) == len(cargs), (name, args, cargs)
        for ja, ca in zip(args, cargs):
            assert ca in C_TO_JL, (name, ca)
            assert ja in C_TO_JL[ca], (name, ca, ja)


def _c_struct_fields(name):
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\}" % name, text, re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(const\s+)?(\w+)\s*(\*?)\s*(\w+)(\[(\d+)\])?$", decl)
        fields.append((m.group(4), m.group(2) + (" *" if m.group(3) else ""), int(m.group(6) or 1)))
    return fields


def _jl_struct_fields(name):
    text = open(JL).read()
    body = re.search(r"struct %s\n(.*?)\nend" % name, text, re.S).group(1)
    return [tuple(x.strip() for x in line.split("::")) for line in body.strip().splitlines()]


def test_parameter_structs_match():
    jl_of = {("int32_t", 1): "Int32", ("double", 1): "Float64", ("double *", 1): "Ptr{Float64}"}
    for cname, jname in (("dns_fv_params", "FVParams"), ("dns_spec_params", "SpecParams")):
        cf, jf = _c_struct_fields(cname), _jl_struct_fields(jname)
        assert [f[0] for f in cf] == [f[0] for f in jf], (cname, cf, jf)
        for (fname, ctype, count), (_, jtype) in zip(cf, jf):
            base = "Int32" if ctype == "int32_t" else "Float64" if ctype == "double" else "Ptr{Float64}"
            want = base if count == 1 else f"NTuple{{{count},{base}}}"
            assert jtype == want or jtype == jl_of.get((ctype, count)), (cname, fname, ctype, jtype)


def test_binding_defines_the_methods_the_reference_loops_call():
    """solve! calls scalarfield, vectorfield, exchange_halo!, project!,
    cfl_timestep and step! on the setup and fields (src/timestep.jl:92-111);
    spectral_solve! calls specvelocity, spectral_project!, spec_to_phys!,
    vsummax and spectral_step!, and shellforcing calls shell_energies and
    rescale_shell! (src/spectral/step.jl:118-170, src/spectral/forcing.jl:47-56);
    channelstats/spectrumstats samples and the diagnostics complete the API."""
    text, _ = julia_ccalls()
    defined = set(re.findall(r"(?:function |^)DNS\.(\w+!?)\(", text, re.M))
    needed = {"scalarfield", "vectorfield", "velocityfield!", "exchange_halo!", "project!", "cfl_timestep", "step!",
              "maxdiv", "kinetic_energy", "channelstats", "sample_channelstats!", "spectrumstats",
              "sample_spectrum!", "specvelocity", "spectral_project!", "spec_to_phys!", "phys_to_spec!", "vsummax",
              "spectral_rhs!", "spectral_step!", "shell_energies", "rescale_shell!", "spectral_energy",
              "spectral_dissipation", "spectral_maxdiv", "spectral_velocityfield!", "spectral_randomfield!"}
    assert needed <= defined, needed - defined
    # keyword-only entry points of the reference are not redefined
    assert not re.search(r"function DNS\.(setup|spectral_setup|solve!|spectral_solve!)\(", text)
