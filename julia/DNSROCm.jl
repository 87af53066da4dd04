"""
    DNSROCm

MI355X backend of DistributedNavierStokes.jl on one GPU.  `DNSROCm.setup`
and `DNSROCm.spectral_setup` take the keywords of the reference's `setup`
(src/setup.jl:34-48) and `spectral_setup` (src/spectral/setup.jl:29-40) and
return a `ROCmFVSetup` / `ROCmSpecSetup` that owns a `libdnsrocm.so` handle
(include/dns_rocm.h).  The package's generic functions get methods on these
types, each forwarding to the C ABI with `ccall`, so the reference's own,
unchanged host loops drive the GPU:

    using DistributedNavierStokes, DNSROCm
    s = DNSROCm.setup(; n = (256, 192, 256), lims, bc, stretch, visc, bodyforce)
    u = vectorfield(s)
    velocityfield!(u, s; x = (x, y, z) -> 1 - y^2)
    cs = channelstats(s)
    state = solve!(; u, setup = s, tlims = (0.0, 1.0), cfl = 0.5, processors = (; cs = cs.sample))
    prof = channelprofiles(cs, s)

Keyword-only entry points (`setup`, `spectral_setup`, `solve!`,
`spectral_solve!`) are not redefined: `setup`/`spectral_setup` here are
functions of this module, and `solve!`/`spectral_solve!` are the reference's
methods, which reach the device through the methods below (`scalarfield`,
`vectorfield`, `exchange_halo!`, `project!`, `cfl_timestep`, `step!`,
`specvelocity`, `spectral_project!`, `spec_to_phys!`, `vsummax`,
`spectral_step!`, `shell_energies`, `rescale_shell!`, ...).

A handle holds one velocity state: every `vectorfield(s)` / `specvelocity(s)`
names it (the RK registers and the pressure are the handle's own, so the
`q`, `p`, `δ` and `cache` arguments of `step!`/`spectral_step!` are
placeholders).  `Array(u.x)` downloads a ghosted component;
`convert(Array{Complex{T},4}, uh)` and `copyto!(uh, A)` move the spectral
state, which is what `spectral_save`/`spectral_load!` (src/spectral/io.jl)
use, so they work unchanged.  One GPU: `procgrid` must be `(1, 1)`.
"""
module DNSROCm

using DistributedNavierStokes
const DNS = DistributedNavierStokes
using KernelAbstractions: CPU
using MPI: MPI

const libdns = get(ENV, "DNS_ROCM_LIB",
    normpath(joinpath(@__DIR__, "..", "distributednavierstokes_amd", "libdnsrocm.so")))

const DNS_OK = Cint(0)
const DNS_FP64 = Int32(0)
const DNS_MIXED = Int32(1)
const DNS_PLAIN_FP64 = Int32(2)
const LAYOUT_INTERIOR = Int32(0)
const LAYOUT_GHOSTED = Int32(1)
const FIELD_UX, FIELD_UY, FIELD_UZ, FIELD_P = Int32(0), Int32(1), Int32(2), Int32(3)
const FIELD_DELTA = Int32(7)

"Marker backend: `backend = ROCmNative()` selects this library."
struct ROCmNative end

"Throw the library's message for a nonzero status."
function check(rc::Cint)
    rc == DNS_OK && return nothing
    error(unsafe_string(ccall((:dns_last_error, libdns), Cstring, ())))
end

# ------------------------------------------------------------------ FV

"`dns_fv_params` (include/dns_rocm.h); same field order and C layout."
struct FVParams
    n::NTuple{3,Int32}
    lims::NTuple{6,Float64}
    ywall::Int32
    yfaces::Ptr{Float64}
    visc::Float64
    bodyforce::NTuple{3,Float64}
    ydiffusion::Int32
    precision::Int32
end

"""
Setup of the finite-volume solver on the GPU: the fields of the reference's
setup NamedTuple that its generic host code reads (`n`, `lims`, `bc`, `visc`,
`w`, `T`, `topo`, `lay`, `grid`, `ranges`, `imexy`, ...) plus the handle.
`T` is the host element type of diagnostics (Float64); `precision` is the
device form (:fp64, :mixed or :plain).
"""
mutable struct ROCmFVSetup
    handle::Ptr{Cvoid}
    n::NTuple{3,Int}
    lims::Any
    bc::NTuple{3,Symbol}
    visc::Float64
    bodyforce::NTuple{3,Float64}
    order::Int
    w::Int
    T::DataType
    backend::ROCmNative
    topo::Any
    axes::NTuple{3,Int}
    lay::Any
    grid::Any
    ranges::Any
    imexy::Bool
    εy::Float64
    precision::Symbol
end

"""
    setup(; n, lims, bc, stretch, visc, bodyforce, procgrid, ydiffusion,
          mpibuf, order, backend = ROCmNative(), T = Float64, comm,
          precision = T == Float32 ? :mixed : :fp64)

The reference's `setup` (src/setup.jl:34-113) on one GPU.  The `stretch`
closure is evaluated here, on the host, exactly as `griddata` does
(src/setup.jl:155); the library builds the device metrics from those faces.
Any other `backend` falls through to the reference's `setup`.
"""
function setup(;
    n,
    lims = ((0.0, 2π), (0.0, 2π), (0.0, 2π)),
    bc = (:periodic, :periodic, :periodic),
    stretch = identity,
    visc = 1e-3,
    bodyforce = (0.0, 0.0, 0.0),
    procgrid = nothing,
    ydiffusion = :explicit,
    mpibuf = :auto,
    order = 2,
    backend = ROCmNative(),
    T = Float64,
    comm = MPI.COMM_WORLD,
    precision = T == Float32 ? :mixed : :fp64,
)
    backend isa ROCmNative || return DNS.setup(;
        n, lims, bc, stretch, visc, bodyforce, procgrid, ydiffusion, mpibuf, order, backend, T, comm)
    bc[1] == bc[3] == :periodic || error("x and z must be periodic")
    bc[2] in (:periodic, :wall) || error("bc[2] must be :periodic or :wall")
    ydiffusion in (:explicit, :implicit) || error("ydiffusion must be :explicit or :implicit")
    order == 2 || error("only order 2 is implemented so far")
    something(procgrid, (1, 1)) == (1, 1) || error("the ROCm backend runs on one GPU: procgrid must be (1, 1)")
    precision in (:fp64, :mixed, :plain) || error("precision must be :fp64, :mixed or :plain")
    MPI.Initialized() || MPI.Init()
    w = order ÷ 2
    axes = DNS.AXES.x
    periodic = ntuple(a -> bc[findfirst(==(a), axes)] == :periodic, 2)
    topo = DNS.topology(MPI.COMM_SELF, (1, 1), periodic)
    lay = DNS.layout(n, axes, (1, 1), topo.coords)
    grid = DNS.griddata(n, lims, bc, stretch, lay, w, CPU(), Float64)
    ranges = (;
        x = DNS.activeranges(lay, n, bc, w, 1),
        y = DNS.activeranges(lay, n, bc, w, 2),
        z = DNS.activeranges(lay, n, bc, w, 3),
        p = DNS.activeranges(lay, n, bc, w, 0),
    )
    yfaces = collect(Float64, grid.yfg)   # y faces 0:ny
    prec = precision == :mixed ? DNS_MIXED : precision == :plain ? DNS_PLAIN_FP64 : DNS_FP64
    params = FVParams(
        Int32.(Tuple(n)),
        Float64.((lims[1]..., lims[2]..., lims[3]...)),
        Int32(bc[2] == :wall),
        pointer(yfaces),
        Float64(visc),
        Float64.(Tuple(bodyforce)),
        Int32(ydiffusion == :implicit),
        prec,
    )
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve yfaces check(ccall((:dns_fv_create, libdns), Cint, (Ref{FVParams}, Ref{Ptr{Cvoid}}), params, h))
    imexy = ydiffusion == :implicit
    s = ROCmFVSetup(h[], Tuple(n), lims, Tuple(bc), Float64(visc), Float64.(Tuple(bodyforce)), order, w, Float64,
        ROCmNative(), topo, axes, lay, grid, ranges, imexy, imexy ? 0.0 : 1.0, precision)
    finalizer(s) do s
        s.handle == C_NULL || ccall((:dns_fv_destroy, libdns), Cint, (Ptr{Cvoid},), s.handle)
        s.handle = C_NULL
    end
    s
end

"A velocity component of the handle's state (ghosted host views on demand)."
struct ROCmComponent
    s::ROCmFVSetup
    id::Int32
end

"The handle's velocity state, `(; x, y, z)` by property."
struct ROCmVelocity
    s::ROCmFVSetup
end

function Base.getproperty(u::ROCmVelocity, c::Symbol)
    s = getfield(u, :s)
    c === :x ? ROCmComponent(s, FIELD_UX) : c === :y ? ROCmComponent(s, FIELD_UY) :
    c === :z ? ROCmComponent(s, FIELD_UZ) : c === :s ? s : error("no component $c")
end
Base.getindex(u::ROCmVelocity, c::Symbol) = getproperty(u, c)

"""
A scalar field of the handle: the pressure p or the projection increment δ.
`solve!` allocates p (and δ with implicit y-diffusion) before the initial
projection; the field it names is bound by the first `project!`/`step!`.
"""
mutable struct ROCmScalar
    s::ROCmFVSetup
    id::Int32
end

function setfield(s::ROCmFVSetup, id::Int32, a::Array{Float64,3}, layout::Int32)
    check(ccall((:dns_fv_set_field, libdns), Cint, (Ptr{Cvoid}, Int32, Ptr{Cdouble}, Int32), s.handle, id, a, layout))
end

function getfield_host(s::ROCmFVSetup, id::Int32)
    a = zeros(Float64, s.n .+ 2s.w)
    check(ccall((:dns_fv_get_field, libdns), Cint, (Ptr{Cvoid}, Int32, Ptr{Cdouble}, Int32), s.handle, id, a,
        LAYOUT_GHOSTED))
    a
end

"Ghosted host copy of a component, halos filled as exchange_halo! would."
Base.Array(f::ROCmComponent) = getfield_host(f.s, f.id)
Base.Array(φ::ROCmScalar) = getfield_host(φ.s, φ.id < 0 ? FIELD_P : φ.id)
Base.copyto!(f::ROCmComponent, a::AbstractArray) = (setfield(f.s, f.id, Array{Float64,3}(a), LAYOUT_GHOSTED); f)

"scalarfield(setup) (src/fields.jl:4-5): zeroes the handle's p (and δ) like a fresh allocation."
function DNS.scalarfield(s::ROCmFVSetup)
    z = zeros(Float64, s.n)
    setfield(s, FIELD_P, z, LAYOUT_INTERIOR)
    s.imexy && setfield(s, FIELD_DELTA, z, LAYOUT_INTERIOR)
    ROCmScalar(s, Int32(-1))
end

"vectorfield(setup) (src/fields.jl:7-9): the handle's velocity state."
DNS.vectorfield(s::ROCmFVSetup) = ROCmVelocity(s)

"velocityfield!(u, setup; x, y, z) (src/fields.jl:28-35): host evaluation at the staggered points."
function DNS.velocityfield!(u::ROCmVelocity, s::ROCmFVSetup; x = nothing, y = nothing, z = nothing)
    g, w = s.grid, s.w
    for (id, f, co, rng) in ((FIELD_UX, x, (g.xf, g.yc, g.zc), s.ranges.x),
                             (FIELD_UY, y, (g.xc, g.yf, g.zc), s.ranges.y),
                             (FIELD_UZ, z, (g.xc, g.yc, g.zf), s.ranges.z))
        isnothing(f) && continue
        h = zeros(Float64, s.n .+ 2w)
        for k in rng[3], j in rng[2], i in rng[1]
            h[i, j, k] = f(co[1][i-w], co[2][j-w], co[3][k-w])
        end
        setfield(s, id, h, LAYOUT_GHOSTED)
    end
    u
end

"Ghosts are index rules inside the kernels: nothing to exchange (src/halo.jl:129-133)."
DNS.exchange_halo!(u::ROCmVelocity, s::ROCmFVSetup) = u
DNS.exchange_halo!(φ::ROCmScalar, stag::Symbol, s::ROCmFVSetup) = φ

"project!(u, ψ, setup) (src/timestep.jl:25-31)."
function DNS.project!(u::ROCmVelocity, ψ::ROCmScalar, s::ROCmFVSetup)
    ψ.id < 0 && (ψ.id = s.imexy ? FIELD_DELTA : FIELD_P)
    check(ccall((:dns_fv_project, libdns), Cint, (Ptr{Cvoid}, Int32), s.handle, ψ.id))
    u
end

"step!(u, p, q, δ, setup, Δt) (src/timestep.jl:41-63); the handle owns p, q and δ."
function DNS.step!(u::ROCmVelocity, p::ROCmScalar, q::ROCmVelocity, δ::ROCmScalar, s::ROCmFVSetup, Δt)
    p.id = FIELD_P
    s.imexy && (δ.id = FIELD_DELTA)
    check(ccall((:dns_fv_step, libdns), Cint, (Ptr{Cvoid}, Cdouble), s.handle, Float64(Δt)))
    u
end

"cfl_timestep(u, setup; cfl) (src/timestep.jl:69-80)."
function DNS.cfl_timestep(u::ROCmVelocity, s::ROCmFVSetup; cfl)
    dt = Ref{Cdouble}(0.0)
    check(ccall((:dns_fv_cfl_timestep, libdns), Cint, (Ptr{Cvoid}, Cdouble, Ref{Cdouble}), s.handle, Float64(cfl), dt))
    dt[]
end

"maxdiv(u, setup) (src/fields.jl:38-42)."
function DNS.maxdiv(u::ROCmVelocity, s::ROCmFVSetup)
    out = Ref{Cdouble}(0.0)
    check(ccall((:dns_fv_maxdiv, libdns), Cint, (Ptr{Cvoid}, Ref{Cdouble}), s.handle, out))
    out[]
end

"kinetic_energy(u, setup) (src/fields.jl:45-60)."
function DNS.kinetic_energy(u::ROCmVelocity, s::ROCmFVSetup)
    out = Ref{Cdouble}(0.0)
    check(ccall((:dns_fv_kinetic_energy, libdns), Cint, (Ptr{Cvoid}, Ref{Cdouble}), s.handle, out))
    out[]
end

"""
    advance_cfl!(s, cfl, nsteps, tend, t) -> (t, Δts)

At most `nsteps` CFL steps of `solve!`'s loop (src/timestep.jl:102-107)
with Δt computed on the device (no host round trip per step in the explicit
form); stops at `t >= tend - 1e-10(|tend| + 1)`.  Use it between processor
calls when processors run every k steps.
"""
function advance_cfl!(s::ROCmFVSetup, cfl, nsteps, tend, t)
    tt = Ref{Cdouble}(Float64(t))
    dts = zeros(Float64, max(nsteps, 1))
    nd = Ref{Int32}(0)
    check(ccall((:dns_fv_advance_cfl, libdns), Cint,
        (Ptr{Cvoid}, Cdouble, Int32, Cdouble, Ref{Cdouble}, Ptr{Cdouble}, Ref{Int32}),
        s.handle, Float64(cfl), Int32(nsteps), Float64(tend), tt, dts, nd))
    tt[], dts[1:nd[]]
end

"channelstats(setup; nupdate) (src/stats.jl:15-37): the same accumulator; samples come from the GPU."
function DNS.channelstats(s::ROCmFVSetup; nupdate = 1)
    ny = s.n[2]
    nf = s.bc[2] == :wall ? ny - 1 : ny
    acc = (;
        nsamples = Ref(0),
        Uc = zeros(Float64, ny), Wc = zeros(Float64, ny), uuc = zeros(Float64, ny), wwc = zeros(Float64, ny),
        Vf = zeros(Float64, nf), vvf = zeros(Float64, nf), Uf = zeros(Float64, nf), uvf = zeros(Float64, nf),
    )
    sample = (state, s2) -> begin
        state.n % nupdate == 0 && DNS.sample_channelstats!(acc, state.u, s2)
        nothing
    end
    (; acc..., sample)
end

"sample_channelstats!(acc, u, s) (src/stats.jl:39-58) on the device (compensated plane sums)."
function DNS.sample_channelstats!(acc, u::ROCmVelocity, s::ROCmFVSetup)
    ny = length(acc.Uc)
    nf = length(acc.Vf)
    buf = zeros(Float64, 4ny + 4nf)
    check(ccall((:dns_fv_channelstats_sample, libdns), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), s.handle, buf))
    for (k, v) in enumerate((acc.Uc, acc.Wc, acc.uuc, acc.wwc))
        v .+= view(buf, (k-1)*ny+1:k*ny)
    end
    for (k, v) in enumerate((acc.Vf, acc.vvf, acc.Uf, acc.uvf))
        v .+= view(buf, 4ny+(k-1)*nf+1:4ny+k*nf)
    end
    acc.nsamples[] += 1
    nothing
end

"spectrumstats(setup; nupdate) (src/stats.jl:108-118)."
function DNS.spectrumstats(s::ROCmFVSetup; nupdate = 1)
    s.bc == (:periodic, :periodic, :periodic) || error("spectra require a fully periodic box")
    nbins = ceil(Int, sqrt(sum(abs2, s.n .÷ 2))) + 1
    acc = (; nsamples = Ref(0), E = zeros(Float64, nbins))
    sample = (state, s2) -> begin
        state.n % nupdate == 0 && DNS.sample_spectrum!(acc, state.u, s2)
        nothing
    end
    (; acc..., sample)
end

"sample_spectrum!(acc, u, s) (src/stats.jl:120-135) on the device."
function DNS.sample_spectrum!(acc, u::ROCmVelocity, s::ROCmFVSetup)
    check(ccall((:dns_fv_spectrum_sample, libdns), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Int32), s.handle, acc.E,
        Int32(length(acc.E))))
    acc.nsamples[] += 1
    nothing
end

# ------------------------------------------------------------ spectral

"`dns_spec_params` (include/dns_rocm.h)."
struct SpecParams
    n::Int32
    kcut::Int32
    l::Float64
    visc::Float64
    precision::Int32
end

"The physical buffer `s.fft.v` of a spectral handle."
struct ROCmPhys
    handle::Ptr{Cvoid}
end

"""
Setup of the pseudo-spectral solver on the GPU with the fields the
reference's generic host code reads (`n`, `l`, `visc`, `kcut`, `m`, `T`,
`topo`, `lspec`, `lphys`, `fft.v`).  `T = Float32` is the mixed form (FP32
state and FFTs, FP64 reductions).
"""
mutable struct ROCmSpecSetup
    handle::Ptr{Cvoid}
    n::Int
    l::Any
    visc::Any
    kcut::Int
    m::Int
    T::DataType
    backend::ROCmNative
    topo::Any
    lphys::Any
    lspec::Any
    fft::Any
end

"""
    spectral_setup(; n, l, visc, kcut, procgrid, mpibuf, backend = ROCmNative(), T, comm)

The reference's `spectral_setup` (src/spectral/setup.jl:33-158) on one GPU;
any other `backend` falls through to it.
"""
function spectral_setup(;
    n,
    l = 2π,
    visc = 1e-3,
    kcut = n ÷ 3,
    procgrid = nothing,
    mpibuf = :auto,
    backend = ROCmNative(),
    T = Float64,
    comm = MPI.COMM_WORLD,
)
    backend isa ROCmNative ||
        return DNS.spectral_setup(; n, l, visc, kcut, procgrid, mpibuf, backend, T, comm)
    n % 2 == 0 || error("n must be even")
    1 ≤ kcut || error("kcut must be at least 1")
    3kcut ≤ n || error("aliasing: 3kcut ≤ n is required (2/3 rule)")
    something(procgrid, (1, 1)) == (1, 1) || error("the ROCm backend runs on one GPU: procgrid must be (1, 1)")
    T in (Float64, Float32) || error("T must be Float64 or Float32")
    MPI.Initialized() || MPI.Init()
    topo = DNS.topology(MPI.COMM_SELF, (1, 1), (true, true))
    m = 2kcut + 1
    lphys = DNS.layout((n, n, n), DNS.AXES.x, (1, 1), topo.coords)
    lspec = DNS.layout((kcut + 1, m, m), DNS.AXES.y, (1, 1), topo.coords)
    params = SpecParams(Int32(n), Int32(kcut), Float64(l), Float64(visc), T == Float32 ? DNS_MIXED : DNS_FP64)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:dns_spec_create, libdns), Cint, (Ref{SpecParams}, Ref{Ptr{Cvoid}}), params, h))
    s = ROCmSpecSetup(h[], n, T(l), T(visc), kcut, m, T, ROCmNative(), topo, lphys, lspec,
        (; nb = 3, v = ROCmPhys(h[])))
    finalizer(s) do s
        s.handle == C_NULL || ccall((:dns_spec_destroy, libdns), Cint, (Ptr{Cvoid},), s.handle)
        s.handle = C_NULL
    end
    s
end

"The handle's spectral state, (kcut+1, m, m, 3) complex on the device."
struct ROCmSpecState
    s::ROCmSpecSetup
    which::Int32   # 0 the state uh, 2 the last spectral_rhs! result
end

"specvelocity(s) (src/spectral/setup.jl:161-162): the handle's state."
DNS.specvelocity(s::ROCmSpecSetup) = ROCmSpecState(s, Int32(0))

function getstate(uh::ROCmSpecState)
    s = uh.s
    a = zeros(ComplexF64, s.kcut + 1, s.m, s.m, 3)
    check(ccall((:dns_spec_get_state, libdns), Cint, (Ptr{Cvoid}, Int32, Ptr{ComplexF64}), s.handle, uh.which, a))
    a
end

Base.convert(::Type{Array{Complex{T},4}}, uh::ROCmSpecState) where {T} = Array{Complex{T},4}(getstate(uh))
Base.Array(uh::ROCmSpecState) = convert(Array{Complex{uh.s.T},4}, uh)

function Base.copyto!(uh::ROCmSpecState, a::AbstractArray{<:Complex,4})
    b = Array{ComplexF64,4}(a)
    check(ccall((:dns_spec_set_state, libdns), Cint, (Ptr{Cvoid}, Int32, Ptr{ComplexF64}), uh.s.handle, uh.which, b))
    uh
end

"spectral_project!(uh, s) (src/spectral/operators.jl:25-29)."
function DNS.spectral_project!(uh::ROCmSpecState, s::ROCmSpecSetup)
    check(ccall((:dns_spec_project, libdns), Cint, (Ptr{Cvoid},), s.handle))
    uh
end

"spec_to_phys!(v, uh, s) (src/spectral/fft.jl:200-214)."
function DNS.spec_to_phys!(v::ROCmPhys, uh::ROCmSpecState, s::ROCmSpecSetup)
    check(ccall((:dns_spec_spec_to_phys, libdns), Cint, (Ptr{Cvoid},), s.handle))
    v
end

"phys_to_spec!(uh, w, s) (src/spectral/fft.jl:221-232)."
function DNS.phys_to_spec!(uh::ROCmSpecState, w::ROCmPhys, s::ROCmSpecSetup)
    check(ccall((:dns_spec_phys_to_spec, libdns), Cint, (Ptr{Cvoid},), s.handle))
    uh
end

"vsummax(v) (src/spectral/step.jl:54-60) of the physical buffer."
function DNS.vsummax(v::ROCmPhys)
    out = Ref{Cdouble}(0.0)
    check(ccall((:dns_spec_vsummax, libdns), Cint, (Ptr{Cvoid}, Ref{Cdouble}), v.handle, out))
    out[]
end

"spectral_rhs!(du, uh, s) (src/spectral/operators.jl:60-79): du names the handle's RHS buffer."
function DNS.spectral_rhs!(du::ROCmSpecState, uh::ROCmSpecState, s::ROCmSpecSetup)
    check(ccall((:dns_spec_rhs, libdns), Cint, (Ptr{Cvoid},), s.handle))
    ROCmSpecState(s, Int32(2))
end

"spectral_step!(uh, s, Δt, cache) (src/spectral/step.jl:68-92); returns vsummax of stage 1."
function DNS.spectral_step!(uh::ROCmSpecState, s::ROCmSpecSetup, Δt, cache)
    vmax = Ref{Cdouble}(0.0)
    check(ccall((:dns_spec_step, libdns), Cint, (Ptr{Cvoid}, Cdouble, Ref{Cdouble}), s.handle, Float64(Δt), vmax))
    s.T(vmax[])
end

"shell_energies(uh, s, shells) (src/spectral/forcing.jl:12-24)."
function DNS.shell_energies(uh::ROCmSpecState, s::ROCmSpecSetup, shells)
    sh = collect(Int, shells)
    isempty(sh) && return s.T[]
    s0, s1 = minimum(sh), maximum(sh)
    E = zeros(Float64, s1 - s0 + 1)
    check(ccall((:dns_spec_shell_energies, libdns), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Cdouble}), s.handle,
        Int32(s0), Int32(s1 - s0 + 1), E))
    s.T[E[k-s0+1] for k in sh]
end

"rescale_shell!(uh, s, sh, factor) (src/spectral/forcing.jl:27-35)."
function DNS.rescale_shell!(uh::ROCmSpecState, s::ROCmSpecSetup, sh, factor)
    f = Ref{Cdouble}(Float64(factor))
    check(ccall((:dns_spec_rescale_shells, libdns), Cint, (Ptr{Cvoid}, Int32, Int32, Ref{Cdouble}), s.handle,
        Int32(sh), Int32(1), f))
    uh
end

"spectral_energy(uh, s) (src/spectral/operators.jl:91-92), compensated FP64 sum."
function DNS.spectral_energy(uh::ROCmSpecState, s::ROCmSpecSetup)
    out = Ref{Cdouble}(0.0)
    check(ccall((:dns_spec_energy, libdns), Cint, (Ptr{Cvoid}, Ref{Cdouble}), s.handle, out))
    out[]
end

"spectral_dissipation(uh, s) (src/spectral/operators.jl:95-101)."
function DNS.spectral_dissipation(uh::ROCmSpecState, s::ROCmSpecSetup)
    out = Ref{Cdouble}(0.0)
    check(ccall((:dns_spec_dissipation, libdns), Cint, (Ptr{Cvoid}, Ref{Cdouble}), s.handle, out))
    out[]
end

"spectral_maxdiv(uh, s) (src/spectral/operators.jl:104-110)."
function DNS.spectral_maxdiv(uh::ROCmSpecState, s::ROCmSpecSetup)
    out = Ref{Cdouble}(0.0)
    check(ccall((:dns_spec_maxdiv, libdns), Cint, (Ptr{Cvoid}, Ref{Cdouble}), s.handle, out))
    out[]
end

"spectral_velocityfield!(uh, s; x, y, z, project) (src/spectral/ic.jl:13-25)."
function DNS.spectral_velocityfield!(uh::ROCmSpecState, s::ROCmSpecSetup; x, y, z, project = true)
    n, l = s.n, Float64(s.l)
    coord(g) = l / n * (g - 0.5)
    h = zeros(Float64, n, n, n, 3)
    for (c, f) in enumerate((x, y, z)), k = 1:n, j = 1:n, i = 1:n
        h[i, j, k, c] = f(coord(i), coord(j), coord(k))
    end
    check(ccall((:dns_spec_set_physical, libdns), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), s.handle, h))
    DNS.phys_to_spec!(uh, s.fft.v, s)
    project && DNS.spectral_project!(uh, s)
    uh
end

"spectral_randomfield!(uh, s; totalenergy, kpeak, seed, profile) (src/spectral/ic.jl:59-84)."
function DNS.spectral_randomfield!(uh::ROCmSpecState, s::ROCmSpecSetup; totalenergy = 1, kpeak = 4, seed = 0,
    profile = k -> k^4 * exp(-2 * (k / kpeak)^2))
    check(ccall((:dns_spec_noise, libdns), Cint, (Ptr{Cvoid}, UInt64), s.handle, UInt64(seed)))
    DNS.phys_to_spec!(uh, s.fft.v, s)
    DNS.spectral_project!(uh, s)
    kdiag = floor(Int, sqrt(3) * s.kcut)
    E = DNS.shell_energies(uh, s, 0:kdiag)
    totalprofile = sum(k -> profile(k), 0:kdiag)
    for (sh, e) in zip(0:kdiag, E)
        e0 = totalenergy * profile(sh) / totalprofile
        DNS.rescale_shell!(uh, s, sh, e > 0 ? sqrt(e0 / e) : zero(s.T))
    end
    uh
end

end # module
