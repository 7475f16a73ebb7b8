"""
    KitBaseAMD

Julia glue for libkbgpu.so: device-resident KitBase solves stepped on an MI355X through the C ABI of
`include/kbgpu.h`. A `DeviceSolve` replaces KitBase's `ctr` / `face` arrays in the calls to
`reconstruct!`, `evolve!`, `update!` and `solve!`; `download!` writes the state back into the host
`ControlVolume` / `Interface` arrays (for `extract_sol`, `write_jld`, plotting).

Supported: 1D1F1V, 1D2F1V and 1D1F3V (`Interface1F/2F`, `ControlVolume1F/2F`), 2D2F2V with
`a1face`/`a2face`; fluxes :kfvs, :kcu, :ugks; collisions :bgk, :shakhov; boundaries :fix, :maxwell.
"""
module KitBaseAMD

using KitBase

export DeviceContext, DeviceSolve, upload!, download!, batch_step!

const libkbgpu = get(ENV, "KBG_LIBRARY", joinpath(@__DIR__, "..", "lib", "libkbgpu.so"))

# kbg_desc (include/kbgpu.h), field for field
struct KbgDesc
    dim::Cint; nf::Cint; vdim::Cint; nv::Cint
    nx::Cint; ny::Cint; ng::Cint; precision::Cint
    flux::Cint; collision::Cint; limiter::Cint; interp_order::Cint
    bc::NTuple{4,Cint}
    K::Cdouble; gamma::Cdouble; mu_ref::Cdouble; omega::Cdouble; Pr::Cdouble
    u::Ptr{Cdouble}; v::Ptr{Cdouble}; w::Ptr{Cdouble}; weights::Ptr{Cdouble}
    dx::Ptr{Cdouble}; dy::Ptr{Cdouble}; bc_prim::Ptr{Cdouble}
    xgeo::Ptr{Cdouble}; ygeo::Ptr{Cdouble}; xwall::Ptr{Cdouble}; ywall::Ptr{Cdouble}
end

const FLUX = Dict(:kfvs => 0, :kcu => 1, :ugks => 2)
const COLL = Dict(:bgk => 0, :shakhov => 1)
const BCS = Dict(:fix => 0, :maxwell => 1)
const LIMITER = Dict(:vanleer => 0, :minmod => 1, :superbee => 2, :vanalbaba => 3, :linear => 4)

function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:kbg_last_error, libkbgpu), Cstring, ()))
    error("kbgpu error $rc: $msg")
end

mutable struct DeviceContext
    handle::Ptr{Cvoid}
    function DeviceContext(device::Integer=0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:kbg_ctx_create, libkbgpu), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h))
        ctx = new(h[])
        finalizer(c -> ccall((:kbg_ctx_destroy, libkbgpu), Cint, (Ptr{Cvoid},), c.handle), ctx)
        return ctx
    end
end

"""
A solve resident in HBM. Keeps the host arrays the descriptor points into alive (`keep`).
"""
mutable struct DeviceSolve
    handle::Ptr{Cvoid}
    ctx::DeviceContext
    ks::SolverSet
    keep::Vector{Any}
end

# host arrays in KitBase's own layout: velocity-contiguous per cell, ghost cells included
flat(x) = collect(Float64, vec(x))

function DeviceSolve(ks::SolverSet, ctr, face; precision=:fp32, ctx=DeviceContext(0))
    keep = Any[]
    ptr(a) = (a = flat(a); push!(keep, a); pointer(a))
    ps, vs, gas = ks.ps, ks.vs, ks.gas
    dim = parse(Int, ks.set.space[1:1])
    nf = ks.set.space[3] == '2' ? 2 : 1
    vdim = parse(Int, ks.set.space[end-1:end-1])
    bcs = ks.set.boundary isa Symbol || ks.set.boundary isa AbstractString ? fill(Symbol(ks.set.boundary), 4) : Symbol.(ks.set.boundary)
    u = ptr(vs.u); w = ptr(vs.weights)
    v = vdim >= 2 ? ptr(vs.v) : Ptr{Cdouble}(C_NULL)
    ww = vdim == 3 ? ptr(vs.w) : Ptr{Cdouble}(C_NULL)
    if dim == 1
        dx = ptr(parent(ps.dx))
        bcp = ptr(vcat(ks.ib.bc(ps.x0), ks.ib.bc(ps.x1)))
        d = KbgDesc(1, nf, vdim, length(vs.u), ps.nx, 1, ps.nxg, precision == :fp64 ? 1 : 0,
                    FLUX[Symbol(ks.set.flux)], COLL[Symbol(ks.set.collision)], LIMITER[Symbol(ks.set.limiter)],
                    ks.set.interpOrder, Tuple(Cint(BCS[b]) for b in bcs), gas.K, gas.γ, gas.μᵣ, gas.ω, gas.Pr,
                    u, v, ww, w, dx, C_NULL, bcp, C_NULL, C_NULL, C_NULL, C_NULL)
    else
        error("2D: build xgeo/ygeo (face normals, lengths) and the wall states as kitbase.jl_amd/geometry.py does")
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep check(ccall((:kbg_solve_create, libkbgpu), Cint, (Ptr{Cvoid}, Ref{KbgDesc}, Ref{Ptr{Cvoid}}), ctx.handle, Ref(d), h))
    dev = DeviceSolve(h[], ctx, ks, keep)
    finalizer(s -> ccall((:kbg_solve_destroy, libkbgpu), Cint, (Ptr{Cvoid},), s.handle), dev)
    upload!(dev, ctr, face)
    return dev
end

cellvec(ctr, field) = reduce(vcat, [flat(getfield(c, field)) for c in ctr])
cellf(ctr) = reduce(vcat, [hasproperty(c, :h) ? vcat(flat(c.h), flat(c.b)) : flat(c.f) for c in ctr])
facevec(face, field) = reduce(vcat, [flat(getfield(f, field)) for f in face])
faceff(face) = reduce(vcat, [hasproperty(f, :fh) ? vcat(flat(f.fh), flat(f.fb)) : flat(f.ff) for f in face])

function upload!(dev::DeviceSolve, ctr, face=nothing)
    w, prim, f = cellvec(ctr, :w), cellvec(ctr, :prim), cellf(ctr)
    GC.@preserve w prim f check(ccall((:kbg_upload, libkbgpu), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                                      dev.handle, w, prim, f))
    if face !== nothing
        fw, ff = facevec(face, :fw), faceff(face)
        GC.@preserve fw ff check(ccall((:kbg_upload_faces, libkbgpu), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                                       dev.handle, fw, ff))
    end
    return dev
end

function download!(ctr, dev::DeviceSolve)
    w, prim, f = cellvec(ctr, :w), cellvec(ctr, :prim), cellf(ctr)
    GC.@preserve w prim f check(ccall((:kbg_download, libkbgpu), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                                      dev.handle, w, prim, f))
    nw = length(first(ctr).w)
    nvf = length(f) ÷ length(ctr)
    for (k, c) in enumerate(ctr)
        c.w .= reshape(w[(k-1)*nw+1:k*nw], size(c.w))
        c.prim .= reshape(prim[(k-1)*nw+1:k*nw], size(c.prim))
        fk = f[(k-1)*nvf+1:k*nvf]
        if hasproperty(c, :h)
            n = length(c.h)
            c.h .= reshape(fk[1:n], size(c.h)); c.b .= reshape(fk[n+1:end], size(c.b))
        else
            c.f .= reshape(fk, size(c.f))
        end
    end
    return ctr
end

# KitBase's entry points on device-resident state (solver_reconstruction.jl:56, solver_evolution.jl:39,
# solver_update.jl:61, solver_main.jl:12). `mode` / `coll` / `bc` must match the SolverSet the solve
# was created from: the device solve carries them.
KitBase.reconstruct!(KS::SolverSet, dev::DeviceSolve) = check(ccall((:kbg_reconstruct, libkbgpu), Cint, (Ptr{Cvoid},), dev.handle))

function KitBase.evolve!(KS::SolverSet, dev::DeviceSolve, face, dt; mode=Symbol(KS.set.flux), bc=nothing)
    mode == Symbol(KS.set.flux) || error("mode $mode differs from the device solve's flux $(KS.set.flux)")
    check(ccall((:kbg_evolve, libkbgpu), Cint, (Ptr{Cvoid}, Cdouble), dev.handle, dt))
end

function KitBase.update!(KS::SolverSet, dev::DeviceSolve, face, dt, residual; coll=Symbol(KS.set.collision), bc=nothing)
    coll == Symbol(KS.set.collision) || error("coll $coll differs from the device solve's collision $(KS.set.collision)")
    r = zeros(Cdouble, 5)
    check(ccall((:kbg_update, libkbgpu), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Cdouble}), dev.handle, dt, r))
    residual .= r[1:length(residual)]
    return nothing
end

# timestep(KS, ctr, simTime) (solver_main.jl:240-342) from the device-resident prim
function KitBase.timestep(KS::SolverSet, dev::DeviceSolve, simTime=0.0)
    v1 = hasproperty(KS.vs, :v1) ? KS.vs.v1 : 0.0
    dt = Ref{Cdouble}(0.0)
    check(ccall((:kbg_timestep, libkbgpu), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Ref{Cdouble}),
                dev.handle, KS.set.cfl, KS.vs.u1, v1, simTime, KS.set.maxTime, dt))
    return dt[]
end

"""
    step!(dev, dt, nsteps) -> residual

`nsteps` fused time steps (reconstruct! + evolve! + update!), state kept in HBM.
"""
function step!(dev::DeviceSolve, dt, nsteps::Integer)
    r = zeros(Cdouble, 5)
    check(ccall((:kbg_step, libkbgpu), Cint, (Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble}), dev.handle, dt, nsteps, r))
    return r
end

"""
    batch_step!(devs, dts, nsteps)

Steps many independent device solves together (one kernel launch per stream and step).
"""
function batch_step!(devs::Vector{DeviceSolve}, dts::Vector{Float64}, nsteps::Integer)
    ctx = first(devs).ctx
    hs = [d.handle for d in devs]
    r = zeros(Cdouble, 5 * length(devs))
    check(ccall((:kbg_batch_step, libkbgpu), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}),
                ctx.handle, hs, length(devs), dts, nsteps, r))
    return reshape(r, 5, :)
end

end # module
