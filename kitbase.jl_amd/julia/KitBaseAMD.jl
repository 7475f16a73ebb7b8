"""
    KitBaseAMD

Julia glue for libkbgpu.so: device-resident KitBase solves stepped on an MI355X through the C ABI of
`include/kbgpu.h`. A `DeviceSolve` replaces KitBase's `ctr` / `face` arrays in the calls to
`reconstruct!`, `evolve!`, `update!`, `timestep` and `solve!`; `download!` writes the state back into
the host `ControlVolume` / `Interface` arrays (for `extract_sol`, `write_jld`, plotting).

Supported: 1D1F1V, 1D2F1V and 1D1F3V (also with KitBase's fast spectral collision, `collision =
"fsm"`, its `Gas.fsm` kernel modes passed through) (`ControlVolume1F/2F` over `eachindex(ps.x)`, ghosts included;
`Interface1F/2F` faces 1..nx+1), 2D1F2V and 2D2F2V without ghost cells (`ControlVolume1F/2F`
matrices with `a1face` (nx+1) x ny and `a2face` nx x (ny+1)); fluxes :kfvs, :kcu, :ugks; collisions
:bgk, :shakhov (not 1D1F3V, whose KitBase step! cannot run it, step_1f.jl:91, nor one-distribution
2D); boundaries :fix, :maxwell.  Unstructured triangle meshes (`UnstructPSpace`, `ControlVolumeUS1F/2F`,
`Interface2D1F/2F`): :kfvs at interpOrder 1, walls from `cellType` 2.

KitBase v0.9.31 fields read here (file:line of the struct that defines each):
- `KS.set.space, flux, collision, limiter, interpOrder, boundary, cfl, maxTime`  Setup, struct_setup.jl:106-119
  (`boundary` is a vector of 2 (1D) or 4 (2D) strings, struct_setup.jl:134-142)
- `KS.gas.K, γ, μᵣ, ω, Pr`                                                    Gas, struct_property.jl:51-64
- `KS.vs.u, v, w, weights, u1, v1`                         VSpace1D/2D/3D, phase_velocity.jl:14-21, 128-140, 341-358
- `KS.ps.x0, x1, nx, x, dx` (1D)                                                PSpace1D, geo_struct.jl:10-16
- `KS.ps.nx, ny, x, dx, dy, vertices, areas, n` (2D)                            PSpace2D, geo_struct.jl:47-61
- `KS.ib.bc, KS.ib.p`: `bc(x, p)` in 1D (cfg_sod.jl:32-38), `bc(x, y, p)` in 2D (cfg_cavity.jl:29-36),
  called as evolve_boundary.jl:19,35 and :77,100,123,147 call them            IB1F/IB2F, struct_ib.jl:33-61
- `ctr[i].w, prim, f` / `h, b` (and `sw, sf` / `sh, sb` for slopes)      ControlVolume1F/2F, struct_ctr.jl:51-58, 94-103
- `face[k].fw, ff` / `fh, fb`                                             Interface1F/2F, struct_face.jl:37-40, 56-60
- `KS.ps.cellid, faceCells, cellFaces, cellNeighbors, cellType, cellArea, faceCenter` (unstructured)
                                                                          UnstructPSpace, geo_unstruct.jl:10-32
- `ctr[i].n, dx, w, prim, f` / `h, b`; `face[k].n, len, fw, ff` / `fh, fb` (unstructured)
                                        ControlVolumeUS1F/2F, struct_ctr.jl:1074-1185; Interface2D1F/2F, struct_face.jl:349-400
"""
module KitBaseAMD

using KitBase
using LinearAlgebra: dot

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
    fsm_nm::Cint; fsm_shape::NTuple{3,Cint}; fsm_kn::Cdouble
    fsm_phi::Ptr{Cdouble}; fsm_psi::Ptr{Cdouble}; fsm_phipsi::Ptr{Cdouble}
end

# kbg_udesc (include/kbgpu.h), field for field
struct KbgUDesc
    nf::Cint; nv::Cint; ncell::Cint; nface::Cint
    precision::Cint; collision::Cint; interp_order::Cint
    K::Cdouble; gamma::Cdouble; mu_ref::Cdouble; omega::Cdouble; Pr::Cdouble
    u::Ptr{Cdouble}; v::Ptr{Cdouble}; weights::Ptr{Cdouble}
    face_cells::Ptr{Cint}; face_kind::Ptr{Cint}; face_geo::Ptr{Cdouble}; face_wall::Ptr{Cdouble}
    cell_faces::Ptr{Cint}; cell_dirc::Ptr{Cdouble}; cell_area::Ptr{Cdouble}; cell_type::Ptr{Cint}
    cell_nb::Ptr{Cint}; cell_dmin::Ptr{Cdouble}
end

const FLUX = Dict(:kfvs => 0, :kcu => 1, :ugks => 2)
const COLL = Dict(:bgk => 0, :shakhov => 1, :fsm => 2)
const BCS = Dict(:fix => 0, :maxwell => 1)
const LIMITER = Dict(:vanleer => 0, :minmod => 1, :superbee => 2, :vanalbaba => 3, :linear => 4)

function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:kbg_last_error, libkbgpu), Cstring, ()))
    error("kbgpu error $rc: $msg")
end

lookup(table, key, what) = haskey(table, Symbol(key)) ? table[Symbol(key)] : error("$what $(key) is not supported by libkbgpu")

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
    dim::Int
    keep::Vector{Any}
end

# host arrays in KitBase's own layout: column-major, velocity-contiguous per cell, ghost cells included
flat(x) = collect(Float64, vec(parent(x)))

# the four boundary symbols of the descriptor: Setup.boundary holds 2 (1D) or 4 (2D) entries
function boundary_codes(set)
    b = set.boundary isa AbstractVector ? set.boundary : [set.boundary]
    codes = [Cint(lookup(BCS, s, "boundary")) for s in b]
    length(codes) == 1 && (codes = fill(codes[1], 4))
    return ntuple(k -> k <= length(codes) ? codes[k] : Cint(0), 4)
end

"""
Per-face [n1, n2, len] of the faces the 2D evolve! visits, in the library's face order
(x-faces (i, j) at (i-1) + (nx+1)(j-1), y-faces (i, j) at (i-1) + nx(j-1)):
interior faces take the normal and edge length of the cell on their low side
(solver_evolution.jl:645-646, 663-664); walls take the outward normal of the boundary cell and its
dy / dx (evolve_boundary.jl:72-152).
"""
function face_geometry_2d(ps)
    nx, ny = ps.nx, ps.ny
    xg = zeros(3, nx + 1, ny)
    yg = zeros(3, nx, ny + 1)
    for j in 1:ny
        for i in 2:nx
            xg[1:2, i, j] .= ps.n[i-1, j, 2]
            xg[3, i, j] = ps.areas[i-1, j, 2]
        end
        xg[1:2, 1, j] .= -ps.n[1, j, 4]
        xg[3, 1, j] = ps.dy[1, j]
        xg[1:2, nx+1, j] .= ps.n[nx, j, 2]
        xg[3, nx+1, j] = ps.dy[nx, j]
    end
    for i in 1:nx
        for j in 2:ny
            yg[1:2, i, j] .= ps.n[i, j-1, 3]
            yg[3, i, j] = ps.areas[i, j-1, 3]
        end
        yg[1:2, i, 1] .= -ps.n[i, 1, 1]
        yg[3, i, 1] = ps.dx[i, 1]
        yg[1:2, i, ny+1] .= ps.n[i, ny, 3]
        yg[3, i, ny+1] = ps.dx[i, ny]
    end
    return xg, yg
end

"""
Global-frame wall primitive vectors of the boundary faces, from `KS.ib.bc(xc, yc, KS.ib.p)` at the
edge midpoints, as evolve_boundary.jl:74-76, 97-99, 120-122, 144-146 evaluate them (the library
rotates them into each face's frame, as `local_frame` does there).
"""
function wall_prims_2d(KS)
    ps = KS.ps
    nx, ny = ps.nx, ps.ny
    vt = ps.vertices
    xw = zeros(4, nx + 1, ny)
    yw = zeros(4, nx, ny + 1)
    mid(i, j, a, b) = ((vt[i, j, a, 1] + vt[i, j, b, 1]) / 2, (vt[i, j, a, 2] + vt[i, j, b, 2]) / 2)
    for j in 1:ny
        xw[:, 1, j] .= KS.ib.bc(mid(1, j, 1, 4)..., KS.ib.p)
        xw[:, nx+1, j] .= KS.ib.bc(mid(nx, j, 2, 3)..., KS.ib.p)
    end
    for i in 1:nx
        yw[:, i, 1] .= KS.ib.bc(mid(i, 1, 1, 2)..., KS.ib.p)
        yw[:, i, ny+1] .= KS.ib.bc(mid(i, ny, 3, 4)..., KS.ib.p)
    end
    return xw, yw
end

function make_desc(ks::SolverSet, precision, keep)
    ptr(a) = (a = flat(a); push!(keep, a); pointer(a))
    ps, vs, gas = ks.ps, ks.vs, ks.gas
    space = ks.set.space
    dim = parse(Int, space[1:1])
    nf = space[3] == '2' ? 2 : 1
    vdim = parse(Int, space[end-1:end-1])
    nul = Ptr{Cdouble}(C_NULL)
    u = ptr(vs.u); w = ptr(vs.weights)
    v = vdim >= 2 ? ptr(vs.v) : nul
    ww = vdim == 3 ? ptr(vs.w) : nul
    prec = precision == :fp64 ? Cint(1) : Cint(0)
    common = (lookup(FLUX, ks.set.flux, "flux"), lookup(COLL, ks.set.collision, "collision"),
              lookup(LIMITER, ks.set.limiter, "limiter"), ks.set.interpOrder, boundary_codes(ks.set))
    # collision :fsm: Gas.fsm = (Kn, nm, ϕ, ψ, χ) of gas_property(vs; ..., collision = "fsm")
    # (struct_property.jl:516-567), kernel modes [u, v, w, idx] column-major as the library reads them
    fsm = if Symbol(ks.set.collision) == :fsm
        (Cint(gas.fsm.nm), (Cint(vs.nu), Cint(vs.nv), Cint(vs.nw)), gas.fsm.Kn, ptr(gas.fsm.ϕ), ptr(gas.fsm.ψ), ptr(gas.fsm.χ))
    else
        (Cint(0), (Cint(0), Cint(0), Cint(0)), 0.0, nul, nul, nul)
    end
    if dim == 1
        ng = 1 - first(axes(ps.x, 1))                  # ghost layers of the OffsetArray, update_boundary.jl:40
        dx = ptr(ps.dx)                                # parent: cells 1-ng .. nx+ng
        bcp = ptr(vcat(ks.ib.bc(ps.x0, ks.ib.p), ks.ib.bc(ps.x1, ks.ib.p)))
        return KbgDesc(1, nf, vdim, length(vs.u), ps.nx, 1, ng, prec, common...,
                       gas.K, gas.γ, gas.μᵣ, gas.ω, gas.Pr, u, v, ww, w, dx, nul, bcp, nul, nul, nul, nul, fsm...)
    elseif dim == 2
        (first(axes(ps.x, 1)) == 1 && first(axes(ps.x, 2)) == 1) ||
            error("libkbgpu steps 2D solves without ghost cells (PSpace2D with NGX = NGY = 0)")
        xg, yg = face_geometry_2d(ps)
        xw, yw = wall_prims_2d(ks)
        return KbgDesc(2, nf, vdim, length(vs.u), ps.nx, ps.ny, 0, prec, common...,
                       gas.K, gas.γ, gas.μᵣ, gas.ω, gas.Pr, u, v, ww, w, ptr(ps.dx), ptr(ps.dy), nul,
                       ptr(xg), ptr(yg), ptr(xw), ptr(yw), fsm...)
    end
    error("space $(space) is not supported by libkbgpu")
end

function create(ks::SolverSet, precision, ctx)
    keep = Any[]
    d = make_desc(ks, precision, keep)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep check(ccall((:kbg_solve_create, libkbgpu), Cint, (Ptr{Cvoid}, Ref{KbgDesc}, Ref{Ptr{Cvoid}}),
                                  ctx.handle, Ref(d), h))
    dev = DeviceSolve(h[], ctx, ks, d.dim, keep)
    finalizer(s -> ccall((:kbg_solve_destroy, libkbgpu), Cint, (Ptr{Cvoid},), s.handle), dev)
    return dev
end

"""
    DeviceSolve(ks, ctr, face; precision=:fp32, ctx)            # 1D
    DeviceSolve(ks, ctr, a1face, a2face; precision=:fp32, ctx)  # 2D

Upload the state of `init_fvm` (or of a checkpoint) once; it stays resident in HBM.
"""
function DeviceSolve(ks::SolverSet, ctr, face; precision=:fp32, ctx=DeviceContext(0))
    dev = create(ks, precision, ctx)
    dev.dim == 1 || error("2D solves take DeviceSolve(ks, ctr, a1face, a2face)")
    return upload!(dev, ctr, face)
end

function DeviceSolve(ks::SolverSet, ctr, a1face, a2face; precision=:fp32, ctx=DeviceContext(0))
    dev = create(ks, precision, ctx)
    dev.dim == 2 || error("1D solves take DeviceSolve(ks, ctr, face)")
    return upload!(dev, ctr, a1face, a2face)
end

# ---- unstructured triangle meshes (UnstructPSpace, geo_unstruct.jl:10-32)

# 0-based indices, -1 kept for "no cell"
zb(a) = Cint[x == -1 ? Cint(-1) : Cint(x - 1) for x in a]

"""
The descriptor of an unstructured solve: ps.faceCells / cellFaces / cellNeighbors (row-major, 0-based),
face kinds (0 interior, 1 Maxwell wall: a boundary face of a cellType-2 cell, 2 a boundary face whose
flux evolve! leaves alone, solver_evolution.jl:883-935), face[i].n and len, the wall state
KS.ib.bc(faceCenter, p) (solver_evolution.jl:906-910), the dirc signs of update! (solver_update.jl:375),
ps.cellArea, ps.cellType and the least of ctr[i].dx (timestep, solver_main.jl:400-426).
"""
function make_udesc(ks::SolverSet, ctr, face, precision, keep)
    ps, vs, gas = ks.ps, ks.vs, ks.gas
    keepd(a) = (a = collect(Float64, a); push!(keep, a); pointer(a))
    keepi(a) = (a = collect(Cint, a); push!(keep, a); pointer(a))
    nc, nfc = size(ps.cellid, 1), size(ps.faceCells, 1)
    kind = Cint[]
    geo = Float64[]
    wall = zeros(Float64, 4 * nfc)
    for i in 1:nfc
        c1, c2 = ps.faceCells[i, 1], ps.faceCells[i, 2]
        append!(geo, (face[i].n[1], face[i].n[2], face[i].len))
        if c1 != -1 && c2 != -1
            push!(kind, 0)
        elseif ps.cellType[c1 != -1 ? c1 : c2] == 2
            push!(kind, 1)
            wall[4i-3:4i] .= ks.ib.bc(ps.faceCenter[i, 1], ps.faceCenter[i, 2], ks.ib.p)
        else
            push!(kind, 2)
        end
    end
    dirc = [sign(dot(ctr[i].n[j], face[ps.cellFaces[i, j]].n)) for j in 1:3, i in 1:nc]
    nf = hasproperty(first(ctr), :h) ? 2 : 1
    return KbgUDesc(nf, length(vs.u), nc, nfc, precision == :fp64 ? Cint(1) : Cint(0),
                    lookup(COLL, ks.set.collision, "collision"), ks.set.interpOrder,
                    gas.K, gas.γ, gas.μᵣ, gas.ω, gas.Pr,
                    keepd(vec(vs.u)), keepd(vec(vs.v)), keepd(vec(vs.weights)),
                    keepi(vec(permutedims(zb(ps.faceCells)))), keepi(kind), keepd(geo), keepd(wall),
                    keepi(vec(permutedims(zb(ps.cellFaces)))), keepd(vec(dirc)), keepd(ps.cellArea),
                    keepi(ps.cellType), keepi(vec(permutedims(zb(ps.cellNeighbors)))),
                    keepd([minimum(c.dx) for c in ctr]))
end

"""
    DeviceSolve(ks, ctr, face; precision=:fp32, ctx)   # ks.ps an UnstructPSpace

An unstructured triangle solve (2D1F2V / 2D2F2V, interpOrder 1) from init_fvm's ctr and face
(solver_init.jl:387-640); evolve!, update!, timestep, solve! and step! then work on it as on the others.
"""
function DeviceSolve(ks::SolverSet{<:Any,<:KitBase.UnstructPSpace}, ctr, face; precision=:fp32,
                     ctx=DeviceContext(0))
    keep = Any[]
    d = make_udesc(ks, ctr, face, precision, keep)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep check(ccall((:kbg_usolve_create, libkbgpu), Cint, (Ptr{Cvoid}, Ref{KbgUDesc}, Ref{Ptr{Cvoid}}),
                                  ctx.handle, Ref(d), h))
    dev = DeviceSolve(h[], ctx, ks, 0, keep)
    finalizer(s -> ccall((:kbg_solve_destroy, libkbgpu), Cint, (Ptr{Cvoid},), s.handle), dev)
    return upload!(dev, ctr, face)
end

cellvec(ctr, field) = reduce(vcat, [flat(getfield(c, field)) for c in ctr])
cellf(ctr) = reduce(vcat, [hasproperty(c, :h) ? vcat(flat(c.h), flat(c.b)) : flat(c.f) for c in ctr])
facevec(face) = reduce(vcat, [flat(f.fw) for f in face])
faceff(face) = reduce(vcat, [hasproperty(f, :fh) ? vcat(flat(f.fh), flat(f.fb)) : flat(f.ff) for f in face])

function upload_cells!(dev, ctr)
    w, prim, f = cellvec(ctr, :w), cellvec(ctr, :prim), cellf(ctr)
    GC.@preserve w prim f check(ccall((:kbg_upload, libkbgpu), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                                      dev.handle, w, prim, f))
end

function upload_faces!(dev, fw, ff)
    GC.@preserve fw ff check(ccall((:kbg_upload_faces, libkbgpu), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                                   dev.handle, fw, ff))
end

function upload!(dev::DeviceSolve, ctr, face=nothing)
    upload_cells!(dev, ctr)
    face === nothing || upload_faces!(dev, facevec(face), faceff(face))
    return dev
end

# 2D: x-faces (a1face, column-major) then y-faces (a2face)
function upload!(dev::DeviceSolve, ctr, a1face, a2face)
    upload_cells!(dev, ctr)
    upload_faces!(dev, vcat(facevec(a1face), facevec(a2face)), vcat(faceff(a1face), faceff(a2face)))
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

# KitBase's entry points on device-resident state (solver_reconstruction.jl:56, solver_evolution.jl:39,608,
# solver_update.jl:61,193, solver_main.jl:12,90). `mode` / `coll` / `bc` must match the SolverSet the solve
# was created from: the device solve carries them.
KitBase.reconstruct!(KS::SolverSet, dev::DeviceSolve) = check(ccall((:kbg_reconstruct, libkbgpu), Cint, (Ptr{Cvoid},), dev.handle))

function check_mode(KS, mode, coll)
    mode === nothing || Symbol(mode) == Symbol(KS.set.flux) || error("mode $mode differs from the device solve's flux $(KS.set.flux)")
    coll === nothing || Symbol(coll) == Symbol(KS.set.collision) ||
        error("coll $coll differs from the device solve's collision $(KS.set.collision)")
end

function KitBase.evolve!(KS::SolverSet, dev::DeviceSolve, face, dt; mode=nothing, bc=nothing)
    check_mode(KS, mode, nothing)
    check(ccall((:kbg_evolve, libkbgpu), Cint, (Ptr{Cvoid}, Cdouble), dev.handle, dt))
end

KitBase.evolve!(KS::SolverSet, dev::DeviceSolve, a1face, a2face, dt; mode=nothing, bc=nothing) =
    KitBase.evolve!(KS, dev, nothing, dt; mode=mode, bc=bc)

function KitBase.update!(KS::SolverSet, dev::DeviceSolve, face, dt, residual; coll=nothing, bc=nothing, kwargs...)
    check_mode(KS, nothing, coll)
    r = zeros(Cdouble, 5)
    check(ccall((:kbg_update, libkbgpu), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Cdouble}), dev.handle, dt, r))
    residual .= r[1:length(residual)]
    return nothing
end

KitBase.update!(KS::SolverSet, dev::DeviceSolve, a1face, a2face, dt, residual; coll=nothing, bc=nothing, kwargs...) =
    KitBase.update!(KS, dev, nothing, dt, residual; coll=coll, bc=bc)

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
    solve!(KS, dev, simTime) -> simTime + nt dt

`solve!`'s loop (solver_main.jl:12-76, 90-150) on the device: dt from `timestep` once, `nt =
floor(maxTime / dt) + 1` fused steps. The JLD2 checkpoints of the reference loop are the caller's:
`download!(ctr, dev)` then `write_jld(KS, ctr, t)`.
"""
function KitBase.solve!(KS::SolverSet, dev::DeviceSolve, simTime; steady=false)
    dt = KitBase.timestep(KS, dev, simTime)
    nt = Int(floor(KS.set.maxTime / dt)) + 1
    step!(dev, dt, nt)
    return simTime + nt * dt
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
