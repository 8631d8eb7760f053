# FarCUR.jl — Julia bindings of libfarcur.so (the MI355X far-field CUR library).
#
# Drop-in: the reference's pure_cuda_compression2
# (codes/GPU_implementation/seed/compres_single_block_spheres_for_sweep2.jl:591-660)
# calls `ccall((:cuda_complete_compression, "./functionsExternal_cuda.so"), ...)`.
# Pointing that ccall at libfarcur.so is the whole change; the argument list is
# identical (66 arguments). The native API below uploads the geometry once and
# runs the adaptive loop and the batched far field on the device.
module FarCUR

const LIB = get(ENV, "FARCUR_LIB", joinpath(@__DIR__, "..", "libfarcur.so"))

struct Opts
    tol::Cdouble
    pinv_tol::Cdouble
    pinv_rtol::Cdouble
    seed::UInt64
    precision::Cint
    row_mode::Cint
    r_start::Cint
    max_rank::Cint
    block_id::Cint
    heldout::Cint
    oversample::Cint
    label_eps::Cdouble
    heldout_safety::Cdouble
    growth::Cdouble
    learned_core::Cint
end

struct BlockInfo
    r::Cint
    rounds::Cint
    kept::Cint
    converged::Cint
    err::Cdouble
    heldout_err::Cdouble
    sigma_max::Cdouble
    time_ms::Cdouble
    rank_eps::Cint
    status::Cint
    sweeps::Cint
    pad::Cint
end

function check(rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:farcur_last_error, LIB), Cstring, ()))
    error("farcur error $rc: $msg")
end

function default_opts(; kw...)
    o = Ref{Opts}()
    ccall((:farcur_opts_default, LIB), Cvoid, (Ref{Opts},), o)
    x = o[]
    vals = Dict(f => getfield(x, f) for f in fieldnames(Opts))
    for (k, v) in kw
        vals[k] = v
    end
    Opts((convert(fieldtype(Opts, f), vals[f]) for f in fieldnames(Opts))...)
end

mutable struct Context
    h::Ptr{Cvoid}
end

function Context(device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:farcur_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h))
    ctx = Context(h[])
    finalizer(c -> ccall((:farcur_destroy, LIB), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    ctx
end

set_em!(ctx::Context, k, eta) = check(ccall((:farcur_set_em, LIB), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), ctx.h, k, eta))

"Geometry from the obj_struct tables (1-based ids, Float64), as SW2:513-545 converts them."
function set_geometry!(ctx::Context, obj)
    vertex = convert(Array{Cdouble,2}, obj.vertex)
    topol = convert(Array{Cdouble,2}, obj.topol)
    edges = convert(Array{Cdouble,2}, obj.edges)
    ln = convert(Array{Cdouble,1}, vec(obj.ln))
    check(ccall((:farcur_set_geometry, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}),
                ctx.h, vertex, size(vertex, 2), topol, size(topol, 2), edges, size(edges, 2), ln))
end

"""
    compress_block(ctx, r1, r2; tol=1e-4, seed=12345) -> (C, U, R, row_samples, col_samples, info)

The adaptive loop of cuda_cur_compress3 (SW2:878-973) on the device. `r1`, `r2`
are 1-based RWG ids like the reference's; returned samples are 1-based too.
"""
function compress_block(ctx::Context, r1, r2; kw...)
    rows = Cint.(collect(r1) .- 1)
    cols = Cint.(collect(r2) .- 1)
    o = default_opts(; kw...)
    info = Ref{BlockInfo}()
    check(ccall((:farcur_compress_block, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cint}, Cint, Ptr{Cint}, Cint, Ref{Opts}, Ref{BlockInfo}),
                ctx.h, rows, length(rows), cols, length(cols), o, info))
    m, n, r = length(rows), length(cols), info[].r
    C = zeros(ComplexF64, m, r)
    U = zeros(ComplexF64, r, r)
    R = zeros(ComplexF64, r, n)
    rs = zeros(Cint, r)
    cs = zeros(Cint, r)
    y = zeros(ComplexF64, m)
    check(ccall((:farcur_block_result, LIB), Cint,
                (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{Cint}, Ptr{Cint}, Ptr{ComplexF64}),
                ctx.h, C, U, R, rs, cs, y))
    return C, U, R, rs .+ 1, cs .+ 1, info[]
end

"Dense panel Z[rows, cols] (1-based ids), FP64 (precision=0) or complex FP32 (1)."
function fill(ctx::Context, r1, r2; precision = 0)
    rows = Cint.(collect(r1) .- 1)
    cols = Cint.(collect(r2) .- 1)
    Z = zeros(ComplexF64, length(rows), length(cols))
    check(ccall((:farcur_fill, LIB), Cint, (Ptr{Cvoid}, Ptr{Cint}, Cint, Ptr{Cint}, Cint, Cint, Ptr{ComplexF64}),
                ctx.h, rows, length(rows), cols, length(cols), precision, Z))
    Z
end

mutable struct Batch
    h::Ptr{Cvoid}
    ctx::Context
    row_off::Vector{Cint}
    col_off::Vector{Cint}
end

"Far-field block list: blocks[b] = (rows, cols) with 1-based RWG ids."
function Batch(ctx::Context, blocks)
    ro, co = Cint[0], Cint[0]
    ri, ci = Cint[], Cint[]
    for (r, c) in blocks
        append!(ri, Cint.(collect(r) .- 1)); push!(ro, length(ri))
        append!(ci, Cint.(collect(c) .- 1)); push!(co, length(ci))
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:farcur_batch_create, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ref{Ptr{Cvoid}}),
                ctx.h, length(blocks), ro, ri, co, ci, h))
    b = Batch(h[], ctx, ro, co)
    finalizer(x -> ccall((:farcur_batch_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.h), b)
    b
end

"mode 0: per-block loop (the reference's), 1: batched rounds, 2: learned rank. Returns wall ms."
function compress!(b::Batch; mode = 1, kw...)
    o = default_opts(; kw...)
    wall = Ref{Cdouble}(0)
    check(ccall((:farcur_batch_compress, LIB), Cint, (Ptr{Cvoid}, Ref{Opts}, Cint, Ref{Cdouble}), b.h, o, mode, wall))
    wall[]
end

function info(b::Batch)
    out = Vector{BlockInfo}(undef, length(b.row_off) - 1)
    check(ccall((:farcur_batch_info, LIB), Cint, (Ptr{Cvoid}, Ptr{BlockInfo}), b.h, out))
    out
end

function block(b::Batch, i::Integer)
    m = b.row_off[i+1] - b.row_off[i]
    n = b.col_off[i+1] - b.col_off[i]
    r = info(b)[i].r
    C = zeros(ComplexF64, m, r); U = zeros(ComplexF64, r, r); R = zeros(ComplexF64, r, n)
    check(ccall((:farcur_batch_block, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{Cint}, Ptr{Cint}),
                b.h, i - 1, C, U, R, C_NULL, C_NULL))
    C, U, R
end

end # module
