# ScaleTrackHIP.jl -- the Julia side of libscaletrack: `ccall` bindings of the C-ABI in
# include/scaletrack.h, for the reference's tracker (run/icoJuliaParcelFoam/cavity3D/).
#
#     include("ScaleTrackHIP.jl"); using .ScaleTrackHIP
#
# The library path defaults to the in-tree build next to this file
# (../scaletrack/libscaletrack.so); the environment variable SCALETRACK_LIB overrides it.
#
# What each call replaces in asyncParallelTracking.jl (A) and syncSerialTracking.jl (S):
#   evolve             the per-chunk kernel launch A:1098-1125 (compiled at A:963-972) of
#                      evolve_on_device!/evolve_particle! A:1257-1365, with the bounding
#                      box of A:749-797 and A:1245-1255 as its result
#   zero_source!       fill!(UTrans, 0), A:1064-1067
#   upload_U!          copyto! of U, A:1076-1083
#   download_source!   copyto! of UTrans, A:1148-1152
#   cloud_create / cloud_upload! / cloud_download!
#                      the CuArray fields of Chunk, A:329-364
#   evolve!            evolve!(chunk, eulerian, mesh, dt, ::GPU), S:325-345
# The source-misfit rrule (ChainRulesCore) is defined when ChainRulesCore is loaded
# before this file is included.
module ScaleTrackHIP

const lib = get(ENV, "SCALETRACK_LIB", joinpath(@__DIR__, "..", "scaletrack", "libscaletrack.so"))

const ST_FP64 = Int32(0)
const ST_FP32 = Int32(1)
const ST_WALL = Int32(0)
const ST_PERIODIC = Int32(1)
const ST_CHANNEL = Int32(2)
const ST_METHOD_PLAIN = Int32(0)
const ST_METHOD_SORTED = Int32(1)
const ST_FLUSH_OWNER = Int32(0)
const ST_FLUSH_ATOMIC = Int32(1)

# C layouts of st_mesh, st_params and st_tune_result (isbits structs follow C alignment)
struct StMesh
    n::NTuple{3,Int32}
    origin::NTuple{3,Float64}
    ending::NTuple{3,Float64}
end

struct StParams
    boundary::Int32
    method::Int32
    flush::Int32
    substeps_per_pass::Int32
    cd::Float64
    brick::NTuple{3,Int32}
    block::Int32
    resort_fraction::Float64
    passes_per_launch::Int32
end

struct StTuneResult
    target_parcels_per_brick::Int32
    block::Int32
    blocks_per_cu::Int32
    from_cache::Int32
    ms_per_pass::Float64
    runner_up_ms_per_pass::Float64
end

struct StError <: Exception
    status::Int32
    msg::String
end
Base.showerror(io::IO, e::StError) = print(io, "libscaletrack status ", e.status, ": ", e.msg)

function check(r::Integer)
    r == 0 && return nothing
    throw(StError(Int32(r), unsafe_string(ccall((:st_last_error, lib), Cstring, ()))))
end

version() = Int(ccall((:st_version, lib), Cint, ()))

function device_count()
    n = Ref{Int32}(0)
    check(ccall((:st_device_count, lib), Cint, (Ref{Int32},), n))
    Int(n[])
end

"""params(; kw...) -- st_params_default() with the named fields replaced."""
function params(; kw...)
    p = Ref{StParams}()
    ccall((:st_params_default, lib), Cvoid, (Ref{StParams},), p)
    d = p[]
    vals = map(fieldnames(StParams)) do f
        v = get(kw, f, getfield(d, f))
        convert(fieldtype(StParams, f), f === :brick ? Tuple(Int32.(v)) : v)
    end
    StParams(vals...)
end

precision_of(::Type{Float64}) = ST_FP64
precision_of(::Type{Float32}) = ST_FP32

# ---- context ------------------------------------------------------------------------

mutable struct Context
    h::Ptr{Cvoid}
end
function Context(device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_ctx_create, lib), Cint, (Int32, Ref{Ptr{Cvoid}}), device, h))
    Context(h[])
end
synchronize(ctx::Context) = check(ccall((:st_ctx_synchronize, lib), Cint, (Ptr{Cvoid},), ctx.h))
function destroy!(ctx::Context)
    ctx.h == C_NULL || check(ccall((:st_ctx_destroy, lib), Cint, (Ptr{Cvoid},), ctx.h))
    ctx.h = C_NULL
    nothing
end

# ---- parcel cloud: the reference's Chunk ----------------------------------------------

mutable struct Cloud{T<:Union{Float32,Float64}}
    h::Ptr{Cvoid}
end

"""cloud_create(ctx, T, N, μᶜ, ρᵈ, ρᵈByρᶜ; rho_c, per_parcel_rho) -- the scalars of
allocate_chunk(executor, N, μᶜ, ρᵈ, ρᵈ/ρᶜ)."""
function cloud_create(ctx::Context, ::Type{T}, N::Integer, μ, ρ, ρByρc; rho_c=0.0,
                      per_parcel_rho=false) where {T<:Union{Float32,Float64}}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_cloud_create, lib), Cint,
                (Ptr{Cvoid}, Int32, Int64, Float64, Float64, Float64, Float64, Int32, Ref{Ptr{Cvoid}}),
                ctx.h, precision_of(T), N, μ, ρ, ρByρc, rho_c, per_parcel_rho ? 1 : 0, h))
    Cloud{T}(h[])
end
function destroy!(c::Cloud)
    c.h == C_NULL || check(ccall((:st_cloud_destroy, lib), Cint, (Ptr{Cvoid},), c.h))
    c.h = C_NULL
    nothing
end

"""cloud_upload!(c, chunk) -- chunk.d, X, Y, Z, U, V, W as host vectors of T."""
function cloud_upload!(c::Cloud{T}, ch; rho::Union{Nothing,Vector{T}}=nothing) where {T}
    check(ccall((:st_cloud_upload, lib), Cint,
                (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}),
                c.h, ch.d, ch.X, ch.Y, ch.Z, ch.U, ch.V, ch.W, rho === nothing ? Ptr{T}(C_NULL) : rho))
end

"""cloud_download!(c, chunk) -- parcels in upload order."""
function cloud_download!(c::Cloud{T}, ch) where {T}
    check(ccall((:st_cloud_download, lib), Cint,
                (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}),
                c.h, ch.d, ch.X, ch.Y, ch.Z, ch.U, ch.V, ch.W))
end

function count(c::Cloud)
    n = Ref{Int64}(0)
    check(ccall((:st_cloud_count, lib), Cint, (Ptr{Cvoid}, Ref{Int64}), c.h, n))
    Int(n[])
end

reserve!(c::Cloud, capacity::Integer) =
    check(ccall((:st_cloud_reserve, lib), Cint, (Ptr{Cvoid}, Int64), c.h, capacity))

"""inject!(c, d, X, Y, Z, U, V, W; rho) -- appends parcels (inlet injection)."""
function inject!(c::Cloud{T}, d::Vector{T}, X::Vector{T}, Y::Vector{T}, Z::Vector{T}, U::Vector{T},
                 V::Vector{T}, W::Vector{T}; rho::Union{Nothing,Vector{T}}=nothing) where {T}
    k = length(d)
    all(length(a) == k for a in (X, Y, Z, U, V, W)) || throw(ArgumentError("parcel arrays differ in length"))
    check(ccall((:st_cloud_inject, lib), Cint,
                (Ptr{Cvoid}, Int64, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}),
                c.h, k, d, X, Y, Z, U, V, W, rho === nothing ? Ptr{T}(C_NULL) : rho))
end

"""compact!(c) -- drops the parcels the channel outlet removed; returns the count left."""
function compact!(c::Cloud)
    n = Ref{Int64}(0)
    check(ccall((:st_cloud_compact, lib), Cint, (Ptr{Cvoid}, Ref{Int64}), c.h, n))
    Int(n[])
end

function stray_fraction(c::Cloud)
    f = Ref{Float64}(0.0)
    check(ccall((:st_cloud_stray_fraction, lib), Cint, (Ptr{Cvoid}, Ref{Float64}), c.h, f))
    f[]
end

function resort_count(c::Cloud)
    n = Ref{Int64}(0)
    check(ccall((:st_cloud_resort_count, lib), Cint, (Ptr{Cvoid}, Ref{Int64}), c.h, n))
    Int(n[])
end

# ---- Eulerian fields: the reference's TwoWayEulerian ----------------------------------

mutable struct Field{T<:Union{Float32,Float64}}
    h::Ptr{Cvoid}
    ncell::Int
end

"""field_create(ctx, mesh, T) -- mesh.N, mesh.origin, mesh.ending as in construct_mesh."""
function field_create(ctx::Context, mesh, ::Type{T}) where {T<:Union{Float32,Float64}}
    m = Ref(StMesh(Tuple(Int32.(mesh.N)), Tuple(Float64.(mesh.origin)), Tuple(Float64.(mesh.ending))))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_field_create, lib), Cint, (Ptr{Cvoid}, Ref{StMesh}, Int32, Ref{Ptr{Cvoid}}),
                ctx.h, m, precision_of(T), h))
    Field{T}(h[], prod(Int.(mesh.N)))
end
function destroy!(f::Field)
    f.h == C_NULL || check(ccall((:st_field_destroy, lib), Cint, (Ptr{Cvoid},), f.h))
    f.h = C_NULL
    nothing
end

# U and UTrans are AoS xyz per cell, x fastest: a Vector of 3-element vectors (the
# reference's VectorField) or a flat vector of 3 * ncell numbers.
_check_len(f::Field, A, what) =
    sizeof(A) == 3 * f.ncell * sizeof(_elt(A)) || throw(ArgumentError("$what holds $(sizeof(A)) bytes, not 3 x $(f.ncell) values"))
_elt(A::AbstractArray{<:Real}) = eltype(A)
_elt(A::AbstractArray) = eltype(eltype(A))

function upload_U!(f::Field, U)
    _check_len(f, U, "U")
    check(ccall((:st_field_upload_U, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32), f.h, U, sizeof(_elt(U))))
end
zero_source!(f::Field) = check(ccall((:st_field_zero_source, lib), Cint, (Ptr{Cvoid},), f.h))
merge_source!(f::Field) = check(ccall((:st_field_merge_source, lib), Cint, (Ptr{Cvoid},), f.h))
function download_source!(f::Field, S)
    _check_len(f, S, "UTrans")
    check(ccall((:st_field_download_source, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32), f.h, S, sizeof(_elt(S))))
end

# ---- the parcel step ------------------------------------------------------------------

"""evolve(ctx, c, f, Δt; nsub=10, p=params()) -- nsub substeps of Δt / nsub; returns
(min xyz, max xyz) of the new positions (the reference's chunk bounding box)."""
function evolve(ctx::Context, c::Cloud, f::Field, Δt; nsub::Integer=10, p::StParams=params())
    bb = zeros(6)
    check(ccall((:st_evolve, lib), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Int32, Ref{StParams}, Ptr{Float64}),
                ctx.h, c.h, f.h, Δt, nsub, Ref(p), bb))
    (bb[1:3], bb[4:6])
end

"""evolve_noalloc(ctx, c, f, Δt; ...) -- the same step without device allocation (for
code under graph capture); the cloud must already be sorted (sort!)."""
function evolve_noalloc(ctx::Context, c::Cloud, f::Field, Δt; nsub::Integer=10, p::StParams=params(),
                        bbox::Bool=false)
    bb = zeros(6)
    check(ccall((:st_evolve_noalloc, lib), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Int32, Ref{StParams}, Ptr{Float64}),
                ctx.h, c.h, f.h, Δt, nsub, Ref(p), bbox ? bb : Ptr{Float64}(C_NULL)))
    bbox ? (bb[1:3], bb[4:6]) : nothing
end

sort!(ctx::Context, c::Cloud, f::Field; p::StParams=params()) =
    check(ccall((:st_cloud_sort, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{StParams}),
                ctx.h, c.h, f.h, Ref(p)))

function last_kernel_ms(ctx::Context)
    ms = Ref{Float64}(0.0)
    n = Ref{Int32}(0)
    check(ccall((:st_last_kernel_ms, lib), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Int32}), ctx.h, ms, n))
    (ms[], Int(n[]))
end

"""evolve!(ctx, c, f, chunk, eulerian, Δt) -- evolve!(chunk, eulerian, mesh, Δt, ::GPU)
(S:325-345): upload U, zero and fill UTrans, step, download UTrans."""
function evolve!(ctx::Context, c::Cloud, f::Field, eulerian, Δt; nsub::Integer=10, p::StParams=params())
    upload_U!(f, eulerian.U)
    zero_source!(f)
    bb = evolve(ctx, c, f, Δt; nsub=nsub, p=p)
    download_source!(f, eulerian.UTrans)
    bb
end

# ---- HIP graph of one fluid step ------------------------------------------------------

mutable struct Graph
    h::Ptr{Cvoid}
    keep::Any  # the cloud(s) and field(s) it steps stay referenced while it lives
end
function graph_capture(ctx::Context, c::Cloud, f::Field, Δt; nsub::Integer=10, p::StParams=params(),
                       zero_source::Bool=true)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_graph_capture, lib), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Int32, Ref{StParams}, Int32, Ref{Ptr{Cvoid}}),
                ctx.h, c.h, f.h, Δt, nsub, Ref(p), zero_source ? 1 : 0, h))
    Graph(h[], (c, f))
end
launch(g::Graph) = check(ccall((:st_graph_launch, lib), Cint, (Ptr{Cvoid},), g.h))
function destroy!(g::Graph)
    g.h == C_NULL || check(ccall((:st_graph_destroy, lib), Cint, (Ptr{Cvoid},), g.h))
    g.h = C_NULL
    g.keep = nothing
    nothing
end

# ---- batches of clouds (one persistent launch per pass) --------------------------------

mutable struct Batch
    h::Ptr{Cvoid}
    keep::Any
end
function batch_create(ctx::Context, clouds::Vector{<:Cloud}, fields::Vector{<:Field}; p::StParams=params(),
                      target_parcels_per_brick::Integer=4096, blocks_per_cu::Integer=0)
    length(clouds) == length(fields) || throw(ArgumentError("one field per cloud"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_batch_create, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ref{StParams}, Int32, Int32, Ref{Ptr{Cvoid}}),
                ctx.h, length(clouds), [c.h for c in clouds], [f.h for f in fields], Ref(p),
                target_parcels_per_brick, blocks_per_cu, h))
    Batch(h[], (clouds, fields))
end
evolve(b::Batch, Δt; nsub::Integer=10) =
    check(ccall((:st_batch_evolve, lib), Cint, (Ptr{Cvoid}, Float64, Int32), b.h, Δt, nsub))
sort!(b::Batch) = check(ccall((:st_batch_sort, lib), Cint, (Ptr{Cvoid},), b.h))
function graph_capture(b::Batch, Δt; nsub::Integer=10, zero_source::Bool=true)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_batch_graph_capture, lib), Cint, (Ptr{Cvoid}, Float64, Int32, Int32, Ref{Ptr{Cvoid}}),
                b.h, Δt, nsub, zero_source ? 1 : 0, h))
    Graph(h[], b)
end
function destroy!(b::Batch)
    b.h == C_NULL || check(ccall((:st_batch_destroy, lib), Cint, (Ptr{Cvoid},), b.h))
    b.h = C_NULL
    b.keep = nothing
    nothing
end

"""autotune_batch(ctx, clouds, fields; ...) -- the tuner's choice of parcels per brick,
block and workgroups per CU (pass to batch_create), cached in `cache` (JSON lines)."""
function autotune_batch(ctx::Context, clouds::Vector{<:Cloud}, fields::Vector{<:Field}; p::StParams=params(),
                        Δt=1e-3, trials::Integer=5, cache::AbstractString="")
    r = Ref{StTuneResult}()
    check(ccall((:st_autotune_batch, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ref{StParams}, Float64, Int32, Cstring,
                 Ref{StTuneResult}),
                ctx.h, length(clouds), [c.h for c in clouds], [f.h for f in fields], Ref(p), Δt, trials, cache, r))
    r[]
end

# ---- adjoint (FP64) ---------------------------------------------------------------------

mutable struct Adjoint
    h::Ptr{Cvoid}
    n::Int
    keep::Any
end
function adjoint_create(ctx::Context, c::Cloud{Float64}, f::Field{Float64}; nsub::Integer, K::Integer,
                        p::StParams=params())
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_adjoint_create, lib), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Ref{StParams}, Ref{Ptr{Cvoid}}),
                ctx.h, c.h, f.h, nsub, K, Ref(p), h))
    Adjoint(h[], count(c), (c, f))
end
forward!(a::Adjoint, Δt) = check(ccall((:st_adjoint_forward, lib), Cint, (Ptr{Cvoid}, Float64), a.h, Δt))
function source_misfit!(a::Adjoint, target::Vector{Float64})
    L = Ref{Float64}(0.0)
    check(ccall((:st_adjoint_source_misfit, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}), a.h, target, L))
    L[]
end
set_source_cotangent!(a::Adjoint, g::Vector{Float64}) =
    check(ccall((:st_adjoint_set_source_cotangent, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), a.h, g))
set_state_cotangent!(a::Adjoint, xbar, vbar) =
    check(ccall((:st_adjoint_set_state_cotangent, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), a.h,
                xbar === nothing ? Ptr{Float64}(C_NULL) : xbar, vbar === nothing ? Ptr{Float64}(C_NULL) : vbar))
backward!(a::Adjoint) = check(ccall((:st_adjoint_backward, lib), Cint, (Ptr{Cvoid},), a.h))
"""gradients(a; per_parcel_rho, gd, gρ) -- (dL/dd, dL/dρ or nothing, dL/dc_D); `gd` and
`gρ` may be preallocated (for instance with `pinned_array`) and are filled in place."""
function gradients(a::Adjoint; per_parcel_rho::Bool=false, gd::Vector{Float64}=zeros(a.n),
                   gρ::Union{Vector{Float64},Nothing}=per_parcel_rho ? zeros(a.n) : nothing)
    length(gd) == a.n || throw(ArgumentError("gd must have $(a.n) elements"))
    gρ === nothing || length(gρ) == a.n || throw(ArgumentError("gρ must have $(a.n) elements"))
    gc = Ref{Float64}(0.0)
    check(ccall((:st_adjoint_gradients, lib), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}, Ptr{Float64}, Ptr{Float64}),
                a.h, gd, gρ === nothing ? Ptr{Float64}(C_NULL) : gρ, gc, C_NULL, C_NULL))
    (gd, gρ, gc[])
end
function destroy!(a::Adjoint)
    a.h == C_NULL || check(ccall((:st_adjoint_destroy, lib), Cint, (Ptr{Cvoid},), a.h))
    a.h = C_NULL
    a.keep = nothing
    nothing
end

"""pinned_array(T, dims...) -- an uninitialised Array on page-locked host memory
(st_host_alloc), freed by the library when the array is collected.  Copies between it
and the device run by DMA at the bus's rate; pinning costs more than one transfer of
the same size, so allocate once and reuse."""
function pinned_array(::Type{T}, dims::Integer...) where {T}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:st_host_alloc, lib), Cint, (Int64, Ref{Ptr{Cvoid}}), max(prod(dims) * sizeof(T), 1), p))
    a = unsafe_wrap(Array, Ptr{T}(p[]), Tuple(dims); own=false)
    ptr = p[]
    finalizer(_ -> ccall((:st_host_free, lib), Cint, (Ptr{Cvoid},), ptr), a)
    a
end

"""source_misfit(ctx, c, f, target, Δt; nsub, K, p) -- L = Σ (UTrans - target)², the
loss whose rrule is defined below."""
function source_misfit(ctx::Context, c::Cloud{Float64}, f::Field{Float64}, target::Vector{Float64}, Δt;
                       nsub::Integer=10, K::Integer=nsub, p::StParams=params())
    a = adjoint_create(ctx, c, f; nsub=nsub, K=K, p=p)
    try
        forward!(a, Δt)
        source_misfit!(a, target)
    finally
        destroy!(a)
    end
end

if isdefined(Main, :ChainRulesCore)
    const CRC = Main.ChainRulesCore
    # Gradients with respect to the cloud's parameters: per-parcel d and ρ and the
    # drag-coefficient multiplier, returned as a named tuple in the tangent of `c`.
    function CRC.rrule(::typeof(source_misfit), ctx::Context, c::Cloud{Float64}, f::Field{Float64},
                       target::Vector{Float64}, Δt; nsub::Integer=10, K::Integer=nsub, p::StParams=params(),
                       per_parcel_rho::Bool=false)
        a = adjoint_create(ctx, c, f; nsub=nsub, K=K, p=p)
        local L
        try
            forward!(a, Δt)
            L = source_misfit!(a, target)
        catch
            destroy!(a)
            rethrow()
        end
        function source_misfit_pullback(L̄)
            # the adjoint pins the cloud and field: release it once the sweep is done
            gd, gρ, gc = try
                backward!(a)
                gradients(a; per_parcel_rho=per_parcel_rho)
            finally
                destroy!(a)
            end
            dc = CRC.Tangent{Cloud{Float64}}(; d=L̄ .* gd, rho=gρ === nothing ? CRC.NoTangent() : L̄ .* gρ,
                                             cd=L̄ * gc)
            (CRC.NoTangent(), CRC.NoTangent(), dc, CRC.NoTangent(), CRC.NoTangent(), CRC.NoTangent())
        end
        L, source_misfit_pullback
    end
end

end # module
