# PTHip.jl — Julia binding of libpt_hip.so (C ABI: include/pt_hip.h).
#
# NOT EXECUTED: Julia is not installed where this library was built and
# tested.  The module is written against include/pt_hip.h, field for field;
# the Python mirror (particle_transport_amd/_lib.py) binds the same entry
# points and is what the test suite runs.
#
# Each wrapper replaces one call site of the reference's trajectory and
# lag-statistics path and keeps that function's argument meaning:
#
#   reference (file:line)                                   wrapper here
#   ------------------------------------------------------  -----------------------------------
#   CuArray uploads, phase_space_gpu_runner.jl:506-520      ctx_create, field_create
#   reference_B0_T, time_units.jl:18-20                     reference_B0_T, field_mean_abs_b
#   init_particles, phase_space_gpu_runner.jl:196-245       ens_create (host state) or
#                                                           ens_create_injected / inject! (device)
#   step loop + advance_particles_kernel!,                  advance!
#     phase_space_gpu_runner.jl:562-591;
#     run_mu_cache_ensemble step loop +
#     reconstruct_mu_step_kernel!,
#     multimode_multienergy_cache_pipeline.jl:1467-1498
#   reconstruct_mu_chunk_gpu, compute_delta_mu2_curve.jl:186-225
#                                                           reconstruct_mu_chunk_gpu!,
#                                                           reconstruct_mu_chunk_device!
#   read_particle_batch + CuArray,                          upload_particle_range!
#     compute_delta_mu2_dmumu_full.jl:411-437
#   process_mu_chunk_dmumu_cpu!,                            process_mu_chunk_dmumu!
#     compute_delta_mu2_dmumu_full.jl:656-704
#   process_momenta_chunk_dpp_cpu!, compute_dpp_full.jl:247 process_momenta_chunk_dpp!
#   fill_energy_snapshots!, compute_dpp_full.jl:290         fill_energy_snapshots!
#   (no counterpart: one campaign at a time,               serve_* (continuous batching)
#     run_mode_campaigns, pipeline:2924)
#
# Errors: every entry point returns a status; `check` turns a non-zero status
# into `error(message)`, the ErrorException the pipeline's per-product
# try/catch (multimode_multienergy_cache_pipeline.jl:2031-2059) records.
# Host arrays are preserved with GC.@preserve for the duration of each ccall.
module PTHip

const LIB = joinpath(@__DIR__, "..", "particle_transport_amd", "libpt_hip.so")

const PT_F64, PT_F32 = Cint(0), Cint(1)
const STORE_F64_PLANAR, STORE_F32_PLANAR, STORE_F32_BRICK, STORE_BF16_BRICK, STORE_F16_BRICK = Cint.(0:4)
const PUSH_REF_F64, PUSH_PLAIN_F32, PUSH_FAST = Cint.(0:2)
const OUT_NONE, OUT_PHASE, OUT_MU, OUT_PMAG, OUT_DPREL = Cint.(0:4)
const REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_RELEASED = Cint.(0:3)

abi_version() = ccall((:pt_abi_version, LIB), Cint, ())
last_error() = unsafe_string(ccall((:pt_last_error, LIB), Cstring, ()))
check(rc::Cint) = rc == 0 || error("libpt_hip: " * last_error())

dtype_code(::Type{Float64}) = PT_F64
dtype_code(::Type{Float32}) = PT_F32

# ---------------------------------------------------------------- structs --

# pt_advance_params
struct AdvanceParams
    dt_s::Float64
    charge_C::Float64
    mass_kg::Float64
    first_step::Int64        # 1-based si of the call's first step
    n_steps::Int64
    total_steps::Int64       # n_integration_steps: no push at si == total_steps
    save_stride::Int64       # saves at (si - 1) % stride == 0
    out_kind::Int32          # OUT_NONE / OUT_PHASE / OUT_MU
    steps_per_launch::Int32  # PUSH_FAST: pushes per kernel launch (0 = library default)
    sort_every_steps::Int64  # Morton re-sort period of split-precision ensembles (0 = never)
end

# pt_inject_params: the injection settings of init_particles
struct InjectParams
    seed::UInt64
    fixed_position::Int32
    fixed_mu::Int32
    position_m::NTuple{3,Float64}
    mu0::Float64
    p_magnitude::Float64     # pscale = gamma0 * M_P * v0 (phase_space_gpu_runner.jl:212)
end

# pt_serve_request
struct ServeRequest
    field_id::Int32
    out_kind::Int32          # OUT_MU (Float32), OUT_PMAG (Float64) or OUT_DPREL (Float32)
    n::Int64
    dt_s::Float64
    charge_C::Float64
    mass_kg::Float64
    total_steps::Int64
    save_stride::Int64
    periodic::Int32
    inject_fixed_mu::Int32
    mu0::Float64
    p_magnitude::Float64
    x::NTuple{3,Ptr{Float64}}
    p::NTuple{3,Ptr{Float64}}
    phase::Ptr{Float64}
    inject_on_device::Int32
    reserved0::Int32
    inject_first_index::Int64
    inject::InjectParams
    out_p_ref::Float64
end

# ---------------------------------------------------------------- context --

function ctx_create(device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pt_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h))
    return h[]
end
ctx_destroy(ctx) = check(ccall((:pt_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), ctx))

function bytes_in_use(ctx)
    b = Ref{Int64}(0)
    check(ccall((:pt_ctx_bytes_in_use, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), ctx, b))
    return b[]
end

function device_memory(ctx)
    f, t = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:pt_device_memory, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), ctx, f, t))
    return (free=f[], total=t[])
end

# device buffers owned by the context (caches kept in HBM, on_device = 1)
function buffer_alloc(ctx, bytes::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pt_buffer_alloc, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Ptr{Cvoid}}), ctx, bytes, p))
    return p[]
end
buffer_free(ctx, p, bytes::Integer) =
    check(ccall((:pt_buffer_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), ctx, p, bytes))
# direction: 0 host -> device, 1 device -> host, 2 device -> device
memcpy!(ctx, dst, src, bytes::Integer, direction::Integer) =
    check(ccall((:pt_memcpy, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint),
                ctx, dst, src, bytes, direction))

# Page-locked host arrays (hipHostMalloc) for caches kept in host memory: the
# copy engine reads them in place.  Release with host_free, not by the GC.
function host_alloc(::Type{T}, dims::Integer...) where {T}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pt_host_alloc, LIB), Cint, (Int64, Ref{Ptr{Cvoid}}), prod(dims) * sizeof(T), p))
    return unsafe_wrap(Array, Ptr{T}(p[]), Int.(dims); own=false)
end
host_free(a::Array) = check(ccall((:pt_host_free, LIB), Cint, (Ptr{Cvoid},), pointer(a)))

# read_particle_batch + CuArray (compute_delta_mu2_dmumu_full.jl:411-437;
# compute_dpp_full.jl:200-220) for a contiguous range: particles
# first:first+count-1 of a cadence cache (n, 3, nsave) or (n, nsave), best in
# host_alloc memory, to a device buffer holding (count, 3, nsave) or
# (count, nsave), by one strided copy of the copy engine.
function upload_particle_range!(ctx, dev::Ptr{Cvoid}, cache::Array{T}, first::Integer, count::Integer) where {T}
    n = size(cache, 1)
    (first >= 1 && count >= 0 && first + count - 1 <= n) || error("upload_particle_range!: range outside the cache")
    w = Int64(count) * sizeof(T)
    GC.@preserve cache begin
        check(ccall((:pt_memcpy_2d, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Cint),
                    ctx, dev, w, pointer(cache, first), Int64(n) * sizeof(T), w, div(length(cache), n), 0))
    end
    return nothing
end

# ----------------------------------------------------------------- fields --

# fields = (Bx, By, Bz, vx, vy, vz) as load_static_fields returns them
# (phase_space_gpu_runner.jl:95-123): tesla and m/s, each (nx, ny, nz).
function field_create(ctx, fields, box_length_m::Real; storage=STORE_F64_PLANAR)
    T = eltype(fields[1])
    nx, ny, nz = size(fields[1])
    ptrs = Ptr{Cvoid}[pointer(f) for f in fields]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve fields ptrs begin
        check(ccall((:pt_field_create, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Int64, Float64, Cint, Cint, Cint, Ptr{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}),
                    ctx, nx, ny, nz, box_length_m, storage, dtype_code(T), 0, ptrs, h))
    end
    return h[]
end
field_destroy(field) = check(ccall((:pt_field_destroy, LIB), Cint, (Ptr{Cvoid},), field))
# Optional cell table (192 bytes per cell): the production pushers then load a
# cell's coefficients on a cell change instead of rebuilding them; results are
# bit-identical.  No reference counterpart.
field_cell_table(field, enable::Bool) =
    check(ccall((:pt_field_cell_table, LIB), Cint, (Ptr{Cvoid}, Cint), field, enable ? 1 : 0))

function field_mean_abs_b(field)
    b = Ref{Float64}(0.0)
    check(ccall((:pt_field_mean_abs_b, LIB), Cint, (Ptr{Cvoid}, Ref{Float64}), field, b))
    return b[]
end

# reference_B0_T (time_units.jl:18-20), an FP64 reduction on the device
function reference_B0_T(ctx, Bx::AbstractArray{T}, By::AbstractArray{T}, Bz::AbstractArray{T}) where {T<:Union{Float32,Float64}}
    b = Ref{Float64}(0.0)
    GC.@preserve Bx By Bz begin
        check(ccall((:pt_mean_abs_b, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}),
                    ctx, length(Bx), dtype_code(T), 0, Bx, By, Bz, b))
    end
    return b[]
end

# -------------------------------------------------------------- ensembles --

function ens_create(ctx, field, x1, x2, x3, p1, p2, p3; pusher=PUSH_REF_F64, periodic::Bool=true)
    xs = Ptr{Cvoid}[pointer(x1), pointer(x2), pointer(x3)]
    ps = Ptr{Cvoid}[pointer(p1), pointer(p2), pointer(p3)]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x1 x2 x3 p1 p2 p3 xs ps begin
        check(ccall((:pt_ens_create, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Cint, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint,
                     Ref{Ptr{Cvoid}}),
                    ctx, field, pusher, length(x1), dtype_code(eltype(x1)), 0, xs, ps, periodic, h))
    end
    return h[]
end

function ens_create_injected(ctx, field, n::Integer, prm::InjectParams; pusher=PUSH_FAST, periodic::Bool=true)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pt_ens_create_injected, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Ref{InjectParams}, Cint, Ref{Ptr{Cvoid}}),
                ctx, field, pusher, n, prm, periodic, h))
    return h[]
end
ens_destroy(ens) = check(ccall((:pt_ens_destroy, LIB), Cint, (Ptr{Cvoid},), ens))

# init_particles on the device into a host (6, n) Float64 matrix [x1 x2 x3 p1 p2 p3]
function inject!(out::Matrix{Float64}, ctx, field, first_index::Integer, prm::InjectParams)
    size(out, 1) == 6 || error("inject!: out must be (6, n)")
    # the C layout is [x1 x2 x3 p1 p2 p3][n], i.e. the transpose of Julia's (6, n)
    tmp = Vector{Float64}(undef, 6 * size(out, 2))
    GC.@preserve tmp begin
        check(ccall((:pt_inject, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ref{InjectParams}, Cint, Ptr{Float64}),
                    ctx, field, first_index, size(out, 2), prm, 0, tmp))
    end
    out .= permutedims(reshape(tmp, size(out, 2), 6))
    return out
end

# Steps si = prm.first_step .. + prm.n_steps - 1.  out_a/out_b: Julia (n, 3,
# nsave_call) positions/momenta (OUT_PHASE) or (n, nsave_call) mu in out_a
# (OUT_MU); alive_counts[k] = sum(dalive) after step first_step + k - 1.
function advance!(ens, field, prm::AdvanceParams, out_a, out_b, alive_counts::Vector{Int64})
    nsaved = Ref{Int64}(0)
    T = out_a === nothing ? Float32 : eltype(out_a)
    cap = out_a === nothing ? 0 : size(out_a, ndims(out_a))
    pa = out_a === nothing ? C_NULL : pointer(out_a)
    pb = out_b === nothing ? C_NULL : pointer(out_b)
    GC.@preserve out_a out_b alive_counts begin
        check(ccall((:pt_ens_advance, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ref{AdvanceParams}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Int64,
                     Ptr{Int64}, Ref{Int64}),
                    ens, field, prm, pa, pb, dtype_code(T), 0, cap, alive_counts, nsaved))
    end
    return nsaved[]
end

function ens_state(ens, n::Integer)
    x = [Vector{Float64}(undef, n) for _ in 1:3]
    p = [Vector{Float64}(undef, n) for _ in 1:3]
    alive = Vector{UInt8}(undef, n)
    xp = Ptr{Float64}[pointer(v) for v in x]
    pp = Ptr{Float64}[pointer(v) for v in p]
    GC.@preserve x p alive xp pp begin
        check(ccall((:pt_ens_get_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{UInt8}),
                    ens, xp, pp, alive))
    end
    return x, p, alive .!= 0
end

ens_sort!(ens) = check(ccall((:pt_ens_sort, LIB), Cint, (Ptr{Cvoid},), ens))

function ens_last_timing(ens)
    ms, nl = Ref{Float64}(0.0), Ref{Int64}(0)
    check(ccall((:pt_ens_last_timing, LIB), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}), ens, ms, nl))
    return (kernel_ms=ms[], launches=nl[])
end

# ----------------------------------------------------- mu reconstruction --

# reconstruct_mu_chunk_gpu (compute_delta_mu2_curve.jl:186-225): `field` is a
# STORE_F32_PLANAR (compute_precision Float32) or STORE_F64_PLANAR field;
# positions/momenta are Julia (chunk, 3, nsave); returns (nsave, chunk).
function reconstruct_mu_chunk_gpu!(mu::Matrix{T}, ctx, field, positions::Array{S,3}, momenta::Array{S,3}) where {T<:Union{Float32,Float64},S<:Union{Float32,Float64}}
    chunk, _, nsave = size(positions)
    size(mu) == (nsave, chunk) || error("reconstruct_mu_chunk_gpu!: mu must be (nsave, chunk)")
    GC.@preserve mu positions momenta begin
        check(ccall((:pt_mu_from_phase, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}),
                    ctx, field, chunk, nsave, dtype_code(S), 0, positions, momenta, dtype_code(T), 0, mu))
    end
    return mu
end

# The same reconstruction from device phase space (upload_particle_range!)
# into a device buffer, sample-major (C [nsave][chunk]): the layout
# process_mu_cache_dmumu! walks with n = chunk.  μ stays in HBM between the
# two passes (run_combined_full, compute_delta_mu2_dmumu_full.jl:1131-1167).
function reconstruct_mu_chunk_device!(mu_dev::Ptr{Cvoid}, ctx, field, pos_dev::Ptr{Cvoid}, mom_dev::Ptr{Cvoid},
                                      chunk::Integer, nsave::Integer, ::Type{S}) where {S<:Union{Float32,Float64}}
    check(ccall((:pt_mu_from_phase_layout, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Cint),
                ctx, field, chunk, nsave, dtype_code(S), 1, pos_dev, mom_dev, PT_F32, 1, mu_dev, 1))
    return nothing
end

# ----------------------------------------------------------- statistics --

# process_mu_chunk_dmumu_cpu! (compute_delta_mu2_dmumu_full.jl:656-704): the
# same arguments after the context, accumulators updated in place.
function process_mu_chunk_dmumu!(ctx, mu_chunk::Matrix{Float32}, lag_steps::Vector{Int64}, mu_edges::Vector{Float64},
                                 particle_counts::Vector{Int64}, particle_means::Vector{Float64},
                                 particle_m2s::Vector{Float64}, pair_sum_squares::Vector{Float64},
                                 pair_counts::Vector{Int64}, dmumu_counts::Matrix{Int64},
                                 dmumu_sum_delta::Matrix{Float64}, dmumu_sum_delta2::Matrix{Float64},
                                 start_mode::Symbol, bin_abs::Bool)
    nsave, chunk = size(mu_chunk)
    mode = start_mode == :injection ? Cint(1) : start_mode == :sliding ? Cint(0) :
           error("Unknown dmumu_start_mode: $(start_mode)")
    GC.@preserve mu_chunk lag_steps mu_edges particle_counts particle_means particle_m2s pair_sum_squares pair_counts dmumu_counts dmumu_sum_delta dmumu_sum_delta2 begin
        check(ccall((:pt_dmumu_accumulate, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Float32}, Cint, Int64, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64,
                     Ptr{Float64}, Int64, Cint, Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
                    ctx, mu_chunk, 0, nsave, chunk, 1, nsave, C_NULL, lag_steps, length(lag_steps), mu_edges,
                    length(mu_edges) - 1, mode, bin_abs, particle_counts, particle_means, particle_m2s,
                    pair_sum_squares, pair_counts, dmumu_counts, dmumu_sum_delta, dmumu_sum_delta2))
    end
    return nothing
end

# The same pass over a device mu cache (n, nsave) written by advance! into a
# buffer_alloc buffer (C [save][n]), the first n_sel particles.
function process_mu_cache_dmumu!(ctx, mu_dev::Ptr{Cvoid}, nsave::Integer, n::Integer, n_sel::Integer,
                                 lag_steps::Vector{Int64}, mu_edges::Vector{Float64}, acc, start_mode::Symbol,
                                 bin_abs::Bool)
    mode = start_mode == :injection ? Cint(1) : Cint(0)
    (; particle_counts, particle_means, particle_m2s, pair_sum_squares, pair_counts, dmumu_counts,
       dmumu_sum_delta, dmumu_sum_delta2) = acc
    GC.@preserve lag_steps mu_edges acc begin
        check(ccall((:pt_dmumu_accumulate, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64,
                     Ptr{Float64}, Int64, Cint, Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
                    ctx, mu_dev, 1, nsave, n_sel, n, 1, C_NULL, lag_steps, length(lag_steps), mu_edges,
                    length(mu_edges) - 1, mode, bin_abs, particle_counts, particle_means, particle_m2s,
                    pair_sum_squares, pair_counts, dmumu_counts, dmumu_sum_delta, dmumu_sum_delta2))
    end
    return nothing
end

# process_mu_cache_dmumu! for callers that cannot allocate device memory
# (code under graph capture, serving loops): no counterpart in the reference.
# Device work areas live in the library's fixed 32 MiB array per GPU; limits
# and the race between overlapping calls on one GPU as in include/pt_hip.h.
function process_mu_cache_dmumu_noalloc!(ctx, mu_dev::Ptr{Cvoid}, nsave::Integer, n::Integer, n_sel::Integer,
                                         lag_steps::Vector{Int64}, mu_edges::Vector{Float64}, acc,
                                         start_mode::Symbol, bin_abs::Bool)
    mode = start_mode == :injection ? Cint(1) : Cint(0)
    (; particle_counts, particle_means, particle_m2s, pair_sum_squares, pair_counts, dmumu_counts,
       dmumu_sum_delta, dmumu_sum_delta2) = acc
    GC.@preserve lag_steps mu_edges acc begin
        check(ccall((:pt_dmumu_accumulate_noalloc, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64,
                     Ptr{Float64}, Int64, Cint, Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
                    ctx, mu_dev, nsave, n_sel, n, 1, C_NULL, lag_steps, length(lag_steps), mu_edges,
                    length(mu_edges) - 1, mode, bin_abs, particle_counts, particle_means, particle_m2s,
                    pair_sum_squares, pair_counts, dmumu_counts, dmumu_sum_delta, dmumu_sum_delta2))
    end
    return nothing
end

# process_momenta_chunk_dpp_cpu! (compute_dpp_full.jl:247): momenta Julia
# (chunk, 3, nsave) Float32, accumulators updated in place.
function process_momenta_chunk_dpp!(ctx, momenta::Array{Float32,3}, lag_steps::Vector{Int64}, p0::Float64,
                                    counts::Vector{Int64}, sum_delta_p::Vector{Float64}, sum_delta_p2::Vector{Float64},
                                    sum_delta_p_norm::Vector{Float64}, sum_delta_p_norm2::Vector{Float64})
    chunk, _, nsave = size(momenta)
    GC.@preserve momenta lag_steps counts sum_delta_p sum_delta_p2 sum_delta_p_norm sum_delta_p_norm2 begin
        check(ccall((:pt_dpp_accumulate, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Float32}, Cint, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Float64,
                     Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    ctx, momenta, 0, nsave, chunk, chunk, C_NULL, lag_steps, length(lag_steps), p0, counts,
                    sum_delta_p, sum_delta_p2, sum_delta_p_norm, sum_delta_p_norm2))
    end
    return nothing
end

# D_pp over a serving cache: |p| Float64 (OUT_PMAG) or |p|/p_ref - 1 Float32
# (OUT_DPREL), device pointer [save][n], the first n_sel particles.
function process_pmag_cache_dpp!(ctx, cache_dev::Ptr{Cvoid}, nsave::Integer, n::Integer, n_sel::Integer,
                                 lag_steps::Vector{Int64}, p0::Float64, acc; dprel_p_ref::Union{Nothing,Float64}=nothing)
    (; counts, sum_delta_p, sum_delta_p2, sum_delta_p_norm, sum_delta_p_norm2) = acc
    GC.@preserve lag_steps acc begin
        if dprel_p_ref === nothing
            check(ccall((:pt_dpp_accumulate_pmag, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Float64,
                         Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                        ctx, cache_dev, 1, nsave, n, n_sel, C_NULL, lag_steps, length(lag_steps), p0, counts,
                        sum_delta_p, sum_delta_p2, sum_delta_p_norm, sum_delta_p_norm2))
        else
            check(ccall((:pt_dpp_accumulate_dprel, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Float64,
                         Float64, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                        ctx, cache_dev, 1, nsave, n, n_sel, C_NULL, lag_steps, length(lag_steps), dprel_p_ref, p0,
                        counts, sum_delta_p, sum_delta_p2, sum_delta_p_norm, sum_delta_p_norm2))
        end
    end
    return nothing
end

# fill_energy_snapshots! (compute_dpp_full.jl:290-299): snapshot_indices are
# the reference's 1-based save indices, particle_offset its 1-based column.
function fill_energy_snapshots!(ctx, energy_snapshots::Matrix{Float64}, momenta::Array{Float32,3},
                                snapshot_indices::AbstractVector{<:Integer}, particle_offset::Integer)
    chunk, _, nsave = size(momenta)
    snap0 = Int64[i - 1 for i in snapshot_indices]
    GC.@preserve energy_snapshots momenta snap0 begin
        check(ccall((:pt_energy_snapshots, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Float32}, Cint, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64,
                     Ptr{Float64}, Int64),
                    ctx, momenta, 0, nsave, chunk, chunk, C_NULL, snap0, length(snap0), energy_snapshots,
                    particle_offset - 1))
    end
    return nothing
end

# ---------------------------------------------------------------- serving --

function serve_create(ctx, hbm_budget_bytes::Integer; steps_per_tick::Integer=200, max_active::Integer=1024)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pt_serve_create, LIB), Cint, (Ptr{Cvoid}, Int64, Int32, Int32, Ref{Ptr{Cvoid}}),
                ctx, hbm_budget_bytes, steps_per_tick, max_active, h))
    return h[]
end
serve_destroy(srv) = check(ccall((:pt_serve_destroy, LIB), Cint, (Ptr{Cvoid},), srv))

function serve_add_field(srv, field)
    id = Ref{Int32}(0)
    check(ccall((:pt_serve_add_field, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Int32}), srv, field, id))
    return id[]
end

# A request drawn on the device at admission (inject_on_device = 1): no host arrays.
function serve_submit(srv, req::ServeRequest)
    id = Ref{Int64}(0)
    check(ccall((:pt_serve_submit, LIB), Cint, (Ptr{Cvoid}, Ref{ServeRequest}, Ref{Int64}), srv, req, id))
    return id[]
end

function serve_step(srv)
    r, q, d = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:pt_serve_step, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}), srv, r, q, d))
    return (running=r[], queued=q[], done=d[])
end
serve_drain(srv) = check(ccall((:pt_serve_drain, LIB), Cint, (Ptr{Cvoid},), srv))

function serve_take_done(srv; capacity::Integer=1024)
    ids = Vector{Int64}(undef, capacity)
    cnt = Ref{Int64}(0)
    GC.@preserve ids begin
        check(ccall((:pt_serve_take_done, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64, Ref{Int64}),
                    srv, ids, capacity, cnt))
    end
    return ids[1:cnt[]]
end

function serve_poll(srv, id::Integer)
    st = Ref{Int32}(0)
    check(ccall((:pt_serve_poll, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Int32}), srv, id, st))
    return st[]
end

# (device cache pointer [save][n], nsave, n, particles alive at the end)
function serve_fetch(srv, id::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    ns, n, a = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:pt_serve_fetch, LIB), Cint,
                (Ptr{Cvoid}, Int64, Ref{Ptr{Cvoid}}, Ref{Int64}, Ref{Int64}, Ref{Int64}), srv, id, p, ns, n, a))
    return (cache=p[], nsave=ns[], n=n[], alive_final=a[])
end

# The cache in the reference's Julia layout (n, nsave).
function serve_copy_cache!(dst::Matrix{T}, srv, id::Integer) where {T<:Union{Float32,Float64}}
    GC.@preserve dst begin
        check(ccall((:pt_serve_copy_cache, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{T}, Int64),
                    srv, id, dst, sizeof(dst)))
    end
    return dst
end

serve_release(srv, id::Integer) = check(ccall((:pt_serve_release, LIB), Cint, (Ptr{Cvoid}, Int64), srv, id))

function serve_stats(srv)
    b, ps, ms = Ref{Int64}(0), Ref{Int64}(0), Ref{Float64}(0.0)
    check(ccall((:pt_serve_stats, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Float64}), srv, b, ps, ms))
    return (bytes_reserved=b[], particle_steps=ps[], kernel_ms=ms[])
end

end # module
