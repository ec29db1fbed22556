# PhotonPropHIP.jl — Julia binding of libphotonprop.so (include/photonprop.h).
#
# Drop-in for the propagation entry points of PhotonProp.jl's
# PhotonPropagationCuda module (src/photon_prop_cuda.jl): the reference's own
# Medium / Spectral / Emission / Detection structs are accepted unchanged and
# `propagate_distance(distance, medium, n_ph_gen)` returns the same DataFrame
# columns (:588-644).  Compute runs in the HIP library through ccall.
#
# Usage, inside the reference's package namespace (after its src/medium.jl,
# src/spectrum.jl, src/emission.jl and src/detection.jl are included):
#     include("PhotonPropHIP.jl"); using .PhotonPropHIP
#     ENV["PHOTONPROP_LIB"] = "/path/to/libphotonprop.so"   # optional
module PhotonPropHIP

using StaticArrays
using DataFrames
using ..Medium
using ..Spectral
using ..Emission
using ..Detection

export propagate_distance, propagate, propagate_batch, PPHit, PPSummary, HIPContext,
       HIPServer, submit!, wait_done, take_result!, server_stats, PPServerStats

const LIB = get(ENV, "PHOTONPROP_LIB", joinpath(@__DIR__, "..", "lib", "libphotonprop.so"))
const C_VAC = 0.299792458  # m/ns

# ---- isbits mirrors of the C structs (layouts asserted in __init__) ----------

struct PPMediumTables
    lambda_min::Float32
    lambda_max::Float32
    n_points::UInt32
    reserved::UInt32
    n_phase::Ptr{Float32}
    abs_len::Ptr{Float32}
    sca_len::Ptr{Float32}
    g::Float32
    reserved2::Float32
end

struct PPSpectrum
    kind::UInt32
    n_knots::UInt32
    lo::Float32
    hi::Float32
    mono::Float32
    reserved::Float32
    knots::Ptr{Float32}
end

struct PPEmitter
    pos::NTuple{3,Float32}
    t0::Float32
    axis::NTuple{3,Float32}
    beta::Float32
    n_photons::UInt64
    profile::UInt32
    reserved::UInt32
end

struct PPSphere
    center::NTuple{3,Float32}
    radius::Float32
end

struct PPHit
    pos::NTuple{3,Float32}
    time::Float32
    dir::NTuple{3,Float32}
    wavelength::Float32
    init_dir::NTuple{3,Float32}
    weight::Float32
    dist::Float32
    request::UInt32
    photon_id::UInt64
end

struct PPRequest
    emitter::PPEmitter
    target::PPSphere
    seed::UInt64
    stream::UInt32
    max_steps::UInt32
end

struct PPServerConfig
    max_slots::UInt32
    max_tiles_per_request::UInt32
    hit_pool_records::UInt64
    idle_exit_us::UInt32
    device_wait_us::UInt32
    flags::UInt32
    reserved::UInt32
end

struct PPServerStats
    submitted::UInt64
    completed::UInt64
    in_flight::UInt64
    pending::UInt64
    kernel_launches::UInt64
    pool_records_free::UInt64
    tiles_published::UInt64
    grid_blocks::UInt64
    device_stops::UInt64
    kernel_running::UInt64
    host_pool::UInt64
end

struct PPSummary
    n_sim::UInt64
    n_steps::UInt64
    n_hits::UInt64
    n_stored::UInt64
    sum_w::Float64
    sum_wt::Float64
    sum_wt2::Float64
    status::Int32
    reserved::UInt32
end

const PP_ERR_OVERFLOW = -2
const PP_ERR_UNSUPPORTED = -5
const SPEC_MONO, SPEC_KNOTS = UInt32(0), UInt32(1)
const EMIT_CONE, EMIT_UNIFORM, EMIT_REF_ISOTROPIC = UInt32(0), UInt32(1), UInt32(2)

function __init__()
    @assert sizeof(PPHit) == 64 && sizeof(PPSummary) == 64 && sizeof(PPEmitter) == 48
    @assert sizeof(PPMediumTables) == 48 && sizeof(PPSpectrum) == 32
    @assert sizeof(PPRequest) == 80 && sizeof(PPServerConfig) == 32
    @assert sizeof(PPServerStats) == 88
end

last_error() = unsafe_string(ccall((:pp_last_error, LIB), Cstring, ()))

function check(rc::Integer)
    rc >= 0 && return rc
    rc == PP_ERR_UNSUPPORTED && throw(ArgumentError(last_error()))   # as src/photon_prop_cuda.jl:67, :70
    error(last_error())
end

# ---- medium -> tables ------------------------------------------------------

# The reference kernel hard-codes g = 0.99 (src/photon_prop_cuda.jl:229).
hg_g(::MediumProperties) = 0.99f0

function medium_tables(medium::MediumProperties; range=(300f0, 800f0), n_points=1001)
    grid = range[1] .+ (0:n_points-1) .* ((range[2] - range[1]) / (n_points - 1))
    n = Float32[get_refractive_index(wl, medium) for wl in grid]
    a = Float32[get_absorption_length(wl, medium) for wl in grid]
    s = Float32[get_scattering_length(wl, medium) for wl in grid]
    return (Float32(range[1]), Float32(range[2]), n, a, s, Float32(hg_g(medium)))
end

mutable struct HIPContext
    handle::Ptr{Cvoid}
    medium::MediumProperties
    function HIPContext(medium::MediumProperties; device::Integer=0, kw...)
        lo, hi, n, a, s, g = medium_tables(medium; kw...)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve n a s begin
            t = PPMediumTables(lo, hi, UInt32(length(n)), 0, pointer(n), pointer(a), pointer(s), g, 0f0)
            check(ccall((:pp_create, LIB), Cint, (Ptr{Ptr{Cvoid}}, Cint, Ref{PPMediumTables}),
                        h, device, t))
        end
        ctx = new(h[], medium)
        finalizer(c -> ccall((:pp_destroy, LIB), Cint, (Ptr{Cvoid},), c.handle), ctx)
        return ctx
    end
end

# ---- spectrum / emission dispatch (same tags as the reference) -------------

spectrum_struct(s::Monochromatic) = (PPSpectrum(SPEC_MONO, 0, 0f0, 0f0, Float32(s.wavelength), 0f0, C_NULL), nothing)
function spectrum_struct(s::CherenkovSpectrum)
    k = Vector{Float32}(collect(s.knots))
    return (k, PPSpectrum(SPEC_KNOTS, UInt32(length(k)), 0f0, 0f0, 0f0, 0f0, pointer(k)))
end
spectrum_struct(s::Spectrum) = throw(ArgumentError("Cannot initialize $(typeof(s))"))

profile_code(::AngularEmissionProfile{:IsotropicEmission,T}) where {T} = EMIT_REF_ISOTROPIC
profile_code(::AngularEmissionProfile{:CherenkovCone,T}) where {T} = EMIT_CONE
profile_code(::AngularEmissionProfile{:UniformSphere,T}) where {T} = EMIT_UNIFORM
profile_code(::AngularEmissionProfile{U,T}) where {U,T} = throw(ArgumentError("Cannot initialize $U"))

emitter_struct(source::PhotonSource, beta::Real) =
    PPEmitter(Tuple(Float32.(source.position)), Float32(source.time),
              Tuple(Float32.(source.direction)), Float32(beta), UInt64(source.photons),
              profile_code(source.emission_profile), 0)
sphere_struct(target::DetectionSphere) = PPSphere(Tuple(Float32.(target.position)), Float32(target.radius))

# (keep-alive, PPSpectrum) for a spectrum; the first element must outlive the call
function spectrum_arg(s::Spectrum)
    sk = spectrum_struct(s)
    return sk isa Tuple{PPSpectrum,Nothing} ? (nothing, sk[1]) : sk
end

"""
    propagate(ctx, source, target; seed=0, stream=0, max_steps=10, hits_cap=source.photons)

Propagate `source.photons` photons from a PhotonSource towards a
DetectionSphere.  Returns (Vector{PPHit}, PPSummary).
"""
function propagate(ctx::HIPContext, source::PhotonSource, target::DetectionSphere;
                   seed::Integer=0, stream::Integer=0, max_steps::Integer=10,
                   hits_cap::Integer=min(source.photons, 1 << 24), beta::Real=1.0)
    em = emitter_struct(source, beta)
    sp = sphere_struct(target)
    hits = Vector{PPHit}(undef, hits_cap)
    summ = Ref{PPSummary}()
    keep, spec = spectrum_arg(source.spectrum)
    rc = GC.@preserve keep hits ccall((:pp_propagate, LIB), Cint,
        (Ptr{Cvoid}, Ref{PPEmitter}, Ref{PPSpectrum}, Ref{PPSphere}, UInt64, UInt32, UInt32,
         Ptr{PPHit}, UInt64, Ref{PPSummary}),
        ctx.handle, em, spec, sp, UInt64(seed), UInt32(stream), UInt32(max_steps),
        pointer(hits), UInt64(hits_cap), summ)
    rc == PP_ERR_OVERFLOW || check(rc)
    resize!(hits, Int(summ[].n_stored))
    return hits, summ[]
end

"""
    propagate_batch(ctx, sources, targets; seeds=0:…, streams=0, max_steps=10,
                    hits_cap=2^24, per_request_cap=0)

Many (source, target) pairs in one persistent launch (pp_propagate_batch; the
batching that src/photon_prop_cuda.jl:474-531 began).  All sources share the
first source's spectrum.  Returns (hits, offsets, summaries): pair r owns
hits[offsets[r]+1 : offsets[r+1]].
"""
function propagate_batch(ctx::HIPContext, sources::AbstractVector{<:PhotonSource},
                         targets::AbstractVector{<:DetectionSphere};
                         seeds=UInt64.(0:length(sources)-1), streams=zeros(UInt32, length(sources)),
                         max_steps::Integer=10, hits_cap::Integer=1 << 24,
                         per_request_cap::Integer=0, beta::Real=1.0)
    n = length(sources)
    @assert length(targets) == n
    reqs = [PPRequest(emitter_struct(sources[i], beta), sphere_struct(targets[i]),
                      UInt64(seeds[i]), UInt32(streams[i]), UInt32(max_steps)) for i in 1:n]
    out_cap = per_request_cap > 0 ? min(hits_cap, per_request_cap * n) : hits_cap
    hits = Vector{PPHit}(undef, out_cap)
    offsets = zeros(UInt64, n + 1)
    sums = Vector{PPSummary}(undef, n)
    keep, spec = spectrum_arg(sources[1].spectrum)
    rc = GC.@preserve keep reqs hits offsets sums ccall((:pp_propagate_batch, LIB), Cint,
        (Ptr{Cvoid}, Ptr{PPRequest}, UInt32, Ref{PPSpectrum}, Ptr{PPHit}, UInt64, UInt64,
         Ptr{UInt64}, Ptr{PPSummary}),
        ctx.handle, pointer(reqs), UInt32(n), spec, pointer(hits), UInt64(hits_cap),
        UInt64(per_request_cap), pointer(offsets), pointer(sums))
    rc == PP_ERR_OVERFLOW || check(rc)
    resize!(hits, Int(offsets[end]))
    return hits, offsets, sums
end

# ---- serving loop (pp_server_*) ---------------------------------------------

"""
    HIPServer(ctx, spectrum; max_slots=256, max_tiles_per_request=16384,
              hit_pool_records=2^24, idle_exit_us=20000, device_wait_us=0, host_pool=false,
              leave_room=false)

Continuous batching of propagation requests into a persistent kernel.  Queue
with `submit!`, poll with `wait_done`, collect with `take_result!`.
"""
mutable struct HIPServer
    handle::Ptr{Cvoid}
    ctx::HIPContext
    keep::Any
    function HIPServer(ctx::HIPContext, spectrum::Spectrum; max_slots::Integer=256,
                       max_tiles_per_request::Integer=16384, hit_pool_records::Integer=1 << 24,
                       idle_exit_us::Integer=20000, device_wait_us::Integer=0,
                       host_pool::Bool=false, leave_room::Bool=false)
        cfg = PPServerConfig(UInt32(max_slots), UInt32(max_tiles_per_request),
                             UInt64(hit_pool_records), UInt32(idle_exit_us),
                             UInt32(device_wait_us),
                             (host_pool ? UInt32(1) : UInt32(0)) | (leave_room ? UInt32(2) : UInt32(0)), 0)
        keep, spec = spectrum_arg(spectrum)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve keep check(ccall((:pp_server_create, LIB), Cint,
            (Ptr{Cvoid}, Ref{PPSpectrum}, Ref{PPServerConfig}, Ptr{Ptr{Cvoid}}),
            ctx.handle, spec, cfg, h))
        srv = new(h[], ctx, keep)
        finalizer(close, srv)
        return srv
    end
end

"""close(srv): drain and free the server now (before its context is freed)"""
function Base.close(srv::HIPServer)
    if srv.handle != C_NULL
        ccall((:pp_server_destroy, LIB), Cint, (Ptr{Cvoid},), srv.handle)
        srv.handle = C_NULL
    end
    return nothing
end

"""submit!(srv, source, target; seed, stream, max_steps=10, hits_cap=0) -> ticket (never blocks)"""
function submit!(srv::HIPServer, source::PhotonSource, target::DetectionSphere;
                 seed::Integer=0, stream::Integer=0, max_steps::Integer=10,
                 hits_cap::Integer=0, beta::Real=1.0)
    req = PPRequest(emitter_struct(source, beta), sphere_struct(target), UInt64(seed),
                    UInt32(stream), UInt32(max_steps))
    t = Ref{UInt64}(0)
    check(ccall((:pp_server_submit, LIB), Cint, (Ptr{Cvoid}, Ref{PPRequest}, UInt64, Ref{UInt64}),
                srv.handle, req, UInt64(hits_cap), t))
    return t[]
end

"""wait_done(srv, ticket; timeout_ms=10000) -> true when the request has finished"""
function wait_done(srv::HIPServer, ticket::Integer; timeout_ms::Integer=10000)
    done = Ref{Cint}(0)
    check(ccall((:pp_server_wait, LIB), Cint, (Ptr{Cvoid}, UInt64, UInt32, Ref{Cint}),
                srv.handle, UInt64(ticket), UInt32(timeout_ms), done))
    return done[] != 0
end

"""take_result!(srv, ticket; hits_cap=0) -> (hits, summary, latency_us); frees the ticket"""
function take_result!(srv::HIPServer, ticket::Integer; hits_cap::Integer=0)
    hits = Vector{PPHit}(undef, hits_cap)
    summ = Ref{PPSummary}()
    n = Ref{UInt64}(0)
    lat = Ref{Float64}(0.0)
    rc = GC.@preserve hits ccall((:pp_server_result, LIB), Cint,
        (Ptr{Cvoid}, UInt64, Ref{PPSummary}, Ptr{PPHit}, UInt64, Ref{UInt64}, Ref{Float64}),
        srv.handle, UInt64(ticket), summ, hits_cap > 0 ? pointer(hits) : C_NULL,
        UInt64(hits_cap), n, lat)
    rc == PP_ERR_OVERFLOW || check(rc)
    resize!(hits, Int(n[]))
    return hits, summ[], lat[]
end

"""server_stats(srv) -> PPServerStats (queue, kernel and pool counters; pp_server_stats)"""
function server_stats(srv::HIPServer)
    st = Ref{PPServerStats}()
    check(ccall((:pp_server_stats_get, LIB), Cint, (Ptr{Cvoid}, Ref{PPServerStats}), srv.handle, st))
    return st[]
end

const _ctx = Dict{UInt64,HIPContext}()

"""
    propagate_distance(distance, medium, n_ph_gen)

Drop-in for src/photon_prop_cuda.jl:588-644: source at the origin with
CherenkovSpectrum((300f0, 800f0), 20, medium) and :IsotropicEmission, a
DetectionSphere of radius 0.21 at (0, 0, distance).  Returns a DataFrame with
columns tres, initial_theta, ref_ix, abs_weight, dist_travelled, wavelength
(empty when no photon arrives).
"""
function propagate_distance(distance::Float32, medium::MediumProperties, n_ph_gen::Int64)
    ctx = get!(() -> HIPContext(medium), _ctx, objectid(medium))
    target_radius = 0.21f0
    source = PhotonSource(@SVector[0.0f0, 0.0f0, 0.0f0], @SVector[0.0f0, 0.0f0, 1.0f0], 0.0f0,
                          n_ph_gen, CherenkovSpectrum((300.0f0, 800.0f0), 20, medium),
                          AngularEmissionProfile{:IsotropicEmission,Float32}())
    target = DetectionSphere(@SVector[0.0f0, 0.0f0, distance], target_radius)
    hits, _ = propagate(ctx, source, target)
    if isempty(hits)
        println("No photons survived (distance=$distance)")
        return DataFrame()
    end
    wls = Float64[h.wavelength for h in hits]
    distt = Float64[h.dist for h in hits]
    ref_ix = get_refractive_index.(wls, Ref(medium))
    photon_times = distt ./ (C_VAC ./ ref_ix)
    tgeo = (distance - target_radius) ./ (C_VAC / get_refractive_index(800.0, medium))
    DataFrame(tres=photon_times .- tgeo, initial_theta=[acos(h.init_dir[3]) for h in hits],
              ref_ix=ref_ix, abs_weight=Float64[h.weight for h in hits],
              dist_travelled=distt, wavelength=wls)
end

end # module
