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

export propagate_distance, propagate, PPHit, PPSummary, HIPContext

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

"""
    propagate(ctx, source, target; seed=0, stream=0, max_steps=10, hits_cap=source.photons)

Propagate `source.photons` photons from a PhotonSource towards a
DetectionSphere.  Returns (Vector{PPHit}, PPSummary).
"""
function propagate(ctx::HIPContext, source::PhotonSource, target::DetectionSphere;
                   seed::Integer=0, stream::Integer=0, max_steps::Integer=10,
                   hits_cap::Integer=min(source.photons, 1 << 24), beta::Real=1.0)
    em = PPEmitter(Tuple(Float32.(source.position)), Float32(source.time),
                   Tuple(Float32.(source.direction)), Float32(beta), UInt64(source.photons),
                   profile_code(source.emission_profile), 0)
    sp = PPSphere(Tuple(Float32.(target.position)), Float32(target.radius))
    hits = Vector{PPHit}(undef, hits_cap)
    summ = Ref{PPSummary}()
    sk = spectrum_struct(source.spectrum)
    keep, spec = sk isa Tuple{PPSpectrum,Nothing} ? (nothing, sk[1]) : sk
    rc = GC.@preserve keep hits ccall((:pp_propagate, LIB), Cint,
        (Ptr{Cvoid}, Ref{PPEmitter}, Ref{PPSpectrum}, Ref{PPSphere}, UInt64, UInt32, UInt32,
         Ptr{PPHit}, UInt64, Ref{PPSummary}),
        ctx.handle, em, spec, sp, UInt64(seed), UInt32(stream), UInt32(max_steps),
        pointer(hits), UInt64(hits_cap), summ)
    rc == PP_ERR_OVERFLOW || check(rc)
    resize!(hits, Int(summ[].n_stored))
    return hits, summ[]
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
