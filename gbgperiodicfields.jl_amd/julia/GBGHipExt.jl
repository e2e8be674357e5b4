# GBGHipExt.jl -- the reference-side binding of libgbg: a package extension a
# GBGPeriodicFields.jl maintainer adds so that the package's public calls, with
# their signatures unchanged (src/GBGPeriodicFields.jl:281-430), run on the
# MI355X kernels for fields stored in libgbg device memory.  Not run in this
# repository (no Julia in the build image); the Python mirror exercises the
# same C ABI in the tests.
#
# How a solver reaches it -- nothing but the array family changes:
#
#   grid  = Grid3D((512, 512, 512), (2π, 2π, 2π))
#   û     = VectorFieldSpectral(GBGHipArray, grid)          # device storage
#   plans = FieldTransformPlans(GBGHipArray, grid)          # libgbg transform plans
#   m     = FieldSpectralMultipliers(GBGHipArray, grid)     # canonical multipliers
#   curl!(ω̂, û, m); to_physical!(u, û, plans, scratch); cross!(n, u, ω); ...
#
# The reference's own constructors build the plans and multipliers: they call
# plan_rfft / plan_irfft on one-component GBGHipArray prototypes (answered
# below with plans holding a libgbg plan) and `similar` + `copyto!` for the
# wavenumber vectors (src/Fields.jl:1431-1488, 1975-2041).  The public methods
# below are more specific than the reference's in exactly one argument each:
# the transform plans' FFT-plan type, the multipliers' array family, or a
# field's storage.  Every refusal is the reference's own check or exception.
module GBGHipExt

using GBGPeriodicFields
using GBGPeriodicFields: Grid3D, AbstractFieldPhysical, AbstractFieldSpectral, ScalarFieldPhysical,
    ScalarFieldSpectral, VectorFieldPhysical, VectorFieldSpectral, TensorFieldPhysical,
    TensorFieldSpectral, FieldTransformPlans, FieldSpectralMultipliers, RadialKernel, SpectralKernel,
    RealFourierLayout, field_data, field_grid, field_layout, field_tensor_order, radial_kernel
import GBGPeriodicFields: to_spectral!, to_physical!, gradient!, symmetric_gradient!, divergence!, curl!, inverse_curl!,
    laplacian!, inverse_laplacian!, helmholtz_projection!, convolve!, spectral_band,
    mean_free_projection!, cross!, multiply!, inner, exponential_convolve!, radial_spectrum,
    reality_residual, reality_projection!
import AbstractFFTs, Adapt, KernelAbstractions
import LinearAlgebra

# the reference's own refusals (argument checks), reused so errors match
const Fields = GBGPeriodicFields.Fields
const Conv = GBGPeriodicFields.Convolution

const libgbg = get(ENV, "GBG_LIB", joinpath(@__DIR__, "..", "libgbg.so"))

# ---- status -> the reference's exceptions -----------------------------------------
function check(status::Cint)
    status == 0 && return nothing
    msg = unsafe_string(ccall((:gbg_last_error, libgbg), Cstring, ()))
    status == 1 && throw(ArgumentError(msg))
    status == 2 && throw(DimensionMismatch(msg))
    status == 3 && throw(ArgumentError("not supported by libgbg: " * msg))
    status == 4 && throw(ArgumentError(msg))                     # scratch === source
    status == 5 && throw(OutOfMemoryError())
    throw(ErrorException("libgbg device error: " * msg))
end

# ---- the device array family ---------------------------------------------------------
# A column-major DenseArray over a libgbg device pointer: the reference's Array
# layout (src/Fields.jl:239, 369, 388), so every field type is built on it
# unchanged (`_allocated_field_storage` calls GBGHipArray{S}(undef, shape),
# src/Fields.jl:922-925).  Elements are not addressable from the host.
mutable struct GBGHipArray{T,N} <: DenseArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    function GBGHipArray{T,N}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:gbg_malloc, libgbg), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p,
                    prod(dims; init = 1) * sizeof(T)))
        a = new{T,N}(Ptr{T}(p[]), dims)
        finalizer(x -> ccall((:gbg_free, libgbg), Cint, (Ptr{Cvoid},), x.ptr), a)
        return a
    end
end
GBGHipArray{T}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N} = GBGHipArray{T,N}(undef, dims)
GBGHipArray{T}(::UndefInitializer, dims::Int...) where {T} = GBGHipArray{T}(undef, dims)
GBGHipArray(a::Array{T,N}) where {T,N} = copyto!(GBGHipArray{T,N}(undef, size(a)), a)
Base.Array(a::GBGHipArray{T,N}) where {T,N} = copyto!(Array{T,N}(undef, size(a)), a)

Base.size(a::GBGHipArray) = a.dims
Base.pointer(a::GBGHipArray) = a.ptr
Base.unsafe_convert(::Type{Ptr{T}}, a::GBGHipArray{T}) where {T} = a.ptr
Base.elsize(::Type{<:GBGHipArray{T}}) where {T} = sizeof(T)
Base.similar(::GBGHipArray, ::Type{S}, dims::Dims{M}) where {S,M} = GBGHipArray{S,M}(undef, dims)
Base.getindex(::GBGHipArray, ::Int...) =
    throw(ArgumentError("scalar indexing of libgbg device memory: copy to an Array first"))
Base.setindex!(::GBGHipArray, v, ::Int...) =
    throw(ArgumentError("scalar indexing of libgbg device memory: copy from an Array instead"))

memcpy(f, d, s, n) = check(ccall((f, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                                 d, s, n, C_NULL))
function Base.copyto!(d::GBGHipArray{T}, s::Array{T}) where {T}
    length(d) == length(s) || throw(DimensionMismatch("copyto! needs equal lengths"))
    memcpy(:gbg_memcpy_h2d, d.ptr, s, sizeof(s)); d
end
function Base.copyto!(d::Array{T}, s::GBGHipArray{T}) where {T}
    length(d) == length(s) || throw(DimensionMismatch("copyto! needs equal lengths"))
    memcpy(:gbg_memcpy_d2h, d, s.ptr, sizeof(d)); d
end
function Base.copyto!(d::GBGHipArray{T}, s::GBGHipArray{T}) where {T}
    length(d) == length(s) || throw(DimensionMismatch("copyto! needs equal lengths"))
    memcpy(:gbg_memcpy_d2d, d.ptr, s.ptr, length(s) * sizeof(T)); d
end
function Base.fill!(a::GBGHipArray{T}, x) where {T}
    iszero(x) || throw(ArgumentError("fill! of libgbg device memory supports zero only"))
    check(ccall((:gbg_memset, libgbg), Cint, (Ptr{Cvoid}, Cint, Csize_t, Ptr{Cvoid}),
                a.ptr, 0, length(a) * sizeof(T), C_NULL))
    a
end

# adapt / adapt! between the host and this family (src/Fields.jl:2810-2847)
Adapt.adapt_storage(::Type{<:GBGHipArray}, a::Array) = GBGHipArray(a)
Adapt.adapt_storage(::Type{Array}, a::GBGHipArray) = Array(a)

# The launch layer of the reference's generic paths.  Only the calls bound to
# libgbg below run on this family; a generic KernelAbstractions launch on it
# is refused by KernelAbstractions itself.
struct GBGBackend <: KernelAbstractions.GPU end
KernelAbstractions.get_backend(::GBGHipArray) = GBGBackend()

# ---- libgbg plans: one per (grid, box, precision, layout), cached ----------------------
mutable struct GBGPlan
    handle::Ptr{Cvoid}
end
const PLANS = Dict{Tuple{NTuple{3,Int},NTuple{3,Float64},Int,NTuple{3,Int}},GBGPlan}()
const PLANS_LOCK = ReentrantLock()

function plan_for(n::NTuple{3,Int}, L::NTuple{3,Float64}, ::Type{T},
                  dims::NTuple{3,Int}=(1, 2, 3)) where {T<:Union{Float32,Float64}}
    lock(PLANS_LOCK) do
        get!(PLANS, (n, L, sizeof(T), dims)) do
            h = Ref{Ptr{Cvoid}}(C_NULL)
            if dims == (1, 2, 3)
                check(ccall((:gbg_plan_create, libgbg), Cint,
                            (Ptr{Cint}, Ptr{Cdouble}, Cint, Ref{Ptr{Cvoid}}),
                            Cint[n...], Cdouble[L...], sizeof(T), h))
            else
                check(ccall((:gbg_plan_create_layout, libgbg), Cint,
                            (Ptr{Cint}, Ptr{Cdouble}, Cint, Ptr{Cint}, Ref{Ptr{Cvoid}}),
                            Cint[n...], Cdouble[L...], sizeof(T), Cint[dims...], h))
            end
            p = GBGPlan(h[])
            finalizer(x -> ccall((:gbg_plan_destroy, libgbg), Cint, (Ptr{Cvoid},), x.handle), p)
            p
        end
    end
end
plan_for(grid::Grid3D{T}, layout::RealFourierLayout=RealFourierLayout()) where {T} =
    plan_for(grid.grid_size, Float64.(grid.domain_size), T, layout.fft_dims)

ncomp(f) = 3^field_tensor_order(f)
dptr(f) = Ptr{Cvoid}(pointer(field_data(f)))

# ---- AbstractFFTs plans on one-component prototypes ------------------------------------
# FieldTransformPlans' constructors plan one component through
# plan_rfft(prototype, fft_dims) / plan_irfft(prototype, n, fft_dims)
# (src/Fields.jl:1463-1488).  The transform needs only the point counts, so
# these plans hold a libgbg plan of a 2 pi box of that shape.
struct GBGForwardPlan{T} <: AbstractFFTs.Plan{T}
    plan::GBGPlan
    n::NTuple{3,Int}
end
struct GBGInversePlan{T} <: AbstractFFTs.Plan{Complex{T}}
    plan::GBGPlan
    n::NTuple{3,Int}
end
# region = fft_dims: any permutation of (1, 2, 3), its first axis folded
function check_dims(region)
    r = Tuple(Int.(collect(region)))
    (length(r) == 3 && Fields._validate_fft_dims(r)) ||
        throw(ArgumentError("fft_dims must be a permutation of (1, 2, 3), got $(Tuple(region))"))
    r
end
function AbstractFFTs.plan_rfft(x::GBGHipArray{T,3}, region; kws...) where {T<:Union{Float32,Float64}}
    dims = check_dims(region)
    GBGForwardPlan{T}(plan_for(size(x), (2π, 2π, 2π), T, dims), size(x))
end
function AbstractFFTs.plan_irfft(x::GBGHipArray{Complex{T},3}, d::Integer, region;
                                 kws...) where {T<:Union{Float32,Float64}}
    dims = check_dims(region)
    n = ntuple(a -> a == dims[1] ? Int(d) : size(x, a), 3)
    GBGInversePlan{T}(plan_for(n, (2π, 2π, 2π), T, dims), n)
end
function LinearAlgebra.mul!(y::GBGHipArray{Complex{T},3}, p::GBGForwardPlan{T},
                            x::GBGHipArray{T,3}) where {T}
    check(ccall((:gbg_to_spectral, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                p.plan.handle, y.ptr, x.ptr, 1, C_NULL)); y
end
# c2r consumes its input, as the device FFT libraries' inverse does
function LinearAlgebra.mul!(y::GBGHipArray{T,3}, p::GBGInversePlan{T},
                            x::GBGHipArray{Complex{T},3}) where {T}
    check(ccall((:gbg_to_physical, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                p.plan.handle, y.ptr, x.ptr, 1, C_NULL, C_NULL)); y
end

# ---- transforms: every component in one call (src/Fields.jl:1589-1657) -----------------
const GBGTransformPlans{T} = FieldTransformPlans{T,<:GBGForwardPlan}

function to_spectral!(spectral::AbstractFieldSpectral{T,R}, physical::AbstractFieldPhysical{T,R},
                      plans::GBGTransformPlans{T}) where {T<:AbstractFloat,R}
    Fields._check_field_matches_context(spectral, plans)
    Fields._check_field_matches_context(physical, plans)
    check(ccall((:gbg_to_spectral, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                plan_for(plans.grid, plans.fourier_layout).handle, dptr(spectral), dptr(physical), ncomp(physical), C_NULL))
    spectral
end
function to_physical!(physical::AbstractFieldPhysical{T,R}, spectral::AbstractFieldSpectral{T,R},
                      plans::GBGTransformPlans{T}) where {T<:AbstractFloat,R}
    Fields._check_field_matches_context(physical, plans)
    Fields._check_field_matches_context(spectral, plans)
    check(ccall((:gbg_to_physical, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                plan_for(plans.grid, plans.fourier_layout).handle, dptr(physical), dptr(spectral), ncomp(spectral), C_NULL,
                C_NULL))
    physical
end
function to_physical!(physical::AbstractFieldPhysical{T,R}, spectral::AbstractFieldSpectral{T,R},
                      plans::GBGTransformPlans{T},
                      spectral_scratch::AbstractFieldSpectral{T,R}) where {T<:AbstractFloat,R}
    spectral_scratch === spectral &&
        throw(ArgumentError("spectral_scratch must not be the spectral source"))
    Fields._check_field_matches_context(physical, plans)
    Fields._check_field_matches_context(spectral, plans)
    Fields._check_field_matches_context(spectral_scratch, plans)
    check(ccall((:gbg_to_physical, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                plan_for(plans.grid, plans.fourier_layout).handle, dptr(physical), dptr(spectral), ncomp(spectral),
                dptr(spectral_scratch), C_NULL))
    physical
end

# ---- operators on libgbg multipliers (src/DifferentialOperators.jl) ---------------------
# Multipliers whose arrays live in this family; libgbg applies the canonical
# wavenumbers of the grid (its plan's FP64 tables), so caller-supplied arrays
# (Canonical = false) are refused rather than silently replaced.
const GBGMultipliers{T} = FieldSpectralMultipliers{T,<:GBGHipArray{T,1}}

function plan_of(m::GBGMultipliers, fields...)
    Fields._is_canonical(m) ||
        throw(ArgumentError("libgbg applies the canonical wavenumbers of the grid; these " *
                            "multipliers were built from caller-supplied arrays"))
    for f in fields
        Fields._check_multipliers_compatibility(f, m)
    end
    plan_for(m.grid, m.fourier_layout)
end

function gradient!(destination::VectorFieldSpectral{T}, field::ScalarFieldSpectral{T},
                   multipliers::GBGMultipliers{T}) where {T<:AbstractFloat}
    Fields._check_field_domain_compatibility(destination, field)
    p = plan_of(multipliers, field)
    check(ccall((:gbg_gradient, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                p.handle, dptr(destination), dptr(field), 0, C_NULL)); destination
end
function gradient!(destination::TensorFieldSpectral{2,T}, field::VectorFieldSpectral{T},
                   multipliers::GBGMultipliers{T}) where {T<:AbstractFloat}
    Fields._check_field_domain_compatibility(destination, field)
    p = plan_of(multipliers, field)
    check(ccall((:gbg_gradient, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                p.handle, dptr(destination), dptr(field), 1, C_NULL)); destination
end

function symmetric_gradient!(destination::TensorFieldSpectral{2,T}, field::VectorFieldSpectral{T},
                             multipliers::GBGMultipliers{T}) where {T<:AbstractFloat}
    Fields._check_field_domain_compatibility(destination, field)
    p = plan_of(multipliers, field)
    check(ccall((:gbg_symmetric_gradient, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                p.handle, dptr(destination), dptr(field), C_NULL)); destination
end

# A band kernel built on libgbg multipliers: spectral_band(m, lo, hi) and, through
# it, spectral_cutoff(m, kc) (src/Convolution.jl:1832-1882).  The symbol keeps the
# reference's per-mode definition for any generic use; libgbg receives the edges.
struct GBGBand{T} <: Function
    lo::Float64
    hi::Float64
    lo2::T
    hi2::T
end
(b::GBGBand{T})(squared) where {T} = ifelse(b.lo2 <= squared <= b.hi2, one(T), zero(T))
const GBGBandKernel{T} = RadialKernel{T,<:GBGBand}

function spectral_band(multipliers::GBGMultipliers{T}, wavenumber_low::Real,
                       wavenumber_high::Real) where {T<:AbstractFloat}
    (isfinite(wavenumber_low) && isfinite(wavenumber_high)) ||
        throw(ArgumentError("band edges must be finite, got ($wavenumber_low, $wavenumber_high)"))
    0 <= wavenumber_low <= wavenumber_high ||
        throw(ArgumentError("band edges must satisfy 0 <= low <= high, got ($wavenumber_low, $wavenumber_high)"))
    radial_kernel(GBGBand{T}(Float64(wavenumber_low), Float64(wavenumber_high),
                             T(wavenumber_low)^2, T(wavenumber_high)^2), multipliers)
end

band_args(::Nothing) = (Cint(0), 0.0, 0.0)
band_args(k::GBGBandKernel) = (Cint(1), k.sigma.lo, k.sigma.hi)
band_args(::SpectralKernel) =
    throw(ArgumentError("libgbg applies radial band kernels (spectral_band, spectral_cutoff built " *
                        "on its multipliers)"))

function div_call(destination, field, multipliers, kernel, α, β, order)
    Fields._check_field_domain_compatibility(destination, field)
    kernel === nothing || Conv._check_kernel_compatibility(field, kernel)
    p = plan_of(multipliers, field)
    hb, lo, hi = band_args(kernel)
    check(ccall((:gbg_divergence, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble,
                 Ptr{Cvoid}),
                p.handle, dptr(destination), dptr(field), order, hb, lo, hi, α, β, C_NULL))
    destination
end
divergence!(d::ScalarFieldSpectral{T}, f::VectorFieldSpectral{T}, m::GBGMultipliers{T}) where {T<:AbstractFloat} =
    div_call(d, f, m, nothing, 1.0, 0.0, 1)
divergence!(d::VectorFieldSpectral{T}, f::TensorFieldSpectral{2,T}, m::GBGMultipliers{T}) where {T<:AbstractFloat} =
    div_call(d, f, m, nothing, 1.0, 0.0, 2)
divergence!(d::ScalarFieldSpectral{T}, f::VectorFieldSpectral{T}, m::GBGMultipliers{T},
            k::SpectralKernel{T}) where {T<:AbstractFloat} = div_call(d, f, m, k, 1.0, 0.0, 1)
divergence!(d::ScalarFieldSpectral{T}, f::VectorFieldSpectral{T}, m::GBGMultipliers{T},
            k::SpectralKernel{T}, α::Real, β::Real) where {T<:AbstractFloat} = div_call(d, f, m, k, α, β, 1)
divergence!(d::VectorFieldSpectral{T}, f::TensorFieldSpectral{2,T}, m::GBGMultipliers{T},
            k::SpectralKernel{T}) where {T<:AbstractFloat} = div_call(d, f, m, k, 1.0, 0.0, 2)
divergence!(d::VectorFieldSpectral{T}, f::TensorFieldSpectral{2,T}, m::GBGMultipliers{T},
            k::SpectralKernel{T}, α::Real, β::Real) where {T<:AbstractFloat} = div_call(d, f, m, k, α, β, 2)

for (jl, c) in ((:curl!, :gbg_curl), (:inverse_curl!, :gbg_inverse_curl))
    @eval function $jl(destination::VectorFieldSpectral{T}, field::VectorFieldSpectral{T},
                       multipliers::GBGMultipliers{T}) where {T<:AbstractFloat}
        Fields._check_field_domain_compatibility(destination, field)
        p = plan_of(multipliers, field)
        check(ccall(($(QuoteNode(c)), libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    p.handle, dptr(destination), dptr(field), C_NULL))
        destination
    end
end

for (jl, c) in ((:laplacian!, :gbg_laplacian), (:inverse_laplacian!, :gbg_inverse_laplacian))
    @eval function $jl(destination::AbstractFieldSpectral{T,R}, field::AbstractFieldSpectral{T,R},
                       multipliers::GBGMultipliers{T}) where {T<:AbstractFloat,R}
        Fields._check_field_compatibility(destination, field)
        p = plan_of(multipliers, field)
        check(ccall(($(QuoteNode(c)), libgbg), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                    p.handle, dptr(destination), dptr(field), ncomp(field), C_NULL))
        destination
    end
end

function helm_call(destination, field, multipliers, kernel, α, β)
    Fields._check_field_domain_compatibility(destination, field)
    kernel === nothing || Conv._check_kernel_compatibility(field, kernel)
    p = plan_of(multipliers, field)
    hb, lo, hi = band_args(kernel)
    check(ccall((:gbg_helmholtz, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                p.handle, dptr(destination), dptr(field), hb, lo, hi, α, β, C_NULL))
    destination
end
helmholtz_projection!(d::VectorFieldSpectral{T}, f::VectorFieldSpectral{T}, m::GBGMultipliers{T}) where {T<:AbstractFloat} =
    helm_call(d, f, m, nothing, 1.0, 0.0)
helmholtz_projection!(d::VectorFieldSpectral{T}, f::VectorFieldSpectral{T}, m::GBGMultipliers{T},
                      k::SpectralKernel{T}) where {T<:AbstractFloat} = helm_call(d, f, m, k, 1.0, 0.0)
helmholtz_projection!(d::VectorFieldSpectral{T}, f::VectorFieldSpectral{T}, m::GBGMultipliers{T},
                      k::SpectralKernel{T}, α::Real, β::Real) where {T<:AbstractFloat} =
    helm_call(d, f, m, k, α, β)

# convolve!(dst, src, kernel[, α, β]) with a band kernel on libgbg multipliers
# (src/Convolution.jl:1035-1065)
function convolve!(destination::AbstractFieldSpectral{T,R}, field::AbstractFieldSpectral{T,R},
                   kernel::GBGBandKernel{T}, α::Real, β::Real) where {T<:AbstractFloat,R}
    Fields._check_field_compatibility(destination, field)
    Conv._check_kernel_compatibility(field, kernel)
    p = plan_of(kernel.multipliers, field)
    check(ccall((:gbg_convolve_band, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                p.handle, dptr(destination), dptr(field), ncomp(field), kernel.sigma.lo,
                kernel.sigma.hi, α, β, C_NULL))
    destination
end
convolve!(destination::AbstractFieldSpectral{T,R}, field::AbstractFieldSpectral{T,R},
          kernel::GBGBandKernel{T}) where {T<:AbstractFloat,R} =
    convolve!(destination, field, kernel, one(T), zero(T))

# exponential_convolve! (src/Convolution.jl:1750-1775); the offset is a stored
# spectral scalar field (a radial-kernel offset is refused)
offset_ptr(::Nothing) = C_NULL
offset_ptr(o::ScalarFieldSpectral) = dptr(o)
offset_ptr(::Any) = throw(ArgumentError("libgbg takes a stored spectral offset field"))

function exponential_convolve!(destination::AbstractFieldSpectral{T,R}, field::AbstractFieldSpectral{T,R},
                               multipliers::GBGMultipliers{T}, α::Real, β::Real; coefficient::Real,
                               offset = nothing, offset_scale::Real = one(T)) where {T<:AbstractFloat,R}
    Fields._check_field_compatibility(destination, field)
    p = plan_of(multipliers, field)
    check(ccall((:gbg_exponential_convolve, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Ptr{Cvoid}, Cdouble, Cdouble,
                 Cdouble, Ptr{Cvoid}),
                p.handle, dptr(destination), dptr(field), ncomp(field), coefficient,
                offset_ptr(offset), offset_scale, α, β, C_NULL))
    destination
end
exponential_convolve!(destination::AbstractFieldSpectral{T,R}, field::AbstractFieldSpectral{T,R},
                      multipliers::GBGMultipliers{T}; coefficient::Real, offset = nothing,
                      offset_scale::Real = one(T)) where {T<:AbstractFloat,R} =
    exponential_convolve!(destination, field, multipliers, one(T), zero(T); coefficient, offset,
                          offset_scale)
function exponential_convolve!(destination::AbstractFieldSpectral{T,R}, α::Real,
                               a::AbstractFieldSpectral{T,R}, β::Real, b::AbstractFieldSpectral{T,R},
                               multipliers::GBGMultipliers{T}; coefficient::Real, offset = nothing,
                               offset_scale::Real = one(T)) where {T<:AbstractFloat,R}
    Fields._check_field_compatibility(destination, a)
    Fields._check_field_compatibility(destination, b)
    p = plan_of(multipliers, a)
    check(ccall((:gbg_exponential_combine, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Cint, Cdouble,
                 Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
                p.handle, dptr(destination), α, dptr(a), β, dptr(b), ncomp(a), coefficient,
                offset_ptr(offset), offset_scale, C_NULL))
    destination
end

# radial_spectrum(f, m) / radial_spectrum(f, g, m) (src/Reductions.jl:1440-1455)
function radial_call(f, g, multipliers)
    p = plan_of(multipliers, f)
    n = Ref{Cint}(0); dk = Ref{Cdouble}(0.0)
    check(ccall((:gbg_radial_shells, libgbg), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cdouble}),
                p.handle, n, dk))
    spectrum = zeros(Float64, n[])
    check(ccall((:gbg_radial_spectrum, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Ptr{Cvoid}),
                p.handle, dptr(f), g === nothing ? C_NULL : dptr(g), ncomp(f), spectrum, n[], C_NULL))
    return (0:(n[] - 1)) .* dk[], spectrum
end
radial_spectrum(f::AbstractFieldSpectral{T}, multipliers::GBGMultipliers{T}) where {T<:AbstractFloat} =
    radial_call(f, nothing, multipliers)
function radial_spectrum(f::AbstractFieldSpectral{T}, g::AbstractFieldSpectral{T},
                         multipliers::GBGMultipliers{T}) where {T<:AbstractFloat}
    Fields._check_field_domain_compatibility(f, g)
    field_tensor_order(f) == field_tensor_order(g) ||
        throw(ArgumentError("radial_spectrum(f, g, multipliers) needs equal tensor orders " *
                            "(got $(field_tensor_order(f)) and $(field_tensor_order(g)))"))
    radial_call(f, g, multipliers)
end

# ---- calls without multipliers: dispatch on the field storage ---------------------------
const HipSpec{T} = Union{ScalarFieldSpectral{T,<:GBGHipArray},VectorFieldSpectral{T,<:GBGHipArray},
                         TensorFieldSpectral{2,T,<:GBGHipArray}}
const HipPhys{T} = Union{ScalarFieldPhysical{T,<:GBGHipArray},VectorFieldPhysical{T,<:GBGHipArray},
                         TensorFieldPhysical{2,T,<:GBGHipArray}}
plan_of(f) = plan_for(field_grid(f), field_layout(f))

function mean_free_projection!(field::HipSpec{T}) where {T<:AbstractFloat}
    check(ccall((:gbg_mean_free, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                plan_of(field).handle, dptr(field), ncomp(field), C_NULL)); field
end

# cross!(dst, u, v) (src/FieldAlgebra.jl:1088-1092) and the scalar product
# multiply!(dst, a, b) (src/FieldAlgebra.jl:778); both alias-safe
function cross!(destination::VectorFieldPhysical{T,<:GBGHipArray}, u::VectorFieldPhysical{T,<:GBGHipArray},
                v::VectorFieldPhysical{T,<:GBGHipArray}) where {T<:AbstractFloat}
    Fields._check_field_compatibility(destination, u)
    Fields._check_field_compatibility(destination, v)
    check(ccall((:gbg_cross, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                plan_of(u).handle, dptr(destination), dptr(u), dptr(v), C_NULL)); destination
end
function multiply!(destination::ScalarFieldPhysical{T,<:GBGHipArray}, a::ScalarFieldPhysical{T,<:GBGHipArray},
                   b::ScalarFieldPhysical{T,<:GBGHipArray}) where {T<:AbstractFloat}
    Fields._check_field_compatibility(destination, a)
    Fields._check_field_compatibility(destination, b)
    check(ccall((:gbg_multiply, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                plan_of(a).handle, dptr(destination), dptr(a), dptr(b), C_NULL)); destination
end

# inner(f, g): Parseval on spectral fields (src/Reductions.jl:883-925), the cell
# average on physical ones (:838-853); l2norm(f) = sqrt(inner(f, f)) (:947-948)
# reaches these through the reference's own definition.
function inner(f::HipSpec{T}, g::HipSpec{T}) where {T<:AbstractFloat}
    Fields._check_field_compatibility(f, g)
    out = Ref{Cdouble}(0.0)
    check(ccall((:gbg_inner_spectral, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}, Ptr{Cvoid}),
                plan_of(f).handle, dptr(f), dptr(g), ncomp(f), out, C_NULL))
    return T(out[])
end
function inner(f::HipPhys{T}, g::HipPhys{T}) where {T<:AbstractFloat}
    Fields._check_field_compatibility(f, g)
    out = Ref{Cdouble}(0.0)
    check(ccall((:gbg_inner_physical, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}, Ptr{Cvoid}),
                plan_of(f).handle, dptr(f), dptr(g), ncomp(f), out, C_NULL))
    return T(out[])
end

# reality_residual(f), reality_projection!(dst, src), reality_projection!(f)
# (src/Fields.jl:2918-3094)
function reality_residual(field::HipSpec{T}) where {T<:AbstractFloat}
    r = Ref{Cdouble}(0.0)
    check(ccall((:gbg_reality_residual, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}, Ptr{Cvoid}),
                plan_of(field).handle, dptr(field), ncomp(field), r, C_NULL))
    return T(r[])
end
function reality_projection!(destination::HipSpec{T}, source::HipSpec{T}) where {T<:AbstractFloat}
    destination === source || Fields._check_field_domain_compatibility(destination, source)
    check(ccall((:gbg_reality_projection, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                plan_of(source).handle, dptr(destination), dptr(source), ncomp(source), C_NULL))
    destination
end
reality_projection!(field::HipSpec) = reality_projection!(field, field)

# ---- beyond the reference's calls -------------------------------------------------------
"""
    rhs_uxw!(N̂, û, m, kc) -> (E, Z)

The dealiased rotational nonlinear term of a solver's right-hand side
(`to_velocity_vorticity!` + `rhs!`, examples/HIST_CPU.jl:229-247) in one call:
N̂ = Q M P F(F⁻¹(M û) × F⁻¹(∇×M û)) with M = spectral_cutoff(m, kc), and the
energy and enstrophy ½⟨u,u⟩, ½⟨ω,ω⟩ of the truncated fields.
"""
function rhs_uxw!(Nh::VectorFieldSpectral{T}, uh::VectorFieldSpectral{T}, m::GBGMultipliers{T},
                  kc::Real) where {T<:AbstractFloat}
    Fields._check_field_domain_compatibility(Nh, uh)
    p = plan_of(m, uh)
    E = Ref{Cdouble}(0.0); Z = Ref{Cdouble}(0.0)
    check(ccall((:gbg_uxw, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ref{Cdouble}, Ref{Cdouble}, Cuint, Ptr{Cvoid}),
                p.handle, dptr(uh), dptr(Nh), kc, E, Z, 0, C_NULL))
    return E[], Z[]
end

"""
    dealiased_product!(p̂, â, b̂, m, kc)

p̂ = M F(F⁻¹(M â) F⁻¹(M b̂)) for scalar fields, the consumer's composition of
convolve!/to_physical!/multiply!/to_spectral! fused into five passes.
"""
function dealiased_product!(ph::ScalarFieldSpectral{T}, ah::ScalarFieldSpectral{T},
                            bh::ScalarFieldSpectral{T}, m::GBGMultipliers{T}, kc::Real) where {T<:AbstractFloat}
    Fields._check_field_domain_compatibility(ph, ah)
    Fields._check_field_domain_compatibility(ph, bh)
    p = plan_of(m, ah)
    check(ccall((:gbg_dealiased_product, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
                p.handle, dptr(ph), dptr(ah), dptr(bh), kc, C_NULL))
    ph
end

"""
    autotune!(plans) -> ms

Per-axis pass choice for the grid of `plans`, the role FFTW's planner flags
play on the host (src/Fields.jl:1352-1362); returns forward + inverse ms.
"""
function autotune!(plans::GBGTransformPlans)
    ms = Ref{Cdouble}(0.0)
    check(ccall((:gbg_plan_autotune, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}),
                plan_for(plans.grid, plans.fourier_layout).handle, C_NULL, ms))
    return ms[]
end

export GBGHipArray, rhs_uxw!, dealiased_product!, autotune!

end # module
