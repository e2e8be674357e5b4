# GBGHipExt.jl -- the reference-side binding of libgbg (not run in this
# repository: it needs Julia and GBGPeriodicFields.jl, which the build image
# lacks).  It shows the `ccall` stubs a GBGPeriodicFields.jl maintainer adds so
# that the package's public calls on fields stored in device memory owned by
# libgbg dispatch to the MI355X kernels.  Names and argument meanings follow
# include/gbg.h; every status is rethrown as the reference's own exception.
module GBGHipExt

using GBGPeriodicFields
import GBGPeriodicFields: to_spectral!, to_physical!, gradient!, divergence!, curl!,
    inverse_curl!, laplacian!, inverse_laplacian!, helmholtz_projection!, convolve!,
    mean_free_projection!, cross!, multiply!, inner, l2norm, exponential_convolve!,
    radial_spectrum, reality_residual, reality_projection!

const libgbg = get(ENV, "GBG_LIB", joinpath(@__DIR__, "..", "libgbg.so"))

# ---- status -> exception -------------------------------------------------------
function check(status::Cint)
    status == 0 && return nothing
    msg = unsafe_string(ccall((:gbg_last_error, libgbg), Cstring, ()))
    status == 1 && throw(ArgumentError(msg))
    status == 2 && throw(DimensionMismatch(msg))
    status == 3 && throw(ErrorException("unsupported: " * msg))   # MethodError-class absence
    status == 4 && throw(ArgumentError(msg))                     # scratch === source
    status == 5 && throw(OutOfMemoryError())
    throw(ErrorException("libgbg device error: " * msg))
end

# ---- device array family ----------------------------------------------------------
# A column-major DenseArray over a libgbg device pointer: the layout of the
# reference's Array storage (src/Fields.jl:239, 369, 388), so fields built on it
# keep every shape and component-block convention.
mutable struct GBGHipArray{T,N} <: DenseArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    function GBGHipArray{T,N}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:gbg_malloc, libgbg), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, prod(dims) * sizeof(T)))
        a = new{T,N}(Ptr{T}(p[]), dims)
        finalizer(x -> ccall((:gbg_free, libgbg), Cint, (Ptr{Cvoid},), x.ptr), a)
        return a
    end
end
Base.size(a::GBGHipArray) = a.dims
Base.pointer(a::GBGHipArray) = a.ptr

# adapt / adapt!: host <-> device moves (src/Fields.jl:2810-2847)
function Base.copyto!(d::GBGHipArray{T}, s::Array{T}) where {T}
    check(ccall((:gbg_memcpy_h2d, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                d.ptr, s, sizeof(s), C_NULL))
    d
end
function Base.copyto!(d::Array{T}, s::GBGHipArray{T}) where {T}
    check(ccall((:gbg_memcpy_d2h, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                d, s.ptr, sizeof(d), C_NULL))
    d
end

# ---- plans: FieldTransformPlans + canonical FieldSpectralMultipliers ----------------
mutable struct GBGPlan
    handle::Ptr{Cvoid}
end
function GBGPlan(grid::Grid3D, ::Type{T}) where {T<:Union{Float32,Float64}}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    n = Cint[grid.grid_size...]
    L = Cdouble[grid.domain_size...]
    check(ccall((:gbg_plan_create, libgbg), Cint, (Ptr{Cint}, Ptr{Cdouble}, Cint, Ref{Ptr{Cvoid}}),
                n, L, sizeof(T), h))
    p = GBGPlan(h[])
    finalizer(x -> ccall((:gbg_plan_destroy, libgbg), Cint, (Ptr{Cvoid},), x.handle), p)
    return p
end

ncomp(f) = 3^field_tensor_order(f)
dptr(f) = Ptr{Cvoid}(pointer(field_data(f)))

# ---- transforms (src/Fields.jl:1589-1657) ---------------------------------------------
to_spectral!(dst::AbstractFieldSpectral, src::AbstractFieldPhysical, plan::GBGPlan) =
    (check(ccall((:gbg_to_spectral, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), ncomp(src), C_NULL)); dst)
to_physical!(dst::AbstractFieldPhysical, src::AbstractFieldSpectral, plan::GBGPlan) =
    (check(ccall((:gbg_to_physical, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), ncomp(src), C_NULL, C_NULL)); dst)
function to_physical!(dst::AbstractFieldPhysical, src::AbstractFieldSpectral, plan::GBGPlan,
                      scratch::AbstractFieldSpectral)
    scratch === src && throw(ArgumentError("scratch must not be the spectral source"))
    check(ccall((:gbg_to_physical, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                plan.handle, dptr(dst), dptr(src), ncomp(src), dptr(scratch), C_NULL))
    dst
end

# ---- operators (src/DifferentialOperators.jl) -----------------------------------------
gradient!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_gradient, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), field_tensor_order(src), C_NULL)); dst)
divergence!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_divergence, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble,
                  Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), field_tensor_order(src), 0, 0.0, 0.0, 1.0, 0.0,
                 C_NULL)); dst)
# divergence!(dst, src, m, spectral_band(m, klo, khi), α, β): band kernel and α/β fused
divergence!(dst, src, plan::GBGPlan, klo, khi, α = 1.0, β = 0.0) =
    (check(ccall((:gbg_divergence, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble,
                  Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), field_tensor_order(src), 1, klo, khi, α, β,
                 C_NULL)); dst)
curl!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_curl, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), C_NULL)); dst)
inverse_curl!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_inverse_curl, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), C_NULL)); dst)
laplacian!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_laplacian, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), ncomp(src), C_NULL)); dst)
inverse_laplacian!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_inverse_laplacian, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), ncomp(src), C_NULL)); dst)
helmholtz_projection!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_helmholtz, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), 0, 0.0, 0.0, 1.0, 0.0, C_NULL)); dst)
# helmholtz_projection!(dst, src, m, spectral_band(m, klo, khi), α, β) (:1589-1593)
helmholtz_projection!(dst, src, plan::GBGPlan, klo, khi, α = 1.0, β = 0.0) =
    (check(ccall((:gbg_helmholtz, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), 1, klo, khi, α, β, C_NULL)); dst)
# convolve!(dst, src, spectral_cutoff(m, kc)) -> band [0, kc]
convolve_band!(dst, src, plan::GBGPlan, klo, khi, α = 1.0, β = 0.0) =
    (check(ccall((:gbg_convolve_band, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), ncomp(src), klo, khi, α, β, C_NULL)); dst)
mean_free_projection!(f, plan::GBGPlan) =
    (check(ccall((:gbg_mean_free, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                 plan.handle, dptr(f), ncomp(f), C_NULL)); f)
cross!(dst, u, v, plan::GBGPlan) =
    (check(ccall((:gbg_cross, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(u), dptr(v), C_NULL)); dst)
multiply!(dst, a, b, plan::GBGPlan) =
    (check(ccall((:gbg_multiply, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(a), dptr(b), C_NULL)); dst)
function inner(f::AbstractFieldPhysical, g::AbstractFieldPhysical, plan::GBGPlan)
    out = Ref{Cdouble}(0.0)
    check(ccall((:gbg_inner_physical, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}, Ptr{Cvoid}),
                plan.handle, dptr(f), dptr(g), ncomp(f), out, C_NULL))
    return out[]
end
function inner(f::AbstractFieldSpectral, g::AbstractFieldSpectral, plan::GBGPlan)
    out = Ref{Cdouble}(0.0)
    check(ccall((:gbg_inner_spectral, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}, Ptr{Cvoid}),
                plan.handle, dptr(f), dptr(g), ncomp(f), out, C_NULL))
    return out[]
end

# l2norm(f) = sqrt(inner(f, f)) (src/Reductions.jl:947-948), either representation
l2norm(f, plan::GBGPlan) = sqrt(inner(f, f, plan))

# ---- exponential kernel (src/Convolution.jl:1750) ----------------------------------------
optr(offset) = offset === nothing ? C_NULL : dptr(offset)
exponential_convolve!(dst, f, plan::GBGPlan, α::Real = 1, β::Real = 0;
                      coefficient::Real, offset = nothing, offset_scale::Real = 1) =
    (check(ccall((:gbg_exponential_convolve, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Ptr{Cvoid}, Cdouble, Cdouble,
                  Cdouble, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(f), ncomp(f), coefficient, optr(offset),
                 offset_scale, α, β, C_NULL)); dst)
exponential_convolve!(dst, α::Real, a, β::Real, b, plan::GBGPlan;
                      coefficient::Real, offset = nothing, offset_scale::Real = 1) =
    (check(ccall((:gbg_exponential_combine, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Cint, Cdouble,
                  Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
                 plan.handle, dptr(dst), α, dptr(a), β, dptr(b), ncomp(a), coefficient,
                 optr(offset), offset_scale, C_NULL)); dst)

# ---- radial spectrum (src/Reductions.jl:1309-1455) ---------------------------------------
function radial_spectrum(f, g, plan::GBGPlan)
    n = Ref{Cint}(0); dk = Ref{Cdouble}(0.0)
    check(ccall((:gbg_radial_shells, libgbg), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cdouble}),
                plan.handle, n, dk))
    spectrum = zeros(Float64, n[])
    check(ccall((:gbg_radial_spectrum, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Ptr{Cvoid}),
                plan.handle, dptr(f), g === nothing ? C_NULL : dptr(g), ncomp(f), spectrum, n[],
                C_NULL))
    return (0:(n[] - 1)) .* dk[], spectrum
end
radial_spectrum(f, plan::GBGPlan) = radial_spectrum(f, nothing, plan)

# ---- reality condition (src/Fields.jl:2858-3094) -----------------------------------------
function reality_residual(f, plan::GBGPlan)
    r = Ref{Cdouble}(0.0)
    check(ccall((:gbg_reality_residual, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}, Ptr{Cvoid}),
                plan.handle, dptr(f), ncomp(f), r, C_NULL))
    return r[]
end
reality_projection!(dst, src, plan::GBGPlan) =
    (check(ccall((:gbg_reality_projection, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                 plan.handle, dptr(dst), dptr(src), ncomp(src), C_NULL)); dst)
reality_projection!(f, plan::GBGPlan) = reality_projection!(f, f, plan)

# ---- the consumer's dealiased product and the planner's role -----------------------------
# p^ = M F(F^-1(M a^) F^-1(M b^)) for scalar fields (fused on power-of-two grids)
dealiased_product!(ph, ah, bh, plan::GBGPlan, kc) =
    (check(ccall((:gbg_dealiased_product, libgbg), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
                 plan.handle, dptr(ph), dptr(ah), dptr(bh), kc, C_NULL)); ph)
# per-axis pass choice, the role of FFTW's planner flags; returns forward + inverse ms
function autotune!(plan::GBGPlan)
    ms = Ref{Cdouble}(0.0)
    check(ccall((:gbg_plan_autotune, libgbg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}),
                plan.handle, C_NULL, ms))
    return ms[]
end

# ---- the fused right-hand side of examples/HIST_CPU.jl:229-247 --------------------------
function rhs_uxw!(Nh, uh, plan::GBGPlan, kc)
    E = Ref{Cdouble}(0.0); Z = Ref{Cdouble}(0.0)
    check(ccall((:gbg_uxw, libgbg), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ref{Cdouble}, Ref{Cdouble}, Cuint, Ptr{Cvoid}),
                plan.handle, dptr(uh), dptr(Nh), kc, E, Z, 0, C_NULL))
    return E[], Z[]
end

end # module
