"""
    SPHtoGridAMD

Julia host side of libsph2grid: the deposition engines of SPHtoGrid.jl
(`cic_mapping_2D`, `cic_mapping_3D`, the particle loop of `healpix_map`) run
as HIP kernels on one MI355X through `ccall`.  Everything before the deposit
(centring, filtering, gathering the particle arrays) is SPHtoGrid.jl's own
code and is unchanged; this module only replaces the engine calls.

Binding points in SPHtoGrid.jl (src/ paths):
  cic_interpolation/cic_interpolation.jl:152-155, 191-196  -> cic_mapping_2D
  cic_interpolation/cic_interpolation.jl:219-221, 248-252  -> cic_mapping_3D
  healpix_interpolation/main.jl:146-213                    -> the particle loop

The shared library path is taken from ENV["SPH2GRID_LIB"] or the default
build location next to this file.
"""
module SPHtoGridAMD

using SPHKernels

const libsph2grid = get(ENV, "SPH2GRID_LIB",
                        joinpath(@__DIR__, "..", "sph2grid", "libsph2grid.so"))

const S2G_OK = Cint(0)
const S2G_F32 = Cint(0)
const S2G_F64 = Cint(1)
const S2G_FLAG_CALC_MEAN = UInt32(0x1)
const S2G_FLAG_EXACT_NORM = UInt32(0x2)
const S2G_FLAG_PERIODIC = UInt32(0x4)
const S2G_FLAG_REDUCE = UInt32(0x8)

struct S2GParams
    npixels::Int64
    len2pix::Float64
end

kernel_id(::Cubic) = Cint(0)
kernel_id(::Quintic) = Cint(1)
kernel_id(::WendlandC2) = Cint(2)
kernel_id(::WendlandC4) = Cint(3)
kernel_id(::WendlandC6) = Cint(4)
kernel_id(::WendlandC8) = Cint(5)
kernel_id(k) = error("kernel $(typeof(k)) is not supported by libsph2grid")

mutable struct Context
    handle::Ptr{Cvoid}
end

const _ctx = Ref{Union{Nothing,Context}}(nothing)

function context()
    if _ctx[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:s2g_create, libsph2grid), Cint, (Ptr{Ptr{Cvoid}}, Cint), h, 0)
        rc == S2G_OK || error("s2g_create failed with code $rc (no usable GPU?)")
        c = Context(h[])
        finalizer(c) do c
            ccall((:s2g_destroy, libsph2grid), Cvoid, (Ptr{Cvoid},), c.handle)
        end
        _ctx[] = c
    end
    return _ctx[]
end

function check(c::Context, rc::Cint, what::String)
    if rc != S2G_OK
        msg = unsafe_string(ccall((:s2g_last_error, libsph2grid), Cstring, (Ptr{Cvoid},),
                                  c.handle))
        error("$what failed ($rc): $msg")
    end
end

# one element type for all arrays of a call: Float64 if any is Float64
# (exact, since SPHtoGrid promotes every Float32 input to Float64 anyway)
_common(arrays...) = any(a -> eltype(a) == Float64, arrays) ? Float64 : Float32
_dtype(::Type{Float32}) = S2G_F32
_dtype(::Type{Float64}) = S2G_F64

"""
    cic_mapping_2D(Pos, HSML, M, Rho, Bin_Q, Weights, RM=nothing;
                   param, kernel, show_progress=false, calc_mean=true, stokes=false)

Drop-in for SPHtoGrid.cic_mapping_2D (cic_2D.jl:103-244): returns the flat
unreduced image `Matrix{Float64}(N_distr, Nq+1)` in deposit order, last column
= weights.  `RM` has no effect, as on the reference's mapping path.
"""
function cic_mapping_2D(Pos, HSML, M, Rho, Bin_Q, Weights, RM=nothing;
                        param, kernel, show_progress::Bool=false,
                        calc_mean::Bool=true, stokes::Bool=false)
    T = _common(Pos, HSML, M, Rho, Bin_Q, Weights)
    x = convert(Matrix{T}, Pos)
    hs, m, rho, w = (convert(Vector{T}, a) for a in (HSML, M, Rho, Weights))
    nq = ndims(Bin_Q) == 1 ? 1 : size(Bin_Q, 1)
    bq = convert(Array{T}, Bin_Q)
    n = length(m)
    npix = param.Npixels[1]
    image = zeros(Float64, npix * npix, nq + 1)
    p = Ref(S2GParams(npix, param.len2pix))
    flags = calc_mean ? S2G_FLAG_CALC_MEAN : UInt32(0)
    c = context()
    GC.@preserve x hs m rho bq w image begin
        rc = ccall((:s2g_deposit, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Cint,
                    Ptr{T}, Int64, Ref{S2GParams}, Cint, Cint, UInt32, Ptr{Float64},
                    Ptr{Cvoid}),
                   c.handle, Cint(2), _dtype(T), x, hs, m, rho, bq, Cint(nq), w, n, p,
                   kernel_id(kernel), Cint(kernel.dim in (2, 3) ? kernel.dim : 2), flags,
                   image, C_NULL)
    end
    check(c, rc, "s2g_deposit")
    return image
end

"""
    cic_mapping_3D(Pos, HSML, M, Rho, Bin_Q, Weights; param, kernel,
                   show_progress=false, calc_mean=false)

Drop-in for SPHtoGrid.cic_mapping_3D (cic_3D.jl:110-209): the flat image
`Matrix{Float64}(N_distr, 2)`.
"""
function cic_mapping_3D(Pos, HSML, M, Rho, Bin_Q, Weights; param, kernel,
                        show_progress::Bool=false, calc_mean=false)
    T = _common(Pos, HSML, M, Rho, Bin_Q, Weights)
    x = convert(Matrix{T}, Pos)
    hs, m, rho, bq, w = (convert(Vector{T}, a) for a in (HSML, M, Rho, Bin_Q, Weights))
    n = length(m)
    npix = param.Npixels[1]
    image = zeros(Float64, npix^3, 2)
    p = Ref(S2GParams(npix, param.len2pix))
    flags = calc_mean ? S2G_FLAG_CALC_MEAN : UInt32(0)
    c = context()
    GC.@preserve x hs m rho bq w image begin
        rc = ccall((:s2g_deposit, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Cint,
                    Ptr{T}, Int64, Ref{S2GParams}, Cint, Cint, UInt32, Ptr{Float64},
                    Ptr{Cvoid}),
                   c.handle, Cint(3), _dtype(T), x, hs, m, rho, bq, Cint(1), w, n, p,
                   kernel_id(kernel), Cint(kernel.dim in (2, 3) ? kernel.dim : 3), flags,
                   image, C_NULL)
    end
    check(c, rc, "s2g_deposit")
    return image
end

"""
    healpix_deposit!(allsky_map, weight_map, pos, hsml, m, rho, bin_q, weights,
                     nside, kernel, calc_mean)

The particle loop of SPHtoGrid.healpix_map (main.jl:146-213) over the already
filtered and sorted particles (filter_particles.jl:17-54); accumulates into
two RING-ordered Float64 vectors of length 12*nside^2.
"""
function healpix_deposit!(allsky_map::Vector{Float64}, weight_map::Vector{Float64},
                          pos, hsml, m, rho, bin_q, weights, nside::Integer, kernel,
                          calc_mean::Bool)
    T = Float64  # healpix_map is Float64-only in the reference (pixel_weights.jl:87-91)
    x = convert(Matrix{T}, pos)
    hs, mm, r, bq, w = (convert(Vector{T}, a) for a in (hsml, m, rho, bin_q, weights))
    c = context()
    flags = calc_mean ? S2G_FLAG_CALC_MEAN : UInt32(0)
    GC.@preserve x hs mm r bq w allsky_map weight_map begin
        rc = ccall((:s2g_healpix_deposit, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Int64,
                    Int64, Cint, Cint, UInt32, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                   c.handle, S2G_F64, x, hs, mm, r, bq, w, length(hs), Int64(nside),
                   kernel_id(kernel), Cint(kernel.dim in (2, 3) ? kernel.dim : 2), flags,
                   allsky_map, weight_map, C_NULL)
    end
    check(c, rc, "s2g_healpix_deposit")
    return allsky_map, weight_map
end

end # module
