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
  effects/mass_density.jl:30-99                            -> mass_density
  tsc_interpolation/tsc_interpolation.jl:131-133           -> TSCInterpolation

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

"""
    mass_density(pos, mass; kernel=Cubic(Float64, 3), Nneighbors=32,
                 boxsize=zeros(3), verbose=false)

Drop-in for SPHtoGrid.mass_density (effects/mass_density.jl:30-99): SPH
density and smoothing length (distance to the (Nneighbors+1)-th nearest
particle, itself included) with the Dehnen & Aly bias correction; a boxsize
of zeros means non-periodic.  Returns `(rho, hsml)`.
"""
function mass_density(pos::Matrix{<:Real}, mass::Vector{<:Real};
                      kernel=Cubic(Float64, 3), Nneighbors::Integer=32,
                      boxsize::Vector{<:Real}=zeros(3), verbose::Bool=false)
    T = _common(pos, mass)
    x = convert(Matrix{T}, pos)
    m = convert(Vector{T}, mass)
    n = length(m)
    box = convert(Vector{Float64}, boxsize)
    rho = Vector{Float64}(undef, n)
    hsml = Vector{Float64}(undef, n)
    c = context()
    GC.@preserve x m box rho hsml begin
        rc = ccall((:s2g_mass_density, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{T}, Ptr{T}, Int64, Cint, Cint, Int64,
                    Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                   c.handle, _dtype(T), x, m, n, kernel_id(kernel),
                   Cint(kernel.dim in (2, 3) ? kernel.dim : 3), Int64(Nneighbors), box,
                   rho, hsml, C_NULL)
    end
    check(c, rc, "s2g_mass_density")
    return rho, hsml
end

"""
    TSCInterpolation(bin_q, pos_tsc, Npixels; average=true, periodic=false,
                     len2pix=1.0)

The grid deposit tsc_interpolation.jl:131-133 calls (that file is commented
out in SPHtoGrid.jl): textbook TSC of `bin_q` at `pos_tsc` (3 x N, centred)
onto an Npixels^3 grid in grid units x*len2pix + Npixels/2.  Returns the
grid [z, y, x] as Σ q W / Σ W (0 where no weight) with `average`, else Σ q W.
"""
function TSCInterpolation(bin_q, pos_tsc, Npixels::Integer; average::Bool=true,
                          periodic::Bool=false, len2pix::Real=1.0)
    T = _common(bin_q, pos_tsc)
    x = convert(Matrix{T}, pos_tsc)
    q = convert(Vector{T}, bin_q)
    nd = Npixels^3
    grid = Vector{Float64}(undef, 2 * nd)
    p = Ref(S2GParams(Npixels, Float64(len2pix)))
    c = context()
    GC.@preserve x q grid begin
        rc = ccall((:s2g_tsc3d, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{T}, Ptr{T}, Int64, Ref{S2GParams}, UInt32,
                    Ptr{Float64}, Ptr{Cvoid}),
                   c.handle, _dtype(T), x, q, length(q), p,
                   periodic ? S2G_FLAG_PERIODIC : UInt32(0), grid, C_NULL)
    end
    check(c, rc, "s2g_tsc3d")
    gq, gw = view(grid, 1:nd), view(grid, nd+1:2nd)
    out = average ? [w > 0 ? a / w : 0.0 for (a, w) in zip(gq, gw)] : collect(gq)
    return reshape(out, Npixels, Npixels, Npixels)
end

end # module
