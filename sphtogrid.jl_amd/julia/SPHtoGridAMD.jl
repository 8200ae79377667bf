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

`sphMapping` and `healpix_map` below keep the reference's surfaces
(cic_interpolation.jl:35-273, main.jl:92-227) around those engines, using
SPHtoGrid.jl's own centring, filtering and shell-sorting functions; with
`parallel = true` they make the same single GPU call on the calling process
instead of the reference's `@spawnat` fan-out (cic_interpolation.jl:171-215,
236-271), since one call already uses the whole GPU.

The module is included inside SPHtoGrid (it imports the parent module's
functions).  The shared library path is taken from ENV["SPH2GRID_LIB"] or the
default build location next to this file.
"""
module SPHtoGridAMD

using SPHKernels
import ..mappingParameters, ..center_particles, ..filter_particles_in_image,
       ..filter_sort_particles, ..HealpixMap, ..RingOrder, ..ustrip

const libsph2grid = get(ENV, "SPH2GRID_LIB",
                        joinpath(@__DIR__, "..", "sph2grid", "libsph2grid.so"))

const S2G_OK = Cint(0)
const S2G_F32 = Cint(0)
const S2G_F64 = Cint(1)
const S2G_FLAG_CALC_MEAN = UInt32(0x1)
const S2G_FLAG_EXACT_NORM = UInt32(0x2)
const S2G_FLAG_PERIODIC = UInt32(0x4)
const S2G_FLAG_REDUCE = UInt32(0x8)

# s2g_params of include/sph2grid.h
struct S2GParams
    npixels::Int64
    len2pix::Float64
end

# s2g_map_desc of include/sph2grid.h (device pointers)
struct S2GMapDesc
    dims::Int32
    nq::Int32
    n::Int64
    params::S2GParams
    kernel::Int32
    flags::UInt32
    d_pos::Ptr{Cvoid}
    d_hsml::Ptr{Cvoid}
    d_m::Ptr{Cvoid}
    d_rho::Ptr{Cvoid}
    d_bin_q::Ptr{Cvoid}
    d_weights::Ptr{Cvoid}
    d_out::Ptr{Float64}
end

# s2g_tuning of include/sph2grid.h (the 2D deposit's work split)
struct S2GTuning
    small_side::Int32
    small_item_min::Int32
    small_items_per_cu::Int32
    gather_item_min::Int32
    gather_items_per_cu::Int32
    gather_planes::Int32
    small_planes::Int32
    gather_item_max::Int32
end
S2GTuning(; small_side=34, small_item_min=256, small_items_per_cu=16, gather_item_min=64,
          gather_items_per_cu=16, gather_planes=4, small_planes=4, gather_item_max=4096) =
    S2GTuning(small_side, small_item_min, small_items_per_cu, gather_item_min,
              gather_items_per_cu, gather_planes, small_planes, gather_item_max)

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
    reduce_image_2D(image, npix, reduce_image)

SPHtoGrid.reduce_image_2D (reduce_image.jl:8-31) on the GPU: the flat image
(N_distr x (Nq+1), deposit order) to the `Array{Float64,3}` (npix, npix, Nq)
in Julia's [x, y, q] order, divided by the weights where they are > 0 iff
`reduce_image`.
"""
function reduce_image_2D(image::Matrix{Float64}, npix::Integer, reduce_image::Bool)
    nq = size(image, 2) - 1
    out = Array{Float64}(undef, npix, npix, nq)
    c = context()
    GC.@preserve image out begin
        rc = ccall((:s2g_reduce, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Int64, Cint, UInt32, Ptr{Float64}, Ptr{Float64}),
                   c.handle, Cint(2), Int64(npix), Cint(nq),
                   reduce_image ? S2G_FLAG_REDUCE : UInt32(0), image, out)
    end
    check(c, rc, "s2g_reduce")
    return out
end

"""
    reduce_image_3D(image, npix, reduce_image)

SPHtoGrid.reduce_image_3D (reduce_image.jl:39-55) on the GPU, with the
`image[:, 2] .= 1.0` of `reduce_image = false` (cic_interpolation.jl:230-232)
folded in: the cube [z, y, x].
"""
function reduce_image_3D(image::Matrix{Float64}, npix::Integer, reduce_image::Bool)
    out = Array{Float64}(undef, npix, npix, npix)
    c = context()
    GC.@preserve image out begin
        rc = ccall((:s2g_reduce, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Int64, Cint, UInt32, Ptr{Float64}, Ptr{Float64}),
                   c.handle, Cint(3), Int64(npix), Cint(1),
                   reduce_image ? S2G_FLAG_REDUCE : UInt32(0), image, out)
    end
    check(c, rc, "s2g_reduce")
    return out
end

_strip(a) = eltype(a) <: AbstractFloat ? a : ustrip.(a)

"""
    sphMapping(Pos, HSML, M, Rho, Bin_Quant, Weights=Rho, RM=nothing;
               param, kernel, show_progress=true, parallel=false,
               reduce_image=true, return_both_maps=false, dimensions=2,
               calc_mean=false, stokes=false, sort_z=false)

SPHtoGrid.sphMapping (cic_interpolation.jl:35-273) with the deposit and the
reduce on the GPU.  The host steps are the reference's own: center_particles
(mutates `Pos`), `stokes` implies `sort_z` (:74-81), filter_particles_in_image
(with its sort_z selection), unit stripping (:95-111).  `parallel = true`
takes the same single GPU call on this process (the reference's Distributed
fan-out, :171-215 and :236-271, is not needed: one call uses the whole GPU).
`RM` has no effect on the maps, as in the reference, where `stokes` is not
passed on to cic_mapping_2D (:152-155) and the parallel path drops RM.
"""
function sphMapping(Pos::Array{<:Real}, HSML::Array{<:Real}, M::Array{<:Real},
                    Rho::Array{<:Real}, Bin_Quant::Array{<:Real},
                    Weights::Array{<:Real}=Rho,
                    RM::Union{Array{<:Real},Nothing}=nothing;
                    param::mappingParameters, kernel::AbstractSPHKernel,
                    show_progress::Bool=true, parallel::Bool=false,
                    reduce_image::Bool=true, return_both_maps::Bool=false,
                    dimensions::Int=2, calc_mean::Bool=false,
                    stokes::Bool=false, sort_z::Bool=false)
    Pos, par = center_particles(Pos, param)
    if stokes
        sort_z = true
        parallel && @warn "Mapping stokes parameters only works in serial!"
    end
    p_in_image = filter_particles_in_image(Pos, par, sort_z)
    x = _strip(Pos[:, p_in_image])
    hsml, m, rho, weights = (_strip(a[p_in_image]) for a in (HSML, M, Rho, Weights))
    bin_q = _strip(ndims(Bin_Quant) == 1 ? Bin_Quant[p_in_image] : Bin_Quant[:, p_in_image])
    if dimensions == 2
        image = cic_mapping_2D(x, hsml, m, rho, bin_q, weights; param=par, kernel=kernel,
                               calc_mean=calc_mean)
        return_both_maps && return image
        return reduce_image_2D(image, param.Npixels[1], reduce_image)
    elseif dimensions == 3
        # calc_mean is not forwarded to cic_mapping_3D (cic_interpolation.jl:219-221)
        image = cic_mapping_3D(x, hsml, m, rho, bin_q, weights; param=par, kernel=kernel)
        return reduce_image_3D(image, param.Npixels[1], reduce_image)
    end
    error("dimensions must be 2 or 3")
end

"""
    set_tuning!(t::Union{S2GTuning,Nothing}=nothing)

The work split of later 2D deposits (s2g_set_tuning); `nothing` restores the
library defaults.  It never changes a pixel weight's arithmetic.
"""
function set_tuning!(t::Union{S2GTuning,Nothing}=nothing)
    c = context()
    rc = if t === nothing
        ccall((:s2g_set_tuning, libsph2grid), Cint, (Ptr{Cvoid}, Ptr{S2GTuning}), c.handle,
              C_NULL)
    else
        ccall((:s2g_set_tuning, libsph2grid), Cint, (Ptr{Cvoid}, Ref{S2GTuning}), c.handle, t)
    end
    check(c, rc, "s2g_set_tuning")
    return nothing
end

"""
    deposit_batch!(descs::Vector{S2GMapDesc}; dtype=Float32, stream=C_NULL)

Many independent 2D maps in one device pass (s2g_deposit_batch): inputs and
outputs are device pointers (e.g. `pointer(::ROCArray)` from AMDGPU.jl), all
of element type `dtype`; see include/sph2grid.h for the descriptor fields.
"""
function deposit_batch!(descs::Vector{S2GMapDesc}; dtype::Type=Float32,
                        stream::Ptr{Cvoid}=C_NULL)
    c = context()
    GC.@preserve descs begin
        rc = ccall((:s2g_deposit_batch, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{S2GMapDesc}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                   c.handle, _dtype(dtype), descs, Cint(length(descs)), stream, C_NULL)
    end
    check(c, rc, "s2g_deposit_batch")
    return nothing
end

"""
    BatchPlan(descs::Vector{S2GMapDesc}; dtype=Float32, headroom=1.25)

A batch captured once in a hipGraph (s2g_plan_create); `run!(plan)` replays it
on the inputs' current contents (same device arrays), `close!(plan)` frees it.
A replay whose inputs outgrew the plan's capacities raises an error
(S2G_ECAPACITY; the maps are then incomplete).
"""
mutable struct BatchPlan
    handle::Ptr{Cvoid}
    descs::Vector{S2GMapDesc}
end

function BatchPlan(descs::Vector{S2GMapDesc}; dtype::Type=Float32, headroom::Real=1.25)
    c = context()
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve descs begin
        rc = ccall((:s2g_plan_create, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{S2GMapDesc}, Cint, Cdouble, Ptr{Ptr{Cvoid}}),
                   c.handle, _dtype(dtype), descs, Cint(length(descs)), Float64(headroom), h)
    end
    check(c, rc, "s2g_plan_create")
    plan = BatchPlan(h[], descs)
    finalizer(close!, plan)
    return plan
end

function run!(plan::BatchPlan; stream::Ptr{Cvoid}=C_NULL)
    plan.handle == C_NULL && error("the plan is closed")
    rc = ccall((:s2g_plan_run, libsph2grid), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
               plan.handle, stream, C_NULL)
    check(context(), rc, "s2g_plan_run")
    return nothing
end

function close!(plan::BatchPlan)
    if plan.handle != C_NULL
        ccall((:s2g_plan_destroy, libsph2grid), Cvoid, (Ptr{Cvoid},), plan.handle)
        plan.handle = C_NULL
    end
    return nothing
end

# the pixel vector of a map: a Vector{Float64} itself, or the `pixels` field
# of the HealpixMap{Float64,RingOrder} the reference allocates (main.jl:110-111)
_pixels(v::Vector{Float64}) = v
function _pixels(v::AbstractVector{Float64})
    hasproperty(v, :pixels) || error("expected a Vector{Float64} or a HealpixMap{Float64}")
    p = getproperty(v, :pixels)
    p isa Vector{Float64} || error("the map's pixels are not a Vector{Float64}")
    return p
end

"""
    healpix_deposit!(allsky_map, weight_map, pos, hsml, m, rho, bin_q, weights,
                     nside, kernel, calc_mean)

The particle loop of SPHtoGrid.healpix_map (main.jl:146-213) over the already
filtered and sorted particles (filter_particles.jl:17-54).  The maps are the
reference's two `HealpixMap{Float64,RingOrder}(Nside)` (main.jl:110-111) or
plain `Vector{Float64}`s of 12*nside^2 RING-ordered pixels; both are
overwritten with the deposit (the reference fills freshly allocated, zeroed
maps).
"""
function healpix_deposit!(allsky_map::AbstractVector{Float64},
                          weight_map::AbstractVector{Float64},
                          pos, hsml, m, rho, bin_q, weights, nside::Integer, kernel,
                          calc_mean::Bool)
    am, wm = _pixels(allsky_map), _pixels(weight_map)
    length(am) == length(wm) == 12 * Int64(nside)^2 ||
        error("maps must hold 12 * nside^2 = $(12 * Int64(nside)^2) pixels")
    T = Float64  # healpix_map is Float64-only in the reference (pixel_weights.jl:87-91)
    x = convert(Matrix{T}, pos)
    hs, mm, r, bq, w = (convert(Vector{T}, a) for a in (hsml, m, rho, bin_q, weights))
    c = context()
    flags = calc_mean ? S2G_FLAG_CALC_MEAN : UInt32(0)
    GC.@preserve x hs mm r bq w am wm begin
        rc = ccall((:s2g_healpix_deposit, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Int64,
                    Int64, Cint, Cint, UInt32, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                   c.handle, S2G_F64, x, hs, mm, r, bq, w, Int64(length(hs)), Int64(nside),
                   kernel_id(kernel), Cint(kernel.dim in (2, 3) ? kernel.dim : 2), flags,
                   am, wm, C_NULL)
    end
    check(c, rc, "s2g_healpix_deposit")
    return allsky_map, weight_map
end

"""
    healpix_deposit_maps!(allsky_maps, weight_maps, pos, hsml, m, rho, bin_qs, weightss,
                          nside, kernel, calc_mean)

Several all-sky maps of one particle set in one call (s2g_healpix_deposit_maps,
at most 8): map k is what `healpix_deposit!` would give for `bin_qs[k]` and
`weightss[k]` (e.g. the ρ map with weights 1 and the ρ-weighted T map), but
the disc queries, the kernel normalisation and the ring lists are computed
once for the particle set.  Every map is overwritten.
"""
function healpix_deposit_maps!(allsky_maps::AbstractVector, weight_maps::AbstractVector,
                               pos, hsml, m, rho, bin_qs::AbstractVector,
                               weightss::AbstractVector, nside::Integer, kernel,
                               calc_mean::Bool)
    nm = length(allsky_maps)
    1 <= nm <= 8 || error("1 to 8 maps per call")
    length(weight_maps) == length(bin_qs) == length(weightss) == nm ||
        error("one weight map, quantity and weight vector per map")
    am = [_pixels(v) for v in allsky_maps]
    wm = [_pixels(v) for v in weight_maps]
    all(length(v) == 12 * Int64(nside)^2 for v in vcat(am, wm)) ||
        error("maps must hold 12 * nside^2 = $(12 * Int64(nside)^2) pixels")
    T = Float64
    x = convert(Matrix{T}, pos)
    hs, mm, r = (convert(Vector{T}, a) for a in (hsml, m, rho))
    bq = [convert(Vector{T}, a) for a in bin_qs]
    w = [convert(Vector{T}, a) for a in weightss]
    c = context()
    flags = calc_mean ? S2G_FLAG_CALC_MEAN : UInt32(0)
    GC.@preserve x hs mm r bq w am wm begin
        bqp, wp = [pointer(a) for a in bq], [pointer(a) for a in w]
        amp, wmp = [pointer(a) for a in am], [pointer(a) for a in wm]
        rc = ccall((:s2g_healpix_deposit_maps, libsph2grid), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Cint, Ptr{Ptr{T}},
                    Ptr{Ptr{T}}, Int64, Int64, Cint, Cint, UInt32, Ptr{Ptr{Float64}},
                    Ptr{Ptr{Float64}}, Ptr{Cvoid}),
                   c.handle, S2G_F64, x, hs, mm, r, Cint(nm), bqp, wp, Int64(length(hs)),
                   Int64(nside), kernel_id(kernel), Cint(kernel.dim in (2, 3) ? kernel.dim : 2),
                   flags, amp, wmp, C_NULL)
    end
    check(c, rc, "s2g_healpix_deposit_maps")
    return allsky_maps, weight_maps
end

"""
    healpix_map(Pos, Hsml, M, Rho, Bin_q, Weights; center=[0.0, 0.0, 0.0],
                radius_limits=[0.0, Inf], Nside=1024, kernel, show_progress=true,
                output_from_all_workers=false, calc_mean=true)

SPHtoGrid.healpix_map (main.jl:92-227) with its particle loop on the GPU:
the same early return (:114-121), `filter_sort_particles` (mutates `Pos`, as
the reference does) and two `HealpixMap{Float64,RingOrder}` results.
"""
function healpix_map(Pos, Hsml, M, Rho, Bin_q, Weights;
                     center::Vector{<:Real}=[0.0, 0.0, 0.0],
                     radius_limits::Vector{<:Real}=[0.0, Inf],
                     Nside::Integer=1024, kernel::AbstractSPHKernel,
                     show_progress::Bool=true, output_from_all_workers::Bool=false,
                     calc_mean::Bool=true)
    allsky_map = HealpixMap{Float64,RingOrder}(Nside)
    weight_map = HealpixMap{Float64,RingOrder}(Nside)
    if !calc_mean && iszero(sum(Bin_q))
        return allsky_map, weight_map
    end
    pos, hsml, m, rho, bin_q, weights = filter_sort_particles(Pos, Hsml, M, Rho, Bin_q, Weights,
                                                              center, radius_limits, calc_mean)
    healpix_deposit!(allsky_map, weight_map, pos, hsml, m, rho, bin_q, weights, Nside, kernel,
                     calc_mean)
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
