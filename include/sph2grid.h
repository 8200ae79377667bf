/*
 * sph2grid.h -- C ABI of libsph2grid, the MI355X (gfx950) implementation of
 * SPHtoGrid.jl's kernel-weighted particle deposition.
 *
 * The reference is a pure-Julia call graph with no FFI of its own; the entry
 * points below replace its deposition engines at the call sites a Julia
 * `ccall` would bind (paths are relative to the SPHtoGrid.jl source tree):
 *
 *   s2g_deposit          <- cic_mapping_2D  (cic_interpolation/cic_2D.jl:103-244),
 *                           called at cic_interpolation/cic_interpolation.jl:152-155
 *                           (serial) and :191-196 (per worker)
 *                        <- cic_mapping_3D  (cic_interpolation/cic_3D.jl:110-209),
 *                           called at cic_interpolation.jl:219-221 and :248-252
 *   s2g_reduce           <- reduce_image_2D / reduce_image_3D
 *                           (cic_interpolation/reduce_image.jl:8-31, 39-55)
 *   s2g_map              <- the deposit + reduce tail of sphMapping
 *                           (cic_interpolation.jl:148-272)
 *   s2g_healpix_deposit  <- the particle loop of healpix_map
 *                           (healpix_interpolation/main.jl:146-213)
 *   s2g_healpix_deposit_maps
 *                        <- several healpix_map calls on one particle set
 *   s2g_tsc3d            <- TSCInterpolation(..., average=true)
 *                           (tsc_interpolation/tsc_interpolation.jl:131-133,
 *                           commented out in the reference)
 *   s2g_deposit_batch    <- many sphMapping calls (no single reference entry;
 *                           the closest is distributed_mapping/cic.jl:24-110)
 *   s2g_mass_density     <- mass_density (effects/mass_density.jl:30-99)
 *   s2g_accumulate_device,
 *   s2g_healpix_reduce_device
 *                        <- the per-subfile summation and the final reduce of
 *                           distributed_cic_map / distributed_allsky_map
 *                           (distributed_mapping/cic.jl:63-73,
 *                           distributed_mapping/healpix.jl:50-73)
 *
 * Conventions
 *   - All arrays are plain pointers with explicit sizes; no ownership is
 *     transferred and no pointer is retained after a call returns.
 *   - Particle positions are 3 x n interleaved (Julia's 3 x N column-major
 *     `Pos`), already centred and filtered by the caller exactly as
 *     center_particles / filter_particles_in_image do
 *     (cic_interpolation/filter_shift.jl:6-67).
 *   - bin_q is nq x n with the quantity index fastest (Julia's Nq x N matrix),
 *     or n values when nq == 1.
 *   - A flat image is npix^dims x (nq+1) doubles, column-major, deposit order
 *     (2D index i*npix + j, 3D index i*npix*npix + j*npix + k, with i, j, k
 *     the x, y, z pixel indices); the last column holds the weights.
 *   - Return value: S2G_OK (0) or a negative S2G_E* code; the message is
 *     available from s2g_last_error().  No C++ exception crosses the ABI.
 *   - A context is not thread safe; use one per host thread.
 *   - *_device variants and s2g_deposit_batch take device pointers and a
 *     hipStream_t (as void*; NULL = the context's stream, a blocking stream,
 *     so work the caller queued on the null stream is ordered before it and
 *     work queued there afterwards waits for it).  The deposits
 *     (s2g_deposit_device, s2g_healpix_deposit_device, s2g_deposit_batch)
 *     size their work lists from device-side counts, so they synchronise the
 *     stream and return with the result complete; so does s2g_tsc3d_device
 *     (its particle order is sorted in the context's buffers);
 *     s2g_reduce_device only enqueues.
 */
#ifndef SPH2GRID_H
#define SPH2GRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define S2G_OK 0
#define S2G_EINVAL (-1)
#define S2G_EHIP (-2)
#define S2G_ENOMEM (-3)
#define S2G_EUNSUPPORTED (-4)
#define S2G_ECAPACITY (-5)  /* s2g_plan_run: the inputs outgrew the plan's capacities */

/* SPHKernels.jl kernel types (the `kernel::AbstractSPHKernel` argument of
 * sphMapping, cic_interpolation.jl:40).  Their normalisation and h^-dim
 * factor cancel in every deposit (pw = w V wk A / sum(wk A)), so kernel_dim
 * (Cubic(2) vs Cubic(3)) does not change results; it is validated only. */
enum s2g_kernel {
    S2G_KERNEL_CUBIC = 0,
    S2G_KERNEL_QUINTIC = 1,
    S2G_KERNEL_WENDLAND_C2 = 2,
    S2G_KERNEL_WENDLAND_C4 = 3,
    S2G_KERNEL_WENDLAND_C6 = 4,
    S2G_KERNEL_WENDLAND_C8 = 5
};

/* element type of every floating-point input array of one call */
enum s2g_dtype { S2G_F32 = 0, S2G_F64 = 1 };

/* flags */
#define S2G_FLAG_CALC_MEAN 0x1u  /* calc_mean keyword (cic_2D.jl:110, main.jl:99) */
#define S2G_FLAG_EXACT_NORM 0x2u /* always sum the footprint for the per-particle
                                    normalisation (no closed-form shortcut for
                                    large unclipped footprints) */
#define S2G_FLAG_PERIODIC 0x4u   /* TSC only: wrap cells around the grid */
#define S2G_FLAG_REDUCE 0x8u     /* s2g_map / s2g_reduce: reduce_image=true */
#define S2G_FLAG_FP32_ACCUM 0x10u /* 2D deposit: the tile gather accumulates each work
                                     item's pixels in FP32 (the items' partials are
                                     still summed in FP64); faster with three or four
                                     planes, rel-L1 about 1e-6 instead of 1e-8.  No
                                     counterpart in the reference (Float64 there) */

typedef struct s2g_ctx s2g_ctx;

/* The two fields of mappingParameters the deposit reads
 * (shared/parameters.jl:14-27): Npixels[1] (all three are equal,
 * parameters.jl:105) and len2pix of the centred parameters
 * (parameters.jl:115, re-derived by filter_shift.jl:29-31). */
typedef struct s2g_params {
    int64_t npixels;
    double len2pix;
} s2g_params;

/* optional per-call statistics */
typedef struct s2g_stats {
    double ms_device;         /* device time of the deposit (HIP events) */
    double ms_h2d, ms_d2h;    /* host<->device copies of the host entry points */
    int64_t n_particles;      /* particles passed in */
    int64_t n_skipped;        /* zero quantity without calc_mean, empty footprint */
    int64_t n_small;          /* particles deposited by the tile scatter path */
    int64_t n_large;          /* particles deposited by the tile gather path */
    int64_t n_large_exact;    /* ... of which normalised by an explicit sum */
    int64_t n_pairs;          /* (tile, particle) pairs of the gather path */
    double sum_footprint;     /* reserved (0) */
    /* device time per stage (HIP events), ms.  In the 2D / 3D deposits the
     * normalisation sums and the small-footprint scatter run on a side stream
     * beside the pair sort and the gather, so those stages overlap: ms_norm
     * and ms_small are the spans on the side stream (they include its wait
     * for a share of the GPU), ms_device the whole call. */
    double ms_setup;          /* per-particle setup + class partition */
    double ms_norm;           /* explicit footprint normalisation sums */
    double ms_sort;           /* pair counting sort + work items */
    double ms_gather;         /* tile gather kernels only */
    double ms_finalize;       /* fixed-order partial sums -> image */
    double ms_small;          /* small-footprint scatter */
    int64_t n_gather_launches;
} s2g_stats;

/* Work split of the 2D deposit (s2g_deposit, s2g_map, s2g_deposit_batch in 2D).
 * No counterpart in the reference: these choose which kernel deposits a
 * particle and how the work is cut into blocks, not the arithmetic of a pixel
 * weight, so every setting gives the same maps up to rounding.  Set per
 * context; sph2grid/autotune.py picks them per shape class. */
typedef struct s2g_tuning {
    int32_t small_side;          /* footprint box side (px) up to which a particle
                                    takes the small-footprint scatter; 5..34, default 34
                                    (>= 5 keeps every particle of hsml < 2 px, whose
                                    weights need FP64, in the scatter) */
    int32_t small_item_min;      /* least particles per small-footprint work item;
                                    16..4096, default 256 */
    int32_t small_items_per_cu;  /* small-footprint work items per CU aimed at;
                                    1..64, default 16 */
    int32_t gather_item_min;     /* least (tile, particle) pairs per gather work item;
                                    8..4096, default 64 */
    int32_t gather_items_per_cu; /* gather work items per CU aimed at; 1..256, default 16 */
    int32_t gather_planes;       /* image planes per gather launch; 1..4, default 4 */
    int32_t small_planes;        /* image planes per small-footprint launch; 1..4, default 4 */
    int32_t gather_item_max;     /* most pairs per gather work item, so that a large pass
                                    is cut into more items than gather_items_per_cu asks
                                    (a shorter tail of the launch); 0 (no bound) or
                                    gather_item_min..2^30, default 4096 */
} s2g_tuning;

/* Context: owns the device buffers (grown monotonically) and one blocking
 * stream (default priority) on `device`. */
int s2g_create(s2g_ctx **out, int device);
void s2g_destroy(s2g_ctx *ctx);
const char *s2g_last_error(const s2g_ctx *ctx);
/* library version string, e.g. "sph2grid 0.1.0 gfx950" */
const char *s2g_version(void);
/* Work split of later 2D deposits on this context (NULL: the defaults);
 * out-of-range fields give S2G_EINVAL and leave the setting unchanged.  No
 * counterpart in the reference (its deposit is one loop per worker). */
int s2g_set_tuning(s2g_ctx *ctx, const s2g_tuning *tuning);
int s2g_get_tuning(s2g_ctx *ctx, s2g_tuning *tuning);

/* Deposit (cic_mapping_2D for dims == 2, cic_mapping_3D for dims == 3).
 * Host pointers.  image_out receives the flat unreduced image
 * (npix^dims x (nq+1), or x 2 in 3D where nq must be 1).
 * A particle with hsml <= 0 or a NaN coordinate or hsml deposits nothing.
 * (The reference writes NaN into an h = 0 particle's pixel in 2D, where
 * area_norm = 0 * Inf, cic_2D.jl:187-222; it deposits nothing for h = 0 in
 * 3D, and it throws InexactError on NaN in floor(Integer, x - hsml),
 * cic_shared.jl:46-52.) */
int s2g_deposit(s2g_ctx *ctx, int dims, int dtype, const void *pos,
                const void *hsml, const void *m, const void *rho,
                const void *bin_q, int nq, const void *weights, int64_t n,
                const s2g_params *params, int kernel, int kernel_dim,
                uint32_t flags, double *image_out, s2g_stats *stats);

/* Same on device pointers; d_image_out must hold the flat image. */
int s2g_deposit_device(s2g_ctx *ctx, int dims, int dtype, const void *d_pos,
                       const void *d_hsml, const void *d_m, const void *d_rho,
                       const void *d_bin_q, int nq, const void *d_weights,
                       int64_t n, const s2g_params *params, int kernel,
                       int kernel_dim, uint32_t flags, double *d_image_out,
                       void *stream, s2g_stats *stats);

/* reduce_image_2D / reduce_image_3D on a flat image (host pointers).
 * 2D output: Julia [x,y,q] column-major, out[q*N*N + y*N + x], divided by
 * the weight column iff (flags & S2G_FLAG_REDUCE) and weight > 0.
 * 3D output: [z,y,x] in deposit order, divided iff the value is > 0; without
 * S2G_FLAG_REDUCE the weights are taken as 1 (cic_interpolation.jl:230-232). */
int s2g_reduce(s2g_ctx *ctx, int dims, int64_t npix, int nq, uint32_t flags,
               const double *image, double *out);
int s2g_reduce_device(s2g_ctx *ctx, int dims, int64_t npix, int nq,
                      uint32_t flags, const double *d_image, double *d_out,
                      void *stream);

/* Deposit + reduce in one call, returning only the reduced map (what
 * sphMapping returns when return_both_maps == false). */
int s2g_map(s2g_ctx *ctx, int dims, int dtype, const void *pos,
            const void *hsml, const void *m, const void *rho,
            const void *bin_q, int nq, const void *weights, int64_t n,
            const s2g_params *params, int kernel, int kernel_dim,
            uint32_t flags, double *map_out, s2g_stats *stats);

/* HEALPix all-sky deposit (the particle loop of healpix_map,
 * healpix_interpolation/main.jl:146-213) over particles already centred on
 * the observer, shell-filtered and ordered by the caller
 * (filter_particles.jl:17-54).  Outputs two RING-ordered maps of 12*nside^2
 * doubles: sum of q*pw and sum of pw (unreduced, as the reference returns). */
int s2g_healpix_deposit(s2g_ctx *ctx, int dtype, const void *pos,
                        const void *hsml, const void *m, const void *rho,
                        const void *bin_q, const void *weights, int64_t n,
                        int64_t nside, int kernel, int kernel_dim,
                        uint32_t flags, double *map_out, double *weight_out,
                        s2g_stats *stats);
/* Same on device pointers (maps zeroed and filled on `stream`). */
int s2g_healpix_deposit_device(s2g_ctx *ctx, int dtype, const void *d_pos,
                               const void *d_hsml, const void *d_m,
                               const void *d_rho, const void *d_bin_q,
                               const void *d_weights, int64_t n, int64_t nside,
                               int kernel, int kernel_dim, uint32_t flags,
                               double *d_map_out, double *d_weight_out,
                               void *stream, s2g_stats *stats);

/* Several all-sky maps of one particle set: map k deposits bin_q[k] weighted
 * by weights[k] (both n values) into map_out[k] / weight_out[k], exactly as
 * one s2g_healpix_deposit call per map would (the reference makes each with a
 * healpix_map call on the same particles, main.jl:92-227, e.g. the density
 * and the density-weighted temperature of config 5).  The disc queries, the
 * normalisation sums (pixel_weights.jl:87-140, which depend only on positions
 * and hsml) and the per-ring particle lists are computed once, and each
 * pixel's weight once for every two maps.  1 <= nmaps <= 8.  Host pointers
 * (arrays of nmaps pointers for the per-map arguments). */
int s2g_healpix_deposit_maps(s2g_ctx *ctx, int dtype, const void *pos, const void *hsml,
                             const void *m, const void *rho, int nmaps,
                             const void *const *bin_q, const void *const *weights, int64_t n,
                             int64_t nside, int kernel, int kernel_dim, uint32_t flags,
                             double *const *map_out, double *const *weight_out,
                             s2g_stats *stats);
/* Same on device pointers (every pixel of every map written on `stream`;
 * returns with the maps complete). */
int s2g_healpix_deposit_maps_device(s2g_ctx *ctx, int dtype, const void *d_pos,
                                    const void *d_hsml, const void *d_m, const void *d_rho,
                                    int nmaps, const void *const *d_bin_q,
                                    const void *const *d_weights, int64_t n, int64_t nside,
                                    int kernel, int kernel_dim, uint32_t flags,
                                    double *const *d_map_out, double *const *d_weight_out,
                                    void *stream, s2g_stats *stats);

/* Triangular-shaped-cloud deposit onto an npix^3 grid (positions centred as
 * for the 3D deposit, grid units x*len2pix + npix/2, cell c centred at c+0.5).
 * grid_out: 2 x npix^3 doubles in 3D deposit order, plane 0 = sum q*W,
 * plane 1 = sum W.  S2G_FLAG_PERIODIC wraps, otherwise outside cells drop. */
int s2g_tsc3d(s2g_ctx *ctx, int dtype, const void *pos, const void *bin_q,
              int64_t n, const s2g_params *params, uint32_t flags,
              double *grid_out, s2g_stats *stats);
/* Same on device pointers, on `stream`; returns with the grid complete (the
 * particles are ordered by grid brick in the context's buffers first). */
int s2g_tsc3d_device(s2g_ctx *ctx, int dtype, const void *d_pos,
                     const void *d_bin_q, int64_t n, const s2g_params *params,
                     uint32_t flags, double *d_grid_out, void *stream);

/* Allocation-free variant of s2g_tsc3d_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result as s2g_tsc3d_device, same kernels, but its work
 * space (sort keys, brick offsets, sorted particle data) is a fixed 80 MiB
 * device array per GPU that the library holds until the process ends, so it
 * calls no hipMalloc / hipFree; limits: n <= 2^20 particles and npixels <= 512
 * (larger calls fail with S2G_EUNSUPPORTED before any device work).  It only
 * enqueues on `stream` (NULL: the context's stream).  Every call on one GPU
 * uses that GPU's one array: calls that overlap in time on one GPU (from
 * other streams, contexts or threads) race; calls ordered on one stream do
 * not. */
int s2g_tsc3d_device_noalloc(s2g_ctx *ctx, int dtype, const void *d_pos, const void *d_bin_q,
                             int64_t n, const s2g_params *params, uint32_t flags,
                             double *d_grid_out, void *stream);

/* One map of a batch (device pointers; inputs centred and filtered as for
 * s2g_deposit). */
typedef struct s2g_map_desc {
    int32_t dims;      /* 2 (other values: S2G_EUNSUPPORTED) */
    int32_t nq;
    int64_t n;
    s2g_params params;
    int32_t kernel;
    uint32_t flags;    /* S2G_FLAG_CALC_MEAN / S2G_FLAG_EXACT_NORM / S2G_FLAG_REDUCE */
    const void *d_pos, *d_hsml, *d_m, *d_rho, *d_bin_q, *d_weights;
    double *d_out;     /* flat image (nq + 1 planes), or with S2G_FLAG_REDUCE the
                          reduced map of reduce_image_2D(reduce_image = true) */
} s2g_map_desc;

/* Many independent 2D maps, all with the same dtype.  Replaces a loop of
 * sphMapping calls (cic_interpolation.jl:35-273, one call per map, as in
 * distributed_mapping/cic.jl:24-110).  Maps with the same kernel and nq share
 * one device pass: their particles and tiles are concatenated, and the host
 * reads back twice per pass, not per map.  Synchronous
 * on return; `stream` may be NULL (the context's stream).  Errors name the
 * offending map. */
int s2g_deposit_batch(s2g_ctx *ctx, int dtype, const s2g_map_desc *maps,
                      int nmaps, void *stream, s2g_stats *stats);

/* Replayable batch (no counterpart in the reference; for callers that
 * deposit the same maps again and again, e.g. every snapshot of a run or a
 * serving loop, and for callers that cannot afford host round trips inside
 * a deposit).  s2g_plan_create fixes the descriptors (device pointers,
 * particle counts, image sizes, kernels, flags) and the context's current
 * s2g_tuning, sizes every work list from one run of the class and pair-count
 * stages on the current inputs (capacities = that run's counts times
 * `headroom`, 1..64), and captures the whole batch in one hipGraph.
 * s2g_plan_run replays it on `stream` (NULL: the context's stream) with the
 * inputs' current contents and returns when the maps are written; nothing in
 * the replay reads back to the host, and the run reads back only the
 * capacity flags (and the counts for `stats`) after it.  When the inputs
 * need more large particles or (tile, particle) pairs than the capacities,
 * the run writes nothing past them and returns S2G_ECAPACITY (the maps are
 * then incomplete): create a new plan.  Results equal s2g_deposit_batch's up
 * to the order of the small-footprint additions.  Plans hold their own
 * device memory; a plan and its context are used from one thread. */
typedef struct s2g_plan s2g_plan;
int s2g_plan_create(s2g_ctx *ctx, int dtype, const s2g_map_desc *maps, int nmaps,
                    double headroom, s2g_plan **out);
int s2g_plan_run(s2g_plan *plan, void *stream, s2g_stats *stats);
void s2g_plan_destroy(s2g_plan *plan);

/* Test export (no counterpart in the reference; for bit-exact checks of the
 * deposit's integer work against a CPU restatement).  Runs the setup, class
 * partition, (tile, particle) counting sort and small-footprint bucketing of
 * s2g_deposit on the same host arguments and returns, in host arrays:
 *   box        n x 4 (2D: i0, i1, j0, j1) or n x 6 (3D: + k0, k1) int32, the
 *              footprint of pix_index_min_max (cic_shared.jl:46-52) clamped to
 *              the image; (0, -1, ...) for a skipped particle
 *   cls        n int32: 0 skipped (zero quantity without calc_mean, empty
 *              footprint, hsml <= 0 or NaN), 1 small (box side <= 34 px in 2D,
 *              or the context's s2g_tuning.small_side; <= 8 in 3D), 2 large
 *              with the closed-form normalisation, 3 large
 *              with an explicit footprint sum
 *   fb         n int32: 1 iff no footprint pixel centre has u < 1 (the
 *              fallback of calculate_weights, cic_2D.jl:49-69, cic_3D.jl:55-75)
 *   tile_off   ntiles + 1 int64 (ntiles = ceil(npix/64)^2 in 2D, ceil(npix/16)^3
 *              in 3D, row-major): the large particles whose box meets tile t are
 *              pairs[tile_off[t] .. tile_off[t+1]), in the order f_j =
 *              e_{(j r) mod L} of the large particles e_0 .. e_{L-1} (particle
 *              order), r = the integer nearest to L / phi (golden ratio) raised
 *              to the first value coprime to L
 *   pairs      the pair lists, copied when their total is <= pair_cap
 *   small_off  2D: ceil(npix/32)^2 + 1 int64, small_order[small_off[t] ..
 *              small_off[t+1]) are the small particles whose box starts in
 *              32 x 32 tile t, in particle order; 3D: {0, nsmall}
 *   small_order n int32 (the first nsmall are written)
 *   counts     6 int64: nsmall, nlarge, nlarge_explicit, npairs, ntiles,
 *              small tiles */
int s2g_footprints(s2g_ctx *ctx, int dims, int dtype, const void *pos, const void *hsml,
                   const void *m, const void *rho, const void *bin_q, int nq,
                   const void *weights, int64_t n, const s2g_params *params, int kernel,
                   uint32_t flags, int32_t *box, int32_t *cls, int32_t *fb, int64_t *tile_off,
                   int32_t *pairs, int64_t pair_cap, int64_t *small_off, int32_t *small_order,
                   int64_t *counts);

/* Subfile reductions of distributed_cic_map / distributed_allsky_map: add
 * one subfile's map into a resident sum, skipping NaN and Inf values
 * (distributed_mapping/cic.jl:63-73, distributed_mapping/healpix.jl:50-58).
 * d_dst[i] += d_src[i] for finite d_src[i].  Enqueues only. */
int s2g_accumulate_device(s2g_ctx *ctx, const double *d_src, double *d_dst, int64_t count,
                          void *stream);

/* reduce_image of distributed_allsky_map (distributed_mapping/healpix.jl:62-70):
 * d_map[i] /= d_weights[i] where the weight is finite and non-zero.
 * Enqueues only. */
int s2g_healpix_reduce_device(s2g_ctx *ctx, double *d_map, const double *d_weights,
                              int64_t npix, void *stream);

/* mass_density (effects/mass_density.jl:30-99): SPH density and smoothing
 * length of an arbitrary particle set.  hsml_i = distance to the farthest of
 * the nneighbors + 1 nearest particles (the particle itself included);
 * rho_i = sum_j m_j W(r_ij / hsml_i, 1 / hsml_i) over them, with the kernel's
 * own normalisation in kernel_dim (2 or 3) dimensions, then the Wendland
 * kernels' bias correction of Dehnen & Aly (2012, eqs. 18-19).  pos is 3 x n
 * interleaved; boxsize = NULL or all zero: Euclidean distances, all positive:
 * periodic box (minimum image per dimension).  nneighbors + 1 <= n and
 * nneighbors <= 471 (the search buffer per particle). */
int s2g_mass_density(s2g_ctx *ctx, int dtype, const void *pos, const void *mass, int64_t n,
                     int kernel, int kernel_dim, int64_t nneighbors, const double *boxsize,
                     double *rho_out, double *hsml_out, s2g_stats *stats);
/* Same on device pointers (synchronises the stream before returning). */
int s2g_mass_density_device(s2g_ctx *ctx, int dtype, const void *d_pos, const void *d_mass,
                            int64_t n, int kernel, int kernel_dim, int64_t nneighbors,
                            const double *boxsize, double *d_rho_out, double *d_hsml_out,
                            void *stream, s2g_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* SPH2GRID_H */
