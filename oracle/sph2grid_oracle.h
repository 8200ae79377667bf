/*
 * sph2grid_oracle.h -- FP64 CPU restatement of SPHtoGrid.jl's deposition path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links,
 * loads or calls anything in oracle/.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the SPHtoGrid.jl source tree).  All arithmetic is IEEE binary64 with FMA
 * contraction disabled at compile time (-ffp-contract=off), so the two-step
 * roundings of the reference (e.g. x*len2pix then +0.5*Npix) are preserved.
 */
#ifndef SPH2GRID_ORACLE_H
#define SPH2GRID_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel ids; identical numbering to include/sph2grid.h */
enum {
    ORC_KERNEL_CUBIC = 0,
    ORC_KERNEL_QUINTIC = 1,
    ORC_KERNEL_WENDLAND_C2 = 2,
    ORC_KERNEL_WENDLAND_C4 = 3,
    ORC_KERNEL_WENDLAND_C6 = 4,
    ORC_KERNEL_WENDLAND_C8 = 5
};

/* ---- SPHKernels.jl (external dependency, compat "2"): 𝒲(kernel,u,h_inv) --- */
double orc_kernel_shape(int kernel, double u);
double orc_kernel_norm(int kernel, int dim);
double orc_kernel_value(int kernel, int dim, double u, double h_inv);

/* ---- shared/parameters.jl:14-27 ---------------------------------------- */
typedef struct {
    double x_lim[2], y_lim[2], z_lim[2];
    double center[3];
    double halfsize[3];
    double len2pix;
    double pixelSideLength;
    int64_t Npixels[3];
    double boxsize;
    int32_t periodic;
    int32_t _pad;
} orc_params;

/* shared/parameters.jl:44-125.  Unset arguments carry the reference's
 * sentinels ([-1,-1] limits, [-1,-1,-1] centre, -1 sizes / pixel length,
 * 0 Npixels).  Returns 0, 1 ("Giving a center position requires extent in x,
 * y and z direction.") or 2 ("Please specify pixelSideLength or number of
 * pixels!"). */
int orc_mapping_parameters(orc_params *out, const double x_lim[2],
                           const double y_lim[2], const double z_lim[2],
                           const double center[3], double x_size,
                           double y_size, double z_size,
                           double pixelSideLength, int64_t Npixels,
                           double boxsize);

/* cic_interpolation/filter_shift.jl:6-32 (mutates pos, 3 x n column-major) */
void orc_center_particles(void *pos, int is_f32, int64_t n,
                          const orc_params *par, orc_params *par_out);
/* cic_interpolation/filter_shift.jl:40-67 (sort_z = false branch) */
void orc_filter_particles(const void *pos, int is_f32, int64_t n,
                          const orc_params *par, uint8_t *keep);

/* cic_interpolation/cic_2D.jl:103-244.  Inputs are the filtered particles,
 * already promoted to Float64.  bin_q is nq x n (quantity fastest) or n when
 * nq == 1.  image (npix^2 x (nq+1), column-major, deposit order, last column
 * = weights) is zeroed here.  Returns 0. */
int orc_cic_mapping_2D(const double *pos, const double *hsml,
                       const double *m, const double *rho,
                       const double *bin_q, int nq, const double *weights,
                       int64_t n, int64_t npix, double len2pix, int kernel,
                       int kernel_dim, int calc_mean, double *image);

/* cic_interpolation/cic_3D.jl:110-209 (scalar quantity, image npix^3 x 2).
 * grid_mass/particle_mass receive the diagnostic sums of cic_3D.jl:186,192
 * (may be NULL). */
int orc_cic_mapping_3D(const double *pos, const double *hsml,
                       const double *m, const double *rho,
                       const double *bin_q, const double *weights, int64_t n,
                       int64_t npix, double len2pix, int kernel,
                       int kernel_dim, int calc_mean, double *image,
                       double *grid_mass, double *particle_mass);

/* The reference's parallel=true layout (cic_interpolation.jl:171-215,
 * 236-271 + parallel/domain_decomp.jl:7-17): contiguous batches, one private
 * image per worker (here: per thread), images summed in worker order. */
int orc_cic_mapping_2D_parallel(int nworkers, const double *pos,
                                const double *hsml, const double *m,
                                const double *rho, const double *bin_q,
                                int nq, const double *weights, int64_t n,
                                int64_t npix, double len2pix, int kernel,
                                int kernel_dim, int calc_mean, double *image);
int orc_cic_mapping_3D_parallel(int nworkers, const double *pos,
                                const double *hsml, const double *m,
                                const double *rho, const double *bin_q,
                                const double *weights, int64_t n,
                                int64_t npix, double len2pix, int kernel,
                                int kernel_dim, int calc_mean, double *image);
/* parallel/domain_decomp.jl:7-17: start (0-based) and length of batch i */
void orc_domain_decomposition(int64_t n, int nworkers, int i, int64_t *start,
                              int64_t *len);

/* cic_interpolation/reduce_image.jl:8-31.  out has Julia layout [x,y,q]
 * column-major: out[q*npix*npix + y*npix + x]. */
void orc_reduce_image_2D(const double *image, int64_t npix, int nq,
                         int reduce_image, double *out);
/* cic_interpolation/reduce_image.jl:39-55 (no transpose; [z,y,x]). */
void orc_reduce_image_3D(const double *image, int64_t npix, double *out);

/* ---- HEALPix RING scheme (Healpix.jl compat "4"; Gorski et al. 2005) ---- */
int64_t orc_hp_npix(int64_t nside);
void orc_hp_pix2ang_ring(int64_t nside, int64_t pix, double *theta,
                         double *phi);
int64_t orc_hp_ang2pix_ring(int64_t nside, double theta, double phi);
void orc_hp_pix2vec_ring(int64_t nside, int64_t pix, double vec[3]);
void orc_hp_vec2ang(const double vec[3], double *theta, double *phi);
int64_t orc_hp_ring_above(int64_t nside, double z);
/* query_disc (non-inclusive): pixels whose centre lies within radius.
 * Writes up to cap ids (ascending ring order), returns the total count. */
int64_t orc_hp_query_disc_ring(int64_t nside, double theta, double phi,
                               double radius, int64_t *out, int64_t cap);

/* ---- integer work of a deposit (for bit-exact checks of the GPU's) ----
 * Per particle: the footprint box of pix_index_min_max (cic_shared.jl:46-52)
 * from get_xyz's two roundings (cic_shared.jl:85-100), (0, -1) per axis for a
 * skipped particle; the fallback flag of calculate_weights (distr_weight ==
 * 0, cic_2D.jl:49-69, cic_3D.jl:55-75) by the reference's own pixel loop; and
 * the class of the GPU deposit's work split, which has no reference
 * counterpart: skipped (zero quantity without calc_mean, hsml <= 0, NaN,
 * empty footprint), small (every box side <= 34 px in 2D, <= 8 voxels in 3D),
 * large with the closed-form normalisation (no ORC_FLAG_EXACT_NORM, footprint
 * not clipped by the image border, no fallback, hsml >= a per-kernel
 * threshold), or large with an explicit sum.  box is n x 2*dims. */
#define ORC_CLS_SKIP 0
#define ORC_CLS_SMALL 1
#define ORC_CLS_LARGE_CLOSED 2
#define ORC_CLS_LARGE_EXPLICIT 3
#define ORC_SMALL_SIDE_2D 34
#define ORC_SMALL_SIDE_3D 8
#define ORC_FLAG_CALC_MEAN 0x1u
#define ORC_FLAG_EXACT_NORM 0x2u
int orc_footprints(int dims, const double *pos, const double *hsml, const double *bin_q,
                   int nq, int64_t n, int64_t npix, double len2pix, int kernel,
                   uint32_t flags, int32_t *box, int32_t *cls, int32_t *fb);

/* The per-pixel weights of one particle (constributing_pixels.jl:7-22,
 * pixel_weights.jl:6-140), without depositing: up to cap entries of the
 * pixel list (RING ids, 0-based), the area A (including its /(ang_pix*Dx)^2)
 * and the kernel weight wk (1 everywhere in the fallback).  info[0..5] =
 * Dx, proj_hsml, ang_pix, weight_per_pix, n_distr_pix, fallback flag.
 * Returns the pixel count (0: skipped, Dx < hsml; -1: out of memory). */
int64_t orc_hp_particle_weights(const double pos[3], double hsml, int64_t nside, int kernel,
                                int kernel_dim, int64_t cap, int64_t *pix, double *A,
                                double *wk, double *info);

/* healpix_interpolation/main.jl:92-227 + filter_particles.jl:17-54.
 * pos (3 x n, Float64) is mutated (centre subtracted) as in the reference.
 * Returns 0, or -1 where the reference throws a BoundsError (calc_mean=false
 * with any particle removed by the shell / positivity mask). */
int orc_healpix_map(double *pos, const double *hsml, const double *m,
                    const double *rho, const double *bin_q,
                    const double *weights, int64_t n, const double center[3],
                    const double radius_limits[2], int64_t nside, int kernel,
                    int kernel_dim, int calc_mean, double *allsky_map,
                    double *weight_map);
/* Same deposit over already-centred particles, split over nworkers threads
 * in contiguous batches with per-thread private maps summed in order (CPU
 * baseline for the all-sky map; the reference itself runs it serially). */
int orc_healpix_deposit_parallel(int nworkers, const double *pos,
                                 const double *hsml, const double *m,
                                 const double *rho, const double *bin_q,
                                 const double *weights, int64_t n,
                                 int64_t nside, int kernel, int kernel_dim,
                                 int calc_mean, double *allsky_map,
                                 double *weight_map);

/* ---- TSC (textbook Hockney & Eastwood; the reference's TSC path is fully
 * commented out in tsc_interpolation/tsc_interpolation.jl:1-182) ---------- */
int orc_tsc_3d(const double *pos, const double *bin_q, int64_t n,
               int64_t npix, double len2pix, int periodic, double *grid);

/* ---- mass_density (effects/mass_density.jl:30-99) ----------------------
 * Brute-force k nearest neighbours (nneighbors + 1 of them, the particle
 * itself included), hsml = the largest of their distances, the SPH density
 * sum m_j W(r_ij / h, 1 / h) and the kernel bias correction of Dehnen & Aly
 * (2012, eqs. 18-19; SPHKernels.jl bias_correction).  pos is 3 x n
 * interleaved; boxsize[d] > 0 for all d selects the periodic (minimum image)
 * metric, all zero the Euclidean one.  O(n^2). */
double orc_bias_correction(int kernel, int dim, double density, double m, double h_inv,
                           int64_t nneighbors);
int orc_mass_density(const double *pos, const double *mass, int64_t n, int kernel, int dim,
                     int64_t nneighbors, const double *boxsize, double *rho, double *hsml);
/* The same for the particles idx[0..nidx) only (against all n); results in
 * rho[a], hsml[a] for a < nidx. */
int orc_mass_density_subset(const double *pos, const double *mass, int64_t n,
                            const int64_t *idx, int64_t nidx, int kernel, int dim,
                            int64_t nneighbors, const double *boxsize, double *rho,
                            double *hsml);

#ifdef __cplusplus
}
#endif
#endif
