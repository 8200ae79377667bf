/*
 * sph2grid_oracle.c -- FP64 CPU restatement of SPHtoGrid.jl's deposition path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
 * product library never links or loads this file.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off, no -ffast-math).
 *
 * Citations are file:line into the SPHtoGrid.jl v0.5.3 source tree.
 * Kernel shapes restate SPHKernels.jl (compat "2", not vendored) from the
 * published closed forms; HEALPix restates the RING scheme of Gorski et al.
 * (2005, ApJ 622, 759) as implemented by Healpix.jl (compat "4", not
 * vendored).
 */
#include "sph2grid_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* SPHKernels.jl shapes.  W(u) = norm(dim) * h_inv^dim * shape(u), 0 for u>=1 */
/* ------------------------------------------------------------------------ */

double orc_kernel_shape(int kernel, double u)
{
    if (!(u < 1.0))
        return 0.0;
    double t = 1.0 - u;
    switch (kernel) {
    case ORC_KERNEL_CUBIC:
        if (u < 0.5)
            return 1.0 + 6.0 * (u - 1.0) * u * u;
        return 2.0 * t * t * t;
    case ORC_KERNEL_QUINTIC: {
        double t5 = t * t * t * t * t;
        double v = t5;
        if (u < 2.0 / 3.0) {
            double a = 2.0 / 3.0 - u;
            v -= 6.0 * a * a * a * a * a;
        }
        if (u < 1.0 / 3.0) {
            double b = 1.0 / 3.0 - u;
            v += 15.0 * b * b * b * b * b;
        }
        return v;
    }
    case ORC_KERNEL_WENDLAND_C2: {
        double t2 = t * t;
        return t2 * t2 * (1.0 + 4.0 * u);
    }
    case ORC_KERNEL_WENDLAND_C4: {
        double t2 = t * t;
        double t6 = t2 * t2 * t2;
        return t6 * (1.0 + 6.0 * u + (35.0 / 3.0) * u * u);
    }
    case ORC_KERNEL_WENDLAND_C6: {
        double t2 = t * t;
        double t4 = t2 * t2;
        double t8 = t4 * t4;
        return t8 * (1.0 + 8.0 * u + 25.0 * u * u + 32.0 * u * u * u);
    }
    case ORC_KERNEL_WENDLAND_C8: {
        double t2 = t * t;
        double t4 = t2 * t2;
        double t10 = t4 * t4 * t2;
        return t10 * (1.0 + 10.0 * u + 42.0 * u * u + 90.0 * u * u * u +
                      (429.0 / 5.0) * u * u * u * u);
    }
    default:
        return 0.0;
    }
}

double orc_kernel_norm(int kernel, int dim)
{
    /* normalisations such that the integral of W over the d-ball of radius
     * h is one (d = 2 or 3) */
    int d3 = (dim == 3);
    switch (kernel) {
    case ORC_KERNEL_CUBIC: return d3 ? 8.0 / M_PI : 40.0 / (7.0 * M_PI);
    case ORC_KERNEL_QUINTIC:
        return d3 ? 2187.0 / (40.0 * M_PI) : 15309.0 / (478.0 * M_PI);
    case ORC_KERNEL_WENDLAND_C2: return d3 ? 21.0 / (2.0 * M_PI) : 7.0 / M_PI;
    case ORC_KERNEL_WENDLAND_C4:
        return d3 ? 495.0 / (32.0 * M_PI) : 9.0 / M_PI;
    case ORC_KERNEL_WENDLAND_C6:
        return d3 ? 1365.0 / (64.0 * M_PI) : 78.0 / (7.0 * M_PI);
    case ORC_KERNEL_WENDLAND_C8:
        return d3 ? 1785.0 / (64.0 * M_PI) : 40.0 / (3.0 * M_PI);
    default: return 0.0;
    }
}

double orc_kernel_value(int kernel, int dim, double u, double h_inv)
{
    double hd = (dim == 3) ? h_inv * h_inv * h_inv : h_inv * h_inv;
    return orc_kernel_norm(kernel, dim) * hd * orc_kernel_shape(kernel, u);
}

/* ------------------------------------------------------------------------ */
/* mappingParameters: shared/parameters.jl:44-125                            */
/* ------------------------------------------------------------------------ */

static int is_unset2(const double v[2]) { return v[0] == -1.0 && v[1] == -1.0; }

int orc_mapping_parameters(orc_params *o, const double x_lim[2],
                           const double y_lim[2], const double z_lim[2],
                           const double center[3], double x_size,
                           double y_size, double z_size,
                           double pixelSideLength, int64_t Npixels,
                           double boxsize)
{
    double xl[2] = {x_lim[0], x_lim[1]}, yl[2] = {y_lim[0], y_lim[1]},
           zl[2] = {z_lim[0], z_lim[1]};
    double c[3] = {center[0], center[1], center[2]};
    int c_unset = (c[0] == -1.0 && c[1] == -1.0 && c[2] == -1.0);

    /* parameters.jl:58-70 */
    if (is_unset2(xl) && is_unset2(yl) && is_unset2(zl)) {
        if (!c_unset && x_size != -1.0 && y_size != -1.0 && z_size != -1.0) {
            xl[0] = c[0] - 0.5 * x_size; xl[1] = c[0] + 0.5 * x_size;
            yl[0] = c[1] - 0.5 * y_size; yl[1] = c[1] + 0.5 * y_size;
            zl[0] = c[2] - 0.5 * z_size; zl[1] = c[2] + 0.5 * z_size;
        } else {
            return 1;
        }
    }
    /* parameters.jl:73-81 */
    if (x_size == -1.0) x_size = xl[1] - xl[0];
    if (y_size == -1.0) y_size = yl[1] - yl[0];
    if (z_size == -1.0) z_size = zl[1] - zl[0];
    /* parameters.jl:83-87 */
    if (c_unset) {
        c[0] = xl[0] + 0.5 * x_size;
        c[1] = yl[0] + 0.5 * y_size;
        c[2] = zl[0] + 0.5 * z_size;
    }
    /* parameters.jl:90-100 */
    double max_size = x_size > y_size ? x_size : y_size;
    if (pixelSideLength == -1.0 && Npixels != 0) {
        pixelSideLength = max_size / (double)Npixels;
    } else if (pixelSideLength != -1.0 && Npixels == 0) {
        Npixels = (int64_t)floor(max_size / pixelSideLength);
        pixelSideLength = max_size / (double)Npixels;
    } else {
        return 2;
    }
    memcpy(o->x_lim, xl, sizeof xl);
    memcpy(o->y_lim, yl, sizeof yl);
    memcpy(o->z_lim, zl, sizeof zl);
    memcpy(o->center, c, sizeof c);
    o->halfsize[0] = 0.5 * x_size;
    o->halfsize[1] = 0.5 * y_size;
    o->halfsize[2] = 0.5 * z_size;
    o->len2pix = 1.0 / pixelSideLength;
    o->pixelSideLength = pixelSideLength;
    o->Npixels[0] = o->Npixels[1] = o->Npixels[2] = Npixels;
    o->boxsize = boxsize;
    o->periodic = (boxsize != -1.0);
    o->_pad = 0;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* filter_shift.jl                                                           */
/* ------------------------------------------------------------------------ */

void orc_center_particles(void *pos, int is_f32, int64_t n,
                          const orc_params *par, orc_params *par_out)
{
    /* filter_shift.jl:14-23: x[dim,i] -= cen[dim]; the Float64 result is
     * stored back into the caller's array (re-rounded for Float32). */
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double v = is_f32 ? (double)((float *)pos)[3 * i + d]
                              : ((double *)pos)[3 * i + d];
            v = v - par->center[d];
            if (is_f32)
                ((float *)pos)[3 * i + d] = (float)v;
            else
                ((double *)pos)[3 * i + d] = v;
            if (par->periodic) {
                /* the shift is boxsize/2, as in the reference */
                double w = is_f32 ? (double)((float *)pos)[3 * i + d]
                                  : ((double *)pos)[3 * i + d];
                if (fabs(w) > par->boxsize / 2.0) {
                    w = (w > 0) ? w - par->boxsize / 2.0 : w + par->boxsize / 2.0;
                    if (is_f32)
                        ((float *)pos)[3 * i + d] = (float)w;
                    else
                        ((double *)pos)[3 * i + d] = w;
                }
            }
        }
    }
    /* filter_shift.jl:25-31: new parameters centred on the origin */
    double xl[2] = {par->x_lim[0] - par->center[0], par->x_lim[1] - par->center[0]};
    double yl[2] = {par->y_lim[0] - par->center[1], par->y_lim[1] - par->center[1]};
    double zl[2] = {par->z_lim[0] - par->center[2], par->z_lim[1] - par->center[2]};
    double c0[3] = {0.0, 0.0, 0.0};
    int64_t np = par->Npixels[0];
    if (par->Npixels[1] > np) np = par->Npixels[1];
    if (par->Npixels[2] > np) np = par->Npixels[2];
    orc_mapping_parameters(par_out, xl, yl, zl, c0, -1.0, -1.0, -1.0, -1.0, np,
                           par->boxsize);
}

void orc_filter_particles(const void *pos, int is_f32, int64_t n,
                          const orc_params *par, uint8_t *keep)
{
    /* filter_shift.jl:44-58: centre inside [c-halfsize, c+halfsize], inclusive */
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = par->center[d] - par->halfsize[d];
        hi[d] = par->center[d] + par->halfsize[d];
    }
    for (int64_t i = 0; i < n; ++i) {
        int in = 1;
        for (int d = 0; d < 3; ++d) {
            double v = is_f32 ? (double)((const float *)pos)[3 * i + d]
                              : ((const double *)pos)[3 * i + d];
            if (!(lo[d] <= v && v <= hi[d]))
                in = 0;
        }
        keep[i] = (uint8_t)in;
    }
}

/* ------------------------------------------------------------------------ */
/* cic_shared.jl helpers                                                     */
/* ------------------------------------------------------------------------ */

/* cic_shared.jl:46-52 */
static void pix_index_min_max(double x, double h, int64_t npix, int64_t *lo,
                              int64_t *hi)
{
    int64_t a = (int64_t)floor(x - h);
    int64_t b = (int64_t)floor(x + h);
    *lo = a > 0 ? a : 0;
    *hi = b < npix - 1 ? b : npix - 1;
}

/* cic_shared.jl:60-62 (get_dxyz) and :70-76 (get_x_dx) */
static inline double overlap_1d(double x, double h, int64_t i)
{
    double a = x + h, b = (double)(i + 1);
    double c = x - h, d = (double)i;
    return (a < b ? a : b) - (c > d ? c : d);
}

static inline double centre_offset(double x, int64_t i)
{
    return (x - (double)i) - 0.5;
}

/* cic_shared.jl:9-39 */
typedef struct {
    double distr_area, distr_weight;
    int64_t n_tot, n_distr;
} wsum;

static inline void weight_per_pixel(wsum *s, double dxdy, double u,
                                    double h_inv, int kernel, int kdim,
                                    double *A_out, double *wk_out)
{
    *A_out = dxdy;
    s->distr_area += dxdy;
    s->n_tot += 1;
    if (u <= 1.0) {
        double w = orc_kernel_value(kernel, kdim, u, h_inv);
        s->distr_weight += w * dxdy;
        s->n_distr += 1;
        *wk_out = w;
    } else {
        *wk_out = 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* cic_2D.jl:103-244                                                         */
/* ------------------------------------------------------------------------ */

static void map2d_range(const double *pos, const double *hsml, const double *m,
                        const double *rho, const double *bin_q, int nq,
                        const double *weights, int64_t p0, int64_t p1,
                        int64_t npix, double len2pix, int kernel, int kdim,
                        int calc_mean, double *image, double *wk, double *A)
{
    const int64_t nd = npix * npix;
    double *wimg = image + (int64_t)nq * nd; /* last column: weights */
    for (int64_t p = p0; p < p1; ++p) {
        /* cic_2D.jl:155-168: skip zero quantities unless calc_mean */
        int allzero = 1;
        for (int q = 0; q < nq; ++q)
            if (bin_q[p * nq + q] != 0.0)
                allzero = 0;
        if (allzero && !calc_mean)
            continue;

        /* get_quantities_2D, cic_2D.jl:80-91 */
        double h = hsml[p] * len2pix;
        double h_inv = 1.0 / h;
        double area = (2.0 * h) * (2.0 * h);
        double r = rho[p] * (1.0 / (len2pix * len2pix * len2pix));
        double dz = m[p] / r / area;
        double w_los = weights[p];

        /* get_xyz, cic_shared.jl:85-100: two roundings, no FMA */
        double x = pos[3 * p + 0] * len2pix;
        double y = pos[3 * p + 1] * len2pix;
        x = x + 0.5 * (double)npix;
        y = y + 0.5 * (double)npix;

        int64_t iMin, iMax, jMin, jMax;
        pix_index_min_max(x, h, npix, &iMin, &iMax);
        pix_index_min_max(y, h, npix, &jMin, &jMax);

        /* calculate_weights, cic_2D.jl:11-72 */
        wsum s = {0.0, 0.0, 0, 0};
        for (int64_t i = iMin; i <= iMax; ++i) {
            double xd = centre_offset(x, i), dx = overlap_1d(x, h, i);
            for (int64_t j = jMin; j <= jMax; ++j) {
                double yd = centre_offset(y, j), dy = overlap_1d(y, h, j);
                double u = sqrt(xd * xd + yd * yd) * h_inv;
                double dxdy = dx * dy;
                int64_t idx = i * npix + j;
                weight_per_pixel(&s, dxdy, u, h_inv, kernel, kdim, &A[idx],
                                 &wk[idx]);
            }
        }
        double wpp;
        if (s.distr_weight == 0.0) {
            s.n_distr = s.n_tot;
            for (int64_t i = iMin; i <= iMax; ++i)
                for (int64_t j = jMin; j <= jMax; ++j)
                    wk[i * npix + j] = 1.0;
            wpp = (s.distr_area != 0.0) ? (double)s.n_distr / s.distr_area : 1.0;
        } else {
            wpp = (double)s.n_distr / s.distr_weight;
        }

        /* cic_2D.jl:187-222 + update_image!, cic_shared.jl:111-121 */
        double kernel_norm = area / (double)s.n_distr;
        double area_norm = kernel_norm * wpp * w_los * dz;
        for (int64_t i = iMin; i <= iMax; ++i) {
            for (int64_t j = jMin; j <= jMax; ++j) {
                int64_t idx = i * npix + j;
                double pw = wk[idx] * A[idx] * area_norm;
                if (pw != 0.0) {
                    wimg[idx] += pw;
                    if (allzero) {
                        /* cic_2D.jl:160-162: bin_q collapses to scalar 0 */
                        image[idx] += 0.0 * pw;
                    } else {
                        for (int q = 0; q < nq; ++q)
                            image[(int64_t)q * nd + idx] += bin_q[p * nq + q] * pw;
                    }
                }
            }
        }
    }
}

int orc_cic_mapping_2D(const double *pos, const double *hsml,
                       const double *m, const double *rho,
                       const double *bin_q, int nq, const double *weights,
                       int64_t n, int64_t npix, double len2pix, int kernel,
                       int kernel_dim, int calc_mean, double *image)
{
    const int64_t nd = npix * npix;
    memset(image, 0, sizeof(double) * (size_t)(nd * (nq + 1)));
    /* cic_2D.jl:139-142: full-image-sized scratch, only the footprint used */
    double *wk = (double *)calloc((size_t)nd, sizeof(double));
    double *A = (double *)malloc(sizeof(double) * (size_t)nd);
    if (!wk || !A) {
        free(wk);
        free(A);
        return -1;
    }
    map2d_range(pos, hsml, m, rho, bin_q, nq, weights, 0, n, npix, len2pix,
                kernel, kernel_dim, calc_mean, image, wk, A);
    free(wk);
    free(A);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* cic_3D.jl:110-209                                                         */
/* ------------------------------------------------------------------------ */

static void map3d_range(const double *pos, const double *hsml, const double *m,
                        const double *rho, const double *bin_q,
                        const double *weights, int64_t p0, int64_t p1,
                        int64_t npix, double len2pix, int kernel, int kdim,
                        int calc_mean, double *image, double *wk, double *V,
                        double *grid_mass, double *particle_mass)
{
    const int64_t nd = npix * npix * npix;
    double gm = 0.0, pm = 0.0;
    const double l3 = len2pix * len2pix * len2pix;
    for (int64_t p = p0; p < p1; ++p) {
        double bq = bin_q[p];
        if (bq == 0.0 && !calc_mean)
            continue;
        /* get_quantities_3D, cic_3D.jl:87-97 */
        double h = hsml[p] * len2pix;
        double h_inv = 1.0 / h;
        double r = rho[p] / l3;
        double vol = m[p] / r;
        double w_los = weights[p];

        double x = pos[3 * p + 0] * len2pix;
        double y = pos[3 * p + 1] * len2pix;
        double z = pos[3 * p + 2] * len2pix;
        x = x + 0.5 * (double)npix;
        y = y + 0.5 * (double)npix;
        z = z + 0.5 * (double)npix;

        int64_t iMin, iMax, jMin, jMax, kMin, kMax;
        pix_index_min_max(x, h, npix, &iMin, &iMax);
        pix_index_min_max(y, h, npix, &jMin, &jMax);
        pix_index_min_max(z, h, npix, &kMin, &kMax);

        /* calculate_weights 3D, cic_3D.jl:13-78 */
        wsum s = {0.0, 0.0, 0, 0};
        for (int64_t i = iMin; i <= iMax; ++i) {
            double xd = centre_offset(x, i), dx = overlap_1d(x, h, i);
            for (int64_t j = jMin; j <= jMax; ++j) {
                double yd = centre_offset(y, j), dy = overlap_1d(y, h, j);
                for (int64_t k = kMin; k <= kMax; ++k) {
                    double zd = centre_offset(z, k), dzz = overlap_1d(z, h, k);
                    int64_t idx = i * npix * npix + j * npix + k;
                    double dxdydz = dx * dy * dzz;
                    double u = sqrt(xd * xd + yd * yd + zd * zd) * h_inv;
                    weight_per_pixel(&s, dxdydz, u, h_inv, kernel, kdim,
                                     &V[idx], &wk[idx]);
                }
            }
        }
        double wpp;
        if (s.distr_weight == 0.0) {
            s.n_distr = s.n_tot;
            for (int64_t i = iMin; i <= iMax; ++i)
                for (int64_t j = jMin; j <= jMax; ++j)
                    for (int64_t k = kMin; k <= kMax; ++k)
                        wk[i * npix * npix + j * npix + k] = 1.0;
            wpp = (s.distr_area != 0.0) ? (double)s.n_distr / s.distr_area : 1.0;
        } else {
            wpp = (double)s.n_distr / s.distr_weight;
        }
        /* cic_3D.jl:167-192 */
        double kernel_norm = vol / (double)s.n_distr;
        double volume_norm = kernel_norm * wpp * w_los * len2pix;
        for (int64_t i = iMin; i <= iMax; ++i)
            for (int64_t j = jMin; j <= jMax; ++j)
                for (int64_t k = kMin; k <= kMax; ++k) {
                    int64_t idx = i * npix * npix + j * npix + k;
                    double pw = wk[idx] * V[idx] * volume_norm;
                    if (pw != 0.0) {
                        image[nd + idx] += pw;
                        image[idx] += bq * pw;
                    }
                    gm += rho[p] * wk[idx] * V[idx] / l3;
                }
        pm += m[p];
    }
    if (grid_mass) *grid_mass = gm;
    if (particle_mass) *particle_mass = pm;
}

int orc_cic_mapping_3D(const double *pos, const double *hsml,
                       const double *m, const double *rho,
                       const double *bin_q, const double *weights, int64_t n,
                       int64_t npix, double len2pix, int kernel,
                       int kernel_dim, int calc_mean, double *image,
                       double *grid_mass, double *particle_mass)
{
    const int64_t nd = npix * npix * npix;
    memset(image, 0, sizeof(double) * (size_t)(2 * nd));
    double *wk = (double *)calloc((size_t)nd, sizeof(double));
    double *V = (double *)malloc(sizeof(double) * (size_t)nd);
    if (!wk || !V) {
        free(wk);
        free(V);
        return -1;
    }
    map3d_range(pos, hsml, m, rho, bin_q, weights, 0, n, npix, len2pix, kernel,
                kernel_dim, calc_mean, image, wk, V, grid_mass, particle_mass);
    free(wk);
    free(V);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Integer work of a deposit, for bit-exact checks of the GPU's               */
/* ------------------------------------------------------------------------ */

/* calculate_weights' fallback test (cic_2D.jl:49-69, cic_3D.jl:55-75):
 * distr_weight == 0, i.e. no footprint pixel with u <= 1 has a non-zero
 * kernel weight times overlap.  Every term is >= 0, so the loop stops at the
 * first positive one. */
static int fallback_2d(double x, double y, double h, int kernel, int64_t i0, int64_t i1,
                       int64_t j0, int64_t j1)
{
    const double h_inv = 1.0 / h;
    for (int64_t i = i0; i <= i1; ++i) {
        double xd = centre_offset(x, i), dx = overlap_1d(x, h, i);
        for (int64_t j = j0; j <= j1; ++j) {
            double yd = centre_offset(y, j), dy = overlap_1d(y, h, j);
            double u = sqrt(xd * xd + yd * yd) * h_inv;
            if (u <= 1.0 && orc_kernel_value(kernel, 2, u, h_inv) * (dx * dy) != 0.0)
                return 0;
        }
    }
    return 1;
}

static int fallback_3d(const double c[3], double h, int kernel, const int64_t lo[3],
                       const int64_t hi[3])
{
    const double h_inv = 1.0 / h;
    for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        double xd = centre_offset(c[0], i), dx = overlap_1d(c[0], h, i);
        for (int64_t j = lo[1]; j <= hi[1]; ++j) {
            double yd = centre_offset(c[1], j), dy = overlap_1d(c[1], h, j);
            for (int64_t k = lo[2]; k <= hi[2]; ++k) {
                double zd = centre_offset(c[2], k), dz = overlap_1d(c[2], h, k);
                double u = sqrt(xd * xd + yd * yd + zd * zd) * h_inv;
                if (u <= 1.0 && orc_kernel_value(kernel, 3, u, h_inv) * (dx * dy * dz) != 0.0)
                    return 0;
            }
        }
    }
    return 1;
}

/* Smallest hsml (pixels) for the closed-form normalisation of large
 * unclipped footprints, per kernel (Cubic, Quintic, WendlandC2..C8). */
static const double kMinH2[6] = {64.0, 24.0, 64.0, 28.0, 16.0, 12.0};
static const double kMinH3[6] = {40.0, 24.0, 32.0, 20.0, 14.0, 11.0};

int orc_footprints(int dims, const double *pos, const double *hsml, const double *bin_q,
                   int nq, int64_t n, int64_t npix, double len2pix, int kernel,
                   uint32_t flags, int32_t *box, int32_t *cls, int32_t *fb)
{
    const int nb = 2 * dims;
    for (int64_t p = 0; p < n; ++p) {
        int32_t *b = box + nb * p;
        for (int d = 0; d < dims; ++d) {
            b[2 * d] = 0;
            b[2 * d + 1] = -1;
        }
        cls[p] = ORC_CLS_SKIP;
        fb[p] = 0;
        /* cic_2D.jl:155-168, cic_3D.jl:140-144: skip all-zero quantities
         * unless calc_mean */
        int allzero = 1;
        for (int q = 0; q < nq; ++q)
            if (bin_q[p * nq + q] != 0.0)
                allzero = 0;
        if (allzero && !(flags & ORC_FLAG_CALC_MEAN))
            continue;
        /* get_quantities_2D/3D and get_xyz (cic_shared.jl:85-100) */
        const double h = hsml[p] * len2pix;
        double c[3];
        for (int d = 0; d < dims; ++d)
            c[d] = pos[3 * p + d] * len2pix + 0.5 * (double)npix;
        int bad = !(h > 0.0);
        int64_t a[3] = {0, 0, 0}, e[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < dims && !bad; ++d) {
            double fa = floor(c[d] - h), fe = floor(c[d] + h);
            if (fa != fa || fe != fe) {
                bad = 1;
                break;
            }
            /* beyond the image by more than a pixel: the same footprint */
            fa = fa < -1.0 ? -1.0 : fa > (double)npix + 1.0 ? (double)npix + 1.0 : fa;
            fe = fe < -1.0 ? -1.0 : fe > (double)npix + 1.0 ? (double)npix + 1.0 : fe;
            a[d] = (int64_t)fa;
            e[d] = (int64_t)fe;
            /* pix_index_min_max (cic_shared.jl:46-52) */
            lo[d] = a[d] > 0 ? a[d] : 0;
            hi[d] = e[d] < npix - 1 ? e[d] : npix - 1;
            if (lo[d] > hi[d])
                bad = 1;
        }
        if (bad) /* hsml <= 0, NaN, or an empty footprint: nothing deposited */
            continue;
        for (int d = 0; d < dims; ++d) {
            b[2 * d] = (int32_t)lo[d];
            b[2 * d + 1] = (int32_t)hi[d];
        }
        fb[p] = dims == 2 ? fallback_2d(c[0], c[1], h, kernel, lo[0], hi[0], lo[1], hi[1])
                          : fallback_3d(c, h, kernel, lo, hi);
        int small = 1, clipped = 0;
        const int64_t side = dims == 2 ? ORC_SMALL_SIDE_2D : ORC_SMALL_SIDE_3D;
        for (int d = 0; d < dims; ++d) {
            small = small && hi[d] - lo[d] + 1 <= side;
            clipped = clipped || a[d] < 0 || e[d] > npix - 1;
        }
        const double hmin = dims == 2 ? kMinH2[kernel] : kMinH3[kernel];
        if (small)
            cls[p] = ORC_CLS_SMALL;
        else if (!(flags & ORC_FLAG_EXACT_NORM) && !clipped && !fb[p] && h >= hmin)
            cls[p] = ORC_CLS_LARGE_CLOSED;
        else
            cls[p] = ORC_CLS_LARGE_EXPLICIT;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* parallel=true layout                                                      */
/* ------------------------------------------------------------------------ */

void orc_domain_decomposition(int64_t n, int nworkers, int i, int64_t *start,
                              int64_t *len)
{
    /* domain_decomp.jl:7-17: size = floor(N/P); the last batch takes the rest */
    int64_t size = n / nworkers;
    *start = (int64_t)i * size;
    *len = (i == nworkers - 1) ? n - (int64_t)(nworkers - 1) * size : size;
}

typedef struct {
    int dims;
    const double *pos, *hsml, *m, *rho, *bin_q, *weights;
    int nq;
    int64_t p0, p1, npix;
    double len2pix;
    int kernel, kdim, calc_mean;
    double *image; /* private image */
    int rc;
} job_t;

static void *run_job(void *arg)
{
    job_t *j = (job_t *)arg;
    int64_t nd = j->dims == 2 ? j->npix * j->npix : j->npix * j->npix * j->npix;
    int ncol = j->dims == 2 ? j->nq + 1 : 2;
    memset(j->image, 0, sizeof(double) * (size_t)(nd * ncol));
    double *wk = (double *)calloc((size_t)nd, sizeof(double));
    double *A = (double *)malloc(sizeof(double) * (size_t)nd);
    if (!wk || !A) {
        j->rc = -1;
        free(wk);
        free(A);
        return NULL;
    }
    if (j->dims == 2)
        map2d_range(j->pos, j->hsml, j->m, j->rho, j->bin_q, j->nq, j->weights,
                    j->p0, j->p1, j->npix, j->len2pix, j->kernel, j->kdim,
                    j->calc_mean, j->image, wk, A);
    else
        map3d_range(j->pos, j->hsml, j->m, j->rho, j->bin_q, j->weights, j->p0,
                    j->p1, j->npix, j->len2pix, j->kernel, j->kdim,
                    j->calc_mean, j->image, wk, A, NULL, NULL);
    free(wk);
    free(A);
    j->rc = 0;
    return NULL;
}

static int run_parallel(int dims, int nworkers, const double *pos,
                        const double *hsml, const double *m, const double *rho,
                        const double *bin_q, int nq, const double *weights,
                        int64_t n, int64_t npix, double len2pix, int kernel,
                        int kdim, int calc_mean, double *image)
{
    if (nworkers < 1)
        nworkers = 1;
    int64_t nd = dims == 2 ? npix * npix : npix * npix * npix;
    int ncol = dims == 2 ? nq + 1 : 2;
    size_t isz = (size_t)(nd * ncol);
    job_t *jobs = (job_t *)calloc((size_t)nworkers, sizeof(job_t));
    pthread_t *th = (pthread_t *)calloc((size_t)nworkers, sizeof(pthread_t));
    if (!jobs || !th) {
        free(jobs);
        free(th);
        return -1;
    }
    int rc = 0;
    for (int w = 0; w < nworkers; ++w) {
        int64_t s, l;
        orc_domain_decomposition(n, nworkers, w, &s, &l);
        job_t *j = &jobs[w];
        j->dims = dims; j->pos = pos; j->hsml = hsml; j->m = m; j->rho = rho;
        j->bin_q = bin_q; j->weights = weights; j->nq = nq;
        j->p0 = s; j->p1 = s + l; j->npix = npix; j->len2pix = len2pix;
        j->kernel = kernel; j->kdim = kdim; j->calc_mean = calc_mean;
        j->image = (double *)malloc(sizeof(double) * isz);
        j->rc = -1;
        if (!j->image)
            rc = -1;
    }
    if (rc == 0) {
        for (int w = 0; w < nworkers; ++w)
            pthread_create(&th[w], NULL, run_job, &jobs[w]);
        for (int w = 0; w < nworkers; ++w)
            pthread_join(th[w], NULL);
        for (int w = 0; w < nworkers; ++w)
            if (jobs[w].rc != 0)
                rc = -1;
    }
    if (rc == 0) {
        /* sum(fetch.(futures)): images summed left to right in worker order */
        memcpy(image, jobs[0].image, sizeof(double) * isz);
        for (int w = 1; w < nworkers; ++w) {
            const double *src = jobs[w].image;
            for (size_t t = 0; t < isz; ++t)
                image[t] = image[t] + src[t];
        }
    }
    for (int w = 0; w < nworkers; ++w)
        free(jobs[w].image);
    free(jobs);
    free(th);
    return rc;
}

int orc_cic_mapping_2D_parallel(int nworkers, const double *pos,
                                const double *hsml, const double *m,
                                const double *rho, const double *bin_q,
                                int nq, const double *weights, int64_t n,
                                int64_t npix, double len2pix, int kernel,
                                int kernel_dim, int calc_mean, double *image)
{
    return run_parallel(2, nworkers, pos, hsml, m, rho, bin_q, nq, weights, n,
                        npix, len2pix, kernel, kernel_dim, calc_mean, image);
}

int orc_cic_mapping_3D_parallel(int nworkers, const double *pos,
                                const double *hsml, const double *m,
                                const double *rho, const double *bin_q,
                                const double *weights, int64_t n,
                                int64_t npix, double len2pix, int kernel,
                                int kernel_dim, int calc_mean, double *image)
{
    return run_parallel(3, nworkers, pos, hsml, m, rho, bin_q, 1, weights, n,
                        npix, len2pix, kernel, kernel_dim, calc_mean, image);
}

/* ------------------------------------------------------------------------ */
/* reduce_image.jl                                                           */
/* ------------------------------------------------------------------------ */

void orc_reduce_image_2D(const double *image, int64_t npix, int nq,
                         int reduce_image, double *out)
{
    /* reduce_image.jl:8-31: k runs over the flat deposit index i*N+j; the
     * final transpose puts pixel (i=x, j=y) at Julia out[x,y,q]. */
    const int64_t nd = npix * npix;
    const double *w = image + (int64_t)nq * nd;
    for (int64_t i = 0; i < npix; ++i)
        for (int64_t j = 0; j < npix; ++j) {
            int64_t k = i * npix + j;
            for (int q = 0; q < nq; ++q) {
                double v = image[(int64_t)q * nd + k];
                if (reduce_image && w[k] > 0.0)
                    v /= w[k];
                out[(int64_t)q * nd + j * npix + i] = v;
            }
        }
}

void orc_reduce_image_3D(const double *image, int64_t npix, double *out)
{
    /* reduce_image.jl:39-55: same order as the deposit, [z,y,x] */
    const int64_t nd = npix * npix * npix;
    for (int64_t t = 0; t < nd; ++t) {
        double v = image[t];
        if (v > 0.0)
            v /= image[nd + t];
        out[t] = v;
    }
}

/* ------------------------------------------------------------------------ */
/* HEALPix RING scheme                                                       */
/* ------------------------------------------------------------------------ */

static int64_t isqrt64(int64_t v)
{
    int64_t r = (int64_t)sqrt((double)v + 0.5);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

int64_t orc_hp_npix(int64_t nside) { return 12 * nside * nside; }

static void hp_pix2zphi(int64_t nside, int64_t pix, double *z, double *phi)
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    const double fact2 = 4.0 / (double)npix;
    if (pix < ncap) {
        int64_t iring = (1 + isqrt64(1 + 2 * pix)) >> 1;
        int64_t iphi = (pix + 1) - 2 * iring * (iring - 1);
        *z = 1.0 - (double)(iring * iring) * fact2;
        *phi = ((double)iphi - 0.5) * (0.5 * M_PI) / (double)iring;
    } else if (pix < npix - ncap) {
        int64_t ip = pix - ncap;
        int64_t tmp = ip / (4 * nside);
        int64_t iring = tmp + nside;
        int64_t iphi = ip - tmp * 4 * nside + 1;
        double fodd = ((iring + nside) & 1) ? 1.0 : 0.5;
        *z = (double)(2 * nside - iring) * (2.0 / (3.0 * (double)nside));
        *phi = ((double)iphi - fodd) * M_PI / (2.0 * (double)nside);
    } else {
        int64_t ip = npix - pix;
        int64_t iring = (1 + isqrt64(2 * ip - 1)) >> 1;
        int64_t iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        *z = -1.0 + (double)(iring * iring) * fact2;
        *phi = ((double)iphi - 0.5) * (0.5 * M_PI) / (double)iring;
    }
}

void orc_hp_pix2ang_ring(int64_t nside, int64_t pix, double *theta, double *phi)
{
    double z;
    hp_pix2zphi(nside, pix, &z, phi);
    *theta = acos(z);
}

void orc_hp_pix2vec_ring(int64_t nside, int64_t pix, double vec[3])
{
    double z, phi;
    hp_pix2zphi(nside, pix, &z, &phi);
    double sth = sqrt((1.0 - z) * (1.0 + z));
    vec[0] = sth * cos(phi);
    vec[1] = sth * sin(phi);
    vec[2] = z;
}

void orc_hp_vec2ang(const double v[3], double *theta, double *phi)
{
    *theta = atan2(sqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
    double p = atan2(v[1], v[0]);
    if (p < 0.0)
        p += 2.0 * M_PI;
    *phi = p;
}

int64_t orc_hp_ang2pix_ring(int64_t nside, double theta, double phi)
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    double z = cos(theta), za = fabs(z);
    double tt = fmod(phi, 2.0 * M_PI);
    if (tt < 0.0)
        tt += 2.0 * M_PI;
    tt = tt / (0.5 * M_PI); /* in [0,4) */
    if (za <= 2.0 / 3.0) {
        double temp1 = (double)nside * (0.5 + tt);
        double temp2 = (double)nside * z * 0.75;
        int64_t jp = (int64_t)(temp1 - temp2);
        int64_t jm = (int64_t)(temp1 + temp2);
        int64_t ir = nside + 1 + jp - jm;
        int64_t kshift = 1 - (ir & 1);
        int64_t ip = (jp + jm - nside + kshift + 1) / 2;
        ip = ((ip % (4 * nside)) + 4 * nside) % (4 * nside);
        return ncap + (ir - 1) * 4 * nside + ip;
    }
    double tp = tt - (double)(int64_t)tt;
    double tmp = (double)nside * sqrt(3.0 * (1.0 - za));
    int64_t jp = (int64_t)(tp * tmp);
    int64_t jm = (int64_t)((1.0 - tp) * tmp);
    int64_t ir = jp + jm + 1;
    int64_t ip = (int64_t)(tt * (double)ir);
    ip = ((ip % (4 * ir)) + 4 * ir) % (4 * ir);
    if (z > 0)
        return 2 * ir * (ir - 1) + ip;
    return npix - 2 * ir * (ir + 1) + ip;
}

int64_t orc_hp_ring_above(int64_t nside, double z)
{
    double az = fabs(z);
    if (az <= 2.0 / 3.0)
        return (int64_t)((double)nside * (2.0 - 1.5 * z));
    int64_t iring = (int64_t)((double)nside * sqrt(3.0 * (1.0 - az)));
    return (z > 0) ? iring : 4 * nside - iring - 1;
}

static double hp_ring2z(int64_t nside, int64_t ring)
{
    const double fact2 = 4.0 / (double)(12 * nside * nside);
    if (ring < nside)
        return 1.0 - (double)(ring * ring) * fact2;
    if (ring <= 3 * nside)
        return (double)(2 * nside - ring) * (2.0 / (3.0 * (double)nside));
    int64_t r = 4 * nside - ring;
    return -1.0 + (double)(r * r) * fact2;
}

static void hp_ring_info(int64_t nside, int64_t ring, int64_t *start,
                         int64_t *nr, int *shifted)
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    if (ring < nside) {
        *shifted = 1;
        *nr = 4 * ring;
        *start = 2 * ring * (ring - 1);
    } else if (ring < 3 * nside) {
        *shifted = ((ring - nside) & 1) == 0;
        *nr = 4 * nside;
        *start = ncap + (ring - nside) * 4 * nside;
    } else {
        int64_t r = 4 * nside - ring;
        *shifted = 1;
        *nr = 4 * r;
        *start = npix - 2 * r * (r + 1);
    }
}

typedef struct {
    int64_t *out;
    int64_t cap, cnt;
} pixlist;

static void pl_range(pixlist *l, int64_t a, int64_t b) /* [a,b) */
{
    for (int64_t p = a; p < b; ++p) {
        if (l->cnt < l->cap)
            l->out[l->cnt] = p;
        l->cnt++;
    }
}

int64_t orc_hp_query_disc_ring(int64_t nside, double theta, double phi,
                               double radius, int64_t *out, int64_t cap)
{
    /* RING-scheme disc query (non-inclusive: pixel centres within radius),
     * ring by ring: rings fully inside around a pole are taken whole; every
     * other ring contributes the centres with |dphi| <= acos(...) */
    const int64_t npix = 12 * nside * nside;
    pixlist l = {out, cap, 0};
    if (radius >= M_PI) {
        pl_range(&l, 0, npix);
        return l.cnt;
    }
    double cosr = cos(radius);
    double z0 = cos(theta);
    double xa = 1.0 / sqrt((1.0 - z0) * (1.0 + z0));
    double rlat1 = theta - radius;
    double zmax = cos(rlat1);
    int64_t irmin = orc_hp_ring_above(nside, zmax) + 1;
    if (rlat1 <= 0.0 && irmin > 1) {
        int64_t sp, rp;
        int sh;
        hp_ring_info(nside, irmin - 1, &sp, &rp, &sh);
        pl_range(&l, 0, sp + rp);
    }
    double rlat2 = theta + radius;
    double zmin = cos(rlat2);
    int64_t irmax = orc_hp_ring_above(nside, zmin);
    for (int64_t iz = irmin; iz <= irmax; ++iz) {
        double z = hp_ring2z(nside, iz);
        double x = (cosr - z * z0) * xa;
        double ysq = 1.0 - z * z - x * x;
        double dphi;
        if (ysq <= 0.0)
            dphi = (x < 0.0) ? M_PI : 0.0;
        else
            dphi = atan2(sqrt(ysq), x);
        if (dphi > 0.0) {
            int64_t ipix1, nr;
            int sh;
            hp_ring_info(nside, iz, &ipix1, &nr, &sh);
            double shift = sh ? 0.5 : 0.0;
            int64_t ipix2 = ipix1 + nr - 1;
            double inv2pi = 1.0 / (2.0 * M_PI);
            int64_t ip_lo = (int64_t)floor((double)nr * inv2pi * (phi - dphi) - shift) + 1;
            int64_t ip_hi = (int64_t)floor((double)nr * inv2pi * (phi + dphi) - shift);
            if (ip_lo <= ip_hi) {
                if (ip_hi >= nr) {
                    ip_lo -= nr;
                    ip_hi -= nr;
                }
                if (ip_lo < 0) {
                    pl_range(&l, ipix1, ipix1 + ip_hi + 1);
                    pl_range(&l, ipix1 + ip_lo + nr, ipix2 + 1);
                } else {
                    pl_range(&l, ipix1 + ip_lo, ipix1 + ip_hi + 1);
                }
            }
        }
    }
    if (rlat2 >= M_PI && irmax + 1 < 4 * nside) {
        int64_t sp, rp;
        int sh;
        hp_ring_info(nside, irmax + 1, &sp, &rp, &sh);
        pl_range(&l, sp, npix);
    }
    return l.cnt;
}

/* ------------------------------------------------------------------------ */
/* healpix_interpolation                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t *pix;
    double *wk, *A;
    int64_t cap;
} hp_scratch;

static int hp_grow(hp_scratch *s, int64_t need)
{
    if (need <= s->cap)
        return 0;
    int64_t c = s->cap ? s->cap : 1024;
    while (c < need) c *= 2;
    int64_t *p = (int64_t *)realloc(s->pix, sizeof(int64_t) * (size_t)c);
    if (!p) return -1;
    s->pix = p;
    double *a = (double *)realloc(s->wk, sizeof(double) * (size_t)c);
    if (!a) return -1;
    s->wk = a;
    double *b = (double *)realloc(s->A, sizeof(double) * (size_t)c);
    if (!b) return -1;
    s->A = b;
    s->cap = c;
    return 0;
}

/* The weights of one particle (constributing_pixels.jl:7-22 and
 * pixel_weights.jl:6-140): its pixel list (the disc query plus the pixel
 * holding the particle, deduplicated), per pixel the area A (with the
 * /(ang_pix*Dx)^2 of weight_per_index) and the kernel weight wk (1 on every
 * pixel in the fallback), the fallback flag, the pixel count n_distr and
 * weight_per_pix.  Returns the pixel count, 0 for a skipped particle
 * (Dx < hsml, main.jl:172-179), -1 if out of memory. */
static int64_t hp_particle(const double pos[3], double hsml, int64_t nside, double ang_pix,
                           int kernel, int kdim, hp_scratch *s, double *Dx_out,
                           double *proj_out, double *wpp_out, int64_t *n_distr_out,
                           int *fallback_out)
{
    /* shared.jl:1-10 */
    double dx2 = 0.0;
    for (int d = 0; d < 3; ++d)
        dx2 += pos[d] * pos[d];
    double Dx = sqrt(dx2);
    *Dx_out = Dx;
    if (Dx < hsml) /* main.jl:172-179 */
        return 0;
    double proj_hsml = asin(hsml / Dx); /* main.jl:182 */
    *proj_out = proj_hsml;

    /* constributing_pixels.jl:7-22 */
    double theta, phi;
    orc_hp_vec2ang(pos, &theta, &phi);
    int64_t cnt = orc_hp_query_disc_ring(nside, theta, phi, proj_hsml, s->pix, s->cap);
    if (cnt + 1 > s->cap) {
        if (hp_grow(s, cnt + 1))
            return -1;
        cnt = orc_hp_query_disc_ring(nside, theta, phi, proj_hsml, s->pix, s->cap);
    }
    int64_t cpix = orc_hp_ang2pix_ring(nside, theta, phi);
    int found = 0;
    for (int64_t t = 0; t < cnt; ++t)
        if (s->pix[t] == cpix) { found = 1; break; }
    if (!found)
        s->pix[cnt++] = cpix; /* push! then unique! keeps first occurrences */

    double ap = ang_pix * Dx;
    /* pixel_weights.jl:87-140 */
    double h_inv = 1.0 / proj_hsml;
    wsum sm = {0.0, 0.0, 0, 0};
    for (int64_t t = 0; t < cnt; ++t) {
        double pc[3];
        orc_hp_pix2vec_ring(nside, s->pix[t], pc);
        double d = 0.0;
        for (int k = 0; k < 3; ++k)
            d += pos[k] * pc[k];
        double c = d / Dx;
        double dist = acos(c < 1.0 ? c : 1.0); /* pixel_weights.jl:16-22 */
        double u = dist * h_inv;
        /* pixel_weights.jl:6-8 */
        double a = fabs(proj_hsml - (dist - 0.5 * ang_pix));
        a = a < ang_pix ? a : ang_pix;
        a = a > 0.0 ? a : 0.0;
        a = a / ang_pix;
        a /= ap * ap; /* pixel_weights.jl:52 */
        weight_per_pixel(&sm, a, u, h_inv, kernel, kdim, &s->A[t], &s->wk[t]);
    }
    double wpp;
    int fb = sm.distr_weight == 0.0;
    if (fb) {
        sm.n_distr = sm.n_tot;
        for (int64_t t = 0; t < cnt; ++t)
            s->wk[t] = 1.0;
        wpp = (sm.distr_area != 0.0) ? (double)sm.n_distr / sm.distr_area : 1.0;
    } else {
        wpp = (double)sm.n_distr / sm.distr_weight;
    }
    *wpp_out = wpp;
    *n_distr_out = sm.n_distr;
    *fallback_out = fb;
    return cnt;
}

int64_t orc_hp_particle_weights(const double pos[3], double hsml, int64_t nside, int kernel,
                                int kernel_dim, int64_t cap, int64_t *pix, double *A,
                                double *wk, double *info)
{
    hp_scratch s = {NULL, NULL, NULL, 0};
    if (hp_grow(&s, 4096))
        return -1;
    const double ang_pix = sqrt(4.0 * M_PI / (double)orc_hp_npix(nside));
    double Dx = 0.0, proj = 0.0, wpp = 0.0;
    int64_t nd = 0;
    int fb = 0;
    int64_t cnt = hp_particle(pos, hsml, nside, ang_pix, kernel, kernel_dim, &s, &Dx, &proj,
                              &wpp, &nd, &fb);
    if (cnt > 0 && cnt <= cap) {
        memcpy(pix, s.pix, sizeof(int64_t) * (size_t)cnt);
        memcpy(A, s.A, sizeof(double) * (size_t)cnt);
        memcpy(wk, s.wk, sizeof(double) * (size_t)cnt);
    }
    if (info) {
        info[0] = Dx;
        info[1] = proj;
        info[2] = ang_pix;
        info[3] = wpp;
        info[4] = (double)nd;
        info[5] = (double)fb;
    }
    free(s.pix);
    free(s.wk);
    free(s.A);
    return cnt;
}

/* one particle of the loop at main.jl:158-213 */
static int hp_deposit_one(const double pos[3], double hsml, double m,
                          double rho, double bq, double w, int64_t nside,
                          double ang_pix, int kernel, int kdim, int calc_mean,
                          double *allsky, double *wmap, hp_scratch *s)
{
    if (!calc_mean && bq == 0.0)
        return 0;
    double Dx = 0.0, proj = 0.0, wpp = 0.0;
    int64_t n_distr = 0;
    int fb = 0;
    int64_t cnt = hp_particle(pos, hsml, nside, ang_pix, kernel, kdim, s, &Dx, &proj, &wpp,
                              &n_distr, &fb);
    if (cnt <= 0)
        return cnt < 0 ? -1 : 0;
    /* main.jl:56-63, 193 */
    double dz = 2.0 * hsml;
    double area = (m / rho) / dz;
    double ap = ang_pix * Dx;
    dz /= ap * ap;
    /* main.jl:25-45 */
    double kernel_norm = area / (double)n_distr;
    double area_norm = kernel_norm * wpp * w * dz;
    for (int64_t t = 0; t < cnt; ++t) {
        double pw = area_norm * s->wk[t] * s->A[t];
        allsky[s->pix[t]] += bq * pw;
        wmap[s->pix[t]] += pw;
    }
    return 0;
}

typedef struct {
    double r;
    int64_t i;
} ridx;

static int cmp_ridx(const void *a, const void *b)
{
    const ridx *x = (const ridx *)a, *y = (const ridx *)b;
    if (x->r < y->r) return -1;
    if (x->r > y->r) return 1;
    return (x->i < y->i) ? -1 : (x->i > y->i);
}

int orc_healpix_map(double *pos, const double *hsml, const double *m,
                    const double *rho, const double *bin_q,
                    const double *weights, int64_t n, const double center[3],
                    const double radius_limits[2], int64_t nside, int kernel,
                    int kernel_dim, int calc_mean, double *allsky_map,
                    double *weight_map)
{
    const int64_t npix = 12 * nside * nside;
    memset(allsky_map, 0, sizeof(double) * (size_t)npix);
    memset(weight_map, 0, sizeof(double) * (size_t)npix);
    /* main.jl:114-121 */
    if (!calc_mean) {
        double sq = 0.0;
        for (int64_t i = 0; i < n; ++i) sq += bin_q[i];
        if (sq == 0.0)
            return 0;
    }
    /* filter_particles.jl:17-54 */
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d)
            pos[3 * i + d] = pos[3 * i + d] - center[d];
    ridx *ord = (ridx *)malloc(sizeof(ridx) * (size_t)(n ? n : 1));
    uint8_t *sel = (uint8_t *)malloc((size_t)(n ? n : 1));
    if (!ord || !sel) {
        free(ord);
        free(sel);
        return -2;
    }
    int64_t nsel = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double *p = pos + 3 * i;
        double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        ord[i].r = r;
        ord[i].i = i;
        sel[i] = (radius_limits[0] <= r && r <= radius_limits[1]);
        nsel += sel[i];
    }
    if (!calc_mean) {
        /* sel[Bin_q[sel] .> 0] is only in bounds when nothing was removed,
         * and sorted[sel] then needs every bin_q > 0 */
        int ok = (nsel == n);
        for (int64_t i = 0; ok && i < n; ++i)
            if (!(bin_q[i] > 0.0)) ok = 0;
        if (!ok) {
            free(ord);
            free(sel);
            return -1;
        }
    }
    /* sortperm (stable ascending) then reverse: descending radius */
    qsort(ord, (size_t)n, sizeof(ridx), cmp_ridx);
    hp_scratch s = {NULL, NULL, NULL, 0};
    if (hp_grow(&s, 4096)) {
        free(ord);
        free(sel);
        return -2;
    }
    const double ang_pix = sqrt(4.0 * M_PI / (double)npix); /* main.jl:144 */
    int rc = 0;
    /* sorted[sel]: the original-order mask applied to the sorted order */
    for (int64_t t = 0; t < n && rc == 0; ++t) {
        if (!sel[t])
            continue;
        int64_t i = ord[n - 1 - t].i;
        rc = hp_deposit_one(pos + 3 * i, hsml[i], m[i], rho[i], bin_q[i],
                            weights[i], nside, ang_pix, kernel, kernel_dim,
                            calc_mean, allsky_map, weight_map, &s);
    }
    free(s.pix);
    free(s.wk);
    free(s.A);
    free(ord);
    free(sel);
    return rc ? -2 : 0;
}

typedef struct {
    const double *pos, *hsml, *m, *rho, *bin_q, *weights;
    int64_t p0, p1, nside;
    int kernel, kdim, calc_mean;
    double *map, *wmap;
    int rc;
} hp_job;

static void *hp_run(void *arg)
{
    hp_job *j = (hp_job *)arg;
    const int64_t npix = 12 * j->nside * j->nside;
    const double ang_pix = sqrt(4.0 * M_PI / (double)npix);
    memset(j->map, 0, sizeof(double) * (size_t)npix);
    memset(j->wmap, 0, sizeof(double) * (size_t)npix);
    hp_scratch s = {NULL, NULL, NULL, 0};
    j->rc = hp_grow(&s, 4096);
    for (int64_t i = j->p0; i < j->p1 && j->rc == 0; ++i)
        j->rc = hp_deposit_one(j->pos + 3 * i, j->hsml[i], j->m[i], j->rho[i],
                               j->bin_q[i], j->weights[i], j->nside, ang_pix,
                               j->kernel, j->kdim, j->calc_mean, j->map,
                               j->wmap, &s);
    free(s.pix);
    free(s.wk);
    free(s.A);
    return NULL;
}

int orc_healpix_deposit_parallel(int nworkers, const double *pos,
                                 const double *hsml, const double *m,
                                 const double *rho, const double *bin_q,
                                 const double *weights, int64_t n,
                                 int64_t nside, int kernel, int kernel_dim,
                                 int calc_mean, double *allsky_map,
                                 double *weight_map)
{
    if (nworkers < 1)
        nworkers = 1;
    const int64_t npix = 12 * nside * nside;
    hp_job *jobs = (hp_job *)calloc((size_t)nworkers, sizeof(hp_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nworkers, sizeof(pthread_t));
    int rc = (jobs && th) ? 0 : -1;
    for (int w = 0; rc == 0 && w < nworkers; ++w) {
        int64_t s0, l;
        orc_domain_decomposition(n, nworkers, w, &s0, &l);
        hp_job *j = &jobs[w];
        j->pos = pos; j->hsml = hsml; j->m = m; j->rho = rho; j->bin_q = bin_q;
        j->weights = weights; j->p0 = s0; j->p1 = s0 + l; j->nside = nside;
        j->kernel = kernel; j->kdim = kernel_dim; j->calc_mean = calc_mean;
        j->map = (double *)malloc(sizeof(double) * (size_t)npix);
        j->wmap = (double *)malloc(sizeof(double) * (size_t)npix);
        if (!j->map || !j->wmap) rc = -1;
    }
    if (rc == 0) {
        for (int w = 0; w < nworkers; ++w)
            pthread_create(&th[w], NULL, hp_run, &jobs[w]);
        for (int w = 0; w < nworkers; ++w)
            pthread_join(th[w], NULL);
        for (int w = 0; w < nworkers; ++w)
            if (jobs[w].rc) rc = -1;
    }
    if (rc == 0) {
        memcpy(allsky_map, jobs[0].map, sizeof(double) * (size_t)npix);
        memcpy(weight_map, jobs[0].wmap, sizeof(double) * (size_t)npix);
        for (int w = 1; w < nworkers; ++w)
            for (int64_t t = 0; t < npix; ++t) {
                allsky_map[t] += jobs[w].map[t];
                weight_map[t] += jobs[w].wmap[t];
            }
    }
    if (jobs)
        for (int w = 0; w < nworkers; ++w) {
            free(jobs[w].map);
            free(jobs[w].wmap);
        }
    free(jobs);
    free(th);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* TSC                                                                       */
/* ------------------------------------------------------------------------ */

static inline double tsc_w(double d)
{
    double a = fabs(d);
    if (a < 0.5)
        return 0.75 - a * a;
    if (a < 1.5) {
        double b = 1.5 - a;
        return 0.5 * b * b;
    }
    return 0.0;
}

int orc_tsc_3d(const double *pos, const double *bin_q, int64_t n, int64_t npix,
               double len2pix, int periodic, double *grid)
{
    /* grid: 2 planes of npix^3 in the 3D deposit order (i*N*N + j*N + k):
     * plane 0 = sum q*W, plane 1 = sum W.  Positions map to grid units as in
     * get_xyz (cic_shared.jl:85-100); cell c has its centre at c + 0.5. */
    const int64_t nd = npix * npix * npix;
    memset(grid, 0, sizeof(double) * (size_t)(2 * nd));
    for (int64_t p = 0; p < n; ++p) {
        double g[3];
        int64_t c[3];
        double w[3][3];
        for (int d = 0; d < 3; ++d) {
            g[d] = pos[3 * p + d] * len2pix;
            g[d] = g[d] + 0.5 * (double)npix;
            c[d] = (int64_t)floor(g[d]);
            for (int o = -1; o <= 1; ++o)
                w[d][o + 1] = tsc_w(g[d] - ((double)(c[d] + o) + 0.5));
        }
        for (int a = -1; a <= 1; ++a) {
            int64_t i = c[0] + a;
            if (periodic) i = ((i % npix) + npix) % npix;
            else if (i < 0 || i >= npix) continue;
            for (int b = -1; b <= 1; ++b) {
                int64_t j = c[1] + b;
                if (periodic) j = ((j % npix) + npix) % npix;
                else if (j < 0 || j >= npix) continue;
                for (int e = -1; e <= 1; ++e) {
                    int64_t k = c[2] + e;
                    if (periodic) k = ((k % npix) + npix) % npix;
                    else if (k < 0 || k >= npix) continue;
                    double ww = w[0][a + 1] * w[1][b + 1] * w[2][e + 1];
                    int64_t idx = i * npix * npix + j * npix + k;
                    grid[idx] += bin_q[p] * ww;
                    grid[nd + idx] += ww;
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* mass_density (effects/mass_density.jl:30-99)                             */
/* ------------------------------------------------------------------------ */

/* Dehnen & Aly (2012, MNRAS 425, 1068) eqs. 18-19: the self-contribution
 * bias of the Wendland kernels, eps = eps100 * (N / 100)^(-alpha), subtracted
 * as eps * m * W(0) when that leaves a positive density (SPHKernels.jl
 * bias_correction; called at mass_density.jl:86).  Other kernels: none. */
double orc_bias_correction(int kernel, int dim, double density, double m, double h_inv,
                           int64_t nneighbors)
{
    double e100, alpha;
    switch (kernel) {
    case ORC_KERNEL_WENDLAND_C2: e100 = 0.0294; alpha = 0.977; break;
    case ORC_KERNEL_WENDLAND_C4: e100 = 0.01342; alpha = 1.579; break;
    case ORC_KERNEL_WENDLAND_C6: e100 = 0.0116; alpha = 2.236; break;
    default: return density;
    }
    double hd = h_inv;
    for (int d = 1; d < dim; ++d)
        hd *= h_inv;
    const double corr = e100 * pow((double)nneighbors * 0.01, -alpha) * m *
                        orc_kernel_norm(kernel, dim) * hd;
    return density - corr > 0.0 ? density - corr : density;
}

static double md_dist(const double *a, const double *b, const double *box, int periodic)
{
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        double x = a[d] - b[d];
        if (periodic) {
            /* Distances.jl PeriodicEuclidean: minimum image per dimension */
            x = fabs(x);
            x = fmod(x, box[d]);
            if (x > 0.5 * box[d])
                x = box[d] - x;
        }
        s += x * x;
    }
    return sqrt(s);
}

/* k-th smallest (0-based) of v[0..n) by quickselect (v is reordered) */
static double md_select(double *v, int64_t n, int64_t k)
{
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const double piv = v[lo + (hi - lo) / 2];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (v[i] < piv) ++i;
            while (v[j] > piv) --j;
            if (i <= j) {
                const double t = v[i];
                v[i] = v[j];
                v[j] = t;
                ++i;
                --j;
            }
        }
        if (k <= j)
            hi = j;
        else if (k >= i)
            lo = i;
        else
            return v[k];
    }
    return v[k];
}

int orc_mass_density(const double *pos, const double *mass, int64_t n, int kernel, int dim,
                     int64_t nneighbors, const double *boxsize, double *rho, double *hsml)
{
    return orc_mass_density_subset(pos, mass, n, NULL, n, kernel, dim, nneighbors, boxsize, rho,
                                   hsml);
}

int orc_mass_density_subset(const double *pos, const double *mass, int64_t n,
                            const int64_t *idx, int64_t nidx, int kernel, int dim,
                            int64_t nneighbors, const double *boxsize, double *rho,
                            double *hsml)
{
    const int64_t k1 = nneighbors + 1;
    if (n < k1 || nneighbors < 1)
        return -1;
    const int periodic = boxsize && boxsize[0] > 0.0 && boxsize[1] > 0.0 && boxsize[2] > 0.0;
    double *d = (double *)malloc(sizeof(double) * (size_t)n);
    double *work = (double *)malloc(sizeof(double) * (size_t)n);
    if (!d || !work) {
        free(d);
        free(work);
        return -2;
    }
    for (int64_t a = 0; a < nidx; ++a) {
        const int64_t i = idx ? idx[a] : a;
        for (int64_t j = 0; j < n; ++j)
            d[j] = md_dist(pos + 3 * i, pos + 3 * j, boxsize, periodic);
        memcpy(work, d, sizeof(double) * (size_t)n);
        /* mass_density.jl:64-68: hsml = the largest of the k1 nearest distances */
        const double h = md_select(work, n, k1 - 1);
        const double h_inv = 1.0 / h;
        /* mass_density.jl:73-81: sum over the neighbours; those at distance h
           have u = 1 and W = 0, so ties there do not matter */
        double r = 0.0;
        for (int64_t j = 0; j < n; ++j)
            if (d[j] < h)
                r += mass[j] * orc_kernel_value(kernel, dim, d[j] * h_inv, h_inv);
        hsml[a] = h;
        rho[a] = orc_bias_correction(kernel, dim, r, mass[i], h_inv, nneighbors);
    }
    free(d);
    free(work);
    return 0;
}
