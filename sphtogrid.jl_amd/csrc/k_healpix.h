// k_healpix.h -- launchers of the all-sky (HEALPix RING) deposit kernels.
#pragma once

#include "s2g_internal.h"

namespace s2g {

// per-particle record of the all-sky deposit
struct HpRec {
    double z0, phi0, theta0, sth0, xa;  // direction (cos, phi, colatitude, sin) and 1/sin
    double ch0, sh0;                    // cos, sin of theta0 / 2
    double theta, cosr;                 // projected hsml asin(h/dx) (main.jl:182), its cos
    double s;                           // area*dz, then divided by the normalisation sum
                                        // (a map's pixel scale is s times its weight)
    int64_t cpix;                       // RING pixel containing the particle
    int64_t unit0;                      // first work unit
    int32_t ra, rb;                     // ring range incl. polar caps fully inside
    int32_t irmin, irmax;               // rings with a partial phi interval
    int32_t cring, ck;                  // ring (1-based) and ring index of cpix
    int32_t fb, ok;
};

// maps of one particle set deposited together (the disc queries, the
// normalisation and the ring lists depend only on positions and hsml)
constexpr int kHpMaxMaps = 8;

struct HpArgs {
    int dtype, kernel;
    const void *pos, *hsml, *m, *rho;
    int nmaps;
    const void *binq[kHpMaxMaps], *w[kHpMaxMaps];  // per map
    int64_t n, nside;
    uint32_t flags;
};

// maps [m0, m0 + nm) of an HpArgs, nm <= 2, and their outputs
struct HpGroup {
    const void *binq[2], *w[2];
    double *map[2], *wmap[2];
    int nm;
    const double4 *rtab;  // per-ring terms (launch_hp_ring_table)
};

// nunits: n+1 int64 (last = 0, the total after an exclusive scan)
void launch_hp_setup(const HpArgs &a, HpRec *rec, int64_t *nunits, hipStream_t s);
void launch_hp_units(const int64_t *ubase, int64_t n, HpRec *rec, int2 *units, hipStream_t s);
// rtab: 4 Nside double4, ring iz at rtab[iz] = (sin(theta/2), cos(theta/2),
// sin(theta), cos(theta)) of the ring's colatitude, for iz in [1, 4 Nside - 1]
void launch_hp_ring_table(int64_t nside, double4 *rtab, hipStream_t s);
void launch_hp_norm(int kernel, const int2 *units, int64_t nu, const HpRec *rec, int64_t nside,
                    const double4 *rtab, double4 *part, hipStream_t s);
void launch_hp_finish(const int64_t *ubase, int64_t n, const double4 *part, HpRec *rec,
                      hipStream_t s);
// per-unit atomic deposit of a group of maps (maps zeroed by the caller)
void launch_hp_deposit(const HpArgs &a, const HpGroup &g, const int2 *units, int64_t nu,
                       const HpRec *rec, hipStream_t s);

// ring-major deposit (no global atomics) for 4 Nside <= its LDS ring size
bool hp_ring_deposit_fits(int64_t nside);
// per-ring particle lists: entry keys (units + one extra entry per particle,
// nu + n of each), sorted with sort_pairs over hp_entry_key_bits(nside) bits
void launch_hp_entry_keys(const int2 *units, int64_t nu, const HpRec *rec, int64_t n,
                          int64_t nside, uint32_t *keys, int32_t *vals, hipStream_t s);
int hp_entry_key_bits(int64_t nside);
// off: 4 Nside + 1 first-ring offsets into the sorted entries
void launch_hp_entry_offsets(const uint32_t *keys, int64_t ne, int64_t nside, int64_t *off,
                             hipStream_t s);
// count: 4 Nside + 1 int64 (entries touching each ring; ends 0)
void launch_hp_ring_count(const uint32_t *keys, const int64_t *eoff, int64_t nside,
                          int64_t *count, hipStream_t s);
// loff: exclusive scan of count; vals: the sorted entries' particles; list:
// the particles per ring
void launch_hp_ring_fill(const uint32_t *keys, const int32_t *vals, const int64_t *eoff,
                         const int64_t *loff, int64_t nside, int32_t *list, hipStream_t s);
// writes every pixel of the group's maps
void launch_hp_ring_deposit(const HpArgs &a, const HpGroup &g, const int32_t *list,
                            const int64_t *off, const HpRec *rec, hipStream_t s);

}  // namespace s2g
