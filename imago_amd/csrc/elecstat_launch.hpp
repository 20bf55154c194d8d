// elecstat_launch.hpp - host interface of the potential-side electrostatics
// (elecstat.hip): neutralAndNuclearQPot and residualQ of elecStat.f90.
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

// What the kernels read of the system; device pointers unless noted.
struct EsArgs {
    int nsites, npt, pot_dim, ncells, maxpa;
    const double *site_pos;     // [site][3]
    const int *site_type;       // [site]
    const int *type_first;      // [type] first site (sites of a type are contiguous)
    const int *type_mult;       // [type] sites of the type
    const int *type_nalpha;     // [type]
    const double *type_alphas;  // [type][maxpa] ascending
    const double *type_z;       // [type] nuclear charge
    const double *type_nuc;     // [type] nuclear alpha
    const int *type_term0;      // [type] first term in potDim
    const int *term_type, *term_alpha;   // [potDim]
    const double *cell_dims;    // [cell][3] real-space cells, |R|^2 ascending
    const double *cell_sizes;   // [cell]
    double rv[9], inv[9];       // real vectors (columns) and their inverse (rows), by value
    double th;                  // -ln(electroConvg)
    double4 *msep;              // scratch [site][site]: minimal separation, |r_i - r_j|^2
};

// neutralAndNuclearQPot (elecStat.f90:59-664): ovl, nl_neut, l_neut are
// potDim x potDim column-major (element (term of site j, term of site i));
// nl_nuc, l_nuc [potDim].  Every element is written.
hipError_t launch_es_neutral_nuclear(const EsArgs &A, double *ovl, double *nl_neut,
                                     double *l_neut, double *nl_nuc, double *l_nuc,
                                     hipStream_t stream);

// residualQ (elecStat.f90:665-1043).
struct EsRecip {
    int ncr;                    // reciprocal cells summed (the origin left out)
    const double *cdr;          // [ncr][3]
    const double *csr;          // [ncr] |G|^2
    const double *cw;           // [ncr] 2 for a cell standing for +G and -G, else 1
    int nsrc;                   // element runs: the types whose alphas key the tables
    const int *src_type;        // [nsrc] type of each run's first site
    const int *site_src;        // [site] run of the site
    const int *ord_type;        // [npt] type of each ordinal (types in site order)
    int nblk;                   // blocks of consecutive sites within one run
    const int *blk;             // [nblk][3] first site, sites, run
    double min_pot, volume;     // minPotAlpha, realCellVolume
    // scratch (device): E [nsrc][ncr][nk], CS [ncr][nsites] (cos, sin),
    // S [chunks][nsites][nk][nsites]; sizes from es_recip_*_doubles()
    double *E;
    double2 *CS;
    double *S;
    int nk, chunks;
};

// Plan of the reciprocal sum: alphas padded to nk in the tables, bi sites
// i per block (bi * nk accumulators a lane), cell chunks summed apart.
inline int es_recip_nk(int maxpa) { return maxpa <= 16 ? 16 : maxpa <= 32 ? 32 : 64; }
inline int es_recip_bi(int nk) { return 64 / nk; }
inline int es_recip_chunks(int nsites, int nblk, int ncr)
{
    const long waves = (long)((nsites + 63) / 64) * nblk;
    long z = (4096 + waves - 1) / waves;
    const long zmax = ncr / 64 > 1 ? ncr / 64 : 1;
    z = z < 1 ? 1 : z > zmax ? zmax : z;
    return (int)(z > 65535 ? 65535 : z);
}
inline size_t es_recip_e_doubles(int nsrc, int ncr, int nk) { return (size_t)nsrc * ncr * nk; }
inline size_t es_recip_s_doubles(int chunks, int nsites, int nk)
{
    return (size_t)chunks * nsites * nk * nsites;
}

// out: numPotTypes x potDim column-major, every element written.  The
// events (may be null) are recorded around the reciprocal-sum kernel.
hipError_t launch_es_residual_q(const EsArgs &A, const EsRecip &R, double *out,
                                hipStream_t stream, hipEvent_t recip0, hipEvent_t recip1);

}  // namespace imago
