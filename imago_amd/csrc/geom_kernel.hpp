// geom_kernel.hpp - nearest lattice vector and the 3-centre candidate
// images of every (atom-pair image, potential site).
//
// For the nuclear (nuclearPE, integrals.F90:2533-2697) and electronic
// (electronicPE :2700-2856) 3-centre terms the reference walks, for every
// alpha pair of an image, the lattice images C + L2 + R of every potential
// site with L2 = findLatticeVector(G - C) and keeps |((G-C)-L2)-R|^2 <=
// threeAlphaDist.  G lies on the segment [B', A], so the images that can
// pass for any alpha pair are those within sqrt(bound) of that segment,
// bound being the largest threshold of the site's element; k_geom_cands
// lists exactly those (a superset of every alpha pair's survivors), and the
// assembly kernels apply the reference's test to each.
#pragma once
#include "assemble.hpp"

namespace imago {

// findLatticeVector (lattice.f90:784-890): the lattice point nearest to v.
// Fractional coordinates are rounded half away from zero, then the closest
// of the cells (sorted by squared length, cell 0 the origin) to the
// remainder r is taken, the first of equals.  A cell R with |R| > |r| +
// sqrt(best) is farther from r than the best so far, and so are all the
// cells after it, so the scan stops there (and after at most 125 cells).
IMAGO_HD void find_lattice_vector(const double *rv, const double *inv, const double *cell_dims,
                                  const double *cell_sizes, int ncells, const double *v,
                                  double *L)
{
#pragma clang fp contract(off)
    double L0[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const double f = inv[3 * i] * v[0] + inv[3 * i + 1] * v[1] + inv[3 * i + 2] * v[2];
        double rnd = (double)(long)(fabs(f) + 0.5);
        if (f < 0.0) rnd = -rnd;
        for (int x = 0; x < 3; ++x) L0[x] += rv[3 * i + x] * rnd;
    }
    const double rem[3] = {v[0] - L0[0], v[1] - L0[1], v[2] - L0[2]};
    const double rr = sqrt(rem[0] * rem[0] + rem[1] * rem[1] + rem[2] * rem[2]);
    double stop = (2.0 * rr + 1e-9) * (2.0 * rr + 1e-9);
    double best = 1e300;
    int bi = 0;
    const int lim = ncells < 125 ? ncells : 125;
    for (int k = 0; k < lim; ++k) {
        if (cell_sizes[k] > stop) break;
        const double *R = cell_dims + 3 * k;
        double d = 0;
        for (int x = 0; x < 3; ++x) d += (rem[x] - R[x]) * (rem[x] - R[x]);
        if (d < best) {
            best = d;
            bi = k;
            const double s = rr + sqrt(best) + 1e-9;
            if (s * s < stop) stop = s * s;
        }
    }
    for (int x = 0; x < 3; ++x) L[x] = L0[x] + cell_dims[3 * bi + x];
}

struct GeomArgs {
    const PairRec *pairs;
    const ImgRec *imgs;
    const int *img_pair;       // [nimgs] pair of each image
    const double *atom_pos;    // [atoms][3]
    const int *atom_type, *type_element;
    const double *site_pos;    // [sites][3]
    const int *site_type, *pot_el;
    const double *adist;       // [e1][e2][a1][a2]
    const double *bound0;      // [pot element][e1][e2] largest 3-centre threshold
    const double *cell_dims, *cell_sizes;
    double rv[9], inv[9];
    int ncells, nsites, nel, maxa;
    long nimgs;
};

// One thread per (image, site).  FILL == false: cnt[img * nsites + site] =
// number of candidate images, *err |= 1 when the cell list is too short
// (integrals.F90:378).  FILL == true: the candidates at off[...].
template <bool FILL>
__global__ __launch_bounds__(256) void k_geom_cands(GeomArgs G, int *__restrict__ cnt,
                                                    const int *__restrict__ off,
                                                    CandRec *__restrict__ out,
                                                    int *__restrict__ err)
{
#pragma clang fp contract(off)
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G.nimgs * G.nsites) return;
    const long img = t / G.nsites;
    const int s2 = (int)(t % G.nsites);
    const PairRec pr = G.pairs[G.img_pair[img]];
    const ImgRec ir = G.imgs[img];
    const int e1 = G.type_element[G.atom_type[pr.i]], e2 = G.type_element[G.atom_type[pr.j]];
    const double *A = G.atom_pos + 3 * pr.i;
    const double Bp[3] = {ir.bx, ir.by, ir.bz};
    const double ad00 = G.adist[(((size_t)e1 * G.nel + e2) * G.maxa) * G.maxa];
    const double fac = 1.0 - ir.sep / ad00;
    double mid[3], seg[3], half = 0;
    for (int x = 0; x < 3; ++x) {
        mid[x] = 0.5 * (A[x] + Bp[x]);
        seg[x] = A[x] - Bp[x];
        half += seg[x] * seg[x];
    }
    const double ss = half;
    half = 0.5 * sqrt(half);
    const int pel = G.pot_el[G.site_type[s2]];
    double bound = G.bound0[((size_t)pel * G.nel + e1) * G.nel + e2];
    bound = bound * (fac > 0.0 ? fac : 0.0) * (1.0 + 1e-9) + 1e-12;
    const double rb = sqrt(bound);
    const double *C = G.site_pos + 3 * s2;
    const double mv[3] = {mid[0] - C[0], mid[1] - C[1], mid[2] - C[2]};
    double L2[3];
    find_lattice_vector(G.rv, G.inv, G.cell_dims, G.cell_sizes, G.ncells, mv, L2);
    const double cm[3] = {mv[0] - L2[0], mv[1] - L2[1], mv[2] - L2[2]};
    const double rreq = sqrt(cm[0] * cm[0] + cm[1] * cm[1] + cm[2] * cm[2]) + half + rb;
    if (rreq * rreq > G.cell_sizes[G.ncells - 1]) {
        if (!FILL) {
            atomicOr(err, 1);
            cnt[t] = 0;
        }
        return;
    }
    int c = 0;
    const long base = FILL ? off[t] : 0;
    for (int n = 0; n < G.ncells; ++n) {
        if (G.cell_sizes[n] > rreq * rreq) break;
        const double *Rn = G.cell_dims + 3 * n;
        double P[3], w[3];
        for (int x = 0; x < 3; ++x) {
            P[x] = C[x] + L2[x] + Rn[x];
            w[x] = P[x] - Bp[x];
        }
        // squared distance from P to the segment [B', A]
        double tpar = ss > 0 ? (w[0] * seg[0] + w[1] * seg[1] + w[2] * seg[2]) / ss : 0.0;
        tpar = tpar < 0.0 ? 0.0 : (tpar > 1.0 ? 1.0 : tpar);
        double dd = 0;
        for (int x = 0; x < 3; ++x) {
            const double q2 = w[x] - tpar * seg[x];
            dd += q2 * q2;
        }
        if (dd > bound) continue;
        if (FILL) {
            CandRec cr;
            cr.lx = L2[0];
            cr.ly = L2[1];
            cr.lz = L2[2];
            cr.px = P[0];
            cr.py = P[1];
            cr.pz = P[2];
            cr.site = s2;
            cr.cell = n;
            out[base + c] = cr;
        }
        ++c;
    }
    if (!FILL) cnt[t] = c;
}

}  // namespace imago
