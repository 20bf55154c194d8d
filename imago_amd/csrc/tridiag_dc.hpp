// tridiag_dc.hpp - divide and conquer for batches of real symmetric
// tridiagonal eigenproblems T = Z diag(w) Z^T, all eigenvectors.
//
// This is the tridiagonal stage of the eigensolve behind solveZHEGV /
// solveDSYGV (src/imago/interfaces.F90:73-200): after the reduction of the
// transformed Hamiltonian to a real tridiagonal T, every eigenpair of T is
// needed.  The method is Cuppen's divide and conquer with the deflation of
// Dongarra-Sorensen and the eigenvector formula of Gu-Eisenstat, the
// numerical scheme of LAPACK dstedc / dlaed0-dlaed4:
//
//   * T is torn at every off-diagonal: each leaf is one diagonal entry
//     d'_i = d_i - |e_{i-1}| - |e_i| with eigenvector [1]; a merge of two
//     children L (n1) and R (n2) undoes the tear at s = r0 + n1:
//     T = diag(T_L', T_R') + rho v v^T, v = e_{s-1} + sgn(e_{s-1}) e_s,
//     rho = |e_{s-1}|.
//   * In the children's eigenbases the merged problem is D + rho z z^T with
//     z = [last row of Q_L; sgn(e_{s-1}) first row of Q_R], normalised to
//     |z| = 1 (rho doubled).  D is sorted; entries with rho |z_j| <= tol
//     deflate, and neighbours closer than tol after a Givens rotation in
//     the (z_p, z_j) plane deflate the first of the pair.
//   * The K remaining eigenvalues are the roots of the secular equation
//     1/rho + sum_j z_j^2 / (d_j - lambda) = 0, one in each interval
//     (d_i, d_{i+1}) and the last in (d_K, d_K + rho |z|^2].  Each root is
//     kept as an origin pole plus an offset, so d_j - lambda is formed from
//     exact pole differences; the iteration fits the two nearest poles
//     (value and derivative) and falls back to bisection of a bracket.
//   * z is recomputed from the roots (Loewner's formula), which makes the
//     eigenvectors u_j = z_j / (d_j - lambda) numerically orthogonal.
//   * The merged eigenvectors are blockdiag(Q_L, Q_R) P G U, U holding the
//     rank-one problem's vectors and unit vectors for the deflated entries,
//     G the deflating rotations and P the sort: two GEMMs per merge.
//
// The functions below are the per-merge and per-root numerics.  They
// compile for the host as well (a host build checks them against LAPACK
// without a GPU) and are parameterised by a lane policy: `Seq` runs a sum
// as one loop, the device's wave policy strides it over 64 lanes and
// broadcasts lane 0's reduction so every lane takes the same branch.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <vector>

namespace imago {
namespace dc {

#define DC_HD __host__ __device__ __forceinline__

constexpr double kEps = 1.1102230246251565e-16;   // 2^-53, unit roundoff
constexpr int kMaxSecularIter = 200;

// one merge of the tree: rows [r0, r0 + n1 + n2) of a matrix, children of
// heights hL and hR, the merge itself of height h (leaves have height 0)
struct Merge {
    int r0, n1, n2, hL, hR, h;
};

// per-merge results of the deflation, kept at the merge's row offset r0
// in arrays of n entries per matrix
struct MergeMeta {
    int K;       // eigenvalues left to the secular equation
    int nrot;    // deflating rotations
    double rho;  // 2 |e_{s-1}| (z normalised to unit length)
};

struct Seq {
    DC_HD int lane() const { return 0; }
    DC_HD int width() const { return 1; }
    DC_HD double sum(double x) const { return x; }
    DC_HD double prod(double x) const { return x; }
};

// The deflation scan of one merge over its m entries in sorted order: Ds
// (the children's eigenvalues, ascending) and zs (z in the same order) are
// updated in place by the rotations.  Writes the non-deflated entries
// (dl ascending, zz, their sorted positions kpos), the deflated ones
// (dpos, dval) and the rotations (rp, rq, rc, rs: columns p, q of the
// sorted basis, x_p' = c x_p + s x_q, x_q' = c x_q - s x_p).  Returns K.
// The neighbour test of dlaed2 between the pending entry (zp, Dp) and the
// next one (zt, Dt): the Givens rotation that zeroes zp (c, s, and the
// combined z, tau) and whether it deflates, |(Dt - Dp) c s| <= tol.
DC_HD bool pair_deflates(double zp, double Dp, double zt, double Dt, double tol, double *c_out,
                         double *s_out, double *tau_out)
{
    double s = zp, c = zt;
    const double tau = sqrt(c * c + s * s);
    const double gap = Dt - Dp;
    c /= tau;
    s = -s / tau;
    *c_out = c;
    *s_out = s;
    *tau_out = tau;
    return fabs(gap * c * s) <= tol;
}

DC_HD int deflate_scan(int m, double rho, double tol, double *Ds, double *zs, double *dl,
                       double *zz, int *kpos, int *dpos, double *dval, int *rp, int *rq,
                       double *rc, double *rs, int *nrot_out)
{
    int K = 0, nd = 0, nrot = 0, pj = -1;
    for (int t = 0; t < m; ++t) {
        if (rho * fabs(zs[t]) <= tol) {
            dpos[nd] = t;
            dval[nd] = Ds[t];
            ++nd;
            continue;
        }
        if (pj < 0) {
            pj = t;
            continue;
        }
        double c, s, tau;
        if (pair_deflates(zs[pj], Ds[pj], zs[t], Ds[t], tol, &c, &s, &tau)) {
            zs[t] = tau;
            zs[pj] = 0.0;
            rp[nrot] = pj;
            rq[nrot] = t;
            rc[nrot] = c;
            rs[nrot] = s;
            ++nrot;
            const double dp = Ds[pj] * c * c + Ds[t] * s * s;
            Ds[t] = Ds[pj] * s * s + Ds[t] * c * c;
            Ds[pj] = dp;
            dpos[nd] = pj;
            dval[nd] = dp;
            ++nd;
        } else {
            dl[K] = Ds[pj];
            zz[K] = zs[pj];
            kpos[K] = pj;
            ++K;
        }
        pj = t;
    }
    if (pj >= 0) {
        dl[K] = Ds[pj];
        zz[K] = zs[pj];
        kpos[K] = pj;
        ++K;
    }
    *nrot_out = nrot;
    return K;
}

// Root i (0-based) of 1/rho + sum_j zz_j^2 / (dl_j - lambda) = 0 over K
// strictly ascending poles dl, rho > 0.  lambda = dl[origin] + tau.
template <class L>
DC_HD void secular_root(const L &ln, int i, int K, const double *dl, const double *zz,
                        double rho, int *origin, double *tau_out)
{
    const double rinv = 1.0 / rho;
    int o;
    double lo, hi, tau;
    if (i < K - 1) {
        const double gap = dl[i + 1] - dl[i];
        const double mid = 0.5 * gap;
        double f = 0.0;
        for (int j = ln.lane(); j < K; j += ln.width()) f += zz[j] * zz[j] / ((dl[j] - dl[i]) - mid);
        f = ln.sum(f) + rinv;
        if (f >= 0.0) {   // the root lies in (dl_i, dl_i + mid]
            o = i;
            lo = 0.0;
            hi = mid;
            tau = 0.5 * mid;
        } else {          // in (dl_i + mid, dl_{i+1})
            o = i + 1;
            lo = (dl[i] - dl[i + 1]) + mid;
            hi = 0.0;
            tau = 0.5 * lo;
        }
    } else {
        double s = 0.0;
        for (int j = ln.lane(); j < K; j += ln.width()) s += zz[j] * zz[j];
        s = ln.sum(s);
        o = K - 1;
        lo = 0.0;
        hi = rho * s * (1.0 + 8.0 * kEps);
        tau = 0.5 * hi;
    }
    const double dlo = dl[o];
    for (int it = 0; it < kMaxSecularIter; ++it) {
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
        for (int j = ln.lane(); j < K; j += ln.width()) {
            const double del = (dl[j] - dlo) - tau;
            const double t = zz[j] / del;
            if (j <= i) {
                psi += zz[j] * t;
                dpsi += t * t;
            } else {
                phi += zz[j] * t;
                dphi += t * t;
            }
        }
        psi = ln.sum(psi);
        dpsi = ln.sum(dpsi);
        phi = ln.sum(phi);
        dphi = ln.sum(dphi);
        const double w = rinv + psi + phi;
        const double dw = dpsi + dphi;
        const double erretm = 8.0 * (phi - psi) + 2.0 * rinv + 3.0 * fabs(tau) * dw;
        if (fabs(w) <= kEps * erretm) break;
        if (w < 0.0)
            lo = tau;
        else
            hi = tau;
        // two-pole model c + s/(dlt - eta) + S/(drt - eta) through (w, w')
        const double dlt = (dl[i] - dlo) - tau;
        double eta;
        if (i < K - 1) {
            const double drt = (dl[i + 1] - dlo) - tau;
            const double s = dlt * dlt * dpsi, S = drt * drt * dphi;
            const double c = w - dlt * dpsi - drt * dphi;
            const double a = c * (dlt + drt) + s + S;
            const double b = c * dlt * drt + s * drt + S * dlt;
            if (c == 0.0) {
                eta = b / a;
            } else {
                const double sq = sqrt(fmax(a * a - 4.0 * b * c, 0.0));
                eta = a >= 0.0 ? 2.0 * b / (a + sq) : (a - sq) / (2.0 * c);
            }
        } else {
            const double s = dlt * dlt * dpsi;
            const double c = w - dlt * dpsi;
            eta = c > 0.0 ? dlt + s / c : 0.5 * (lo + hi) - tau;
        }
        double tn = tau + eta;
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == tau) break;
        tau = tn;
    }
    *origin = o;
    *tau_out = tau;
}

// z_k recomputed from the roots (Loewner): z_k^2 = prod_i (lambda_i - dl_k)
// / (rho prod_{i != k} (dl_i - dl_k)), with lambda_i - dl_k formed from the
// root's origin and offset; the sign of the original zz_k.
template <class L>
DC_HD double loewner_z(const L &ln, int k, int K, const double *dl, const double *zz,
                       const int *origin, const double *tau, double rho)
{
    double p = 1.0;
    for (int i = ln.lane(); i < K; i += ln.width()) {
        const double lam_minus_dk = tau[i] - (dl[k] - dl[origin[i]]);
        p *= i == k ? lam_minus_dk / rho : lam_minus_dk / (dl[i] - dl[k]);
    }
    p = ln.prod(p);
    return copysign(sqrt(fabs(p)), zz[k]);
}

// the rotations of the deflation applied to a column of the sorted basis,
// last to first: v = G_1 ... G_r v
DC_HD void apply_rotations(double *v, int nrot, const int *rp, const int *rq, const double *rc,
                           const double *rs)
{
    for (int r = nrot - 1; r >= 0; --r) {
        const int p = rp[r], q = rq[r];
        const double vp = v[p], vq = v[q];
        if (vp == 0.0 && vq == 0.0) continue;
        v[p] = rc[r] * vp - rs[r] * vq;
        v[q] = rs[r] * vp + rc[r] * vq;
    }
}

// The merge tree of an order-n problem: halves (floor, ceil) down to single
// rows.  Returns the merges ordered by height (all of height 1 first), and
// the root's height.
inline int build_tree(int n, std::vector<Merge> &merges)
{
    merges.clear();
    struct Node {
        int r0, m;
    };
    // post-order recursion with an explicit stack: heights of children first
    std::vector<std::pair<Node, int>> stack;   // (node, state)
    std::vector<int> hstack;
    stack.push_back({{0, n}, 0});
    while (!stack.empty()) {
        auto &top = stack.back();
        const Node nd = top.first;
        if (nd.m == 1) {
            hstack.push_back(0);
            stack.pop_back();
            continue;
        }
        const int n1 = nd.m / 2, n2 = nd.m - n1;
        if (top.second == 0) {
            top.second = 1;
            stack.push_back({{nd.r0 + n1, n2}, 0});
            stack.push_back({{nd.r0, n1}, 0});
            continue;
        }
        // both children done: left's height pushed first, right's on top
        const int hR = hstack.back();
        hstack.pop_back();
        const int hL = hstack.back();
        hstack.pop_back();
        const int h = 1 + std::max(hL, hR);
        merges.push_back({nd.r0, n1, n2, hL, hR, h});
        hstack.push_back(h);
        stack.pop_back();
    }
    std::stable_sort(merges.begin(), merges.end(),
                     [](const Merge &a, const Merge &b) { return a.h < b.h; });
    return hstack.empty() ? 0 : hstack.back();
}

}   // namespace dc
}   // namespace imago
