// populate.hpp - populateStandard (src/imago/populate.F90:190-485) and the
// order of its merge sort (src/imago/sortSubs.f90:18-177), host code.
//
// Spin 1 or 2 with no excited core state and no thermal smearing (the
// configuration makeinput writes by default, THERMAL_SMEARING_SIGMA 0):
// states are filled in ascending energy with kPointWeight / spin electrons
// each until numElectrons are placed; the states within smallThresh of the
// highest occupied one share the charge that was placed on them in
// proportion to their weights; the band gap and its kind come from the
// next state up.
#pragma once

#include <cmath>
#include <vector>

namespace imago {

// Index order of mergeSort with one element per initial segment: adjacent
// runs merged bottom-up, on ties the element of the second run first (the
// comparison is a strict '<', sortSubs.f90:130).
inline std::vector<int> merge_sort_order(const std::vector<double> &x)
{
    const int n = (int)x.size();
    std::vector<int> a(n), b(n);
    for (int i = 0; i < n; ++i) a[i] = i;
    std::vector<int> borders(n + 1);
    for (int i = 0; i <= n; ++i) borders[i] = i;
    int nseg = n;
    while (nseg > 1) {
        int to = 0;
        std::vector<int> nb;
        nb.push_back(0);
        for (int s = 0; s + 1 < nseg; s += 2) {
            int p1 = borders[s], e1 = borders[s + 1], p2 = borders[s + 1], e2 = borders[s + 2];
            while (p1 < e1 && p2 < e2) b[to++] = x[a[p1]] < x[a[p2]] ? a[p1++] : a[p2++];
            while (p1 < e1) b[to++] = a[p1++];
            while (p2 < e2) b[to++] = a[p2++];
            nb.push_back(e2);
        }
        if (nseg % 2) {
            for (int i = borders[nseg - 1]; i < borders[nseg]; ++i) b[to++] = a[i];
            nb.push_back(borders[nseg]);
        }
        borders = nb;
        nseg = (int)borders.size() - 1;
        a.swap(b);
    }
    return a;
}

struct PopulateResult {
    double occupied_energy = 0.0;   // highest occupied level (Ha)
    double band_gap = 0.0;          // 0 for a metal (gap <= 1e-3 Ha)
    int gap_kind = 0;               // 0 none, 1 direct, 2 indirect
    int occupied_index = 0;         // 1-based in the sorted list
};

// eigenvalues [nk][spin][ns] (the reference's energyEigenValues(k, i, j)
// visited k-point, then spin, then state); pop gets the same layout.
// Returns -1 when the states cannot hold numElectrons.
inline int populate_standard(int nk, int spin, int ns, const double *eigenvalues,
                             const double *kweights, double num_electrons, double *pop,
                             PopulateResult *res)
{
    const double small = 1.0e-8;           // smallThresh, constants.f90:49
    const double metal_gap = 1.0e-3;       // populate.F90:223
    const int N = nk * spin * ns;
    std::vector<double> e(eigenvalues, eigenvalues + N);
    const std::vector<int> idx = merge_sort_order(e);
    auto w = [&](int q) { return kweights[q / (ns * spin)] / (double)spin; };
    for (int q = 0; q < N; ++q) pop[q] = 0.0;
    double placed = 0.0;
    int occ = -1;
    for (int i = 0; i < N; ++i) {
        const int q = idx[i];
        pop[q] = w(q);
        placed += pop[q];
        if (placed > num_electrons) {
            pop[q] = w(q) + num_electrons - placed;
            placed = num_electrons;
        }
        if (std::fabs(num_electrons - placed) < small) {
            occ = i;
            break;
        }
    }
    if (occ < 0) return -1;
    const double ef = e[idx[occ]];
    int lo = occ, hi = occ;
    double degen = pop[idx[occ]], possible = w(idx[occ]);
    while (hi + 1 < N && std::fabs(ef - e[idx[hi + 1]]) <= small) {
        ++hi;
        possible += w(idx[hi]);
    }
    while (lo > 0 && std::fabs(ef - e[idx[lo - 1]]) <= small) {
        --lo;
        possible += w(idx[lo]);
        degen += w(idx[lo]);
    }
    for (int i = lo; i <= hi; ++i) pop[idx[i]] = degen / possible * w(idx[i]);
    res->occupied_index = hi + 1;
    res->occupied_energy = e[idx[hi]];
    if (hi + 1 >= N) {
        res->band_gap = 0.0;
        res->gap_kind = 0;
    } else {
        const double gap = e[idx[hi + 1]] - e[idx[hi]];
        if (gap <= metal_gap) {
            res->band_gap = 0.0;
            res->gap_kind = 0;
        } else {
            res->band_gap = gap;
            res->gap_kind = idx[hi] / (ns * spin) == idx[hi + 1] / (ns * spin) ? 1 : 2;
        }
    }
    return 0;
}

}  // namespace imago
