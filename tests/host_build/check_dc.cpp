// Host build of the tridiagonal divide and conquer's numerics
// (imago_amd/csrc/tridiag_dc.hpp): the same tree, per-merge deflation,
// secular roots, Loewner vectors and rotations the device kernels run,
// driven merge by merge with plain loops for the two GEMMs of each merge.
//
// Usage: check_dc in.bin out.bin.  in.bin holds int32 count, then per
// problem int32 n, n doubles d, n-1 doubles e; out.bin receives per
// problem n eigenvalues (ascending) and the n x n eigenvectors (column-
// major).
#include "../../imago_amd/csrc/tridiag_dc.hpp"

#include <cmath>
#include <cstdio>
#include <vector>

using namespace imago::dc;

static void solve_host(int n, const double *d, const double *e, double *w_out, double *Z_out)
{
    double scale = 0.0;
    for (int i = 0; i < n; ++i) scale = std::max(scale, std::fabs(d[i]));
    for (int i = 0; i + 1 < n; ++i) scale = std::max(scale, std::fabs(e[i]));
    if (scale == 0.0) scale = 1.0;
    std::vector<double> es(std::max(n - 1, 1));
    for (int i = 0; i + 1 < n; ++i) es[i] = e[i] / scale;
    std::vector<Merge> merges;
    const int H = build_tree(n, merges);
    const size_t nn = (size_t)n * n;
    std::vector<double> Zb[3], wb[3];
    for (int q = 0; q < 3; ++q) {
        Zb[q].assign(nn, 0.0);
        wb[q].assign(n, 0.0);
    }
    for (int i = 0; i < n; ++i) {
        wb[0][i] = d[i] / scale - (i > 0 ? std::fabs(es[i - 1]) : 0.0) -
                   (i + 1 < n ? std::fabs(es[i]) : 0.0);
        Zb[0][(size_t)i * n + i] = 1.0;
    }
    std::vector<double> Vu(nn, 0.0);
    std::vector<double> D(n), z(n), Ds(n), zs(n), dl(n), zz(n), dval(n), rc(n), rs(n), tau(n),
        val(n), zh(n), col(n);
    std::vector<int> perm(n), kpos(n), dpos(n), rp(n), rq(n), org(n), rank(n);
    for (const Merge &mg : merges) {
        const int r0 = mg.r0, n1 = mg.n1, n2 = mg.n2, m = n1 + n2, s = r0 + n1;
        const double beta = es[s - 1], rho = 2.0 * std::fabs(beta);
        const double sigma = beta >= 0.0 ? 1.0 : -1.0;
        const std::vector<double> &ZL = Zb[mg.hL % 3], &ZR = Zb[mg.hR % 3];
        const std::vector<double> &wL = wb[mg.hL % 3], &wR = wb[mg.hR % 3];
        for (int x = 0; x < m; ++x) {
            const bool left = x < n1;
            D[x] = left ? wL[r0 + x] : wR[r0 + x];
            z[x] = (left ? ZL[(size_t)(r0 + x) * n + s - 1] : sigma * ZR[(size_t)(r0 + x) * n + s]) *
                   std::sqrt(0.5);
        }
        double dmax = 0.0, zmax = 0.0;
        for (int x = 0; x < m; ++x) {
            int pos;
            if (x < n1) {
                int c = 0;
                for (int j = n1; j < m; ++j) c += D[j] < D[x];
                pos = x + c;
            } else {
                int c = 0;
                for (int j = 0; j < n1; ++j) c += D[j] <= D[x];
                pos = (x - n1) + c;
            }
            Ds[pos] = D[x];
            zs[pos] = z[x];
            perm[pos] = x;
            dmax = std::max(dmax, std::fabs(D[x]));
            zmax = std::max(zmax, std::fabs(z[x]));
        }
        const double tol = 8.0 * kEps * std::max(dmax, zmax);
        int nrot = 0;
        const int K = deflate_scan(m, rho, tol, Ds.data(), zs.data(), dl.data(), zz.data(),
                                   kpos.data(), dpos.data(), dval.data(), rp.data(), rq.data(),
                                   rc.data(), rs.data(), &nrot);
        Seq ln;
        for (int x = 0; x < K; ++x) {
            secular_root(ln, x, K, dl.data(), zz.data(), rho, &org[x], &tau[x]);
            val[x] = dl[org[x]] + tau[x];
        }
        for (int q = 0; q < m - K; ++q) val[K + q] = dval[q];
        for (int k = 0; k < K; ++k)
            zh[k] = loewner_z(ln, k, K, dl.data(), zz.data(), org.data(), tau.data(), rho);
        for (int x = 0; x < m; ++x) {
            int c = 0;
            for (int y = 0; y < m; ++y) c += val[y] < val[x] || (val[y] == val[x] && y < x);
            rank[x] = c;
        }
        for (int x = 0; x < m; ++x) {
            std::fill(col.begin(), col.begin() + m, 0.0);
            if (x < K) {
                double nrm = 0.0;
                for (int k = 0; k < K; ++k) {
                    const double u = zh[k] / ((dl[k] - dl[org[x]]) - tau[x]);
                    col[kpos[k]] = u;
                    nrm += u * u;
                }
                nrm = 1.0 / std::sqrt(nrm);
                for (int k = 0; k < K; ++k) col[kpos[k]] *= nrm;
            } else {
                col[dpos[x - K]] = 1.0;
            }
            apply_rotations(col.data(), nrot, rp.data(), rq.data(), rc.data(), rs.data());
            for (int t = 0; t < m; ++t) Vu[(size_t)(r0 + rank[x]) * n + r0 + perm[t]] = col[t];
        }
        std::vector<double> &Zo = Zb[mg.h % 3];
        for (int c = 0; c < m; ++c) {
            const double *v = &Vu[(size_t)(r0 + c) * n + r0];
            for (int i = 0; i < n1; ++i) {
                double acc = 0.0;
                for (int j = 0; j < n1; ++j) acc += ZL[(size_t)(r0 + j) * n + r0 + i] * v[j];
                Zo[(size_t)(r0 + c) * n + r0 + i] = acc;
            }
            for (int i = 0; i < n2; ++i) {
                double acc = 0.0;
                for (int j = 0; j < n2; ++j)
                    acc += ZR[(size_t)(s + j) * n + s + i] * v[n1 + j];
                Zo[(size_t)(r0 + c) * n + s + i] = acc;
            }
        }
        for (int x = 0; x < m; ++x) wb[mg.h % 3][r0 + rank[x]] = val[x];
    }
    for (int i = 0; i < n; ++i) w_out[i] = wb[H % 3][i] * scale;
    std::copy(Zb[H % 3].begin(), Zb[H % 3].end(), Z_out);
}

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    FILE *fi = std::fopen(argv[1], "rb"), *fo = std::fopen(argv[2], "wb");
    if (!fi || !fo) return 2;
    int count = 0;
    if (std::fread(&count, 4, 1, fi) != 1) return 2;
    for (int p = 0; p < count; ++p) {
        int n = 0;
        if (std::fread(&n, 4, 1, fi) != 1 || n < 1) return 2;
        std::vector<double> d(n), e(std::max(n - 1, 1)), w(n), Z((size_t)n * n);
        if (std::fread(d.data(), 8, n, fi) != (size_t)n) return 2;
        if (n > 1 && std::fread(e.data(), 8, n - 1, fi) != (size_t)(n - 1)) return 2;
        solve_host(n, d.data(), e.data(), w.data(), Z.data());
        std::fwrite(w.data(), 8, n, fo);
        std::fwrite(Z.data(), 8, (size_t)n * n, fo);
    }
    std::fclose(fo);
    std::fclose(fi);
    return 0;
}
