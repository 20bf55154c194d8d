// Host build of the device primitive evaluators (imago_amd/csrc/prim.hpp)
// checked against the oracle's restatement (oracle/liboracle.so) for all
// four kinds and all 16 (l_a, l_b) classes on seeded random tuples.
// Prints the worst normwise relative error per kind; exits 1 above 1e-12
// (1e-9 for the nuclear term, whose Boys closed form cancels near T = 0.1).
#include "../../imago_amd/csrc/prim.hpp"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <random>

using namespace imago;

struct HostAcc {
    double *o;
    void add(int i, int j, double x) { o[i + 16 * j] += x; }
};
typedef void (*prim2)(double, double, const double *, const double *, int, double *);
typedef void (*prim3)(double, double, double, const double *, const double *, const double *,
                      int, double *);

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    void *h = dlopen(argv[1], RTLD_NOW);
    if (!h) return 2;
    prim2 ol = (prim2)dlsym(h, "orc_overlap2c"), ke = (prim2)dlsym(h, "orc_kinetic2c");
    prim3 ep = (prim3)dlsym(h, "orc_electron3c"), np = (prim3)dlsym(h, "orc_nuclear3c");
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> lg(std::log(0.1), std::log(5e3)), u(-1.5, 1.5);
    double worst[5] = {0, 0, 0, 0, 0};
    for (int trial = 0; trial < 40; ++trial)
        for (int la = 0; la < 4; ++la)
            for (int lb = 0; lb < 4; ++lb)
                for (int kind = 1; kind <= 4; ++kind) {
                    const double a1 = std::exp(lg(rng)), a2 = std::exp(lg(rng)), a3 = std::exp(lg(rng));
                    double A[3], B[3], C[3];
                    for (int x = 0; x < 3; ++x) {
                        A[x] = u(rng);
                        B[x] = A[x] + 0.6 * u(rng);
                        C[x] = 0.5 * (A[x] + B[x]) + 0.4 * u(rng);
                    }
                    // keep the tuple inside the region the assembly evaluates:
                    // xi |A-B|^2 <= L and (zeta a3/zeta3)|P-C|^2 <= L at the
                    // reference's default threshold L = -ln(1e-16)
                    const double L = 36.84, xi = a1 * a2 / (a1 + a2);
                    double d2 = 0;
                    for (int x = 0; x < 3; ++x) d2 += (B[x] - A[x]) * (B[x] - A[x]);
                    if (xi * d2 > L) {
                        const double f = std::sqrt(L / (xi * d2));
                        for (int x = 0; x < 3; ++x) B[x] = A[x] + f * (B[x] - A[x]);
                    }
                    {
                        const double z = a1 + a2, w = z * a3 / (z + a3);
                        double P[3], c2 = 0;
                        for (int x = 0; x < 3; ++x) {
                            P[x] = (a1 * A[x] + a2 * B[x]) / z;
                            c2 += (C[x] - P[x]) * (C[x] - P[x]);
                        }
                        if (w * c2 > L) {
                            const double f = std::sqrt(L / (w * c2));
                            for (int x = 0; x < 3; ++x) C[x] = P[x] + f * (C[x] - P[x]);
                        }
                    }
                    const int sw = (1 << la) + (16 << lb);
                    double ref[256] = {0}, got[256] = {0};
                    HostAcc acc{got};
                    const Vec3 a{A[0], A[1], A[2]}, b{B[0], B[1], B[2]}, c{C[0], C[1], C[2]};
                    switch (kind) {
                    case 1: ol(a1, a2, A, B, sw, ref); prim_dispatch<K_OL, 3>(la, lb, a1, a2, 0, a, b, c, 1.0, acc); break;
                    case 2: ke(a1, a2, A, B, sw, ref); prim_dispatch<K_KE, 3>(la, lb, a1, a2, 0, a, b, c, 1.0, acc); break;
                    case 3: np(a1, a2, a3, A, B, C, sw, ref); prim_dispatch<K_NP, 3>(la, lb, a1, a2, a3, a, b, c, 1.0, acc); break;
                    default: ep(a1, a2, a3, A, B, C, sw, ref); prim_dispatch<K_EP, 3>(la, lb, a1, a2, a3, a, b, c, 1.0, acc); break;
                    }
                    double md = 0, mx = 0;
                    for (int q = 0; q < 256; ++q) {
                        md = std::fmax(md, std::fabs(got[q] - ref[q]));
                        mx = std::fmax(mx, std::fabs(ref[q]));
                    }
                    if (mx > 1e-250) worst[kind] = std::fmax(worst[kind], md / mx);
                }
    int bad = 0;
    const char *names[5] = {"", "ol", "ke", "np", "ep"};
    for (int k = 1; k <= 4; ++k) {
        printf("%s %.3e\n", names[k], worst[k]);
        if (worst[k] > (k == 3 ? 1e-9 : 1e-12)) bad = 1;
    }
    return bad;
}
