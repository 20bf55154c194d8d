// prim.hpp - primitive Gaussian integrals on CDNA4, FP64 VALU.
//
// Device restatement of the primitive seam of the reference
// (src/imago/gaussIntegrals.f90): overlap2CIntg (:26), kinetic2CIntg
// (:4422), electron3CIntg (:11624), nuclear3CIntg (:16223) and boys
// (:16080).  The reference evaluates generated straight-line Obara-Saika
// code per (l_a,l_b) class; here every class is a template instance of
//   * a 1-D Obara-Saika table per Cartesian direction for the overlap-type
//     integrals (2-centre overlap, kinetic via the second derivative of
//     the ket, 3-centre overlap about the three-Gaussian centre G), and
//   * McMurchie-Davidson Hermite expansion about G with the Hermite
//     Coulomb integrals R_tuv for the Gaussian-damped nuclear term.
// Outputs are the 16 unnormalised real solid-harmonic combinations of
// sh(16,16) (gaussIntegrals.f90:33-35), delivered entry by entry to an
// accumulator so callers can sum over lattice images in registers.
#pragma once
#include <hip/hip_runtime.h>

// Primitive evaluators compile for the host as well, so a CPU build can be
// checked against the oracle without a GPU.
#define IMAGO_HD __host__ __device__ __forceinline__

namespace imago {

constexpr double kPi = 3.141592653589793;

// Cartesian monomials up to l = 3: exponents (x,y,z)
__host__ __device__ constexpr int cart_e(int c, int d)
{
    constexpr int E[20][3] = {
        {0, 0, 0},
        {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
        {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1},
        {1, 1, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1}, {1, 0, 2},
        {0, 1, 2}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}};
    return E[c][d];
}

// sh combinations: up to 3 (monomial, weight) terms per component
__host__ __device__ constexpr int sh_n(int lm)
{
    constexpr int N[16] = {1, 1, 1, 1, 1, 1, 1, 2, 3, 1, 2, 2, 2, 3, 3, 3};
    return N[lm];
}
__host__ __device__ constexpr int sh_c(int lm, int t)
{
    constexpr int C[16][3] = {
        {0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {3, 0, 0},
        {7, 0, 0}, {8, 0, 0}, {9, 0, 0}, {4, 5, 0}, {6, 4, 5},
        {10, 0, 0}, {12, 14, 0}, {17, 13, 0}, {11, 18, 0},
        {19, 12, 14}, {15, 17, 13}, {16, 11, 18}};
    return C[lm][t];
}
__host__ __device__ constexpr double sh_w(int lm, int t)
{
    constexpr double W[16][3] = {
        {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0},
        {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, -1, 0}, {2, -1, -1},
        {1, 0, 0}, {1, -1, 0}, {1, -3, 0}, {3, -1, 0},
        {2, -3, -3}, {4, -1, -1}, {4, -1, -1}};
    return W[lm][t];
}

// ---------------------------------------------------------------------------
// Boys function F_0..F_M(T), formula of gaussIntegrals.f90:16080-16221:
// below T <= 10E-2 (a single-precision literal) the 7-term series with
// single-precision Gamma(n+1/2) constants, above it the closed forms in
// erf(sqrt T) and exp(-T).
IMAGO_HD double boys_gamma_half(int k)
{
    // Gamma(k + 1/2), k = 0..13, rounded to single precision as the
    // reference's default-real literals are.
    switch (k) {
    case 0: return (double)1.7724538509055161f;
    case 1: return (double)0.88622692545275805f;
    case 2: return (double)1.3293403881791370f;
    case 3: return (double)3.3233509704478426f;
    case 4: return (double)11.631728396567450f;
    case 5: return (double)52.342777784553533f;
    case 6: return (double)287.88527781504416f;
    case 7: return (double)1871.2543057977896f;
    case 8: return (double)14034.407293483402f;
    case 9: return (double)119292.46199460901f;
    case 10: return (double)1133278.3889487833f;
    case 11: return (double)11899423.083962219f;
    case 12: return (double)136843365.46556622f;
    default: return (double)1710542068.3195722f;
    }
}

template <int M>
IMAGO_HD void boys_ref(double T, double (&F)[M + 1])
{
#pragma clang fp contract(off)
    const double exp_T = exp(-T);
    if (T <= (double)10E-2f) {
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, T5 = T4 * T, T6 = T5 * T;
#pragma unroll
        for (int N = 0; N <= M; ++N) {
            F[N] = 0.5 * exp_T * boys_gamma_half(N) *
                   ((1 / boys_gamma_half(N + 1)) + (T / boys_gamma_half(N + 2)) +
                    (T2 / boys_gamma_half(N + 3)) + (T3 / boys_gamma_half(N + 4)) +
                    (T4 / boys_gamma_half(N + 5)) + (T5 / boys_gamma_half(N + 6)) +
                    (T6 / boys_gamma_half(N + 7)));
        }
    } else {
        const double sqrt_pi = sqrt(kPi);
        const double erf_T = erf(sqrt(T));
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, T5 = T4 * T, T6 = T5 * T;
        F[0] = 0.5 * ((sqrt_pi / sqrt(T)) * erf_T);
        if (M >= 1)
            F[1 <= M ? 1 : 0] = ((sqrt_pi / (4 * pow(T, 1.5))) * erf_T - exp_T * (1 / (2.0 * T)));
        if (M >= 2)
            F[2 <= M ? 2 : 0] = 6 * ((sqrt_pi / (16 * pow(T, 2.5))) * erf_T -
                                     exp_T * (2 / (24.0 * T) + 1 / ((4) * 2.0 * T2)));
        if (M >= 3)
            F[3 <= M ? 3 : 0] = 60 * ((sqrt_pi / (64 * pow(T, 3.5))) * erf_T -
                                      exp_T * (6 / (720.0 * T) + 2 / ((4) * 24.0 * T2) +
                                               1 / ((16) * 2.0 * T3)));
        if (M >= 4)
            F[4 <= M ? 4 : 0] = 840 * ((sqrt_pi / (256 * pow(T, 4.5))) * erf_T -
                                       exp_T * (24 / (40320.0 * T) + 6 / ((4) * 720.0 * T2) +
                                                2 / ((16) * 24.0 * T3) + 1 / ((64) * 2.0 * T4)));
        if (M >= 5)
            F[5 <= M ? 5 : 0] = 15120 * ((sqrt_pi / (1024 * pow(T, 5.5))) * erf_T -
                                         exp_T * (120 / (3628800.0 * T) +
                                                  24 / ((4) * 40320.0 * T2) +
                                                  6 / ((16) * 720.0 * T3) +
                                                  2 / ((64) * 24.0 * T4) +
                                                  1 / ((256) * 2.0 * T5)));
        if (M >= 6)
            F[6 <= M ? 6 : 0] = 332640 * ((sqrt_pi / (4096 * pow(T, 6.5))) * erf_T -
                                          exp_T * (720 / (479001600.0 * T) +
                                                   120 / ((4) * 3628800.0 * T2) +
                                                   24 / ((16) * 40320.0 * T3) +
                                                   6 / ((64) * 720.0 * T4) +
                                                   2 / ((256) * 24.0 * T5) +
                                                   1 / ((1024) * 2.0 * T6)));
    }
}

// ---------------------------------------------------------------------------
// 1-D Obara-Saika overlap table relative to the s-s value:
//   s(i+1,j) = PA s(i,j) + (i s(i-1,j) + j s(i,j-1)) / (2p)
//   s(i,j+1) = PB s(i,j) + (i s(i-1,j) + j s(i,j-1)) / (2p)
template <int LA, int LB>
IMAGO_HD void os1d(double i2p, double PA, double PB,
                                     double (&s)[LA + 1][LB + 1])
{
    s[0][0] = 1.0;
#pragma unroll
    for (int i = 1; i <= LA; ++i)
        s[i][0] = PA * s[i - 1][0] + (i > 1 ? (i - 1) * i2p * s[i > 1 ? i - 2 : 0][0] : 0.0);
#pragma unroll
    for (int j = 1; j <= LB; ++j) {
#pragma unroll
        for (int i = 0; i <= LA; ++i) {
            double v = PB * s[i][j - 1];
            if (i > 0) v += i * i2p * s[i > 0 ? i - 1 : 0][j - 1];
            if (j > 1) v += (j - 1) * i2p * s[i][j > 1 ? j - 2 : 0];
            s[i][j] = v;
        }
    }
}

// Accumulates sh(lm1,lm2) += scale * pref * sum combos over Cartesian values
// given by a functor cart(ca, cb).
template <int LA, int LB, class Cart, class Acc>
IMAGO_HD void combine_sh(double f, const Cart &cart, Acc &acc)
{
    constexpr int NA = (LA + 1) * (LA + 1), NB = (LB + 1) * (LB + 1);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            double v = 0.0;
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                if (p >= sh_n(i)) continue;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (q >= sh_n(j)) continue;
                    v += (sh_w(i, p) * sh_w(j, q)) * cart(sh_c(i, p), sh_c(j, q));
                }
            }
            acc.add(i, j, f * v);
        }
    }
}

struct Vec3 {
    double x, y, z;
};

// overlap2CIntg: <a|b>, prefactor (pi/zeta)^1.5 exp(-xi |A-B|^2)
template <int LA, int LB, class Acc>
IMAGO_HD void prim_overlap(double a1, double a2, Vec3 A, Vec3 B,
                                             double scale, Acc &acc)
{
    const double zeta = a1 + a2, xi = a1 * a2 / zeta, i2p = 0.5 / zeta;
    const double Px = (a1 * A.x + a2 * B.x) / zeta, Py = (a1 * A.y + a2 * B.y) / zeta,
                 Pz = (a1 * A.z + a2 * B.z) / zeta;
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double pref = (kPi / zeta) * sqrt(kPi / zeta) * exp(-xi * d2);
    double sx[LA + 1][LB + 1], sy[LA + 1][LB + 1], sz[LA + 1][LB + 1];
    os1d<LA, LB>(i2p, Px - A.x, Px - B.x, sx);
    os1d<LA, LB>(i2p, Py - A.y, Py - B.y, sy);
    os1d<LA, LB>(i2p, Pz - A.z, Pz - B.z, sz);
    auto cart = [&](int ca, int cb) {
        return sx[cart_e(ca, 0)][cart_e(cb, 0)] * sy[cart_e(ca, 1)][cart_e(cb, 1)] *
               sz[cart_e(ca, 2)][cart_e(cb, 2)];
    };
    combine_sh<LA, LB>(scale * pref, cart, acc);
}

// kinetic2CIntg: <a| -1/2 nabla^2 |b> as T = tx sy sz + sx ty sz + sx sy tz
// with the 1-D Obara-Saika kinetic recursion per direction (relative to the
// s-s overlap, X = A - B):
//   t(0,0)   = xi - 2 xi^2 X^2
//   t(i+1,j) = PA t(i,j) + (i t(i-1,j) + j t(i,j-1))/(2p)
//              + 2 xi (s(i+1,j) - i/(2 a1) s(i-1,j))
//   t(i,j+1) = PB t(i,j) + (i t(i-1,j) + j t(i,j-1))/(2p)
//              + 2 xi (s(i,j+1) - j/(2 a2) s(i,j-1))
// (stable for very different exponents, unlike derivative formulas)
template <int LA, int LB>
IMAGO_HD void os1d_kinetic(double i2p, double xi, double a1, double a2, double X, double PA,
                           double PB, const double (&s)[LA + 1][LB + 1],
                           double (&t)[LA + 1][LB + 1])
{
    t[0][0] = xi - 2.0 * xi * xi * X * X;
#pragma unroll
    for (int i = 1; i <= LA; ++i) {
        double v = PA * t[i - 1][0] + 2.0 * xi * s[i][0];
        if (i > 1) {
            v += (i - 1) * i2p * t[i > 1 ? i - 2 : 0][0];
            v -= 2.0 * xi * ((i - 1) / (2.0 * a1)) * s[i > 1 ? i - 2 : 0][0];
        }
        t[i][0] = v;
    }
#pragma unroll
    for (int j = 1; j <= LB; ++j)
#pragma unroll
        for (int i = 0; i <= LA; ++i) {
            double v = PB * t[i][j - 1] + 2.0 * xi * s[i][j];
            if (i > 0) v += i * i2p * t[i > 0 ? i - 1 : 0][j - 1];
            if (j > 1) {
                v += (j - 1) * i2p * t[i][j > 1 ? j - 2 : 0];
                v -= 2.0 * xi * ((j - 1) / (2.0 * a2)) * s[i][j > 1 ? j - 2 : 0];
            }
            t[i][j] = v;
        }
}

template <int LA, int LB, class Acc>
IMAGO_HD void prim_kinetic(double a1, double a2, Vec3 A, Vec3 B, double scale, Acc &acc)
{
    const double zeta = a1 + a2, xi = a1 * a2 / zeta, i2p = 0.5 / zeta;
    const double Px = (a1 * A.x + a2 * B.x) / zeta, Py = (a1 * A.y + a2 * B.y) / zeta,
                 Pz = (a1 * A.z + a2 * B.z) / zeta;
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double pref = (kPi / zeta) * sqrt(kPi / zeta) * exp(-xi * d2);
    double sx[LA + 1][LB + 1], sy[LA + 1][LB + 1], sz[LA + 1][LB + 1];
    os1d<LA, LB>(i2p, Px - A.x, Px - B.x, sx);
    os1d<LA, LB>(i2p, Py - A.y, Py - B.y, sy);
    os1d<LA, LB>(i2p, Pz - A.z, Pz - B.z, sz);
    double tx[LA + 1][LB + 1], ty[LA + 1][LB + 1], tz[LA + 1][LB + 1];
    os1d_kinetic<LA, LB>(i2p, xi, a1, a2, dx, Px - A.x, Px - B.x, sx, tx);
    os1d_kinetic<LA, LB>(i2p, xi, a1, a2, dy, Py - A.y, Py - B.y, sy, ty);
    os1d_kinetic<LA, LB>(i2p, xi, a1, a2, dz, Pz - A.z, Pz - B.z, sz, tz);
    auto cart = [&](int ca, int cb) {
        const int ax = cart_e(ca, 0), ay = cart_e(ca, 1), az = cart_e(ca, 2);
        const int bx = cart_e(cb, 0), by = cart_e(cb, 1), bz = cart_e(cb, 2);
        return tx[ax][bx] * sy[ay][by] * sz[az][bz] + sx[ax][bx] * ty[ay][by] * sz[az][bz] +
               sx[ax][bx] * sy[ay][by] * tz[az][bz];
    };
    combine_sh<LA, LB>(scale * pref, cart, acc);
}

// electron3CIntg: <a| exp(-a3 |r-C|^2) |b>, the Gaussian product about
// G = (a1 A + a2 B + a3 C)/zeta3; prefactor
// (pi/zeta3)^1.5 exp(-xi |A-B|^2 - (zeta a3/zeta3) |P-C|^2).
template <int LA, int LB, class Acc>
IMAGO_HD void prim_elec3c(double a1, double a2, double a3, Vec3 A,
                                            Vec3 B, Vec3 C, double scale, Acc &acc)
{
    const double zeta = a1 + a2, z3 = zeta + a3, xi = a1 * a2 / zeta, i2p = 0.5 / z3;
    const double Px = (a1 * A.x + a2 * B.x) / zeta, Py = (a1 * A.y + a2 * B.y) / zeta,
                 Pz = (a1 * A.z + a2 * B.z) / zeta;
    const double Gx = (a1 * A.x + a2 * B.x + a3 * C.x) / z3,
                 Gy = (a1 * A.y + a2 * B.y + a3 * C.y) / z3,
                 Gz = (a1 * A.z + a2 * B.z + a3 * C.z) / z3;
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double cx = Px - C.x, cy = Py - C.y, cz = Pz - C.z;
    const double pc2 = cx * cx + cy * cy + cz * cz;
    const double pref =
        (kPi / z3) * sqrt(kPi / z3) * exp(-xi * d2 - (zeta * a3 / z3) * pc2);
    double sx[LA + 1][LB + 1], sy[LA + 1][LB + 1], sz[LA + 1][LB + 1];
    os1d<LA, LB>(i2p, Gx - A.x, Gx - B.x, sx);
    os1d<LA, LB>(i2p, Gy - A.y, Gy - B.y, sy);
    os1d<LA, LB>(i2p, Gz - A.z, Gz - B.z, sz);
    auto cart = [&](int ca, int cb) {
        return sx[cart_e(ca, 0)][cart_e(cb, 0)] * sy[cart_e(ca, 1)][cart_e(cb, 1)] *
               sz[cart_e(ca, 2)][cart_e(cb, 2)];
    };
    combine_sh<LA, LB>(scale * pref, cart, acc);
}

// Hermite expansion coefficients E^{ij}_t about G (McMurchie-Davidson):
//   E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + GA E^{ij}_t + (t+1) E^{ij}_{t+1}
//   E^{i,j+1}_t = E^{ij}_{t-1}/(2p) + GB E^{ij}_t + (t+1) E^{ij}_{t+1}
template <int LA, int LB>
IMAGO_HD void hermite_e(double i2p, double GA, double GB,
                                          double (&E)[LA + 1][LB + 1][LA + LB + 1])
{
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= LA + LB; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
#pragma unroll
    for (int i = 1; i <= LA; ++i)
#pragma unroll
        for (int t = 0; t <= i; ++t) {
            double v = GA * E[i - 1][0][t];
            if (t > 0) v += i2p * E[i - 1][0][t > 0 ? t - 1 : 0];
            if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1 <= LA + LB ? t + 1 : 0];
            E[i][0][t] = v;
        }
#pragma unroll
    for (int j = 1; j <= LB; ++j)
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int t = 0; t <= i + j; ++t) {
                double v = (t <= i + j - 1) ? GB * E[i][j - 1][t] : 0.0;
                if (t > 0) v += i2p * E[i][j - 1][t > 0 ? t - 1 : 0];
                if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1 <= LA + LB ? t + 1 : 0];
                E[i][j][t] = v;
            }
}

// Hermite Coulomb integrals R^0_{tuv}, t+u+v <= L, stored compactly by
// total order s = t+u+v:
//   idx(t,u,v) = s(s+1)(s+2)/6 + (s-t)(s-t+1)/2 + v
// Recursion (McMurchie-Davidson):
//   R^n_{000} = (-2p)^n F_n(T);  R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{tuv}
__host__ __device__ constexpr int tuv_index(int t, int u, int v)
{
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 +
           (u + v) * (u + v + 1) / 2 + v;
}
__host__ __device__ constexpr int tuv_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

template <int L>
struct HermiteR {
    double r[tuv_count(L)];
    IMAGO_HD double operator()(int t, int u, int v) const { return r[tuv_index(t, u, v)]; }
};

template <int L>
IMAGO_HD void hermite_r(double p, double X, double Y, double Z, const double (&F)[L + 1],
                        HermiteR<L> &out)
{
    constexpr int NC = tuv_count(L);
    double nxt[NC], cur[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) nxt[q] = cur[q] = 0.0;
    const double m2p = -2.0 * p;
    double pw[L + 1];
    pw[0] = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * m2p;
    nxt[0] = pw[L] * F[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
        const int lev = L - n;
#pragma unroll
        for (int s = 0; s <= L; ++s) {
            if (s > lev) continue;
#pragma unroll
            for (int t = s; t >= 0; --t) {
#pragma unroll
                for (int v = 0; v <= s - t; ++v) {
                    const int u = s - t - v;
                    double val;
                    if (s == 0) {
                        val = pw[n] * F[n];
                    } else if (t > 0) {
                        val = X * nxt[tuv_index(t - 1, u, v)];
                        if (t > 1) val += (t - 1) * nxt[tuv_index(t > 1 ? t - 2 : 0, u, v)];
                    } else if (u > 0) {
                        val = Y * nxt[tuv_index(0, u - 1, v)];
                        if (u > 1) val += (u - 1) * nxt[tuv_index(0, u > 1 ? u - 2 : 0, v)];
                    } else {
                        val = Z * nxt[tuv_index(0, 0, v - 1)];
                        if (v > 1) val += (v - 1) * nxt[tuv_index(0, 0, v > 1 ? v - 2 : 0)];
                    }
                    cur[tuv_index(t, u, v)] = val;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) nxt[q] = cur[q];
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) out.r[q] = nxt[q];
}

// nuclear3CIntg: int G_A G_B exp(-a3|r-C|^2) / |r-C|, i.e.
// 2 pi/zeta3 exp(-(a1a2|A-B|^2 + a1a3|A-C|^2 + a2a3|B-C|^2)/zeta3)
//   * sum_tuv E_t E_u E_v R_tuv(zeta3, G-C)
template <int LA, int LB, class Acc>
IMAGO_HD void prim_nuc3c(double a1, double a2, double a3, Vec3 A,
                                           Vec3 B, Vec3 C, double scale, Acc &acc)
{
    constexpr int L = LA + LB;
    const double zeta = a1 + a2, z3 = zeta + a3, i2p = 0.5 / z3;
    const double Gx = (a1 * A.x + a2 * B.x + a3 * C.x) / z3,
                 Gy = (a1 * A.y + a2 * B.y + a3 * C.y) / z3,
                 Gz = (a1 * A.z + a2 * B.z + a3 * C.z) / z3;
    const double GCx = Gx - C.x, GCy = Gy - C.y, GCz = Gz - C.z;
    const double d1 = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y) +
                      (A.z - B.z) * (A.z - B.z);
    const double d2 = (A.x - C.x) * (A.x - C.x) + (A.y - C.y) * (A.y - C.y) +
                      (A.z - C.z) * (A.z - C.z);
    const double d3 = (B.x - C.x) * (B.x - C.x) + (B.y - C.y) * (B.y - C.y) +
                      (B.z - C.z) * (B.z - C.z);
    const double U = z3 * (GCx * GCx + GCy * GCy + GCz * GCz);
    double F[L + 1];
    boys_ref<L>(U, F);
    const double pref = 2.0 * (kPi / z3) * exp(-(a1 * a2 * d1 + a1 * a3 * d2 + a2 * a3 * d3) / z3);
    double ex[LA + 1][LB + 1][L + 1], ey[LA + 1][LB + 1][L + 1], ez[LA + 1][LB + 1][L + 1];
    hermite_e<LA, LB>(i2p, Gx - A.x, Gx - B.x, ex);
    hermite_e<LA, LB>(i2p, Gy - A.y, Gy - B.y, ey);
    hermite_e<LA, LB>(i2p, Gz - A.z, Gz - B.z, ez);
    HermiteR<L> R;
    hermite_r<L>(z3, GCx, GCy, GCz, F, R);
    auto cart = [&](int ca, int cb) {
        const int ax = cart_e(ca, 0), ay = cart_e(ca, 1), az = cart_e(ca, 2);
        const int bx = cart_e(cb, 0), by = cart_e(cb, 1), bz = cart_e(cb, 2);
        double s = 0.0;
#pragma unroll
        for (int t = 0; t <= LA + LB; ++t) {
            if (t > ax + bx) continue;
#pragma unroll
            for (int u = 0; u <= LA + LB; ++u) {
                if (u > ay + by || t + u > L) continue;
                double exy = ex[ax][bx][t] * ey[ay][by][u];
#pragma unroll
                for (int v = 0; v <= LA + LB; ++v) {
                    if (v > az + bz || t + u + v > L) continue;
                    s += exy * ez[az][bz][v] * R(t, u, v);
                }
            }
        }
        return s;
    };
    combine_sh<LA, LB>(scale * pref, cart, acc);
}

enum PrimKind { K_OL = 1, K_KE = 2, K_NP = 3, K_EP = 4 };

// Dispatch on the runtime (l_a, l_b) class to the template instance.
template <int KIND, int LMAX, class Acc>
IMAGO_HD void prim_dispatch(int la, int lb, double a1, double a2,
                                              double a3, Vec3 A, Vec3 B, Vec3 C,
                                              double scale, Acc &acc)
{
#define IMAGO_PRIM_CASE(LA_, LB_)                                                   \
    case (LA_) * 4 + (LB_):                                                         \
        if constexpr ((LA_) <= LMAX && (LB_) <= LMAX) {                             \
            if constexpr (KIND == K_OL) prim_overlap<LA_, LB_>(a1, a2, A, B, scale, acc); \
            else if constexpr (KIND == K_KE) prim_kinetic<LA_, LB_>(a1, a2, A, B, scale, acc); \
            else if constexpr (KIND == K_EP) prim_elec3c<LA_, LB_>(a1, a2, a3, A, B, C, scale, acc); \
            else prim_nuc3c<LA_, LB_>(a1, a2, a3, A, B, C, scale, acc);             \
        }                                                                           \
        break;
    switch (la * 4 + lb) {
        IMAGO_PRIM_CASE(0, 0) IMAGO_PRIM_CASE(0, 1) IMAGO_PRIM_CASE(0, 2) IMAGO_PRIM_CASE(0, 3)
        IMAGO_PRIM_CASE(1, 0) IMAGO_PRIM_CASE(1, 1) IMAGO_PRIM_CASE(1, 2) IMAGO_PRIM_CASE(1, 3)
        IMAGO_PRIM_CASE(2, 0) IMAGO_PRIM_CASE(2, 1) IMAGO_PRIM_CASE(2, 2) IMAGO_PRIM_CASE(2, 3)
        IMAGO_PRIM_CASE(3, 0) IMAGO_PRIM_CASE(3, 1) IMAGO_PRIM_CASE(3, 2) IMAGO_PRIM_CASE(3, 3)
    default: break;
    }
#undef IMAGO_PRIM_CASE
}

}  // namespace imago
