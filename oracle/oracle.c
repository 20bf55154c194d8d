/*
 * oracle.c - CPU restatement of the reference's integral + secular path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker (see oracle.h).  It is a
 * plain, single-threaded restatement that keeps the reference's loop
 * order, screening comparisons and storage; it is never used by the
 * product, which must fail rather than fall back to it.
 *
 * Reference: UMKC-CPG/imago, src/imago/*.
 *   primitives   gaussIntegrals.f90 (algorithm restated from Obara-Saika
 *                1986 / Helgaker ch. 9, not from the generated text)
 *   Boys         gaussIntegrals.f90:16080-16221 (formula restated verbatim,
 *                including the default-real (float) literals)
 *   screening    gaussRelations.f90:64-278
 *   assembly     integrals.F90:151-652 (OL), 656-1087 (KE), 1531-1942 +
 *                2533-2697 (NP), 1946-2353 + 2357-2530 + 2700-2856 (EP)
 *   saving       intgSaving.F90:56-89, 132-189, 241-292, 346-509, 704-880
 *   packing      integrals.F90:3747-3910 (ortho with coreDim = 0)
 *   H build      secularEqn.F90:148-171, matrixSubs.F90:339-376
 *   traces       matrixSubs.F90:66-112
 *   lattice      lattice.f90:784-890 (findLatticeVector)
 *   electrostatics  elecStat.f90:59-664 (neutralAndNuclearQPot),
 *                665-1043 (residualQ, restated with its real-space loop
 *                stopping at the first cell beyond range in |R| order);
 *                error_fn mathSubs.f90:22-161
 *   XC mesh      exchCorr.f90:63-342 (getECMeshParameters), 345-875
 *                (makeECMeshAndOverlap), 920-980 (makeSampleVectors)
 *   XC potential potentialUpdate.F90:833-1008, 2835-2894 (Wigner, LDA)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define PI 3.141592653589793
#define SMALL_THRESH 1.0e-8

/* ------------------------------------------------------------------ */
/* Cartesian monomials up to l = 3 and the 16 real-harmonic combinations
 * of sh(16,16) (gaussIntegrals.f90:33-35). */
static const int CART[20][3] = {
    {0, 0, 0},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
    {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1},
    {1, 1, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1}, {1, 0, 2},
    {0, 1, 2}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}};
static const int NCART[4] = {1, 4, 10, 20};

typedef struct {
    int n;
    int c[3];
    double w[3];
} combo_t;

static const combo_t SHC[16] = {
    {1, {0}, {1}},
    {1, {1}, {1}}, {1, {2}, {1}}, {1, {3}, {1}},
    {1, {7}, {1}},                    /* xy          */
    {1, {8}, {1}},                    /* xz          */
    {1, {9}, {1}},                    /* yz          */
    {2, {4, 5}, {1, -1}},             /* xx-yy       */
    {3, {6, 4, 5}, {2, -1, -1}},      /* 2zz-xx-yy   */
    {1, {10}, {1}},                   /* xyz         */
    {2, {12, 14}, {1, -1}},           /* xxz-yyz     */
    {2, {17, 13}, {1, -3}},           /* xxx-3yyx    */
    {2, {11, 18}, {3, -1}},           /* 3xxy-yyy    */
    {3, {19, 12, 14}, {2, -3, -3}},   /* 2zzz-3xxz-3yyz */
    {3, {15, 17, 13}, {4, -1, -1}},   /* 4zzx-xxx-yyx   */
    {3, {16, 11, 18}, {4, -1, -1}}};  /* 4zzy-xxy-yyy   */

static int g_init = 0;
static int g_lcart[20];
static int g_dir[20];          /* direction used to build monomial c   */
static int g_minus[20][3];     /* index of c - 1_i or -1               */

static void cart_init(void)
{
    if (g_init) return;
    for (int c = 0; c < 20; ++c) {
        g_lcart[c] = CART[c][0] + CART[c][1] + CART[c][2];
        g_dir[c] = -1;
        for (int i = 0; i < 3; ++i) {
            if (g_dir[c] < 0 && CART[c][i] > 0) g_dir[c] = i;
            g_minus[c][i] = -1;
            if (CART[c][i] == 0) continue;
            int e[3] = {CART[c][0], CART[c][1], CART[c][2]};
            e[i] -= 1;
            for (int d = 0; d < 20; ++d)
                if (CART[d][0] == e[0] && CART[d][1] == e[1] &&
                    CART[d][2] == e[2])
                    g_minus[c][i] = d;
        }
    }
    g_init = 1;
}

/* l1l2switch = 2^p(lma1) + 16*2^p(lma2) (integrals.F90:533-536) */
static int nib_to_l(int nib)
{
    switch (nib) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    default: return 3;
    }
}

static void decode_switch(int sw, int *La, int *Lb)
{
    *La = nib_to_l(sw & 15);
    *Lb = nib_to_l((sw >> 4) & 15);
}

static void cart_to_sh(int La, int Lb, double pc[20][20], double pref,
                       double *sh)
{
    memset(sh, 0, 256 * sizeof(double));
    int na = (La + 1) * (La + 1), nb = (Lb + 1) * (Lb + 1);
    for (int j = 0; j < nb; ++j)
        for (int i = 0; i < na; ++i) {
            double v = 0.0;
            for (int p = 0; p < SHC[i].n; ++p)
                for (int q = 0; q < SHC[j].n; ++q)
                    v += SHC[i].w[p] * SHC[j].w[q] *
                         pc[SHC[i].c[p]][SHC[j].c[q]];
            sh[i + 16 * j] = v * pref;
        }
}

/* 3-D Obara-Saika overlap recursion, relative to the s-s value:
 * S(a+1_i,b) = PA_i S(a,b) + 1/(2p) [a_i S(a-1_i,b) + b_i S(a,b-1_i)]
 * S(a,b+1_i) = PB_i S(a,b) + 1/(2p) [a_i S(a-1_i,b) + b_i S(a,b-1_i)] */
static void os_overlap(double p, const double *PA, const double *PB, int La,
                       int Lb, double S[20][20])
{
    double i2p = 1.0 / (2.0 * p);
    S[0][0] = 1.0;
    for (int a = 1; a < NCART[La]; ++a) {
        int i = g_dir[a], am = g_minus[a][i];
        double v = PA[i] * S[am][0];
        if (CART[am][i] > 0) v += CART[am][i] * i2p * S[g_minus[am][i]][0];
        S[a][0] = v;
    }
    for (int b = 1; b < NCART[Lb]; ++b) {
        int i = g_dir[b], bm = g_minus[b][i];
        for (int a = 0; a < NCART[La]; ++a) {
            double v = PB[i] * S[a][bm];
            if (CART[a][i] > 0) v += CART[a][i] * i2p * S[g_minus[a][i]][bm];
            if (CART[bm][i] > 0)
                v += CART[bm][i] * i2p * S[a][g_minus[bm][i]];
            S[a][b] = v;
        }
    }
}

/* Obara-Saika kinetic recursion (Obara & Saika 1986, eq. for T):
 * T(a+1_i,b) = PA_i T(a,b) + 1/(2p)[a_i T(a-1_i,b) + b_i T(a,b-1_i)]
 *              + 2 xi [S(a+1_i,b) - a_i/(2 a1) S(a-1_i,b)]
 * and the mirror for b; T(0,0) = xi (3 - 2 xi d^2) S(0,0). */
static void os_kinetic(double p, double xi, double a1, double a2, double d2,
                       const double *PA, const double *PB, int La, int Lb,
                       double S[20][20], double T[20][20])
{
    double i2p = 1.0 / (2.0 * p);
    os_overlap(p, PA, PB, La, Lb, S);
    T[0][0] = xi * (3.0 - 2.0 * xi * d2);
    for (int a = 1; a < NCART[La]; ++a) {
        int i = g_dir[a], am = g_minus[a][i];
        double v = PA[i] * T[am][0] + 2.0 * xi * S[a][0];
        if (CART[am][i] > 0) {
            int amm = g_minus[am][i];
            v += CART[am][i] * i2p * T[amm][0];
            v -= 2.0 * xi * CART[am][i] / (2.0 * a1) * S[amm][0];
        }
        T[a][0] = v;
    }
    for (int b = 1; b < NCART[Lb]; ++b) {
        int i = g_dir[b], bm = g_minus[b][i];
        for (int a = 0; a < NCART[La]; ++a) {
            double v = PB[i] * T[a][bm] + 2.0 * xi * S[a][b];
            if (CART[a][i] > 0) v += CART[a][i] * i2p * T[g_minus[a][i]][bm];
            if (CART[bm][i] > 0) {
                int bmm = g_minus[bm][i];
                v += CART[bm][i] * i2p * T[a][bmm];
                v -= 2.0 * xi * CART[bm][i] / (2.0 * a2) * S[a][bmm];
            }
            T[a][b] = v;
        }
    }
}

/* Obara-Saika nuclear-attraction recursion with auxiliary index m, about
 * the three-Gaussian centre G with exponent p = a1+a2+a3:
 * Th^m(a+1_i,b) = GA_i Th^m(a,b) - GC_i Th^{m+1}(a,b)
 *     + 1/(2p) a_i [Th^m(a-1_i,b) - Th^{m+1}(a-1_i,b)]
 *     + 1/(2p) b_i [Th^m(a,b-1_i) - Th^{m+1}(a,b-1_i)]     (and mirror) */
static void os_nuclear(double p, const double *PA, const double *PB,
                       const double *PC, int La, int Lb, const double *F,
                       double out[20][20])
{
    double th[7][20][20];
    double i2p = 1.0 / (2.0 * p);
    int M = La + Lb;
    for (int m = 0; m <= M; ++m) th[m][0][0] = F[m];
    for (int a = 1; a < NCART[La]; ++a) {
        int i = g_dir[a], am = g_minus[a][i];
        for (int m = 0; m <= M - g_lcart[a]; ++m) {
            double v = PA[i] * th[m][am][0] - PC[i] * th[m + 1][am][0];
            if (CART[am][i] > 0) {
                int amm = g_minus[am][i];
                v += CART[am][i] * i2p * (th[m][amm][0] - th[m + 1][amm][0]);
            }
            th[m][a][0] = v;
        }
    }
    for (int b = 1; b < NCART[Lb]; ++b) {
        int i = g_dir[b], bm = g_minus[b][i];
        for (int a = 0; a < NCART[La]; ++a) {
            for (int m = 0; m <= M - g_lcart[a] - g_lcart[b]; ++m) {
                double v = PB[i] * th[m][a][bm] - PC[i] * th[m + 1][a][bm];
                if (CART[a][i] > 0) {
                    int a1 = g_minus[a][i];
                    v += CART[a][i] * i2p * (th[m][a1][bm] - th[m + 1][a1][bm]);
                }
                if (CART[bm][i] > 0) {
                    int bmm = g_minus[bm][i];
                    v += CART[bm][i] * i2p * (th[m][a][bmm] - th[m + 1][a][bmm]);
                }
                th[m][a][b] = v;
            }
        }
    }
    for (int a = 0; a < NCART[La]; ++a)
        for (int b = 0; b < NCART[Lb]; ++b) out[a][b] = th[0][a][b];
}

/* ------------------------------------------------------------------ */
/* Boys function F_0..F_6, gaussIntegrals.f90:16080-16221.  The series
 * branch and its threshold use the reference's default-real literals,
 * i.e. single-precision values widened to double. */
static const float BOYS_GAMMA_HALF[14] = {
    1.7724538509055161f, 0.88622692545275805f, 1.3293403881791370f,
    3.3233509704478426f, 11.631728396567450f, 52.342777784553533f,
    287.88527781504416f, 1871.2543057977896f, 14034.407293483402f,
    119292.46199460901f, 1133278.3889487833f, 11899423.083962219f,
    136843365.46556622f, 1710542068.3195722f};

void orc_boys(double T, double *F)
{
    double exp_T = exp(-T);
    double sqrt_pi = sqrt(PI);
    double erf_T = erf(sqrt(T));
    if (T <= (double)10E-2f) {
        for (int N = 0; N <= 6; ++N) {
            double a[8];
            for (int i = 0; i < 8; ++i) a[i] = (double)BOYS_GAMMA_HALF[N + i];
            double T2 = T * T, T3 = T2 * T, T4 = T3 * T, T5 = T4 * T,
                   T6 = T5 * T;
            F[N] = 0.5 * exp_T * a[0] *
                   ((1 / a[1]) + (T / a[2]) + (T2 / a[3]) + (T3 / a[4]) +
                    (T4 / a[5]) + (T5 / a[6]) + (T6 / a[7]));
        }
    } else {
        double T2 = T * T, T3 = T2 * T, T4 = T3 * T, T5 = T4 * T, T6 = T5 * T;
        F[0] = 0.5 * ((sqrt_pi / pow(T, 0.5)) * erf_T);
        F[1] = ((sqrt_pi / (4 * pow(T, 1.5))) * erf_T - exp_T * (1 / (2.0 * T)));
        F[2] = 6 * ((sqrt_pi / (16 * pow(T, 2.5))) * erf_T -
                    exp_T * (2 / (24.0 * T) + 1 / ((4) * 2.0 * T2)));
        F[3] = 60 * ((sqrt_pi / (64 * pow(T, 3.5))) * erf_T -
                     exp_T * (6 / (720.0 * T) + 2 / ((4) * 24.0 * T2) +
                              1 / ((16) * 2.0 * T3)));
        F[4] = 840 * ((sqrt_pi / (256 * pow(T, 4.5))) * erf_T -
                      exp_T * (24 / (40320.0 * T) + 6 / ((4) * 720.0 * T2) +
                               2 / ((16) * 24.0 * T3) + 1 / ((64) * 2.0 * T4)));
        F[5] = 15120 * ((sqrt_pi / (1024 * pow(T, 5.5))) * erf_T -
                        exp_T * (120 / (3628800.0 * T) +
                                 24 / ((4) * 40320.0 * T2) +
                                 6 / ((16) * 720.0 * T3) +
                                 2 / ((64) * 24.0 * T4) +
                                 1 / ((256) * 2.0 * T5)));
        F[6] = 332640 * ((sqrt_pi / (4096 * pow(T, 6.5))) * erf_T -
                         exp_T * (720 / (479001600.0 * T) +
                                  120 / ((4) * 3628800.0 * T2) +
                                  24 / ((16) * 40320.0 * T3) +
                                  6 / ((64) * 720.0 * T4) +
                                  2 / ((256) * 24.0 * T5) +
                                  1 / ((1024) * 2.0 * T6)));
    }
}

/* ------------------------------------------------------------------ */
/* Primitive integrals. */

/* overlap2CIntg, gaussIntegrals.f90:26-4420 */
void orc_overlap2c(double a1, double a2, const double *A, const double *B,
                   int sw, double *sh)
{
    cart_init();
    int La, Lb;
    decode_switch(sw, &La, &Lb);
    double zeta = a1 + a2, xi = a1 * a2 / zeta, P[3], PA[3], PB[3], d2 = 0;
    for (int i = 0; i < 3; ++i) {
        P[i] = (a1 * A[i] + a2 * B[i]) / zeta;
        PA[i] = P[i] - A[i];
        PB[i] = P[i] - B[i];
        d2 += (A[i] - B[i]) * (A[i] - B[i]);
    }
    double pref = pow(PI / zeta, 1.5) * exp(-xi * d2);
    double S[20][20];
    os_overlap(zeta, PA, PB, La, Lb, S);
    cart_to_sh(La, Lb, S, pref, sh);
}

/* kinetic2CIntg, gaussIntegrals.f90:4422-11622 */
void orc_kinetic2c(double a1, double a2, const double *A, const double *B,
                   int sw, double *sh)
{
    cart_init();
    int La, Lb;
    decode_switch(sw, &La, &Lb);
    double zeta = a1 + a2, xi = a1 * a2 / zeta, P[3], PA[3], PB[3], d2 = 0;
    for (int i = 0; i < 3; ++i) {
        P[i] = (a1 * A[i] + a2 * B[i]) / zeta;
        PA[i] = P[i] - A[i];
        PB[i] = P[i] - B[i];
        d2 += (A[i] - B[i]) * (A[i] - B[i]);
    }
    double pref = pow(PI / zeta, 1.5) * exp(-xi * d2);
    double S[20][20], T[20][20];
    os_kinetic(zeta, xi, a1, a2, d2, PA, PB, La, Lb, S, T);
    cart_to_sh(La, Lb, T, pref, sh);
}

/* electron3CIntg, gaussIntegrals.f90:11624-16023 */
void orc_electron3c(double a1, double a2, double a3, const double *A,
                    const double *B, const double *C, int sw, double *sh)
{
    cart_init();
    int La, Lb;
    decode_switch(sw, &La, &Lb);
    double zeta = a1 + a2, z3 = zeta + a3, xi = a1 * a2 / zeta;
    double G[3], GA[3], GB[3], d2 = 0, pc2 = 0;
    for (int i = 0; i < 3; ++i) {
        double P = (a1 * A[i] + a2 * B[i]) / zeta;
        G[i] = (a1 * A[i] + a2 * B[i] + a3 * C[i]) / z3;
        GA[i] = G[i] - A[i];
        GB[i] = G[i] - B[i];
        d2 += (A[i] - B[i]) * (A[i] - B[i]);
        pc2 += (P - C[i]) * (P - C[i]);
    }
    double pref = pow(PI / z3, 1.5) * exp(-xi * d2 - (zeta * a3 / z3) * pc2);
    double S[20][20];
    os_overlap(z3, GA, GB, La, Lb, S);
    cart_to_sh(La, Lb, S, pref, sh);
}

/* nuclear3CIntg, gaussIntegrals.f90:16223-23039: the Gaussian-damped
 * nuclear term  int G_A G_B exp(-a3 |r-C|^2) / |r-C|. */
void orc_nuclear3c(double a1, double a2, double a3, const double *A,
                   const double *B, const double *C, int sw, double *sh)
{
    cart_init();
    int La, Lb;
    decode_switch(sw, &La, &Lb);
    double zeta = a1 + a2, z3 = zeta + a3;
    double G[3], GA[3], GB[3], GC[3], dAB = 0, dAC = 0, dBC = 0, gc2 = 0;
    for (int i = 0; i < 3; ++i) {
        G[i] = (a1 * A[i] + a2 * B[i] + a3 * C[i]) / z3;
        GA[i] = G[i] - A[i];
        GB[i] = G[i] - B[i];
        GC[i] = G[i] - C[i];
        dAB += (A[i] - B[i]) * (A[i] - B[i]);
        dAC += (A[i] - C[i]) * (A[i] - C[i]);
        dBC += (B[i] - C[i]) * (B[i] - C[i]);
        gc2 += GC[i] * GC[i];
    }
    double F[7];
    orc_boys(z3 * gc2, F);
    double pref = 2.0 * (PI / z3) *
                  exp(-(a1 * a2 * dAB + a1 * a3 * dAC + a2 * a3 * dBC) / z3);
    double N[20][20];
    os_nuclear(z3, GA, GB, GC, La, Lb, F, N);
    cart_to_sh(La, Lb, N, pref, sh);
}

/* ------------------------------------------------------------------ */
static orc_prim2c_fn g_ext_ol, g_ext_ke;
static orc_prim3c_fn g_ext_ep, g_ext_np;
static int g_use_ext = 0;

void orc_set_external_primitives(orc_prim2c_fn ol, orc_prim2c_fn ke,
                                 orc_prim3c_fn ep, orc_prim3c_fn np)
{
    g_ext_ol = ol;
    g_ext_ke = ke;
    g_ext_ep = ep;
    g_ext_np = np;
}

void orc_use_external_primitives(int on) { g_use_ext = on && g_ext_ol; }

static long g_tuples = 0;
long orc_last_tuple_count(void) { return g_tuples; }

/* Threads of the assembly loops (the outer atom loop of each integral
 * kind runs in parallel; every (i, j) pair writes its own output block and
 * its own interaction-mask words).  1 = the reference's serial order. */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* ------------------------------------------------------------------ */
/* System helpers */

typedef struct {
    const imago_system_desc *s;
    int nel;
    int maxa;                  /* max alphas                           */
    int *el_nalpha;            /* [nel]                                */
    double *el_alpha;          /* [nel][maxa]                          */
    int *el_npot;              /* [nel]                                */
    double *el_pot_alpha;      /* [nel][maxpa]                         */
    double *el_nuc_alpha;      /* [nel]                                */
    int *lma;                  /* [ntypes][maxa] lmAlphaIndex          */
    int *cumul;                /* [natoms] cumulValeStates             */
    int n;                     /* valeDim                              */
    double inv[9];             /* inverse of real vectors (rows)       */
} ctx_t;

static int ctx_init(ctx_t *c, const imago_system_desc *s)
{
    memset(c, 0, sizeof(*c));
    c->s = s;
    int nel = 0;
    for (int t = 0; t < s->num_atom_types; ++t)
        if (s->type_element[t] + 1 > nel) nel = s->type_element[t] + 1;
    for (int t = 0; t < s->num_pot_types; ++t)
        if (s->pot_type_element[t] + 1 > nel) nel = s->pot_type_element[t] + 1;
    c->nel = nel;
    c->maxa = s->max_alphas;
    int maxpa = s->max_pot_alphas > 0 ? s->max_pot_alphas : 1;
    c->el_nalpha = calloc(nel, sizeof(int));
    c->el_alpha = calloc((size_t)nel * c->maxa, sizeof(double));
    c->el_npot = calloc(nel, sizeof(int));
    c->el_pot_alpha = calloc((size_t)nel * maxpa, sizeof(double));
    c->el_nuc_alpha = calloc(nel, sizeof(double));
    /* makeAlphaDist: first type of each element (gaussRelations.f90:98-106) */
    for (int t = 0; t < s->num_atom_types; ++t) {
        int e = s->type_element[t];
        if (c->el_nalpha[e] == 0) {
            c->el_nalpha[e] = s->type_num_orb_alphas[4 * t];
            for (int a = 0; a < c->el_nalpha[e]; ++a)
                c->el_alpha[e * c->maxa + a] = s->type_alphas[t * c->maxa + a];
        }
    }
    /* makeAlphaPotDist / makeAlphaNucDist: first pot type of each element */
    for (int t = 0; t < s->num_pot_types; ++t) {
        int e = s->pot_type_element[t];
        if (c->el_npot[e] == 0) {
            c->el_npot[e] = s->pot_type_num_alphas[t];
            for (int a = 0; a < c->el_npot[e]; ++a)
                c->el_pot_alpha[e * maxpa + a] =
                    s->pot_type_alphas[t * s->max_pot_alphas + a];
        }
        if (c->el_nuc_alpha[e] == 0.0) c->el_nuc_alpha[e] = s->pot_type_nuc_alpha[t];
    }
    /* lmAlphaIndex (basis.f90:124-148) */
    c->lma = calloc((size_t)s->num_atom_types * c->maxa, sizeof(int));
    for (int t = 0; t < s->num_atom_types; ++t) {
        const int *noa = s->type_num_orb_alphas + 4 * t;
        for (int a = 0; a < noa[0]; ++a) {
            int v;
            if (a < noa[3]) v = 16;
            else if (a < noa[2]) v = 9;
            else if (a < noa[1]) v = 4;
            else v = 1;
            c->lma[t * c->maxa + a] = v;
        }
    }
    c->cumul = calloc(s->num_atoms, sizeof(int));
    int n = 0;
    for (int i = 0; i < s->num_atoms; ++i) {
        c->cumul[i] = n;
        n += s->type_num_states[s->atom_type[i]];
    }
    c->n = n;
    /* inverse of the matrix whose columns are the real vectors */
    const double *r = s->real_vectors;
    double m[3][3];
    for (int i = 0; i < 3; ++i)
        for (int x = 0; x < 3; ++x) m[x][i] = r[3 * i + x];
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                 m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double id = 1.0 / det;
    c->inv[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    c->inv[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    c->inv[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    c->inv[3] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
    c->inv[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    c->inv[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    c->inv[6] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
    c->inv[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    c->inv[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    return 0;
}

static void ctx_free(ctx_t *c)
{
    free(c->el_nalpha);
    free(c->el_alpha);
    free(c->el_npot);
    free(c->el_pot_alpha);
    free(c->el_nuc_alpha);
    free(c->lma);
    free(c->cumul);
}

/* alphaCenter(a1,a2,e1,e2) and alphaDist(a1,a2,e1,e2), gaussRelations.f90:
 * 123-129: center = al1/(al1+al2), dist = L / al2 / center. */
static double alpha_center(const ctx_t *c, int a1, int a2, int e1, int e2)
{
    double x = c->el_alpha[e1 * c->maxa + a1], y = c->el_alpha[e2 * c->maxa + a2];
    return x / (x + y);
}

static double alpha_dist(const ctx_t *c, int a1, int a2, int e1, int e2)
{
    double y = c->el_alpha[e2 * c->maxa + a2];
    return c->s->log_basis_fn_thresh / y / alpha_center(c, a1, a2, e1, e2);
}

/* alphaNucDist(a1,a2,ep,e1,e2), gaussRelations.f90:265-268 */
static double alpha_nuc_dist(const ctx_t *c, int a1, int a2, int ep, int e1,
                             int e2)
{
    double x = c->el_alpha[e1 * c->maxa + a1], y = c->el_alpha[e2 * c->maxa + a2];
    double an = c->el_nuc_alpha[ep];
    return c->s->log_basis_fn_thresh * (x + y + an) / (x + y) / an;
}

/* alphaPotDist(a1,a2,j,e1,e2,ep), gaussRelations.f90:194-197 */
static double alpha_pot_dist(const ctx_t *c, int a1, int a2, int j, int e1,
                             int e2, int ep)
{
    int maxpa = c->s->max_pot_alphas > 0 ? c->s->max_pot_alphas : 1;
    double x = c->el_alpha[e1 * c->maxa + a1], y = c->el_alpha[e2 * c->maxa + a2];
    double ap = c->el_pot_alpha[ep * maxpa + j];
    return c->s->log_basis_fn_thresh * (x + y + ap) / (x + y) / ap;
}

int orc_alpha_dist(const imago_system_desc *s, double *alpha_dist_out,
                   double *alpha_center_out, int *num_elements)
{
    ctx_t c;
    ctx_init(&c, s);
    int M = c.maxa, E = c.nel;
    for (int e2 = 0; e2 < E; ++e2)
        for (int e1 = 0; e1 < E; ++e1)
            for (int a2 = 0; a2 < M; ++a2)
                for (int a1 = 0; a1 < M; ++a1) {
                    size_t idx = a1 + (size_t)M * (a2 + (size_t)M * (e1 + (size_t)E * e2));
                    int ok = a1 < c.el_nalpha[e1] && a2 < c.el_nalpha[e2];
                    if (alpha_dist_out)
                        alpha_dist_out[idx] = ok ? alpha_dist(&c, a1, a2, e1, e2) : 0.0;
                    if (alpha_center_out)
                        alpha_center_out[idx] = ok ? alpha_center(&c, a1, a2, e1, e2) : 0.0;
                }
    if (num_elements) *num_elements = E;
    ctx_free(&c);
    return 0;
}

/* findLatticeVector, lattice.f90:784-890: round the fractional coordinates
 * half away from zero, then take the closest point among the nearby cells
 * (the reference consults its Wigner-cell tables for the same search). */
static void find_lattice_vector(const ctx_t *c, const double *v, double *L)
{
    const imago_system_desc *s = c->s;
    const double *r = s->real_vectors;
    double L0[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        double f = c->inv[3 * i] * v[0] + c->inv[3 * i + 1] * v[1] +
                   c->inv[3 * i + 2] * v[2];
        double rnd = (double)(long)(fabs(f) + 0.5);
        if (f < 0.0) rnd = -rnd;
        for (int x = 0; x < 3; ++x) L0[x] += r[3 * i + x] * rnd;
    }
    double rem[3] = {v[0] - L0[0], v[1] - L0[1], v[2] - L0[2]};
    double best = 1e300;
    int bi = 0;
    int lim = s->num_cells < 125 ? s->num_cells : 125;
    for (int k = 0; k < lim; ++k) {
        const double *R = s->cell_dims + 3 * k;
        double d = 0;
        for (int x = 0; x < 3; ++x) d += (rem[x] - R[x]) * (rem[x] - R[x]);
        if (d < best) {
            best = d;
            bi = k;
        }
    }
    for (int x = 0; x < 3; ++x) L[x] = L0[x] + s->cell_dims[3 * bi + x];
}

void orc_find_lattice_vector(const imago_system_desc *s, const double *v,
                             double *L)
{
    ctx_t c;
    ctx_init(&c, s);
    find_lattice_vector(&c, v, L);
    ctx_free(&c);
}

/* ------------------------------------------------------------------ */
/* Assembly (integrals.F90) */

typedef struct {
    int kind;
    /* EP term state (integrals.F90:61-71 module globals) */
    int pot_type;
    int alpha_number;
    double pot_alpha;
    int pot_element;
    int first_site;
    int multiplicity;
    int elec_pot_interaction;
} term_t;

static void prim2(int kind, double a1, double a2, const double *A,
                  const double *B, int sw, double *sh)
{
#pragma omp atomic
    ++g_tuples;
    if (kind == IMAGO_OL)
        (g_use_ext ? g_ext_ol : orc_overlap2c)(a1, a2, A, B, sw, sh);
    else
        (g_use_ext ? g_ext_ke : orc_kinetic2c)(a1, a2, A, B, sw, sh);
}

static void prim3(int kind, double a1, double a2, double a3, const double *A,
                  const double *B, const double *C, int sw, double *sh)
{
#pragma omp atomic
    ++g_tuples;
    if (kind == IMAGO_EP)
        (g_use_ext ? g_ext_ep : orc_electron3c)(a1, a2, a3, A, B, C, sw, sh);
    else
        (g_use_ext ? g_ext_np : orc_nuclear3c)(a1, a2, a3, A, B, C, sw, sh);
}

static int pow2idx(int lma)
{
    return lma == 1 ? 0 : lma == 4 ? 1 : lma == 9 ? 2 : 3;
}

/* nuclearPE, integrals.F90:2533-2697 */
static int nuclear_pe(const ctx_t *c, int a1, int a2, int t1, int t2, int e1,
                      int e2, double sep, const double *A, const double *Bp,
                      double *out)
{
    const imago_system_desc *s = c->s;
    int sw = (1 << pow2idx(c->lma[t1 * c->maxa + a1])) +
             (16 << pow2idx(c->lma[t2 * c->maxa + a2]));
    double al1 = s->type_alphas[t1 * c->maxa + a1];
    double al2 = s->type_alphas[t2 * c->maxa + a2];
    double ac = alpha_center(c, a1, a2, e1, e2);
    double oc[3];
    for (int x = 0; x < 3; ++x) oc[x] = ac * (A[x] - Bp[x]) + Bp[x];
    double acc[256], sh[256];
    memset(acc, 0, sizeof(acc));
    int contrib = 0;
    for (int m = 0; m < s->num_pot_sites; ++m) {
        int pt = s->pot_site_type[m];
        int pel = s->pot_type_element[pt];
        double Z = s->pot_type_nuc_charge[pt];
        double nucAlpha = s->pot_type_nuc_alpha[pt];
        const double *C = s->pot_site_pos + 3 * m;
        if (Z < SMALL_THRESH) continue;
        double tad = (1 - sep / alpha_dist(c, a1, a2, e1, e2)) *
                     alpha_nuc_dist(c, a1, a2, pel, e1, e2);
        double dv[3], L2[3], cov[3];
        for (int x = 0; x < 3; ++x) dv[x] = oc[x] - C[x];
        find_lattice_vector(c, dv, L2);
        double cos2 = 0;
        for (int x = 0; x < 3; ++x) {
            cov[x] = oc[x] - C[x] - L2[x];
            cos2 += cov[x] * cov[x];
        }
        if (cos2 > tad) continue;
        contrib = 1;
        double maxLR2 = cos2 + tad + 2.0 * sqrt(cos2 * tad);
        for (int n = 0; n < s->num_cells; ++n) {
            if (s->cell_sizes[n] > maxLR2) break;
            const double *R = s->cell_dims + 3 * n;
            double sc = 0;
            for (int x = 0; x < 3; ++x) sc += (cov[x] - R[x]) * (cov[x] - R[x]);
            if (sc > tad) continue;
            double Cp[3];
            for (int x = 0; x < 3; ++x) Cp[x] = C[x] + L2[x] + R[x];
            prim3(IMAGO_NP, al1, al2, nucAlpha, A, Bp, Cp, sw, sh);
            for (int q = 0; q < 256; ++q) acc[q] -= sh[q] * Z;
        }
    }
    memcpy(out, acc, sizeof(acc));
    return contrib;
}

/* electronicPE, integrals.F90:2700-2856 */
static int electronic_pe(const ctx_t *c, term_t *tm, int a1, int a2, int t1,
                         int t2, int e1, int e2, double sep, const double *A,
                         const double *Bp, double *out)
{
    const imago_system_desc *s = c->s;
    int sw = (1 << pow2idx(c->lma[t1 * c->maxa + a1])) +
             (16 << pow2idx(c->lma[t2 * c->maxa + a2]));
    double al1 = s->type_alphas[t1 * c->maxa + a1];
    double al2 = s->type_alphas[t2 * c->maxa + a2];
    double ac = alpha_center(c, a1, a2, e1, e2);
    double oc[3];
    for (int x = 0; x < 3; ++x) oc[x] = ac * (A[x] - Bp[x]) + Bp[x];
    double tad = (1 - sep / alpha_dist(c, a1, a2, e1, e2)) *
                 alpha_pot_dist(c, a1, a2, tm->alpha_number, e1, e2,
                                tm->pot_element);
    double acc[256], sh[256];
    memset(acc, 0, sizeof(acc));
    int contrib = 0;
    for (int m = 0; m < tm->multiplicity; ++m) {
        const double *C = s->pot_site_pos + 3 * (tm->first_site + m);
        double dv[3], L2[3], cov[3];
        for (int x = 0; x < 3; ++x) dv[x] = oc[x] - C[x];
        find_lattice_vector(c, dv, L2);
        double cos2 = 0;
        for (int x = 0; x < 3; ++x) {
            cov[x] = oc[x] - C[x] - L2[x];
            cos2 += cov[x] * cov[x];
        }
        if (cos2 > tad) continue;
        contrib = 1;
        tm->elec_pot_interaction = 1;
        double maxLR2 = cos2 + tad + 2.0 * sqrt(cos2 * tad);
        for (int n = 0; n < s->num_cells; ++n) {
            if (s->cell_sizes[n] > maxLR2) break;
            const double *R = s->cell_dims + 3 * n;
            double sc = 0;
            for (int x = 0; x < 3; ++x) sc += (cov[x] - R[x]) * (cov[x] - R[x]);
            if (sc > tad) continue;
            double Cp[3];
            for (int x = 0; x < 3; ++x) Cp[x] = C[x] + L2[x] + R[x];
            prim3(IMAGO_EP, al1, al2, tm->pot_alpha, A, Bp, Cp, sw, sh);
            for (int q = 0; q < 256; ++q) acc[q] += sh[q];
        }
    }
    memcpy(out, acc, sizeof(acc));
    return contrib;
}

/* The atom-pair / lattice / alpha loop nest shared by gaussOverlapOL/KE/NP
 * and gaussOverlapEP.  out: packed [nk][2*npack] or [npack] (gamma). */
static int assemble_one(const ctx_t *c, const term_t *tm0, unsigned long *mask,
                        int nmaskw, double *out)
{
    const imago_system_desc *s = c->s;
    int nk = s->gamma ? 1 : s->num_kpoints;
    int n = c->n;
    long npack = (long)n * (n + 1) / 2;
    int maxa = c->maxa, maxs = s->max_states;
    /* computePhaseFactors, kpoints.f90:578-609 */
    double *phase = malloc(sizeof(double) * 2 * nk * s->num_cells);
    if (!s->gamma)
        for (int k = 0; k < s->num_cells; ++k)
            for (int q = 0; q < nk; ++q) {
                const double *kp = s->kpoints + 3 * q, *R = s->cell_dims + 3 * k;
                double dot = kp[0] * R[0] + kp[1] * R[1] + kp[2] * R[2];
                phase[2 * (q + nk * k)] = cos(dot);
                phase[2 * (q + nk * k) + 1] = sin(dot);
            }
    int rc = 0;
#pragma omp parallel num_threads(g_threads)
    {
    term_t tmv = *tm0;
    term_t *tm = &tmv;
    double *P2 = malloc(sizeof(double) * 16 * maxa * maxs);
    double *P12 = malloc(sizeof(double) * maxs * maxs);
    double *cpr = malloc(sizeof(double) * 2 * maxs * maxs * nk); /* currentPair */
    double sh[256];
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < s->num_atoms; ++i) {
        int stop;
#pragma omp atomic read
        stop = rc;
        if (stop) continue;
        int t1 = s->atom_type[i], e1 = s->type_element[t1];
        int n1 = s->type_num_orb_alphas[4 * t1], ns1 = s->type_num_states[t1];
        const double *A = s->atom_pos + 3 * i;
        for (int j = i; j < s->num_atoms; ++j) {
            int t2 = s->atom_type[j], e2 = s->type_element[t2];
            int n2 = s->type_num_orb_alphas[4 * t2], ns2 = s->type_num_states[t2];
            const double *B = s->atom_pos + 3 * j;
            double dv[3], L[3];
            for (int x = 0; x < 3; ++x) dv[x] = A[x] - B[x];
            find_lattice_vector(c, dv, L);
            double sepsq = 0;
            for (int x = 0; x < 3; ++x)
                sepsq += (A[x] - B[x] - L[x]) * (A[x] - B[x] - L[x]);
            double negl = alpha_dist(c, 0, 0, e1, e2);
            if (sepsq > negl) continue;
            double maxLR = sepsq + negl + 2.0 * sqrt(sepsq * negl);
            if (maxLR > s->cell_sizes[s->num_cells - 1]) {
#pragma omp atomic write
                rc = IMAGO_E_LATTICE;
                break;
            }
            memset(cpr, 0, sizeof(double) * 2 * maxs * maxs * nk);
            for (int k = 0; k < s->num_cells; ++k) {
                if (s->cell_sizes[k] > maxLR) break;
                size_t mi = 0;
                int bit = 0;
                if (tm->kind == IMAGO_EP) {
                    /* checkElecPotInteraction, integrals.F90:2488-2530 */
                    mi = ((size_t)i * s->num_atoms + j) * nmaskw + k / 64;
                    bit = k % 64;
                    if (!((mask[mi] >> bit) & 1UL)) continue;
                    tm->elec_pot_interaction = 0;
                }
                const double *R = s->cell_dims + 3 * k;
                double Bp[3], sep = 0;
                for (int x = 0; x < 3; ++x) {
                    Bp[x] = B[x] + L[x] + R[x];
                    sep += (A[x] - Bp[x]) * (A[x] - Bp[x]);
                }
                if (sep > negl) continue;
                memset(P2, 0, sizeof(double) * 16 * maxa * maxs);
                int maxA1 = 0;
                int lmin = n1 < n2 ? n1 : n2;
                for (int l = 0; l < lmin; ++l) {
                    int mode = 0, ai1 = l, ai2 = l;
                    if (alpha_dist(c, l, l, e1, e2) < sep) break;
                    for (;;) {
                        if (mode == 1) {
                            ++ai1;
                            if (ai1 >= n1) {
                                mode = 2;
                                ai1 = l;
                                continue;
                            }
                        } else if (mode == 2) {
                            ++ai2;
                            if (ai2 >= n2) break;
                        }
                        if (alpha_dist(c, ai1, ai2, e1, e2) < sep) {
                            if (mode == 1) {
                                mode = 2;
                                ai1 = l;
                                continue;
                            }
                            break;
                        }
                        int contrib = 1;
                        double al1 = s->type_alphas[t1 * maxa + ai1];
                        double al2 = s->type_alphas[t2 * maxa + ai2];
                        int lma1 = c->lma[t1 * maxa + ai1];
                        int lma2 = c->lma[t2 * maxa + ai2];
                        int sw = (1 << pow2idx(lma1)) + (16 << pow2idx(lma2));
                        if (tm->kind == IMAGO_OL || tm->kind == IMAGO_KE)
                            prim2(tm->kind, al1, al2, A, Bp, sw, sh);
                        else if (tm->kind == IMAGO_NP)
                            contrib = nuclear_pe(c, ai1, ai2, t1, t2, e1, e2, sep,
                                                 A, Bp, sh);
                        else
                            contrib = electronic_pe(c, tm, ai1, ai2, t1, t2, e1,
                                                    e2, sep, A, Bp, sh);
                        if (contrib) {
                            for (int m = 0; m < ns2; ++m) {
                                int lm2 = s->type_lm_index[t2 * maxs + m] - 1;
                                double cf = s->type_basis[((size_t)t2 * maxs + m) * maxa + ai2];
                                for (int q = 0; q < lma1; ++q)
                                    P2[q + 16 * (ai1 + maxa * m)] += sh[q + 16 * lm2] * cf;
                            }
                            if (ai1 + 1 > maxA1) maxA1 = ai1 + 1;
                        }
                        if (mode == 0) mode = 1;
                    }
                }
                if (!(tm->kind == IMAGO_OL && maxA1 == 0)) {
                    /* multWithBasisFn1, intgSaving.F90:56-89 */
                    for (int l2 = 0; l2 < ns2; ++l2)
                        for (int m1 = 0; m1 < ns1; ++m1) {
                            int lm1 = s->type_lm_index[t1 * maxs + m1] - 1;
                            double v = 0;
                            for (int a = 0; a < maxA1; ++a)
                                v += s->type_basis[((size_t)t1 * maxs + m1) * maxa + a] *
                                     P2[lm1 + 16 * (a + maxa * l2)];
                            P12[m1 + maxs * l2] = v;
                        }
                    /* applyPhaseFactors(Gamma), intgSaving.F90:132-189, 704-734 */
                    for (int q = 0; q < nk; ++q) {
                        double pr = 1.0, pi = 0.0;
                        if (!s->gamma) {
                            pr = phase[2 * (q + nk * k)];
                            pi = phase[2 * (q + nk * k) + 1];
                        }
                        for (int l2 = 0; l2 < ns2; ++l2)
                            for (int m1 = 0; m1 < ns1; ++m1) {
                                double v = P12[m1 + maxs * l2];
                                size_t o = 2 * ((size_t)q * maxs * maxs + m1 + maxs * l2);
                                cpr[o] += pr * v;
                                cpr[o + 1] += pi * v;
                            }
                    }
                }
                if (tm->kind == IMAGO_EP && tm->elec_pot_interaction == 0)
                    mask[mi] &= ~(1UL << bit);
            }
            if (!s->gamma) {
                /* kPointLatticeOriginShift, intgSaving.F90:241-292 */
                for (int q = 0; q < nk; ++q) {
                    const double *kp = s->kpoints + 3 * q;
                    double dot = kp[0] * L[0] + kp[1] * L[1] + kp[2] * L[2];
                    if (fabs(dot) >= SMALL_THRESH) {
                        double cr = cos(dot), ci = sin(dot);
                        for (int l2 = 0; l2 < ns2; ++l2)
                            for (int m1 = 0; m1 < ns1; ++m1) {
                                size_t o = 2 * ((size_t)q * maxs * maxs + m1 + maxs * l2);
                                double re = cpr[o], im = cpr[o + 1];
                                cpr[o] = cr * re - ci * im;
                                cpr[o + 1] = cr * im + ci * re;
                            }
                    }
                }
            }
            /* valeValeSaving + packing (intgSaving.F90:463-509, integrals.F90:
             * 3868-3886): upper triangle only; the diagonal block's strict
             * lower triangle is zeroed. */
            for (int q = 0; q < nk; ++q)
                for (int l2 = 0; l2 < ns2; ++l2)
                    for (int m1 = 0; m1 < ns1; ++m1) {
                        long r = c->cumul[i] + m1, col = c->cumul[j] + l2;
                        if (r > col) continue;
                        long idx = col * (col + 1) / 2 + r;
                        size_t o = 2 * ((size_t)q * maxs * maxs + m1 + maxs * l2);
                        if (s->gamma)
                            out[idx] = cpr[o];
                        else {
                            out[2 * (q * npack + idx)] = cpr[o];
                            out[2 * (q * npack + idx) + 1] = cpr[o + 1];
                        }
                    }
        }
    }
    free(P2);
    free(P12);
    free(cpr);
    }
    free(phase);
    return rc;
}

int orc_assemble(const imago_system_desc *s, int kind, int term_begin,
                 int term_end, double *out)
{
    ctx_t c;
    ctx_init(&c, s);
    g_tuples = 0;
    long npack = (long)c.n * (c.n + 1) / 2;
    int nk = s->gamma ? 1 : s->num_kpoints;
    size_t per = (size_t)npack * (s->gamma ? 1 : 2) * nk;
    int rc = 0;
    term_t tm;
    memset(&tm, 0, sizeof(tm));
    tm.kind = kind;
    if (kind != IMAGO_EP) {
        memset(out, 0, per * sizeof(double));
        rc = assemble_one(&c, &tm, NULL, 0, out);
    } else {
        /* elecPotGaussOverlap, integrals.F90:2357-2485: terms by potential
         * type (in the order of each type's first site), then alpha; the
         * interaction bitmask is reset at the start of each type. */
        int nmaskw = (s->num_cells + 63) / 64;
        size_t nmask = (size_t)s->num_atoms * s->num_atoms * nmaskw;
        unsigned long *mask = malloc(nmask * sizeof(unsigned long));
        int term = 0;
        for (int site = 0; site < s->num_pot_sites && rc == 0; ++site) {
            int pt = s->pot_site_type[site];
            if (s->pot_type_first_site[pt] != site) continue;
            int na = s->pot_type_num_alphas[pt];
            int reset = 1;
            for (int a = 0; a < na && rc == 0; ++a, ++term) {
                if (term < term_begin || term >= term_end) continue;
                if (reset) {
                    for (size_t q = 0; q < nmask; ++q) mask[q] = ~0UL;
                    reset = 0;
                }
                tm.pot_type = pt;
                tm.alpha_number = a;
                tm.pot_alpha = s->pot_type_alphas[pt * s->max_pot_alphas + a];
                tm.pot_element = s->pot_type_element[pt];
                tm.first_site = site;
                tm.multiplicity = s->pot_type_multiplicity[pt];
                double *o = out + (size_t)(term - term_begin) * per;
                memset(o, 0, per * sizeof(double));
                rc = assemble_one(&c, &tm, mask, nmaskw, o);
            }
        }
        free(mask);
    }
    ctx_free(&c);
    return rc;
}

/* ------------------------------------------------------------------ */
/* secularEqn.F90:148-171 with readPackedMatrixAccum (matrixSubs.F90:
 * 339-376): a factor of exactly 0.0 adds the matrix unscaled. */
void orc_build_hamiltonian(long len, const double *np, const double *ke,
                           int pot_dim, const double *const *v,
                           const double *coeffs, double *h)
{
    for (long q = 0; q < len; ++q) h[q] = np[q];
    for (long q = 0; q < len; ++q) h[q] = h[q] + ke[q];
    for (int j = 0; j < pot_dim; ++j) {
        double f = coeffs[j];
        if (f == 0.0)
            for (long q = 0; q < len; ++q) h[q] = h[q] + v[j][q];
        else
            for (long q = 0; q < len; ++q) h[q] = h[q] + v[j][q] * f;
    }
}

/* matrixElementMult(Gamma), matrixSubs.F90:66-112 */
double orc_matrix_element_mult(int dim1, int n, const double *m1,
                               const double *m2)
{
    double t = 0.0;
    long diag = 0;
    for (int i = 1; i <= n; ++i) {
        for (long j = diag + 1; j <= diag + i - 1; ++j) {
            if (dim1 == 2)
                t += m1[2 * (j - 1)] * m2[2 * (j - 1)] +
                     m1[2 * (j - 1) + 1] * m2[2 * (j - 1) + 1];
            else
                t += m1[j - 1] * m2[j - 1];
        }
        diag += i;
        t += m1[dim1 * (diag - 1)] * m2[dim1 * (diag - 1)] / 2.0;
    }
    return t * 2.0;
}

/* ------------------------------------------------------------------ */
/* Electrostatics of the potential-side Gaussians (elecStat.f90).      */

/* error_fn, mathSubs.f90:22-161: the Sandia rational approximation of
 * erf (SAND77-1441; W. J. Cody, Math. Comp. 23 (1969) 631), three ranges
 * |x| <= 0.46875, <= 4, <= 6, and sign(x) beyond; 11 significant digits.
 * The coefficients are the published ones, which the reference uses. */
double orc_error_fn(double x)
{
    const double ax = fabs(x);
    if (ax > 6.0) return x / ax;
    const double x2 = ax * ax;
    if (ax <= 0.46875) {
        double a = ax * (242.6679552305318 +
                         x2 * (21.97926161829415 +
                               x2 * (6.996383488619136 + x2 * -0.03560984370181539)));
        a = a / (215.0588758698612 +
                 x2 * (91.16490540451490 + x2 * (15.08279763040779 + x2 * 1.0)));
        return x < 0.0 ? -a : a;
    }
    double a;
    if (ax <= 4.0) {
        a = exp(-x2) *
            (22.898992851659 +
             ax * (26.094746956075 +
                   ax * (14.571898596926 +
                         ax * (4.2677201070898 +
                               ax * (0.56437160686381 + ax * -0.0000060858151959688)))));
        a = a / (22.898985749891 +
                 ax * (51.933570687552 +
                       ax * (50.273202863803 +
                             ax * (26.288795758761 + ax * (7.5688482293618 + ax * 1.0)))));
    } else {
        const double ix2 = 1.0 / x2;
        a = ix2 * (-0.0121308276389978 +
                   ix2 * (-0.1199039552681460 +
                          ix2 * (-0.243911029488626 + ix2 * -0.0324319519277746))) /
            (0.0430026643452770 +
             ix2 * (0.489552441961437 + ix2 * (1.43771227937118 + ix2 * 1.0)));
        a = exp(-x2) * (0.564189583547756 + a) / ax;
    }
    /* sign(1 - a, x): |1 - a| with the sign of x (1 - a > 0 here) */
    const double v = fabs(1.0 - a);
    return x < 0.0 ? -v : v;
}

/* First term (0-based, in potDim) of every potential type: types in the
 * order of their first site, alphas in order (initAlphaIndex in
 * elecStat.f90:319-331). */
static int *es_type_term0(const imago_system_desc *s)
{
    int *t0 = calloc(s->num_pot_types, sizeof(int));
    int acc = 0;
    for (int i = 0; i < s->num_pot_sites; ++i) {
        const int t = s->pot_site_type[i];
        if (i == 0 || s->pot_site_type[i - 1] != t) {
            t0[t] = acc;
            acc += s->pot_type_num_alphas[t];
        }
    }
    return t0;
}

static int es_pot_dim(const imago_system_desc *s)
{
    int p = 0;
    for (int t = 0; t < s->num_pot_types; ++t) p += s->pot_type_num_alphas[t];
    return p;
}

/* Adds the (a, b) alpha-pair contribution of one site pair and cell
 * (the three branches of elecStat.f90:449-537 share this form). */
static void es_neut_add(int P, int col, int row, double r2, double r, double nqc,
                        double s1, double red, double sred, double poc, double x2isp,
                        double *ovl, double *nl_neut, double *l_neut)
{
    const long at = (long)col * P + row;
    if (fabs(r2) >= SMALL_THRESH)
        nl_neut[at] = nl_neut[at] +
                      nqc / r * (-orc_error_fn(s1 * r) + orc_error_fn(sred * r));
    else
        l_neut[at] = l_neut[at] + nqc * x2isp * sred;
    ovl[at] = ovl[at] + poc * exp(-red * r2);
}

/* neutralAndNuclearQPot, elecStat.f90:59-664.  Column-major potDim x
 * potDim outputs with element (term of site j, term of site i), as the
 * reference's arrays; vectors indexed by the term of site i. */
int orc_elecstat(const imago_system_desc *s, double th, double *ovl, double *nl_neut,
                 double *l_neut, double *nl_nuc, double *l_nuc)
{
    ctx_t c;
    ctx_init(&c, s);
    const int P = es_pot_dim(s);
    const int mpa = s->max_pot_alphas;
    int *t0 = es_type_term0(s);
    memset(ovl, 0, sizeof(double) * (size_t)P * P);
    memset(nl_neut, 0, sizeof(double) * (size_t)P * P);
    memset(l_neut, 0, sizeof(double) * (size_t)P * P);
    memset(nl_nuc, 0, sizeof(double) * P);
    memset(l_nuc, 0, sizeof(double) * P);
    const double pi = PI, x2pi = 2.0 * PI, sqrt_pi = sqrt(PI);
    const double pi3 = PI * PI * PI, pi32 = pow(PI, 1.5), x2isp = 2.0 / sqrt_pi;
    double *red = malloc(sizeof(double) * mpa * mpa);
    double *sred = malloc(sizeof(double) * mpa * mpa);
    double *nqc = malloc(sizeof(double) * mpa * mpa);
    double *poc = malloc(sizeof(double) * mpa * mpa);
    double *s1 = malloc(sizeof(double) * mpa), *s2 = malloc(sizeof(double) * mpa);
    (void)pi;
    for (int i = 0; i < s->num_pot_sites; ++i) {
        const int t1 = s->pot_site_type[i];
        const int n1 = s->pot_type_num_alphas[t1];
        const double *a1 = s->pot_type_alphas + (size_t)t1 * mpa;
        for (int l = 0; l < n1; ++l) s1[l] = sqrt(a1[l]);
        for (int j = 0; j < s->num_pot_sites; ++j) {
            const int t2 = s->pot_site_type[j];
            const int n2 = s->pot_type_num_alphas[t2];
            const double *a2 = s->pot_type_alphas + (size_t)t2 * mpa;
            for (int m = 0; m < n2; ++m) s2[m] = sqrt(a2[m]);
            const int minN = n1 < n2 ? n1 : n2, maxN = n1 > n2 ? n1 : n2;
            /* elecStat.f90:362-393, element (l, k) at [l * mpa + k] */
            for (int k = 0; k < n2; ++k)
                for (int l = 0; l < n1; ++l) {
                    red[l * mpa + k] = a1[l] * a2[k] / (a1[l] + a2[k]);
                    nqc[l * mpa + k] = pi3 / (a1[l] * s1[l] * a2[k] * s2[k]);
                    poc[l * mpa + k] = pow(PI / (a1[l] + a2[k]), 1.5);
                    sred[l * mpa + k] = sqrt(red[l * mpa + k]);
                }
            const double *pi_ = s->pot_site_pos + 3 * i, *pj = s->pot_site_pos + 3 * j;
            const double sep[3] = {pi_[0] - pj[0], pi_[1] - pj[1], pi_[2] - pj[2]};
            double L[3];
            find_lattice_vector(&c, sep, L);
            const double ms[3] = {sep[0] - L[0], sep[1] - L[1], sep[2] - L[2]};
            const double nucA = s->pot_type_nuc_alpha[t2], Z2 = s->pot_type_nuc_charge[t2];
            for (int k = 0; k < s->num_cells; ++k) {
                const double *R = s->cell_dims + 3 * k;
                const double d0 = ms[0] - R[0], d1 = ms[1] - R[1], d2 = ms[2] - R[2];
                const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                const double r = sqrt(r2);
                int above = 0, l;
                for (l = 0; l < minN - 1; ++l) {
                    if (red[l * mpa + l] * r2 > th) {
                        above = 1;
                        break;
                    }
                    for (int m = l; m < n2; ++m)
                        es_neut_add(P, t0[t1] + l, t0[t2] + m, r2, r, nqc[l * mpa + m], s1[l],
                                    red[l * mpa + m], sred[l * mpa + m], poc[l * mpa + m],
                                    x2isp, ovl, nl_neut, l_neut);
                    for (int m = l + 1; m < n1; ++m)
                        es_neut_add(P, t0[t1] + m, t0[t2] + l, r2, r, nqc[m * mpa + l], s1[m],
                                    red[m * mpa + l], sred[m * mpa + l], poc[m * mpa + l],
                                    x2isp, ovl, nl_neut, l_neut);
                }
                if (!above) {
                    /* l = minN - 1 here (the loop variable after a loop
                     * that ran to completion, elecStat.f90:500-537) */
                    for (int m = l; m < maxN; ++m) {
                        const int la = n1 <= n2 ? l : m, lb = n1 <= n2 ? m : l;
                        es_neut_add(P, t0[t1] + la, t0[t2] + lb, r2, r, nqc[la * mpa + lb],
                                    s1[la], red[la * mpa + lb], sred[la * mpa + lb],
                                    poc[la * mpa + lb], x2isp, ovl, nl_neut, l_neut);
                    }
                }
                /* gaussian-screened nuclear charge, elecStat.f90:541-578 */
                for (int m = 0; m < n1; ++m) {
                    const int at = t0[t1] + m;
                    const double nf = a1[m] + nucA;
                    const double scr = r2 * a1[m] * nucA / nf;
                    if (scr < th) {
                        if (fabs(scr) >= SMALL_THRESH) {
                            const double cf = a1[m] * r / sqrt(nf);
                            nl_nuc[at] = nl_nuc[at] -
                                         Z2 * pi32 * exp(-scr) * orc_error_fn(cf) / cf / nf;
                        } else {
                            nl_nuc[at] = nl_nuc[at] - Z2 * exp(-scr) * x2pi / nf;
                        }
                    } else {
                        break;
                    }
                }
            }
            /* bare nuclear local part, elecStat.f90:583-596 */
            if (sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2] < SMALL_THRESH) {
                const double Z1 = s->pot_type_nuc_charge[t1];
                for (int m = 0; m < n1; ++m)
                    l_nuc[t0[t1] + m] = l_nuc[t0[t1] + m] - x2pi * Z1 / a1[m];
            }
        }
    }
    free(red);
    free(sred);
    free(nqc);
    free(poc);
    free(s1);
    free(s2);
    free(t0);
    ctx_free(&c);
    return 0;
}

/* realCellVolume as initializeLattice leaves it (lattice.f90:195-216):
 * (a x b) . c from the last cofactor column, made positive. */
static double es_cell_volume(const double *rv)
{
    const double *a = rv, *b = rv + 3, *cc = rv + 6;
    double v = 0;
    for (int j = 0; j < 3; ++j) {
        const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        v = v + (a[j1] * b[j2] - a[j2] * b[j1]) * cc[j];
    }
    return fabs(v);
}

/* residualQ, elecStat.f90:665-1043: the real-space and reciprocal-space
 * (Ewald) sums of every (potential type of site j, term of site i) against
 * the system's smallest potential alpha.  out is column-major numPotTypes x
 * potDim; the row is the ordinal of site j's type in site order.
 * Reciprocal cells [ncr][3] with sizes |G|^2 ascending, cell 0 the origin.
 * Returns IMAGO_E_ARG when a type has more alphas than the type at the
 * start of its element run (the reference would read stale table rows). */
int orc_residual_q(const imago_system_desc *s, double th, int ncr, const double *cdr,
                   const double *csr, double *out)
{
    return orc_residual_q_sites(s, th, ncr, cdr, csr, 0, s->num_pot_sites, out);
}

/* orc_residual_q restricted to the sites i in [i_begin, i_end) of the
 * outer loop (the columns of their types; others are left 0).  The
 * element-run tables are kept as the full loop would have them. */
int orc_residual_q_sites(const imago_system_desc *s, double th, int ncr, const double *cdr,
                         const double *csr, int i_begin, int i_end, double *out)
{
    ctx_t c;
    ctx_init(&c, s);
    const int P = es_pot_dim(s);
    const int mpa = s->max_pot_alphas;
    int npt_seen = 0;
    for (int i = 0; i < s->num_pot_sites; ++i)
        if (i == 0 || s->pot_site_type[i - 1] != s->pot_site_type[i]) ++npt_seen;
    int *t0 = es_type_term0(s);
    memset(out, 0, sizeof(double) * (size_t)npt_seen * P);
    const double pi32 = pow(PI, 1.5), x2isp = 2.0 / sqrt(PI), sqth = sqrt(th);
    double min_pot = 10000.0;
    for (int t = 0; t < s->num_pot_types; ++t)
        if (s->pot_type_alphas[(size_t)t * mpa] < min_pot)
            min_pot = s->pot_type_alphas[(size_t)t * mpa];
    const double vol = es_cell_volume(s->real_vectors);
    double *caf1 = calloc(mpa, sizeof(double)), *caf2 = calloc(mpa, sizeof(double));
    double *sred0 = calloc(mpa, sizeof(double));   /* reducedPotAlphaSqrt(:, 1) */
    double *mras = calloc(mpa, sizeof(double));
    double *rce = calloc((size_t)mpa * ncr, sizeof(double));
    double *phase = calloc(ncr, sizeof(double));
    int cur_el = -1, src_n = 0, rc = 0;
    for (int i = 0; i < s->num_pot_sites && rc == 0; ++i) {
        const int t1 = s->pot_site_type[i];
        const int n1 = s->pot_type_num_alphas[t1];
        const double *a1 = s->pot_type_alphas + (size_t)t1 * mpa;
        if (i == 0 || s->pot_site_type[i - 1] != t1) {
            for (int k = 0; k < n1; ++k) caf1[k] = pi32 / (a1[k] * sqrt(a1[k]));
            const int last_el = cur_el;
            cur_el = s->pot_type_element[t1];
            if (last_el != cur_el) {
                for (int k = 0; k < n1; ++k) {
                    const double mra = a1[k] * min_pot / (a1[k] + min_pot);
                    mras[k] = sqrt(mra);
                    const double inv4 = 0.25 / mra;
                    for (int q = 1; q < ncr; ++q)
                        rce[(size_t)k * ncr + q] = exp(-csr[q] * inv4) / csr[q];
                }
                src_n = n1;
            }
            if (n1 > src_n) {
                rc = IMAGO_E_ARG;
                break;
            }
        }
        if (i < i_begin || i >= i_end) continue;
        int ord = -1;
        for (int j = 0; j < s->num_pot_sites; ++j) {
            const int t2 = s->pot_site_type[j];
            const double *a2 = s->pot_type_alphas + (size_t)t2 * mpa;
            if (j == 0 || s->pot_site_type[j - 1] != t2) {
                ++ord;
                const int n2 = s->pot_type_num_alphas[t2];
                for (int k = 0; k < n2; ++k) caf2[k] = pi32 / (a2[k] * sqrt(a2[k]));
                for (int k = 0; k < n1; ++k) sred0[k] = sqrt(a1[k] * a2[0] / (a1[k] + a2[0]));
            }
            const double *pi_ = s->pot_site_pos + 3 * i, *pj = s->pot_site_pos + 3 * j;
            const double sep[3] = {pi_[0] - pj[0], pi_[1] - pj[1], pi_[2] - pj[2]};
            for (int q = 0; q < ncr; ++q) {
                const double *G = cdr + 3 * q;
                phase[q] = cos(G[0] * sep[0] + G[1] * sep[1] + G[2] * sep[2]);
            }
            double L[3];
            find_lattice_vector(&c, sep, L);
            const double ms[3] = {sep[0] - L[0], sep[1] - L[1], sep[2] - L[2]};
            for (int k = 0; k < n1; ++k) {
                const long at = (long)(t0[t1] + k) * npt_seen + ord;
                const double af = caf1[k] * caf2[0];
                const double asf = x2isp * af * (sred0[k] - mras[k]);
                const double avf = 4.0 * PI * af / vol;
                for (int l = 0; l < s->num_cells; ++l) {
                    const double *R = s->cell_dims + 3 * l;
                    const double d0 = ms[0] - R[0], d1 = ms[1] - R[1], d2 = ms[2] - R[2];
                    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                    if (r2 < SMALL_THRESH) {
                        out[at] = out[at] + asf;
                    } else {
                        const double r = sqrt(r2);
                        if (mras[k] * r > sqth && sred0[k] * r > sqth) break;
                        out[at] = out[at] + af * (-orc_error_fn(mras[k] * r) +
                                                  orc_error_fn(sred0[k] * r)) / r;
                    }
                }
                double sum = 0.0;
                for (int q = 1; q < ncr; ++q) sum = sum + rce[(size_t)k * ncr + q] * phase[q];
                out[at] = out[at] + avf * (sum - 0.25 / min_pot);
            }
        }
    }
    free(caf1);
    free(caf2);
    free(sred0);
    free(mras);
    free(rce);
    free(phase);
    free(t0);
    ctx_free(&c);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Exchange-correlation mesh (exchCorr.f90).                           */

/* makeSampleVectors, exchCorr.f90:920-980: a spiral of n unit vectors. */
void orc_xc_sample_vectors(int n, double *v)
{
    double *h = malloc(sizeof(double) * n), *th = malloc(sizeof(double) * n);
    double *ph = malloc(sizeof(double) * n);
    h[0] = -1.0;
    h[n - 1] = 1.0;
    for (int k = 2; k <= n - 1; ++k) h[k - 1] = -1.0 + 2.0 * (k - 1.0) / (n - 1.0);
    for (int k = 0; k < n; ++k) th[k] = acos(h[k]);
    ph[0] = 0.0;
    ph[n - 1] = 0.0;
    for (int k = 2; k <= n - 1; ++k) {
        ph[k - 1] = ph[k - 2] + 3.6 / sqrt((double)n) * 1.0 / sqrt(1.0 - h[k - 1] * h[k - 1]);
        while (ph[k - 1] > 2.0 * PI) ph[k - 1] = ph[k - 1] - 2.0 * PI;
    }
    for (int k = 0; k < n; ++k) {
        v[3 * k] = sin(th[k]) * cos(ph[k]);
        v[3 * k + 1] = sin(th[k]) * sin(ph[k]);
        v[3 * k + 2] = cos(th[k]);
    }
    free(h);
    free(th);
    free(ph);
}

static double xc_min_pot_alpha(const imago_system_desc *s)
{
    double m = 10000.0;
    for (int t = 0; t < s->num_pot_types; ++t)
        if (s->pot_type_alphas[(size_t)t * s->max_pot_alphas] < m)
            m = s->pot_type_alphas[(size_t)t * s->max_pot_alphas];
    return m;
}

/* getECMeshParameters, exchCorr.f90:63-342.  vec[nvec][3] from
 * orc_xc_sample_vectors; out: pls[site][3] (potLatticeSep), maxrad
 * [site][nvec] (siteSampleMaxRadius), npts[site] (ray points).  Returns
 * maxNumRayPoints. */
int orc_xc_mesh(const imago_system_desc *s, const imago_xc_mesh_params *p,
                const double *cov, const double *vec, double *pls, double *maxrad,
                int *npts)
{
    ctx_t c;
    ctx_init(&c, s);
    const int ns = s->num_pot_sites, nv = p->num_sample_vectors;
    const int mpa = s->max_pot_alphas;
    const int nreps = s->num_cells < 64 ? s->num_cells : 64;
    const double angw[2] = {p->weight_in, p->weight_out};
    memset(pls, 0, sizeof(double) * 3 * ns);
    memset(maxrad, 0, sizeof(double) * (size_t)ns * nv);
    memset(npts, 0, sizeof(int) * ns);
    for (int i = 0; i < ns; ++i) {
        if (cov[s->pot_site_type[i]] < SMALL_THRESH) continue;
        double L[3];
        find_lattice_vector(&c, s->pot_site_pos + 3 * i, L);
        for (int x = 0; x < 3; ++x) pls[3 * i + x] = s->pot_site_pos[3 * i + x] - L[x];
    }
    const double max_radius = sqrt(p->r_sample_out / xc_min_pot_alpha(s));
    int most = 0;
    for (int i = 0; i < ns; ++i) {
        const int ti = s->pot_site_type[i];
        if (cov[ti] < SMALL_THRESH) continue;
        const double cov_half = cov[ti] / 2.0;
        const double min_radius =
            sqrt(p->r_sample_in / s->pot_type_alphas[(size_t)ti * mpa + s->pot_type_num_alphas[ti] - 1]);
        int cur = 0;
        for (int j = 0; j < nv; ++j) {
            const double *v = vec + 3 * j;
            double cmax = max_radius;
            for (int k = 0; k < ns; ++k) {
                const int tk = s->pot_site_type[k];
                if (cov[tk] < SMALL_THRESH) continue;
                const double f = cov[ti] / (cov[ti] + cov[tk]);
                for (int l = 0; l < nreps; ++l) {
                    double sep[3];
                    for (int x = 0; x < 3; ++x) {
                        const double xl = pls[3 * k + x] + s->cell_dims[3 * l + x];
                        const double x1 = pls[3 * i + x] + s->cell_dims[x];
                        sep[x] = f * (xl - x1);
                    }
                    const double dot = sep[0] * v[0] + sep[1] * v[1] + sep[2] * v[2];
                    if (dot <= SMALL_THRESH) continue;
                    const double q = (sep[0] * sep[0] + sep[1] * sep[1] + sep[2] * sep[2]) / dot;
                    if (q < cmax) cmax = q;
                }
            }
            maxrad[(size_t)i * nv + j] = cmax;
            double r = cmax;
            int w = 1;
            for (;;) {
                ++cur;
                if (r < cov_half || r < min_radius) w = 0;
                if (angw[w] < SMALL_THRESH) break;
                if (r < min_radius) {
                    ++cur;
                    break;
                }
                r = r * p->r_sample_space;
            }
        }
        npts[i] = cur;
        if (cur > most) most = cur;
    }
    ctx_free(&c);
    return most;
}

/* The operator of one ray point (exchCorr.f90:587-760): for every site l
 * and cell within logElecThresh/alpha_1 of the point, exp(-alpha r^2) of
 * the site's alphas up to the first beyond the threshold, added into
 * op[o][k][term] (o = 0 .. numOp-1).  The GGA values follow the
 * reference's recurrences as written. */
static void xc_point_op(const ctx_t *c, const int *t0, double th, const double *pt, int gga,
                        int maxpts, int k, int P, double *op)
{
    const imago_system_desc *s = c->s;
    const int mpa = s->max_pot_alphas;
    for (int l = 0; l < s->num_pot_sites; ++l) {
        const int t = s->pot_site_type[l];
        const int na = s->pot_type_num_alphas[t];
        const double *a = s->pot_type_alphas + (size_t)t * mpa;
        const double negl = th / a[0];
        const double *pl = s->pot_site_pos + 3 * l;
        const double sv[3] = {pt[0] - pl[0], pt[1] - pl[1], pt[2] - pl[2]};
        double L[3];
        find_lattice_vector(c, sv, L);
        const double off[3] = {sv[0] - L[0], sv[1] - L[1], sv[2] - L[2]};
        const double osq = off[0] * off[0] + off[1] * off[1] + off[2] * off[2];
        const double negl_plus = negl + osq + 2.0 * sqrt(negl * osq);
        for (int m = 0; m < s->num_cells; ++m) {
            if (s->cell_sizes[m] > negl_plus) break;
            const double *R = s->cell_dims + 3 * m;
            const double dx = off[0] - R[0], dy = off[1] - R[1], dz = off[2] - R[2];
            const double rad = dx * dx + dy * dy + dz * dz;
            if (rad > negl) continue;
            const double dd[9] = {dx, dy, dz, dx * dx, dx * dy, dx * dz, dy * dy, dy * dz, dz * dz};
            for (int n = 0; n < na; ++n) {
                const double am = a[n] * rad;
                if (am > th) break;
                const long at = (long)k * P + t0[t] + n;
                op[at] = op[at] + exp(-am);
                if (gga) {
                    for (int o = 1; o < 10; ++o) {
                        const long ao = ((long)o * maxpts + k) * P + t0[t] + n;
                        const double f = o < 4 ? a[n] : a[n] * a[n];
                        op[ao] = f * op[ao] + dd[o - 1] * exp(-am);
                    }
                }
            }
        }
    }
}

/* The mesh of one site as makeECMeshAndOverlap builds it (exchCorr.f90:
 * 492-775): its ray points and radial weights, and the operator of every
 * point.  rw[maxpts], op[numOp][maxpts][P], both zeroed first.  Returns
 * the site's number of ray points. */
int orc_xc_site(const imago_system_desc *s, const imago_xc_mesh_params *p, double th,
                const double *cov, const double *vec, const double *pls,
                const double *maxrad, int maxpts, int site, int gga, double *rw, double *op)
{
    ctx_t c;
    ctx_init(&c, s);
    const int P = es_pot_dim(s);
    const int nop = gga ? 10 : 1;
    const int mpa = s->max_pot_alphas, nv = p->num_sample_vectors;
    int *t0 = es_type_term0(s);
    memset(rw, 0, sizeof(double) * maxpts);
    memset(op, 0, sizeof(double) * (size_t)nop * maxpts * P);
    const int ti = s->pot_site_type[site];
    int np = 0;
    if (!(cov[ti] < SMALL_THRESH)) {
        const double angw[2] = {p->weight_in, p->weight_out};
        const double cov_half = cov[ti] / 2.0;
        const double wf = (1.0 - p->r_sample_space) / sqrt(p->r_sample_space);
        const double sqsp = sqrt(p->r_sample_space);
        const double min_radius =
            sqrt(p->r_sample_in / s->pot_type_alphas[(size_t)ti * mpa + s->pot_type_num_alphas[ti] - 1]);
        const double *ls = pls + 3 * site;
        double *pts = malloc(sizeof(double) * 3 * (maxpts + 1));
        for (int j = 0; j < nv; ++j) {
            const double *v = vec + 3 * j;
            double r = maxrad[(size_t)site * nv + j];
            int w = 1;
            const int start = np;
            double epf = 0.5;
            for (;;) {
                for (int x = 0; x < 3; ++x) pts[3 * np + x] = r * v[x] + ls[x];
                rw[np] = angw[w] * wf * epf * (r * r * r);
                ++np;
                epf = 1.0;
                if (r < cov_half || r < min_radius) w = 0;
                if (angw[w] < SMALL_THRESH) break;
                if (r < min_radius) {
                    r = r * sqsp;
                    for (int x = 0; x < 3; ++x) pts[3 * np + x] = 0.75 * r * v[x] + ls[x];
                    rw[np] = angw[0] * (r * r * r) / 3.0;
                    ++np;
                    break;
                }
                r = r * p->r_sample_space;
            }
            for (int k = start; k < np; ++k)
                xc_point_op(&c, t0, th, pts + 3 * k, gga, maxpts, k, P, op);
        }
        free(pts);
    }
    free(t0);
    ctx_free(&c);
    return np;
}

/* exchCorrOverlap, exchCorr.f90:805-810: over every site's ray points in
 * order, column k gets op(k) op(1:k) radialWeight (upper triangle). */
int orc_xc_overlap(const imago_system_desc *s, const imago_xc_mesh_params *p, double th,
                   const double *cov, const double *vec, const double *pls,
                   const double *maxrad, int maxpts, double *ovl)
{
    return orc_xc_overlap_sites(s, p, th, cov, vec, pls, maxrad, maxpts, 0, s->num_pot_sites,
                                ovl);
}

/* The same sum over the sites [i_begin, i_end) only. */
int orc_xc_overlap_sites(const imago_system_desc *s, const imago_xc_mesh_params *p, double th,
                         const double *cov, const double *vec, const double *pls,
                         const double *maxrad, int maxpts, int i_begin, int i_end, double *ovl)
{
    const int P = es_pot_dim(s);
    memset(ovl, 0, sizeof(double) * (size_t)P * P);
    double *rw = malloc(sizeof(double) * maxpts);
    double *op = malloc(sizeof(double) * (size_t)maxpts * P);
    for (int i = i_begin; i < i_end; ++i) {
        const int np = orc_xc_site(s, p, th, cov, vec, pls, maxrad, maxpts, i, 0, rw, op);
        for (int j = 0; j < np; ++j) {
            const double *o = op + (size_t)j * P;
            for (int k = 0; k < P; ++k) {
                if (o[k] == 0.0) continue;   /* adds only zeros */
                double *col = ovl + (size_t)k * P;
                for (int r = 0; r <= k; ++r) col[r] = col[r] + o[k] * o[r] * rw[j];
            }
        }
    }
    free(rw);
    free(op);
    return 0;
}

/* wignerXC and wignerXCEnergy, potentialUpdate.F90:2835-2894 (Hartree). */
static void xc_wigner(double rho, double *pot, double *en)
{
    const double r13 = pow(rho, 0.3333333333333333333);
    const double f = 1.0 + 12.57 * r13;
    *pot = -r13 * (0.984 + (0.943656 + 8.8963 * r13) / (f * f));
    *en = -0.738 * r13 * (1.0 + 0.959 / f);
}

void orc_xc_wigner(double rho, double *pot, double *en) { xc_wigner(rho, pot, en); }

static double xc_wigner_energy(double rho)
{
    const double r13 = pow(rho, 0.3333333333333333333);
    return -0.738 * r13 * (1.0 + 0.959 / (1.0 + 12.57 * r13));
}

/* The exchange-correlation potential terms of makeSCFPot for the Wigner
 * functional (xcCode 100/300), spin 1, LDA (potentialUpdate.F90:833-1008):
 * at every ray point the total and core densities op . rho (negative
 * values set to 0), then exchCorrPot(:, k) += radialWeight * v_k * op for
 * v = (potential, energy density) of the total and the energy density of
 * the core.  out[3][potDim] is exchCorrPot(:, 1:3) before the solve. */
int orc_xc_potential(const imago_system_desc *s, const imago_xc_mesh_params *p, double th,
                     const double *cov, const double *vec, const double *pls,
                     const double *maxrad, int maxpts, const double *rho_total,
                     const double *rho_core, double *out)
{
    const int P = es_pot_dim(s);
    memset(out, 0, sizeof(double) * 3 * (size_t)P);
    double *rw = malloc(sizeof(double) * maxpts);
    double *op = malloc(sizeof(double) * (size_t)maxpts * P);
    for (int i = 0; i < s->num_pot_sites; ++i) {
        const int np = orc_xc_site(s, p, th, cov, vec, pls, maxrad, maxpts, i, 0, rw, op);
        for (int j = 0; j < np; ++j) {
            const double *o = op + (size_t)j * P;
            double core = 0.0, total = 0.0;
            for (int k = 0; k < P; ++k) core = core + o[k] * rho_core[k];
            for (int k = 0; k < P; ++k) total = total + o[k] * rho_total[k];
            if (core < 0.0) core = 0.0;
            if (total < 0.0) total = 0.0;
            double v[3];
            xc_wigner(total, &v[0], &v[1]);
            v[2] = xc_wigner_energy(core);
            for (int k = 0; k < 3; ++k) {
                const double f = rw[j] * v[k];
                double *col = out + (size_t)k * P;
                for (int t = 0; t < P; ++t) col[t] = col[t] + f * o[t];
            }
        }
    }
    free(rw);
    free(op);
    return 0;
}
