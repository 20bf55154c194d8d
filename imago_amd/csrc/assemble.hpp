// assemble.hpp - lattice-summed, contracted, Bloch-phased integral blocks.
//
// One wavefront owns one (atom pair, potential term) output block and
// produces every k-point's packed entries of that block exactly once, so
// the packed panels need no clearing and no atomics.  Restates the loop
// nest of src/imago/integrals.F90 (gaussOverlapOL :151-652, KE :656-1087,
// NP :1531-1942 with nuclearPE :2533-2697, EP :1946-2353 with
// electronicPE :2700-2856) and intgSaving.F90 (multWithBasisFn1 :56-89,
// applyPhaseFactors :132-189, kPointLatticeOriginShift :241-292,
// valeValeSaving :463-509):
//   * the lattice images of atom 2 and their Bloch phases (including the
//     origin-shift rule |k.L| < 1e-8 -> no shift) come from the host-built
//     image table, with the shifted positions formed as B + L + R exactly
//     like the reference;
//   * the alpha-pair walk of integrals.F90:451-568 visits exactly the
//     pairs with alphaDist(a1,a2) >= |A-B'|^2 because alphaDist is
//     non-increasing in both indices (checked on the host); the pairs are
//     compacted with a wavefront ballot;
//   * 3-centre terms loop over host-built candidate potential images and
//     apply the reference's test |((G-C)-L2)-R|^2 <= threeAlphaDist with
//     the same operation order (fp contraction off), which makes the EP
//     interaction bitmask of integrals.F90:2488-2530 unnecessary (it only
//     skips images whose test fails for every alpha pair);
//   * contraction with the basis coefficients of both atoms goes through
//     LDS in rounds of 16 primitive blocks; Bloch phases are applied once
//     per lattice image to the contracted block.
#pragma once
#include <type_traits>
#include "prim.hpp"

namespace imago {

struct PairRec {
    int i, j;          // atoms, i <= j
    int img_off, img_cnt;
    int blk;           // first entry of the pair's block in the block layout
};

struct ImgRec {
    double bx, by, bz;  // B + L + R_k
    double sep;         // |A - B'|^2
};

struct CandRec {
    double lx, ly, lz;  // lattice vector L2 (min image of G - C)
    double px, py, pz;  // C + L2 + R_n
    int site, cell;     // potential site, n
};

struct DevSys {
    int num_types, max_alphas, max_states, num_elements, max_pot_alphas;
    int num_atoms, num_pot_sites, num_pot_types, num_cells, nk, vale_dim, pot_dim, gamma;
    long nbe;         // block-layout entries per k-point
    const double *type_alphas;    // [t][a]
    const double *type_basis;     // [t][s][a]
    const int *type_lm;           // [t][s] 1..16
    const int *type_lma;          // [t][a] 1,4,9,16
    const int *type_nalpha, *type_nstates, *type_element;
    const double *adist;          // [e1][e2][a1][a2]
    const double *acenter;        // [e1][e2][a1][a2]
    const double *andist;         // [pe][e1][e2][a1][a2]
    const double *apdist;         // [pe][j][e1][e2][a1][a2]
    const int *atom_type, *cumul;
    const double *atom_pos;       // [atom][3]
    const double *pot_pos;        // [site][3]
    const int *pot_site_type;
    const double *pot_z, *pot_nuc_alpha, *pot_alphas;  // per pot type
    const int *pot_el, *pot_first, *pot_mult;
    const int *term_type, *term_alpha;                   // per EP term
    const double *cell_dims;      // [cell][3]
    const int *pot_nalpha;        // per pot type
    __device__ __forceinline__ int pot_nalpha_of(int t) const { return pot_nalpha[t]; }
};

__device__ __forceinline__ int lma_to_l(int lma)
{
    return lma == 1 ? 0 : lma == 4 ? 1 : lma == 9 ? 2 : 3;
}

template <int SH>
struct ShAcc {
    double v[SH][SH];
    __device__ __forceinline__ void add(int i, int j, double x) { v[i][j] += x; }
    __device__ __forceinline__ double get(int i, int j) const { return v[i][j]; }
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int i = 0; i < SH; ++i)
#pragma unroll
            for (int j = 0; j < SH; ++j) v[i][j] = 0.0;
    }
};

// A lane's primitive block in device memory instead of registers, for the
// nuclear kind with f shells: its 16 x 16 block summed over every nuclear
// site and image needs more registers than a lane has (the register
// version spills 5.8 KB per lane).  Lane-interleaved, so that one entry of
// the 64 lanes is one coalesced access: entry (i, j) of lane l at
// o[(i * SH + j) * 64 + l]; each work-group owns SH * SH * 64 doubles.
template <int SH>
struct LaneAcc {
    double *o;   // the work-group's slice + lane
    __device__ __forceinline__ void add(int i, int j, double x) { o[(i * SH + j) * 64] += x; }
    __device__ __forceinline__ double get(int i, int j) const { return o[(i * SH + j) * 64]; }
    __device__ __forceinline__ void zero()
    {
        for (int q = 0; q < SH * SH; ++q) o[q * 64] = 0.0;
    }
};

// the nuclear kind with f shells accumulates in device memory
template <int KIND, int LMAX>
constexpr bool kLaneAccInMemory = KIND == K_NP && LMAX == 3;
template <int KIND, int LMAX, int SH>
using AssembleAcc =
    typename std::conditional<kLaneAccInMemory<KIND, LMAX>, LaneAcc<SH>, ShAcc<SH>>::type;

// LDS bytes needed by k_assemble for given sizes
__host__ __device__ inline size_t assemble_lds_bytes(int lmax, int ma, int ms, int nk)
{
    const int sh = (lmax + 1) * (lmax + 1);
    size_t d = (size_t)2 * ma * ms + (size_t)16 * sh * sh + (size_t)nk * ms * ms * 2;
    size_t bytes = d * 8 + (size_t)(2 * ms + 16) * 4 + (size_t)ma * ma * 2;
    return (bytes + 15) & ~(size_t)15;
}

template <int KIND, int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_assemble(DevSys S, const PairRec *__restrict__ pairs,
                                                 const ImgRec *__restrict__ imgs,
                                                 const double2 *__restrict__ phase,
                                                 const CandRec *__restrict__ cands,
                                                 const int *__restrict__ cand_off,
                                                 int term0, double *__restrict__ out,
                                                 long term_stride, double *__restrict__ acc_work)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int RQ = 16;  // primitive blocks per LDS round
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const PairRec pr = pairs[blockIdx.x];
    const int term = term0 + (int)blockIdx.y;
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int e1 = S.type_element[t1], e2 = S.type_element[t2];
    const int n1 = S.type_nalpha[t1], n2 = S.type_nalpha[t2];
    const int ns1 = S.type_nstates[t1], ns2 = S.type_nstates[t2];
    const int MA = S.max_alphas, MS = S.max_states, NEL = S.num_elements;
    const int nk = GAMMA ? 1 : S.nk;

    double *C1 = smem;                                  // [MA][MS]
    double *C2 = C1 + MA * MS;                          // [MA][MS]
    double *stage = C2 + MA * MS;                       // [RQ][SH*SH]
    double *Q = stage + RQ * SH * SH;                   // [nk][MS*MS][2]
    int *lm1s = (int *)(Q + (size_t)nk * MS * MS * 2);  // [MS]
    int *lm2s = lm1s + MS;                              // [MS]
    int *st_a = lm2s + MS;                              // [RQ] a1*256+a2
    unsigned short *alist = (unsigned short *)(st_a + RQ);  // [MA*MA]

    for (int idx = lane; idx < n1 * ns1; idx += 64) {
        int a = idx / ns1, m = idx % ns1;
        C1[a * MS + m] = S.type_basis[((size_t)t1 * MS + m) * MA + a];
    }
    for (int idx = lane; idx < n2 * ns2; idx += 64) {
        int a = idx / ns2, m = idx % ns2;
        C2[a * MS + m] = S.type_basis[((size_t)t2 * MS + m) * MA + a];
    }
    for (int m = lane; m < ns1; m += 64) lm1s[m] = S.type_lm[t1 * MS + m] - 1;
    for (int m = lane; m < ns2; m += 64) lm2s[m] = S.type_lm[t2 * MS + m] - 1;
    const int nent = ns1 * ns2;
    for (int idx = lane; idx < nk * nent * 2; idx += 64) Q[idx] = 0.0;
    __syncthreads();

    // owned entries e = lane + 64 r: (m1, m2) = (e % ns1, e / ns1)
    int em1[NE], em2[NE], elm1[NE], elm2[NE];
#pragma unroll
    for (int r = 0; r < NE; ++r) {
        int e = lane + 64 * r;
        em1[r] = e < nent ? e % ns1 : 0;
        em2[r] = e < nent ? e / ns1 : 0;
        elm1[r] = e < nent ? lm1s[em1[r]] : 0;
        elm2[r] = e < nent ? lm2s[em2[r]] : 0;
    }

    double a3 = 0.0;
    int site_lo = 0, site_hi = 0, term_alpha = 0, term_pel = 0;
    if (KIND == K_EP) {
        int pt = S.term_type[term];
        term_alpha = S.term_alpha[term];
        a3 = S.pot_alphas[pt * S.max_pot_alphas + term_alpha];
        term_pel = S.pot_el[pt];
        site_lo = S.pot_first[pt];
        site_hi = site_lo + S.pot_mult[pt];
    } else if (KIND == K_NP) {
        site_hi = S.num_pot_sites;
    }
    const Vec3 A = {S.atom_pos[3 * pr.i], S.atom_pos[3 * pr.i + 1], S.atom_pos[3 * pr.i + 2]};
    const size_t eoff = ((size_t)(e1 * NEL + e2) * MA) * MA;
    const double *ad = S.adist + eoff;
    const double *ac = S.acenter + eoff;

    for (int im = 0; im < pr.img_cnt; ++im) {
        const int img = pr.img_off + im;
        const ImgRec ir = imgs[img];
        const Vec3 Bp = {ir.bx, ir.by, ir.bz};
        const double sep = ir.sep;
        int cnt = 0;
        for (int base = 0; base < n1 * n2; base += 64) {
            int q = base + lane;
            int a1 = q / n2, a2 = q % n2;
            bool ok = q < n1 * n2 && !(ad[a1 * MA + a2] < sep);
            unsigned long long m = __ballot(ok);
            int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
            if (ok) alist[pos] = (unsigned short)(a1 * 256 + a2);
            cnt += __popcll(m);
        }
        __syncthreads();
        double P[NE];
#pragma unroll
        for (int r = 0; r < NE; ++r) P[r] = 0.0;

        for (int c0 = 0; c0 < cnt; c0 += 64) {
            const int q = c0 + lane;
            const bool act = q < cnt;
            AssembleAcc<KIND, LMAX, SH> sh;
            if constexpr (kLaneAccInMemory<KIND, LMAX>)
                sh.o = acc_work + (size_t)blockIdx.x * SH * SH * 64 + lane;
            sh.zero();
            int code = 0;
            if (act) {
                code = alist[q];
                const int a1 = code >> 8, a2 = code & 255;
                const double al1 = S.type_alphas[t1 * MA + a1];
                const double al2 = S.type_alphas[t2 * MA + a2];
                const int la = lma_to_l(S.type_lma[t1 * MA + a1]);
                const int lb = lma_to_l(S.type_lma[t2 * MA + a2]);
                if (KIND == K_OL || KIND == K_KE) {
                    prim_dispatch<KIND, LMAX>(la, lb, al1, al2, 0.0, A, Bp, Bp, 1.0, sh);
                } else {
                    double ocx, ocy, ocz, tad_base;
                    {
#pragma clang fp contract(off)
                        const double c = ac[a1 * MA + a2];
                        ocx = c * (A.x - Bp.x) + Bp.x;
                        ocy = c * (A.y - Bp.y) + Bp.y;
                        ocz = c * (A.z - Bp.z) + Bp.z;
                        tad_base = 1 - sep / ad[a1 * MA + a2];
                    }
                    const int clo = cand_off[(size_t)img * S.num_pot_sites + site_lo];
                    const int chi = cand_off[(size_t)img * S.num_pot_sites + site_hi];
                    for (int ci = clo; ci < chi; ++ci) {
                        const CandRec cr = cands[ci];
                        const int pt = S.pot_site_type[cr.site];
                        // nuclearPE skips sites with Z < smallThresh (integrals.F90:2619)
                        if (KIND == K_NP && S.pot_z[pt] < 1.0e-8) continue;
                        bool take;
                        {
#pragma clang fp contract(off)
                            const double *C = S.pot_pos + 3 * cr.site;
                            const double *R = S.cell_dims + 3 * cr.cell;
                            double tad;
                            if (KIND == K_NP)
                                tad = tad_base *
                                      S.andist[(((size_t)S.pot_el[pt] * NEL + e1) * NEL + e2) * MA * MA +
                                               a1 * MA + a2];
                            else
                                tad = tad_base *
                                      S.apdist[((((size_t)term_pel * S.max_pot_alphas + term_alpha) * NEL + e1) *
                                                    NEL + e2) * MA * MA + a1 * MA + a2];
                            const double cx = ocx - C[0] - cr.lx;
                            const double cy = ocy - C[1] - cr.ly;
                            const double cz = ocz - C[2] - cr.lz;
                            const double dx = cx - R[0], dy = cy - R[1], dz = cz - R[2];
                            const double d2 = dx * dx + dy * dy + dz * dz;
                            take = !(d2 > tad);
                        }
                        if (!take) continue;
                        const Vec3 Cp = {cr.px, cr.py, cr.pz};
                        if (KIND == K_NP)
                            prim_dispatch<K_NP, LMAX>(la, lb, al1, al2, S.pot_nuc_alpha[pt], A, Bp,
                                                      Cp, -S.pot_z[pt], sh);
                        else
                            prim_dispatch<K_EP, LMAX>(la, lb, al1, al2, a3, A, Bp, Cp, 1.0, sh);
                    }
                }
            }
            // contraction in rounds of RQ primitive blocks through LDS
            const int nq = min(64, cnt - c0);
            for (int r = 0; r * RQ < nq; ++r) {
                if (act && (lane / RQ) == r) {
                    double *dst = stage + (lane % RQ) * SH * SH;
#pragma unroll
                    for (int i = 0; i < SH; ++i)
#pragma unroll
                        for (int j = 0; j < SH; ++j) dst[i * SH + j] = sh.get(i, j);
                    st_a[lane % RQ] = code;
                }
                __syncthreads();
                const int nr = min(RQ, nq - r * RQ);
                for (int qq = 0; qq < nr; ++qq) {
                    const int cd = st_a[qq];
                    const double *c1 = C1 + (cd >> 8) * MS;
                    const double *c2 = C2 + (cd & 255) * MS;
                    const double *blk = stage + qq * SH * SH;
#pragma unroll
                    for (int e = 0; e < NE; ++e)
                        P[e] += c1[em1[e]] * c2[em2[e]] * blk[elm1[e] * SH + elm2[e]];
                }
                __syncthreads();
            }
        }
        // Bloch sum: Q(k) += phase(k, image) * P
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int ee = lane + 64 * e;
            if (ee >= nent) continue;
            if (GAMMA) {
                Q[2 * ee] += P[e];
            } else {
                for (int k = 0; k < nk; ++k) {
                    const double2 ph = phase[(size_t)img * nk + k];
                    Q[2 * ((size_t)k * nent + ee)] += ph.x * P[e];
                    Q[2 * ((size_t)k * nent + ee) + 1] += ph.y * P[e];
                }
            }
        }
        __syncthreads();
    }

    // store: the pair's block, column-major over (row state, column state),
    // at its offset in the block layout; the strict lower triangle of a
    // diagonal block is zero (valeValeSaving, intgSaving.F90:433-446)
    double *o = out + (size_t)term * term_stride;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int ee = lane + 64 * e;
        if (ee >= nent) continue;
        const bool zero = pr.i == pr.j && em1[e] > em2[e];
        const long idx = (long)pr.blk + ee;
        if (GAMMA) {
            o[idx] = zero ? 0.0 : Q[2 * ee];
        } else {
            for (int k = 0; k < nk; ++k) {
                double2 v = zero ? make_double2(0.0, 0.0)
                                 : make_double2(Q[2 * ((size_t)k * nent + ee)],
                                                Q[2 * ((size_t)k * nent + ee) + 1]);
                reinterpret_cast<double2 *>(o)[(size_t)k * S.nbe + idx] = v;
            }
        }
    }
}

}  // namespace imago
