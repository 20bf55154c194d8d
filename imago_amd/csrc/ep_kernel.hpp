// ep_kernel.hpp - electronic-potential terms V_j(k), one wavefront per
// (atom pair, potential type).
//
// The reference builds every potential term j with a full pass of the
// atom-pair / image / alpha-pair loop nest (elecPotGaussOverlap,
// integrals.F90:2357-2485, and electronicPE :2700-2856).  For a typical
// cell almost every (pair, term) block receives only a handful of
// primitive tuples, so this kernel gives a wavefront all terms of one
// potential type for one atom pair:
//   for each exponent alpha_j of the type:
//     for each lattice image of atom 2:
//       flatten (screened alpha pair) x (candidate potential image) and
//       apply the reference's exact test |((G-C)-L2)-R|^2 <= threeAlphaDist,
//       ballot-compact the survivors, one primitive block per lane, stage
//       16 blocks at a time through LDS and contract with both atoms'
//       coefficients, then apply the image's Bloch phase;
//     write the block of term j (every k) once.
// The screening tables, candidate images and phases are the same as for
// k_assemble (assemble.hpp), so the included tuples are identical.
#pragma once
#include "assemble.hpp"

namespace imago {

__host__ __device__ inline size_t ep_lds_bytes(int lmax, int ma, int ms, int nk)
{
    const int sh = (lmax + 1) * (lmax + 1);
    size_t d = (size_t)2 * ma * ms + (size_t)16 * sh * sh + (size_t)nk * ms * ms * 2;
    size_t bytes = d * 8 + (size_t)(2 * ms + 16 + 64) * 4 + (size_t)ma * ma * 2;
    return (bytes + 15) & ~(size_t)15;
}

template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_assemble_ep(DevSys S, const PairRec *__restrict__ pairs,
                                                    const ImgRec *__restrict__ imgs,
                                                    const double2 *__restrict__ phase,
                                                    const CandRec *__restrict__ cands,
                                                    const int *__restrict__ cand_off,
                                                    const int *__restrict__ type_term0,
                                                    double *__restrict__ out, long term_stride)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int RQ = 16;
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const PairRec pr = pairs[blockIdx.x];
    const int pt = blockIdx.y;  // potential type
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int e1 = S.type_element[t1], e2 = S.type_element[t2];
    const int n1 = S.type_nalpha[t1], n2 = S.type_nalpha[t2];
    const int ns1 = S.type_nstates[t1], ns2 = S.type_nstates[t2];
    const int MA = S.max_alphas, MS = S.max_states, NEL = S.num_elements;
    const int nk = GAMMA ? 1 : S.nk;

    double *C1 = smem;
    double *C2 = C1 + MA * MS;
    double *stage = C2 + MA * MS;
    double *Q = stage + RQ * SH * SH;
    int *lm1s = (int *)(Q + (size_t)nk * MS * MS * 2);
    int *lm2s = lm1s + MS;
    int *st_a = lm2s + MS;            // [RQ] staged tuple: a1*256+a2
    int *tl = st_a + RQ;              // [64] compacted tuples of a chunk
    unsigned short *alist = (unsigned short *)(tl + 64);

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
    __syncthreads();
    const int nent = ns1 * ns2;
    int em1[NE], em2[NE], elm1[NE], elm2[NE];
#pragma unroll
    for (int r = 0; r < NE; ++r) {
        int e = lane + 64 * r;
        em1[r] = e < nent ? e % ns1 : 0;
        em2[r] = e < nent ? e / ns1 : 0;
        elm1[r] = e < nent ? lm1s[em1[r]] : 0;
        elm2[r] = e < nent ? lm2s[em2[r]] : 0;
    }

    const int pel = S.pot_el[pt];
    const int site_lo = S.pot_first[pt], site_hi = site_lo + S.pot_mult[pt];
    const int na3 = S.pot_nalpha_of(pt);
    const Vec3 A = {S.atom_pos[3 * pr.i], S.atom_pos[3 * pr.i + 1], S.atom_pos[3 * pr.i + 2]};
    const size_t eoff = ((size_t)(e1 * NEL + e2) * MA) * MA;
    const double *ad = S.adist + eoff;
    const double *ac = S.acenter + eoff;
    const int r0 = S.cumul[pr.i], c0g = S.cumul[pr.j];

    for (int aj = 0; aj < na3; ++aj) {
        const double a3 = S.pot_alphas[pt * S.max_pot_alphas + aj];
        const double *apd =
            S.apdist + ((((size_t)pel * S.max_pot_alphas + aj) * NEL + e1) * NEL + e2) * MA * MA;
        for (int idx = lane; idx < nk * nent * 2; idx += 64) Q[idx] = 0.0;
        __syncthreads();
        for (int im = 0; im < pr.img_cnt; ++im) {
            const int img = pr.img_off + im;
            const ImgRec ir = imgs[img];
            const Vec3 Bp = {ir.bx, ir.by, ir.bz};
            const double sep = ir.sep;
            const int clo = cand_off[(size_t)img * (S.num_pot_sites + 1) + site_lo];
            const int chi = cand_off[(size_t)img * (S.num_pot_sites + 1) + site_hi];
            const int ncand = chi - clo;
            if (ncand == 0) continue;
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
            const int total = cnt * ncand;
            for (int base = 0; base < total; base += 64) {
                const int q = base + lane;
                bool pass = false;
                if (q < total) {
#pragma clang fp contract(off)
                    const int code = alist[q / ncand];
                    const int a1 = code >> 8, a2 = code & 255;
                    const CandRec cr = cands[clo + q % ncand];
                    const double c = ac[a1 * MA + a2];
                    const double ocx = c * (A.x - Bp.x) + Bp.x;
                    const double ocy = c * (A.y - Bp.y) + Bp.y;
                    const double ocz = c * (A.z - Bp.z) + Bp.z;
                    const double tad = (1 - sep / ad[a1 * MA + a2]) * apd[a1 * MA + a2];
                    const double *C = S.pot_pos + 3 * cr.site;
                    const double *R = S.cell_dims + 3 * cr.cell;
                    const double cx = ocx - C[0] - cr.lx;
                    const double cy = ocy - C[1] - cr.ly;
                    const double cz = ocz - C[2] - cr.lz;
                    const double dx = cx - R[0], dy = cy - R[1], dz = cz - R[2];
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    pass = !(d2 > tad);
                }
                const unsigned long long m = __ballot(pass);
                if (m == 0ull) continue;
                const int nq = __popcll(m);
                if (pass) tl[__popcll(m & ((1ull << lane) - 1ull))] = q;
                __syncthreads();
                ShAcc<SH> sh;
#pragma unroll
                for (int i = 0; i < SH; ++i)
#pragma unroll
                    for (int j = 0; j < SH; ++j) sh.v[i][j] = 0.0;
                int code = 0;
                if (lane < nq) {
                    const int qq = tl[lane];
                    code = alist[qq / ncand];
                    const int a1 = code >> 8, a2 = code & 255;
                    const CandRec cr = cands[clo + qq % ncand];
                    const double al1 = S.type_alphas[t1 * MA + a1];
                    const double al2 = S.type_alphas[t2 * MA + a2];
                    const int la = lma_to_l(S.type_lma[t1 * MA + a1]);
                    const int lb = lma_to_l(S.type_lma[t2 * MA + a2]);
                    const Vec3 Cp = {cr.px, cr.py, cr.pz};
                    prim_dispatch<K_EP, LMAX>(la, lb, al1, al2, a3, A, Bp, Cp, 1.0, sh);
                }
                for (int r = 0; r * RQ < nq; ++r) {
                    if (lane < nq && (lane / RQ) == r) {
                        double *dst = stage + (lane % RQ) * SH * SH;
#pragma unroll
                        for (int i = 0; i < SH; ++i)
#pragma unroll
                            for (int j = 0; j < SH; ++j) dst[i * SH + j] = sh.v[i][j];
                        st_a[lane % RQ] = code;
                    }
                    __syncthreads();
                    const int nr = min(RQ, nq - r * RQ);
                    for (int q2 = 0; q2 < nr; ++q2) {
                        const int cd = st_a[q2];
                        const double *c1 = C1 + (cd >> 8) * MS;
                        const double *c2 = C2 + (cd & 255) * MS;
                        const double *blk = stage + q2 * SH * SH;
#pragma unroll
                        for (int e = 0; e < NE; ++e)
                            P[e] += c1[em1[e]] * c2[em2[e]] * blk[elm1[e] * SH + elm2[e]];
                    }
                    __syncthreads();
                }
            }
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
        // store the block of term (type pt, alpha aj)
        double *o = out + (size_t)(type_term0[pt] + aj) * term_stride;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int ee = lane + 64 * e;
            if (ee >= nent) continue;
            const long row = r0 + em1[e], col = c0g + em2[e];
            if (row > col) continue;
            const long idx = col * (col + 1) / 2 + row;
            if (GAMMA) {
                o[idx] = Q[2 * ee];
            } else {
                for (int k = 0; k < nk; ++k)
                    reinterpret_cast<double2 *>(o)[(size_t)k * S.npack + idx] =
                        make_double2(Q[2 * ((size_t)k * nent + ee)],
                                     Q[2 * ((size_t)k * nent + ee) + 1]);
            }
        }
        __syncthreads();
    }
}

}  // namespace imago
