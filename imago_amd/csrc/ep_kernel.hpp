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

// LDS of k_assemble_ep: C1, C2, staging slots, Q(k), lm tables, the tuple
// list of one term and the per-image alpha-pair cache.
constexpr int kEpMaxAp = 256;   // cached screened alpha pairs over all images

__host__ __device__ inline int ep_ap_capacity(int ma) { return ma * ma > kEpMaxAp ? ma * ma : kEpMaxAp; }

__host__ __device__ inline size_t ep_lds_bytes(int lmax, int ma, int ms, int nk)
{
    const int sh = (lmax + 1) * (lmax + 1);
    const int apc = ep_ap_capacity(ma);
    size_t d = (size_t)2 * ma * ms + (size_t)16 * sh * sh + (size_t)nk * ms * ms * 2 +
               (size_t)apc * 4;
    size_t ints = (size_t)2 * ms + 16 + 64 * 2 + apc + 2 * 257;
    size_t bytes = d * 8 + ints * 4;
    return (bytes + 15) & ~(size_t)15;
}

// Writes one primitive block straight into an LDS staging slot.
template <int SH>
struct SlotAcc {
    double *o;
    __device__ __forceinline__ void add(int i, int j, double x) { o[i * SH + j] = x; }
};

// Work-group -> (potential type, pair) mapping.  Work-groups are dealt to
// the 8 XCDs round-robin; each XCD is given a contiguous range of the
// sequence (type, pairs ordered by column atom j then row atom i), so the
// 224-byte row chunks that neighbouring pairs write into the same packed
// columns meet in the same L2 and leave it as whole lines.
constexpr int kNumXcd = 8;

__device__ __forceinline__ bool ep_task(int npairs, int ntypes, const int *__restrict__ order,
                                        int &pair, int &type)
{
    const int total = npairs * ntypes;
    const int per = (total + kNumXcd - 1) / kNumXcd;
    const int g = blockIdx.x;
    const int logical = (g % kNumXcd) * per + g / kNumXcd;
    if (logical >= total) return false;
    type = logical / npairs;
    pair = order[logical % npairs];
    return true;
}

template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_assemble_ep(DevSys S, const PairRec *__restrict__ pairs,
                                                    const ImgRec *__restrict__ imgs,
                                                    const double2 *__restrict__ phase,
                                                    const CandRec *__restrict__ cands,
                                                    const int *__restrict__ cand_off,
                                                    const int *__restrict__ type_term0,
                                                    const int *__restrict__ order, int npairs,
                                                    double *__restrict__ out, long term_stride)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int RQ = 16;
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    int pidx, pt;
    if (!ep_task(npairs, S.num_pot_types, order, pidx, pt)) return;
    const PairRec pr = pairs[pidx];
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
    const int APC = ep_ap_capacity(MA);
    double *apd_c = Q + (size_t)nk * MS * MS * 2;       // [APC][4] oc xyz, tad base
    int *lm1s = (int *)(apd_c + (size_t)APC * 4);       // [MS]
    int *lm2s = lm1s + MS;                              // [MS]
    int *st_q = lm2s + MS;                              // [RQ] staged tuple index
    int *tl_a = st_q + RQ;                              // [64] tuple: image << 16 | alpha pair
    int *tl_c = tl_a + 64;                              // [64] tuple: candidate index
    int *ap_code = tl_c + 64;                           // [APC] a1 << 8 | a2
    int *img_ap = ap_code + APC;                        // [257] per-image offsets (+ count)

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
    const size_t csz = (size_t)S.num_pot_sites + 1;

    // Screened alpha pairs of image im with their overlap centre and the
    // factor (1 - |A-B'|^2/alphaDist) of threeAlphaDist, in the reference's
    // operation order (integrals.F90:2765-2775), stored from offset base.
    auto build_ap = [&](int im, int base) -> int {
        const ImgRec ir = imgs[pr.img_off + im];
        int cnt = 0;
        for (int b = 0; b < n1 * n2; b += 64) {
            const int q = b + lane;
            const int a1 = q / n2, a2 = q % n2;
            const bool ok = q < n1 * n2 && !(ad[a1 * MA + a2] < ir.sep);
            const unsigned long long m = __ballot(ok);
            const int pos = base + cnt + __popcll(m & ((1ull << lane) - 1ull));
            if (ok) {
#pragma clang fp contract(off)
                const double c = ac[a1 * MA + a2];
                ap_code[pos] = a1 * 256 + a2;
                apd_c[4 * pos + 0] = c * (A.x - ir.bx) + ir.bx;
                apd_c[4 * pos + 1] = c * (A.y - ir.by) + ir.by;
                apd_c[4 * pos + 2] = c * (A.z - ir.bz) + ir.bz;
                apd_c[4 * pos + 3] = 1 - ir.sep / ad[a1 * MA + a2];
            }
            cnt += __popcll(m);
        }
        return cnt;
    };

    // cache every image's alpha pairs once when they fit
    bool cached = pr.img_cnt <= 256;
    if (cached) {
        int tot = 0;
        for (int im = 0; im < pr.img_cnt && cached; ++im) {
            const ImgRec ir = imgs[pr.img_off + im];
            int cnt = 0;
            for (int b = 0; b < n1 * n2; b += 64) {
                const int q = b + lane;
                const bool ok = q < n1 * n2 && !(ad[(q / n2) * MA + q % n2] < ir.sep);
                cnt += __popcll(__ballot(ok));
            }
            if (lane == 0) img_ap[im] = tot;
            tot += cnt;
            if (tot > APC) cached = false;
        }
        if (cached) {
            if (lane == 0) img_ap[pr.img_cnt] = tot;
            __syncthreads();
            for (int im = 0; im < pr.img_cnt; ++im) build_ap(im, img_ap[im]);
        }
        __syncthreads();
    }

    double P[NE];
    int cur_img = -1;
    // P -> Q(k) with the Bloch phase of image cur_img
    auto flush = [&]() {
        if (cur_img >= 0) {
            const int img = pr.img_off + cur_img;
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const int ee = lane + 64 * e;
                if (ee < nent) {
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
                P[e] = 0.0;
            }
        }
    };

    for (int aj = 0; aj < na3; ++aj) {
        const double a3 = S.pot_alphas[pt * S.max_pot_alphas + aj];
        const double *apd =
            S.apdist + ((((size_t)pel * S.max_pot_alphas + aj) * NEL + e1) * NEL + e2) * MA * MA;
        for (int idx = lane; idx < nk * nent * 2; idx += 64) Q[idx] = 0.0;
#pragma unroll
        for (int e = 0; e < NE; ++e) P[e] = 0.0;
        cur_img = -1;
        int nlist = 0;
        __syncthreads();

        // contract the nlist gathered tuples (tl_a/tl_c) into Q
        auto process = [&]() {
            for (int r = 0; r * RQ < nlist; ++r) {
                const int q = r * RQ + (lane % RQ);
                if (lane < RQ && q < nlist) {
                    const int im = tl_a[q] >> 16, ap = tl_a[q] & 65535;
                    const int code = ap_code[ap];
                    const int a1 = code >> 8, a2 = code & 255;
                    const CandRec cr = cands[tl_c[q]];
                    const ImgRec ir = imgs[pr.img_off + im];
                    const Vec3 Bp = {ir.bx, ir.by, ir.bz};
                    const Vec3 Cp = {cr.px, cr.py, cr.pz};
                    const int la = lma_to_l(S.type_lma[t1 * MA + a1]);
                    const int lb = lma_to_l(S.type_lma[t2 * MA + a2]);
                    double *slot = stage + lane * SH * SH;
                    for (int z = 0; z < SH * SH; ++z) slot[z] = 0.0;
                    SlotAcc<SH> acc{slot};
                    prim_dispatch<K_EP, LMAX>(la, lb, S.type_alphas[t1 * MA + a1],
                                              S.type_alphas[t2 * MA + a2], a3, A, Bp, Cp, 1.0,
                                              acc);
                    st_q[lane] = q;
                }
                __syncthreads();
                const int nr = min(RQ, nlist - r * RQ);
                for (int q2 = 0; q2 < nr; ++q2) {
                    const int q = st_q[q2];
                    const int im = tl_a[q] >> 16;
                    if (im != cur_img) {
                        flush();
                        cur_img = im;
                    }
                    const int code = ap_code[tl_a[q] & 65535];
                    const double *c1 = C1 + (code >> 8) * MS;
                    const double *c2 = C2 + (code & 255) * MS;
                    const double *blk = stage + q2 * SH * SH;
#pragma unroll
                    for (int e = 0; e < NE; ++e)
                        P[e] += c1[em1[e]] * c2[em2[e]] * blk[elm1[e] * SH + elm2[e]];
                }
                __syncthreads();
            }
            nlist = 0;
        };

        for (int im = 0; im < pr.img_cnt; ++im) {
            const int img = pr.img_off + im;
            const int clo = cand_off[(size_t)img * csz + site_lo];
            const int chi = cand_off[(size_t)img * csz + site_hi];
            const int ncand = chi - clo;
            if (ncand == 0) continue;
            int apbase, cnt;
            if (cached) {
                apbase = img_ap[im];
                cnt = img_ap[im + 1] - apbase;
            } else {
                // the list of the previous image must be consumed first
                process();
                apbase = 0;
                cnt = build_ap(im, 0);
                __syncthreads();
            }
            const int total = cnt * ncand;
            for (int b = 0; b < total; b += 64) {
                const int q = b + lane;
                bool pass = false;
                int ap = 0, ci = 0;
                if (q < total) {
#pragma clang fp contract(off)
                    ap = apbase + q / ncand;
                    ci = clo + q % ncand;
                    const int code = ap_code[ap];
                    const CandRec cr = cands[ci];
                    const double tad = apd_c[4 * ap + 3] * apd[(code >> 8) * MA + (code & 255)];
                    const double *C = S.pot_pos + 3 * cr.site;
                    const double *R = S.cell_dims + 3 * cr.cell;
                    const double cx = apd_c[4 * ap + 0] - C[0] - cr.lx;
                    const double cy = apd_c[4 * ap + 1] - C[1] - cr.ly;
                    const double cz = apd_c[4 * ap + 2] - C[2] - cr.lz;
                    const double dx = cx - R[0], dy = cy - R[1], dz = cz - R[2];
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    pass = !(d2 > tad);
                }
                const unsigned long long m = __ballot(pass);
                if (m == 0ull) continue;
                const int np = __popcll(m);
                if (nlist + np > 64) {
                    process();
                }
                if (pass) {
                    const int pos = nlist + __popcll(m & ((1ull << lane) - 1ull));
                    tl_a[pos] = (im << 16) | ap;
                    tl_c[pos] = ci;
                }
                nlist += np;
                __syncthreads();
            }
        }
        process();
        flush();
        __syncthreads();
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
