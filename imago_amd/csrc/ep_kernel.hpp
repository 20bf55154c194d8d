// ep_kernel.hpp - electronic-potential terms V_j(k) in two stages.
//
// The reference builds every potential term j with a full pass of the
// atom-pair / image / alpha-pair loop nest (elecPotGaussOverlap,
// integrals.F90:2357-2485, and electronicPE :2700-2856): for term j of
// potential type t it walks the lattice images of atom 2, the alpha pairs
// within alphaDist, the sites of type t and their images, and keeps the
// tuples with |((G-C)-L2)-R|^2 <= threeAlphaDist(j).  Only threeAlphaDist
// depends on j, so here the walk runs once per (atom pair, potential type):
//
//   k_ep_screen  one wavefront per (pair, type): every (image, alpha pair,
//                candidate site image) with the distance evaluated once and
//                compared against the threshold of each exponent of the type;
//                survivors are written as EpTuple {image, alpha pair,
//                candidate, 64-bit exponent mask}.  Run twice (count, then
//                fill at the scanned offsets), once per system.
//   k_ep_eval    one wavefront per (pair, type): the (tuple, exponent) items
//                in exponent-major order, RQ primitive blocks evaluated side
//                by side and staged in LDS, contracted with both atoms'
//                coefficients, Bloch-phased per image, and every term's block
//                written once.
//
// The tuples are the ones the reference evaluates (same tables, candidate
// images, operation order of the distance test), and each term sums them in
// the reference's image order.
#pragma once
#include "assemble.hpp"

namespace imago {

// One screened (image, alpha pair, candidate) of an (atom pair, potential
// type); bit a of mask is set when exponent a of the type passes.  The
// primitive blocks of its set bits are items item, item + 1, ... in
// ascending exponent order.
struct EpTuple {
    int key;                 // image << 16 | a1 << 8 | a2
    int cand;                // index into the candidate table
    unsigned long long mask;
    int item;                // first item of this tuple
    int task;                // potential type * npairs + pair
};

constexpr int kEpMaxPotAlphas = 64;
constexpr int kEpdTraceChunks = 8;   // parts of a potential type's items in the trace sums
constexpr int kNumXcd = 8;

__host__ __device__ inline size_t ep_screen_lds_bytes(int ma)
{
    return (size_t)ma * ma * (4 * sizeof(double) + sizeof(int));
}

// Screened alpha pairs of one image (ballot-compacted into LDS) with the
// overlap centre and the factor (1 - |A-B'|^2/alphaDist) of threeAlphaDist,
// in the reference's operation order (integrals.F90:2765-2775).
__device__ __forceinline__ int ep_build_ap(const DevSys &S, const ImgRec &ir, const Vec3 &A,
                                           const double *__restrict__ ad,
                                           const double *__restrict__ ac, int n1, int n2,
                                           double *apd_c, int *ap_code)
{
    const int lane = threadIdx.x;
    const int MA = S.max_alphas;
    int cnt = 0;
    for (int b = 0; b < n1 * n2; b += 64) {
        const int q = b + lane;
        const int a1 = q / n2, a2 = q % n2;
        const bool ok = q < n1 * n2 && !(ad[a1 * MA + a2] < ir.sep);
        const unsigned long long m = __ballot(ok);
        const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
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
}

// task = type * npairs + pair.
// FILL == false: tcnt[task] = number of tuples, icnt[task] = number of items.
// FILL == true:  the tuples, in image order, from offset toff[task], their
//                items numbered from ioff[task], and item_tuple[item] = the
//                index of the tuple the item belongs to.
template <bool FILL>
__global__ __launch_bounds__(64) void k_ep_screen(DevSys S, const PairRec *__restrict__ pairs,
                                                  const ImgRec *__restrict__ imgs,
                                                  const CandRec *__restrict__ cands,
                                                  const int *__restrict__ cand_off, int npairs,
                                                  long *__restrict__ tseg,
                                                  long *__restrict__ iseg,
                                                  EpTuple *__restrict__ out,
                                                  int *__restrict__ item_tuple)
{
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const long task = blockIdx.x;
    const int pidx = (int)(task % npairs), pt = (int)(task / npairs);
    const PairRec pr = pairs[pidx];
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int e1 = S.type_element[t1], e2 = S.type_element[t2];
    const int n1 = S.type_nalpha[t1], n2 = S.type_nalpha[t2];
    const int MA = S.max_alphas, NEL = S.num_elements;
    double *apd_c = smem;                                    // [MA*MA][4]
    int *ap_code = (int *)(apd_c + (size_t)MA * MA * 4);     // [MA*MA]

    const int pel = S.pot_el[pt];
    const int site_lo = S.pot_first[pt], site_hi = site_lo + S.pot_mult[pt];
    const int na3 = S.pot_nalpha_of(pt);
    const Vec3 A = {S.atom_pos[3 * pr.i], S.atom_pos[3 * pr.i + 1], S.atom_pos[3 * pr.i + 2]};
    const size_t eoff = ((size_t)(e1 * NEL + e2) * MA) * MA;
    const double *ad = S.adist + eoff;
    const double *ac = S.acenter + eoff;
    // alphaPotDist of exponent a: apd0[a * aj_stride + a1 * MA + a2]
    const size_t aj_stride = (size_t)NEL * NEL * MA * MA;
    const double *apd0 = S.apdist + (size_t)pel * S.max_pot_alphas * aj_stride + eoff;
    // candidates of (image, site) at cand_off[image * sites + site]; the
    // entry after the last site of an image is the next image's first
    const size_t csz = (size_t)S.num_pot_sites;

    long tpos = FILL ? tseg[task] : 0;
    long ipos = FILL ? iseg[task] : 0;
    for (int im = 0; im < pr.img_cnt; ++im) {
        const int img = pr.img_off + im;
        const int clo = cand_off[(size_t)img * csz + site_lo];
        const int chi = cand_off[(size_t)img * csz + site_hi];
        const int ncand = chi - clo;
        if (ncand == 0) continue;
        const ImgRec ir = imgs[img];
        __syncthreads();
        const int cnt = ep_build_ap(S, ir, A, ad, ac, n1, n2, apd_c, ap_code);
        __syncthreads();
        const int total = cnt * ncand;
        for (int b = 0; b < total; b += 64) {
            const int q = b + lane;
            unsigned long long mask = 0ull;
            int code = 0, ci = 0;
            if (q < total) {
#pragma clang fp contract(off)
                const int ap = q / ncand;
                ci = clo + q % ncand;
                code = ap_code[ap];
                const CandRec cr = cands[ci];
                const double *C = S.pot_pos + 3 * cr.site;
                const double *R = S.cell_dims + 3 * cr.cell;
                const double cx = apd_c[4 * ap + 0] - C[0] - cr.lx;
                const double cy = apd_c[4 * ap + 1] - C[1] - cr.ly;
                const double cz = apd_c[4 * ap + 2] - C[2] - cr.lz;
                const double dx = cx - R[0], dy = cy - R[1], dz = cz - R[2];
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double f = apd_c[4 * ap + 3];
                const double *apq = apd0 + (code >> 8) * MA + (code & 255);
                for (int a = 0; a < na3; ++a) {
                    const double tad = f * apq[a * aj_stride];
                    if (!(d2 > tad)) mask |= 1ull << a;
                }
            }
            const unsigned long long m = __ballot(mask != 0ull);
            // items of the lanes below: inclusive wave scan of popcount(mask)
            int nit = __popcll(mask);
            int incl = nit;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int v = __shfl_up(incl, d, 64);
                if (lane >= d) incl += v;
            }
            const int wave_items = __shfl(incl, 63, 64);
            if (FILL && mask != 0ull) {
                const long tidx = tpos + __popcll(m & ((1ull << lane) - 1ull));
                const long first = ipos + incl - nit;
                EpTuple t;
                t.key = (im << 16) | code;
                t.cand = ci;
                t.mask = mask;
                t.item = (int)first;
                t.task = (int)task;
                out[tidx] = t;
                for (int r = 0; r < nit; ++r) item_tuple[first + r] = (int)tidx;
            }
            tpos += __popcll(m);
            ipos += wave_items;
        }
    }
    if (!FILL && lane == 0) {
        tseg[task] = tpos;
        iseg[task] = ipos;
    }
}

// k_ep_prims: one lane per item (tuple, exponent): the primitive block of
// electron3CIntg (gaussIntegrals.f90:11624) for the tuple's alpha pair,
// image and candidate site image with the exponent's alpha, written whole
// (SH x SH, lm1-major) to blocks + item * SH * SH.  A wavefront's 64 items
// are consecutive, so their blocks are one contiguous range: a quarter of
// the lanes at a time evaluates into an LDS tile, which the whole
// wavefront then copies out with coalesced stores (a block written
// directly by its lane would be 81 stores 648 B apart).
template <int SH>
struct BlockAcc {
    double *o;
    __device__ __forceinline__ void add(int i, int j, double x) { o[i * SH + j] = x; }
};

constexpr int kEpPrimsTile = 16;   // blocks per LDS tile

__host__ __device__ inline size_t ep_prims_lds_bytes(int lmax)
{
    const int sh = (lmax + 1) * (lmax + 1);
    return (size_t)kEpPrimsTile * sh * sh * sizeof(double);
}

template <int LMAX>
__global__ __launch_bounds__(64) void k_ep_prims(DevSys S, const PairRec *__restrict__ pairs,
                                                 const ImgRec *__restrict__ imgs,
                                                 const CandRec *__restrict__ cands,
                                                 const EpTuple *__restrict__ tuples,
                                                 const int *__restrict__ item_tuple, long nitems,
                                                 int npairs, double *__restrict__ blocks)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    extern __shared__ double tile[];   // [kEpPrimsTile][BS]
    const int lane = threadIdx.x;
    const long base = (long)blockIdx.x * 64;
    const long it = base + lane;
    for (int q = 0; q < 64 / kEpPrimsTile; ++q) {
        const long t0 = base + q * kEpPrimsTile;
        if (t0 >= nitems) break;
        if (lane / kEpPrimsTile == q && it < nitems) {
            const EpTuple t = tuples[item_tuple[it]];
            // exponent: the (it - t.item)-th set bit of the mask
            unsigned long long m = t.mask;
            for (long r = it - t.item; r > 0; --r) m &= m - 1ull;
            const int aj = __ffsll((long long)m) - 1;
            const int pidx = t.task % npairs, pt = t.task / npairs;
            const PairRec pr = pairs[pidx];
            const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
            const int MA = S.max_alphas;
            const int im = t.key >> 16, a1 = (t.key >> 8) & 255, a2 = t.key & 255;
            const CandRec cr = cands[t.cand];
            const ImgRec ir = imgs[pr.img_off + im];
            const Vec3 A = {S.atom_pos[3 * pr.i], S.atom_pos[3 * pr.i + 1],
                            S.atom_pos[3 * pr.i + 2]};
            const Vec3 Bp = {ir.bx, ir.by, ir.bz};
            const Vec3 Cp = {cr.px, cr.py, cr.pz};
            const int la = lma_to_l(S.type_lma[t1 * MA + a1]);
            const int lb = lma_to_l(S.type_lma[t2 * MA + a2]);
            double *o = tile + (lane % kEpPrimsTile) * BS;
            for (int z = 0; z < BS; ++z) o[z] = 0.0;
            BlockAcc<SH> acc{o};
            prim_dispatch<K_EP, LMAX>(la, lb, S.type_alphas[t1 * MA + a1],
                                      S.type_alphas[t2 * MA + a2],
                                      S.pot_alphas[(size_t)pt * S.max_pot_alphas + aj], A, Bp, Cp,
                                      1.0, acc);
        }
        __syncthreads();
        const long nb = min((long)kEpPrimsTile, nitems - t0);
        double *dst = blocks + (size_t)t0 * BS;
        for (int z = lane; z < nb * BS; z += 64) dst[z] = tile[z];
        __syncthreads();
    }
}

// Work-group -> (potential type, pair) mapping.  Work-groups are dealt to
// the 8 XCDs round-robin; each XCD is given a contiguous range of the
// sequence (type, pairs ordered by column atom j then row atom i), so the
// row chunks that neighbouring pairs write into the same packed columns
// meet in the same L2 and leave it as whole lines.
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

// k_ep_store: one wavefront per (potential term, pair).  Wavefronts are
// dealt to the XCDs in contiguous runs of the sequence (term, pairs in
// block-layout order), so consecutive wavefronts write consecutive blocks
// of one term's panel.  A wavefront contracts the items of its term's
// exponent (tuples in image order) with both atoms' coefficients
// (multWithBasisFn1, intgSaving.F90:56-89), sums them per image, applies
// the Bloch phase (applyPhaseFactors, intgSaving.F90:132-189) and writes
// the block once.  It has no LDS and issues no load after its stores, so
// many wavefronts per CU overlap their load latencies.
// NKR > 0: Q(k) for nk <= NKR k-points in registers; NKR == 0: in LDS.
__host__ __device__ inline size_t ep_store_lds_bytes(int ms, int nk, bool qreg)
{
    return qreg ? 0 : (size_t)nk * ms * ms * 2 * sizeof(double);
}

template <int LMAX, int NE, int NKR, bool GAMMA>
__global__ __launch_bounds__(64) void k_ep_store(
    DevSys S, const PairRec *__restrict__ pairs, const double2 *__restrict__ phase,
    const EpTuple *__restrict__ tuples, const long *__restrict__ tseg,
    const double *__restrict__ blocks, const int *__restrict__ order, int npairs,
    double *__restrict__ out, long term_stride, int term0, int nterms)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    constexpr int NQ = GAMMA ? 1 : (NKR > 0 ? NKR : 1);
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const long total = (long)nterms * npairs;
    const long per = (total + kNumXcd - 1) / kNumXcd;
    // grid-stride over (term, pair) tasks: a dispatch describes at most
    // 2^32 - 1 work-items; the stride is a multiple of kNumXcd, so each
    // work-group keeps its XCD's contiguous range
    for (long g = blockIdx.x; g < per * kNumXcd; g += gridDim.x) {
    const long logical = (g % kNumXcd) * per + g / kNumXcd;
    if (logical >= total) continue;
    const int term = term0 + (int)(logical / npairs);
    const int pidx = order[logical % npairs];
    const int pt = S.term_type[term], aj = S.term_alpha[term];
    const long task = (long)pt * npairs + pidx;
    const PairRec pr = pairs[pidx];
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int ns1 = S.type_nstates[t1], ns2 = S.type_nstates[t2];
    const int MA = S.max_alphas, MS = S.max_states;
    const int nk = GAMMA ? 1 : S.nk;
    const int nent = ns1 * ns2;
    const double *B1 = S.type_basis + (size_t)t1 * MS * MA;   // [state][alpha]
    const double *B2 = S.type_basis + (size_t)t2 * MS * MA;
    double *Ql = smem;                                         // [nk][nent][2] (NKR == 0)

    int em1[NE], em2[NE], eo[NE];
    // entry state: 0 outside the block, 1 stored, 2 stored as zero (strict
    // lower triangle of a diagonal block, valeValeSaving intgSaving.F90:433)
    int est[NE];
#pragma unroll
    for (int r = 0; r < NE; ++r) {
        const int e = lane + 64 * r;
        em1[r] = e < nent ? e % ns1 : 0;
        em2[r] = e < nent ? e / ns1 : 0;
        eo[r] = e < nent ? (S.type_lm[t1 * MS + em1[r]] - 1) * SH + S.type_lm[t2 * MS + em2[r]] - 1
                         : 0;
        est[r] = e >= nent ? 0 : (pr.i == pr.j && em1[r] > em2[r]) ? 2 : 1;
    }

    const long toff = tseg[task];
    const int cnt = (int)(tseg[task + 1] - toff);
    const unsigned long long below = (1ull << aj) - 1ull;
    double P[NE];
    double2 Q[NE][NQ];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        P[e] = 0.0;
#pragma unroll
        for (int k = 0; k < NQ; ++k) Q[e][k] = make_double2(0.0, 0.0);
    }
    // each lane clears, accumulates and stores only its own entries of Ql
    if (NKR == 0 && !GAMMA)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int ee = lane + 64 * e;
            if (ee < nent)
                for (int k = 0; k < nk; ++k) {
                    Ql[2 * ((size_t)k * nent + ee)] = 0.0;
                    Ql[2 * ((size_t)k * nent + ee) + 1] = 0.0;
                }
        }
    int cur_img = -1;
    auto flush = [&]() {
        if (cur_img < 0) return;
        const double2 *ph = phase + (size_t)(pr.img_off + cur_img) * nk;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            if (GAMMA) {
                Q[e][0].x += P[e];
            } else if (NKR > 0) {
#pragma unroll
                for (int k = 0; k < NQ; ++k)
                    if (k < nk) {
                        const double2 f = ph[k];
                        Q[e][k].x += f.x * P[e];
                        Q[e][k].y += f.y * P[e];
                    }
            } else {
                const int ee = lane + 64 * e;
                if (ee < nent)
                    for (int k = 0; k < nk; ++k) {
                        const double2 f = ph[k];
                        Ql[2 * ((size_t)k * nent + ee)] += f.x * P[e];
                        Ql[2 * ((size_t)k * nent + ee) + 1] += f.y * P[e];
                    }
            }
            P[e] = 0.0;
        }
        cur_img = -1;
    };
    for (int b = 0; b < cnt; b += 64) {
        EpTuple t{};
        if (b + lane < cnt) t = tuples[toff + b + lane];
        unsigned long long m = __ballot((t.mask >> aj) & 1ull);
        while (m) {
            const int src = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            const int key = __builtin_amdgcn_readlane(t.key, src);
            const int item0 = __builtin_amdgcn_readlane(t.item, src);
            const unsigned lo = __builtin_amdgcn_readlane((unsigned)t.mask, src);
            const unsigned hi = __builtin_amdgcn_readlane((unsigned)(t.mask >> 32), src);
            const unsigned long long tm = ((unsigned long long)hi << 32) | lo;
            const int item = item0 + __popcll(tm & below);
            const int im = key >> 16, a1 = (key >> 8) & 255, a2 = key & 255;
            if (im != cur_img) {
                flush();
                cur_img = im;
            }
            const double *blk = blocks + (size_t)item * BS;
#pragma unroll
            for (int e = 0; e < NE; ++e)
                P[e] += B1[em1[e] * MA + a1] * B2[em2[e] * MA + a2] * blk[eo[e]];
        }
    }
    flush();
    // store the block: contiguous entries pr.blk + (row state + ns1 *
    // column state) of every k
    double *o = out + (size_t)(term - term0) * term_stride;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        if (est[e] == 0) continue;
        const bool zero = est[e] == 2;
        const long idx = (long)pr.blk + lane + 64 * e;
        if (GAMMA) {
            o[idx] = zero ? 0.0 : Q[e][0].x;
        } else if (NKR > 0) {
#pragma unroll
            for (int k = 0; k < NQ; ++k)
                if (k < nk)
                    reinterpret_cast<double2 *>(o)[(size_t)k * S.nbe + idx] =
                        zero ? make_double2(0.0, 0.0) : Q[e][k];
        } else {
            const int ee = lane + 64 * e;
            for (int k = 0; k < nk; ++k) {
                const double *q = Ql + 2 * ((size_t)k * nent + ee);
                reinterpret_cast<double2 *>(o)[(size_t)k * S.nbe + idx] =
                    zero ? make_double2(0.0, 0.0) : make_double2(q[0], q[1]);
            }
        }
    }
    }
}

}  // namespace imago
