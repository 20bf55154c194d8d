// ep_direct.hpp - electronic-potential contributions rebuilt from the
// primitive blocks each SCF iteration, without the potDim panels.
//
// The reference stores every potential term V_j(k) and, each iteration,
// forms H(k) = NP + KE + sum_j c_j V_j (readPackedMatrixAccum,
// secularEqn.F90:148-171; a factor of exactly 0.0 adds the term unscaled,
// matrixSubs.F90:369-373) and the traces Tr(rho V_j) (matrixElementMult,
// matrixSubs.F90:66-112, in makeValenceRho valeCharge.F90:230-393).  Its
// post-SCF path folds the coefficients into one accumulation instead
// (integrals.F90:2860-3605).  Here both per-iteration quantities come from
// the primitive blocks of k_ep_prims, once per item ((tuple, exponent)
// pair of one potential type and atom pair):
//
//   set-up (once per system)  every pair's items sorted by lattice image
//                (stable, so types stay in order within an image) into
//                pitem, cut into runs of at most kEpdRunItems (EpdSub).
//   k_epd_hsum   one wavefront per run: sum over its items of c'_j C1 C2
//                block, summed per image and Bloch-phased, written as the
//                run's partial block; k_epd_hfinal adds NP, KE and the
//                pair's partials in run order.
//   k_epd_fold   (k-points) X(image)_e = sum_k Re(conj(phase_k) rw_k,e), the
//                weighted density folded onto each lattice image.
//   k_epd_trace  one wavefront per run: per item sum_e C1 C2 block_e X_e,
//                one value per item; k_epd_trace_final adds them per term
//                in item order.
// A wavefront decodes 64 items at once (item -> tuple -> image, exponents,
// type) and then streams their blocks a few items at a time, so the block
// loads of several items are in flight together.  The hsum accumulators of
// 4 k-points (or Gamma) stay in registers; more k-points take one hsum
// launch per group of four.
#pragma once
#include "ep_kernel.hpp"
#include "lane_xor.hpp"

namespace imago {

constexpr int kEpdGroup = 4;   // items whose block loads are issued together (NE <= 4)

// per-lane entry tables of one pair block: em1/em2 row/column states, eo
// the offset in a primitive block, est 0 outside / 1 stored / 2 stored zero
template <int NE, int SH>
struct EpdEntries {
    int em1[NE], em2[NE], eo[NE], est[NE];
    __device__ __forceinline__ void init(const DevSys &S, const PairRec &pr, int t1, int t2,
                                         int ns1, int nent, int lane)
    {
        const int MS = S.max_states;
#pragma unroll
        for (int r = 0; r < NE; ++r) {
            const int e = lane + 64 * r;
            em1[r] = e < nent ? e % ns1 : 0;
            em2[r] = e < nent ? e / ns1 : 0;
            eo[r] = e < nent ? (S.type_lm[t1 * MS + em1[r]] - 1) * SH + S.type_lm[t2 * MS + em2[r]] - 1
                             : 0;
            est[r] = e >= nent ? 0 : (pr.i == pr.j && em1[r] > em2[r]) ? 2 : 1;
        }
    }
};

// The pair's two contraction tables in LDS: B1[state][alpha] of atom i then
// B2 of atom j (2 * max_states * max_alphas doubles per wavefront).
__device__ __forceinline__ void epd_stage_basis(const DevSys &S, int t1, int t2, double *lds,
                                                int lane)
{
    const int n = S.max_states * S.max_alphas;
    const double *B1 = S.type_basis + (size_t)t1 * n, *B2 = S.type_basis + (size_t)t2 * n;
    for (int x = lane; x < n; x += 64) {
        lds[x] = B1[x];
        lds[n + x] = B2[x];
    }
    __syncthreads();
}

// Decode of item pitem[q] for this lane: the item, its lattice image, its
// exponent pair key (a1 << 8 | a2), its potential exponent j and type.
struct EpdItem {
    int item, img, a12, j, type;
};
__device__ __forceinline__ EpdItem epd_decode(const EpTuple *__restrict__ tuples,
                                              const int *__restrict__ item_tuple,
                                              const int *__restrict__ pitem, long q, int npairs)
{
    EpdItem d;
    d.item = pitem[q];
    const EpTuple t = tuples[item_tuple[d.item]];
    unsigned long long m = t.mask;
    for (int r = d.item - t.item; r > 0; --r) m &= m - 1ull;
    d.j = __ffsll((long long)m) - 1;
    d.img = t.key >> 16;
    d.a12 = t.key & 0xffff;
    d.type = t.task / npairs;
    return d;
}

__device__ __forceinline__ double epd_readlane(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// wave sum of a double (every lane gets the total)
__device__ __forceinline__ double epd_wave_sum(double v) { return wave_sum_xor(v); }

// ---------------------------------------------------------------------------
// set-up: image keys, pair ranges, runs

// key[item] = global lattice image of the item (pairs[p].img_off + image),
// val[item] = item, jexp[item] = its potential exponent.
__global__ void k_epd_item_keys(const EpTuple *__restrict__ tuples,
                                const int *__restrict__ item_tuple,
                                const PairRec *__restrict__ pairs, long nitems, int npairs,
                                int *__restrict__ key, int *__restrict__ val,
                                unsigned char *__restrict__ jexp)
{
    const long it = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nitems) return;
    const EpTuple t = tuples[item_tuple[it]];
    unsigned long long m = t.mask;
    for (int r = (int)(it - t.item); r > 0; --r) m &= m - 1ull;
    key[it] = pairs[t.task % npairs].img_off + (t.key >> 16);
    val[it] = (int)it;
    jexp[it] = (unsigned char)(__ffsll((long long)m) - 1);
}

__device__ __forceinline__ long epd_lower_bound(const int *__restrict__ a, long n, int x)
{
    long lo = 0, hi = n;
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Runs of every pair's image-sorted items [lo, hi) of pitem, C items each.
// H sum: at least one run per pair (an empty one for a pair without items),
// partial blocks of blk doubles from woff[p]; traces: ceil(n / C) runs.
// count: hcnt[p], wcnt[p] = hcnt[p] * blk, tcnt[p]; fill: the runs.
template <bool FILL>
__global__ void k_epd_pair_runs(const int *__restrict__ skey, long nitems,
                                const PairRec *__restrict__ pairs, int npairs, const DevSys S,
                                long C, long *__restrict__ hcnt, long *__restrict__ wcnt,
                                long *__restrict__ tcnt, const long *__restrict__ hoff,
                                const long *__restrict__ woff, const long *__restrict__ toff,
                                EpdSub *__restrict__ hruns, EpdSub *__restrict__ truns)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const PairRec pr = pairs[p];
    const long lo = epd_lower_bound(skey, nitems, pr.img_off);
    const long hi = epd_lower_bound(skey, nitems, pr.img_off + pr.img_cnt);
    const long blk = (long)S.type_nstates[S.atom_type[pr.i]] * S.type_nstates[S.atom_type[pr.j]] *
                     S.nk * (S.gamma ? 1 : 2);
    const long nt = (hi - lo + C - 1) / C, nh = nt > 0 ? nt : 1;
    if (!FILL) {
        hcnt[p] = nh;
        wcnt[p] = nh * blk;
        tcnt[p] = nt;
        return;
    }
    for (long r = 0; r < nh; ++r) {
        const long a = lo + r * C, b = min(hi, a + C);
        hruns[hoff[p] + r] = EpdSub{p, 0, 0, 0, a, a < b ? b : a, woff[p] + r * blk};
        if (r < nt) truns[toff[p] + r] = EpdSub{p, 0, 0, 0, a, b, 0};
    }
}

// ---------------------------------------------------------------------------
// H sum

// part[sub.out ...] = [k][nent] (complex, or real for Gamma): the run's
// sum of c'_j V_j(k) over its items, on its pair's block, for the k-points
// k0 .. k0 + 3 (one launch per group of four).
template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_epd_hsum(DevSys S, const PairRec *__restrict__ pairs,
                                                 const double2 *__restrict__ phase,
                                                 const EpTuple *__restrict__ tuples,
                                                 const int *__restrict__ item_tuple,
                                                 const int *__restrict__ pitem,
                                                 const double *__restrict__ blocks,
                                                 const double *__restrict__ coef_eff,
                                                 const int *__restrict__ type_term0,
                                                 const EpdSub *__restrict__ subs, long nsubs,
                                                 int npairs, double *__restrict__ part, int k0)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    constexpr int NQ = GAMMA ? 1 : 4;
    constexpr int G = NE <= 4 ? kEpdGroup : 1;   // large blocks: one item at a time
    extern __shared__ double bas[];   // B1, B2 of the pair
    const int lane = threadIdx.x;
    const long sidx = blockIdx.x;
    if (sidx >= nsubs) return;
    const EpdSub sub = subs[sidx];
    const PairRec pr = pairs[sub.pair];
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int ns1 = S.type_nstates[t1], ns2 = S.type_nstates[t2];
    const int MA = S.max_alphas, nb2 = S.max_states * MA;
    const int nk = GAMMA ? 1 : S.nk;
    const int nent = ns1 * ns2;
    EpdEntries<NE, SH> E;
    E.init(S, pr, t1, t2, ns1, nent, lane);
    epd_stage_basis(S, t1, t2, bas, lane);

    double P[NE];
    double2 Q[NE][NQ];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        P[e] = 0.0;
#pragma unroll
        for (int k = 0; k < NQ; ++k) Q[e][k] = make_double2(0.0, 0.0);
    }
    int cur_img = -1;
    auto flush = [&]() {
        if (cur_img < 0) return;
        const double2 *ph = phase + (size_t)(pr.img_off + cur_img) * nk + k0;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            if (GAMMA) {
                Q[e][0].x += P[e];
            } else {
#pragma unroll
                for (int k = 0; k < NQ; ++k)
                    if (k0 + k < nk) {
                        const double2 f = ph[k];
                        Q[e][k].x += f.x * P[e];
                        Q[e][k].y += f.y * P[e];
                    }
            }
            P[e] = 0.0;
        }
    };

    for (long q0 = sub.i_lo; q0 < sub.i_hi; q0 += 64) {
        const int nb = (int)min(64L, sub.i_hi - q0);
        EpdItem d{0, 0, 0, 0, 0};
        double cf = 0.0;
        if (lane < nb) {
            d = epd_decode(tuples, item_tuple, pitem, q0 + lane, npairs);
            cf = coef_eff[type_term0[d.type] + d.j];
        }
        for (int u = 0; u < nb;) {
            int it[G], im[G], a12[G];
            double c[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int src = min(u + g, nb - 1);
                it[g] = __builtin_amdgcn_readlane(d.item, src);
                im[g] = __builtin_amdgcn_readlane(d.img, src);
                a12[g] = __builtin_amdgcn_readlane(d.a12, src);
                c[g] = epd_readlane(cf, src);
            }
            // items are image-ordered: this group is the prefix of m items
            // on the first item's image
            int m = 1;
#pragma unroll
            for (int g = 1; g < G; ++g)
                if (u + g < nb && im[g] == im[0]) m = g + 1;
            if (im[0] != cur_img) {
                flush();
                cur_img = im[0];
            }
            double v[G][NE];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const double *bl = blocks + (size_t)it[g] * BS;
#pragma unroll
                for (int e = 0; e < NE; ++e) v[g][e] = bl[E.eo[e]];
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int a1 = a12[g] >> 8, a2 = a12[g] & 255;
                const double cg = g < m ? c[g] : 0.0;
#pragma unroll
                for (int e = 0; e < NE; ++e)
                    P[e] += cg * (bas[E.em1[e] * MA + a1] * bas[nb2 + E.em2[e] * MA + a2]) *
                            v[g][e];
            }
            u += m;
        }
    }
    flush();
    double *o = part + sub.out;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        if (E.est[e] == 0) continue;
        const bool zero = E.est[e] == 2;
        const int ee = lane + 64 * e;
        if (GAMMA) {
            o[ee] = zero ? 0.0 : Q[e][0].x;
        } else {
#pragma unroll
            for (int k = 0; k < NQ; ++k)
                if (k0 + k < nk)
                    reinterpret_cast<double2 *>(o)[(size_t)(k0 + k) * nent + ee] =
                        zero ? make_double2(0.0, 0.0) : Q[e][k];
        }
    }
}

// H = NP + KE + the pair's run partials, one work-group per pair; the runs
// of pair p are psub[p] .. psub[p + 1] and their partial blocks lie back to
// back from woff[p], added in run order.
template <bool GAMMA>
__global__ __launch_bounds__(256) void k_epd_hfinal(const DevSys S, const PairRec *__restrict__ pairs,
                                                    const double *__restrict__ np,
                                                    const double *__restrict__ ke,
                                                    const double *__restrict__ part,
                                                    const long *__restrict__ psub,
                                                    const long *__restrict__ woff,
                                                    double *__restrict__ H)
{
    const int p = blockIdx.x;
    const PairRec pr = pairs[p];
    const int ns1 = S.type_nstates[S.atom_type[pr.i]], ns2 = S.type_nstates[S.atom_type[pr.j]];
    const int nk = GAMMA ? 1 : S.nk, dz = GAMMA ? 1 : 2;
    const int blk = ns1 * ns2 * nk * dz;
    const int nr = (int)(psub[p + 1] - psub[p]);
    const double *src = part + woff[p];
    for (int idx = threadIdx.x; idx < blk; idx += blockDim.x) {
        const int k = idx / (ns1 * ns2 * dz), r = idx % (ns1 * ns2 * dz);
        const size_t o = ((size_t)k * S.nbe + pr.blk) * dz + r;
        double a = np[o] + ke[o];
#pragma unroll 4
        for (int q = 0; q < nr; ++q) a += src[(size_t)q * blk + idx];
        H[o] = a;
    }
}

// ---------------------------------------------------------------------------
// traces

// X[(img_off + image) * ent + e] = sum_k Re(conj(phase_k(image)) rw_k,e) for
// every image of the pair (one work-group per pair); rw is the weighted
// density in the block layout, [k][nbe] complex.  Four k-points at a time
// are held in registers, the images summed for each group in turn.
__global__ __launch_bounds__(256) void k_epd_fold(const DevSys S, const PairRec *__restrict__ pairs,
                                                  const double2 *__restrict__ phase,
                                                  const double2 *__restrict__ rw, int ent,
                                                  double *__restrict__ X)
{
    const PairRec pr = pairs[blockIdx.x];
    const int nent = S.type_nstates[S.atom_type[pr.i]] * S.type_nstates[S.atom_type[pr.j]];
    const int nk = S.nk;
    for (int e = threadIdx.x; e < nent; e += blockDim.x) {
        for (int k0 = 0; k0 < nk; k0 += 4) {
            double2 R[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                R[k] = k0 + k < nk ? rw[(size_t)(k0 + k) * S.nbe + pr.blk + e]
                                   : make_double2(0.0, 0.0);
            for (int im = 0; im < pr.img_cnt; ++im) {
                const double2 *ph = phase + (size_t)(pr.img_off + im) * nk + k0;
                double x = k0 == 0 ? 0.0 : X[(size_t)(pr.img_off + im) * ent + e];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (k0 + k < nk) x += ph[k].x * R[k].x + ph[k].y * R[k].y;
                X[(size_t)(pr.img_off + im) * ent + e] = x;
            }
        }
    }
}

// ival[item] = sum_e C1 C2 block_e X_e(image of the item) for every item of
// the run; X from k_epd_fold ([image][ent]), or for Gamma the weighted
// density rw itself (block layout, the same for every image).
template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_epd_trace(DevSys S, const PairRec *__restrict__ pairs,
                                                  const EpTuple *__restrict__ tuples,
                                                  const int *__restrict__ item_tuple,
                                                  const int *__restrict__ pitem,
                                                  const double *__restrict__ blocks,
                                                  const double *__restrict__ X, int ent,
                                                  const EpdSub *__restrict__ subs, long nsubs,
                                                  int npairs, double *__restrict__ ival)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    constexpr int G = NE <= 4 ? kEpdGroup : 1;   // large blocks: one item at a time
    extern __shared__ double bas[];
    const int lane = threadIdx.x;
    const long sidx = blockIdx.x;
    if (sidx >= nsubs) return;
    const EpdSub sub = subs[sidx];
    const PairRec pr = pairs[sub.pair];
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int ns1 = S.type_nstates[t1], ns2 = S.type_nstates[t2];
    const int MA = S.max_alphas, nb2 = S.max_states * MA;
    const int nent = ns1 * ns2;
    EpdEntries<NE, SH> E;
    E.init(S, pr, t1, t2, ns1, nent, lane);
    epd_stage_basis(S, t1, t2, bas, lane);
    // the density of the current image (for Gamma the same for all images)
    double Xc[NE];
    int cur_img = -1;
    if (GAMMA) {
#pragma unroll
        for (int e = 0; e < NE; ++e) Xc[e] = E.est[e] != 0 ? X[pr.blk + lane + 64 * e] : 0.0;
    }

    for (long q0 = sub.i_lo; q0 < sub.i_hi; q0 += 64) {
        const int nb = (int)min(64L, sub.i_hi - q0);
        EpdItem d{0, 0, 0, 0, 0};
        if (lane < nb) d = epd_decode(tuples, item_tuple, pitem, q0 + lane, npairs);
        for (int u = 0; u < nb;) {
            int it[G], im[G], a12[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int src = min(u + g, nb - 1);
                it[g] = __builtin_amdgcn_readlane(d.item, src);
                im[g] = __builtin_amdgcn_readlane(d.img, src);
                a12[g] = __builtin_amdgcn_readlane(d.a12, src);
            }
            // items are image-ordered: this group is the prefix of m items
            // on the first item's image
            int m = 1;
#pragma unroll
            for (int g = 1; g < G; ++g)
                if (u + g < nb && im[g] == im[0]) m = g + 1;
            if (!GAMMA && im[0] != cur_img) {
                cur_img = im[0];
                const double *xi = X + (size_t)(pr.img_off + cur_img) * ent;
#pragma unroll
                for (int e = 0; e < NE; ++e) Xc[e] = E.est[e] != 0 ? xi[lane + 64 * e] : 0.0;
            }
            double v[G][NE];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const double *bl = blocks + (size_t)it[g] * BS;
#pragma unroll
                for (int e = 0; e < NE; ++e) v[g][e] = bl[E.eo[e]];
            }
            double s[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int a1 = a12[g] >> 8, a2 = a12[g] & 255;
                s[g] = 0.0;
#pragma unroll
                for (int e = 0; e < NE; ++e)
                    s[g] += bas[E.em1[e] * MA + a1] * bas[nb2 + E.em2[e] * MA + a2] * Xc[e] *
                            v[g][e];
            }
            double tot;
            int slot;
            if (G == 4) {
                // four wave sums at once: halve the values per lane at the
                // xor-32 and xor-16 steps, then sum within 16 lanes; lane
                // 16 g ends with item g's total
                const bool hi32 = lane >= 32, hi16 = (lane & 16) != 0;
                const double x0 = xor_lane<32>(hi32 ? s[0] : s[2 % G]);
                const double x1 = xor_lane<32>(hi32 ? s[1 % G] : s[3 % G]);
                const double a0 = (hi32 ? s[2 % G] : s[0]) + x0;
                const double a1v = (hi32 ? s[3 % G] : s[1 % G]) + x1;
                const double y = xor_lane<16>(hi16 ? a0 : a1v);
                tot = sum16_xor((hi16 ? a1v : a0) + y);
                slot = (lane >> 4) & 3;
            } else {
                tot = epd_wave_sum(s[0]);
                slot = 0;
            }
            const bool writer = G == 4 ? (lane & 15) == 0 : lane == 0;
            if (writer && slot < m) {
                int iv = it[0];
#pragma unroll
                for (int g = 1; g < G; ++g)
                    if (slot == g) iv = it[g];
                ival[iv] = tot;
            }
            u += m;
        }
    }
}

// Sum of ival per term in two fixed-order passes.  Pass 1: work-group
// (type, chunk) of 64 lanes takes chunk c of kEpdTraceChunks equal parts of
// the type's items (iseg[type * npairs] .. iseg[(type + 1) * npairs]), each
// lane a fixed stride, T[j][lane] in LDS, then a wave sum per exponent j:
// part[(type * chunks + c) * max_pot_alphas + j].  Pass 2 (one thread per
// term) adds the chunks in order: out[type_term0[type] + j].
__global__ __launch_bounds__(64) void k_epd_trace_chunks(const double *__restrict__ ival,
                                                         const unsigned char *__restrict__ jexp,
                                                         const long *__restrict__ iseg,
                                                         int npairs,
                                                         const int *__restrict__ pot_nalpha,
                                                         int max_pot_alphas,
                                                         double *__restrict__ part)
{
    extern __shared__ double T[];   // [na][64]
    const int type = blockIdx.x / kEpdTraceChunks, ch = blockIdx.x % kEpdTraceChunks;
    const int lane = threadIdx.x;
    const int na = pot_nalpha[type];
    for (int j = 0; j < na; ++j) T[j * 64 + lane] = 0.0;
    const long a0 = iseg[(long)type * npairs], b0 = iseg[(long)(type + 1) * npairs];
    const long len = (b0 - a0 + kEpdTraceChunks - 1) / kEpdTraceChunks;
    const long a = a0 + ch * len, b = min(b0, a + len);
    for (long it = a + lane; it < b; it += 64) T[jexp[it] * 64 + lane] += ival[it];
    double *o = part + ((size_t)type * kEpdTraceChunks + ch) * max_pot_alphas;
    for (int j = 0; j < na; ++j) {
        const double v = epd_wave_sum(T[j * 64 + lane]);
        if (lane == 0) o[j] = v;
    }
}

__global__ void k_epd_trace_final(const double *__restrict__ part,
                                  const int *__restrict__ term_type,
                                  const int *__restrict__ term_alpha, int pot_dim,
                                  int max_pot_alphas, double *__restrict__ out)
{
    const int term = blockIdx.x * blockDim.x + threadIdx.x;
    if (term >= pot_dim) return;
    const int type = term_type[term], j = term_alpha[term];
    const double *p = part + (size_t)type * kEpdTraceChunks * max_pot_alphas + j;
    double s = 0.0;
    for (int c = 0; c < kEpdTraceChunks; ++c) s += p[(size_t)c * max_pot_alphas];
    out[term] = s;
}

}  // namespace imago
