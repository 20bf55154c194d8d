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
// the primitive blocks of k_ep_prims, once per item:
//
//   k_epd_hsum   one wavefront per run of at most a fixed number of a
//                pair's items (EpdSub): sum_items c'_j(item) * C1 C2 block,
//                summed per image and Bloch-phased, written as the run's
//                partial block; k_epd_hfinal adds NP, KE and the pair's
//                partials in run order.
//   k_epd_trace  one wavefront per run of one (pair, potential type): for
//                each image X_e = sum_k Re(phase_k conj(rw_k,e)) with rw the
//                weighted density of the block layout, and per item
//                sum_e C1 C2 block_e X_e added to its term; the runs' sums
//                are added per term in a fixed order by k_epd_trace_final.
// Both require at most 4 k-points (or Gamma) so Q(k) and rw(k) of a lane's
// entries stay in registers.
#pragma once
#include "ep_kernel.hpp"

namespace imago {

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

__device__ __forceinline__ void epd_tuple_fields(const EpTuple &t, int src, int &key, int &item0,
                                                 unsigned long long &mask)
{
    key = __builtin_amdgcn_readlane(t.key, src);
    item0 = __builtin_amdgcn_readlane(t.item, src);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)t.mask, src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(t.mask >> 32), src);
    mask = ((unsigned long long)hi << 32) | lo;
}

// Trace runs: each (type, pair) task's items cut into runs of at most C
// items.  The run's partial goes to slot roff[task] + r (type-major task =
// type * npairs + pair, the order k_epd_trace_final sums per term); the
// runs are stored for execution in pair-major order (roff2[pair * ntypes +
// type] + r), so consecutive work-groups share the pair's rho block in L2.
// count: rcnt[task] = rcnt2[pair * ntypes + type] = runs of the task.
template <bool FILL>
__global__ void k_epd_trace_runs(const long *__restrict__ iseg, long ntask, int npairs, long C,
                                 long *__restrict__ rcnt, long *__restrict__ rcnt2,
                                 const long *__restrict__ roff, const long *__restrict__ roff2,
                                 EpdSub *__restrict__ runs)
{
    const long task = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (task >= ntask) return;
    const long a = iseg[task], b = iseg[task + 1];
    const long n = (b - a + C - 1) / C;
    const int p = (int)(task % npairs), t = (int)(task / npairs);
    const long ntypes = ntask / npairs;
    const long pm = (long)p * ntypes + t;
    if (!FILL) {
        rcnt[task] = n;
        rcnt2[pm] = n;
        return;
    }
    for (long r = 0; r < n; ++r)
        runs[roff2[pm] + r] = EpdSub{p, t, t + 1, 0, a + r * C, min(b, a + (r + 1) * C), roff[task] + r};
}

// H sum: each pair's items over all types cut in order into runs of at
// most C items (a pair without items gets one empty run).  count:
// rcnt[p] = runs, wcnt[p] = runs * block doubles; fill: the runs with
// their partial offsets from woff.
template <bool FILL>
__global__ void k_epd_hsum_runs(const long *__restrict__ iseg, int npairs, int ntypes, long C,
                                const PairRec *__restrict__ pairs, const DevSys S,
                                long *__restrict__ rcnt, long *__restrict__ wcnt,
                                const long *__restrict__ roff, const long *__restrict__ woff,
                                EpdSub *__restrict__ runs)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const PairRec pr = pairs[p];
    const long blk = (long)S.type_nstates[S.atom_type[pr.i]] * S.type_nstates[S.atom_type[pr.j]] *
                     (S.nk) * (S.gamma ? 1 : 2);
    long nr = 0, acc = 0;
    bool open = false;
    EpdSub cur{p, 0, 0, 0, 0, 0, 0};
    const long r0 = FILL ? roff[p] : 0, w0 = FILL ? woff[p] : 0;
    for (int t = 0; t < ntypes; ++t) {
        long a = iseg[(long)t * npairs + p];
        const long b = iseg[(long)t * npairs + p + 1];
        while (a < b) {
            if (!open) {
                cur = EpdSub{p, t, t + 1, 0, a, b, w0 + nr * blk};
                open = true;
                acc = 0;
            }
            const long take = min(b - a, C - acc);
            cur.t_hi = t + 1;
            cur.i_hi = a + take;
            acc += take;
            a += take;
            if (acc == C) {
                if (FILL) runs[r0 + nr] = cur;
                ++nr;
                open = false;
            }
        }
    }
    if (open) {
        if (FILL) runs[r0 + nr] = cur;
        ++nr;
    }
    if (nr == 0) {
        if (FILL) runs[r0] = EpdSub{p, 0, 0, 0, 0, 0, w0};
        nr = 1;
    }
    if (!FILL) {
        rcnt[p] = nr;
        wcnt[p] = nr * blk;
    }
}

// Walks, in order, the items of task [t0, t1) tuples that fall in [a, b):
// f(image, a1, a2, exponent, item).  Tuples are in image order and their
// items are numbered in tuple order, exponents ascending.
template <class F>
__device__ __forceinline__ void epd_walk(const EpTuple *__restrict__ tuples, long t0, long t1,
                                         long a, long b, F &&f)
{
    const int lane = threadIdx.x & 63;
    for (long tb = t0; tb < t1; tb += 64) {
        EpTuple t{};
        if (tb + lane < t1) t = tuples[tb + lane];
        unsigned long long m = __ballot(tb + lane < t1);
        while (m) {
            const int src = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            int key, item0;
            unsigned long long mm;
            epd_tuple_fields(t, src, key, item0, mm);
            if (item0 >= b) return;
            if (item0 + __popcll(mm) <= a) continue;
            f(key >> 16, (key >> 8) & 255, key & 255, item0, mm);
        }
    }
}

// part[sub.out ...] = [k][nent] (complex, or real for Gamma): the run's
// sum of c'_j V_j(k) over its items, on its pair's block.
template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_epd_hsum(DevSys S, const PairRec *__restrict__ pairs,
                                                 const double2 *__restrict__ phase,
                                                 const EpTuple *__restrict__ tuples,
                                                 const long *__restrict__ tseg,
                                                 const long *__restrict__ iseg,
                                                 const double *__restrict__ blocks,
                                                 const double *__restrict__ coef_eff,
                                                 const int *__restrict__ type_term0,
                                                 const EpdSub *__restrict__ subs, long nsubs,
                                                 int npairs, double *__restrict__ part)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    constexpr int NQ = GAMMA ? 1 : 4;
    const int lane = threadIdx.x;
    const long sidx = blockIdx.x;
    if (sidx >= nsubs) return;
    const EpdSub sub = subs[sidx];
    const int pidx = sub.pair;
    const PairRec pr = pairs[pidx];
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int ns1 = S.type_nstates[t1], ns2 = S.type_nstates[t2];
    const int MA = S.max_alphas, MS = S.max_states;
    const int nk = GAMMA ? 1 : S.nk;
    const int nent = ns1 * ns2;
    const double *B1 = S.type_basis + (size_t)t1 * MS * MA;
    const double *B2 = S.type_basis + (size_t)t2 * MS * MA;
    EpdEntries<NE, SH> E;
    E.init(S, pr, t1, t2, ns1, nent, lane);

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
        const double2 *ph = phase + (size_t)(pr.img_off + cur_img) * nk;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            if (GAMMA) {
                Q[e][0].x += P[e];
            } else {
#pragma unroll
                for (int k = 0; k < NQ; ++k)
                    if (k < nk) {
                        const double2 f = ph[k];
                        Q[e][k].x += f.x * P[e];
                        Q[e][k].y += f.y * P[e];
                    }
            }
            P[e] = 0.0;
        }
        cur_img = -1;
    };

    for (int pt = sub.t_lo; pt < sub.t_hi; ++pt) {
        const long task = (long)pt * npairs + pidx;
        long a = iseg[task], b = iseg[task + 1];
        if (pt == sub.t_lo && sub.i_lo > a) a = sub.i_lo;
        if (pt == sub.t_hi - 1 && sub.i_hi < b) b = sub.i_hi;
        if (a >= b) continue;
        const double *cf = coef_eff + type_term0[pt];
        flush();
        epd_walk(tuples, tseg[task], tseg[task + 1], a, b,
                 [&](int im, int a1, int a2, int item0, unsigned long long mm) {
                     if (im != cur_img) {
                         flush();
                         cur_img = im;
                     }
                     double cc[NE];
#pragma unroll
                     for (int e = 0; e < NE; ++e)
                         cc[e] = B1[E.em1[e] * MA + a1] * B2[E.em2[e] * MA + a2];
                     // items of this tuple inside [a, b), two at a time so
                     // their block loads overlap
                     long it = item0;
                     while (mm) {
                         const int ja = __ffsll((long long)mm) - 1;
                         mm &= mm - 1ull;
                         const long ia = it++;
                         if (ia < a) continue;
                         if (ia >= b) break;
                         const bool two = mm != 0ull && it < b;
                         const int jb = two ? __ffsll((long long)mm) - 1 : ja;
                         if (two) {
                             mm &= mm - 1ull;
                             ++it;
                         }
                         const double *ba = blocks + (size_t)ia * BS;
                         const double *bb = blocks + (size_t)(ia + (two ? 1 : 0)) * BS;
                         const double ca = cf[ja], cb = two ? cf[jb] : 0.0;
                         double va[NE], vb[NE];
#pragma unroll
                         for (int e = 0; e < NE; ++e) {
                             va[e] = ba[E.eo[e]];
                             vb[e] = bb[E.eo[e]];
                         }
#pragma unroll
                         for (int e = 0; e < NE; ++e) P[e] += cc[e] * (ca * va[e] + cb * vb[e]);
                     }
                 });
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
                if (k < nk)
                    reinterpret_cast<double2 *>(o)[(size_t)k * nent + ee] =
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

// wave sum of a double (every lane gets the total)
__device__ __forceinline__ double epd_wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// tsub[sub.out][a] = sum over the run's items of exponent a of
// sum_e C1 C2 block_e X_e(image), X from the weighted density rw
// (runs of one potential type).
template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_epd_trace(DevSys S, const PairRec *__restrict__ pairs,
                                                  const double2 *__restrict__ phase,
                                                  const EpTuple *__restrict__ tuples,
                                                  const long *__restrict__ tseg,
                                                  const long *__restrict__ iseg,
                                                  const double *__restrict__ blocks,
                                                  const double *__restrict__ rw,
                                                  const EpdSub *__restrict__ subs, long nsubs,
                                                  int npairs, double *__restrict__ tsub)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    constexpr int NQ = GAMMA ? 1 : 4;
    extern __shared__ double smem[];   // [max_pot_alphas][64] per-lane term sums
    const int lane = threadIdx.x;
    const long sidx = blockIdx.x;
    if (sidx >= nsubs) return;
    const EpdSub sub = subs[sidx];
    const int pidx = sub.pair, pt = sub.t_lo;
    const PairRec pr = pairs[pidx];
    const int t1 = S.atom_type[pr.i], t2 = S.atom_type[pr.j];
    const int ns1 = S.type_nstates[t1], ns2 = S.type_nstates[t2];
    const int MA = S.max_alphas, MS = S.max_states;
    const int nk = GAMMA ? 1 : S.nk;
    const int nent = ns1 * ns2;
    const int na3 = S.pot_nalpha_of(pt);
    const double *B1 = S.type_basis + (size_t)t1 * MS * MA;
    const double *B2 = S.type_basis + (size_t)t2 * MS * MA;
    EpdEntries<NE, SH> E;
    E.init(S, pr, t1, t2, ns1, nent, lane);
    double *T = smem;
    for (int a = 0; a < na3; ++a) T[a * 64 + lane] = 0.0;

    // the pair's weighted density, every k
    double2 R[NE][NQ];
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            R[e][k] = make_double2(0.0, 0.0);
            if (E.est[e] != 0 && k < nk) {
                const size_t o = (size_t)k * S.nbe + pr.blk + lane + 64 * e;
                R[e][k] = GAMMA ? make_double2(rw[o], 0.0) : reinterpret_cast<const double2 *>(rw)[o];
            }
        }
    double X[NE];
    int cur_img = -1;

    const long task = (long)pt * npairs + pidx;
    const long a = sub.i_lo, b = sub.i_hi;
    epd_walk(tuples, tseg[task], tseg[task + 1], a, b,
             [&](int im, int a1, int a2, int item0, unsigned long long mm) {
                 if (im != cur_img) {
                     cur_img = im;
                     const double2 *ph = phase + (size_t)(pr.img_off + im) * nk;
#pragma unroll
                     for (int e = 0; e < NE; ++e) {
                         double x = 0.0;
#pragma unroll
                         for (int k = 0; k < NQ; ++k)
                             if (k < nk) {
                                 const double2 f = GAMMA ? make_double2(1.0, 0.0) : ph[k];
                                 x += f.x * R[e][k].x + f.y * R[e][k].y;
                             }
                         X[e] = x;
                     }
                 }
                 double cx[NE];
#pragma unroll
                 for (int e = 0; e < NE; ++e)
                     cx[e] = B1[E.em1[e] * MA + a1] * B2[E.em2[e] * MA + a2] * X[e];
                 long it = item0;
                 while (mm) {
                     const int ja = __ffsll((long long)mm) - 1;
                     mm &= mm - 1ull;
                     const long ia = it++;
                     if (ia < a) continue;
                     if (ia >= b) break;
                     const bool two = mm != 0ull && it < b;
                     const int jb = two ? __ffsll((long long)mm) - 1 : ja;
                     if (two) {
                         mm &= mm - 1ull;
                         ++it;
                     }
                     const double *ba = blocks + (size_t)ia * BS;
                     const double *bb = blocks + (size_t)(ia + (two ? 1 : 0)) * BS;
                     double sa = 0.0, sb = 0.0;
#pragma unroll
                     for (int e = 0; e < NE; ++e) {
                         sa += cx[e] * ba[E.eo[e]];
                         sb += cx[e] * bb[E.eo[e]];
                     }
                     T[ja * 64 + lane] += sa;
                     if (two) T[jb * 64 + lane] += sb;
                 }
             });
    for (int x = 0; x < na3; ++x) {
        const double v = epd_wave_sum(T[x * 64 + lane]);
        if (lane == 0) tsub[(size_t)sub.out * S.max_pot_alphas + x] = v;
    }
}

// out[term] = sum over the runs of the term's potential type (runs
// roff[type * npairs] .. roff[(type + 1) * npairs]) of their sums, a
// fixed-order tree
__global__ __launch_bounds__(256) void k_epd_trace_final(const double *__restrict__ tsub,
                                                         const long *__restrict__ roff, int npairs,
                                                         const int *__restrict__ term_type,
                                                         const int *__restrict__ term_alpha,
                                                         int max_pot_alphas,
                                                         double *__restrict__ out)
{
    __shared__ double red[256];
    const int term = blockIdx.x;
    const int pt = term_type[term], aj = term_alpha[term];
    const long s0 = roff[(long)pt * npairs], s1 = roff[(long)(pt + 1) * npairs];
    double s = 0.0;
    for (long q = s0 + threadIdx.x; q < s1; q += 256) s += tsub[(size_t)q * max_pot_alphas + aj];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[term] = red[0];
}

}  // namespace imago
