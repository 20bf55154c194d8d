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
//   k_epd_hsum   one wavefront per (pair, group of potential types):
//                sum_items c'_j(item) * C1 C2 block, summed per image and
//                Bloch-phased, written as the group's partial block;
//                k_epd_hfinal adds NP, KE and the groups' partials.
//   k_epd_trace  one wavefront per (pair, potential type): for each image
//                X_e = sum_k Re(phase_k conj(rw_k,e)) with rw the weighted
//                density of the block layout, and per item
//                sum_e C1 C2 block_e X_e added to its term; per-(pair, term)
//                partials are summed over pairs in a fixed order by
//                k_epd_trace_final.
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

// partial[grp][k][nbe] (complex, or real for Gamma) of one pair's block:
// sum over the group's potential terms of c'_j V_j(k).
template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_epd_hsum(DevSys S, const PairRec *__restrict__ pairs,
                                                 const double2 *__restrict__ phase,
                                                 const EpTuple *__restrict__ tuples,
                                                 const long *__restrict__ tseg,
                                                 const double *__restrict__ blocks,
                                                 const double *__restrict__ coef_eff,
                                                 const int *__restrict__ type_term0,
                                                 const int *__restrict__ order, int npairs,
                                                 int ngroups, double *__restrict__ partial,
                                                 long group_stride)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    constexpr int NQ = GAMMA ? 1 : 4;
    const int lane = threadIdx.x;
    const long logical = blockIdx.x;
    if (logical >= (long)npairs * ngroups) return;
    const int grp = (int)(logical % ngroups);
    const int pidx = order[logical / ngroups];
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

    const int pt_lo = grp * kEpdTypesPerGroup;
    const int pt_hi = min(S.num_pot_types, pt_lo + kEpdTypesPerGroup);
    for (int pt = pt_lo; pt < pt_hi; ++pt) {
        const long task = (long)pt * npairs + pidx;
        const long toff = tseg[task];
        const int cnt = (int)(tseg[task + 1] - toff);
        const double *cf = coef_eff + type_term0[pt];
        flush();
        for (int b = 0; b < cnt; b += 64) {
            EpTuple t{};
            if (b + lane < cnt) t = tuples[toff + b + lane];
            unsigned long long m = __ballot(b + lane < cnt);
            while (m) {
                const int src = __ffsll((long long)m) - 1;
                m &= m - 1ull;
                int key, item;
                unsigned long long mm;
                epd_tuple_fields(t, src, key, item, mm);
                const int im = key >> 16, a1 = (key >> 8) & 255, a2 = key & 255;
                if (im != cur_img) {
                    flush();
                    cur_img = im;
                }
                double cc[NE];
#pragma unroll
                for (int e = 0; e < NE; ++e) cc[e] = B1[E.em1[e] * MA + a1] * B2[E.em2[e] * MA + a2];
                // two items at a time so their block loads overlap
                while (mm) {
                    const int ja = __ffsll((long long)mm) - 1;
                    mm &= mm - 1ull;
                    const bool two = mm != 0ull;
                    const int jb = two ? __ffsll((long long)mm) - 1 : ja;
                    if (two) mm &= mm - 1ull;
                    const double *ba = blocks + (size_t)item * BS;
                    const double *bb = blocks + (size_t)(item + (two ? 1 : 0)) * BS;
                    const double ca = cf[ja], cb = two ? cf[jb] : 0.0;
                    double va[NE], vb[NE];
#pragma unroll
                    for (int e = 0; e < NE; ++e) {
                        va[e] = ba[E.eo[e]];
                        vb[e] = bb[E.eo[e]];
                    }
#pragma unroll
                    for (int e = 0; e < NE; ++e) P[e] += cc[e] * (ca * va[e] + cb * vb[e]);
                    item += two ? 2 : 1;
                }
            }
        }
    }
    flush();
    double *o = partial + (size_t)grp * group_stride;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        if (E.est[e] == 0) continue;
        const bool zero = E.est[e] == 2;
        const long idx = (long)pr.blk + lane + 64 * e;
        if (GAMMA) {
            o[idx] = zero ? 0.0 : Q[e][0].x;
        } else {
#pragma unroll
            for (int k = 0; k < NQ; ++k)
                if (k < nk)
                    reinterpret_cast<double2 *>(o)[(size_t)k * S.nbe + idx] =
                        zero ? make_double2(0.0, 0.0) : Q[e][k];
        }
    }
}

// H = NP + KE + sum of the groups' partials, elementwise over the block layout
__global__ __launch_bounds__(256) void k_epd_hfinal(const double *__restrict__ np,
                                                    const double *__restrict__ ke,
                                                    const double *__restrict__ partial, int ngroups,
                                                    long len, double *__restrict__ H)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    double a = np[e] + ke[e];
    for (int g = 0; g < ngroups; ++g) a += partial[(size_t)g * len + e];
    H[e] = a;
}

// wave sum of a double (every lane gets the total)
__device__ __forceinline__ double epd_wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// tpart[pair][term] = sum over the items of (pair, term) of
// sum_e C1 C2 block_e X_e(image), X from the weighted density rw.
template <int LMAX, int NE, bool GAMMA>
__global__ __launch_bounds__(64) void k_epd_trace(DevSys S, const PairRec *__restrict__ pairs,
                                                  const double2 *__restrict__ phase,
                                                  const EpTuple *__restrict__ tuples,
                                                  const long *__restrict__ tseg,
                                                  const double *__restrict__ blocks,
                                                  const double *__restrict__ rw,
                                                  const int *__restrict__ type_term0,
                                                  const int *__restrict__ order, int npairs,
                                                  double *__restrict__ tpart)
{
    constexpr int SH = (LMAX + 1) * (LMAX + 1);
    constexpr int BS = SH * SH;
    constexpr int NQ = GAMMA ? 1 : 4;
    extern __shared__ double smem[];   // [max_pot_alphas][64] per-lane term sums
    const int lane = threadIdx.x;
    const long logical = blockIdx.x;
    if (logical >= (long)npairs * S.num_pot_types) return;
    const int pt = (int)(logical % S.num_pot_types);
    const int pidx = order[logical / S.num_pot_types];
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
    const long toff = tseg[task];
    const int cnt = (int)(tseg[task + 1] - toff);
    for (int b = 0; b < cnt; b += 64) {
        EpTuple t{};
        if (b + lane < cnt) t = tuples[toff + b + lane];
        unsigned long long m = __ballot(b + lane < cnt);
        while (m) {
            const int src = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            int key, item;
            unsigned long long mm;
            epd_tuple_fields(t, src, key, item, mm);
            const int im = key >> 16, a1 = (key >> 8) & 255, a2 = key & 255;
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
            while (mm) {
                const int ja = __ffsll((long long)mm) - 1;
                mm &= mm - 1ull;
                const bool two = mm != 0ull;
                const int jb = two ? __ffsll((long long)mm) - 1 : ja;
                if (two) mm &= mm - 1ull;
                const double *ba = blocks + (size_t)item * BS;
                const double *bb = blocks + (size_t)(item + (two ? 1 : 0)) * BS;
                double sa = 0.0, sb = 0.0;
#pragma unroll
                for (int e = 0; e < NE; ++e) {
                    sa += cx[e] * ba[E.eo[e]];
                    sb += cx[e] * bb[E.eo[e]];
                }
                T[ja * 64 + lane] += sa;
                if (two) T[jb * 64 + lane] += sb;
                item += two ? 2 : 1;
            }
        }
    }
    const int term0 = type_term0[pt];
    for (int a = 0; a < na3; ++a) {
        const double v = epd_wave_sum(T[a * 64 + lane]);
        if (lane == 0) tpart[(size_t)pidx * S.pot_dim + term0 + a] = v;
    }
}

// out[term] = sum over pairs of tpart[pair][term], a fixed-order tree per term
__global__ __launch_bounds__(256) void k_epd_trace_final(const double *__restrict__ tpart,
                                                         int npairs, int pot_dim,
                                                         double *__restrict__ out)
{
    __shared__ double red[256];
    const int term = blockIdx.x;
    double s = 0.0;
    for (int p = threadIdx.x; p < npairs; p += 256) s += tpart[(size_t)p * pot_dim + term];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[term] = red[0];
}

}  // namespace imago
