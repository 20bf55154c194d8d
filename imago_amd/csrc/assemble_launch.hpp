// assemble_launch.hpp - launchers of k_assemble, one translation unit per
// integral kind (asm_ol.hip, asm_ke.hip, asm_np.hip, asm_ep.hip) so the
// template instances compile in parallel.
#pragma once
#include <algorithm>
#include "assemble.hpp"
#include "ep_kernel.hpp"

namespace imago {

struct AssembleLaunch {
    DevSys S;
    const PairRec *pairs;
    const ImgRec *imgs;
    const double2 *phase;
    const CandRec *cands;
    const int *cand_off;
    int term0;
    double *out;
    long term_stride;
    int npairs, nterms;
    size_t lds;
    hipStream_t stream;
    int lmax, ne;
    bool gamma;
    const int *type_term0;  // EP: first term index of each potential type
    int num_pot_types;
    const int *order;       // EP: pairs ordered by column atom, then row atom
    EpTuple *ep_tuples;     // EP: screened tuples of every (type, pair)
    const long *ep_tseg;    // EP: [types * pairs + 1] offsets into ep_tuples
    int *ep_item_tuple;     // EP: item -> tuple
    double *ep_blocks;      // EP: primitive block of every item
    long ep_nitems;
    int nk;
    // NP with f shells: each work-group's 16 x 16 x 64 accumulator block in
    // device memory (kAccSlots work-groups per launch)
    double *acc_work = nullptr;
};

constexpr int kAccSlots = 1024;
constexpr size_t kAccSlotDoubles = 16 * 16 * 64;

hipError_t launch_assemble_ol(const AssembleLaunch &L);
hipError_t launch_assemble_ke(const AssembleLaunch &L);
hipError_t launch_assemble_np(const AssembleLaunch &L);
hipError_t launch_assemble_ep(const AssembleLaunch &L);
hipError_t launch_ep_screen(const AssembleLaunch &L, bool fill, long *tseg, long *iseg,
                            size_t lds);
hipError_t launch_ep_prims(const AssembleLaunch &L);

// direct (panel-free) EP contributions, asm_epd.hip

// A run of items of one atom pair: positions [i_lo, i_hi) of the pair's
// image-sorted items (pitem).  The host cuts every pair's items into runs
// of at most kEpdRunItems, so an on-site pair with thousands of items does
// not hold one wavefront for the length of the whole launch.
constexpr int kEpdRunItems = 128;
struct EpdSub {
    int pair, t_lo, t_hi, pad;
    long i_lo, i_hi;
    long out;   // H sum: offset (doubles) of the run's partial block
};
struct EpdLaunch {
    DevSys S;
    const PairRec *pairs;
    const double2 *phase;
    const EpTuple *tuples;
    const int *item_tuple;    // [items] tuple of each item
    const int *pitem;         // [items] items sorted by (pair, lattice image)
    const double *blocks;
    const double *coef_eff;   // hsum: per term, 0.0 replaced by 1.0
    const double *X;          // trace: folded density [image][ent] (rw for Gamma)
    int ent;                  // entries per image of X (64 * ne)
    const int *type_term0;
    const EpdSub *subs;       // runs of items, one wavefront each
    long nsubs;
    double *out;              // hsum: run partial blocks; trace: one value per item
    int npairs, lmax, ne;
    bool gamma;
    hipStream_t stream;
};
hipError_t launch_epd_hsum(const EpdLaunch &L);
hipError_t launch_epd_trace(const EpdLaunch &L);
hipError_t launch_epd_hfinal(const EpdLaunch &L, const double *np, const double *ke,
                             const long *psub, const long *woff, double *H);
hipError_t launch_epd_fold(const EpdLaunch &L, const double *rw, double *X);
// part: [types][kEpdTraceChunks][max_pot_alphas] scratch of the per-term sums
hipError_t launch_epd_trace_final(const double *ival, const unsigned char *jexp,
                                  const long *iseg, int npairs, int ntypes, int max_pot_alphas,
                                  const int *pot_nalpha, const int *term_type,
                                  const int *term_alpha, int pot_dim, double *part, double *out,
                                  hipStream_t stream);
hipError_t launch_epd_item_keys(const EpTuple *tuples, const int *item_tuple,
                                const PairRec *pairs, long nitems, int npairs, int *key, int *val,
                                unsigned char *jexp, hipStream_t stream);
hipError_t launch_epd_pair_runs(bool fill, const int *skey, long nitems, const PairRec *pairs,
                                int npairs, const DevSys &S, long C, long *hcnt, long *wcnt,
                                long *tcnt, const long *hoff, const long *woff, const long *toff,
                                EpdSub *hruns, EpdSub *truns, hipStream_t stream);
// stable sort of n (key, value) int pairs on keys[0, 2^bits)
hipError_t int_sort_pairs(const int *kin, int *kout, const int *vin, int *vout, long n, int bits,
                          void *tmp, size_t *tmp_bytes, hipStream_t stream);
// exclusive prefix sum of n longs (tmp == nullptr: only sets *tmp_bytes)
hipError_t ep_exclusive_scan(const long *in, long *out, long n, void *tmp, size_t *tmp_bytes,
                             hipStream_t stream);
hipError_t int_exclusive_scan(const int *in, int *out, long n, void *tmp, size_t *tmp_bytes,
                              hipStream_t stream);

template <int KIND, int LMAX, int NE, bool G>
hipError_t launch_one(const AssembleLaunch &L)
{
    if constexpr (kLaneAccInMemory<KIND, LMAX>) {
        // one term; the pairs in launches of at most kAccSlots work-groups,
        // work-group x of a launch owning slot x of acc_work
        if (!L.acc_work || L.nterms != 1) return hipErrorInvalidValue;
        for (int p0 = 0; p0 < L.npairs; p0 += kAccSlots) {
            const int np = std::min(kAccSlots, L.npairs - p0);
            hipLaunchKernelGGL((k_assemble<KIND, LMAX, NE, G>), dim3(np, 1), dim3(64), L.lds,
                               L.stream, L.S, L.pairs + p0, L.imgs, L.phase, L.cands, L.cand_off,
                               L.term0, L.out, L.term_stride, L.acc_work);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    hipLaunchKernelGGL((k_assemble<KIND, LMAX, NE, G>), dim3(L.npairs, L.nterms), dim3(64), L.lds,
                       L.stream, L.S, L.pairs, L.imgs, L.phase, L.cands, L.cand_off, L.term0,
                       L.out, L.term_stride, nullptr);
    return hipGetLastError();
}

template <int LMAX, int NE, int NKR, bool G>
hipError_t launch_ep_one(const AssembleLaunch &L)
{
    const long total = (long)L.npairs * L.nterms;
    const long per = (total + kNumXcd - 1) / kNumXcd;
    // at most 2^24 work-groups of 64 (a multiple of kNumXcd); the kernel
    // strides over the rest
    const long groups = std::min(per * kNumXcd, 1L << 24);
    hipLaunchKernelGGL((k_ep_store<LMAX, NE, NKR, G>), dim3((unsigned)groups), dim3(64),
                       L.lds, L.stream, L.S, L.pairs, L.phase, L.ep_tuples, L.ep_tseg,
                       L.ep_blocks, L.order, L.npairs, L.out, L.term_stride, L.term0, L.nterms);
    return hipGetLastError();
}

template <int LMAX, int NE>
hipError_t launch_ep_ne(const AssembleLaunch &L)
{
    if (L.gamma) return launch_ep_one<LMAX, NE, 1, true>(L);
    if (L.nk <= 4) return launch_ep_one<LMAX, NE, 4, false>(L);
    return launch_ep_one<LMAX, NE, 0, false>(L);
}

template <int UNUSED = 0>
hipError_t launch_ep(const AssembleLaunch &L)
{
#define IMAGO_EP_CASE(LM)                                                        \
    case LM:                                                                     \
        return L.ne == 4 ? launch_ep_ne<LM, 4>(L) : launch_ep_ne<LM, 16>(L);
    switch (L.lmax) {
        IMAGO_EP_CASE(0)
        IMAGO_EP_CASE(1)
        IMAGO_EP_CASE(2)
        IMAGO_EP_CASE(3)
    default: return hipErrorInvalidValue;
    }
#undef IMAGO_EP_CASE
}

// Screening pass of the EP terms: count (fill == false) writes the tuple
// and item counts of every (type, pair) task into tseg / iseg; fill writes
// the tuples and the item -> tuple map from the offsets in tseg / iseg.
template <int UNUSED = 0>
hipError_t launch_ep_screen_pass(const AssembleLaunch &L, bool fill, long *tseg, long *iseg,
                                 size_t lds)
{
    const unsigned tasks = (unsigned)((long)L.npairs * L.num_pot_types);
    if (fill)
        hipLaunchKernelGGL(k_ep_screen<true>, dim3(tasks), dim3(64), lds, L.stream, L.S, L.pairs,
                           L.imgs, L.cands, L.cand_off, L.npairs, tseg, iseg, L.ep_tuples,
                           L.ep_item_tuple);
    else
        hipLaunchKernelGGL(k_ep_screen<false>, dim3(tasks), dim3(64), lds, L.stream, L.S, L.pairs,
                           L.imgs, L.cands, L.cand_off, L.npairs, tseg, iseg, L.ep_tuples,
                           L.ep_item_tuple);
    return hipGetLastError();
}

// Primitive blocks of every item (k_ep_prims), one lane per item.
template <int UNUSED = 0>
hipError_t launch_ep_prims_pass(const AssembleLaunch &L)
{
    if (L.ep_nitems == 0) return hipSuccess;
    const unsigned g = (unsigned)((L.ep_nitems + 63) / 64);
#define IMAGO_EPP_CASE(LM)                                                                    \
    case LM:                                                                                  \
        hipLaunchKernelGGL(k_ep_prims<LM>, dim3(g), dim3(64), ep_prims_lds_bytes(LM),         \
                           L.stream, L.S, L.pairs,                                            \
                           L.imgs, L.cands, L.ep_tuples, L.ep_item_tuple, L.ep_nitems,         \
                           L.npairs, L.ep_blocks);                                            \
        return hipGetLastError();
    switch (L.lmax) {
        IMAGO_EPP_CASE(0)
        IMAGO_EPP_CASE(1)
        IMAGO_EPP_CASE(2)
        IMAGO_EPP_CASE(3)
    default: return hipErrorInvalidValue;
    }
#undef IMAGO_EPP_CASE
}

template <int KIND>
hipError_t launch_kind(const AssembleLaunch &L)
{
#define IMAGO_CASE(LM)                                                          \
    case LM:                                                                    \
        if (L.ne == 4) return L.gamma ? launch_one<KIND, LM, 4, true>(L)        \
                                      : launch_one<KIND, LM, 4, false>(L);      \
        return L.gamma ? launch_one<KIND, LM, 16, true>(L)                      \
                       : launch_one<KIND, LM, 16, false>(L);
    switch (L.lmax) {
        IMAGO_CASE(0)
        IMAGO_CASE(1)
        IMAGO_CASE(2)
        IMAGO_CASE(3)
    default: return hipErrorInvalidValue;
    }
#undef IMAGO_CASE
}

}  // namespace imago
