// assemble_launch.hpp - launchers of k_assemble, one translation unit per
// integral kind (asm_ol.hip, asm_ke.hip, asm_np.hip, asm_ep.hip) so the
// template instances compile in parallel.
#pragma once
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
};

hipError_t launch_assemble_ol(const AssembleLaunch &L);
hipError_t launch_assemble_ke(const AssembleLaunch &L);
hipError_t launch_assemble_np(const AssembleLaunch &L);
hipError_t launch_assemble_ep(const AssembleLaunch &L);

template <int KIND, int LMAX, int NE, bool G>
hipError_t launch_one(const AssembleLaunch &L)
{
    hipLaunchKernelGGL((k_assemble<KIND, LMAX, NE, G>), dim3(L.npairs, L.nterms), dim3(64), L.lds,
                       L.stream, L.S, L.pairs, L.imgs, L.phase, L.cands, L.cand_off, L.term0,
                       L.out, L.term_stride);
    return hipGetLastError();
}

template <int LMAX, int NE, bool G>
hipError_t launch_ep_one(const AssembleLaunch &L)
{
    const long total = (long)L.npairs * L.num_pot_types;
    const long per = (total + kNumXcd - 1) / kNumXcd;
    hipLaunchKernelGGL((k_assemble_ep<LMAX, NE, G>), dim3((unsigned)(per * kNumXcd)), dim3(64),
                       L.lds, L.stream, L.S, L.pairs, L.imgs, L.phase, L.cands, L.cand_off,
                       L.type_term0, L.order, L.npairs, L.out, L.term_stride);
    return hipGetLastError();
}

template <int UNUSED = 0>
hipError_t launch_ep(const AssembleLaunch &L)
{
#define IMAGO_EP_CASE(LM)                                                        \
    case LM:                                                                     \
        if (L.ne == 4) return L.gamma ? launch_ep_one<LM, 4, true>(L)            \
                                      : launch_ep_one<LM, 4, false>(L);          \
        return L.gamma ? launch_ep_one<LM, 16, true>(L) : launch_ep_one<LM, 16, false>(L);
    switch (L.lmax) {
        IMAGO_EP_CASE(0)
        IMAGO_EP_CASE(1)
        IMAGO_EP_CASE(2)
        IMAGO_EP_CASE(3)
    default: return hipErrorInvalidValue;
    }
#undef IMAGO_EP_CASE
}

template <int KIND>
hipError_t launch_kind(const AssembleLaunch &L)
{
    // nuclear integrals with f shells are refused by imago_assemble, so the
    // LMAX = 3 instance is not built for them
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
    case 3:
        if constexpr (KIND != K_NP) {
            if (L.ne == 4) return L.gamma ? launch_one<KIND, 3, 4, true>(L)
                                          : launch_one<KIND, 3, 4, false>(L);
            return L.gamma ? launch_one<KIND, 3, 16, true>(L) : launch_one<KIND, 3, 16, false>(L);
        }
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
#undef IMAGO_CASE
}

}  // namespace imago
