// asm_epd.hip - launchers of the direct (panel-free) EP kernels.
#include "assemble_launch.hpp"
#include "ep_direct.hpp"

namespace imago {

template <int LMAX, int NE, bool G>
static hipError_t hsum_one(const EpdLaunch &L)
{
    if (L.nsubs == 0) return hipSuccess;
    hipLaunchKernelGGL((k_epd_hsum<LMAX, NE, G>), dim3((unsigned)L.nsubs), dim3(64), 0, L.stream,
                       L.S, L.pairs, L.phase, L.tuples, L.tseg, L.iseg, L.blocks, L.coef_eff,
                       L.type_term0, L.subs, L.nsubs, L.npairs, L.out);
    return hipGetLastError();
}

template <int LMAX, int NE, bool G>
static hipError_t trace_one(const EpdLaunch &L)
{
    if (L.nsubs == 0) return hipSuccess;
    hipLaunchKernelGGL((k_epd_trace<LMAX, NE, G>), dim3((unsigned)L.nsubs), dim3(64),
                       (size_t)L.S.max_pot_alphas * 64 * sizeof(double), L.stream, L.S, L.pairs,
                       L.phase, L.tuples, L.tseg, L.iseg, L.blocks, L.rw, L.subs, L.nsubs,
                       L.npairs, L.out);
    return hipGetLastError();
}

#define IMAGO_EPD_LM(FN, LM)                                                                 \
    case LM:                                                                                 \
        if (L.ne == 4) return L.gamma ? FN<LM, 4, true>(L) : FN<LM, 4, false>(L);            \
        return L.gamma ? FN<LM, 16, true>(L) : FN<LM, 16, false>(L);
#define IMAGO_EPD_DISPATCH(FN)                                                               \
    switch (L.lmax) {                                                                        \
        IMAGO_EPD_LM(FN, 0)                                                                  \
        IMAGO_EPD_LM(FN, 1)                                                                  \
        IMAGO_EPD_LM(FN, 2)                                                                  \
        IMAGO_EPD_LM(FN, 3)                                                                  \
    default: return hipErrorInvalidValue;                                                    \
    }

hipError_t launch_epd_hsum(const EpdLaunch &L) { IMAGO_EPD_DISPATCH(hsum_one) }

hipError_t launch_epd_trace(const EpdLaunch &L) { IMAGO_EPD_DISPATCH(trace_one) }
#undef IMAGO_EPD_DISPATCH
#undef IMAGO_EPD_LM

hipError_t launch_epd_hfinal(const EpdLaunch &L, const double *np, const double *ke,
                             const long *psub, const long *woff, double *H)
{
    if (L.npairs <= 0) return hipSuccess;
    if (L.gamma)
        hipLaunchKernelGGL(k_epd_hfinal<true>, dim3(L.npairs), dim3(256), 0, L.stream, L.S,
                           L.pairs, np, ke, L.out, psub, woff, H);
    else
        hipLaunchKernelGGL(k_epd_hfinal<false>, dim3(L.npairs), dim3(256), 0, L.stream, L.S,
                           L.pairs, np, ke, L.out, psub, woff, H);
    return hipGetLastError();
}

hipError_t launch_epd_trace_final(const double *tsub, const long *roff, int npairs,
                                  const int *term_type, const int *term_alpha, int max_pot_alphas,
                                  int pot_dim, double *out, hipStream_t stream)
{
    if (pot_dim <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_epd_trace_final, dim3(pot_dim), dim3(256), 0, stream, tsub, roff, npairs,
                       term_type, term_alpha, max_pot_alphas, out);
    return hipGetLastError();
}

hipError_t launch_epd_trace_runs(bool fill, const long *iseg, long ntask, int npairs, long C,
                                 long *rcnt, long *rcnt2, const long *roff, const long *roff2,
                                 EpdSub *runs, hipStream_t stream)
{
    if (ntask <= 0) return hipSuccess;
    const unsigned g = (unsigned)((ntask + 255) / 256);
    if (fill)
        hipLaunchKernelGGL(k_epd_trace_runs<true>, dim3(g), dim3(256), 0, stream, iseg, ntask,
                           npairs, C, rcnt, rcnt2, roff, roff2, runs);
    else
        hipLaunchKernelGGL(k_epd_trace_runs<false>, dim3(g), dim3(256), 0, stream, iseg, ntask,
                           npairs, C, rcnt, rcnt2, roff, roff2, runs);
    return hipGetLastError();
}

hipError_t launch_epd_hsum_runs(bool fill, const long *iseg, int npairs, int ntypes, long C,
                                const PairRec *pairs, const DevSys &S, long *rcnt, long *wcnt,
                                const long *roff, const long *woff, EpdSub *runs,
                                hipStream_t stream)
{
    if (npairs <= 0) return hipSuccess;
    const unsigned g = (unsigned)((npairs + 63) / 64);
    if (fill)
        hipLaunchKernelGGL(k_epd_hsum_runs<true>, dim3(g), dim3(64), 0, stream, iseg, npairs,
                           ntypes, C, pairs, S, rcnt, wcnt, roff, woff, runs);
    else
        hipLaunchKernelGGL(k_epd_hsum_runs<false>, dim3(g), dim3(64), 0, stream, iseg, npairs,
                           ntypes, C, pairs, S, rcnt, wcnt, roff, woff, runs);
    return hipGetLastError();
}

}  // namespace imago
