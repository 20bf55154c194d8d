// asm_epd.hip - launchers of the direct (panel-free) EP kernels.
#include "assemble_launch.hpp"
#include "ep_direct.hpp"

namespace imago {

template <int LMAX, int NE, bool G>
static hipError_t hsum_one(const EpdLaunch &L)
{
    const long total = (long)L.npairs * L.ngroups;
    hipLaunchKernelGGL((k_epd_hsum<LMAX, NE, G>), dim3((unsigned)total), dim3(64), 0, L.stream, L.S,
                       L.pairs, L.phase, L.tuples, L.tseg, L.blocks, L.coef_eff, L.type_term0,
                       L.order, L.npairs, L.ngroups, L.partial, L.group_stride);
    return hipGetLastError();
}

template <int LMAX, int NE, bool G>
static hipError_t trace_one(const EpdLaunch &L)
{
    const long total = (long)L.npairs * L.S.num_pot_types;
    hipLaunchKernelGGL((k_epd_trace<LMAX, NE, G>), dim3((unsigned)total), dim3(64),
                       (size_t)L.S.max_pot_alphas * 64 * sizeof(double), L.stream, L.S, L.pairs,
                       L.phase, L.tuples, L.tseg, L.blocks, L.rw, L.type_term0, L.order, L.npairs,
                       L.tpart);
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

hipError_t launch_epd_hfinal(const double *np, const double *ke, const double *partial, int ngroups,
                             long len, double *H, hipStream_t stream)
{
    hipLaunchKernelGGL(k_epd_hfinal, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, stream, np,
                       ke, partial, ngroups, len, H);
    return hipGetLastError();
}

hipError_t launch_epd_trace_final(const double *tpart, int npairs, int pot_dim, double *out,
                                  hipStream_t stream)
{
    if (pot_dim <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_epd_trace_final, dim3(pot_dim), dim3(256), 0, stream, tpart, npairs,
                       pot_dim, out);
    return hipGetLastError();
}

}  // namespace imago
