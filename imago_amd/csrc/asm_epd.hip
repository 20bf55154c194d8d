// asm_epd.hip - launchers of the direct (panel-free) EP kernels.
#include "assemble_launch.hpp"
#include "ep_direct.hpp"

namespace imago {

static size_t epd_basis_lds(const EpdLaunch &L)
{
    return (size_t)2 * L.S.max_states * L.S.max_alphas * sizeof(double);
}

template <int LMAX, int NE, bool G>
static hipError_t hsum_one(const EpdLaunch &L)
{
    if (L.nsubs == 0) return hipSuccess;
    const int nk = G ? 1 : L.S.nk;
    for (int k0 = 0; k0 < nk; k0 += 4)
        hipLaunchKernelGGL((k_epd_hsum<LMAX, NE, G>), dim3((unsigned)L.nsubs), dim3(64),
                           epd_basis_lds(L), L.stream, L.S, L.pairs, L.phase, L.tuples,
                           L.item_tuple, L.pitem, L.blocks, L.coef_eff, L.type_term0, L.subs,
                           L.nsubs, L.npairs, L.out, k0);
    return hipGetLastError();
}

template <int LMAX, int NE, bool G>
static hipError_t trace_one(const EpdLaunch &L)
{
    if (L.nsubs == 0) return hipSuccess;
    hipLaunchKernelGGL((k_epd_trace<LMAX, NE, G>), dim3((unsigned)L.nsubs), dim3(64),
                       epd_basis_lds(L), L.stream, L.S, L.pairs, L.tuples, L.item_tuple, L.pitem,
                       L.blocks, L.X, L.ent, L.subs, L.nsubs, L.npairs, L.out);
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

hipError_t launch_epd_fold(const EpdLaunch &L, const double *rw, double *X)
{
    if (L.npairs <= 0 || L.gamma) return hipSuccess;
    hipLaunchKernelGGL(k_epd_fold, dim3(L.npairs), dim3(256), 0, L.stream, L.S, L.pairs, L.phase,
                       reinterpret_cast<const double2 *>(rw), L.ent, X);
    return hipGetLastError();
}

hipError_t launch_epd_trace_final(const double *ival, const unsigned char *jexp,
                                  const long *iseg, int npairs, int ntypes, int max_pot_alphas,
                                  const int *pot_nalpha, const int *term_type,
                                  const int *term_alpha, int pot_dim, double *part, double *out,
                                  hipStream_t stream)
{
    if (ntypes <= 0 || pot_dim <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_epd_trace_chunks, dim3((unsigned)ntypes * kEpdTraceChunks), dim3(64),
                       (size_t)max_pot_alphas * 64 * sizeof(double), stream, ival, jexp, iseg,
                       npairs, pot_nalpha, max_pot_alphas, part);
    hipLaunchKernelGGL(k_epd_trace_final, dim3((unsigned)((pot_dim + 255) / 256)), dim3(256), 0,
                       stream, part, term_type, term_alpha, pot_dim, max_pot_alphas, out);
    return hipGetLastError();
}

hipError_t launch_epd_item_keys(const EpTuple *tuples, const int *item_tuple,
                                const PairRec *pairs, long nitems, int npairs, int *key, int *val,
                                unsigned char *jexp, hipStream_t stream)
{
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_epd_item_keys, dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0,
                       stream, tuples, item_tuple, pairs, nitems, npairs, key, val, jexp);
    return hipGetLastError();
}

hipError_t launch_epd_pair_runs(bool fill, const int *skey, long nitems, const PairRec *pairs,
                                int npairs, const DevSys &S, long C, long *hcnt, long *wcnt,
                                long *tcnt, const long *hoff, const long *woff, const long *toff,
                                EpdSub *hruns, EpdSub *truns, hipStream_t stream)
{
    if (npairs <= 0) return hipSuccess;
    const unsigned g = (unsigned)((npairs + 63) / 64);
    if (fill)
        hipLaunchKernelGGL(k_epd_pair_runs<true>, dim3(g), dim3(64), 0, stream, skey, nitems,
                           pairs, npairs, S, C, hcnt, wcnt, tcnt, hoff, woff, toff, hruns, truns);
    else
        hipLaunchKernelGGL(k_epd_pair_runs<false>, dim3(g), dim3(64), 0, stream, skey, nitems,
                           pairs, npairs, S, C, hcnt, wcnt, tcnt, hoff, woff, toff, hruns, truns);
    return hipGetLastError();
}

}  // namespace imago
