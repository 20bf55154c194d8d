// hetrd.hip - launcher of the cooperative tridiagonal reduction
// (hetrd_coop.hpp).
#include "hetrd_coop.hpp"
#include "hetrd_launch.hpp"

namespace imago {

namespace {

template <int S>
hipError_t trd_one(const TrdLaunch &L)
{
    const void *fn = reinterpret_cast<const void *>(&k_hetrd_coop<S>);
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t st = hipGetDevice(&dev);
    if (st != hipSuccess) return st;
    st = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (st != hipSuccess) return st;
    st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kTrdThreads, 0);
    if (st != hipSuccess) return st;
    const int P = (L.n + kTrdRows - 1) / kTrdRows;
    const int fit = per_cu * ncu / P;   // matrices whose work-groups are co-resident
    if (fit < 1) return hipErrorCooperativeLaunchTooLarge;
    const size_t nn = (size_t)L.n * L.n;
    for (int m0 = 0; m0 < L.nb; m0 += fit) {
        const int nb = std::min(fit, L.nb - m0);
        st = hipMemsetAsync(L.bar, 0, sizeof(unsigned) * nb, L.stream);
        if (st != hipSuccess) return st;
        TrdArgs T;
        T.n = L.n;
        T.P = P;
        T.A = L.A + m0 * nn;
        T.V = L.V + m0 * nn;
        T.d = L.d + (size_t)m0 * L.n;
        T.e = L.e + (size_t)m0 * L.n;
        T.tau = L.tau + (size_t)m0 * L.n;
        T.vbuf = L.vbuf;
        T.pbuf = L.pbuf;
        T.dbuf = L.dbuf;
        T.bar = L.bar;
        T.err = L.err;
        T.spin_limit = 2000000;
        void *args[] = {&T};
        st = hipLaunchCooperativeKernel(fn, dim3((unsigned)(nb * P)), dim3(kTrdThreads), args, 0,
                                        L.stream);
        if (st != hipSuccess) return st;
    }
    return hipSuccess;
}

}  // namespace

bool hetrd_coop_supported(int n) { return n >= 2 && n <= kTrdMaxOrder; }

hipError_t launch_hetrd_coop(const TrdLaunch &L)
{
    if (!hetrd_coop_supported(L.n) || L.nb < 1) return hipErrorInvalidValue;
    if (L.n <= 128) return trd_one<4>(L);
    if (L.n <= 256) return trd_one<8>(L);
    if (L.n <= 512) return trd_one<16>(L);
    return trd_one<32>(L);
}

}  // namespace imago
