// asm_ep.hip - EP screening and evaluation kernel instances.
#include <hipcub/hipcub.hpp>

#include "assemble_launch.hpp"

namespace imago {
hipError_t launch_assemble_ep(const AssembleLaunch &L) { return launch_ep<>(L); }
hipError_t launch_ep_screen(const AssembleLaunch &L, bool fill, long *tseg, long *iseg,
                            size_t lds)
{
    return launch_ep_screen_pass<>(L, fill, tseg, iseg, lds);
}
hipError_t launch_ep_prims(const AssembleLaunch &L) { return launch_ep_prims_pass<>(L); }
hipError_t ep_exclusive_scan(const long *in, long *out, long n, void *tmp, size_t *tmp_bytes,
                             hipStream_t stream)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, *tmp_bytes, in, out, (int)n, stream);
}
hipError_t int_exclusive_scan(const int *in, int *out, long n, void *tmp, size_t *tmp_bytes,
                              hipStream_t stream)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, *tmp_bytes, in, out, (int)n, stream);
}
hipError_t int_sort_pairs(const int *kin, int *kout, const int *vin, int *vout, long n, int bits,
                          void *tmp, size_t *tmp_bytes, hipStream_t stream)
{
    return hipcub::DeviceRadixSort::SortPairs(tmp, *tmp_bytes, kin, kout, vin, vout, (int)n, 0,
                                              bits, stream);
}
}  // namespace imago
