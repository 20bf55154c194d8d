// prim_batch.hip - the primitive seam of gaussIntegrals.f90 evaluated on the
// GPU for arbitrary tuples (imago_primitive_batch, imago_boys_batch).
#include "common.hpp"
#include "prim.hpp"

using namespace imago;

struct imago_ctx;
int imago_ctx_device(imago_ctx *c);
hipStream_t imago_ctx_stream(imago_ctx *c);

// ---------------------------------------------------------------------------
// primitive seam (tests of the device primitives against the reference)

struct GlobalAcc {
    double *o;
    __device__ __forceinline__ void add(int i, int j, double x) { o[i + 16 * j] += x; }
};

template <int KIND>
__global__ __launch_bounds__(64) void k_primitive_batch(int n, const double *__restrict__ a,
                                                        const double *__restrict__ xyz,
                                                        const int *__restrict__ sw,
                                                        double *__restrict__ sh)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double *o = sh + (size_t)256 * t;
    for (int q = 0; q < 256; ++q) o[q] = 0.0;
    const int s = sw[t];
    auto nib = [](int v) { return v == 1 ? 0 : v == 2 ? 1 : v == 4 ? 2 : 3; };
    const int la = nib(s & 15), lb = nib((s >> 4) & 15);
    const Vec3 A = {xyz[9 * t], xyz[9 * t + 1], xyz[9 * t + 2]};
    const Vec3 B = {xyz[9 * t + 3], xyz[9 * t + 4], xyz[9 * t + 5]};
    const Vec3 C = {xyz[9 * t + 6], xyz[9 * t + 7], xyz[9 * t + 8]};
    GlobalAcc acc{o};
    prim_dispatch<KIND, 3>(la, lb, a[3 * t], a[3 * t + 1], a[3 * t + 2], A, B, C, 1.0, acc);
}

__global__ void k_boys_batch(int n, const double *__restrict__ T, double *__restrict__ F)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double f[7];
    boys_ref<6>(T[t], f);
    for (int m = 0; m < 7; ++m) F[7 * t + m] = f[m];
}

extern "C" int imago_primitive_batch(imago_ctx *c, int kind, int n, const double *a,
                                     const double *xyz, const int *l1l2switch, double *sh)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!a || !xyz || !l1l2switch || !sh)))
        return fail(IMAGO_E_ARG, "bad primitive batch arguments");
    if (kind < IMAGO_OL || kind > IMAGO_EP) return fail(IMAGO_E_ARG, "unknown kind %d", kind);
    for (int t = 0; t < n; ++t) {
        int s = l1l2switch[t], lo = s & 15, hi = (s >> 4) & 15;
        auto okn = [](int v) { return v == 1 || v == 2 || v == 4 || v == 8; };
        if (!okn(lo) || !okn(hi) || (s >> 8) != 0)
            return fail(IMAGO_E_ARG, "invalid l1l2switch %d at tuple %d", s, t);
    }
    if (n == 0) return IMAGO_OK;
    HIPCHK(hipSetDevice(imago_ctx_device(c)));
    double *da, *dx, *dsh;
    int *dsw;
    HIPCHK(hipMalloc(&da, sizeof(double) * 3 * n));
    HIPCHK(hipMalloc(&dx, sizeof(double) * 9 * n));
    HIPCHK(hipMalloc(&dsw, sizeof(int) * n));
    HIPCHK(hipMalloc(&dsh, sizeof(double) * 256 * n));
    HIPCHK(hipMemcpy(da, a, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dx, xyz, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dsw, l1l2switch, sizeof(int) * n, hipMemcpyHostToDevice));
    dim3 grid((n + 63) / 64), block(64);
    switch (kind) {
    case IMAGO_OL: hipLaunchKernelGGL(k_primitive_batch<K_OL>, grid, block, 0, imago_ctx_stream(c), n, da, dx, dsw, dsh); break;
    case IMAGO_KE: hipLaunchKernelGGL(k_primitive_batch<K_KE>, grid, block, 0, imago_ctx_stream(c), n, da, dx, dsw, dsh); break;
    case IMAGO_NP: hipLaunchKernelGGL(k_primitive_batch<K_NP>, grid, block, 0, imago_ctx_stream(c), n, da, dx, dsw, dsh); break;
    default: hipLaunchKernelGGL(k_primitive_batch<K_EP>, grid, block, 0, imago_ctx_stream(c), n, da, dx, dsw, dsh); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(imago_ctx_stream(c)));
    HIPCHK(hipMemcpy(sh, dsh, sizeof(double) * 256 * n, hipMemcpyDeviceToHost));
    (void)hipFree(da);
    (void)hipFree(dx);
    (void)hipFree(dsw);
    (void)hipFree(dsh);
    return IMAGO_OK;
}

extern "C" int imago_boys_batch(imago_ctx *c, int n, const double *T, double *F7)
{
    if (!c) return fail(IMAGO_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!T || !F7))) return fail(IMAGO_E_ARG, "bad Boys batch arguments");
    if (n == 0) return IMAGO_OK;
    HIPCHK(hipSetDevice(imago_ctx_device(c)));
    double *dT, *dF;
    HIPCHK(hipMalloc(&dT, sizeof(double) * n));
    HIPCHK(hipMalloc(&dF, sizeof(double) * 7 * n));
    HIPCHK(hipMemcpy(dT, T, sizeof(double) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_boys_batch, dim3((n + 255) / 256), dim3(256), 0, imago_ctx_stream(c), n, dT, dF);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(imago_ctx_stream(c)));
    HIPCHK(hipMemcpy(F7, dF, sizeof(double) * 7 * n, hipMemcpyDeviceToHost));
    (void)hipFree(dT);
    (void)hipFree(dF);
    return IMAGO_OK;
}

