// hetrd_coop.hpp - Householder reduction of a batch of complex Hermitian
// matrices to real symmetric tridiagonal form (LAPACK zhetd2, uplo = 'L':
// A = Q T Q^H, Q = H(0) H(1) ... H(n-2), H(j) = I - tau_j v_j v_j^H) in one
// persistent cooperative kernel.
//
// Each matrix lives in the registers of P = ceil(n / 16) work-groups for the
// whole reduction: work-group w holds rows 16w .. 16w+15, thread t of the
// work-group row 16w + (t & 15) at the columns (t >> 4) + 32 s, s < S.  One
// column step is one exchange between the work-groups of a matrix:
//   every work-group holds the pivot row j of the current matrix in LDS and
//   builds the reflector from conj(row j) (zlarfg) itself, bit-identically;
//   forms p = tau A v on its own rows and the partial p^H v;
//   publishes p, the partial and (the owner of row j+1) row j+1;
//   after the barrier reads p, the partials and row j+1, forms
//   w = p - tau/2 (p^H v) v, applies A -= v w^H + w v^H to its rows and
//   the same update to row j+1, the next pivot row.
// Only p, the dot partials and one row per step go through memory, double
// buffered by step parity.  The tridiagonal (d, e), tau and the reflectors
// (LAPACK's lower layout, consumed by zunmtr) are written by work-group 0.
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

constexpr int kTrdThreads = 512;   // 16 rows x 32 column groups
constexpr int kTrdRows = 16;
constexpr int kTrdGroups = 32;

struct TrdArgs {
    int n, P;             // order; work-groups per matrix (ceil(n / 16))
    int nb_launch;        // matrices of this launch
    const double2 *A;     // [nb][n * n] column-major, full Hermitian
    double2 *V;           // [nb][n * n] reflectors below the subdiagonal
    double *d, *e;        // [nb][n] diagonal, subdiagonal (e[n-1] = 0)
    double2 *tau;         // [nb][n] (tau[n-1] = 0)
    double2 *vbuf;        // [nb][2][n] published pivot rows
    double2 *pbuf;        // [nb][2][n] published p
    double2 *dbuf;        // [nb][2][64] published dot partials
    unsigned *bar;        // [nb][P][kTrdFlagStride] barrier flags, then one abort
                          // word; zero at launch
    int *err;             // set to 1 when a barrier wait times out
    int spin_limit;       // polls before a barrier wait gives up
};

__device__ __forceinline__ double2 trd_zero() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 trd_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 trd_mulc(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 trd_shfl_xor(double2 v, int m)
{
    return make_double2(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
}

// The vectors exchanged between work-groups (v, p, dot partials) are
// written and read with agent-scope atomic accesses, which are coherent
// across the XCDs' L2 caches without a cache-wide write-back or invalidate.
__device__ __forceinline__ void pub_store(double2 *p, double2 v)
{
    auto *q = reinterpret_cast<unsigned long long *>(p);
    __hip_atomic_store(q, (unsigned long long)__double_as_longlong(v.x), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, (unsigned long long)__double_as_longlong(v.y), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double2 pub_load(const double2 *p)
{
    auto *q = reinterpret_cast<const unsigned long long *>(p);
    const unsigned long long x =
        __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long y =
        __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_double2(__longlong_as_double((long long)x), __longlong_as_double((long long)y));
}

// Barrier of the P work-groups of one matrix without a shared counter:
// work-group w stores the number of the step it has published into its own
// flag (one 64-byte line each) after every thread's published stores have
// completed; lanes 0 .. P-1 of wave 0 poll the P flags of the matrix in
// parallel until all have reached the step.  Every 64 polls the abort word
// is read; a wait longer than limit polls sets it (every later wait of the
// launch then passes within 64 polls) and *err, so the kernel always ends.
constexpr int kTrdFlagStride = 16;   // unsigned per flag: one 64-byte line
__device__ __forceinline__ void trd_barrier(unsigned *flags, int w, int P, unsigned step,
                                            unsigned *abortw, int *err, int limit)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int t = threadIdx.x;
    if (t == 0)
        __hip_atomic_store(flags + (size_t)w * kTrdFlagStride, step, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (t < 64) {
        unsigned *f = flags + (size_t)(t < P ? t : 0) * kTrdFlagStride;
        int polls = 0;
        while (true) {
            const unsigned v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__all(v >= step)) break;
            if ((++polls & 63) == 0) {
                if (__hip_atomic_load(abortw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
                    break;
                if (polls > limit) {
                    if (t == 0) {
                        __hip_atomic_store(abortw, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// wave sum of a double; every lane gets the total
__device__ __forceinline__ double trd_wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// grid = nb * P work-groups of kTrdThreads, all co-resident (cooperative
// launch); n <= 32 * S.
template <int S>
__global__ __launch_bounds__(kTrdThreads) void k_hetrd_coop(TrdArgs T)
{
    constexpr int NW = kTrdThreads / 64;
    __shared__ double2 vl[kTrdGroups * S];   // v of this step
    __shared__ double2 pl[kTrdGroups * S];   // p of this step
    __shared__ double2 rl[kTrdGroups * S];   // the pivot row
    __shared__ double2 red[NW][kTrdRows];
    __shared__ double2 prow[kTrdRows];
    __shared__ double nrm[kTrdThreads / 64];  // partial |x(2:)|^2 per wave
    __shared__ double2 bc[4];                // [3]: the dot p^H v
    const int n = T.n, P = T.P;
    const int m = blockIdx.x / P, w = blockIdx.x % P;
    const int t = threadIdx.x, r = t & (kTrdRows - 1), g = t >> 4;
    const int lane = t & 63, wave = t >> 6;
    const int i = kTrdRows * w + r;
    const size_t nn = (size_t)n * n;
    const double2 *A = T.A + m * nn;
    double2 *V = T.V + m * nn;
    double2 *rbuf = T.vbuf + (size_t)m * 2 * n, *pbuf = T.pbuf + (size_t)m * 2 * n;
    double2 *dbuf = T.dbuf + (size_t)m * 2 * 64;
    unsigned *flags = T.bar + (size_t)m * P * kTrdFlagStride;
    unsigned *abortw = T.bar + (size_t)T.nb_launch * P * kTrdFlagStride;
    const bool out = w == 0;   // work-group 0 writes d, e, tau and V

    double2 a[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int c = g + kTrdGroups * s;
        a[s] = (i < n && c < n) ? A[(size_t)c * n + i] : trd_zero();
    }
    // pivot row 0: A(0, c) for every column
    for (int c = t; c < n; c += kTrdThreads) rl[c] = A[(size_t)c * n];
    __syncthreads();
    unsigned nstep = 0;
    for (int j = 0; j + 1 < n; ++j) {
        const int slot = j & 1;
        // ---- reflector from conj(row j): |x(2:)|^2 over the work-group ----
        double x2 = 0.0;
        for (int c = j + 2 + t; c < n; c += kTrdThreads) x2 += rl[c].x * rl[c].x + rl[c].y * rl[c].y;
        x2 = trd_wave_sum(x2);
        if (lane == 0) nrm[wave] = x2;
        __syncthreads();
        // every thread forms the reflector from the same partials, so all
        // work-groups and threads hold bit-identical beta, tau and scale
        double2 tau, scale;
        double beta;
        {
            double xn2 = 0.0;
            for (int q = 0; q < NW; ++q) xn2 += nrm[q];
            const double2 alpha = make_double2(rl[j + 1].x, -rl[j + 1].y);
            if (xn2 == 0.0 && alpha.y == 0.0) {
                beta = alpha.x;
                tau = trd_zero();
                scale = trd_zero();
            } else {
                beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
                tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
                const double zr = alpha.x - beta, zi = alpha.y;
                const double den = zr * zr + zi * zi;
                scale = make_double2(zr / den, -zi / den);
            }
        }
        // v into LDS for the dot, the update and the next pivot row (read
        // after the next work-group barrier); the hemv forms v from rl
        for (int c = j + 1 + t; c < n; c += kTrdThreads) {
            const double2 v = c == j + 1 ? make_double2(1.0, 0.0)
                                         : trd_mul(scale, make_double2(rl[c].x, -rl[c].y));
            vl[c] = v;
            if (out) V[(size_t)j * n + c] = v;
        }
        if (out && t == 0) {
            T.tau[(size_t)m * n + j] = tau;
            T.e[(size_t)m * n + j] = beta;
            T.d[(size_t)m * n + j] = rl[j].x;
        }

        // ---- p = tau A v on this work-group's rows, partial p^H v ----
        const bool row_on = i > j && i < n;
        double2 sp = trd_zero();
        if (row_on) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int c = g + kTrdGroups * s;
                if (c > j && c < n) {
                    const double2 vc = c == j + 1 ? make_double2(1.0, 0.0)
                                                  : trd_mul(scale, make_double2(rl[c].x, -rl[c].y));
                    const double2 q = trd_mul(a[s], vc);
                    sp.x += q.x;
                    sp.y += q.y;
                }
                if ((s & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        sp = make_double2(sp.x + __shfl_xor(sp.x, 16), sp.y + __shfl_xor(sp.y, 16));
        sp = make_double2(sp.x + __shfl_xor(sp.x, 32), sp.y + __shfl_xor(sp.y, 32));
        if (lane < kTrdRows) red[wave][lane] = sp;
        __syncthreads();
        if (t < kTrdRows) {
            double2 p = trd_zero();
            for (int q = 0; q < NW; ++q) {
                p.x += red[q][t].x;
                p.y += red[q][t].y;
            }
            p = row_on ? trd_mul(tau, p) : trd_zero();
            prow[t] = p;
            if (row_on) pub_store(pbuf + slot * n + i, p);
            // conj(p_i) v_i over the 16 rows, the same tree in every work-group
            double2 dp = row_on ? trd_mulc(vl[i], p) : trd_zero();
            for (int o = 8; o >= 1; o >>= 1) {
                const double2 x = trd_shfl_xor(dp, o);
                dp = make_double2(dp.x + x.x, dp.y + x.y);
            }
            if (t == 0) pub_store(dbuf + slot * 64 + w, dp);
        }
        // the owner of row j+1 publishes it (columns j+1 ..), staged in LDS
        // so that every thread issues a few of the stores
        if (w == ((j + 1) >> 4)) {
            if (r == ((j + 1) & (kTrdRows - 1))) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int c = g + kTrdGroups * s;
                    if (c > j && c < n) pl[c] = a[s];
                }
            }
            __syncthreads();
            for (int c = j + 1 + t; c < n; c += kTrdThreads) pub_store(rbuf + slot * n + c, pl[c]);
        }
        trd_barrier(flags, w, P, ++nstep, abortw, T.err, T.spin_limit);

        // ---- w = p - tau/2 (p^H v) v; A -= v w^H + w v^H; next pivot row ----
        for (int c = j + 1 + t; c < n; c += kTrdThreads) {
            pl[c] = pub_load(pbuf + slot * n + c);
            rl[c] = pub_load(rbuf + slot * n + c);
        }
        if (wave == 0) {
            double2 x = lane < P ? pub_load(dbuf + slot * 64 + lane) : trd_zero();
            for (int o = 32; o >= 1; o >>= 1) {
                const double2 y = trd_shfl_xor(x, o);
                x = make_double2(x.x + y.x, x.y + y.y);
            }
            if (lane == 0) bc[3] = x;
        }
        __syncthreads();
        const double2 a2 = trd_mul(make_double2(-0.5 * tau.x, -0.5 * tau.y), bc[3]);
        if (row_on) {
            const double2 vi = vl[i];
            const double2 pi = prow[r];
            const double2 q0 = trd_mul(a2, vi);
            const double2 wi = make_double2(pi.x + q0.x, pi.y + q0.y);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int c = g + kTrdGroups * s;
                if (c > j && c < n) {
                    const double2 vc = vl[c];
                    const double2 q1 = trd_mul(a2, vc);
                    const double2 wc = make_double2(pl[c].x + q1.x, pl[c].y + q1.y);
                    const double2 u1 = trd_mulc(vi, wc), u2 = trd_mulc(wi, vc);
                    a[s].x -= u1.x + u2.x;
                    a[s].y -= u1.y + u2.y;
                }
                if ((s & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        // row j+1 of the updated matrix (v_{j+1} = 1): the next pivot row
        {
            const double2 q0 = a2;   // a2 * v_{j+1}
            const double2 wj1 = make_double2(pl[j + 1].x + q0.x, pl[j + 1].y + q0.y);
            __syncthreads();
            for (int c = j + 1 + t; c < n; c += kTrdThreads) {
                const double2 vc = vl[c];
                const double2 q1 = trd_mul(a2, vc);
                const double2 wc = make_double2(pl[c].x + q1.x, pl[c].y + q1.y);
                const double2 u1 = make_double2(wc.x, -wc.y);   // 1 * conj(w_c)
                const double2 u2 = trd_mulc(wj1, vc);
                rl[c] = make_double2(rl[c].x - (u1.x + u2.x), rl[c].y - (u1.y + u2.y));
            }
        }
        __syncthreads();
    }
    if (out && t == 0) {
        T.d[(size_t)m * n + n - 1] = rl[n - 1].x;
        T.e[(size_t)m * n + n - 1] = 0.0;
        T.tau[(size_t)m * n + n - 1] = trd_zero();
    }
}

}  // namespace imago
