// hetrd_coop.hpp - Householder reduction of a batch of complex Hermitian
// matrices to real symmetric tridiagonal form (LAPACK zhetd2, uplo = 'L':
// A = Q T Q^H, Q = H(0) H(1) ... H(n-2), H(j) = I - tau_j v_j v_j^H) in one
// persistent cooperative kernel.
//
// Each matrix lives in the registers of P = ceil(n / 16) work-groups for the
// whole reduction: work-group w holds rows 16w .. 16w+15, thread t of the
// work-group row 16w + (t & 15) at the columns (t >> 4) + 32 s, s < S.  One
// column step is two per-matrix barriers:
//   A. the work-group holding row j builds the reflector from conj(row j)
//      (zlarfg) and publishes v and tau;
//   B. every work-group forms p = tau A v on its rows and the partial
//      p^H v, and publishes both;
//   C. every work-group forms w = p - tau/2 (p^H v) v and applies
//      A -= v w^H + w v^H to its rows.
// Only v, p and the dot partials go through memory; the matrix does not.
// The tridiagonal (d, e), tau and the reflectors (LAPACK's lower layout,
// consumed by zunmtr) are written as they are produced.
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

constexpr int kTrdThreads = 512;   // 16 rows x 32 column groups
constexpr int kTrdRows = 16;
constexpr int kTrdGroups = 32;

struct TrdArgs {
    int n, P;             // order; work-groups per matrix (ceil(n / 16))
    const double2 *A;     // [nb][n * n] column-major, full Hermitian
    double2 *V;           // [nb][n * n] reflectors below the subdiagonal
    double *d, *e;        // [nb][n] diagonal, subdiagonal (e[n-1] = 0)
    double2 *tau;         // [nb][n] (tau[n-1] = 0)
    double2 *vbuf;        // [nb][n] published v (and tau at index j)
    double2 *pbuf;        // [nb][n] published p
    double2 *dbuf;        // [nb][P] published dot partials
    unsigned *bar;        // [nb] barrier counters, zero at launch
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

// Barrier of the P work-groups of one matrix: monotonic counter, target =
// P * (barriers passed so far).  Every thread first waits until its
// published stores have completed; one thread per work-group then counts
// in and polls.  A wait that exceeds spin_limit polls sets the counter's top
// bit (every later wait of that matrix then passes at once) and *err, so
// the kernel always ends.
__device__ __forceinline__ void trd_barrier(unsigned *cnt, unsigned target, int *err, int limit)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int polls = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++polls > limit) {
                __hip_atomic_fetch_or(cnt, 0x80000000u, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// grid = nb * P work-groups of kTrdThreads, all co-resident (cooperative
// launch); n <= 32 * S.
template <int S>
__global__ __launch_bounds__(kTrdThreads) void k_hetrd_coop(TrdArgs T)
{
    __shared__ double2 vl[kTrdGroups * S];   // v (tau at index j)
    __shared__ double2 pl[kTrdGroups * S];   // p
    __shared__ double2 red[kTrdThreads / 64][kTrdRows];
    __shared__ double2 prow[kTrdRows];
    __shared__ double2 bcast;
    const int n = T.n, P = T.P;
    const int m = blockIdx.x / P, w = blockIdx.x % P;
    const int t = threadIdx.x, r = t & (kTrdRows - 1), g = t >> 4;
    const int lane = t & 63, wave = t >> 6;
    const int i = kTrdRows * w + r;
    const size_t nn = (size_t)n * n;
    const double2 *A = T.A + m * nn;
    double2 *V = T.V + m * nn;
    double2 *vbuf = T.vbuf + (size_t)m * n, *pbuf = T.pbuf + (size_t)m * n;
    double2 *dbuf = T.dbuf + (size_t)m * P;
    unsigned *bar = T.bar + m;

    double2 a[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int c = g + kTrdGroups * s;
        a[s] = (i < n && c < n) ? A[(size_t)c * n + i] : trd_zero();
    }
    unsigned nbar = 0;
    for (int j = 0; j + 1 < n; ++j) {
        // ---- A: reflector from conj(row j) ----
        if (w == (j >> 4)) {
            const bool mine = r == (j & (kTrdRows - 1));
            double2 nd = trd_zero();      // (|x(2:)|^2, A(j,j))
            double2 al = trd_zero();      // alpha = conj(A(j, j+1))
            if (mine) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int c = g + kTrdGroups * s;
                    if (c >= j + 2 && c < n) nd.x += a[s].x * a[s].x + a[s].y * a[s].y;
                    if (c == j + 1) al = make_double2(a[s].x, -a[s].y);
                    if (c == j) nd.y = a[s].x;
                }
            }
            // the 32 threads of row j: lanes rj, rj+16, rj+32, rj+48 of each wave
            nd.x += __shfl_xor(nd.x, 16);
            nd.y += __shfl_xor(nd.y, 16);
            al = make_double2(al.x + __shfl_xor(al.x, 16), al.y + __shfl_xor(al.y, 16));
            nd.x += __shfl_xor(nd.x, 32);
            nd.y += __shfl_xor(nd.y, 32);
            al = make_double2(al.x + __shfl_xor(al.x, 32), al.y + __shfl_xor(al.y, 32));
            if (mine && lane < kTrdRows) {
                red[wave][0] = nd;
                red[wave][1] = al;
            }
            __syncthreads();
            if (mine) {
                double2 s2 = trd_zero(), alpha = trd_zero();
                for (int q = 0; q < kTrdThreads / 64; ++q) {
                    s2.x += red[q][0].x;
                    s2.y += red[q][0].y;
                    alpha.x += red[q][1].x;
                    alpha.y += red[q][1].y;
                }
                const double xn2 = s2.x, djj = s2.y;
                double beta, tr, ti;
                double2 scale;
                if (xn2 == 0.0 && alpha.y == 0.0) {
                    beta = alpha.x;
                    tr = ti = 0.0;
                    scale = trd_zero();
                } else {
                    beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
                    tr = (beta - alpha.x) / beta;
                    ti = -alpha.y / beta;
                    // 1 / (alpha - beta)
                    const double zr = alpha.x - beta, zi = alpha.y;
                    const double den = zr * zr + zi * zi;
                    scale = make_double2(zr / den, -zi / den);
                }
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int c = g + kTrdGroups * s;
                    if (c >= j + 1 && c < n)
                        vl[c] = c == j + 1 ? make_double2(1.0, 0.0)
                                           : trd_mul(scale, make_double2(a[s].x, -a[s].y));
                }
                if (g == 0) {
                    vl[j] = make_double2(tr, ti);
                    bcast = make_double2(beta, djj);
                }
            }
            __syncthreads();
            // publish v and tau; reflector column j (zhetrd 'L' layout)
            for (int c = j + t; c < n; c += kTrdThreads) {
                const double2 v = vl[c];
                pub_store(vbuf + c, v);
                if (c > j) V[(size_t)j * n + c] = v;
            }
            if (t == 0) {
                T.tau[(size_t)m * n + j] = vl[j];
                T.e[(size_t)m * n + j] = bcast.x;
                T.d[(size_t)m * n + j] = bcast.y;
            }
        }
        trd_barrier(bar, (unsigned)P * ++nbar, T.err, T.spin_limit);

        // ---- B: p = tau A v on this work-group's rows, partial p^H v ----
        for (int c = j + t; c < n; c += kTrdThreads) vl[c] = pub_load(vbuf + c);
        __syncthreads();
        const double2 tau = vl[j];
        const bool row_on = i > j && i < n;
        double2 sp = trd_zero();
        if (row_on) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int c = g + kTrdGroups * s;
                if (c > j && c < n) {
                    const double2 q = trd_mul(a[s], vl[c]);
                    sp.x += q.x;
                    sp.y += q.y;
                }
                // keep the scheduler from hoisting every LDS read at once
                if ((s & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        sp = make_double2(sp.x + __shfl_xor(sp.x, 16), sp.y + __shfl_xor(sp.y, 16));
        sp = make_double2(sp.x + __shfl_xor(sp.x, 32), sp.y + __shfl_xor(sp.y, 32));
        if (lane < kTrdRows) red[wave][lane] = sp;
        __syncthreads();
        if (t < kTrdRows) {
            double2 p = trd_zero();
            for (int q = 0; q < kTrdThreads / 64; ++q) {
                p.x += red[q][t].x;
                p.y += red[q][t].y;
            }
            p = row_on ? trd_mul(tau, p) : trd_zero();
            prow[t] = p;
            if (row_on) pub_store(pbuf + i, p);
            // conj(p_i) v_i summed over the 16 rows (lanes 0..15 of wave 0),
            // the same tree in every work-group
            double2 dp = row_on ? trd_mulc(vl[i], p) : trd_zero();
            for (int o = 8; o >= 1; o >>= 1) {
                const double2 x = trd_shfl_xor(dp, o);
                dp = make_double2(dp.x + x.x, dp.y + x.y);
            }
            if (t == 0) pub_store(dbuf + w, dp);
        }
        trd_barrier(bar, (unsigned)P * ++nbar, T.err, T.spin_limit);

        // ---- C: w = p - tau/2 (p^H v) v; A -= v w^H + w v^H ----
        for (int c = j + 1 + t; c < n; c += kTrdThreads) pl[c] = pub_load(pbuf + c);
        if (wave == 0) {
            double2 x = lane < P ? pub_load(dbuf + lane) : trd_zero();
            for (int o = 32; o >= 1; o >>= 1) {
                const double2 y = trd_shfl_xor(x, o);
                x = make_double2(x.x + y.x, x.y + y.y);
            }
            if (lane == 0) bcast = x;
        }
        __syncthreads();
        if (row_on) {
            const double2 dot = bcast;
            const double2 a2 = trd_mul(make_double2(-0.5 * tau.x, -0.5 * tau.y), dot);
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
        __syncthreads();
    }
    // last diagonal entry
    if (i == n - 1) {
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (g + kTrdGroups * s == n - 1) {
                T.d[(size_t)m * n + n - 1] = a[s].x;
                T.e[(size_t)m * n + n - 1] = 0.0;
                T.tau[(size_t)m * n + n - 1] = make_double2(0.0, 0.0);
            }
    }
}

}  // namespace imago
