// hetrd_coop.hpp - Householder reduction of a batch of Hermitian (complex)
// or symmetric (real) matrices to real symmetric tridiagonal form (LAPACK
// zhetd2 / dsytd2, uplo = 'L': A = Q T Q^H, Q = H(0) H(1) ... H(n-2),
// H(j) = I - tau_j v_j v_j^H) in one persistent cooperative kernel.
//
// Each matrix lives in the registers of P = ceil(n / 16) work-groups for the
// whole reduction: work-group w holds rows 16w .. 16w+15, thread t of the
// work-group row 16w + (t & 15) at the columns (t >> 4) + 32 s, s < S.  One
// column step is one exchange between the work-groups of a matrix:
//   every work-group holds the pivot row j of the current matrix in LDS and
//   builds the reflector from conj(row j) (zlarfg / dlarfg) itself,
//   bit-identically;
//   forms p = tau A v on its own rows and the partial p^H v;
//   publishes p, the partial and (the owner of row j+1) row j+1;
//   after the barrier reads p, the partials and row j+1, forms
//   w = p - tau/2 (p^H v) v, applies A -= v w^H + w v^H to its rows and
//   the same update to row j+1, the next pivot row.
// Only p, the dot partials and one row per step go through memory, double
// buffered by step parity.  The tridiagonal (d, e), tau and the reflectors
// (LAPACK's lower layout) are written by work-group 0.
#pragma once
#include <hip/hip_runtime.h>

namespace imago {

constexpr int kTrdThreads = 512;   // 16 rows x 32 column groups
constexpr int kTrdRows = 16;
constexpr int kTrdGroups = 32;
constexpr int kTrdMaxGroups = 128;   // work-groups per matrix (order <= 2048)

// scalar helpers for double (real symmetric) and double2 (complex)
__device__ __forceinline__ double2 trd_zero() { return make_double2(0.0, 0.0); }
template <class T> __device__ __forceinline__ T sc_zero();
template <> __device__ __forceinline__ double sc_zero<double>() { return 0.0; }
template <> __device__ __forceinline__ double2 sc_zero<double2>() { return trd_zero(); }
template <class T> __device__ __forceinline__ T sc_one();
template <> __device__ __forceinline__ double sc_one<double>() { return 1.0; }
template <> __device__ __forceinline__ double2 sc_one<double2>() { return make_double2(1.0, 0.0); }
template <class T> __device__ __forceinline__ T sc_make(double re, double im);
template <> __device__ __forceinline__ double sc_make<double>(double re, double) { return re; }
template <> __device__ __forceinline__ double2 sc_make<double2>(double re, double im)
{
    return make_double2(re, im);
}
__device__ __forceinline__ double sc_re(double a) { return a; }
__device__ __forceinline__ double sc_re(double2 a) { return a.x; }
__device__ __forceinline__ double sc_im(double) { return 0.0; }
__device__ __forceinline__ double sc_im(double2 a) { return a.y; }
__device__ __forceinline__ double sc_conj(double a) { return a; }
__device__ __forceinline__ double2 sc_conj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double sc_abs2(double a) { return a * a; }
__device__ __forceinline__ double sc_abs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double sc_add(double a, double b) { return a + b; }
__device__ __forceinline__ double2 sc_add(double2 a, double2 b)
{
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double sc_sub(double a, double b) { return a - b; }
__device__ __forceinline__ double2 sc_sub(double2 a, double2 b)
{
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double sc_mul(double a, double b) { return a * b; }
__device__ __forceinline__ double2 sc_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double sc_mulc(double a, double b) { return a * b; }
__device__ __forceinline__ double2 sc_mulc(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double sc_scale(double s, double a) { return s * a; }
__device__ __forceinline__ double2 sc_scale(double s, double2 a)
{
    return make_double2(s * a.x, s * a.y);
}
__device__ __forceinline__ double sc_shfl_xor(double v, int m) { return __shfl_xor(v, m); }
__device__ __forceinline__ double2 sc_shfl_xor(double2 v, int m)
{
    return make_double2(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
}

// The vectors exchanged between work-groups (p, pivot rows, dot partials)
// are written and read with agent-scope atomic accesses, which are coherent
// across the XCDs' L2 caches without a cache-wide write-back or invalidate.
__device__ __forceinline__ void pub_word(double *p, double v)
{
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p),
                       (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pub_word_load(const double *p)
{
    const unsigned long long x = __hip_atomic_load(
        reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)x);
}
__device__ __forceinline__ void pub_store(double *p, double v) { pub_word(p, v); }
__device__ __forceinline__ void pub_store(double2 *p, double2 v)
{
    pub_word(reinterpret_cast<double *>(p), v.x);
    pub_word(reinterpret_cast<double *>(p) + 1, v.y);
}
__device__ __forceinline__ double pub_load(const double *p) { return pub_word_load(p); }
__device__ __forceinline__ double2 pub_load(const double2 *p)
{
    const double *q = reinterpret_cast<const double *>(p);
    return make_double2(pub_word_load(q), pub_word_load(q + 1));
}

template <class T>
struct TrdArgs {
    int n, P;             // order; work-groups per matrix (ceil(n / 16))
    int nb_launch;        // matrices of this launch
    const T *A;           // [nb][n * n] column-major, full Hermitian / symmetric
    T *V;                 // [nb][n * n] reflectors below the diagonal, zero above
    double *d, *e;        // [nb][n] diagonal, subdiagonal (e[n-1] = 0)
    T *tau;               // [nb][n] (tau[n-1] = 0)
    T *vbuf;              // [nb][2][n] published pivot rows
    T *pbuf;              // [nb][2][n] published p
    T *dbuf;              // [nb][2][kTrdMaxGroups] published dot partials
    unsigned *bar;        // [nb][P][kTrdFlagStride] barrier flags, then one abort
                          // word; zero at launch
    int *err;             // set to 1 when a barrier wait times out
    int spin_limit;       // polls before a barrier wait gives up
};

// Barrier of the P work-groups of one matrix without a shared counter:
// work-group w stores the number of the step it has published into its own
// flag (one 64-byte line each) after every thread's published stores have
// completed; the lanes of wave 0 poll the P flags of the matrix in
// parallel (two per lane above 64) until all have reached the step.  Every
// 64 polls the abort word is read; a wait longer than limit polls sets it
// (every later wait of the launch then passes within 64 polls) and *err,
// so the kernel always ends.
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
        unsigned *f0 = flags + (size_t)(t < P ? t : 0) * kTrdFlagStride;
        unsigned *f1 = flags + (size_t)(t + 64 < P ? t + 64 : 0) * kTrdFlagStride;
        const bool two = P > 64;
        int polls = 0;
        while (true) {
            const unsigned v0 = __hip_atomic_load(f0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned v1 =
                two ? __hip_atomic_load(f1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : step;
            if (__all(v0 >= step && v1 >= step)) break;
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
// launch); n <= 32 * S and P <= kTrdMaxGroups.
template <class T, int S>
__global__ __launch_bounds__(kTrdThreads) void k_hetrd_coop(TrdArgs<T> A_)
{
    constexpr int NW = kTrdThreads / 64;
    __shared__ T vl[kTrdGroups * S];   // v of this step
    __shared__ T pl[kTrdGroups * S];   // p of this step
    __shared__ T rl[kTrdGroups * S];   // the pivot row
    __shared__ T red[NW][kTrdRows];
    __shared__ T prow[kTrdRows];
    __shared__ double nrm[NW];         // partial |x(2:)|^2 per wave
    __shared__ T dot;                  // p^H v
    const int n = A_.n, P = A_.P;
    const int m = blockIdx.x / P, w = blockIdx.x % P;
    const int t = threadIdx.x, r = t & (kTrdRows - 1), g = t >> 4;
    const int lane = t & 63, wave = t >> 6;
    const int i = kTrdRows * w + r;
    const size_t nn = (size_t)n * n;
    const T *A = A_.A + m * nn;
    T *V = A_.V + m * nn;
    T *rbuf = A_.vbuf + (size_t)m * 2 * n, *pbuf = A_.pbuf + (size_t)m * 2 * n;
    T *dbuf = A_.dbuf + (size_t)m * 2 * kTrdMaxGroups;
    unsigned *flags = A_.bar + (size_t)m * P * kTrdFlagStride;
    unsigned *abortw = A_.bar + (size_t)A_.nb_launch * P * kTrdFlagStride;
    const bool out = w == 0;   // work-group 0 writes d, e, tau and V

    T a[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int c = g + kTrdGroups * s;
        a[s] = (i < n && c < n) ? A[(size_t)c * n + i] : sc_zero<T>();
    }
    // pivot row 0: A(0, c) for every column
    for (int c = t; c < n; c += kTrdThreads) rl[c] = A[(size_t)c * n];
    __syncthreads();
    unsigned nstep = 0;
    for (int j = 0; j + 1 < n; ++j) {
        const int slot = j & 1;
        // ---- reflector from conj(row j): |x(2:)|^2 over the work-group ----
        double x2 = 0.0;
        for (int c = j + 2 + t; c < n; c += kTrdThreads) x2 += sc_abs2(rl[c]);
        x2 = trd_wave_sum(x2);
        if (lane == 0) nrm[wave] = x2;
        __syncthreads();
        // every thread forms the reflector from the same partials, so all
        // work-groups and threads hold bit-identical beta, tau and scale
        T tau, scale;
        double beta;
        {
            double xn2 = 0.0;
            for (int q = 0; q < NW; ++q) xn2 += nrm[q];
            const T alpha = sc_conj(rl[j + 1]);
            const double ar = sc_re(alpha), ai = sc_im(alpha);
            if (xn2 == 0.0 && ai == 0.0) {
                beta = ar;
                tau = sc_zero<T>();
                scale = sc_zero<T>();
            } else {
                beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
                tau = sc_make<T>((beta - ar) / beta, -ai / beta);
                // 1 / (alpha - beta)
                const double zr = ar - beta, zi = ai;
                const double den = zr * zr + zi * zi;
                scale = sc_make<T>(zr / den, -zi / den);
            }
        }
        // v into LDS for the dot, the update and the next pivot row (read
        // after the next work-group barrier); the hemv forms v from rl
        for (int c = j + 1 + t; c < n; c += kTrdThreads) {
            const T v = c == j + 1 ? sc_one<T>() : sc_mul(scale, sc_conj(rl[c]));
            vl[c] = v;
            if (out) V[(size_t)j * n + c] = v;
        }
        if (out && t == 0) {
            A_.tau[(size_t)m * n + j] = tau;
            A_.e[(size_t)m * n + j] = beta;
            A_.d[(size_t)m * n + j] = sc_re(rl[j]);
        }

        // ---- p = tau A v on this work-group's rows, partial p^H v ----
        const bool row_on = i > j && i < n;
        T sp = sc_zero<T>();
        if (row_on) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int c = g + kTrdGroups * s;
                if (c > j && c < n) {
                    const T vc = c == j + 1 ? sc_one<T>() : sc_mul(scale, sc_conj(rl[c]));
                    sp = sc_add(sp, sc_mul(a[s], vc));
                }
                if ((s & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        sp = sc_add(sp, sc_shfl_xor(sp, 16));
        sp = sc_add(sp, sc_shfl_xor(sp, 32));
        if (lane < kTrdRows) red[wave][lane] = sp;
        __syncthreads();
        if (t < kTrdRows) {
            T p = sc_zero<T>();
            for (int q = 0; q < NW; ++q) p = sc_add(p, red[q][t]);
            p = row_on ? sc_mul(tau, p) : sc_zero<T>();
            prow[t] = p;
            if (row_on) pub_store(pbuf + slot * n + i, p);
            // conj(p_i) v_i over the 16 rows, the same tree in every work-group
            T dp = row_on ? sc_mulc(vl[i], p) : sc_zero<T>();
            for (int o = 8; o >= 1; o >>= 1) dp = sc_add(dp, sc_shfl_xor(dp, o));
            if (t == 0) pub_store(dbuf + slot * kTrdMaxGroups + w, dp);
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
        trd_barrier(flags, w, P, ++nstep, abortw, A_.err, A_.spin_limit);

        // ---- w = p - tau/2 (p^H v) v; A -= v w^H + w v^H; next pivot row ----
        for (int c = j + 1 + t; c < n; c += kTrdThreads) {
            pl[c] = pub_load(pbuf + slot * n + c);
            rl[c] = pub_load(rbuf + slot * n + c);
        }
        if (wave == 0) {
            // partials of work-groups lane and lane + 64, then a fixed tree
            T x = lane < P ? pub_load(dbuf + slot * kTrdMaxGroups + lane) : sc_zero<T>();
            if (lane + 64 < P) x = sc_add(x, pub_load(dbuf + slot * kTrdMaxGroups + lane + 64));
            for (int o = 32; o >= 1; o >>= 1) x = sc_add(x, sc_shfl_xor(x, o));
            if (lane == 0) dot = x;
        }
        __syncthreads();
        const T a2 = sc_mul(sc_scale(-0.5, tau), dot);
        if (row_on) {
            const T vi = vl[i];
            const T wi = sc_add(prow[r], sc_mul(a2, vi));
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int c = g + kTrdGroups * s;
                if (c > j && c < n) {
                    const T vc = vl[c];
                    const T wc = sc_add(pl[c], sc_mul(a2, vc));
                    a[s] = sc_sub(a[s], sc_add(sc_mulc(vi, wc), sc_mulc(wi, vc)));
                }
                if ((s & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        // row j+1 of the updated matrix (v_{j+1} = 1): the next pivot row
        {
            const T wj1 = sc_add(pl[j + 1], a2);
            __syncthreads();
            for (int c = j + 1 + t; c < n; c += kTrdThreads) {
                const T vc = vl[c];
                const T wc = sc_add(pl[c], sc_mul(a2, vc));
                rl[c] = sc_sub(rl[c], sc_add(sc_conj(wc), sc_mulc(wj1, vc)));
            }
        }
        __syncthreads();
    }
    if (out && t == 0) {
        A_.d[(size_t)m * n + n - 1] = sc_re(rl[n - 1]);
        A_.e[(size_t)m * n + n - 1] = 0.0;
        A_.tau[(size_t)m * n + n - 1] = sc_zero<T>();
    }
}

}  // namespace imago
